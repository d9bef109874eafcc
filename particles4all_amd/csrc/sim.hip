// Host orchestration of libp4a: the reference's `class Sim` (src/sim.js)
// over HIP.  One handle = one scene on one GPU, with its own stream.
//
// Substep (sim.js:458-596), as launched here:
//   [bodyDrag] predict+count, scan, scatter (stable), ->parity flip,
//   [body centroids], I x (lambda, delta, [shape matching]),
//   xsph(+normals), tension+finalize (or finalize)
// The whole substep is captured once per parity as a hipGraph and
// replayed; the graph is rebuilt when anything it bakes in changes
// (particle count, grid, parameters, iteration count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/p4a.h"
#include "p4a_internal.h"

using namespace p4a;

namespace {

const double kPi = 3.14159265358979323846;

struct TimedLaunch {
  const char* name;
  hipEvent_t a, b;
};

}  // namespace

namespace p4a {

const double kPiU = 3.14159265358979323846;

// Sim.uploadParams (sim.js:371-413): FP64 host arithmetic, then f32.
Uniform make_uniform_from(const SceneConsts& s, double dt) {
  const p4a_params& p = s.params;
  const double h = s.h, d = p.spacing;
  Uniform U;
  std::memset(&U, 0, sizeof(U));
  U.minusOne = -1.0f;
  for (int k = 0; k < 3; k++) {
    U.boxMin[k] = 0.0f;
    U.boxMax[k] = (float)p.box[k];
    U.gridDim[k] = s.gridDim[k];
    U.clampMin[k] = (float)(0.5 * d);
    U.clampMax[k] = (float)(p.box[k] - 0.5 * d);
  }
  U.divGx = FastDiv((uint32_t)s.gridDim[0]);
  U.divGy = FastDiv((uint32_t)s.gridDim[1]);
  U.dt = (float)dt;
  U.h = (float)h;
  U.h2 = (float)(h * h);
  const double poly6 = 315.0 / (64.0 * kPiU * std::pow(h, 9));
  U.poly6 = (float)poly6;
  U.spikyGrad = (float)(-45.0 / (kPiU * std::pow(h, 6)));
  U.gravity = (float)p.gravity;
  U.mass = (float)s.mass;
  U.rho0 = (float)p.rest_density;
  U.invRho0 = (float)(1.0 / p.rest_density);
  U.cfmEps = (float)std::max(1e-9, p.cfm_epsilon_rel * s.denomRest);
  U.sCorrK = (float)(p.scorr_k / (s.denomRest * std::max(1, p.iterations)));
  const double rq = p.scorr_dq * h;
  const double tq = h * h - rq * rq;
  const double wq = poly6 * tq * tq * tq;
  U.sCorrWq = (float)(wq > 0 ? 1.0 / wq : 0.0);
  const double cq = wq > 0 ? poly6 / wq : 0.0;
  U.sCorrFast = (float)(-(p.scorr_k / (s.denomRest * std::max(1, p.iterations))) * cq * cq * cq * cq);
  U.xsphC = (float)p.xsph_c;
  U.n = (uint32_t)s.n;
  U.nCells = (uint32_t)s.nCells;
  U.nBoundary = p.no_boundary ? 0u : (uint32_t)s.nBoundary;
  U.nBodyParts = (uint32_t)s.nBodyParts;
  U.nBodies = (uint32_t)s.nBodies;
  U.bodyScale = (float)s.bodyScale;
  U.covScale = (float)s.covScale;
  U.omega = (float)p.omega;
  U.sorAverage = p.sor_average ? 1u : 0u;
  U.invDt = (float)(dt > 0 ? 1.0 / dt : 0.0);
  U.volume = (float)(s.mass / p.rest_density);
  U.tension = (float)p.surface_tension_k;
  U.cohes = (float)(32.0 / (kPiU * std::pow(h, 9)));
  U.cohesTerm = (float)(std::pow(h, 6) / 64.0);
  return U;
}

// sortBoundary (sim.js:255-286): stable counting sort of the wall samples
// into a simulation grid, in host FP64 arithmetic on the float samples.
void sort_boundary_grid(double h, const int* dim, int nCells, const float* pts, const float* psi,
                        int nb, std::vector<float>& sorted, std::vector<float>& sortedPsi,
                        std::vector<uint32_t>& start) {
  std::vector<uint32_t> counts(nCells + 2, 0), cell(std::max(nb, 1));
  for (int i = 0; i < nb; i++) {
    int c[3];
    for (int k = 0; k < 3; k++) {
      double f = std::floor((double)pts[i * 3 + k] / h);
      c[k] = (int)std::min((double)dim[k] - 1, std::max(0.0, f));
    }
    cell[i] = (uint32_t)((c[2] * dim[1] + c[1]) * dim[0] + c[0]);
    counts[cell[i]]++;
  }
  start.assign(nCells + 2, 0);
  uint32_t run = 0;
  for (int c = 0; c <= nCells; c++) {
    start[c] = run;
    run += counts[c];
  }
  start[nCells + 1] = run;
  std::vector<uint32_t> cursor(start);
  sorted.assign((size_t)std::max(1, nb) * 4, 0.0f);
  sortedPsi.assign((size_t)std::max(1, nb), 0.0f);
  for (int i = 0; i < nb; i++) {
    const uint32_t k = cursor[cell[i]]++;
    sorted[k * 4 + 0] = pts[i * 3 + 0];
    sorted[k * 4 + 1] = pts[i * 3 + 1];
    sorted[k * 4 + 2] = pts[i * 3 + 2];
    sorted[k * 4 + 3] = psi[i];
    sortedPsi[k] = psi[i];
  }
}

}  // namespace p4a

struct p4a_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool ready = false;
  bool ofsValid = false;  // FP16 records match the current pred buffers
  // neighbour masks: built since the last grid build or position change,
  // list launches since the build (their queue alternates)
  bool nbBuilt = false;
  int nbLaunch = 0;
  // solver iterations since the build: the moved queue of iteration k is
  // nbMoved[k & 1], filled by the delta (and shape matching) of k - 1; a
  // queue's counts are clean once a lambda list launch or the build has
  // emptied them
  int nbIter = 0;
  bool nbMovedClean[2] = {false, false};
  // the host state a captured substep leaves
  bool nbBuiltAfter[2] = {false, false};
  int nbLaunchAfter[2] = {0, 0};
  int nbIterAfter[2] = {0, 0};
  bool nbMovedCleanAfter[2][2] = {{false, false}, {false, false}};
  // the stream the moved targets' fallback walks run on beside the list
  // launches, and the events that fork and join it
  hipStream_t side = nullptr;
  hipEvent_t nbFork = nullptr, nbJoin = nullptr;
  int driftCount = 0;     // drift slots written since the last grid build
  p4a_params params{};
  // scene
  double h = 0, d = 0, mass = 0, denomRest = 0;
  int gridDim[3] = {1, 1, 1};
  int nCells = 1;
  int n = 0, cap = 0, nFluid = 0, nBodyParts = 0, nBodies = 0, nBoundary = 0, pourBudget = 0;
  // particle ids whose slot the sort keeps in slotOf: the body particles'
  // range when their ids are contiguous (fluid first, then the bodies, as
  // scenes are built), every id otherwise
  uint32_t slotIdLo = 0, slotIdCount = 0xffffffffu;
  double bodyScale = 0, covScale = 0;
  // state
  int parity = 0, gen = 0, lastSubsteps = 0;
  double simTime = 0, timeBank = 0, lastAdvanced = 0;
  float4* predCur = nullptr;
  Arrays A{};
  std::vector<void*> devAllocs;       // particle/body/stat arrays
  std::vector<void*> cellAllocs;      // grid arrays (depend on nCells)
  std::vector<void*> bndAllocs;       // wall arrays
  size_t bposCount = 0;
  // host mirrors for reads in the reference layout
  std::vector<uint32_t> bodyOfPidH;
  std::vector<float> restOfPidH;
  std::vector<int32_t> bodyCountH;
  // interaction
  int heldBody = -1;
  double heldTarget[3] = {0, 0, 0};
  double heldRate = 10, heldLimit = 4;
  // graphs
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  std::string graphKey[2];
  // stats
  double* statsHost = nullptr;
  bool statsValid = false;
  // timing
  bool timing = false;
  std::vector<TimedLaunch> timed;
  std::vector<hipEvent_t> eventPool;
  size_t eventNext = 0;
  // tile shapes
  TileCfg tile1{0}, tile2{0};
  // anisotropy outputs (allocated on first use)
  float4* smoothPos = nullptr;
  float4* anisoG = nullptr;
  uint32_t* anisoFallbacks = nullptr;
  int anisoCap = 0;
  std::vector<void*> anisoAllocs;
  // density field (SurfaceMesh.configure / generateField)
  bool fieldConfigured = false;
  int fieldVertDim[3] = {0, 0, 0};
  double fieldVoxel = 0.0, fieldOrigin[3] = {0.0, 0.0, 0.0};
  size_t fieldVerts = 0, fieldVertCap = 0;
  float* field = nullptr;
  unsigned long long* fieldAcc = nullptr;
  std::vector<void*> fieldAllocs;
  int fieldRecCap = 0;
  float4* fieldRec = nullptr;
  float4* fieldRecA = nullptr;
  float4* fieldRecB = nullptr;
  std::vector<void*> fieldRecAllocs;
  // surface extraction (smoothField, buildNormalField, march)
  float* field2 = nullptr;
  float* nfield = nullptr;
  uint32_t* meshBlockCount = nullptr;
  uint32_t* meshBlockSum = nullptr;
  uint32_t* meshBlockStart = nullptr;
  uint32_t* meshCounter = nullptr;  // [0] total triangles, [4..7] draw-indirect arguments
  bool normalReady = false;
  uint32_t meshMaxTris = 0;
  size_t meshVertCap = 0;  // triangles the vertex buffer holds
  uint2* meshVerts = nullptr;
  std::vector<void*> meshVertAllocs;
  // renderer exports: the per-slot phase (kept current once requested) and
  // the packed body poses
  bool exportPhase = false;
  uint4* phaseOut = nullptr;
  float4* bodyStat = nullptr;
  float4* bodyPacked = nullptr;
  bool bodyStatSet = false;
  std::vector<void*> exportAllocs;
};

namespace {

#define HIP_TRY(sim, expr)                                                      \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      (sim)->err = std::string(#expr) + ": " + hipGetErrorString(e_);           \
      return P4A_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)

int fail(p4a_sim* s, int code, const std::string& m) {
  if (s) s->err = m;
  return code;
}

template <class T>
int dalloc(p4a_sim* s, std::vector<void*>& owner, T** p, size_t count) {
  size_t bytes = std::max<size_t>(16, count * sizeof(T));
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    s->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
    return P4A_ERR_NOMEM;
  }
  e = hipMemsetAsync(q, 0, bytes, s->stream);
  if (e != hipSuccess) {
    hipFree(q);
    s->err = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
    return P4A_ERR_HIP;
  }
  owner.push_back(q);
  *p = (T*)q;
  return P4A_OK;
}

void free_list(std::vector<void*>& v) {
  for (void* q : v) hipFree(q);
  v.clear();
}

void drop_graphs(p4a_sim* s) {
  // every caller changes positions, the grid or the parameters behind the
  // masks' back
  s->nbBuilt = false;
  for (int k = 0; k < 2; k++) {
    if (s->graph[k]) hipGraphExecDestroy(s->graph[k]);
    s->graph[k] = nullptr;
    s->graphKey[k].clear();
  }
}

Uniform make_uniform(const p4a_sim* s, double dt) {
  SceneConsts c;
  c.params = s->params;
  c.h = s->h;
  c.mass = s->mass;
  c.denomRest = s->denomRest;
  c.bodyScale = s->bodyScale;
  c.covScale = s->covScale;
  for (int k = 0; k < 3; k++) c.gridDim[k] = s->gridDim[k];
  c.n = s->n;
  c.nCells = s->nCells;
  c.nBoundary = s->nBoundary;
  c.nBodyParts = s->nBodyParts;
  c.nBodies = s->nBodies;
  Uniform U = make_uniform_from(c, dt);
  U.slotIdLo = s->slotIdLo;
  U.slotIdCount = s->slotIdCount;
  return U;
}

hipEvent_t take_event(p4a_sim* s) {
  if (s->eventNext == s->eventPool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    s->eventPool.push_back(e);
  }
  return s->eventPool[s->eventNext++];
}

// Launch one kernel group, bracketed by events when timing is on.
template <class F>
int launch(p4a_sim* s, const char* name, F&& f) {
  hipEvent_t a = nullptr, b = nullptr;
  if (s->timing) {
    a = take_event(s);
    b = take_event(s);
    if (a) hipEventRecord(a, s->stream);
  }
  hipError_t e = f();
  if (e != hipSuccess) {
    s->err = std::string(name) + ": " + hipGetErrorString(e);
    return P4A_ERR_HIP;
  }
  if (s->timing && a && b) {
    hipEventRecord(b, s->stream);
    s->timed.push_back({name, a, b});
  }
  return P4A_OK;
}

#define LAUNCH(s, name, call)                          \
  do {                                                 \
    int rc_ = launch((s), (name), [&]() { return (call); }); \
    if (rc_) return rc_;                               \
  } while (0)

bool exact_math(const p4a_sim* s) { return s->params.math == P4A_MATH_EXACT; }

// The FP16-offset form exists in the fast arithmetic only.
const char* form_error(const p4a_params* p) {
  if (p->math != P4A_MATH_FAST && p->math != P4A_MATH_EXACT) return "unknown math";
  if (p->precision != P4A_PREC_FP32 && p->precision != P4A_PREC_FP16_OFFSET)
    return "unknown precision";
  if (p->precision == P4A_PREC_FP16_OFFSET && p->math == P4A_MATH_EXACT)
    return "the FP16-offset form runs with P4A_MATH_FAST only";
  if (!(p->neighbour_skin >= 0.0 && p->neighbour_skin <= 1.0))
    return "neighbour_skin must lie in [0, 1]";
  if (p->neighbour_min_particles < 0) return "neighbour_min_particles must be >= 0";
  return nullptr;
}
bool fp16_form(const p4a_sim* s) { return s->params.precision == P4A_PREC_FP16_OFFSET; }

// FP16 records of a pred buffer: ofs[1] belongs to predX, ofs[0] to the
// pred of the live parity.
HRec* ofs_of(const p4a_sim* s, const float4* pred) {
  return pred == s->A.predX ? s->A.ofs[1] : s->A.ofs[0];
}

// block size of the direct kernel for a tile variant (LDS variants map to 64)
int direct_threads(const TileCfg& c) { return c.variant == 0 ? 256 : c.variant == 1 ? 128 : 64; }

int stale_records(p4a_sim* s) {
  return fail(s, P4A_ERR_ARG,
              "FP16 records are stale: the FP16-offset form starts at a grid build "
              "(p4a_step or P4A_STAGE_GRID)");
}

// Drift slots of the stale grid: each position update records into the
// next slot (none past DRIFT_SLOTS); neighbour launches record nothing.
uint32_t* drift_slot(const p4a_sim* s) {
  return s->driftCount < DRIFT_SLOTS ? s->A.drift + s->driftCount : nullptr;
}

bool nb_on(const p4a_sim* s) {
  return s->params.neighbour_skin > 0.0 && s->n >= s->params.neighbour_min_particles;
}

// The second stream the moved targets' walks run on beside the list launches,
// with the events that fork and join it: made when the scene first uses the
// masks, before any capture (without it those walks run in order on the
// launch stream).
int ensure_side(p4a_sim* s) {
  if (s->side || !nb_on(s)) return P4A_OK;
  hipStream_t st = nullptr;
  hipEvent_t f = nullptr, j = nullptr;
  if (hipStreamCreate(&st) != hipSuccess ||
      hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
    if (f) hipEventDestroy(f);
    if (st) hipStreamDestroy(st);
    return fail(s, P4A_ERR_HIP, "side stream");
  }
  s->side = st;
  s->nbFork = f;
  s->nbJoin = j;
  return P4A_OK;
}

// Mask arguments of the next solver launch: BUILD for the first lambda
// after the masks were dropped, USE once built, OFF otherwise.
NbLaunch nb_launch(p4a_sim* s, bool lambda) {
  NbLaunch L;
  if (!nb_on(s) || (!s->nbBuilt && !lambda)) return L;
  const double S = s->params.neighbour_skin * s->h;
  NbArgs& a = L.args;
  a.mask = s->A.nbMask;
  a.ref = s->A.nbRef;
  a.ctl = s->A.nbCtl;
  a.queue[0] = s->A.nbQueue[0];
  a.queue[1] = s->A.nbQueue[1];
  a.moved[0] = s->A.nbMoved[0];
  a.moved[1] = s->A.nbMoved[1];
  a.mslot = s->nbIter & 1;
  L.side = s->side;
  L.fork = s->nbFork;
  L.join = s->nbJoin;
  a.stride = s->A.nbStride;
  a.segCap = nb_seg_cap((uint32_t)s->n);
  a.reach2 = (float)((s->h + S) * (s->h + S));
  // a margin of 1e-4 h covers the float rounding of the distances at the
  // build and at the use and of the travels.  The FP16-offset form marks
  // and walks quantized positions: its y and z offsets (at most about h
  // from the binned cell's centre for a particle that stays in it) are
  // rounded to half, within sqrt(2) h 2^-11 of the FP32 position, and two
  // particles, each at build and at use, take 4 such margins more.
  const double q = fp16_form(s) ? 4.0 * std::sqrt(2.0) * s->h * std::ldexp(1.0, -11) : 0.0;
  a.lim = (float)(S - 1e-4 * s->h - q);
  if (!s->nbBuilt) {
    L.mode = Nb::BUILD;
    a.slot = 0;
  } else {
    L.mode = Nb::USE;
    a.slot = s->nbLaunch & 1;
  }
  return L;
}

void nb_advance(p4a_sim* s, const NbLaunch& L) {
  if (L.mode == Nb::BUILD) {
    // the build empties every queue
    s->nbBuilt = true;
    s->nbLaunch = 0;
    s->nbIter = 0;
    s->nbMovedClean[0] = s->nbMovedClean[1] = true;
  } else if (L.mode == Nb::USE) {
    s->nbLaunch++;
  }
}

// Before a delta list launch: the moved queue it fills must start empty.
// In a substep the lambda before it has emptied it; a delta that follows
// another (the stage API) empties it with a fill kernel.
int nb_clean_moved(p4a_sim* s, const NbLaunch& L) {
  if (L.mode != Nb::USE) return P4A_OK;
  const int q = L.args.mslot ^ 1;
  if (!s->nbMovedClean[q])
    LAUNCH(s, "delta", launch_fill_u32(s->A.nbCtl + nb_moved_counts(q), 0u, NB_QSEG, s->stream));
  s->nbMovedClean[q] = false;
  return P4A_OK;
}

int launch_lambda_any(p4a_sim* s, const Uniform& U) {
  const Drift d;
  if (fp16_form(s) && !s->ofsValid) return stale_records(s);
  const NbLaunch nb = nb_launch(s, true);
  if (fp16_form(s)) {
    LAUNCH(s, "lambda", launch_lambda16(U, s->A, s->predCur, ofs_of(s, s->predCur),
                                        direct_threads(s->tile1), d, s->stream, nb));
  } else {
    LAUNCH(s, "lambda", launch_lambda(U, s->A, s->predCur, exact_math(s), s->tile1, d,
                                      s->stream, nullptr, nb));
  }
  // a lambda list launch empties the moved queue its iteration's delta fills
  if (nb.mode == Nb::USE) s->nbMovedClean[nb.args.mslot ^ 1] = true;
  nb_advance(s, nb);
  return P4A_OK;
}

int launch_delta_any(p4a_sim* s, const Uniform& U, float4* out) {
  Drift d;
  d.out = drift_slot(s);
  if (fp16_form(s) && !s->ofsValid) return stale_records(s);
  const NbLaunch nb = nb_launch(s, false);
  int rc = nb_clean_moved(s, nb);
  if (rc) return rc;
  if (fp16_form(s)) {
    LAUNCH(s, "delta", launch_delta16(U, s->A, s->predCur, ofs_of(s, s->predCur), out,
                                      ofs_of(s, out), direct_threads(s->tile1), d, s->stream,
                                      nb));
  } else {
    LAUNCH(s, "delta", launch_delta(U, s->A, s->predCur, out, exact_math(s), s->tile1, d,
                                    s->stream, nullptr, nb));
    s->ofsValid = false;
  }
  nb_advance(s, nb);
  if (nb.mode == Nb::USE) s->nbIter++;
  s->driftCount++;
  return P4A_OK;
}

int launch_body_solve_any(p4a_sim* s, const Uniform& U) {
  if (!fp16_form(s)) s->ofsValid = false;
  LAUNCH(s, "body_solve", launch_body_solve(U, s->A, s->predCur,
                                            fp16_form(s) ? ofs_of(s, s->predCur) : nullptr,
                                            drift_slot(s), s->stream,
                                            nb_on(s) && s->nbBuilt ? (s->nbIter & 1) : -1));
  s->driftCount++;
  return P4A_OK;
}

template <class T>
bool all_finite(const T* v, size_t count) {
  for (size_t k = 0; k < count; k++)
    if (!std::isfinite(v[k])) return false;
  return true;
}

float4* other_pred(p4a_sim* s, int par) {
  return s->predCur == s->A.pred[par] ? s->A.predX : s->A.pred[par];
}

// Grid build on the current parity: count is done by predict; scan and
// the stable scatter follow, then the parity flips (sim.js:475-492).
int run_grid(p4a_sim* s, const Uniform& U, bool moveVel) {
  const int par = s->parity;
  LAUNCH(s, "scan", launch_scan(U, s->A, s->stream));
  LAUNCH(s, "scatter", launch_scatter(U, s->A, par, moveVel, fp16_form(s), s->stream));
  s->ofsValid = fp16_form(s);
  s->nbBuilt = false;
  s->driftCount = 0;
  s->parity ^= 1;
  s->predCur = s->A.pred[s->parity];
  return P4A_OK;
}

int run_finish(p4a_sim* s, const Uniform& U) {
  const int par = s->parity;
  const bool ex = exact_math(s);
  const bool tension = s->params.surface_tension_k > 0;
  if (!ex)
    LAUNCH(s, "prep_finish", launch_prep_finish(U, s->A, s->predCur, s->A.pos[par], s->stream));
  const Drift d;
  // under the substep's neighbour masks both walk the marked candidates
  const NbLaunch nbx = nb_launch(s, false);
  LAUNCH(s, "xsph", launch_xsph(U, s->A, s->predCur, s->A.pos[par], ex, tension, s->tile2, d,
                                s->stream, nullptr, nbx));
  nb_advance(s, nbx);
  if (tension) {
    const NbLaunch nbt = nb_launch(s, false);
    LAUNCH(s, "tension", launch_tension(U, s->A, s->predCur, s->A.pos[par], s->A.vel[par], ex,
                                        s->tile2, d, s->stream, nullptr, nbt));
    nb_advance(s, nbt);
  } else
    LAUNCH(s, "finalize", launch_finalize(U, s->A, s->predCur, s->A.pos[par], s->A.vel[par],
                                          s->stream));
  return P4A_OK;
}

// One substep after the optional drag: predict .. finalize.
int run_substep_body(p4a_sim* s, const Uniform& U) {
  LAUNCH(s, "predict", launch_predict_count(U, s->A, s->parity, s->stream));
  int rc = run_grid(s, U, false);
  if (rc) return rc;
  const int par = s->parity;
  if (s->nBodies > 0) LAUNCH(s, "body_pre", launch_body_pre(U, s->A, s->predCur, s->stream));
  const int iters = std::max(1, s->params.iterations);
  for (int it = 0; it < iters; it++) {
    if ((rc = launch_lambda_any(s, U))) return rc;
    float4* out = other_pred(s, par);
    if ((rc = launch_delta_any(s, U, out))) return rc;
    s->predCur = out;
    if (s->nBodies > 0 && (rc = launch_body_solve_any(s, U))) return rc;
  }
  return run_finish(s, U);
}

std::string graph_key(const p4a_sim* s, const Uniform& U) {
  std::string k((const char*)&U, sizeof(U));
  k += std::to_string(s->params.iterations) + "/" + std::to_string(s->params.math) + "/" +
       std::to_string(s->params.precision) + "/" +
       std::to_string(s->params.surface_tension_k > 0) + "/" + std::to_string(s->tile1.variant) +
       "/" + std::to_string(s->tile2.variant) + "/" + std::to_string(s->params.neighbour_skin) +
       "/" + std::to_string(nb_on(s));
  return k;
}

int run_substep(p4a_sim* s, double dt) {
  const Uniform U = make_uniform(s, dt);
  if (s->heldBody >= 1 && s->nBodyParts > 0) {
    DragArgs D;
    for (int k = 0; k < 3; k++) D.goal[k] = (float)s->heldTarget[k];
    D.rate = (float)s->heldRate;
    D.held = (uint32_t)s->heldBody;
    D.blend = (float)(1.0 - std::exp(-std::max(dt, 0.0) / 0.02));
    D.limit = (float)s->heldLimit;
    LAUNCH(s, "body_drag", launch_body_drag(U, s->A, s->A.vel[s->parity], D, s->stream));
  }
  const int par0 = s->parity;
  if (!s->params.use_graph || s->timing) {
    int rc = run_substep_body(s, U);
    if (rc) return rc;
  } else {
    const std::string key = graph_key(s, U);
    if (!s->graph[par0] || s->graphKey[par0] != key) {
      if (s->graph[par0]) {
        hipGraphExecDestroy(s->graph[par0]);
        s->graph[par0] = nullptr;
      }
      HIP_TRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      int rc = run_substep_body(s, U);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s->stream, &g);
      if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(s, P4A_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&s->graph[par0], g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) return fail(s, P4A_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      s->graphKey[par0] = key;
      // capture only recorded the work; the host-side state advanced as if
      // it ran, so rewind it and replay the graph below.
      s->parity = par0;
      s->nbBuiltAfter[par0] = s->nbBuilt;
      s->nbLaunchAfter[par0] = s->nbLaunch;
      s->nbIterAfter[par0] = s->nbIter;
      s->nbMovedCleanAfter[par0][0] = s->nbMovedClean[0];
      s->nbMovedCleanAfter[par0][1] = s->nbMovedClean[1];
    }
    HIP_TRY(s, hipGraphLaunch(s->graph[par0], s->stream));
    // host-side bookkeeping of what the graph does
    s->parity = par0 ^ 1;
    s->nbBuilt = s->nbBuiltAfter[par0];
    s->nbLaunch = s->nbLaunchAfter[par0];
    s->nbIter = s->nbIterAfter[par0];
    s->nbMovedClean[0] = s->nbMovedCleanAfter[par0][0];
    s->nbMovedClean[1] = s->nbMovedCleanAfter[par0][1];
    const int iters = std::max(1, s->params.iterations);
    s->predCur = (iters % 2 == 0) ? s->A.pred[s->parity] : s->A.predX;
  }
  s->simTime += dt;
  return P4A_OK;
}

int run_stats(p4a_sim* s) {
  const Uniform U = make_uniform(s, 0.0);
  LAUNCH(s, "stats", launch_stats(U, s->A, s->A.vel[s->parity], s->stream));
  HIP_TRY(s, hipMemcpyAsync(s->statsHost, s->A.statsOut, 4 * sizeof(double),
                            hipMemcpyDeviceToHost, s->stream));
  s->statsValid = true;
  return P4A_OK;
}

void sort_boundary(const p4a_sim* s, const float* pts, const float* psi, int nb,
                   std::vector<float>& sorted, std::vector<float>& sortedPsi,
                   std::vector<uint32_t>& start) {
  sort_boundary_grid(s->h, s->gridDim, s->nCells, pts, psi, nb, sorted, sortedPsi, start);
}

// Cells per axis max(1, floor(box / h)); -1 when the grid would not fit an
// int32 cell index (NaN boxes included).
long grid_cells(const double box[3], double h, int dim[3]) {
  long cells = 1;
  for (int k = 0; k < 3; k++) {
    const double g = std::floor(box[k] / h);
    if (!(g < 2.0e9)) return -1;
    dim[k] = g >= 1 ? (int)g : 1;
    cells *= dim[k];
    if (cells > (long)INT32_MAX - 4) return -1;
  }
  return cells;
}

int set_grid(p4a_sim* s) {
  const long cells = grid_cells(s->params.box, s->h, s->gridDim);
  if (cells < 0) return fail(s, P4A_ERR_RANGE, "grid too large");
  s->nCells = (int)cells;
  return P4A_OK;
}

int alloc_cells(p4a_sim* s) {
  free_list(s->cellAllocs);
  const size_t nc = (size_t)s->nCells;
  int rc;
  if ((rc = dalloc(s, s->cellAllocs, &s->A.cellCount, nc + 1))) return rc;
  if ((rc = dalloc(s, s->cellAllocs, &s->A.cellStart, nc + 2))) return rc;
  if ((rc = dalloc(s, s->cellAllocs, &s->A.blockSum, nc / 2048 + 2))) return rc;
  if ((rc = dalloc(s, s->cellAllocs, &s->A.bcellStart, nc + 2))) return rc;
  return P4A_OK;
}

int upload_boundary(p4a_sim* s, const float* pts, const float* psi, int nb) {
  std::vector<float> sorted, sortedPsi;
  std::vector<uint32_t> start;
  sort_boundary(s, pts, psi, nb, sorted, sortedPsi, start);
  free_list(s->bndAllocs);
  int rc;
  if ((rc = dalloc(s, s->bndAllocs, &s->A.bpos, (size_t)std::max(1, nb)))) return rc;
  if ((rc = dalloc(s, s->bndAllocs, &s->A.bpsi, (size_t)std::max(1, nb)))) return rc;
  s->bposCount = (size_t)nb;
  s->nBoundary = nb;
  HIP_TRY(s, hipMemcpyAsync(s->A.bpos, sorted.data(), sorted.size() * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(s->A.bpsi, sortedPsi.data(), sortedPsi.size() * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(s->A.bcellStart, start.data(), start.size() * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return P4A_OK;
}

struct BufView {
  void* ptr;
  size_t bytes;
};

BufView buf_view(p4a_sim* s, int id) {
  const size_t n = (size_t)s->n, cap = (size_t)s->cap, nc = (size_t)s->nCells;
  const int par = s->parity;
  (void)cap;
  switch (id) {
    case P4A_BUF_POS: return {s->A.pos[par], n * 16};
    case P4A_BUF_VEL: return {s->A.vel[par], n * 16};
    case P4A_BUF_PRED: return {s->predCur, n * 16};
    case P4A_BUF_LAMBDA: return {s->A.lambda, n * 4};
    case P4A_BUF_DENSITY: return {s->A.density, n * 4};
    case P4A_BUF_SLOT: return {s->A.slot, n * 4};
    case P4A_BUF_CORR: return {s->A.corr, n * 16};
    case P4A_BUF_NORMAL: return {s->A.normal, n * 16};
    case P4A_BUF_CELL_START: return {s->A.cellStart, (nc + 2) * 4};
    case P4A_BUF_BPOS: return {s->A.bpos, s->bposCount * 16};
    case P4A_BUF_BCELL_START: return {s->A.bcellStart, (nc + 2) * 4};
    case P4A_BUF_BODY_CENTRE: return {s->A.bodyCentre, (size_t)s->nBodies * 16};
    case P4A_BUF_BODY_ROT: return {s->A.bodyRot, (size_t)s->nBodies * 48};
    case P4A_BUF_PID: return {s->A.pid[par], n * 4};
    case P4A_BUF_KEY: return {s->A.key, n * 4};
    case P4A_BUF_ORDER: return {s->A.order, n * 4};
    case P4A_BUF_BPSI: return {s->A.bpsi, s->bposCount * 4};
    case P4A_BUF_SMOOTH_POS: return {s->smoothPos, s->smoothPos ? n * 16 : 0};
    case P4A_BUF_ANISO_G: return {s->anisoG, s->anisoG ? n * 32 : 0};
    case P4A_BUF_OFS: return {ofs_of(s, s->predCur), n * 8};
    case P4A_BUF_CELL_OF_SLOT: return {s->A.cellOfSlot, n * 4};
    case P4A_BUF_NEIGHBOUR_STATE: return {s->A.nbCtl, (size_t)NB_CTL_WORDS * 4};
    case P4A_BUF_DRIFT: return {s->A.drift, (size_t)DRIFT_SLOTS * 4};
    case P4A_BUF_FIELD: return {s->field, s->field ? s->fieldVerts * 4 : 0};
    case P4A_BUF_FIELD_NORMAL: return {s->nfield, s->nfield ? s->fieldVerts * 4 : 0};
    case P4A_BUF_BODY: return {s->exportPhase ? s->phaseOut : nullptr, s->exportPhase ? n * 16 : 0};
    case P4A_BUF_BODY_PACKED:
      return {s->bodyPacked, s->bodyPacked ? (size_t)s->nBodies * 96 : (size_t)0};
    case P4A_BUF_MESH_VERTS: return {s->meshVerts, s->meshVerts ? (size_t)s->meshMaxTris * 24 : 0};
    case P4A_BUF_MESH_INDIRECT: return {s->meshCounter ? s->meshCounter + 4 : nullptr,
                                        s->meshCounter ? (size_t)16 : (size_t)0};
    default: return {nullptr, 0};
  }
}

int check(p4a_sim* s) {
  if (!s) return P4A_ERR_ARG;
  if (!s->ready) return fail(s, P4A_ERR_STATE, "reset() has not been called");
  return P4A_OK;
}

// Rewrites the per-slot phase after anything that moves slots, once a caller
// has asked for its device buffer.
int refresh_exports(p4a_sim* s) {
  if (!s->exportPhase) return P4A_OK;
  LAUNCH(s, "export_phase", launch_export_phase((uint32_t)s->n, s->A.pid[s->parity],
                                                s->A.bodyOfPid, s->phaseOut, s->stream));
  return P4A_OK;
}

}  // namespace

extern "C" {

P4A_API const char* p4a_version(void) { return "p4a 0.1 (gfx950)"; }

P4A_API int p4a_copy_probe(int device, size_t bytes, int32_t reps, double* ms) {
  if (!ms || reps < 1 || bytes == 0 || bytes % 16) return P4A_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return P4A_ERR_NODEV;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return P4A_ERR_NOMEM;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return P4A_ERR_NOMEM;
  }
  hipStream_t st;
  hipEvent_t e0, e1;
  int rc = P4A_OK;
  if (hipStreamCreate(&st) != hipSuccess) rc = P4A_ERR_HIP;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = P4A_ERR_HIP;
  if (!rc) {
    const size_t n = bytes / 16;
    hipMemsetAsync(a, 0, bytes, st);
    hipError_t le = launch_copy4((const float4*)a, (float4*)b, n, st);  // warm
    hipEventRecord(e0, st);
    for (int r = 0; r < reps && le == hipSuccess; r++)
      le = launch_copy4((const float4*)a, (float4*)b, n, st);
    hipEventRecord(e1, st);
    if (le != hipSuccess) rc = P4A_ERR_ARG;
    if (hipStreamSynchronize(st) != hipSuccess) rc = P4A_ERR_HIP;
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
  }
  hipFree(a);
  hipFree(b);
  return rc;
}

P4A_API int p4a_create(int device, p4a_sim** out) {
  if (!out) return P4A_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return P4A_ERR_NODEV;
  p4a_sim* s = new (std::nothrow) p4a_sim();
  if (!s) return P4A_ERR_NOMEM;
  s->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess) {
    delete s;
    return P4A_ERR_HIP;
  }
  if (hipHostMalloc((void**)&s->statsHost, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    hipStreamDestroy(s->stream);
    delete s;
    return P4A_ERR_NOMEM;
  }
  for (int k = 0; k < 4; k++) s->statsHost[k] = 0;
  s->statsHost[0] = 1;
  s->statsHost[1] = 1;
  p4a_params_default(&s->params);
  if (const char* e = std::getenv("P4A_TILE1")) s->tile1.variant = std::atoi(e) % tile_variants(1);
  if (const char* e = std::getenv("P4A_TILE2")) s->tile2.variant = std::atoi(e) % tile_variants(2);
  *out = s;
  return P4A_OK;
}

P4A_API int p4a_destroy(p4a_sim* s) {
  if (!s) return P4A_ERR_ARG;
  hipSetDevice(s->device);
  hipStreamSynchronize(s->stream);
  if (s->side) hipStreamSynchronize(s->side);
  drop_graphs(s);
  free_list(s->devAllocs);
  free_list(s->cellAllocs);
  free_list(s->bndAllocs);
  free_list(s->anisoAllocs);
  free_list(s->fieldAllocs);
  free_list(s->fieldRecAllocs);
  free_list(s->meshVertAllocs);
  free_list(s->exportAllocs);
  for (hipEvent_t e : s->eventPool) hipEventDestroy(e);
  if (s->statsHost) hipHostFree(s->statsHost);
  if (s->side) {
    hipEventDestroy(s->nbFork);
    hipEventDestroy(s->nbJoin);
    hipStreamDestroy(s->side);
  }
  hipStreamDestroy(s->stream);
  delete s;
  return P4A_OK;
}

P4A_API const char* p4a_last_error(const p4a_sim* s) {
  return s ? s->err.c_str() : "null handle";
}

P4A_API int p4a_reset(p4a_sim* s, const p4a_params* p, const p4a_scene* sc) {
  if (!s || !p || !sc) return P4A_ERR_ARG;
  if (sc->n < 0 || (sc->n > 0 && (!sc->pos || !sc->vel || !sc->rest || !sc->phase)))
    return fail(s, P4A_ERR_ARG, "scene arrays missing");
  if (sc->n_boundary < 0 || (sc->n_boundary > 0 && (!sc->boundary_pts || !sc->boundary_psi)))
    return fail(s, P4A_ERR_ARG, "boundary arrays missing");
  if (sc->n_bodies < 0 || (sc->n_bodies > 0 && (!sc->body_centre || !sc->body_density)))
    return fail(s, P4A_ERR_ARG, "body arrays missing");
  if (!(sc->h > 0) || !(p->spacing > 0)) return fail(s, P4A_ERR_ARG, "h and spacing must be > 0");
  if (const char* fe = form_error(p)) return fail(s, P4A_ERR_ARG, fe);
  {
    // refused before the old state is released: a failed reset leaves it usable
    int dim[3];
    if (grid_cells(p->box, sc->h, dim) < 0) return fail(s, P4A_ERR_RANGE, "grid too large");
    const long pour = p->pour_budget < 0 ? (long)sc->n_fluid : (long)p->pour_budget;
    if ((long)sc->n + pour > (long)INT32_MAX / 16) return fail(s, P4A_ERR_RANGE, "capacity too large");
    for (int i = 0; i < sc->n; i++)
      if (sc->phase[(size_t)i * 4] > (uint32_t)sc->n_bodies)
        return fail(s, P4A_ERR_ARG, "phase names a body that does not exist");
  }
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  drop_graphs(s);
  free_list(s->devAllocs);
  free_list(s->cellAllocs);
  free_list(s->bndAllocs);
  free_list(s->anisoAllocs);
  free_list(s->fieldRecAllocs);
  free_list(s->exportAllocs);
  s->exportPhase = false;
  s->phaseOut = nullptr;
  s->bodyStat = s->bodyPacked = nullptr;
  s->bodyStatSet = false;
  s->smoothPos = nullptr;
  s->anisoG = nullptr;
  s->anisoFallbacks = nullptr;
  s->anisoCap = 0;
  s->fieldRec = s->fieldRecA = s->fieldRecB = nullptr;
  s->fieldRecCap = 0;
  s->ready = false;
  s->params = *p;
  s->h = sc->h;
  s->d = sc->d;
  s->mass = sc->mass;
  s->denomRest = sc->denom_rest;
  s->simTime = 0;
  s->gen += 1;
  int rc = set_grid(s);
  if (rc) return rc;

  const int n = sc->n;
  s->n = n;
  s->nFluid = sc->n_fluid;
  s->pourBudget = p->pour_budget < 0 ? sc->n_fluid : p->pour_budget;
  const long capL = (long)n + s->pourBudget;
  if (capL > (long)INT32_MAX / 16) return fail(s, P4A_ERR_RANGE, "capacity too large");
  s->cap = (int)capL;
  const size_t cap = (size_t)std::max(1, s->cap);

  Arrays& A = s->A;
  std::memset(&A, 0, sizeof(A));
#define ALLOC(ptr, count) if ((rc = dalloc(s, s->devAllocs, &(ptr), (count)))) return rc
  for (int k = 0; k < 2; k++) {
    ALLOC(A.pos[k], cap);
    ALLOC(A.pred[k], cap);
    ALLOC(A.vel[k], cap);
    ALLOC(A.pid[k], cap);
  }
  ALLOC(A.predX, cap);
  ALLOC(A.drift, DRIFT_SLOTS);
  ALLOC(A.ofs[0], cap);
  ALLOC(A.ofs[1], cap);
  ALLOC(A.cellOfSlot, cap);
  ALLOC(A.key, cap);
  ALLOC(A.tmp, cap);
  ALLOC(A.fatList, fat_list_len(cap));
  ALLOC(A.nbMask, cap * 9);
  A.nbStride = (uint32_t)cap;
  ALLOC(A.nbRef, cap);
  ALLOC(A.nbCtl, NB_CTL_WORDS);
  for (int k = 0; k < 2; k++) {
    ALLOC(A.nbQueue[k], (size_t)NB_QSEG * nb_seg_cap((uint32_t)cap));
    ALLOC(A.nbMoved[k], (size_t)NB_QSEG * nb_seg_cap((uint32_t)cap));
  }
  ALLOC(A.order, cap);
  ALLOC(A.slot, cap);
  ALLOC(A.slotOf, cap);
  ALLOC(A.density, cap);
  ALLOC(A.lambda, cap);
  ALLOC(A.corr, cap);
  ALLOC(A.normal, cap);
  ALLOC(A.velq, cap);
  ALLOC(A.hitMask, cap * 9);
  A.hitStride = (uint32_t)cap;
  ALLOC(A.bodyOfPid, cap);
  ALLOC(A.restOfPid, cap);
  for (int k = 0; k < 3; k++) ALLOC(A.census[k], cap);
  ALLOC(A.statsPart, (size_t)stats_blocks((uint32_t)cap) * 4 + 4);
  ALLOC(A.statsOut, 4);

  // Bodies: particle ids grouped by body, in ascending id order.
  const int nB = sc->n_bodies;
  s->nBodies = nB;
  s->bodyOfPidH.assign(cap, 0u);
  s->restOfPidH.assign(cap * 4, 0.0f);
  std::vector<std::vector<uint32_t>> members(nB);
  int nBodyParts = 0;
  for (int i = 0; i < n; i++) {
    const uint32_t b = sc->phase[(size_t)i * 4];
    if (b == 0) continue;
    if (b > (uint32_t)nB) return fail(s, P4A_ERR_ARG, "phase names a body that does not exist");
    members[b - 1].push_back((uint32_t)i);
    s->bodyOfPidH[i] = b;
    for (int k = 0; k < 3; k++) s->restOfPidH[(size_t)i * 4 + k] = sc->rest[(size_t)i * 4 + k];
    nBodyParts++;
  }
  s->nBodyParts = nBodyParts;
  {
    uint32_t lo = 0xffffffffu, hi = 0;
    for (int b = 0; b < nB; b++)
      for (uint32_t q : members[b]) {
        lo = std::min(lo, q);
        hi = std::max(hi, q + 1);
      }
    s->slotIdLo = 0;
    s->slotIdCount = 0xffffffffu;
    if (nBodyParts == 0) s->slotIdCount = 0;
    else if (hi - lo == (uint32_t)nBodyParts) {
      s->slotIdLo = lo;
      s->slotIdCount = hi - lo;
    }
  }
  std::vector<uint32_t> pids, first(nB + 1, 0);
  s->bodyCountH.assign(nB, 0);
  for (int b = 0; b < nB; b++) {
    first[b] = (uint32_t)pids.size();
    pids.insert(pids.end(), members[b].begin(), members[b].end());
    s->bodyCountH[b] = (int32_t)members[b].size();
  }
  first[nB] = (uint32_t)pids.size();
  ALLOC(A.bodyPids, std::max<size_t>(1, pids.size()));
  ALLOC(A.bodyFirst, (size_t)nB + 1);
  ALLOC(A.bodyCentreD, (size_t)std::max(1, nB) * 4);
  ALLOC(A.bodyRotD, (size_t)std::max(1, nB) * 9);
  ALLOC(A.bodyInvRatio, (size_t)std::max(1, nB));
  ALLOC(A.bodyCentre, (size_t)std::max(1, nB));
  ALLOC(A.bodyRot, (size_t)std::max(1, nB) * 3);
#undef ALLOC
  // Fixed-point scales of the reference (sim.js:163-171), reported only:
  // the body sums here are FP64.
  {
    double maxCount = 1, maxRest = 1e-3;
    for (int b = 0; b < nB; b++) maxCount = std::max(maxCount, (double)s->bodyCountH[b]);
    for (int i = 0; i < n; i++) {
      const double r = std::hypot(std::hypot((double)sc->rest[i * 4], (double)sc->rest[i * 4 + 1]),
                                  (double)sc->rest[i * 4 + 2]);
      if (r > maxRest) maxRest = r;
    }
    const double extent = std::max(p->box[0], std::max(p->box[1], p->box[2]));
    s->bodyScale = 1e9 / (extent * maxCount);
    s->covScale = 1e9 / (extent * maxRest * maxCount);
  }
  if (nB > 0) {
    std::vector<double> cD((size_t)nB * 4, 0.0), rD((size_t)nB * 9, 0.0);
    std::vector<float> cF((size_t)nB * 4, 0.0f), rF((size_t)nB * 12, 0.0f), inv(nB);
    for (int b = 0; b < nB; b++) {
      for (int k = 0; k < 3; k++) {
        cD[b * 4 + k] = sc->body_centre[b * 3 + k];
        cF[b * 4 + k] = (float)sc->body_centre[b * 3 + k];
        rD[b * 9 + k * 4] = 1.0;
        rF[b * 12 + k * 4 + k] = 1.0f;
      }
      inv[b] = (float)(1.0 / std::max(1e-6, sc->body_density[b]));
    }
    HIP_TRY(s, hipMemcpyAsync(A.bodyPids, pids.data(), pids.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(A.bodyFirst, first.data(), first.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(A.bodyCentreD, cD.data(), cD.size() * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(A.bodyRotD, rD.data(), rD.size() * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(A.bodyCentre, cF.data(), cF.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(A.bodyRot, rF.data(), rF.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(A.bodyInvRatio, inv.data(), inv.size() * 4, hipMemcpyHostToDevice, s->stream));
  }
  HIP_TRY(s, hipMemcpyAsync(A.bodyOfPid, s->bodyOfPidH.data(), (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(A.restOfPid, s->restOfPidH.data(), (size_t)n * 16, hipMemcpyHostToDevice, s->stream));

  // Particle state (sim.js:191-195).
  std::vector<uint32_t> iota(cap);
  for (size_t i = 0; i < cap; i++) iota[i] = (uint32_t)i;
  HIP_TRY(s, hipMemcpyAsync(A.pos[0], sc->pos, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(A.vel[0], sc->vel, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(A.pred[0], sc->pos, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(A.pid[0], iota.data(), cap * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(s, hipMemcpyAsync(A.slotOf, iota.data(), cap * 4, hipMemcpyHostToDevice, s->stream));

  if ((rc = alloc_cells(s))) return rc;
  if ((rc = upload_boundary(s, sc->boundary_pts, sc->boundary_psi, sc->n_boundary))) return rc;

  s->parity = 0;
  s->predCur = A.pred[0];
  s->heldBody = -1;
  s->heldTarget[0] = s->heldTarget[1] = s->heldTarget[2] = 0;
  s->heldRate = 10;
  s->heldLimit = 4;
  s->statsValid = false;
  s->ready = true;

  // primeGrid (sim.js:223-253): predict with dt = 0 and one grid build.
  const Uniform U0 = make_uniform(s, 0.0);
  const bool timing = s->timing;
  s->timing = false;
  rc = launch(s, "predict", [&]() { return launch_predict_count(U0, s->A, s->parity, s->stream); });
  if (!rc) rc = run_grid(s, U0, true);
  s->timing = timing;
  if (rc) return rc;
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return P4A_OK;
}

P4A_API int p4a_set_params(p4a_sim* s, const p4a_params* p) {
  int rc = check(s);
  if (rc) return rc;
  if (!p) return P4A_ERR_ARG;
  if (const char* fe = form_error(p)) return fail(s, P4A_ERR_ARG, fe);
  p4a_params& q = s->params;
  q.substeps = p->substeps;
  q.iterations = p->iterations;
  q.cfm_epsilon_rel = p->cfm_epsilon_rel;
  q.scorr_k = p->scorr_k;
  q.scorr_dq = p->scorr_dq;
  q.xsph_c = p->xsph_c;
  q.omega = p->omega;
  q.sor_average = p->sor_average;
  q.surface_tension_k = p->surface_tension_k;
  q.gravity = p->gravity;
  q.no_boundary = p->no_boundary;
  q.math = p->math;
  q.precision = p->precision;
  q.use_graph = p->use_graph;
  q.neighbour_skin = p->neighbour_skin;
  q.neighbour_min_particles = p->neighbour_min_particles;
  s->nbBuilt = false;
  return P4A_OK;
}

P4A_API int p4a_substeps(p4a_sim* s, double dt, int32_t count) {
  int rc = check(s);
  if (rc) return rc;
  if (count < 0) return P4A_ERR_ARG;
  HIP_TRY(s, hipSetDevice(s->device));
  if ((rc = ensure_side(s))) return rc;
  if (s->timing) {
    s->timed.clear();
    s->eventNext = 0;
  }
  for (int k = 0; k < count; k++)
    if ((rc = run_substep(s, dt))) return rc;
  if (count > 0 && (rc = run_stats(s))) return rc;
  return count > 0 ? refresh_exports(s) : P4A_OK;
}

P4A_API int p4a_step(p4a_sim* s, double frame_dt, int32_t* substeps_run) {
  int rc = check(s);
  if (rc) return rc;
  // time bank (sim.js:415-429)
  const int base = std::max(1, s->params.substeps);
  const double dtTarget = (1.0 / 60.0) / base;
  // `(timeBank || 0) + frameDt`: a bank left at NaN by a NaN frame restarts
  // from zero on the next frame
  if (s->timeBank == 0 || std::isnan(s->timeBank)) s->timeBank = 0;
  s->timeBank += frame_dt;
  // the count stays a double until it is known to fit an int
  const double q = std::floor(s->timeBank / dtTarget + 1e-4);
  const int workCap = base * 8;
  int sub = 0;
  if (q > workCap) {
    sub = workCap;
    s->timeBank = 0;
  } else {
    s->timeBank -= q * dtTarget;
    if (q >= 1) sub = (int)q;
  }
  s->lastSubsteps = sub;
  s->lastAdvanced = sub * dtTarget;
  if (substeps_run) *substeps_run = std::max(sub, 0);
  if (sub < 1) {
    s->lastAdvanced = 0;
    return P4A_OK;
  }
  return p4a_substeps(s, dtTarget, sub);
}

P4A_API int p4a_run_stage(p4a_sim* s, int32_t stage, double dt) {
  int rc = check(s);
  if (rc) return rc;
  HIP_TRY(s, hipSetDevice(s->device));
  if ((rc = ensure_side(s))) return rc;
  const Uniform U = make_uniform(s, dt);
  const bool ex = exact_math(s);
  switch (stage) {
    case P4A_STAGE_PREDICT:
      if (s->heldBody >= 1 && s->nBodyParts > 0) {
        DragArgs D;
        for (int k = 0; k < 3; k++) D.goal[k] = (float)s->heldTarget[k];
        D.rate = (float)s->heldRate;
        D.held = (uint32_t)s->heldBody;
        D.blend = (float)(1.0 - std::exp(-std::max(dt, 0.0) / 0.02));
        D.limit = (float)s->heldLimit;
        LAUNCH(s, "body_drag", launch_body_drag(U, s->A, s->A.vel[s->parity], D, s->stream));
      }
      LAUNCH(s, "predict", launch_predict_count(U, s->A, s->parity, s->stream));
      s->nbBuilt = false;
      break;
    case P4A_STAGE_GRID:
      if ((rc = run_grid(s, U, true))) return rc;
      break;
    case P4A_STAGE_BODY_PRE:
      LAUNCH(s, "body_pre", launch_body_pre(U, s->A, s->predCur, s->stream));
      break;
    case P4A_STAGE_LAMBDA:
      if ((rc = launch_lambda_any(s, U))) return rc;
      break;
    case P4A_STAGE_DELTA: {
      float4* out = other_pred(s, s->parity);
      if ((rc = launch_delta_any(s, U, out))) return rc;
      s->predCur = out;
      break;
    }
    case P4A_STAGE_BODY_SOLVE:
      if ((rc = launch_body_solve_any(s, U))) return rc;
      break;
    case P4A_STAGE_FINISH:
      if ((rc = run_finish(s, U))) return rc;
      s->nbBuilt = false;
      s->simTime += dt;
      break;
    case P4A_STAGE_STATS:
      if ((rc = run_stats(s))) return rc;
      break;
    default:
      return fail(s, P4A_ERR_ARG, "unknown stage");
  }
  if ((rc = refresh_exports(s))) return rc;
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return P4A_OK;
}

P4A_API int p4a_resize_box(p4a_sim* s, const double box[3]) {
  int rc = check(s);
  if (rc) return rc;
  if (!box) return P4A_ERR_ARG;
  const double oldBox[3] = {s->params.box[0], s->params.box[1], s->params.box[2]};
  bool moved = false;
  for (int a = 0; a < 3; a++)
    if (std::fabs(box[a] - oldBox[a]) > 1e-9) moved = true;
  if (!moved) return P4A_OK;
  for (int a = 0; a < 3; a++)
    if (!(box[a] > 0)) return fail(s, P4A_ERR_ARG, "box sizes must be > 0");
  {
    // refused before anything changes: the sim keeps its old box and grid
    int dim[3];
    if (grid_cells(box, s->h, dim) < 0) return fail(s, P4A_ERR_RANGE, "grid too large");
  }
  s->ofsValid = false;
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  drop_graphs(s);
  for (int a = 0; a < 3; a++) s->params.box[a] = box[a];
  const int oldCells = s->nCells;
  if ((rc = set_grid(s))) return rc;
  if (s->nCells != oldCells && (rc = alloc_cells(s))) return rc;
  int32_t nb = 0;
  if ((rc = p4a_boundary_build(box, s->params.spacing, s->h, s->params.rest_density, nullptr,
                               nullptr, 0, &nb)))
    return fail(s, rc, "boundary build failed");
  std::vector<float> pts((size_t)std::max(1, nb) * 3), psi((size_t)std::max(1, nb));
  if ((rc = p4a_boundary_build(box, s->params.spacing, s->h, s->params.rest_density, pts.data(),
                               psi.data(), nb, &nb)))
    return fail(s, rc, "boundary build failed");
  if ((rc = upload_boundary(s, pts.data(), psi.data(), nb))) return rc;
  s->gen += 1;
  ResizeArgs R;
  const double halfD = 0.5 * s->params.spacing;
  for (int a = 0; a < 3; a++) {
    R.oldMin[a] = 0.0f;
    R.oldSize[a] = (float)oldBox[a];
    R.newMin[a] = 0.0f;
    R.newSize[a] = (float)box[a];
    R.clampMin[a] = (float)halfD;
    R.clampMax[a] = (float)(box[a] - halfD);
  }
  R.n = (uint32_t)s->n;
  LAUNCH(s, "resize", launch_resize(R, s->A.pos[s->parity], s->stream));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return P4A_OK;
}

P4A_API int p4a_append_fluid(p4a_sim* s, const float* pos3, const float* vel3, int32_t count,
                             int32_t* appended) {
  int rc = check(s);
  if (rc) return rc;
  if (appended) *appended = 0;
  const int room = std::max(0, s->cap - s->n);
  const int k = std::min(count, room);
  if (count > 0 && (!pos3 || !vel3)) return P4A_ERR_ARG;
  if (k > 0 && (!all_finite(pos3, (size_t)k * 3) || !all_finite(vel3, (size_t)k * 3)))
    return fail(s, P4A_ERR_ARG, "appended positions and velocities must be finite");
  s->ofsValid = false;
  if (k <= 0) return P4A_OK;
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  drop_graphs(s);
  const size_t off = (size_t)s->n;
  std::vector<float> p4((size_t)k * 4), v4((size_t)k * 4, 0.0f), zero4((size_t)k * 4, 0.0f),
      zero1(k, 0.0f), rho(k, (float)s->params.rest_density);
  std::vector<uint32_t> ids(k);
  for (int i = 0; i < k; i++) {
    p4[i * 4 + 0] = pos3[i * 3 + 0];
    p4[i * 4 + 1] = pos3[i * 3 + 1];
    p4[i * 4 + 2] = pos3[i * 3 + 2];
    p4[i * 4 + 3] = 1.0f;
    v4[i * 4 + 0] = vel3[i * 3 + 0];
    v4[i * 4 + 1] = vel3[i * 3 + 1];
    v4[i * 4 + 2] = vel3[i * 3 + 2];
    ids[i] = (uint32_t)(off + i);
  }
  const int par = s->parity;
  HIP_TRY(s, hipMemcpy(s->A.pos[par] + off, p4.data(), p4.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.vel[par] + off, v4.data(), v4.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->predCur + off, p4.data(), p4.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.density + off, rho.data(), rho.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.lambda + off, zero1.data(), zero1.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.corr + off, zero4.data(), zero4.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.normal + off, zero4.data(), zero4.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.pid[par] + off, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.slotOf + off, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
  s->n += k;
  if (appended) *appended = k;
  return refresh_exports(s);
}

P4A_API int p4a_apply_ray_impulse(p4a_sim* s, const double origin[3], const double dir[3],
                                  const double impulse[3], double radius, double speed_limit) {
  int rc = check(s);
  if (rc) return rc;
  if (!origin || !dir || !impulse) return P4A_ERR_ARG;
  if (!all_finite(origin, 3) || !all_finite(dir, 3) || !all_finite(impulse, 3) ||
      !std::isfinite(radius) || !std::isfinite(speed_limit))
    return fail(s, P4A_ERR_ARG, "ray impulse arguments must be finite");
  if (!s->n || radius <= 0) return P4A_OK;
  // Math.hypot, as the reference (no overflow or underflow of the squares)
  const double len = std::hypot(impulse[0], impulse[1], impulse[2]);
  if (len <= 0) return P4A_OK;
  double dl = std::hypot(dir[0], dir[1], dir[2]);
  if (!(dl != 0)) dl = 1;  // `|| 1`
  RayArgs R;
  for (int k = 0; k < 3; k++) {
    R.origin[k] = (float)origin[k];
    R.dir[k] = (float)(dir[k] / dl);
    R.impulse[k] = (float)impulse[k];
  }
  R.radius = (float)radius;
  R.speedLimit = (float)speed_limit;
  R.n = (uint32_t)s->n;
  HIP_TRY(s, hipSetDevice(s->device));
  LAUNCH(s, "impulse", launch_impulse(s->A.pos[s->parity], s->A.vel[s->parity], R, s->stream));
  return P4A_OK;
}

P4A_API int p4a_hold_body(p4a_sim* s, int32_t body, const double target[3], double rate,
                          double limit) {
  int rc = check(s);
  if (rc) return rc;
  if (!target) return P4A_ERR_ARG;
  if (!all_finite(target, 3) || !std::isfinite(rate) || !std::isfinite(limit))
    return fail(s, P4A_ERR_ARG, "hold target, rate and limit must be finite");
  s->heldBody = (body >= 1 && body <= s->nBodies) ? body : -1;
  for (int k = 0; k < 3; k++) s->heldTarget[k] = target[k];
  s->heldRate = rate;
  s->heldLimit = limit;
  return P4A_OK;
}

P4A_API int p4a_release_body(p4a_sim* s) {
  int rc = check(s);
  if (rc) return rc;
  s->heldBody = -1;
  return P4A_OK;
}

P4A_API int p4a_read(p4a_sim* s, int32_t id, void* dst, size_t offset, size_t bytes) {
  int rc = check(s);
  if (rc) return rc;
  if (!dst && bytes) return P4A_ERR_ARG;
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  if (id == P4A_BUF_BODY || id == P4A_BUF_REST) {
    // reference layout (uint4 / float4 per live slot) from the id maps
    const size_t total = (size_t)s->n * 16;
    if (offset > total || bytes > total - offset) return fail(s, P4A_ERR_RANGE, "read past end");
    std::vector<uint32_t> pid(s->n);
    HIP_TRY(s, hipMemcpy(pid.data(), s->A.pid[s->parity], (size_t)s->n * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> out((size_t)s->n * 4, 0u);
    for (int i = 0; i < s->n; i++) {
      const uint32_t q = pid[i];
      if (id == P4A_BUF_BODY) out[(size_t)i * 4] = q < s->bodyOfPidH.size() ? s->bodyOfPidH[q] : 0u;
      else if ((size_t)q * 4 + 3 < s->restOfPidH.size())
        std::memcpy(&out[(size_t)i * 4], &s->restOfPidH[(size_t)q * 4], 16);
    }
    std::memcpy(dst, (const char*)out.data() + offset, bytes);
    return P4A_OK;
  }
  BufView v = buf_view(s, id);
  if (!v.ptr) return fail(s, P4A_ERR_ARG, "unknown buffer id");
  if (offset > v.bytes || bytes > v.bytes - offset) return fail(s, P4A_ERR_RANGE, "read past end");
  if (bytes) HIP_TRY(s, hipMemcpy(dst, (const char*)v.ptr + offset, bytes, hipMemcpyDeviceToHost));
  return P4A_OK;
}

P4A_API int p4a_write(p4a_sim* s, int32_t id, const void* src, size_t offset, size_t bytes) {
  int rc = check(s);
  if (rc) return rc;
  if (!src && bytes) return P4A_ERR_ARG;
  // index buffers: the walks and the sort index particle arrays through
  // them without bounds checks, so only the library writes them
  if (id == P4A_BUF_BODY || id == P4A_BUF_REST || id == P4A_BUF_PID || id == P4A_BUF_ORDER ||
      id == P4A_BUF_KEY || id == P4A_BUF_SLOT || id == P4A_BUF_CELL_START ||
      id == P4A_BUF_BCELL_START || id == P4A_BUF_CELL_OF_SLOT || id == P4A_BUF_NEIGHBOUR_STATE)
    return fail(s, P4A_ERR_ARG, "buffer is not writable");
  // particle positions and velocities stay finite (a non-finite position
  // would be binned into a clamped corner cell)
  if (id == P4A_BUF_POS || id == P4A_BUF_VEL || id == P4A_BUF_PRED) {
    if (offset % 4 != 0 || bytes % 4 != 0)
      return fail(s, P4A_ERR_ARG, "float buffers are written in whole floats");
    if (!all_finite((const float*)src, bytes / 4))
      return fail(s, P4A_ERR_ARG, "positions and velocities must be finite");
  }
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  BufView v = buf_view(s, id);
  if (!v.ptr) return fail(s, P4A_ERR_ARG, "unknown buffer id");
  if (offset > v.bytes || bytes > v.bytes - offset) return fail(s, P4A_ERR_RANGE, "write past end");
  s->ofsValid = false;
  if (bytes) HIP_TRY(s, hipMemcpy((char*)v.ptr + offset, src, bytes, hipMemcpyHostToDevice));
  drop_graphs(s);
  return P4A_OK;
}

P4A_API int p4a_device_ptr(p4a_sim* s, int32_t id, void** ptr, size_t* bytes) {
  int rc = check(s);
  if (rc) return rc;
  if (!ptr) return P4A_ERR_ARG;
  if (id == P4A_BUF_BODY && !s->exportPhase) {
    // a renderer binds the per-slot phase: keep it on the device from now on
    HIP_TRY(s, hipSetDevice(s->device));
    if (!s->phaseOut &&
        (rc = dalloc(s, s->exportAllocs, &s->phaseOut, (size_t)std::max(1, s->cap))))
      return rc;
    s->exportPhase = true;
    if ((rc = refresh_exports(s))) return rc;
  }
  BufView v = buf_view(s, id);
  if (!v.ptr) return fail(s, P4A_ERR_ARG, "no device buffer with that id");
  *ptr = v.ptr;
  if (bytes) *bytes = v.bytes;
  return P4A_OK;
}

P4A_API int p4a_get_info(p4a_sim* s, p4a_info* o) {
  if (!s || !o) return P4A_ERR_ARG;
  std::memset(o, 0, sizeof(*o));
  o->n = s->n; o->cap = s->cap; o->n_fluid = s->nFluid; o->n_body_parts = s->nBodyParts;
  o->n_bodies = s->nBodies; o->n_boundary = s->nBoundary; o->n_cells = s->nCells;
  for (int k = 0; k < 3; k++) o->grid_dim[k] = s->gridDim[k];
  o->parity = s->parity; o->gen = s->gen; o->last_substeps = s->lastSubsteps;
  o->pour_budget = s->pourBudget;
  o->h = s->h; o->mass = s->mass; o->denom_rest = s->denomRest; o->sim_time = s->simTime;
  o->last_advanced = s->lastAdvanced; o->time_bank = s->timeBank;
  o->body_scale = s->bodyScale; o->cov_scale = s->covScale;
  return P4A_OK;
}

P4A_API int p4a_stats(p4a_sim* s, double out[4]) {
  int rc = check(s);
  if (rc) return rc;
  if (!out) return P4A_ERR_ARG;
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  for (int k = 0; k < 4; k++) out[k] = s->statsHost[k];
  return P4A_OK;
}

// ---- the stats reduction without allocation --------------------------------
// The per-block partials of p4a_stats_noalloc live in this fixed array, one
// copy per GPU for the life of the process.  Host code never names it
// directly: its device address comes from hipGetSymbolAddress, once per
// device, with that device current.
constexpr uint32_t kNoAllocMaxParticles = P4A_STATS_NOALLOC_MAX_PARTICLES;
constexpr size_t kNoAllocPartials = (size_t)((kNoAllocMaxParticles + 2047u) / 2048u) * 4u;
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device `device`.
int scratch_on(p4a_sim* s, int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(s, P4A_ERR_RANGE, "device ordinal too large");
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(s, P4A_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[device] = (double*)ptr;
  }
  *out = g_scratch[device];
  return P4A_OK;
}
}  // namespace

P4A_API int p4a_stats_noalloc(p4a_sim* s, double* out_dev) {
  int rc = check(s);
  if (rc) return rc;
  if (!out_dev) return fail(s, P4A_ERR_ARG, "out_dev is null");
  if ((uint32_t)s->n > kNoAllocMaxParticles)
    return fail(s, P4A_ERR_RANGE, "p4a_stats_noalloc takes at most "
                                  "P4A_STATS_NOALLOC_MAX_PARTICLES particles");
  HIP_TRY(s, hipSetDevice(s->device));
  double* part = nullptr;
  if ((rc = scratch_on(s, s->device, &part))) return rc;
  const Uniform U = make_uniform(s, 0.0);
  const hipError_t e = launch_stats_into(U, s->A.density, s->A.vel[s->parity], part, out_dev,
                                         s->stream);
  if (e != hipSuccess) return fail(s, P4A_ERR_HIP, std::string("stats: ") + hipGetErrorString(e));
  return P4A_OK;
}

P4A_API int p4a_body_pose(p4a_sim* s, float* centre, float* rot) {
  int rc = check(s);
  if (rc) return rc;
  const int nB = s->nBodies;
  if (nB == 0) return P4A_OK;
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  std::vector<float> c((size_t)nB * 4), r((size_t)nB * 12);
  HIP_TRY(s, hipMemcpy(c.data(), s->A.bodyCentre, c.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(s, hipMemcpy(r.data(), s->A.bodyRot, r.size() * 4, hipMemcpyDeviceToHost));
  for (int b = 0; b < nB; b++) {
    if (centre)
      for (int k = 0; k < 3; k++) centre[b * 3 + k] = c[b * 4 + k];
    if (rot)
      for (int row = 0; row < 3; row++)
        for (int col = 0; col < 3; col++) rot[b * 9 + col * 3 + row] = r[(b * 3 + row) * 4 + col];
  }
  return P4A_OK;
}

P4A_API int p4a_aniso_params_default(p4a_aniso_params* a) {
  if (!a) return P4A_ERR_ARG;
  std::memset(a, 0, sizeof(*a));
  a->radius_scale = 2.0;
  a->lambda = 0.9;
  a->kr = 4.0;
  a->kn = 1.0;
  a->ks = 1.0;
  a->stretch = 2.0;
  a->min_extent = 0.0;
  a->neps = 25;
  a->fp64_fallback = 1;
  a->gap_tol = 1e-4;
  return P4A_OK;
}

P4A_API int p4a_aniso(p4a_sim* s, const p4a_aniso_params* a, int32_t* fp64_solves) {
  int rc = check(s);
  if (rc) return rc;
  if (!a) return P4A_ERR_ARG;
  if (!(a->radius_scale > 0) || a->radius_scale > 2.0)
    return fail(s, P4A_ERR_ARG, "radius_scale must be in (0, 2]: the pass walks +-2 cells");
  HIP_TRY(s, hipSetDevice(s->device));
  // allocated on first use even at capacity 0: the fallback counter is
  // cleared and read back whatever the particle count
  if (!s->anisoFallbacks || s->anisoCap < s->cap) {
    free_list(s->anisoAllocs);
    if ((rc = dalloc(s, s->anisoAllocs, &s->smoothPos, (size_t)s->cap))) return rc;
    if ((rc = dalloc(s, s->anisoAllocs, &s->anisoG, (size_t)s->cap * 2))) return rc;
    if ((rc = dalloc(s, s->anisoAllocs, &s->anisoFallbacks, 1))) return rc;
    s->anisoCap = s->cap;
  }
  const Uniform U = make_uniform(s, 0.0);
  AnisoArgs A;
  const double support = s->h, radius = a->radius_scale * support;
  A.lambda = (float)a->lambda;
  A.invSupport = (float)(1.0 / support);
  A.support = (float)support;
  A.radius = (float)radius;
  A.kr = (float)a->kr;
  A.kn = (float)a->kn;
  A.ks = (float)(a->ks / (0.15 * radius * radius));
  A.stretch = (float)a->stretch;
  A.minExtent = (float)a->min_extent;
  A.neps = a->neps;
  A.fp64Fallback = a->fp64_fallback;
  A.gapTol = (float)a->gap_tol;
  LAUNCH(s, "aniso_clear", launch_fill_u32(s->anisoFallbacks, 0u, 1, s->stream));
  LAUNCH(s, "aniso", launch_aniso(U, s->A, s->A.pos[s->parity], A, s->smoothPos, s->anisoG,
                                  s->anisoFallbacks, exact_math(s), s->stream));
  if (fp64_solves) {
    uint32_t c = 0;
    HIP_TRY(s, hipMemcpyAsync(&c, s->anisoFallbacks, 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    *fp64_solves = (int32_t)c;
  }
  return P4A_OK;
}

P4A_API int p4a_field_configure(p4a_sim* s, const double box_min[3], const double box_max[3],
                                int32_t resolution, int64_t max_triangles) {
  int rc = check(s);
  if (rc) return rc;
  if (!box_min || !box_max || resolution < 1) return P4A_ERR_ARG;
  if (max_triangles <= 0) max_triangles = 3000000;  // SurfaceMesh.maxTriangles (mesh.js:20)
  if (max_triangles > (int64_t)1 << 30) return fail(s, P4A_ERR_RANGE, "max_triangles > 2^30");
  // SurfaceMesh.configure (mesh.js:96-128), JavaScript double arithmetic
  double size[3], longest = 0.0;
  for (int k = 0; k < 3; k++) {
    size[k] = box_max[k] - box_min[k];
    if (!(size[k] > 0)) return fail(s, P4A_ERR_ARG, "field box must have positive size");
    longest = std::max(longest, size[k]);
  }
  const double voxel = longest / resolution;
  size_t nv = 1;
  int vd[3];
  for (int k = 0; k < 3; k++) {
    const double v = std::ceil(size[k] / voxel) + 2.0 + 1.0;
    if (!(v <= 65536.0)) return fail(s, P4A_ERR_RANGE, "field too large");  // NaN too
    vd[k] = (int)v;
    nv *= (size_t)vd[k];
  }
  if (nv > (size_t)INT32_MAX) return fail(s, P4A_ERR_RANGE, "field too large");
  HIP_TRY(s, hipSetDevice(s->device));
  if (nv > s->fieldVertCap) {
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    free_list(s->fieldAllocs);
    s->fieldConfigured = false;  // until every buffer is back
    s->field = s->field2 = s->nfield = nullptr;
    s->fieldAcc = nullptr;
    s->meshBlockCount = s->meshBlockSum = s->meshBlockStart = s->meshCounter = nullptr;
    s->normalReady = false;
    s->fieldVertCap = 0;
    // voxels <= vertices, so nv / 256 + 1 march workgroups always suffice
    const size_t nb = nv / MESH_WG + 1;
    if ((rc = dalloc(s, s->fieldAllocs, &s->field, nv))) return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->fieldAcc, nv))) return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->field2, nv))) return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->nfield, nv))) return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->meshBlockCount, nb))) return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->meshBlockSum, (size_t)scan_blocks((uint32_t)nb))))
      return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->meshBlockStart, nb + 1))) return rc;
    if ((rc = dalloc(s, s->fieldAllocs, &s->meshCounter, 8))) return rc;
    s->fieldVertCap = nv;
  }
  if ((size_t)max_triangles != s->meshVertCap) {
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    free_list(s->meshVertAllocs);
    s->fieldConfigured = false;
    s->meshVerts = nullptr;
    s->meshVertCap = 0;
    if ((rc = dalloc(s, s->meshVertAllocs, &s->meshVerts, (size_t)max_triangles * 3))) return rc;
    s->meshVertCap = (size_t)max_triangles;
  }
  s->meshMaxTris = (uint32_t)max_triangles;
  for (int k = 0; k < 3; k++) {
    s->fieldVertDim[k] = vd[k];
    s->fieldOrigin[k] = box_min[k] - voxel;
  }
  s->fieldVoxel = voxel;
  s->fieldVerts = nv;
  s->fieldConfigured = true;
  return P4A_OK;
}

P4A_API int p4a_field_generate(p4a_sim* s, int32_t kind, int32_t skip_bodies) {
  int rc = check(s);
  if (rc) return rc;
  if (!s->fieldConfigured) return fail(s, P4A_ERR_ARG, "p4a_field_configure first");
  if (kind != P4A_FIELD_ISOTROPIC && kind != P4A_FIELD_ANISOTROPIC) return P4A_ERR_ARG;
  if (kind == P4A_FIELD_ANISOTROPIC && !s->smoothPos)
    return fail(s, P4A_ERR_ARG, "the anisotropic field needs a p4a_aniso pass first");
  HIP_TRY(s, hipSetDevice(s->device));
  if (!s->fieldRec || s->fieldRecCap < s->cap) {
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    free_list(s->fieldRecAllocs);
    const size_t cap = (size_t)std::max(1, s->cap);
    if ((rc = dalloc(s, s->fieldRecAllocs, &s->fieldRec, cap))) return rc;
    if ((rc = dalloc(s, s->fieldRecAllocs, &s->fieldRecA, cap))) return rc;
    if ((rc = dalloc(s, s->fieldRecAllocs, &s->fieldRecB, cap * 3))) return rc;
    s->fieldRecCap = s->cap;
  }
  const Uniform U = make_uniform(s, 0.0);
  FieldArgs F;
  for (int k = 0; k < 3; k++) {
    F.vertDim[k] = s->fieldVertDim[k];
    F.origin[k] = (float)s->fieldOrigin[k];
  }
  F.voxel = (float)s->fieldVoxel;
  F.invSupport = (float)(1.0 / s->h);
  F.support2 = (float)(s->h * s->h);
  F.poly6 = (float)(315.0 / (64.0 * kPi * std::pow(s->h, 9)));
  F.mass = (float)s->mass;
  F.rho0 = (float)s->params.rest_density;
  F.skipBodies = skip_bodies < 0 ? (s->nBodyParts > 0 ? 1u : 0u) : (skip_bodies ? 1u : 0u);
  F.nVerts = (uint32_t)s->fieldVerts;
  const int par = s->parity;
  if (kind == P4A_FIELD_ISOTROPIC)
    LAUNCH(s, "field", launch_field_iso(U, s->A, F, s->A.pos[par], s->A.pid[par], s->fieldRec,
                                        s->field, s->stream));
  else
    LAUNCH(s, "field", launch_field_aniso(U, s->A, F, s->smoothPos, s->anisoG, s->A.pid[par],
                                          s->fieldRecA, s->fieldRecB, s->fieldAcc, s->field,
                                          s->stream));
  return P4A_OK;
}

P4A_API int p4a_field_get_info(p4a_sim* s, p4a_field_info* out) {
  int rc = check(s);
  if (rc) return rc;
  if (!out) return P4A_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  for (int k = 0; k < 3; k++) {
    out->vert_dim[k] = s->fieldVertDim[k];
    out->origin[k] = s->fieldOrigin[k];
  }
  out->voxel = s->fieldVoxel;
  out->n_verts = (int64_t)s->fieldVerts;
  out->max_triangles = (int64_t)s->meshMaxTris;
  return P4A_OK;
}

namespace {

MeshArgs mesh_args(const p4a_sim* s, float iso) {
  MeshArgs M;
  for (int k = 0; k < 3; k++) M.vertDim[k] = s->fieldVertDim[k];
  M.iso = iso;
  M.divVx = FastDiv((uint32_t)M.vertDim[0]);
  M.divVy = FastDiv((uint32_t)M.vertDim[1]);
  M.divNx = FastDiv((uint32_t)(M.vertDim[0] - 1));
  M.divNy = FastDiv((uint32_t)(M.vertDim[1] - 1));
  M.nVerts = (uint32_t)s->fieldVerts;
  M.nVox = (uint32_t)((size_t)(M.vertDim[0] - 1) * (M.vertDim[1] - 1) * (M.vertDim[2] - 1));
  M.maxTris = s->meshMaxTris;
  M.nBlocks = (M.nVox + MESH_WG - 1) / MESH_WG;
  return M;
}

}  // namespace

P4A_API int p4a_body_pack(p4a_sim* s, const float* stat, int32_t count) {
  int rc = check(s);
  if (rc) return rc;
  if (count != s->nBodies) return fail(s, P4A_ERR_ARG, "count must equal the number of bodies");
  if (count == 0) return P4A_OK;
  HIP_TRY(s, hipSetDevice(s->device));
  if (!s->bodyPacked) {
    if ((rc = dalloc(s, s->exportAllocs, &s->bodyStat, (size_t)count * 2))) return rc;
    if ((rc = dalloc(s, s->exportAllocs, &s->bodyPacked, (size_t)count * 6))) return rc;
  }
  if (stat) {
    HIP_TRY(s, hipMemcpyAsync(s->bodyStat, stat, (size_t)count * 32, hipMemcpyHostToDevice,
                              s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    s->bodyStatSet = true;
  }
  if (!s->bodyStatSet) return fail(s, P4A_ERR_ARG, "no body static data uploaded yet");
  LAUNCH(s, "body_pack", launch_body_pack((uint32_t)count, s->A.bodyCentre, s->A.bodyRot,
                                          s->bodyStat, s->bodyPacked, s->stream));
  return P4A_OK;
}

P4A_API int p4a_field_smooth(p4a_sim* s, int32_t passes) {
  int rc = check(s);
  if (rc) return rc;
  if (!s->fieldConfigured) return fail(s, P4A_ERR_ARG, "p4a_field_configure first");
  if (passes <= 0) return P4A_OK;
  HIP_TRY(s, hipSetDevice(s->device));
  LAUNCH(s, "field smooth", launch_mesh_blur(mesh_args(s, 0.0f), passes, s->field, s->field2,
                                             s->field, s->stream));
  return P4A_OK;
}

P4A_API int p4a_mesh_normal_field(p4a_sim* s, int32_t passes) {
  int rc = check(s);
  if (rc) return rc;
  if (!s->fieldConfigured) return fail(s, P4A_ERR_ARG, "p4a_field_configure first");
  s->normalReady = false;
  if (passes <= 0) return P4A_OK;
  HIP_TRY(s, hipSetDevice(s->device));
  LAUNCH(s, "normal field", launch_mesh_blur(mesh_args(s, 0.0f), passes, s->field, s->field2,
                                             s->nfield, s->stream));
  s->normalReady = true;
  return P4A_OK;
}

P4A_API int p4a_mesh_march(p4a_sim* s, float iso) {
  int rc = check(s);
  if (rc) return rc;
  if (!s->fieldConfigured) return fail(s, P4A_ERR_ARG, "p4a_field_configure first");
  HIP_TRY(s, hipSetDevice(s->device));
  const MeshArgs M = mesh_args(s, iso);
  LAUNCH(s, "march", launch_mesh_march(M, s->field, s->normalReady ? s->nfield : s->field,
                                       s->meshBlockCount, s->meshBlockSum, s->meshBlockStart,
                                       s->meshVerts, s->meshCounter, s->meshCounter + 4,
                                       s->stream));
  return P4A_OK;
}

P4A_API int p4a_mesh_triangles(p4a_sim* s, int64_t* total, int64_t* emitted) {
  int rc = check(s);
  if (rc) return rc;
  if (!s->meshCounter) return fail(s, P4A_ERR_ARG, "p4a_field_configure first");
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  uint32_t c[8];
  HIP_TRY(s, hipMemcpy(c, s->meshCounter, sizeof(c), hipMemcpyDeviceToHost));
  if (total) *total = (int64_t)c[0];
  if (emitted) *emitted = (int64_t)(c[4] / 3u);
  return P4A_OK;
}

// ---- state snapshot ------------------------------------------------------
// Layout (native byte order): SnapHeader, then the arrays in this order:
//   pos float4[n], vel float4[n], pid u32[n], density f32[n]   (live order)
//   bodyOfPid u32[n], restOfPid float4[n]                      (by particle id)
//   cellStart u32[nCells + 2]                                  (last grid build)
//   bodyCentreD f64[4 nB], bodyRotD f64[9 nB], bodyCentre float4[nB],
//   bodyRot float4[3 nB], bodyInvRatio f32[nB]
//   bpos float4[nBoundary]                                     (cell-sorted, w = psi)
namespace {

constexpr char kSnapMagic[8] = {'P', '4', 'A', 'S', 'N', 'A', 'P', '\0'};
constexpr uint32_t kSnapVersion = 1;

struct SnapHeader {
  char magic[8];
  uint32_t version, headerBytes, paramsBytes, pad0;
  uint64_t totalBytes;
  p4a_params params;
  double h, d, mass, denomRest, simTime, timeBank, lastAdvanced;
  double heldTarget[3], heldRate, heldLimit;
  int32_t n, nFluid, cap, pourBudget, nBodies, nBoundary, nCells, heldBody;
  int32_t lastSubsteps, tile1, tile2, pad1;
};

size_t snap_bytes(size_t n, size_t nCells, size_t nB, size_t nBnd) {
  return sizeof(SnapHeader) + n * (16 + 16 + 4 + 4 + 4 + 16) + (nCells + 2) * 4 +
         nB * (32 + 72 + 16 + 48 + 4) + nBnd * 16;
}

struct Cursor {
  unsigned char* w = nullptr;
  const unsigned char* r = nullptr;
  void put(const void* src, size_t bytes) {
    if (bytes) std::memcpy(w, src, bytes);
    w += bytes;
  }
  void get(void* dst, size_t bytes) {
    if (bytes) std::memcpy(dst, r, bytes);
    r += bytes;
  }
};

}  // namespace

P4A_API int p4a_snapshot_size(p4a_sim* s, size_t* bytes) {
  int rc = check(s);
  if (rc) return rc;
  if (!bytes) return P4A_ERR_ARG;
  *bytes = snap_bytes((size_t)s->n, (size_t)s->nCells, (size_t)s->nBodies, s->bposCount);
  return P4A_OK;
}

P4A_API int p4a_snapshot_save(p4a_sim* s, void* dst, size_t bytes) {
  int rc = check(s);
  if (rc) return rc;
  const size_t n = (size_t)s->n, nc = (size_t)s->nCells, nB = (size_t)s->nBodies;
  const size_t need = snap_bytes(n, nc, nB, s->bposCount);
  if (!dst) return P4A_ERR_ARG;
  if (bytes < need) return fail(s, P4A_ERR_RANGE, "snapshot buffer too small");
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  SnapHeader H;
  std::memset(&H, 0, sizeof(H));
  std::memcpy(H.magic, kSnapMagic, 8);
  H.version = kSnapVersion;
  H.headerBytes = (uint32_t)sizeof(SnapHeader);
  H.paramsBytes = (uint32_t)sizeof(p4a_params);
  H.totalBytes = need;
  H.params = s->params;
  H.h = s->h; H.d = s->d; H.mass = s->mass; H.denomRest = s->denomRest;
  H.simTime = s->simTime; H.timeBank = s->timeBank; H.lastAdvanced = s->lastAdvanced;
  for (int k = 0; k < 3; k++) H.heldTarget[k] = s->heldTarget[k];
  H.heldRate = s->heldRate; H.heldLimit = s->heldLimit;
  H.n = s->n; H.nFluid = s->nFluid; H.cap = s->cap; H.pourBudget = s->pourBudget;
  H.nBodies = s->nBodies; H.nBoundary = (int32_t)s->bposCount; H.nCells = s->nCells;
  H.heldBody = s->heldBody; H.lastSubsteps = s->lastSubsteps;
  H.tile1 = s->tile1.variant; H.tile2 = s->tile2.variant;
  Cursor c;
  c.w = (unsigned char*)dst;
  c.put(&H, sizeof(H));
  const int par = s->parity;
  auto d2h = [&](const void* src, size_t b) -> hipError_t {
    const hipError_t e = b ? hipMemcpy(c.w, src, b, hipMemcpyDeviceToHost) : hipSuccess;
    c.w += b;
    return e;
  };
  HIP_TRY(s, d2h(s->A.pos[par], n * 16));
  HIP_TRY(s, d2h(s->A.vel[par], n * 16));
  HIP_TRY(s, d2h(s->A.pid[par], n * 4));
  HIP_TRY(s, d2h(s->A.density, n * 4));
  c.put(s->bodyOfPidH.data(), n * 4);
  c.put(s->restOfPidH.data(), n * 16);
  HIP_TRY(s, d2h(s->A.cellStart, (nc + 2) * 4));
  HIP_TRY(s, d2h(s->A.bodyCentreD, nB * 32));
  HIP_TRY(s, d2h(s->A.bodyRotD, nB * 72));
  HIP_TRY(s, d2h(s->A.bodyCentre, nB * 16));
  HIP_TRY(s, d2h(s->A.bodyRot, nB * 48));
  HIP_TRY(s, d2h(s->A.bodyInvRatio, nB * 4));
  HIP_TRY(s, d2h(s->A.bpos, s->bposCount * 16));
  return P4A_OK;
}

P4A_API int p4a_snapshot_load(p4a_sim* s, const void* src, size_t bytes) {
  if (!s || !src) return P4A_ERR_ARG;
  if (bytes < sizeof(SnapHeader)) return fail(s, P4A_ERR_RANGE, "snapshot too short");
  SnapHeader H;
  std::memcpy(&H, src, sizeof(H));
  if (std::memcmp(H.magic, kSnapMagic, 8) != 0) return fail(s, P4A_ERR_ARG, "not a snapshot");
  if (H.version != kSnapVersion || H.headerBytes != sizeof(SnapHeader) ||
      H.paramsBytes != sizeof(p4a_params))
    return fail(s, P4A_ERR_ARG, "snapshot of another version or build");
  if (H.n < 0 || H.cap < H.n || H.nBodies < 0 || H.nBoundary < 0 || H.nCells < 1 ||
      H.nFluid < 0 || H.nFluid > H.n || H.heldBody > H.nBodies || H.cap > INT32_MAX / 16 ||
      !(H.h > 0) || !(H.d > 0))
    return fail(s, P4A_ERR_ARG, "snapshot header out of range");
  {
    // the grid that reset will build from the saved box must be the saved one
    int dim[3];
    if (grid_cells(H.params.box, H.h, dim) != (long)H.nCells)
      return fail(s, P4A_ERR_ARG, "snapshot grid does not match its box");
  }
  const size_t n = (size_t)H.n, nc = (size_t)H.nCells, nB = (size_t)H.nBodies,
               nBnd = (size_t)H.nBoundary;
  if (H.totalBytes != snap_bytes(n, nc, nB, nBnd) || bytes < H.totalBytes)
    return fail(s, P4A_ERR_RANGE, "snapshot size does not match its header");
  if (H.tile1 < 0 || H.tile1 >= tile_variants(1) || H.tile2 < 0 || H.tile2 >= tile_variants(2))
    return fail(s, P4A_ERR_ARG, "snapshot names an unknown traversal");
  Cursor c;
  c.r = (const unsigned char*)src + sizeof(H);
  std::vector<float> pos(n * 4), vel(n * 4), density(n), rest(n * 4), bpos(nBnd * 4);
  std::vector<uint32_t> pid(n), body(n), cellStart(nc + 2);
  std::vector<double> cD(nB * 4), rD(nB * 9);
  std::vector<float> cF(nB * 4), rF(nB * 12), inv(nB);
  c.get(pos.data(), n * 16);
  c.get(vel.data(), n * 16);
  c.get(pid.data(), n * 4);
  c.get(density.data(), n * 4);
  c.get(body.data(), n * 4);
  c.get(rest.data(), n * 16);
  c.get(cellStart.data(), (nc + 2) * 4);
  c.get(cD.data(), nB * 32);
  c.get(rD.data(), nB * 72);
  c.get(cF.data(), nB * 16);
  c.get(rF.data(), nB * 48);
  c.get(inv.data(), nB * 4);
  c.get(bpos.data(), nBnd * 16);
  // the live order must be a permutation of the particle ids
  std::vector<uint32_t> slotOf((size_t)H.cap);
  for (size_t k = 0; k < slotOf.size(); k++) slotOf[k] = (uint32_t)k;
  {
    std::vector<char> seen(n, 0);
    for (size_t i = 0; i < n; i++) {
      if (pid[i] >= n || seen[pid[i]]) return fail(s, P4A_ERR_ARG, "snapshot ids are not a permutation");
      seen[pid[i]] = 1;
      slotOf[pid[i]] = (uint32_t)i;
    }
    for (size_t k = 0; k < n; k++)
      if (body[k] > (uint32_t)nB) return fail(s, P4A_ERR_ARG, "snapshot names a body that does not exist");
    // the neighbour walks index particles through these row bounds
    if (cellStart[0] != 0 || cellStart[nc] != (uint32_t)n)
      return fail(s, P4A_ERR_ARG, "snapshot grid does not hold n particles");
    for (size_t c2 = 0; c2 < nc; c2++)
      if (cellStart[c2] > cellStart[c2 + 1])
        return fail(s, P4A_ERR_ARG, "snapshot grid rows are not in order");
  }
  // a scene in particle-id order, so reset derives the per-id tables
  std::vector<float> posP(n * 4), velP(n * 4);
  std::vector<uint32_t> phaseP(n * 4, 0u);
  for (size_t i = 0; i < n; i++) {
    std::memcpy(&posP[(size_t)pid[i] * 4], &pos[i * 4], 16);
    std::memcpy(&velP[(size_t)pid[i] * 4], &vel[i * 4], 16);
  }
  for (size_t k = 0; k < n; k++) phaseP[k * 4] = body[k];
  std::vector<float> pts(nBnd * 3), psi(nBnd);
  for (size_t k = 0; k < nBnd; k++) {
    for (int a = 0; a < 3; a++) pts[k * 3 + a] = bpos[k * 4 + a];
    psi[k] = bpos[k * 4 + 3];
  }
  std::vector<double> centre(nB * 3), dens(nB);
  std::vector<int32_t> count(nB, 0), spec(nB, 0);
  for (size_t b = 0; b < nB; b++) {
    for (int a = 0; a < 3; a++) centre[b * 3 + a] = cD[b * 4 + a];
    dens[b] = inv[b] > 0 ? 1.0 / inv[b] : 1.0;
    spec[b] = (int32_t)b;
  }
  for (size_t k = 0; k < n; k++)
    if (body[k]) count[body[k] - 1]++;
  p4a_scene sc;
  std::memset(&sc, 0, sizeof(sc));
  sc.n = H.n;
  sc.n_fluid = H.nFluid;
  sc.n_body = H.n - H.nFluid;
  sc.pos = posP.data();
  sc.vel = velP.data();
  sc.rest = rest.data();
  sc.phase = phaseP.data();
  sc.n_boundary = H.nBoundary;
  sc.boundary_pts = pts.data();
  sc.boundary_psi = psi.data();
  sc.n_bodies = H.nBodies;
  sc.body_centre = centre.data();
  sc.body_density = dens.data();
  sc.body_count = count.data();
  sc.body_spec = spec.data();
  sc.mass = H.mass;
  sc.h = H.h;
  sc.d = H.d;
  sc.denom_rest = H.denomRest;
  p4a_params q = H.params;
  q.pour_budget = H.cap - H.n;
  int rc = p4a_reset(s, &q, &sc);
  if (rc) return rc;
  if (s->nCells != H.nCells) return fail(s, P4A_ERR_HIP, "snapshot grid changed in reset");
  // the live state as it was saved (reset's grid build reordered the scene)
  s->params = H.params;
  s->pourBudget = H.pourBudget;
  const int par = s->parity;
  HIP_TRY(s, hipMemcpy(s->A.pos[par], pos.data(), n * 16, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.vel[par], vel.data(), n * 16, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.pred[par], pos.data(), n * 16, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.pid[par], pid.data(), n * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.slotOf, slotOf.data(), slotOf.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.density, density.data(), n * 4, hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemcpy(s->A.cellStart, cellStart.data(), (nc + 2) * 4, hipMemcpyHostToDevice));
  if (nB) {
    HIP_TRY(s, hipMemcpy(s->A.bodyCentreD, cD.data(), nB * 32, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(s->A.bodyRotD, rD.data(), nB * 72, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(s->A.bodyCentre, cF.data(), nB * 16, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(s->A.bodyRot, rF.data(), nB * 48, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(s->A.bodyInvRatio, inv.data(), nB * 4, hipMemcpyHostToDevice));
  }
  s->predCur = s->A.pred[par];
  s->simTime = H.simTime;
  s->timeBank = H.timeBank;
  s->lastAdvanced = H.lastAdvanced;
  s->lastSubsteps = H.lastSubsteps;
  for (int k = 0; k < 3; k++) s->heldTarget[k] = H.heldTarget[k];
  s->heldRate = H.heldRate;
  s->heldLimit = H.heldLimit;
  s->heldBody = H.heldBody;
  s->tile1.variant = H.tile1;
  s->tile2.variant = H.tile2;
  s->ofsValid = false;
  s->statsValid = false;
  drop_graphs(s);
  return P4A_OK;
}

P4A_API int p4a_sync(p4a_sim* s) {
  if (!s) return P4A_ERR_ARG;
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return P4A_OK;
}

P4A_API int p4a_neighbour_census(p4a_sim* s, uint32_t* count, uint32_t* sum_j, uint32_t* sum_j2) {
  int rc = check(s);
  if (rc) return rc;
  HIP_TRY(s, hipSetDevice(s->device));
  const Uniform U = make_uniform(s, 0.0);
  LAUNCH(s, "census", launch_census(U, s->A, s->predCur, s->tile1, s->stream));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  const size_t b = (size_t)s->n * 4;
  if (count) HIP_TRY(s, hipMemcpy(count, s->A.census[0], b, hipMemcpyDeviceToHost));
  if (sum_j) HIP_TRY(s, hipMemcpy(sum_j, s->A.census[1], b, hipMemcpyDeviceToHost));
  if (sum_j2) HIP_TRY(s, hipMemcpy(sum_j2, s->A.census[2], b, hipMemcpyDeviceToHost));
  return P4A_OK;
}

// Autotuner: time each tile variant of the one-record kernels (on lambda)
// and of the two-record kernels (on xsph + normals) on the live scene and
// keep the fastest.  Both timed kernels only rewrite per-step scratch
// (density, lambda, corr, normal), so the state is unchanged.
P4A_API int p4a_tune(p4a_sim* s, int32_t chosen[2], double* ms, int32_t capacity) {
  int rc = check(s);
  if (rc) return rc;
  HIP_TRY(s, hipSetDevice(s->device));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  const double dt = (1.0 / 60.0) / std::max(1, s->params.substeps);
  const Uniform U = make_uniform(s, dt);
  const bool ex = exact_math(s);
  const bool tension = s->params.surface_tension_k > 0;
  // group 2 is timed as the substep runs it: xsph+normals, then tension
  // (which in the two-phase shapes reuses xsph's hit masks); tension's
  // position and velocity outputs go to scratch, so the live state stays
  TunerScratch S;
  if (tension && s->n > 0)
    for (void*& q : S.buf)
      if (hipMalloc(&q, (size_t)s->n * sizeof(float4)) != hipSuccess)
        return fail(s, P4A_ERR_NOMEM, "tuner scratch");
  float4 *sPos = (float4*)S.buf[0], *sVel = (float4*)S.buf[1];
  HIP_TRY(s, hipEventCreate(&S.ev[0]));
  HIP_TRY(s, hipEventCreate(&S.ev[1]));
  hipEvent_t e0 = S.ev[0], e1 = S.ev[1];
  // the FP16-offset form runs its own lambda and delta kernels, which take
  // only the thread count of the group-1 shape (direct_threads): group 1
  // then times the FP16 lambda at the three direct-walk shapes, and the
  // other shapes (same kernels) report -1
  const bool f16 = fp16_form(s);
  if (f16 && !s->ofsValid) return stale_records(s);
  int k = 0;
  for (int group = 1; group <= 2; group++) {
    float best = 1e30f;
    int bestV = 0;
    for (int v = 0; v < tile_variants(group); v++) {
      TileCfg c{v};
      if (f16 && group == 1 && v > 2) {
        if (ms && k < capacity) ms[k] = -1.0;
        k++;
        continue;
      }
      auto run = [&]() -> hipError_t {
        if (group == 1 && f16)
          return launch_lambda16(U, s->A, s->predCur, ofs_of(s, s->predCur), direct_threads(c),
                                 Drift{}, s->stream);
        if (group == 1) return launch_lambda(U, s->A, s->predCur, ex, c, Drift{}, s->stream);
        hipError_t e = launch_xsph(U, s->A, s->predCur, s->A.pos[s->parity], ex, tension, c,
                                   Drift{}, s->stream);
        if (e != hipSuccess || !sPos) return e;
        return launch_tension(U, s->A, s->predCur, sPos, sVel, ex, c, Drift{}, s->stream);
      };
      HIP_TRY(s, run());
      HIP_TRY(s, hipEventRecord(e0, s->stream));
      for (int r = 0; r < 3; r++) HIP_TRY(s, run());
      HIP_TRY(s, hipEventRecord(e1, s->stream));
      HIP_TRY(s, hipEventSynchronize(e1));
      float t = 0;
      HIP_TRY(s, hipEventElapsedTime(&t, e0, e1));
      t /= 3;
      if (ms && k < capacity) ms[k] = t;
      k++;
      if (t < best) {
        best = t;
        bestV = v;
      }
    }
    if (group == 1) s->tile1.variant = bestV;
    else s->tile2.variant = bestV;
  }
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  drop_graphs(s);
  if (chosen) {
    chosen[0] = s->tile1.variant;
    chosen[1] = s->tile2.variant;
  }
  return P4A_OK;
}

P4A_API const char* p4a_tile_variant_name(int32_t group, int32_t v) {
  return tile_variant_name(group, v);
}

P4A_API int p4a_set_timing(p4a_sim* s, int32_t on) {
  if (!s) return P4A_ERR_ARG;
  s->timing = on != 0;
  s->timed.clear();
  s->eventNext = 0;
  return P4A_OK;
}

P4A_API int p4a_timing(p4a_sim* s, char* names, size_t names_bytes, double* ms,
                       int32_t* launches, int32_t capacity, int32_t* count) {
  if (!s || !count) return P4A_ERR_ARG;
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  std::vector<std::string> keys;
  std::vector<double> tot;
  std::vector<int> cnt;
  for (const TimedLaunch& t : s->timed) {
    float e = 0;
    if (hipEventElapsedTime(&e, t.a, t.b) != hipSuccess) continue;
    size_t k = 0;
    while (k < keys.size() && keys[k] != t.name) k++;
    if (k == keys.size()) {
      keys.push_back(t.name);
      tot.push_back(0);
      cnt.push_back(0);
    }
    tot[k] += e;
    cnt[k] += 1;
  }
  *count = (int32_t)keys.size();
  std::string all;
  for (size_t k = 0; k < keys.size(); k++) {
    if ((int32_t)k < capacity) {
      if (ms) ms[k] = cnt[k] ? tot[k] / cnt[k] : 0.0;
      if (launches) launches[k] = cnt[k];
    }
    all += keys[k];
    all += "\n";
  }
  if (names && names_bytes) {
    std::strncpy(names, all.c_str(), names_bytes - 1);
    names[names_bytes - 1] = 0;
  }
  return P4A_OK;
}

}  // extern "C"
