// Batched scenes: many independent simulations stepped together on one GPU.
//
// The reference runs one scene per `Sim` (src/sim.js); this handle holds
// many.  All scenes' particles, grid cells and wall samples are
// concatenated, scene after scene, so:
//  * the grid build is one predict+count launch, one scan over every cell
//    of every scene (the exclusive scan of the concatenated counts yields
//    cell starts that are already global particle indices) and one stable
//    scatter over every particle;
//  * every neighbour stage is one launch over a launch index space in
//    which each scene is padded to a multiple of 256, so a workgroup lies
//    in exactly one scene and reads that scene's parameter block (box, h,
//    mass, gravity, XSPH, ...) from a device table;
//  * rigid bodies of all scenes are one list, one workgroup per body.
// The substep (same kernel sequence for every scene) is captured as one
// hipGraph per parity and replayed; a change of dt only rewrites the
// parameter table, which the graph reads from device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/p4a.h"
#include "p4a_internal.h"

using namespace p4a;

namespace {

struct SceneMeta {
  SceneConsts c;
  uint32_t pBase = 0, cBase = 0, vBase = 0, bBase = 0, bodyBase = 0;
};

}  // namespace

struct p4a_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int S = 0;
  int iterations = 4;
  bool tension = false, exact = false;
  std::vector<SceneMeta> m;
  uint32_t N = 0, C = 0, NB = 0, chunks = 0, nBodies = 0, nBodyParts = 0;
  Arrays A;
  Uniform* dU = nullptr;
  uint32_t* dChunkScene = nullptr;
  BatchTab tab;
  double uploadedDt = -1.0;
  int parity = 0;
  float4* predCur = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // traversal shapes of the one-record (lambda, delta) and two-record
  // (xsph, tension) neighbour kernels; p4a_batch_tune picks them
  TileCfg tile1{2}, tile2{2};
  std::vector<void*> allocs;
};

namespace {

// tile variants with a batched form (the LDS row tiles have none)
constexpr int kBatchVariants[] = {0, 1, 2, 5, 6, 7, 8, 9};

int fail(p4a_batch* b, int code, const std::string& msg) {
  b->err = msg;
  return code;
}

#define BTRY(b, expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) return fail((b), P4A_ERR_HIP, std::string(#expr ": ") + \
                                                          hipGetErrorString(e_)); \
  } while (0)

#define BLAUNCH(b, call)                                                          \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return fail((b), P4A_ERR_HIP, std::string(#call ": ") +  \
                                                           hipGetErrorString(e_)); \
  } while (0)

template <class T>
int balloc(p4a_batch* b, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess)
    return fail(b, P4A_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " +
                                      hipGetErrorString(e));
  b->allocs.push_back(q);
  *p = (T*)q;
  return P4A_OK;
}

void free_all(p4a_batch* b) {
  for (int k = 0; k < 2; k++)
    if (b->graph[k]) hipGraphExecDestroy(b->graph[k]);
  for (void* q : b->allocs) hipFree(q);
  b->allocs.clear();
}

// Parameter table for a substep of dt: one block per scene, with its slice
// bases; written outside any graph (the graph reads the table).
int upload_uniforms(p4a_batch* b, double dt) {
  if (dt == b->uploadedDt) return P4A_OK;
  std::vector<Uniform> u(b->S);
  for (int s = 0; s < b->S; s++) {
    u[s] = make_uniform_from(b->m[s].c, dt);
    u[s].pBase = b->m[s].pBase;
    u[s].cBase = b->m[s].cBase;
    u[s].vBase = b->m[s].vBase;
  }
  BTRY(b, hipMemcpyAsync(b->dU, u.data(), u.size() * sizeof(Uniform), hipMemcpyHostToDevice,
                         b->stream));
  BTRY(b, hipStreamSynchronize(b->stream));
  b->uploadedDt = dt;
  return P4A_OK;
}

// The batch-wide block for the global kernels (scan, scatter, bodies):
// totals only.
Uniform totals(const p4a_batch* b) {
  Uniform U;
  std::memset(&U, 0, sizeof(U));
  U.n = b->N;
  U.nCells = b->C;
  U.nBodies = b->nBodies;
  U.nBodyParts = b->nBodyParts;
  U.minusOne = -1.0f;
  U.slotIdCount = 0xffffffffu;  // every id's slot (bodies of all scenes, not one range)
  // the delta walk forms neighbour counts when any scene averages by them
  for (const auto& m : b->m)
    if (m.c.params.sor_average) U.sorAverage = 1u;
  return U;
}

float4* other_pred(p4a_batch* b) {
  return b->predCur == b->A.pred[b->parity] ? b->A.predX : b->A.pred[b->parity];
}

int run_grid(p4a_batch* b, bool moveVel) {
  const Uniform T = totals(b);
  BLAUNCH(b, launch_scan(T, b->A, b->stream));
  BLAUNCH(b, launch_scatter(T, b->A, b->parity, moveVel, false, b->stream));
  b->parity ^= 1;
  b->predCur = b->A.pred[b->parity];
  return P4A_OK;
}

int run_substep_body(p4a_batch* b) {
  const Uniform T = totals(b);
  const BatchTab* B = &b->tab;
  const TileCfg t = b->tile1, t2 = b->tile2;
  BLAUNCH(b, launch_predict_count(T, b->A, b->parity, b->stream, B));
  int rc = run_grid(b, false);
  if (rc) return rc;
  const int par = b->parity;
  if (b->nBodies > 0) BLAUNCH(b, launch_body_pre(T, b->A, b->predCur, b->stream));
  for (int it = 0; it < std::max(1, b->iterations); it++) {
    BLAUNCH(b, launch_lambda(T, b->A, b->predCur, b->exact, t, Drift{}, b->stream, B));
    float4* out = other_pred(b);
    Drift d;
    d.out = b->A.drift + std::min(2 * it, DRIFT_SLOTS - 1);
    BLAUNCH(b, launch_delta(T, b->A, b->predCur, out, b->exact, t, d, b->stream, B));
    b->predCur = out;
    if (b->nBodies > 0)
      BLAUNCH(b, launch_body_solve(T, b->A, b->predCur, nullptr,
                                   b->A.drift + std::min(2 * it + 1, DRIFT_SLOTS - 1),
                                   b->stream));
  }
  if (!b->exact)
    BLAUNCH(b, launch_prep_finish(T, b->A, b->predCur, b->A.pos[par], b->stream, B));
  BLAUNCH(b, launch_xsph(T, b->A, b->predCur, b->A.pos[par], b->exact, b->tension, t2, Drift{},
                         b->stream, B));
  if (b->tension)
    BLAUNCH(b, launch_tension(T, b->A, b->predCur, b->A.pos[par], b->A.vel[par], b->exact, t2,
                              Drift{}, b->stream, B));
  else
    BLAUNCH(b, launch_finalize(T, b->A, b->predCur, b->A.pos[par], b->A.vel[par], b->stream,
                               B));
  return P4A_OK;
}

int run_substep(p4a_batch* b, double dt) {
  int rc = upload_uniforms(b, dt);
  if (rc) return rc;
  const int par0 = b->parity;
  if (!b->graph[par0]) {
    BTRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = run_substep_body(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return fail(b, P4A_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph[par0], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess)
      return fail(b, P4A_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    b->parity = par0;  // capture recorded the work only
  }
  BTRY(b, hipGraphLaunch(b->graph[par0], b->stream));
  b->parity = par0 ^ 1;
  b->predCur = (std::max(1, b->iterations) % 2 == 0) ? b->A.pred[b->parity] : b->A.predX;
  return P4A_OK;
}

}  // namespace

P4A_API int p4a_batch_create(int32_t device, int32_t n_scenes, const p4a_params* params,
                             const p4a_scene* scenes, p4a_batch** out) {
  if (!out || n_scenes < 1 || !params || !scenes) return P4A_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return P4A_ERR_NODEV;
  p4a_batch* b = new p4a_batch();
  *out = b;
  b->device = device;
  // shapes named in the environment, as for a single scene (only shapes with
  // a batched form)
  for (int g = 1; g <= 2; g++) {
    const char* e = std::getenv(g == 1 ? "P4A_TILE1" : "P4A_TILE2");
    if (!e) continue;
    const int v = std::atoi(e);
    if (std::find(std::begin(kBatchVariants), std::end(kBatchVariants), v) !=
        std::end(kBatchVariants))
      (g == 1 ? b->tile1 : b->tile2).variant = v;
  }
  if (hipSetDevice(device) != hipSuccess) return fail(b, P4A_ERR_HIP, "hipSetDevice");
  BTRY(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  b->S = n_scenes;
  b->iterations = params[0].iterations;
  b->tension = params[0].surface_tension_k > 0;
  b->exact = params[0].math == P4A_MATH_EXACT;
  b->m.resize(n_scenes);
  // slices
  uint32_t N = 0, C = 0, NB = 0, V = 0, nBodies = 0;
  for (int s = 0; s < n_scenes; s++) {
    const p4a_params& p = params[s];
    const p4a_scene& sc = scenes[s];
    if (p.iterations != b->iterations || (p.surface_tension_k > 0) != b->tension ||
        (p.math == P4A_MATH_EXACT) != b->exact)
      return fail(b, P4A_ERR_ARG, "scenes of a batch share iterations, tension on/off and math");
    if (p.precision != P4A_PREC_FP32)
      return fail(b, P4A_ERR_ARG, "batched scenes run the FP32 form");
    if (sc.n < 0 || (sc.n > 0 && (!sc.pos || !sc.vel || !sc.rest || !sc.phase)) ||
        sc.n_boundary < 0 || (sc.n_boundary > 0 && (!sc.boundary_pts || !sc.boundary_psi)) ||
        sc.n_bodies < 0 || (sc.n_bodies > 0 && (!sc.body_centre || !sc.body_density)) ||
        !(sc.h > 0))
      return fail(b, P4A_ERR_ARG, "scene " + std::to_string(s) + ": arrays missing");
    SceneMeta& M = b->m[s];
    M.c.params = p;
    M.c.h = sc.h;
    M.c.mass = sc.mass;
    M.c.denomRest = sc.denom_rest;
    if (!(sc.h > 0)) return fail(b, P4A_ERR_ARG, "scene " + std::to_string(s) + ": h must be > 0");
    long cells = 1;
    for (int k = 0; k < 3; k++) {
      // max(1, floor(box / h)), kept a double until it is known to fit
      const double g = std::floor(p.box[k] / sc.h);
      if (!(g < 2.0e9)) return fail(b, P4A_ERR_RANGE, "batch too large");
      M.c.gridDim[k] = g >= 1 ? (int)g : 1;
      cells *= M.c.gridDim[k];
      if (cells > (long)INT32_MAX) return fail(b, P4A_ERR_RANGE, "batch too large");
    }
    M.c.nCells = (int)cells;
    M.c.n = sc.n;
    M.c.nBoundary = p.no_boundary ? 0 : sc.n_boundary;
    M.c.nBodies = sc.n_bodies;
    int parts = 0;
    for (int i = 0; i < sc.n; i++) parts += sc.phase[(size_t)i * 4] != 0;
    M.c.nBodyParts = parts;
    M.pBase = N;
    M.cBase = C;
    M.vBase = V * 256u;
    M.bBase = NB;
    M.bodyBase = nBodies;
    const double total = (double)N + sc.n + (double)C + cells + 2;
    if (total > 2.0e9) return fail(b, P4A_ERR_RANGE, "batch too large");
    N += (uint32_t)sc.n;
    C += (uint32_t)cells;
    NB += (uint32_t)sc.n_boundary;
    V += (uint32_t)((sc.n + 255) / 256);
    nBodies += (uint32_t)sc.n_bodies;
    b->nBodyParts += (uint32_t)parts;
  }
  b->N = N;
  b->C = C;
  b->NB = NB;
  b->chunks = V;
  b->nBodies = nBodies;
  int rc;
  Arrays& A = b->A;
  std::memset(&A, 0, sizeof(A));
#define BALLOC(ptr, n) if ((rc = balloc(b, &(ptr), (n)))) return rc
  for (int k = 0; k < 2; k++) {
    BALLOC(A.pos[k], N);
    BALLOC(A.pred[k], N);
    BALLOC(A.vel[k], N);
    BALLOC(A.pid[k], N);
  }
  BALLOC(A.predX, N);
  BALLOC(A.drift, DRIFT_SLOTS);
  BALLOC(A.cellOfSlot, N);
  BALLOC(A.key, N);
  BALLOC(A.tmp, N);
  BALLOC(A.fatList, fat_list_len(N));
  BALLOC(A.order, N);
  BALLOC(A.slot, N);
  BALLOC(A.slotOf, N);
  BALLOC(A.density, N);
  BALLOC(A.lambda, N);
  BALLOC(A.corr, N);
  BALLOC(A.normal, N);
  BALLOC(A.velq, N);
  BALLOC(A.hitMask, (size_t)N * 9);  // xsph's first-chunk hit masks for tension
  A.hitStride = N;
  BALLOC(A.bodyOfPid, N);
  BALLOC(A.restOfPid, N);
  BALLOC(A.cellCount, (size_t)C + 1);
  BALLOC(A.cellStart, (size_t)C + 2);
  BALLOC(A.blockSum, (size_t)C / 2048 + 2);
  BALLOC(A.bcellStart, (size_t)C + 2);
  BALLOC(A.bpos, NB);
  BALLOC(A.bpsi, NB);
  BALLOC(A.statsPart, (size_t)V * 4 + 4);
  BALLOC(A.statsOut, (size_t)n_scenes * 4);
  BALLOC(b->dU, (size_t)n_scenes);
  BALLOC(b->dChunkScene, V);
  // host images
  std::vector<float> pos((size_t)N * 4), vel((size_t)N * 4), rest((size_t)N * 4, 0.0f);
  std::vector<uint32_t> iota(N), bodyOf(N, 0u), chunkScene(V);
  std::vector<uint32_t> bcs((size_t)C + 2);
  std::vector<float> bpos((size_t)std::max<uint32_t>(1, NB) * 4), bpsi(std::max<uint32_t>(1, NB));
  std::vector<uint32_t> bodyPids, bodyFirst;
  std::vector<double> cD, rD;
  std::vector<float> cF, rF, inv;
  for (uint32_t i = 0; i < N; i++) iota[i] = i;
  for (int s = 0; s < n_scenes; s++) {
    const p4a_scene& sc = scenes[s];
    SceneMeta& M = b->m[s];
    std::memcpy(&pos[(size_t)M.pBase * 4], sc.pos, (size_t)sc.n * 16);
    std::memcpy(&vel[(size_t)M.pBase * 4], sc.vel, (size_t)sc.n * 16);
    for (uint32_t c = M.vBase / 256; c < M.vBase / 256 + (uint32_t)((sc.n + 255) / 256); c++)
      chunkScene[c] = (uint32_t)s;
    // walls: sorted per scene, indices offset by the scene's slice
    std::vector<float> sorted, sortedPsi;
    std::vector<uint32_t> start;
    sort_boundary_grid(sc.h, M.c.gridDim, M.c.nCells, sc.boundary_pts, sc.boundary_psi,
                       sc.n_boundary, sorted, sortedPsi, start);
    for (int c = 0; c < M.c.nCells; c++) bcs[(size_t)M.cBase + c] = M.bBase + start[c];
    if (sc.n_boundary > 0) {
      std::memcpy(&bpos[(size_t)M.bBase * 4], sorted.data(), (size_t)sc.n_boundary * 16);
      std::memcpy(&bpsi[M.bBase], sortedPsi.data(), (size_t)sc.n_boundary * 4);
    }
    // bodies: global ids (1-based) and particle-id lists
    std::vector<std::vector<uint32_t>> members(sc.n_bodies);
    for (int i = 0; i < sc.n; i++) {
      const uint32_t bid = sc.phase[(size_t)i * 4];
      if (bid == 0) continue;
      if (bid > (uint32_t)sc.n_bodies)
        return fail(b, P4A_ERR_ARG, "phase names a body that does not exist");
      members[bid - 1].push_back(M.pBase + (uint32_t)i);
      bodyOf[M.pBase + i] = M.bodyBase + bid;
      for (int k = 0; k < 3; k++)
        rest[((size_t)M.pBase + i) * 4 + k] = sc.rest[(size_t)i * 4 + k];
    }
    for (int q = 0; q < sc.n_bodies; q++) {
      bodyFirst.push_back((uint32_t)bodyPids.size());
      bodyPids.insert(bodyPids.end(), members[q].begin(), members[q].end());
      for (int k = 0; k < 4; k++) {
        cD.push_back(k < 3 ? sc.body_centre[q * 3 + k] : 0.0);
        cF.push_back(k < 3 ? (float)sc.body_centre[q * 3 + k] : 0.0f);
      }
      for (int k = 0; k < 9; k++) rD.push_back(k % 4 == 0 ? 1.0 : 0.0);
      for (int k = 0; k < 12; k++) rF.push_back((k % 4 == k / 4) ? 1.0f : 0.0f);
      inv.push_back((float)(1.0 / std::max(1e-6, sc.body_density[q])));
    }
  }
  bcs[C] = NB;
  bcs[(size_t)C + 1] = NB;
  bodyFirst.push_back((uint32_t)bodyPids.size());
  hipStream_t st = b->stream;
  BTRY(b, hipMemcpyAsync(A.pos[0], pos.data(), (size_t)N * 16, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.pred[0], pos.data(), (size_t)N * 16, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.vel[0], vel.data(), (size_t)N * 16, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.pid[0], iota.data(), (size_t)N * 4, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.slotOf, iota.data(), (size_t)N * 4, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.bodyOfPid, bodyOf.data(), (size_t)N * 4, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.restOfPid, rest.data(), (size_t)N * 16, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.bcellStart, bcs.data(), bcs.size() * 4, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.bpos, bpos.data(), bpos.size() * 4, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(A.bpsi, bpsi.data(), bpsi.size() * 4, hipMemcpyHostToDevice, st));
  BTRY(b, hipMemcpyAsync(b->dChunkScene, chunkScene.data(), (size_t)V * 4,
                         hipMemcpyHostToDevice, st));
  BTRY(b, hipMemsetAsync(A.cellCount, 0, ((size_t)C + 1) * 4, st));
  if (nBodies > 0) {
    BALLOC(A.bodyPids, bodyPids.size());
    BALLOC(A.bodyFirst, bodyFirst.size());
    BALLOC(A.bodyCentreD, (size_t)nBodies * 4);
    BALLOC(A.bodyRotD, (size_t)nBodies * 9);
    BALLOC(A.bodyInvRatio, nBodies);
    BALLOC(A.bodyCentre, nBodies);
    BALLOC(A.bodyRot, (size_t)nBodies * 3);
    BTRY(b, hipMemcpyAsync(A.bodyPids, bodyPids.data(), bodyPids.size() * 4,
                           hipMemcpyHostToDevice, st));
    BTRY(b, hipMemcpyAsync(A.bodyFirst, bodyFirst.data(), bodyFirst.size() * 4,
                           hipMemcpyHostToDevice, st));
    BTRY(b, hipMemcpyAsync(A.bodyCentreD, cD.data(), cD.size() * 8, hipMemcpyHostToDevice, st));
    BTRY(b, hipMemcpyAsync(A.bodyRotD, rD.data(), rD.size() * 8, hipMemcpyHostToDevice, st));
    BTRY(b, hipMemcpyAsync(A.bodyCentre, cF.data(), cF.size() * 4, hipMemcpyHostToDevice, st));
    BTRY(b, hipMemcpyAsync(A.bodyRot, rF.data(), rF.size() * 4, hipMemcpyHostToDevice, st));
    BTRY(b, hipMemcpyAsync(A.bodyInvRatio, inv.data(), inv.size() * 4, hipMemcpyHostToDevice,
                           st));
  }
#undef BALLOC
  b->tab.U = b->dU;
  b->tab.chunkScene = b->dChunkScene;
  b->tab.chunks = V;
  b->tab.nScenes = (uint32_t)n_scenes;
  b->parity = 0;
  b->predCur = A.pred[0];
  // primeGrid (sim.js:223-253) of every scene: predict with dt = 0, grid
  if ((rc = upload_uniforms(b, 0.0))) return rc;
  BLAUNCH(b, launch_predict_count(totals(b), A, 0, st, &b->tab));
  if ((rc = run_grid(b, true))) return rc;
  BTRY(b, hipStreamSynchronize(st));
  return P4A_OK;
}

P4A_API int p4a_batch_destroy(p4a_batch* b) {
  if (!b) return P4A_ERR_ARG;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  free_all(b);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return P4A_OK;
}

P4A_API const char* p4a_batch_last_error(p4a_batch* b) { return b ? b->err.c_str() : ""; }

P4A_API int p4a_batch_substeps(p4a_batch* b, double dt, int32_t count) {
  if (!b || count < 0 || !(dt >= 0)) return P4A_ERR_ARG;
  BTRY(b, hipSetDevice(b->device));
  for (int k = 0; k < count; k++) {
    int rc = run_substep(b, dt);
    if (rc) return rc;
  }
  if (count > 0) {
    const Uniform T = totals(b);
    BLAUNCH(b, launch_stats(T, b->A, b->A.vel[b->parity], b->stream, &b->tab));
  }
  return P4A_OK;
}

P4A_API int p4a_batch_tune(p4a_batch* b, double dt, int32_t chosen[2], double* ms,
                           int32_t capacity) {
  if (!b || !(dt >= 0)) return P4A_ERR_ARG;
  BTRY(b, hipSetDevice(b->device));
  BTRY(b, hipStreamSynchronize(b->stream));
  int rc = upload_uniforms(b, dt);
  if (rc) return rc;
  const Uniform T = totals(b);
  const BatchTab* B = &b->tab;
  // group 2 as the substep runs it: xsph+normals, then tension, whose
  // position and velocity outputs go to scratch so the scenes' state stays
  TunerScratch S;
  if (b->tension && b->N > 0)
    for (void*& q : S.buf)
      if (hipMalloc(&q, (size_t)b->N * sizeof(float4)) != hipSuccess)
        return fail(b, P4A_ERR_NOMEM, "tuner scratch");
  float4 *sPos = (float4*)S.buf[0], *sVel = (float4*)S.buf[1];
  BTRY(b, hipEventCreate(&S.ev[0]));
  BTRY(b, hipEventCreate(&S.ev[1]));
  hipEvent_t e0 = S.ev[0], e1 = S.ev[1];
  int k = 0;
  for (int group = 1; group <= 2; group++) {
    float best = 1e30f;
    int bestV = 2;
    for (int v : kBatchVariants) {
      const TileCfg c{v};
      auto run = [&]() -> hipError_t {
        if (group == 1)
          return launch_lambda(T, b->A, b->predCur, b->exact, c, Drift{}, b->stream, B);
        hipError_t e = launch_xsph(T, b->A, b->predCur, b->A.pos[b->parity], b->exact,
                                   b->tension, c, Drift{}, b->stream, B);
        if (e != hipSuccess || !sPos) return e;
        return launch_tension(T, b->A, b->predCur, sPos, sVel, b->exact, c, Drift{}, b->stream,
                              B);
      };
      BTRY(b, run());
      BTRY(b, hipEventRecord(e0, b->stream));
      for (int r = 0; r < 3; r++) BTRY(b, run());
      BTRY(b, hipEventRecord(e1, b->stream));
      BTRY(b, hipEventSynchronize(e1));
      float t = 0;
      BTRY(b, hipEventElapsedTime(&t, e0, e1));
      t /= 3;
      if (ms && k < capacity) ms[k] = t;
      k++;
      if (t < best) {
        best = t;
        bestV = v;
      }
    }
    (group == 1 ? b->tile1 : b->tile2).variant = bestV;
  }
  BTRY(b, hipStreamSynchronize(b->stream));
  // the next substep captures its graphs with the chosen shapes
  for (int p = 0; p < 2; p++)
    if (b->graph[p]) {
      hipGraphExecDestroy(b->graph[p]);
      b->graph[p] = nullptr;
    }
  if (chosen) {
    chosen[0] = b->tile1.variant;
    chosen[1] = b->tile2.variant;
  }
  return P4A_OK;
}

P4A_API int p4a_batch_sync(p4a_batch* b) {
  if (!b) return P4A_ERR_ARG;
  BTRY(b, hipStreamSynchronize(b->stream));
  return P4A_OK;
}

P4A_API int p4a_batch_scene_size(p4a_batch* b, int32_t scene, int32_t* n) {
  if (!b || !n || scene < 0 || scene >= b->S) return P4A_ERR_ARG;
  *n = b->m[scene].c.n;
  return P4A_OK;
}

// One scene's slice of a particle buffer in the live order: P4A_BUF_POS,
// VEL, PRED, DENSITY, LAMBDA or PID (particle ids local to the scene).
P4A_API int p4a_batch_read(p4a_batch* b, int32_t scene, int32_t id, void* dst, size_t offset,
                           size_t bytes) {
  if (!b || !dst || scene < 0 || scene >= b->S) return P4A_ERR_ARG;
  const SceneMeta& M = b->m[scene];
  const size_t n = (size_t)M.c.n;
  const void* src = nullptr;
  size_t elem = 16;
  switch (id) {
    case P4A_BUF_POS: src = b->A.pos[b->parity] + M.pBase; break;
    case P4A_BUF_VEL: src = b->A.vel[b->parity] + M.pBase; break;
    case P4A_BUF_PRED: src = b->predCur + M.pBase; break;
    case P4A_BUF_DENSITY: src = b->A.density + M.pBase; elem = 4; break;
    case P4A_BUF_LAMBDA: src = b->A.lambda + M.pBase; elem = 4; break;
    case P4A_BUF_PID: src = b->A.pid[b->parity] + M.pBase; elem = 4; break;
    default: return fail(b, P4A_ERR_ARG, "buffer not readable per scene");
  }
  if (offset > n * elem || bytes > n * elem - offset)
    return fail(b, P4A_ERR_RANGE, "read past the scene's slice");
  if (id == P4A_BUF_PID && (offset % 4 || bytes % 4))
    return fail(b, P4A_ERR_ARG, "ids are read in whole words");
  BTRY(b, hipSetDevice(b->device));
  BTRY(b, hipMemcpyAsync(dst, (const char*)src + offset, bytes, hipMemcpyDeviceToHost,
                         b->stream));
  BTRY(b, hipStreamSynchronize(b->stream));
  if (id == P4A_BUF_PID) {
    uint32_t* q = (uint32_t*)dst;
    for (size_t k = 0; k < bytes / 4; k++) q[k] -= M.pBase;
  }
  return P4A_OK;
}

// Stats of the last p4a_batch_substeps, per scene (as p4a_stats).
P4A_API int p4a_batch_stats(p4a_batch* b, int32_t scene, double out[4]) {
  if (!b || !out || scene < 0 || scene >= b->S) return P4A_ERR_ARG;
  BTRY(b, hipSetDevice(b->device));
  BTRY(b, hipMemcpyAsync(out, b->A.statsOut + (size_t)scene * 4, 32, hipMemcpyDeviceToHost,
                         b->stream));
  BTRY(b, hipStreamSynchronize(b->stream));
  return P4A_OK;
}
