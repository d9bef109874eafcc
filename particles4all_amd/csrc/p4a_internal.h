// Internal definitions shared by the host orchestration (sim.hip) and the
// gfx950 kernels (kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <vector>

#include "../../include/p4a.h"

namespace p4a {

// FP16-offset neighbour record (8 bytes): x as float, y and z as half
// offsets from the centre of the cell the particle was binned into.
struct __align__(8) HRec {
  float x;
  __half y, z;
};

// The reference's `Params` uniform (wgsl.js:4-40) with the values
// Sim.uploadParams derives (sim.js:371-413).  Passed to every kernel by
// value, so a captured graph carries its own copy.
// Unsigned 32-bit division by an invariant divisor d >= 1 with one high
// multiply (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, figure 4.1): q = (t + ((n - t) >> s1)) >> s2,
// t = umulhi(m, n), exact for every 32-bit n.
struct FastDiv {
  uint32_t d, m, s1, s2;
  FastDiv() : d(1), m(1), s1(0), s2(0) {}
  explicit FastDiv(uint32_t div) : d(div) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < div) l++;
    m = (uint32_t)((((1ull << 32) * ((1ull << l) - div)) / div) + 1);
    s1 = l < 1 ? l : 1;
    s2 = l > 1 ? l - 1 : 0;
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    const uint32_t t = __umulhi(m, n);
    return (t + ((n - t) >> s1)) >> s2;
  }
};

struct Uniform {
  float boxMin[3];
  float dt;
  float boxMax[3];
  float h;
  int gridDim[3];
  FastDiv divGx, divGy;  // cells per row, rows per plane (flat cell id -> x, y, z)
  // the sort records the slot of particle ids in [slotIdLo, slotIdLo +
  // slotIdCount) only: the body particles, which the body kernels reach by id
  uint32_t slotIdLo, slotIdCount;
  float h2;
  float clampMin[3];
  float poly6;
  float clampMax[3];
  float spikyGrad;
  float gravity, mass, rho0, invRho0;
  float cfmEps, sCorrK, sCorrWq, xsphC;
  uint32_t n, nCells, nBoundary, nBodyParts, nBodies;
  float bodyScale, covScale, omega;
  uint32_t sorAverage;
  float invDt, volume, tension, cohes, cohesTerm;
  float minusOne;  // -1 as a run-time operand (keeps fused half-to-float FMAs)
  // fast forms: s_corr = sCorrFast (h^2 - r^2)^12, i.e. -k (poly6 / W(dq))^4
  // folded into one constant
  float sCorrFast;
  // Batched scenes: where this scene's slice starts in the concatenated
  // particle and cell arrays, and in the launch index space (each scene
  // padded to a multiple of 256 so no workgroup spans two scenes).  Zero for
  // a single scene.
  uint32_t pBase, cBase, vBase;
};

// Scene table of a batched launch: the parameter block of every scene and
// the scene of every 256-wide chunk of the launch index space.
struct BatchTab {
  const Uniform* U = nullptr;
  const uint32_t* chunkScene = nullptr;
  uint32_t chunks = 0;  // launch index space = chunks * 256
  uint32_t nScenes = 0;
};

// Drift record of the stale grid: every position update after a grid build
// (delta, shape matching) records the largest per-axis displacement of any
// particle into its own slot (float bits, atomicMax); their sum bounds how
// far a particle can be from the cell it was binned into.  Slots are
// cleared by predict.
constexpr int DRIFT_SLOTS = 256;

// The stable sort ranks the particles of a cell by scanning the cell's
// segment when it holds at most FAT_CELL of them; longer segments are
// listed and sorted by a workgroup-wide radix sort (linear in the segment).
constexpr uint32_t FAT_CELL = 64;

// Neighbour masks of one substep.  The first lambda after a grid build tests
// every candidate of every target's 27-cell stencil and marks, per stencil
// row (its 3 cells, at most 64 candidates), those within reach = h + S of
// the target (S, the skin, in units of h from the parameters).  A later
// kernel of the substep takes only the marked candidates, in stencil order,
// when the target is in the cell it was in at the build and its travel
// since the build plus the largest travel of any particle is below the
// skin: an unmarked candidate then stays at least h away, where every pair
// term is zero, so the sums are the full walk's bit for bit.  Other targets
// are walked in full by a fallback launch: a target that has left the cell
// it was in at the build ("moved") is queued by the kernel that moved it
// (delta, shape matching), so its fallback walk runs on a second stream
// beside the list launch that skips it; a target that fails only the
// travel test (or has a row of more than 64 candidates) is walked in full
// by its own wave when the wave has many such targets, else queued by the
// list launch for a fallback launch after it.
struct NbArgs {
  uint64_t* mask;
  float4* ref;
  uint32_t* ctl;
  uint32_t* queue[2];  // left by a list launch, by launch parity
  uint32_t* moved[2];  // moved targets, by solver iteration parity
  uint32_t stride;
  uint32_t segCap;  // entries per queue segment
  float reach2;  // (h + S)^2 (build)
  float lim;     // S less a rounding margin (use)
  int slot;      // list-launch queue of this launch
  int mslot;     // moved queue this launch's targets are in
};
// The queues are NB_QSEG segments of segCap entries.  A list launch appends
// to segment (workgroup mod NB_QSEG); a moved target i goes to segment
// (i / 64) mod NB_QSEG, so no segment can take more than segCap entries.
constexpr uint32_t NB_QSEG = 256;
inline uint32_t nb_seg_cap(uint32_t n) { return ((n + 63u) / 64u + NB_QSEG - 1u) / NB_QSEG * 64u; }
// ctl: [0] travel bound, then the counts of queue[0], queue[1], moved[0],
// moved[1] (NB_QSEG each)
constexpr uint32_t NB_CTL_WORDS = 1 + 4 * NB_QSEG;
__host__ __device__ inline uint32_t nb_queue_counts(int slot) {
  return 1u + (uint32_t)slot * NB_QSEG;
}
__host__ __device__ inline uint32_t nb_moved_counts(int slot) {
  return 1u + (uint32_t)(2 + slot) * NB_QSEG;
}
// nbRef.w bit of a target with a stencil row of more than 64 candidates
constexpr uint32_t NB_UNLISTED = 0x80000000u;
constexpr int NB_FALLBACK_PARTS = 4;  // workgroups per queue segment
inline size_t fat_list_len(size_t cap) { return cap / (FAT_CELL + 1) + 2; }
struct Drift {
  uint32_t* out = nullptr;  // slot the launch records into (null: none)
};

// Device arrays of one scene.  Particle arrays are in the live
// (cell-sorted) order; `pid` names each slot's particle.
struct Arrays {
  float4* pos[2];       // live positions, ping-pong across the grid build
  float4* pred[2];      // predict output (pre-sort) / sorted copy
  float4* predX;        // solver ping-pong partner
  float4* vel[2];       // velocities, same parity as pos
  uint32_t* pid[2];     // particle id per slot
  uint32_t* key;        // cell id per pre-sort index
  uint32_t* tmp;        // unstable scatter target (pre-sort index per slot)
  // cells holding more than FAT_CELL particles, listed by the sort for its
  // segment sort: [0] = count, [1..] = cell ids (fat_list_len(cap) words)
  uint32_t* fatList;
  uint32_t* order;      // pre-sort index per sorted slot
  uint32_t* slot;       // sorted slot per pre-sort index
  uint32_t* slotOf;     // live slot per particle id
  float* density;
  float* lambda;
  float4* corr;
  float4* normal;
  float4* velq;         // fast path: (pred - pos)/dt and m/max(rho, rho0/2)
  // fast path with tension, two-phase walks: the r^2 < h^2 mask of the first
  // 32-candidate chunk of each of a target's 9 stencil rows, written by xsph
  // and read by tension (same positions, same grid); row r of particle i at
  // hitMask[r * hitStride + i]
  uint32_t* hitMask;
  uint32_t hitStride;
  // per-substep neighbour masks (FP32 forms): built by the first lambda
  // after a grid build, read by the later solver kernels of the substep
  uint64_t* nbMask;     // 9 x nbStride: row r's candidates within reach of target i
  float4* nbRef;        // position at the build, w = cell at the build (bits)
  uint32_t* nbCtl;      // NB_CTL_WORDS: travel bound and queue counts
  uint32_t* nbQueue[2]; // targets a list launch left to the full walk
  uint32_t* nbMoved[2]; // targets that left their build cell, per iteration parity
  uint32_t nbStride;
  HRec* ofs[2];           // FP16 form: offsets of pred[parity] / predX
  uint32_t* drift;      // DRIFT_SLOTS displacement bounds (float bits)
  uint32_t* cellOfSlot; // cell each sorted slot was binned into
  uint32_t* cellCount;  // nCells+1, zero between grid builds
  uint32_t* cellStart;  // nCells+2
  uint32_t* blockSum;
  float4* bpos;         // sorted wall samples, w = psi
  float* bpsi;
  uint32_t* bcellStart; // nCells+2
  // bodies (indexed by particle id or body)
  uint32_t* bodyOfPid;  // 1-based body id per particle id, 0 = fluid
  float4* restOfPid;    // body-local rest offset per particle id
  uint32_t* bodyPids;   // particle ids of body particles, grouped by body
  uint32_t* bodyFirst;  // nBodies+1 offsets into bodyPids
  double* bodyCentreD;  // 4 per body
  double* bodyRotD;     // 9 per body, row-major R
  float* bodyInvRatio;  // 1/max(1e-6, density) as the reference's f32 info.x
  float4* bodyCentre;   // reference layout (f32), bodyCentre buffer
  float4* bodyRot;      // reference layout: 3 rows per body
  // stats
  double* statsPart;    // 4 per block
  double* statsOut;     // 4
  // census
  uint32_t* census[3];
};

// AnisoParams uniform (aniso_wgsl.js:4-21) as mesh.js:221-242 fills it,
// plus the FP64 fallback switch.
struct AnisoArgs {
  float lambda, invSupport, support, radius;
  float kr, kn, ks, stretch, minExtent;
  int neps;
  int fp64Fallback;
  float gapTol;  // relative eigenvalue gap below which the FP64 solve runs
};

// Density field on a voxel-vertex lattice (SurfaceMesh.configure, mesh.js:
// 96-128; meshDensityWGSL, mesh_wgsl.js:23-69).
struct FieldArgs {
  int vertDim[3];
  float voxel;
  float origin[3];
  float invSupport;  // f32(1/h)
  float support2;    // f32(h*h)
  float poly6;       // 315/(64 pi h^9)
  float mass, rho0;
  uint32_t skipBodies;
  uint32_t nVerts;
};
// Fixed-point scale of the anisotropic splat: every term is rounded to a
// multiple of 2^-40 and summed with 64-bit integer atomics, so the field is
// independent of the order the atomics resolve in.
constexpr double FIELD_FIXED_SCALE = 1099511627776.0;  // 2^40

// Launch configuration of the tiled neighbour kernels: an index into the
// variant tables of kernels.hip (group 1: one-record kernels lambda, delta,
// census; group 2: two-record kernels xsph, tension).
struct TileCfg {
  int variant;
};
int tile_variants(int group);
const char* tile_variant_name(int group, int v);

// Drag of a held body (wgsl.js:187-214).
struct DragArgs {
  float goal[3];
  float rate;
  uint32_t held;   // 1-based body id
  float blend;
  float limit;
};

// Ray impulse (wgsl.js:216-248).
struct RayArgs {
  float origin[3];
  float radius;
  float dir[3];
  float speedLimit;
  float impulse[3];
  uint32_t n;
};

// Box-wall move (wgsl.js:125-157).
struct ResizeArgs {
  float oldMin[3], oldSize[3], newMin[3], newSize[3], clampMin[3], clampMax[3];
  uint32_t n;
};

// Kernel launchers (kernels.hip).  All enqueue on `s` and return the
// hipError_t of the launch.
hipError_t launch_predict_count(const Uniform& U, const Arrays& A, int par, hipStream_t s,
                                const BatchTab* B = nullptr);
hipError_t launch_scan(const Uniform& U, const Arrays& A, hipStream_t s);
hipError_t launch_scatter(const Uniform& U, const Arrays& A, int par, bool moveVel,
                          bool fp16, hipStream_t s);
// How a solver launch uses the neighbour masks: not at all, building them
// (lambda after a grid build), or taking them (with the queued targets'
// full walks in a second launch; delta also records travel).
enum class Nb { OFF, BUILD, USE };
struct NbLaunch {
  Nb mode = Nb::OFF;
  NbArgs args{};
  // the stream the moved targets' fallback runs on beside the list launch,
  // and the events that fork it from and join it to the launch stream
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
hipError_t launch_lambda16(const Uniform& U, const Arrays& A, const float4* pred,
                           const HRec* ofs, int threads, const Drift& d, hipStream_t s,
                           const NbLaunch& nb = NbLaunch{});
hipError_t launch_delta16(const Uniform& U, const Arrays& A, const float4* pred,
                          const HRec* ofs, float4* out, HRec* ofsOut, int threads,
                          const Drift& d, hipStream_t s, const NbLaunch& nb = NbLaunch{});
hipError_t launch_lambda(const Uniform& U, const Arrays& A, const float4* pred,
                         bool exact, const TileCfg& t, const Drift& d, hipStream_t s,
                         const BatchTab* B = nullptr, const NbLaunch& nb = NbLaunch{});
hipError_t launch_delta(const Uniform& U, const Arrays& A, const float4* pred, float4* out,
                        bool exact, const TileCfg& t, const Drift& d, hipStream_t s,
                        const BatchTab* B = nullptr, const NbLaunch& nb = NbLaunch{});
hipError_t launch_body_pre(const Uniform& U, const Arrays& A, const float4* pred,
                           hipStream_t s);
// movedSlot: the moved queue of the next solver iteration when neighbour
// masks are in use, -1 otherwise
hipError_t launch_body_solve(const Uniform& U, const Arrays& A, float4* pred, HRec* ofs,
                             uint32_t* driftOut, hipStream_t s, int movedSlot = -1);
hipError_t launch_xsph(const Uniform& U, const Arrays& A, const float4* pred, const float4* pos,
                       bool exact, bool normals, const TileCfg& t, const Drift& d,
                       hipStream_t s, const BatchTab* B = nullptr,
                       const NbLaunch& nb = NbLaunch{});
hipError_t launch_tension(const Uniform& U, const Arrays& A, const float4* pred, float4* pos,
                          float4* vel, bool exact, const TileCfg& t, const Drift& d,
                          hipStream_t s, const BatchTab* B = nullptr,
                          const NbLaunch& nb = NbLaunch{});
hipError_t launch_finalize(const Uniform& U, const Arrays& A, const float4* pred, float4* pos,
                           float4* vel, hipStream_t s, const BatchTab* B = nullptr);
hipError_t launch_prep_finish(const Uniform& U, const Arrays& A, const float4* pred,
                              const float4* pos, hipStream_t s, const BatchTab* B = nullptr);
// The stats reduction with caller-given partials (stats_blocks(n) x 4
// doubles) and output (4 doubles).
hipError_t launch_stats_into(const Uniform& U, const float* density, const float4* vel,
                             double* part, double* out, hipStream_t s);
hipError_t launch_stats(const Uniform& U, const Arrays& A, const float4* vel, hipStream_t s,
                        const BatchTab* B = nullptr);
hipError_t launch_body_drag(const Uniform& U, const Arrays& A, float4* vel, const DragArgs& d,
                            hipStream_t s);
hipError_t launch_impulse(const float4* pos, float4* vel, const RayArgs& r, hipStream_t s);
hipError_t launch_resize(const ResizeArgs& r, float4* pos, hipStream_t s);
hipError_t launch_census(const Uniform& U, const Arrays& A, const float4* pred,
                         const TileCfg& t, hipStream_t s);
hipError_t launch_aniso(const Uniform& U, const Arrays& A, const float4* pos,
                        const AnisoArgs& a, float4* smoothPos, float4* g, uint32_t* fallbacks,
                        bool exact, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_field_iso(const Uniform& U, const Arrays& A, const FieldArgs& F,
                            const float4* pos, const uint32_t* pid, float4* rec, float* field,
                            hipStream_t s);
hipError_t launch_field_aniso(const Uniform& U, const Arrays& A, const FieldArgs& F,
                              const float4* smoothPos, const float4* g, const uint32_t* pid,
                              float4* recA, float4* recB, unsigned long long* acc,
                              float* field, hipStream_t s);
hipError_t launch_copy4(const float4* a, float4* b, size_t n, hipStream_t s);
hipError_t launch_export_phase(uint32_t n, const uint32_t* pid, const uint32_t* bodyOfPid,
                               uint4* out, hipStream_t s);
hipError_t launch_body_pack(uint32_t count, const float4* centre, const float4* rot,
                            const float4* stat, float4* out, hipStream_t s);
// Exclusive scan of count[0..n) into out[0..n], out[n] = the total;
// consumed counts are zeroed.  blockSum holds scan_blocks(n) entries.
int scan_blocks(uint32_t n);
hipError_t launch_exclusive_scan(uint32_t* count, uint32_t n, uint32_t* blockSum, uint32_t* out,
                                 hipStream_t s);

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8).  Hand each XCD a contiguous range of tiles so neighbouring
// tiles, which share halo rows, share one L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  int xcd = bid & 7, k = bid >> 3;
  int per = nb >> 3, rem = nb & 7;
  return xcd * per + min(xcd, rem) + k;
}

// Surface extraction on the field lattice (SurfaceMesh.smoothField,
// buildNormalField, generate's march; mesh_wgsl.js:70-244).
struct MeshArgs {
  int vertDim[3];
  float iso;
  FastDiv divVx, divVy;    // vertices per row, rows per plane
  FastDiv divNx, divNy;    // voxels per row, voxel rows per plane
  uint32_t nVerts, nVox;
  uint32_t maxTris;
  uint32_t nBlocks;        // march workgroups (256 voxels each)
};
constexpr int MESH_WG = 256;
hipError_t launch_mesh_blur(const MeshArgs& M, int passes, const float* src, float* tmp,
                            float* dst, hipStream_t s);
hipError_t launch_mesh_march(const MeshArgs& M, const float* field, const float* nfield,
                             uint32_t* blockCount, uint32_t* blockSum, uint32_t* blockStart,
                             uint2* verts, uint32_t* counter, uint32_t* indirect, hipStream_t s);

int stats_blocks(uint32_t n);

// The autotuners' device scratch and timing events, released on every
// return path, an error part way through included.
struct TunerScratch {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  TunerScratch() = default;
  TunerScratch(const TunerScratch&) = delete;
  TunerScratch& operator=(const TunerScratch&) = delete;
  ~TunerScratch() {
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
    for (void* p : buf)
      if (p) hipFree(p);
  }
};

// Host-side restatements shared by the single-scene and batched handles.
struct SceneConsts {
  p4a_params params;
  double h = 0, mass = 0, denomRest = 0, bodyScale = 0, covScale = 0;
  int gridDim[3] = {1, 1, 1};
  int n = 0, nCells = 0, nBoundary = 0, nBodyParts = 0, nBodies = 0;
};
// Sim.uploadParams (sim.js:371-413): FP64 host arithmetic, then f32.
Uniform make_uniform_from(const SceneConsts& s, double dt);
// sortBoundary (sim.js:255-286) into a grid of the given dimensions.
void sort_boundary_grid(double h, const int* dim, int nCells, const float* pts, const float* psi,
                        int nb, std::vector<float>& sorted, std::vector<float>& sortedPsi,
                        std::vector<uint32_t>& start);

}  // namespace p4a
