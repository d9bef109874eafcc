// gfx950 kernels of the Particles4All simulation step.
//
// Reference kernels (src/wgsl.js unless named) and their counterparts here:
//   predict 56-70 + count 352-362        -> k_predict_count (fused)
//   scanBlock/scanBlocks/scanAdd 364-423 -> k_scan_reduce + k_scan_add
//   scatterSlot 425-438 (atomic slot)    -> k_scatter_tmp + k_sort_move
//   scatterMove/scatterPhase 440-475        (stable order, fused move)
//   lambda 477-543                       -> k_direct<LambdaOp/LambdaFast>
//   delta 545-612                        -> k_direct<DeltaOp/DeltaFast>
//   bodyCentre/bodySeed/bodyRefLoad      -> k_body_pre      (FP64)
//   bodyCentre/Seed/Cov/Resolve/Project  -> k_body_solve    (FP64)
//   velFromPos 72-83, xsph 85-123,
//   normals 250-288                      -> k_prep_finish + k_direct<XsphOp/XsphFast>
//   tension 290-335 + finalize 337-350   -> k_direct<TensionOp/TensionFast>
//   finalize 337-350                     -> k_finalize (no tension)
//   stats 785-832                        -> k_stats_partial + k_stats_final
//   bodyDrag 187-214, impulse 216-248, resize 125-157 -> k_body_drag,
//                                           k_impulse, k_resize
//   anisoWGSL (aniso_wgsl.js:3-159)      -> k_tile<AnisoOp> (5x5x5 cells)
//   meshDensityWGSL (mesh_wgsl.js:23-69) -> k_field_records + k_field_iso
//   scatterPhase's body buffer (wgsl.js:460-475) -> k_export_phase (on request)
//   bodyPackWGSL (solid_wgsl.js:3-26)    -> k_body_pack
// Forms of this build without a reference kernel: the FP16-offset solver
// iterations (k_direct16<Lambda16/Delta16T>), the anisotropic field splat
// (k_field_aniso_prep + k_field_splat) and the batched-scene wrappers
// (*_batch).
//
// Neighbour kernels: the direct walk (k_direct) gives each target particle
// one thread, in cell-sorted order, and walks its 27-cell stencil row by
// row straight out of global memory; consecutive lanes share stencils, so
// the gathers are served from L1/L2.  The LDS row-tile form (k_tile: a
// workgroup stages the records of the 3x3 neighbouring rows of TX cells
// once) is kept as an autotuner candidate and runs the anisotropy pass.
// Both visit the same candidates in the same order.
//
// Arithmetic: with EXACT the float operations are those of the reference
// kernels, one IEEE-rounded operation each (the file is compiled with
// -ffp-contract=off), so a float CPU evaluation of the same sequence gives
// identical bits.  The one fixed choice is r^2 = fma(z,z, fma(y,y, x*x)),
// used by every path.  Without EXACT, sqrt/division pairs become the
// hardware reciprocal square root and constants are folded.
#include "p4a_internal.h"

#include <type_traits>

namespace p4a {
namespace {

constexpr int WG = 256;
constexpr int SCAN_CH = 2048;  // cells per scan block (256 threads x 8)

__device__ __forceinline__ float r2_of(float x, float y, float z) {
  return __fmaf_rn(z, z, __fmaf_rn(y, y, x * x));
}

// cellOf + cellIndex (wgsl.js:43-49); the clamp is done on the float so a
// NaN position lands in cell 0 instead of an undefined integer.
struct Cell { int x, y, z; };
__device__ __forceinline__ Cell cell_of(const Uniform& U, float px, float py, float pz) {
  float fx = floorf((px - U.boxMin[0]) / U.h);
  float fy = floorf((py - U.boxMin[1]) / U.h);
  float fz = floorf((pz - U.boxMin[2]) / U.h);
  fx = fminf(fmaxf(fx, 0.0f), (float)(U.gridDim[0] - 1));
  fy = fminf(fmaxf(fy, 0.0f), (float)(U.gridDim[1] - 1));
  fz = fminf(fmaxf(fz, 0.0f), (float)(U.gridDim[2] - 1));
  return Cell{(int)fx, (int)fy, (int)fz};
}
__device__ __forceinline__ uint32_t cell_index(const Uniform& U, Cell c) {
  return (uint32_t)((c.z * U.gridDim[1] + c.y) * U.gridDim[0] + c.x);
}

// Cell coordinates of a flat cell index and the FP16-offset record of a
// position binned into that cell: x kept as float, y and z as half offsets
// from the cell's centre boxMin + (index + 0.5) h, rounded to nearest even.
__device__ __forceinline__ void cell_xyz(const Uniform& U, uint32_t c, int& x, int& y, int& z) {
  // two divisions by invariant integers as multiplications (no integer
  // divide instruction on the GPU)
  const uint32_t gx = (uint32_t)U.gridDim[0], gy = (uint32_t)U.gridDim[1];
  const uint32_t yz = U.divGx.div(c);
  const uint32_t zz = U.divGy.div(yz);
  x = (int)(c - yz * gx);
  y = (int)(yz - zz * gy);
  z = (int)zz;
}
__device__ __forceinline__ float cell_centre(const Uniform& U, int a, int k) {
  return U.boxMin[a] + ((float)k + 0.5f) * U.h;
}
__device__ __forceinline__ HRec encode_ofs(const Uniform& U, uint32_t c, float px, float py,
                                           float pz) {
  int x, y, z;
  cell_xyz(U, c, x, y, z);
  HRec o;
  o.x = px;
  o.y = __float2half_rn(py - cell_centre(U, 1, y));
  o.z = __float2half_rn(pz - cell_centre(U, 2, z));
  return o;
}

// poly6(r2) of wgsl.js:50-53 for r2 < h2: ((poly6 * t) * t) * t.
__device__ __forceinline__ float poly6_w(const Uniform& U, float r2) {
  float t = U.h2 - r2;
  return U.poly6 * t * t * t;
}

// Records the largest per-axis displacement (p -> o) into a drift slot: an
// atomic only from a lane that raises the stored bound.  These records run
// in the divergent tails of the walks (lanes that left early, queued
// targets), so they use no wave-wide reduction, which would read the
// registers of inactive lanes.  Non-negative floats order like their bits;
// a NaN or infinite displacement stores a bound that switches culling off.
__device__ __forceinline__ void raise_bound(uint32_t* out, float d) {
  const uint32_t m = __float_as_uint(d);
  if (m > *(volatile const uint32_t*)out) atomicMax(out, m);
}
__device__ __forceinline__ void record_drift(uint32_t* out, float px, float py, float pz,
                                             float ox, float oy, float oz) {
  if (!out) return;
  raise_bound(out, fmaxf(fabsf(ox - px), fmaxf(fabsf(oy - py), fabsf(oz - pz))));
}

// Appends target i (where `want`) to a moved-target queue, segment
// (i / 64) mod NB_QSEG.  Lanes of a wave that append together into one
// segment take their places with one atomic, others one atomic each.  Any
// subset of a wave's lanes may call it (the ballots see the active lanes).
__device__ __forceinline__ void enqueue_moved(uint32_t* __restrict__ q, uint32_t* cnt,
                                              uint32_t segCap, bool want, uint32_t i) {
  const uint64_t qm = __ballot(want);
  if (!qm) return;
  const int lane = __lane_id();
  const uint32_t seg = (i >> 6) & (NB_QSEG - 1u);
  const int leader = __builtin_ctzll(qm);
  const uint32_t segL = (uint32_t)__shfl((int)seg, leader, 64);
  if (!__ballot(want && seg != segL)) {
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&cnt[seg], (uint32_t)__popcll(qm));
    base = (uint32_t)__shfl((int)base, leader, 64);
    const uint32_t at = base + (uint32_t)__popcll(qm & ((1ull << lane) - 1ull));
    if (want && at < segCap) q[(size_t)seg * segCap + at] = i;
  } else if (want) {
    const uint32_t at = atomicAdd(&cnt[seg], 1u);
    if (at < segCap) q[(size_t)seg * segCap + at] = i;
  }
}

// ---------------------------------------------------------------------------
// Wave helpers for runs of equal keys in a 64-lane wavefront.
// Valid lanes form a prefix of the wave (i < n); each maximal run of equal
// keys among adjacent valid lanes is handled by its first lane (the head).
struct Run {
  bool head;
  int headLane;
  int len;  // valid for heads
};
__device__ __forceinline__ Run wave_runs(uint32_t key, bool ok) {
  const int lane = threadIdx.x & 63;
  uint32_t prev = __shfl_up(key, 1, 64);
  bool head = ok && (lane == 0 || prev != key);
  unsigned long long heads = __ballot(head);
  unsigned long long valid = __ballot(ok);
  Run r;
  r.head = head;
  unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  unsigned long long hl = heads & le;
  r.headLane = hl ? 63 - __builtin_clzll(hl) : 0;
  unsigned long long gt = (lane == 63) ? 0ull : (~0ull << (lane + 1));
  unsigned long long nx = heads & gt;
  int end = nx ? __builtin_ctzll(nx) : __popcll(valid);
  r.len = end - lane;
  return r;
}

// ---------------------------------------------------------------------------
// predict (wgsl.js:56-70) fused with count (wgsl.js:352-362).
// v = vel + dt*(0,-g,0); pred = (pos + dt*v, 1).  The updated v itself is not
// stored: after the grid build the reference overwrites vel with
// (pred - pos)/dt (velFromPos, wgsl.js:72-83) before anything reads it.
__device__ __forceinline__ void
predict_body(const Uniform& U, uint32_t i, bool ok, const float4* __restrict__ pos,
             const float4* __restrict__ vel, float4* __restrict__ pred,
             uint32_t* __restrict__ key, uint32_t* __restrict__ count) {
  uint32_t c = 0xffffffffu;
  if (ok) {
    float4 p = pos[i], v = vel[i];
    float gx = U.dt * 0.0f, gy = U.dt * (-U.gravity), gz = U.dt * 0.0f;
    float vx = v.x + gx, vy = v.y + gy, vz = v.z + gz;
    float qx = p.x + U.dt * vx, qy = p.y + U.dt * vy, qz = p.z + U.dt * vz;
    pred[i] = make_float4(qx, qy, qz, 1.0f);
    c = U.cBase + cell_index(U, cell_of(U, qx, qy, qz));
    key[i] = c;
  }
  Run r = wave_runs(c, ok);
  if (r.head) atomicAdd(&count[c], (uint32_t)r.len);
}

__global__ void __launch_bounds__(WG)
k_predict_count(Uniform U, const float4* __restrict__ pos, const float4* __restrict__ vel,
                float4* __restrict__ pred, uint32_t* __restrict__ key,
                uint32_t* __restrict__ count, uint32_t* __restrict__ drift) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  static_assert(DRIFT_SLOTS <= WG, "block 0 clears the drift slots");
  if (blockIdx.x == 0 && threadIdx.x < DRIFT_SLOTS) drift[threadIdx.x] = 0u;
  predict_body(U, i, i < U.n, pos, vel, pred, key, count);
}

// Batched: one 256-chunk of one scene per workgroup.
__global__ void __launch_bounds__(WG)
k_predict_count_batch(BatchTab B, const float4* __restrict__ pos,
                      const float4* __restrict__ vel, float4* __restrict__ pred,
                      uint32_t* __restrict__ key, uint32_t* __restrict__ count,
                      uint32_t* __restrict__ drift) {
  if (blockIdx.x == 0 && threadIdx.x < DRIFT_SLOTS) drift[threadIdx.x] = 0u;
  const Uniform& U = B.U[B.chunkScene[blockIdx.x]];
  const uint32_t local = blockIdx.x * WG - U.vBase + threadIdx.x;
  predict_body(U, U.pBase + local, local < U.n, pos, vel, pred, key, count);
}

// Exclusive scan of cellCount into cellStart (wgsl.js:364-423), two passes:
// block totals, then each block adds the totals before it and scans its own
// 2048 cells.  Counts are zeroed as they are consumed; they are reused as
// the scatter cursor.
__global__ void __launch_bounds__(WG)
k_scan_reduce(const uint32_t* __restrict__ count, uint32_t nCells, uint32_t* __restrict__ blockSum) {
  __shared__ uint32_t red[WG / 64];
  const uint32_t base = blockIdx.x * SCAN_CH + threadIdx.x * 8;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s += (base + k < nCells) ? count[base + k] : 0u;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < WG / 64; w++) t += red[w];
    blockSum[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(WG)
k_scan_add(uint32_t* __restrict__ count, uint32_t nCells, const uint32_t* __restrict__ blockSum,
           uint32_t* __restrict__ cellStart) {
  __shared__ uint32_t red[WG / 64];
  __shared__ uint32_t woff[WG / 64];
  __shared__ uint32_t sPrefix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Sum of the block totals before this block.
  uint32_t pre = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += WG) pre += blockSum[b];
  for (int off = 32; off > 0; off >>= 1) pre += __shfl_xor(pre, off, 64);
  if (lane == 0) red[wave] = pre;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < WG / 64; w++) t += red[w];
    sPrefix = t;
  }
  // Thread-local counts.
  const uint32_t base = blockIdx.x * SCAN_CH + threadIdx.x * 8;
  uint32_t v[8];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    v[k] = (base + k < nCells) ? count[base + k] : 0u;
    s += v[k];
  }
  // Inclusive wave scan of thread sums.
  uint32_t inc = s;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  __syncthreads();
  if (lane == 63) woff[wave] = inc;
  __syncthreads();
  uint32_t wpre = 0;
  for (int w = 0; w < wave; w++) wpre += woff[w];
  uint32_t run = sPrefix + wpre + inc - s;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (base + k < nCells) {
      cellStart[base + k] = run;
      count[base + k] = 0u;
    }
    run += v[k];
  }
  if (base <= nCells && nCells < base + 8) {
    // the thread owning the last cell writes cellStart[nCells] = total
    cellStart[nCells] = run;
  }
}

// Unstable placement by atomics (the reference's scatterSlot order), one
// atomic per run of equal cells in a wave; k_sort_move then orders every
// cell's segment by pre-sort index, which makes the result the stable
// counting sort whatever order the atomics resolved in.  Thread 0 also
// empties the fat-cell list k_sort_move fills.
__global__ void __launch_bounds__(WG)
k_scatter_tmp(uint32_t n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ cellStart,
              uint32_t* __restrict__ cursor, uint32_t* __restrict__ tmp,
              uint32_t* __restrict__ fatList) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i == 0) fatList[0] = 0u;
  const bool ok = i < n;
  const uint32_t c = ok ? key[i] : 0xffffffffu;
  Run r = wave_runs(c, ok);
  uint32_t base = 0;
  if (r.head) base = atomicAdd(&cursor[c], (uint32_t)r.len);
  base = __shfl(base, r.headLane, 64);
  if (ok) {
    const int lane = threadIdx.x & 63;
    const uint32_t d = cellStart[c] + base + (uint32_t)(lane - r.headLane);
    if (d < n) tmp[d] = i;  // holds unless counts and keys disagree
  }
}

// The moves of scatterMove and scatterPhase (wgsl.js:440-475) for one
// particle: pre-sort index i of cell c goes to sorted slot d.  Writes slot
// (the reference's scatterSlot output) and the inverse map order.
// Velocities move only when `vel` is given: inside a substep they are dead
// until velFromPos rewrites them, so the hot path skips that traffic.
struct MoveArgs {
  const float4* __restrict__ pos;
  const float4* __restrict__ pred;
  const uint32_t* __restrict__ pid;
  float4* __restrict__ posOut;
  float4* __restrict__ predOut;
  uint32_t* __restrict__ pidOut;
  uint32_t* __restrict__ order;
  uint32_t* __restrict__ slot;
  uint32_t* __restrict__ slotOf;
  const float4* __restrict__ vel;
  float4* __restrict__ velOut;
  HRec* __restrict__ ofsOut;
  uint32_t* __restrict__ cellOfSlot;
};
__device__ __forceinline__ void move_one(const Uniform& U, const MoveArgs& M, uint32_t i,
                                         uint32_t d, uint32_t c) {
  M.order[d] = i;
  M.slot[i] = d;
  M.posOut[d] = M.pos[i];
  M.predOut[d] = M.pred[i];
  const uint32_t q = M.pid[i];
  M.pidOut[d] = q;
  // a scattered 4-byte store per particle otherwise: kept for the ids the
  // body kernels look up
  if (q - U.slotIdLo < U.slotIdCount) M.slotOf[q] = d;
  if (M.vel) M.velOut[d] = M.vel[i];
  if (M.ofsOut) {
    const float4 p = M.pred[i];
    M.ofsOut[d] = encode_ofs(U, c, p.x, p.y, p.z);
    M.cellOfSlot[d] = c;
  }
}

// Stable rank inside each cell segment of at most FAT_CELL particles (a
// scan of the segment), then the moves as a gather into sorted slots.  The
// first thread of a longer segment lists its cell for k_sort_fat instead,
// so no thread scans more than FAT_CELL entries whatever the occupancy.
__global__ void __launch_bounds__(WG)
k_sort_move(uint32_t n, const uint32_t* __restrict__ tmp, const uint32_t* __restrict__ key,
            const uint32_t* __restrict__ cellStart, uint32_t* __restrict__ count,
            uint32_t* __restrict__ fatList, uint32_t fatCap, MoveArgs M, Uniform U) {
  const uint32_t s = blockIdx.x * WG + threadIdx.x;
  if (s >= n) return;
  const uint32_t i = tmp[s];
  const uint32_t c = key[i];
  // clamped so a segment stays inside [0, n) even where counts and keys
  // disagree
  const uint32_t b = min(cellStart[c], n), e = max(b, min(cellStart[c + 1], n));
  if (s == b) count[c] = 0u;
  if (e - b > FAT_CELL) {
    if (s == b) {
      const uint32_t f = atomicAdd(&fatList[0], 1u);
      if (f < fatCap) fatList[1 + f] = c;
    }
    return;
  }
  uint32_t rank = 0;
  for (uint32_t t = b; t < e; t++) rank += (tmp[t] < i) ? 1u : 0u;
  const uint32_t d = b + rank;
  if (d >= n) return;  // holds unless counts and keys disagree
  move_one(U, M, i, d, c);
}

// Segments longer than FAT_CELL: one workgroup per listed cell sorts the
// segment's pre-sort indices (distinct values) by an LSD radix sort of
// 8-bit digits over value - min, ping-ponging between the segment's slices
// of tmp and order; each pass is a histogram, an exclusive scan and a
// scatter of 256-entry chunks in order, ranked inside a chunk by digit
// matching over wave ballots, so the pass is stable.  Work is linear in the
// segment length (times the number of digits of its index range).
__global__ void __launch_bounds__(WG)
k_sort_fat(uint32_t n, uint32_t* __restrict__ tmp, const uint32_t* __restrict__ cellStart,
           const uint32_t* __restrict__ fatList, uint32_t fatCap, MoveArgs M, Uniform U) {
  static_assert(WG == 256, "one thread per digit");
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[WG / 64][256];
  __shared__ uint32_t red[2][WG / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t nFat = min(fatList[0], fatCap);
  for (uint32_t f = blockIdx.x; f < nFat; f += gridDim.x) {
    const uint32_t c = fatList[1 + f];
    const uint32_t b = min(cellStart[c], n), e = max(b, min(cellStart[c + 1], n));
    const uint32_t k = e - b;
    // index range of the segment
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (uint32_t t = tid; t < k; t += WG) {
      const uint32_t v = tmp[b + t];
      lo = min(lo, v);
      hi = max(hi, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
      lo = min(lo, (uint32_t)__shfl_xor((int)lo, off, 64));
      hi = max(hi, (uint32_t)__shfl_xor((int)hi, off, 64));
    }
    if (lane == 0) {
      red[0][wave] = lo;
      red[1][wave] = hi;
    }
    __syncthreads();
    lo = red[0][0];
    hi = red[1][0];
    for (int w = 1; w < WG / 64; w++) {
      lo = min(lo, red[0][w]);
      hi = max(hi, red[1][w]);
    }
    const uint32_t span = hi - lo;
    const int bits = span ? 32 - __builtin_clz(span) : 0;
    const int passes = (bits + 7) / 8;
    uint32_t* src = tmp + b;
    uint32_t* dst = M.order + b;
    for (int ps = 0; ps < passes; ps++) {
      const int sh = 8 * ps;
      base[tid] = 0u;
      for (int w = 0; w < WG / 64; w++) wcnt[w][tid] = 0u;
      __syncthreads();
      for (uint32_t t = tid; t < k; t += WG) atomicAdd(&base[((src[t] - lo) >> sh) & 255u], 1u);
      __syncthreads();
      // exclusive scan of the 256 digit counts (one per thread)
      const uint32_t cnt = base[tid];
      uint32_t inc = cnt;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= off) inc += u;
      }
      if (lane == 63) red[0][wave] = inc;
      __syncthreads();
      uint32_t pre = 0;
      for (int w = 0; w < wave; w++) pre += red[0][w];
      base[tid] = pre + inc - cnt;
      __syncthreads();
      for (uint32_t c0 = 0; c0 < k; c0 += WG) {
        const uint32_t t = c0 + (uint32_t)tid;
        const bool valid = t < k;
        const uint32_t v = valid ? src[t] : 0u;
        const uint32_t dg = ((v - lo) >> sh) & 255u;
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
          const bool on = (dg >> bit) & 1u;
          const unsigned long long bal = __ballot(valid && on);
          peers &= on ? bal : ~bal;
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0u) wcnt[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
          uint32_t d = base[dg] + rank;
          for (int w = 0; w < wave; w++) d += wcnt[w][dg];
          dst[d] = v;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < WG / 64; w++) {
          add += wcnt[w][tid];
          wcnt[w][tid] = 0u;
        }
        base[tid] += add;
        __syncthreads();
      }
      uint32_t* sw = src;
      src = dst;
      dst = sw;
    }
    // src holds the segment in ascending pre-sort index order
    for (uint32_t t = tid; t < k; t += WG) move_one(U, M, src[t], b + t, c);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Row-tiled neighbour traversal over a (2CR+1)^3-cell stencil (CR = 1 for
// the solver kernels, 2 for the anisotropy pass).
template <int NREC, int TX, int MAXREC, int CR>
struct TileLds {
  static constexpr int ROWS = (2 * CR + 1) * (2 * CR + 1);
  static constexpr int W = TX + 2 * CR + 1;
  float4 rec[NREC][MAXREC + 4];
  uint32_t cs[ROWS][W];     // cellStart of cells xa..xb of each row
  int off[ROWS + 1];        // LDS offset of each row's first record
  uint32_t base[ROWS];      // global index of each row's first record
  int anyB;
  int nonEmpty;
};

// Default hooks: the cell of a target is cellOf(pred) (wgsl.js:43-46) and
// no row is skipped.
struct OpBase {
  // a lambda list launch empties the moved queue its iteration's delta fills
  static constexpr bool CLEARS_MOVED = false;
  uint32_t* driftOut = nullptr;  // drift slot a position update records into
  // a position update after a neighbour-mask build records the largest
  // distance of any particle from its build position (nbTravel, float
  // bits; a NaN distance stores a bound that fails every later check)
  const float4* nbRef = nullptr;
  uint32_t* nbTravel = nullptr;
  // ... and appends the targets whose new position lies outside the cell
  // they were in at the build to the moved queue of the next iteration
  uint32_t* nbMovedQ = nullptr;
  uint32_t* nbMovedCnt = nullptr;
  uint32_t nbSegCap = 0;
  __device__ void record_nb(const Uniform& U, uint32_t i, float x, float y, float z) const {
    if (!nbTravel) return;
    const float4 r = nbRef[i];
    raise_bound(nbTravel, sqrtf(r2_of(x - r.x, y - r.y, z - r.z)));
    if (nbMovedQ) {
      const uint32_t built = __float_as_uint(r.w) & ~NB_UNLISTED;
      enqueue_moved(nbMovedQ, nbMovedCnt, nbSegCap,
                    cell_index(U, cell_of(U, x, y, z)) != built, i);
    }
  }
  // two-phase walks (3x3 rows): the first chunk's hit mask of each row, kept
  // for a later kernel over the same positions and grid, or taken from an
  // earlier one instead of testing the chunk again (row r of target i at
  // [r * maskStride + i]); null: test every chunk
  uint32_t* maskOut = nullptr;
  const uint32_t* maskIn = nullptr;
  uint32_t maskStride = 0;
  template <class T>
  __device__ Cell cell(const Uniform& U, const T& t) const { return cell_of(U, t.px, t.py, t.pz); }
  template <class T>
  __device__ bool skip_row(const Uniform&, const T&, int, int) const { return false; }
  // The walks call this when the target's own record lay in none of the
  // rows they visited (it moved more than a cell from where it was binned):
  // a fast form that takes its self term inside the sum adds it here.
  template <class T>
  __device__ void self_outside(const Uniform&, T&) const {}
};

template <class Op, int TX, int NT, int MAXREC, int CR>
__global__ void __launch_bounds__(NT)
k_tile(Uniform U, Op op, const uint32_t* __restrict__ cellStart,
       const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
       int tilesPerRow) {
  using L = TileLds<Op::NREC, TX, MAXREC, CR>;
  constexpr int ROWS = L::ROWS, SIDE = 2 * CR + 1, CW = L::W;
  __shared__ L S;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int txi = tile % tilesPerRow;
  const int ty = (tile / tilesPerRow) % gy;
  const int tz = tile / (tilesPerRow * gy);
  const int x0 = txi * TX;
  const int x1 = min(x0 + TX, gx);
  const int xa = max(x0 - CR, 0), xb = min(x1 + CR, gx);
  const int tid = threadIdx.x;

  // Empty tiles leave at once.
  const uint32_t rowC = (uint32_t)((tz * gy + ty) * gx);
  if (tid == 0) S.nonEmpty = cellStart[rowC + x1] != cellStart[rowC + x0];
  __syncthreads();
  if (!S.nonEmpty) return;

  // cellStart slices of the neighbour rows.
  const int W = xb - xa + 1;
  for (int e = tid; e < ROWS * CW; e += NT) {
    const int r = e / CW, k = e - r * CW;
    const int z = tz + r / SIDE - CR, y = ty + r % SIDE - CR;
    if (k < W && z >= 0 && z < gz && y >= 0 && y < gy)
      S.cs[r][k] = cellStart[(uint32_t)((z * gy + y) * gx + xa + k)];
  }
  if (tid == 0) S.anyB = 0;
  __syncthreads();
  if (tid < ROWS) {
    const int r = tid;
    const int z = tz + r / SIDE - CR, y = ty + r % SIDE - CR;
    const bool valid = z >= 0 && z < gz && y >= 0 && y < gy;
    const int cnt = valid ? (int)(S.cs[r][W - 1] - S.cs[r][0]) : 0;
    S.base[r] = valid ? S.cs[r][0] : 0u;
    // exclusive prefix over rows, done by each row's thread
    int o = 0;
    for (int q = 0; q < r; q++) {
      const int zq = tz + q / SIDE - CR, yq = ty + q % SIDE - CR;
      if (zq >= 0 && zq < gz && yq >= 0 && yq < gy) o += (int)(S.cs[q][W - 1] - S.cs[q][0]);
    }
    S.off[r] = o;
    if (r == ROWS - 1) S.off[ROWS] = o + cnt;
    if (valid && U.nBoundary > 0 && Op::BOUNDARY) {
      const uint32_t rb = (uint32_t)((z * gy + y) * gx);
      if (bcellStart[rb + xb] != bcellStart[rb + xa]) atomicOr(&S.anyB, 1);
    }
  }
  __syncthreads();
  const int total = S.off[ROWS];
  const bool ldsOk = total <= MAXREC;
  if (ldsOk) {
    for (int idx = tid; idx < total; idx += NT) {
      int r = 0;
#pragma unroll
      for (int q = 1; q < ROWS; q++) r += (idx >= S.off[q]) ? 1 : 0;
      const uint32_t j = S.base[r] + (uint32_t)(idx - S.off[r]);
      float4 v[Op::NREC];
      op.load(U, j, v);
#pragma unroll
      for (int k = 0; k < Op::NREC; k++) S.rec[k][idx] = v[k];
    }
  }
  __syncthreads();

  constexpr int RC = CR * SIDE + CR;  // the tile's own row
  const uint32_t T0 = S.cs[RC][x0 - xa], T1 = S.cs[RC][x1 - xa];
  const bool anyB = S.anyB != 0;
  for (uint32_t i = T0 + tid; i < T1; i += NT) {
    typename Op::Target t;
    op.target(U, i, t);
    const Cell c = op.cell(U, t);
    const int xlo = max(c.x - CR, 0), xhi = min(c.x + CR, gx - 1);
    const bool inTile = ldsOk && c.y == ty && c.z == tz && c.x >= x0 && c.x < x1;
    const bool doB = Op::BOUNDARY && U.nBoundary > 0 && (anyB || !inTile);
    bool seen = false;  // the target's own record lay in a visited row
    for (int dz = -CR; dz <= CR; dz++) {
      const int z = c.z + dz;
      if (z < 0 || z >= gz) continue;
      for (int dy = -CR; dy <= CR; dy++) {
        const int y = c.y + dy;
        if (y < 0 || y >= gy) continue;
        if (op.skip_row(U, t, z, y)) continue;
        const uint32_t rb = (uint32_t)((z * gy + y) * gx);
        if (inTile) {
          const int r = (dz + CR) * SIDE + (dy + CR);
          const int rel = S.off[r] - (int)S.base[r];
          const int b = (int)S.cs[r][xlo - xa] + rel;
          const int e = (int)S.cs[r][xhi + 1 - xa] + rel;
          seen = seen || ((int)i + rel >= b && (int)i + rel < e);
          if constexpr (Op::FAST) {
            // branchless, four records in flight; the tail pads with
            // records that lie beyond the support and weigh nothing
            int q = b;
            for (; q + 3 < e; q += 4) {
              float4 v[4][Op::NREC];
#pragma unroll
              for (int u = 0; u < 4; u++)
#pragma unroll
                for (int k = 0; k < Op::NREC; k++) v[u][k] = S.rec[k][q + u];
#pragma unroll
              for (int u = 0; u < 4; u++) op.fpair(U, t, v[u]);
            }
            if (q < e) {
              float4 v[3][Op::NREC];
#pragma unroll
              for (int u = 0; u < 3; u++)
#pragma unroll
                for (int k = 0; k < Op::NREC; k++)
                  v[u][k] = (q + u < e) ? S.rec[k][q + u] : Op::sentinel(k);
#pragma unroll
              for (int u = 0; u < 3; u++) op.fpair(U, t, v[u]);
            }
          } else {
            for (int q = b; q < e; q++) {
              float4 v[Op::NREC];
#pragma unroll
              for (int k = 0; k < Op::NREC; k++) v[k] = S.rec[k][q];
              op.pair(U, t, (uint32_t)(q - rel), v);
            }
          }
        } else {
          const uint32_t b = cellStart[rb + xlo], e = cellStart[rb + xhi + 1];
          seen = seen || (i >= b && i < e);
          for (uint32_t j = b; j < e; j++) {
            float4 v[Op::NREC];
            op.load(U, j, v);
            if constexpr (Op::FAST) op.fpair(U, t, v);
            else op.pair(U, t, j, v);
          }
        }
        if (doB) {
          const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
          for (uint32_t k = bb; k < be; k++) op.bpair(U, t, bpos[k]);
        }
      }
    }
    if constexpr (Op::FAST) {
      if (!seen) op.self_outside(U, t);
    }
    op.finish(U, i, t);
  }
}

// Direct traversal: one thread per particle in sorted order, candidates
// read from global memory (the cell-sorted arrays give each workgroup a
// compact neighbourhood that stays in L1/L2).  Workgroups are remapped so
// each XCD walks a contiguous range of the sorted particles.  Wall samples
// exist only in cells with some index 0 or gridDim-1, so a particle whose
// stencil cannot reach such a cell skips the wall loops.
// PH = 1, 2: two phases per chunk of up to 32 consecutive candidates of a
// row.  Phase 1 only forms r^2 (the same operations as the pair term) and
// sets one bit per candidate inside h; phase 2 runs the pair term for the
// set bits, in the same order (PH = 2: four per trip, loads in flight
// together).  About 15% of the candidates lie inside h, and the full term
// costs ~25 issue slots against ~10 for the test, so the walk issues far
// fewer instructions.  The terms dropped are the ones the fast branchless
// term reduces to zero and the ones the exact term returns from (its own
// r^2, the same bits as phase 1's, is at least h^2), so the exact ops keep
// the full walk's bits.
// With LB (lambda, once per substep) the walk also builds the target's
// neighbour masks: bit k of row r's mask marks candidate b_r + k within
// reach (r^2 < nb->reach2), and the build position and cell go to nbRef.
// Bit for candidate k of a row (k < 64) when r^2 is within reach.
// The mask build shifts each candidate's reach test in from the right (one
// shift-and-add per candidate), so the first of a row's len candidates ends
// at bit len - 1; the stored mask has candidate k at bit k.  A row of more
// than 64 candidates sends its target to the full walk, and its mask is
// never read.
__device__ __forceinline__ uint64_t reach_in(uint64_t acc, float r2, float reach2) {
  return (acc << 1) + (r2 < reach2 ? 1ull : 0ull);
}
__device__ __forceinline__ uint64_t reach_mask(uint64_t acc, uint32_t len) {
  if (len == 0u) return 0ull;
  return __builtin_bitreverse64(acc) >> (64u - min(len, 64u));
}

template <class Op, int NT, int CR, int PH = 0, bool LB = false>
__device__ __forceinline__ void
direct_walk(const Uniform& U, const Op& op, uint32_t i, const uint32_t* __restrict__ cellStart,
            const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
            const NbArgs* nb = nullptr) {
  static_assert(!LB || (PH == 0 && CR == 1), "masks are built by the one-phase 27-cell walk");
  cellStart += U.cBase;
  bcellStart += U.cBase;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  typename Op::Target t;
  op.target(U, i, t);
  const Cell c = op.cell(U, t);
  const int xlo = max(c.x - CR, 0), xhi = min(c.x + CR, gx - 1);
  const bool nearWall = c.x <= CR || c.x >= gx - 1 - CR || c.y <= CR || c.y >= gy - 1 - CR ||
                        c.z <= CR || c.z >= gz - 1 - CR;
  const bool doB = Op::BOUNDARY && U.nBoundary > 0 && nearWall;
  bool seen = false;  // the target's own record lay in a visited row
  bool unlisted = false;
  for (int dz = -CR; dz <= CR; dz++) {
    const int z = c.z + dz;
    if (z < 0 || z >= gz) continue;
    for (int dy = -CR; dy <= CR; dy++) {
      const int y = c.y + dy;
      if (y < 0 || y >= gy) continue;
      if (op.skip_row(U, t, z, y)) continue;
      const uint32_t rb = (uint32_t)((z * gy + y) * gx);
      const uint32_t b = cellStart[rb + xlo], e = cellStart[rb + xhi + 1];
      seen = seen || (i >= b && i < e);
      uint64_t rowMask = 0;
      if constexpr (LB) unlisted = unlisted || (e - b > 64u);
      if constexpr (PH >= 1) {
        const size_t slot = (size_t)((dz + CR) * (2 * CR + 1) + (dy + CR)) * op.maskStride + i;
        for (uint32_t c0 = b; c0 < e; c0 += 32u) {
          const uint32_t len = min(e - c0, 32u);
          const float4* P = op.pred + c0;
          // m shifts left once per candidate and takes the test as its low
          // bit (one add with carry): candidate u ends at bit len - 1 - u
          uint32_t m = 0u, k = 0u;
          if (op.maskIn && c0 == b) {
            m = op.maskIn[slot];  // the same test, made by the kernel before
            k = len;
          }
          for (; k + 3u < len; k += 4u) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; u++) q[u] = P[(size_t)k + u];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const float r2 = r2_of(t.px - q[u].x, t.py - q[u].y, t.pz - q[u].z);
              m = (m + m) + (r2 < U.h2 ? 1u : 0u);
            }
          }
          for (; k < len; k++) {
            const float4 q = P[k];
            const float r2 = r2_of(t.px - q.x, t.py - q.y, t.pz - q.z);
            m = (m + m) + (r2 < U.h2 ? 1u : 0u);
          }
          if (op.maskOut && c0 == b) op.maskOut[slot] = m;
          if constexpr (PH == 1) {
            // highest set bit first: candidates in their walk order; the next
            // hit's records are loaded before the current hit's term runs
            if (m) {
              uint32_t top = 31u - (uint32_t)__builtin_clz(m);
              m ^= 1u << top;
              uint32_t jc = c0 + (len - 1u - top);
              float4 cur[Op::NREC];
              op.load(U, jc, cur);
              while (m) {
                top = 31u - (uint32_t)__builtin_clz(m);
                m ^= 1u << top;
                const uint32_t jn = c0 + (len - 1u - top);
                float4 nxt[Op::NREC];
                op.load(U, jn, nxt);
                if constexpr (Op::FAST) op.fpair(U, t, cur);
                else op.pair(U, t, jc, cur);
#pragma unroll
                for (int r = 0; r < Op::NREC; r++) cur[r] = nxt[r];
                jc = jn;
              }
              if constexpr (Op::FAST) op.fpair(U, t, cur);
              else op.pair(U, t, jc, cur);
            }
          } else {
            // four hits per trip with their loads in flight together; a
            // missing hit is the far sentinel, which every fast pair term
            // reduces to zero
            while (m) {
              uint32_t jj[4];
              bool hit[4];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                hit[u] = m != 0u;
                const uint32_t top = hit[u] ? 31u - (uint32_t)__builtin_clz(m) : 0u;
                jj[u] = c0 + (hit[u] ? len - 1u - top : 0u);
                m = hit[u] ? (m ^ (1u << top)) : 0u;
              }
              float4 v[4][Op::NREC];
#pragma unroll
              for (int u = 0; u < 4; u++) op.load(U, jj[u], v[u]);
#pragma unroll
              for (int u = 0; u < 4; u++) {
                if constexpr (Op::FAST) {
#pragma unroll
                  for (int r = 0; r < Op::NREC; r++)
                    if (!hit[u]) v[u][r] = Op::sentinel(r);
                  op.fpair(U, t, v[u]);
                } else if (hit[u]) {
                  op.pair(U, t, jj[u], v[u]);
                }
              }
            }
          }
        }
      } else if constexpr (Op::FAST) {
        // 64-bit candidate index: one address per group of four, the
        // others as immediate offsets
        size_t j = b;
        const size_t eEnd = e;
        for (; j + 3 < eEnd; j += 4) {
          float4 v[4][Op::NREC];
#pragma unroll
          for (int u = 0; u < 4; u++) op.load(U, j + u, v[u]);
          if constexpr (LB) {
#pragma unroll
            for (int u = 0; u < 4; u++)
              rowMask = reach_in(rowMask, r2_of(t.px - v[u][0].x, t.py - v[u][0].y,
                                                t.pz - v[u][0].z), nb->reach2);
          }
#pragma unroll
          for (int u = 0; u < 4; u++) op.fpair(U, t, v[u]);
        }
        for (; j < eEnd; j++) {
          float4 v[Op::NREC];
          op.load(U, j, v);
          if constexpr (LB)
            rowMask = reach_in(rowMask, r2_of(t.px - v[0].x, t.py - v[0].y, t.pz - v[0].z),
                               nb->reach2);
          op.fpair(U, t, v);
        }
      } else {
        // exact ops: candidates strictly in walk order, four loads in flight
        uint32_t j = b;
        for (; j + 3u < e; j += 4u) {
          float4 v[4][Op::NREC];
#pragma unroll
          for (int u = 0; u < 4; u++) op.load(U, (size_t)j + u, v[u]);
          if constexpr (LB) {
#pragma unroll
            for (int u = 0; u < 4; u++)
              rowMask = reach_in(rowMask, r2_of(t.px - v[u][0].x, t.py - v[u][0].y,
                                                t.pz - v[u][0].z), nb->reach2);
          }
#pragma unroll
          for (int u = 0; u < 4; u++) op.pair(U, t, j + u, v[u]);
        }
        for (; j < e; j++) {
          float4 v[Op::NREC];
          op.load(U, j, v);
          if constexpr (LB)
            rowMask = reach_in(rowMask, r2_of(t.px - v[0].x, t.py - v[0].y, t.pz - v[0].z),
                               nb->reach2);
          op.pair(U, t, j, v);
        }
      }
      if constexpr (LB)
        nb->mask[(size_t)((dz + 1) * 3 + (dy + 1)) * nb->stride + i] = reach_mask(rowMask, e - b);
      if (doB) {
        const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
        for (uint32_t k = bb; k < be; k++) op.bpair(U, t, bpos[k]);
      }
    }
  }
  if constexpr (Op::FAST) {
    if (!seen) op.self_outside(U, t);
  }
  // a target whose own record is in none of its rows (it has moved more
  // than a cell since the grid build) stays on the full walk, which adds
  // its self term
  if constexpr (LB)
    nb->ref[i] = make_float4(
        t.px, t.py, t.pz,
        __uint_as_float(cell_index(U, c) | ((unlisted || !seen) ? NB_UNLISTED : 0u)));
  op.finish(U, i, t);
}

template <class Op, int NT, int CR, int PH = 0>
__global__ void __launch_bounds__(NT)
k_direct(Uniform U, Op op, const uint32_t* __restrict__ cellStart,
         const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos) {
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= U.n) return;
  direct_walk<Op, NT, CR, PH>(U, op, i, cellStart, bcellStart, bpos);
}

// The marked hits of one row half (32 candidates from `base`) in stencil
// order, four per trip with their record reads in flight together; a
// trip's missing hits reread its first hit and take the far sentinel (fast
// ops, whose terms it reduces to zero) or are skipped (exact ops).  `rec`
// reads the records of candidate j.
template <class Op, class Rec>
__device__ __forceinline__ void hit_loop(const Uniform& U, const Op& op,
                                         typename Op::Target& t, uint32_t m, uint32_t base,
                                         const Rec& rec) {
  while (m) {
    uint32_t jj[4];
    bool hit[4];
    jj[0] = base + (uint32_t)__builtin_ctz(m);
    hit[0] = true;
    m &= m - 1u;
#pragma unroll
    for (int u = 1; u < 4; u++) {
      hit[u] = m != 0u;
      jj[u] = hit[u] ? base + (uint32_t)__builtin_ctz(m) : jj[0];
      m &= m - 1u;
    }
    float4 v[4][Op::NREC];
#pragma unroll
    for (int u = 0; u < 4; u++) rec(jj[u], v[u]);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if constexpr (Op::FAST) {
#pragma unroll
        for (int r = 0; r < Op::NREC; r++)
          if (!hit[u]) v[u][r] = Op::sentinel(r);
        op.fpair(U, t, v[u]);
      } else {
        if (hit[u]) op.pair(U, t, jj[u], v[u]);
      }
    }
  }
}

// The same, one hit per trip (records read from LDS, whose latency the
// four-wide trips would only pad).
template <class Op, class Rec>
__device__ __forceinline__ void hit_loop1(const Uniform& U, const Op& op,
                                          typename Op::Target& t, uint32_t m, uint32_t base,
                                          const Rec& rec) {
  while (m) {
    const uint32_t j = base + (uint32_t)__builtin_ctz(m);
    m &= m - 1u;
    float4 v[Op::NREC];
    rec(j, v);
    if constexpr (Op::FAST) op.fpair(U, t, v);
    else op.pair(U, t, j, v);
  }
}

// Whether the wave's targets with `on` set all lie in one grid row of cells
// (same y and z cell) with cell x never decreasing from lane to lane: then
// the bounds of each of their stencil rows never decrease either.  Slot
// order follows the cells the targets were binned into, so this holds
// unless some target has left its cell since the grid build (a delta run
// before the mask build through the stage API).
__device__ __forceinline__ bool one_grid_row(bool on, Cell c) {
  const uint64_t m = __ballot(on);
  if (!m) return true;
  const int yz = c.y * 65536 + c.z;
  const int first = __builtin_amdgcn_readlane(yz, __builtin_ctzll(m));
  if (__ballot(on && yz != first)) return false;
  // x against the nearest lane before it with `on` set
  const int lane = __lane_id();
  const uint64_t below = m & ((1ull << lane) - 1ull);
  const int prev = below ? 63 - __builtin_clzll(below) : lane;
  const int px = __shfl(c.x, prev, 64);
  return !__ballot(on && below != 0ull && c.x < px);
}

// The union [ub, ue) of the row ranges [b, e) of the lanes with `has` set
// (ub >= ue when none): with bounds that never decrease from lane to lane
// (one_grid_row), the first such lane's b and the last one's e; otherwise
// a wave minimum and maximum.
__device__ __forceinline__ void row_union(bool has, uint32_t b, uint32_t e, bool mono,
                                          uint32_t& ub, uint32_t& ue) {
  const uint64_t m = __ballot(has);
  if (!m) {
    ub = 0xffffffffu;
    ue = 0u;
  } else if (mono) {
    ub = (uint32_t)__builtin_amdgcn_readlane((int)b, __builtin_ctzll(m));
    ue = (uint32_t)__builtin_amdgcn_readlane((int)e, 63 - __builtin_clzll(m));
  } else {
    ub = __ockl_wfred_min_u32(has ? b : 0xffffffffu);
    ue = __ockl_wfred_max_u32(has ? e : 0u);
  }
}

// A staged walk (one wave per workgroup, all 64 lanes present; `on` marks
// the lanes with a target).  Per stencil row the wave stages the union of
// its targets' candidate ranges in LDS; each target marks its candidates
// within reach (r^2 < reach2) from there and walks the marked ones in
// stencil order with the row's wall samples after it.  With reach at
// least h every unmarked candidate's term is zero (the fast terms select
// zero, the exact ones return), so the sums are the full walk's bit for
// bit.  A target with a row of more than 64 candidates, or none of whose
// rows holds its own record, takes the full walk afterwards.
//
// BUILD (the first lambda after a grid build, reach h + S): the marks are
// the row masks the list launches take and the target's reference is
// stored; the full walk after it writes both, the reference as unlisted.
template <class Op, bool BUILD, uint32_t STAGE>
__device__ __forceinline__ void staged_walk(const Uniform& U, const Op& op, uint32_t i, bool on,
                                            const uint32_t* __restrict__ cellStart0,
                                            const uint32_t* __restrict__ bcellStart0,
                                            const float4* __restrict__ bpos, float reach2,
                                            const NbArgs* nb, float4 (*S)[STAGE]) {
  // One-record ops stage 64 records of slack so that the marks of a lane
  // past its row read inside the stage; two-record ops clamp those reads to
  // the union instead (their stage holds half as many records).
  constexpr bool CLAMP = Op::NREC > 1;
  constexpr uint32_t SLACK = CLAMP ? 0u : 64u;
  const uint32_t* __restrict__ cellStart = cellStart0 + U.cBase;
  const uint32_t* __restrict__ bcellStart = bcellStart0 + U.cBase;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  typename Op::Target t;
  Cell c{0, 0, 0};
  if (on) {
    op.target(U, i, t);
    c = op.cell(U, t);
  }
  const int xlo = max(c.x - 1, 0);
  const int xhi = min(c.x + 1, gx - 1);
  const bool nearWall = c.x <= 1 || c.x >= gx - 2 || c.y <= 1 || c.y >= gy - 2 ||
                        c.z <= 1 || c.z >= gz - 2;
  const bool doB = Op::BOUNDARY && U.nBoundary > 0 && nearWall && on;
  const bool mono = one_grid_row(on, c);
  bool over = false;  // a row of more than 64 candidates
  bool seen = false;  // the target's own record lay in one of its rows
  for (int dz = -1; dz <= 1; dz++) {
    for (int dy = -1; dy <= 1; dy++) {
      const int z = c.z + dz, y = c.y + dy;
      const bool rowOk = on && z >= 0 && z < gz && y >= 0 && y < gy;
      const uint32_t rb = rowOk ? (uint32_t)((z * gy + y) * gx) : 0u;
      uint32_t b = 0, e = 0;
      if (rowOk) {
        b = cellStart[rb + xlo];
        e = cellStart[rb + xhi + 1];
      }
      over = over || e - b > 64u;
      seen = seen || (i >= b && i < e);
      const uint32_t len = min(e - b, 64u);
      uint32_t ub, ue;
      row_union(rowOk, b, e, mono, ub, ue);
      // the marks of candidates b .. b + len - 1, bit k for candidate b + k;
      // a trip tests four candidates of every lane, those past a lane's row
      // are masked off after the loop
      const uint32_t most = __ockl_wfred_max_u32(len);
      uint32_t mlo = 0u, mhi = 0u;
      const auto marks = [&](const auto& pos) {
        for (uint32_t k0 = 0; k0 < most; k0 += 4u) {
          float4 q[4];
#pragma unroll
          for (int u = 0; u < 4; u++) q[u] = pos(k0 + (uint32_t)u);
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const bool in = r2_of(t.px - q[u].x, t.py - q[u].y, t.pz - q[u].z) < reach2;
            const uint32_t k = k0 + (uint32_t)u;
            if (k < 32u) mlo = in ? (mlo | (1u << k)) : mlo;
            else mhi = in ? (mhi | (1u << (k - 32u))) : mhi;
          }
        }
        mlo &= len >= 32u ? ~0u : (1u << len) - 1u;
        mhi &= len >= 64u ? ~0u : len > 32u ? (1u << (len - 32u)) - 1u : 0u;
      };
      if (ub < ue && ue - ub + SLACK <= STAGE) {
        const uint32_t span = ue - ub;
        for (uint32_t q = (uint32_t)__lane_id(); q < span; q += 64u) {
          float4 v[Op::NREC];
          op.load(U, (size_t)ub + q, v);
#pragma unroll
          for (int r = 0; r < Op::NREC; r++) S[r][q] = v[r];
        }
        __syncthreads();
        // the lane's first candidate in the stage (a lane without the row
        // reads from the stage's start)
        const uint32_t at = rowOk ? b - ub : 0u;
        if constexpr (CLAMP) marks([&](uint32_t k) { return S[0][min(at + k, span - 1u)]; });
        else marks([&](uint32_t k) { return S[0][at + k]; });
        const auto rec = [&](uint32_t j, float4* v) {
#pragma unroll
          for (int r = 0; r < Op::NREC; r++) v[r] = S[r][j - ub];
        };
        hit_loop1(U, op, t, mlo, b, rec);
        if (__ballot(mhi != 0u)) hit_loop1(U, op, t, mhi, b + 32u, rec);
        __syncthreads();
      } else if (ub < ue) {
        // a lane past its row reads the union's first candidate instead (its
        // own b can be n: an empty row past the last occupied cell)
        marks([&](uint32_t k) {
          float4 v[Op::NREC];
          op.load(U, k < len ? (size_t)b + k : (size_t)ub, v);
          return v[0];
        });
        const auto rec = [&](uint32_t j, float4* v) { op.load(U, (size_t)j, v); };
        hit_loop(U, op, t, mlo, b, rec);
        hit_loop(U, op, t, mhi, b + 32u, rec);
      }
      if constexpr (BUILD) {
        if (rowOk)
          nb->mask[(size_t)((dz + 1) * 3 + (dy + 1)) * nb->stride + i] =
              (uint64_t)mlo | ((uint64_t)mhi << 32);
      }
      if (doB && rowOk) {
        const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
        for (uint32_t k = bb; k < be; k++) op.bpair(U, t, bpos[k]);
      }
    }
  }
  if (on) {
    if (!over && seen) {
      op.finish(U, i, t);
      if constexpr (BUILD)
        nb->ref[i] = make_float4(t.px, t.py, t.pz, __uint_as_float(cell_index(U, c)));
    } else if constexpr (BUILD) {
      direct_walk<Op, 64, 1, 0, true>(U, op, i, cellStart0, bcellStart0, bpos, nb);
    } else {
      direct_walk<Op, 64, 1>(U, op, i, cellStart0, bcellStart0, bpos);
    }
  }
}

// Records of one stencil row staged per wave: the union of the wave's row
// ranges, when it holds at most NB_STAGE candidates (a wave's targets lie in
// a few neighbouring cells of one grid row, so its rows overlap).
constexpr uint32_t NB_STAGE = 256;
// The records of a plane's three rows staged per wave (fast one-record ops)
constexpr uint32_t NB_STAGE_PLANE = 320;

// A list launch (one wave per workgroup).  Each target whose masks still
// hold -- it is in the cell it was in at the build, and its travel since
// the build plus the largest travel of any particle stays below the skin --
// takes only its marked candidates of each of its 9 stencil rows, in
// stencil order, each row followed by its wall samples as in the full
// walk.  The wave stages each row's candidates in LDS with coalesced loads,
// so the hits, which differ from lane to lane, are read from LDS rather
// than gathered from memory.  A target that has left its build cell is
// skipped: the kernel that moved it queued it, and that queue's fallback
// walk runs beside this launch.  Other failing targets are queued here
// (one atomic per wave) for the fallback walk after it.  Block 0 empties
// the other list-launch queue, which the previous list launch's fallback
// has consumed, and (lambda launches) the moved queue this iteration's
// delta will fill, whose previous contents the last iteration walked.
template <class Op>
// Eight waves per SIMD for the one-record ops: the scalar registers of the
// full-walk path must not cost the list walk a wave (the two-record ops
// would spill at the 64 vector registers that takes, and keep their own)
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(Op::NREC == 1 ? 8 : Op::FAST ? 7 : 1, 8)))
k_listed(Uniform U, Op op, NbArgs nb, const uint32_t* __restrict__ cellStart,
         const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos) {
  // the fast one-record ops stage a plane's three rows at once
  constexpr bool PLANES = Op::FAST && Op::NREC == 1;
  constexpr uint32_t STAGE = PLANES ? NB_STAGE_PLANE : NB_STAGE / Op::NREC;
  __shared__ float4 S[Op::NREC][STAGE];
  if (blockIdx.x == 0) {
    for (uint32_t k = threadIdx.x; k < NB_QSEG; k += 64u)
      nb.ctl[nb_queue_counts(nb.slot ^ 1) + k] = 0u;
    if (Op::CLEARS_MOVED)
      for (uint32_t k = threadIdx.x; k < NB_QSEG; k += 64u)
        nb.ctl[nb_moved_counts(nb.mslot ^ 1) + k] = 0u;
  }
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * 64u + threadIdx.x;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  const int lane = __lane_id();
  typename Op::Target t;
  Cell c{0, 0, 0};
  bool valid = false;  // the masks hold
  bool fails = false;  // in its build cell, but the masks do not hold
  // appends the lanes with q set to this launch's queue (one atomic per wave)
  const auto enqueue = [&](bool q) {
    const uint64_t qm = __ballot(q);
    if (!qm) return;
    const int leader = __builtin_ctzll(qm);
    uint32_t base = 0;
    const uint32_t seg = blockIdx.x % NB_QSEG;
    if (lane == leader)
      base = atomicAdd(&nb.ctl[nb_queue_counts(nb.slot) + seg], (uint32_t)__popcll(qm));
    base = (uint32_t)__shfl((int)base, leader, 64);
    const uint32_t at = base + (uint32_t)__popcll(qm & ((1ull << lane) - 1ull));
    if (q && at < nb.segCap) nb.queue[nb.slot][seg * nb.segCap + at] = i;
  };
  if (i < U.n) {
    op.target(U, i, t);
    c = op.cell(U, t);
    const float4 ref = nb.ref[i];
    const uint32_t built = __float_as_uint(ref.w);
    // moved targets are the moved queue's (walked beside this launch)
    const bool stays = (built & ~NB_UNLISTED) == cell_index(U, c);
    const float travel = sqrtf(r2_of(t.px - ref.x, t.py - ref.y, t.pz - ref.z));
    const float most = __uint_as_float(*(volatile const uint32_t*)nb.ctl);
    valid = stays && built == cell_index(U, c) && travel + most <= nb.lim;
    fails = stays && !valid;
  }
  // A wave with a quarter of its targets or more failing the test (all of
  // them once some particle has travelled past the skin) walks those in
  // full itself, after the list walk of the others: the fallback launch is
  // sized for a few targets per lane.  Sparser failures are queued for it.
  const bool dense = __popcll(__ballot(fails)) >= 16;
  enqueue(fails && !dense);
  if (__ballot(valid)) {
    const int xlo = max(c.x - 1, 0);
    const int xhi = min(c.x + 1, gx - 1);
    const bool nearWall = c.x <= 1 || c.x >= gx - 2 || c.y <= 1 || c.y >= gy - 2 ||
                          c.z <= 1 || c.z >= gz - 2;
    const bool doB = Op::BOUNDARY && U.nBoundary > 0 && nearWall && valid;
    const bool mono = one_grid_row(valid, c);
    // one stencil row on its own: its hits, then its wall samples
    const auto row_walk = [&](int dz, int dy) {
      const int z = c.z + dz, y = c.y + dy;
      const bool rowOk = valid && z >= 0 && z < gz && y >= 0 && y < gy;
      const uint32_t rb = rowOk ? (uint32_t)((z * gy + y) * gx) : 0u;
      uint32_t b = 0, e = 0;
      uint64_t m64 = 0;
      if (rowOk) {
        b = cellStart[rb + xlo];
        e = cellStart[rb + xhi + 1];
        m64 = nb.mask[(size_t)((dz + 1) * 3 + (dy + 1)) * nb.stride + i];
      }
      // union of the wave's ranges that hold a hit (every lane is active
      // here, the others contribute the neutral values)
      const bool has = m64 != 0ull;
      uint32_t ub, ue;
      row_union(has, b, e, mono, ub, ue);
      if (ub < ue && ue - ub <= STAGE) {
        const uint32_t span = ue - ub;
        for (uint32_t q = (uint32_t)lane; q < span; q += 64u) {
          float4 v[Op::NREC];
          op.load(U, (size_t)ub + q, v);
#pragma unroll
          for (int r = 0; r < Op::NREC; r++) S[r][q] = v[r];
        }
        __syncthreads();
        const auto rec = [&](uint32_t j, float4* v) {
#pragma unroll
          for (int r = 0; r < Op::NREC; r++) v[r] = S[r][j - ub];
        };
        hit_loop1(U, op, t, (uint32_t)m64, b, rec);
        // rows of more than 32 candidates are rare: the upper half only
        // when some lane has hits there
        if (__ballot((m64 >> 32) != 0ull))
          hit_loop1(U, op, t, (uint32_t)(m64 >> 32), b + 32u, rec);
        __syncthreads();
      } else if (has) {
        const auto rec = [&](uint32_t j, float4* v) { op.load(U, (size_t)j, v); };
        hit_loop(U, op, t, (uint32_t)m64, b, rec);
        hit_loop(U, op, t, (uint32_t)(m64 >> 32), b + 32u, rec);
      }
      if (doB && rowOk) {
        const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
        for (uint32_t k = bb; k < be; k++) op.bpair(U, t, bpos[k]);
      }
    };
    bool perRow = true;
    if constexpr (PLANES) perRow = __ballot(doB) != 0ull;
    if (perRow) {
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) row_walk(dz, dy);
    } else if constexpr (PLANES) {
      // No wall samples in this wave: the three rows of a plane are staged
      // together and each lane walks its hits of all three in one loop, so
      // the plane runs as many trips as its longest lane has hits there
      // rather than the sum of each row's longest.
      for (int dz = -1; dz <= 1; dz++) {
        const int z = c.z + dz;
        uint32_t lo[3], at[3], first[3], span[3];
        uint32_t seg = 0u;  // stage offset of the next row
        bool upper = false;  // hits past a row's 32nd candidate
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const int y = c.y + r - 1;
          const bool rowOk = valid && z >= 0 && z < gz && y >= 0 && y < gy;
          uint32_t b = 0, e = 0;
          uint64_t m64 = 0;
          if (rowOk) {
            const uint32_t rb = (uint32_t)((z * gy + y) * gx);
            b = cellStart[rb + xlo];
            e = cellStart[rb + xhi + 1];
            m64 = nb.mask[(size_t)((dz + 1) * 3 + r) * nb.stride + i];
          }
          const bool has = m64 != 0ull;
          uint32_t ub, ue;
          row_union(has, b, e, mono, ub, ue);
          if (ub >= ue) ub = ue = 0u;
          first[r] = ub;
          span[r] = ue - ub;
          lo[r] = (uint32_t)m64;
          at[r] = seg + (b - ub);  // stage position of candidate b
          upper = upper || (m64 >> 32) != 0ull;
          seg += ue - ub;
        }
        if (__ballot(upper) || seg > STAGE) {
          for (int dy = -1; dy <= 1; dy++) row_walk(dz, dy);
          continue;
        }
        uint32_t o = 0u;
#pragma unroll
        for (int r = 0; r < 3; r++) {
          for (uint32_t q = (uint32_t)lane; q < span[r]; q += 64u) {
            float4 v[Op::NREC];
            op.load(U, (size_t)first[r] + q, v);
#pragma unroll
            for (int k = 0; k < Op::NREC; k++) S[k][o + q] = v[k];
          }
          o += span[r];
        }
        __syncthreads();
        // the current row is m0 at a0, the next ones m1 at a1 and m2 at
        // a2; rows without hits drop out first, so m0 is empty only when
        // all three are
        uint32_t m0 = lo[0], a0 = at[0], m1 = lo[1], a1 = at[1], m2 = lo[2], a2 = at[2];
        if (!m1) {
          m1 = m2;
          a1 = a2;
          m2 = 0u;
        }
        if (!m0) {
          m0 = m1;
          a0 = a1;
          m1 = m2;
          a1 = a2;
          m2 = 0u;
        }
        while (m0) {
          const uint32_t q = a0 + (uint32_t)__builtin_ctz(m0);
          m0 &= m0 - 1u;
          float4 v[Op::NREC];
#pragma unroll
          for (int k = 0; k < Op::NREC; k++) v[k] = S[k][q];
          op.fpair(U, t, v);
          const bool next = m0 == 0u;
          m0 = next ? m1 : m0;
          a0 = next ? a1 : a0;
          m1 = next ? m2 : m1;
          a1 = next ? a2 : a1;
          m2 = next ? 0u : m2;
        }
        __syncthreads();
      }
    }
    if (valid) op.finish(U, i, t);
  }
  // a dense wave walks its failing targets with the staged walk over the
  // same stage (reach h + S keeps every term the full walk adds); the
  // two-record ops walk them in full (the staged walk's registers would
  // cost them a wave per SIMD)
  if constexpr (Op::FAST || Op::NREC == 1) {
    if (dense)
      staged_walk<Op, false, STAGE>(U, op, i, fails, cellStart, bcellStart, bpos, nb.reach2,
                                    nullptr, S);
  } else if (dense && fails) {
    direct_walk<Op, 64, 1>(U, op, i, cellStart, bcellStart, bpos);
  }
}

// The mask build: the staged walk with reach h + S, writing the masks and
// references.  It also restarts the travel bound and the queues of the
// list launches.
template <class Op>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(Op::NREC == 1 ? 8 : 1, 8)))
k_staged_build(Uniform U, Op op, NbArgs nb, const uint32_t* __restrict__ cellStart,
               const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos) {
  __shared__ float4 S[Op::NREC][NB_STAGE / Op::NREC];
  if (blockIdx.x == 0)
    for (uint32_t k = threadIdx.x; k < NB_CTL_WORDS; k += 64u) nb.ctl[k] = 0u;
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * 64u + threadIdx.x;
  staged_walk<Op, true>(U, op, i, i < U.n, cellStart, bcellStart, bpos, nb.reach2, &nb, S);
}

// The staged walk as a traversal shape (reach h) over one scene,
template <class Op>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(Op::NREC == 1 ? 8 : 1, 8)))
k_staged(Uniform U, Op op, const uint32_t* __restrict__ cellStart,
         const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos) {
  __shared__ float4 S[Op::NREC][NB_STAGE / Op::NREC];
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * 64u + threadIdx.x;
  staged_walk<Op, false>(U, op, i, i < U.n, cellStart, bcellStart, bpos, U.h2, nullptr, S);
}

// and over the batch's scenes (a wave never spans two of them: their slices
// start on 256-slot chunks).
template <class Op>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(Op::NREC == 1 ? 8 : 1, 8)))
k_staged_batch(BatchTab B, Op op, const uint32_t* __restrict__ cellStart,
               const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos) {
  __shared__ float4 S[Op::NREC][NB_STAGE / Op::NREC];
  const uint32_t v0 = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * 64u;
  const Uniform& U = B.U[B.chunkScene[v0 >> 8]];
  const uint32_t local = v0 - U.vBase + threadIdx.x;
  staged_walk<Op, false>(U, op, U.pBase + local, local < U.n, cellStart, bcellStart, bpos, U.h2,
                         nullptr, S);
}

// The 27-cell walk over FP16-offset records (defined with those records'
// ops below).
template <class Op, bool LB = false>
__device__ void direct16_walk(const Uniform& U, const Op& op, uint32_t i,
                              const uint32_t* __restrict__ cellStart,
                              const uint32_t* __restrict__ bcellStart,
                              const float4* __restrict__ bpos, const HRec* __restrict__ ofs,
                              const uint32_t* __restrict__ cellOfSlot,
                              const NbArgs* nb = nullptr);

// The full 27-cell walk of a target a list launch left (k_fallback).  Few
// waves run it (about one per SIMD, each lane on a target of its own), so
// its time is the latency of its loads rather than issue: each row's
// candidates are read 32 at a time (the tail of a row rereads the row's
// last candidate, whose term is skipped), and the next row's bounds are
// read while the current row is summed, so a row costs about one round of
// loads.  The terms are the full walk's, in its order.
template <class Op>
__device__ void wide_walk(const Uniform& U, const Op& op, uint32_t i,
                          const uint32_t* __restrict__ cellStart,
                          const uint32_t* __restrict__ bcellStart,
                          const float4* __restrict__ bpos) {
  constexpr int W = 32;
  cellStart += U.cBase;
  bcellStart += U.cBase;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  typename Op::Target t;
  op.target(U, i, t);
  const Cell c = op.cell(U, t);
  const int xlo = max(c.x - 1, 0), xhi = min(c.x + 1, gx - 1);
  const bool nearWall = c.x <= 1 || c.x >= gx - 2 || c.y <= 1 || c.y >= gy - 2 || c.z <= 1 ||
                        c.z >= gz - 2;
  const bool doB = Op::BOUNDARY && U.nBoundary > 0 && nearWall;
  // bounds of stencil row k (dz = k / 3 - 1, dy = k % 3 - 1), if in the grid
  const auto bounds = [&](int k, uint32_t& rb, uint32_t& b, uint32_t& e) {
    const int z = c.z + k / 3 - 1, y = c.y + k % 3 - 1;
    if (z < 0 || z >= gz || y < 0 || y >= gy) return false;
    rb = (uint32_t)((z * gy + y) * gx);
    b = cellStart[rb + xlo];
    e = cellStart[rb + xhi + 1];
    return true;
  };
  bool seen = false;  // the target's own record lay in a visited row
  uint32_t nrb = 0, nb0 = 0, ne0 = 0;
  bool nok = bounds(0, nrb, nb0, ne0);
  for (int k = 0; k < 9; k++) {
    const uint32_t rb = nrb, b = nb0, e = ne0;
    const bool ok = nok;
    if (k + 1 < 9) nok = bounds(k + 1, nrb, nb0, ne0);
    if (!ok) continue;
    seen = seen || (i >= b && i < e);
    for (uint32_t j0 = b; j0 < e; j0 += W) {
      float4 v[W][Op::NREC];
#pragma unroll
      for (int u = 0; u < W; u++) op.load(U, (size_t)min(j0 + (uint32_t)u, e - 1u), v[u]);
#pragma unroll
      for (int u = 0; u < W; u++) {
        if (j0 + (uint32_t)u < e) {
          if constexpr (Op::FAST) op.fpair(U, t, v[u]);
          else op.pair(U, t, j0 + (uint32_t)u, v[u]);
        }
      }
    }
    if (doB) {
      const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
      for (uint32_t q = bb; q < be; q++) op.bpair(U, t, bpos[q]);
    }
  }
  if constexpr (Op::FAST) {
    if (!seen) op.self_outside(U, t);
  }
  op.finish(U, i, t);
}

// The targets a list launch queued, walked in full.  The queue segments
// fill unevenly (the queued targets cluster where the flow is fastest), so
// every workgroup first forms the prefix sums of the segment counts and the
// whole grid then deals the concatenated queue out one target per thread
// at a time: the launch lasts about one walk however the targets fell.
//
// MOVED: the moved queue of this iteration, whose targets the list launch
// beside it skips.  A target is walked only if it is still outside its
// build cell: shape matching after the delta that queued it may have moved
// it back, and then the list launch takes it.
//
// F16: the FP16-offset form's walk over its records (ofs, cellOfSlot).
template <class Op, int NT, bool MOVED, bool F16 = false>
__global__ void __launch_bounds__(NT)
k_fallback(Uniform U, Op op, NbArgs nb, const uint32_t* __restrict__ cellStart,
           const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
           const HRec* __restrict__ ofs = nullptr,
           const uint32_t* __restrict__ cellOfSlot = nullptr) {
  static_assert(NT >= 64 && NB_QSEG == 256, "one wave scans 4 counts per lane");
  __shared__ uint32_t pre[NB_QSEG + 1];
  const int tid = threadIdx.x;
  const uint32_t* counts = nb.ctl + (MOVED ? nb_moved_counts(nb.mslot) : nb_queue_counts(nb.slot));
  const uint32_t* queue = MOVED ? nb.moved[nb.mslot] : nb.queue[nb.slot];
  if (tid < 64) {
    // exclusive prefix of the 256 counts: 4 per lane, then across the wave
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[k] = min(counts[4 * tid + k], nb.segCap);
      sum += v[k];
    }
    uint32_t inc = sum;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)inc, off, 64);
      if (tid >= off) inc += u;
    }
    uint32_t run = inc - sum;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      pre[4 * tid + k] = run;
      run += v[k];
    }
    if (tid == 63) pre[NB_QSEG] = run;
  }
  __syncthreads();
  const uint32_t total = pre[NB_QSEG];
  for (uint32_t g = blockIdx.x * NT + (uint32_t)tid; g < total; g += gridDim.x * NT) {
    // the segment holding item g: the last s with pre[s] <= g
    uint32_t lo = 0, hi = NB_QSEG;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pre[mid] <= g) lo = mid;
      else hi = mid;
    }
    const uint32_t item = queue[(size_t)lo * nb.segCap + (g - pre[lo])];
    if (MOVED) {
      const float4 p = op.pred[item];
      const uint32_t built = __float_as_uint(nb.ref[item].w) & ~NB_UNLISTED;
      if (cell_index(U, cell_of(U, p.x, p.y, p.z)) == built) continue;
    }
    if constexpr (F16)
      direct16_walk<Op>(U, op, item, cellStart, bcellStart, bpos, ofs, cellOfSlot);
    else
      wide_walk<Op>(U, op, item, cellStart, bcellStart, bpos);
  }
}

// Batched scenes: each workgroup lies in one scene's 256-aligned slice of
// the launch index space; the scene's parameter block is read from the
// table, and its particle and cell slices are offsets into the
// concatenated arrays (cellStart holds global particle indices).
template <class Op, int NT, int PH = 0>
__global__ void __launch_bounds__(NT)
k_direct_batch(BatchTab B, Op op, const uint32_t* __restrict__ cellStart,
               const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos) {
  static_assert(256 % NT == 0, "a workgroup must not span two scenes");
  const uint32_t v0 = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * NT;
  const Uniform& U = B.U[B.chunkScene[v0 >> 8]];
  const uint32_t local = v0 - U.vBase + threadIdx.x;
  if (local >= U.n) return;
  direct_walk<Op, NT, 1, PH>(U, op, U.pBase + local, cellStart, bcellStart, bpos);
}

// --- lambda (wgsl.js:477-543) ----------------------------------------------
template <bool EXACT>
struct LambdaOp : OpBase {
  static constexpr bool CLEARS_MOVED = true;
  static constexpr bool FAST = false;
  static constexpr bool BOUNDARY = true;
  static constexpr int NREC = 1;
  const float4* __restrict__ pred;
  float* __restrict__ density;
  float* __restrict__ lambda;
  struct Target {
    float px, py, pz;
    uint32_t i;
    float rho, gx, gy, gz, sg2;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const { v[0] = pred[j]; }
  __device__ void target(const Uniform& U, uint32_t i, Target& t) const {
    float4 p = pred[i];
    t.px = p.x; t.py = p.y; t.pz = p.z; t.i = i;
    t.rho = U.mass * (U.poly6 * U.h2 * U.h2 * U.h2);  // P.mass * poly6(0.0)
    t.gx = t.gy = t.gz = t.sg2 = 0.0f;
  }
  __device__ void pair(const Uniform& U, Target& t, uint32_t j, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    if (j == t.i || r2 >= U.h2) return;
    if (EXACT) {
      t.rho = t.rho + U.mass * poly6_w(U, r2);
      if (r2 > 1.0e-12f) {
        const float r = sqrtf(r2);
        const float hr = U.h - r;
        const float s = U.volume * U.spikyGrad * hr * hr / r;
        const float gx = s * rx, gy = s * ry, gz = s * rz;
        t.gx = t.gx + gx; t.gy = t.gy + gy; t.gz = t.gz + gz;
        t.sg2 = t.sg2 + (gx * gx + gy * gy + gz * gz);
      }
    } else {
      const float tt = U.h2 - r2;
      t.rho = __fmaf_rn(U.mass * U.poly6, tt * tt * tt, t.rho);
      if (r2 > 1.0e-12f) {
        const float ir = __builtin_amdgcn_rsqf(r2);
        const float hr = __fmaf_rn(-r2, ir, U.h);
        const float s = (U.volume * U.spikyGrad) * hr * hr * ir;
        t.gx = __fmaf_rn(s, rx, t.gx);
        t.gy = __fmaf_rn(s, ry, t.gy);
        t.gz = __fmaf_rn(s, rz, t.gz);
        t.sg2 = __fmaf_rn(s * s, r2, t.sg2);
      }
    }
  }
  __device__ void bpair(const Uniform& U, Target& t, float4 b) const {
    const float rx = t.px - b.x, ry = t.py - b.y, rz = t.pz - b.z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2 || r2 < 1.0e-12f) return;
    const float psi = b.w;
    if (EXACT) {
      t.rho = t.rho + psi * poly6_w(U, r2);
      const float r = sqrtf(r2);
      const float hr = U.h - r;
      const float s = psi * U.invRho0 * U.spikyGrad * hr * hr / r;
      t.gx = t.gx + s * rx; t.gy = t.gy + s * ry; t.gz = t.gz + s * rz;
    } else {
      const float tt = U.h2 - r2;
      t.rho = __fmaf_rn(psi * U.poly6, tt * tt * tt, t.rho);
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float hr = __fmaf_rn(-r2, ir, U.h);
      const float s = psi * (U.invRho0 * U.spikyGrad) * hr * hr * ir;
      t.gx = __fmaf_rn(s, rx, t.gx);
      t.gy = __fmaf_rn(s, ry, t.gy);
      t.gz = __fmaf_rn(s, rz, t.gz);
    }
  }
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    density[i] = t.rho;
    const float C = t.rho * U.invRho0 - 1.0f;
    float lam = 0.0f;
    if (C > 0.0f) {
      const float gg = t.gx * t.gx + t.gy * t.gy + t.gz * t.gz;
      lam = -C / (gg + t.sg2 + U.cfmEps + 1.0e-12f);
    }
    lambda[i] = lam;
  }
};

// --- delta (wgsl.js:545-612) -----------------------------------------------
template <bool EXACT>
struct DeltaOp : OpBase {
  static constexpr bool FAST = false;
  static constexpr bool BOUNDARY = true;
  static constexpr int NREC = 1;
  const float4* __restrict__ pred;
  const float* __restrict__ lambda;
  float4* __restrict__ out;
  struct Target {
    float px, py, pz, li;
    float dx, dy, dz, bx, by, bz, nc;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const {
    float4 p = pred[j];
    v[0] = make_float4(p.x, p.y, p.z, lambda[j]);
  }
  __device__ void target(const Uniform& U, uint32_t i, Target& t) const {
    float4 p = pred[i];
    t.px = p.x; t.py = p.y; t.pz = p.z; t.li = lambda[i];
    t.dx = t.dy = t.dz = t.bx = t.by = t.bz = 0.0f;
    t.nc = 1.0f;
  }
  __device__ void pair(const Uniform& U, Target& t, uint32_t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2 || r2 < 1.0e-12f) return;
    if (EXACT) {
      const float r = sqrtf(r2);
      const float hr = U.h - r;
      const float s = U.spikyGrad * hr * hr / r;
      const float tt = U.h2 - r2;
      const float ratio = U.poly6 * tt * tt * tt * U.sCorrWq;
      const float r2r = ratio * ratio;
      const float sc = -U.sCorrK * r2r * r2r;
      const float f = (t.li + v[0].w + sc) * s;
      t.dx = t.dx + f * rx; t.dy = t.dy + f * ry; t.dz = t.dz + f * rz;
    } else {
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float hr = __fmaf_rn(-r2, ir, U.h);
      const float s = U.spikyGrad * hr * hr * ir;
      const float tt = U.h2 - r2;
      const float ratio = (U.poly6 * U.sCorrWq) * (tt * tt * tt);
      const float r2r = ratio * ratio;
      const float sc = -U.sCorrK * r2r * r2r;
      const float f = (t.li + v[0].w + sc) * s;
      t.dx = __fmaf_rn(f, rx, t.dx);
      t.dy = __fmaf_rn(f, ry, t.dy);
      t.dz = __fmaf_rn(f, rz, t.dz);
    }
    t.nc = t.nc + 1.0f;
  }
  __device__ void bpair(const Uniform& U, Target& t, float4 b) const {
    const float rx = t.px - b.x, ry = t.py - b.y, rz = t.pz - b.z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2 || r2 < 1.0e-12f) return;
    if (EXACT) {
      const float r = sqrtf(r2);
      const float hr = U.h - r;
      const float f = (t.li * b.w) * (U.spikyGrad * hr * hr / r);
      t.bx = t.bx + f * rx; t.by = t.by + f * ry; t.bz = t.bz + f * rz;
    } else {
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float hr = __fmaf_rn(-r2, ir, U.h);
      const float f = (t.li * b.w) * (U.spikyGrad * hr * hr * ir);
      t.bx = __fmaf_rn(f, rx, t.bx);
      t.by = __fmaf_rn(f, ry, t.by);
      t.bz = __fmaf_rn(f, rz, t.bz);
    }
  }
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    const float dx = t.dx * U.volume + t.bx * U.invRho0;
    const float dy = t.dy * U.volume + t.by * U.invRho0;
    const float dz = t.dz * U.volume + t.bz * U.invRho0;
    float scale = U.omega;
    if (U.sorAverage == 1u) scale = U.omega / fmaxf(t.nc, 1.0f);
    float ox = t.px + dx * scale, oy = t.py + dy * scale, oz = t.pz + dz * scale;
    ox = fminf(fmaxf(ox, U.clampMin[0]), U.clampMax[0]);
    oy = fminf(fmaxf(oy, U.clampMin[1]), U.clampMax[1]);
    oz = fminf(fmaxf(oz, U.clampMin[2]), U.clampMax[2]);
    out[i] = make_float4(ox, oy, oz, 1.0f);
    record_drift(driftOut, t.px, t.py, t.pz, ox, oy, oz);
    record_nb(U, i, ox, oy, oz);
  }
};

// --- velFromPos + xsph + normals (wgsl.js:72-123, 250-288) -----------------
// Records: (pred.xyz, max(rho_j, rho0/2)), (vel_j.xyz, 0) with
// vel_j = (pred_j - pos_j) * invDt, the value velFromPos writes.
template <bool EXACT, bool NORMALS>
struct XsphOp : OpBase {
  static constexpr bool FAST = false;
  static constexpr bool BOUNDARY = false;
  static constexpr int NREC = 2;
  const float4* __restrict__ pred;
  float4* __restrict__ pos;
  const float* __restrict__ density;
  float4* __restrict__ corr;
  float4* __restrict__ normal;
  struct Target {
    float px, py, pz, vx, vy, vz;
    float ax, ay, az, nx, ny, nz;
  };
  __device__ void load(const Uniform& U, size_t j, float4* v) const {
    float4 p = pred[j], q = pos[j];
    v[0] = make_float4(p.x, p.y, p.z, fmaxf(density[j], 0.5f * U.rho0));
    v[1] = make_float4((p.x - q.x) * U.invDt, (p.y - q.y) * U.invDt, (p.z - q.z) * U.invDt, 0.0f);
  }
  __device__ void target(const Uniform& U, uint32_t i, Target& t) const {
    float4 p = pred[i], q = pos[i];
    t.px = p.x; t.py = p.y; t.pz = p.z;
    t.vx = (p.x - q.x) * U.invDt; t.vy = (p.y - q.y) * U.invDt; t.vz = (p.z - q.z) * U.invDt;
    t.ax = t.ay = t.az = t.nx = t.ny = t.nz = 0.0f;
  }
  __device__ void pair(const Uniform& U, Target& t, uint32_t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2) return;
    const float den = v[0].w;
    if (EXACT) {
      const float f = U.mass * poly6_w(U, r2) / den;
      t.ax = t.ax + f * (v[1].x - t.vx);
      t.ay = t.ay + f * (v[1].y - t.vy);
      t.az = t.az + f * (v[1].z - t.vz);
      if (NORMALS && r2 >= 1.0e-12f) {
        const float r = sqrtf(r2);
        const float hr = U.h - r;
        const float g = (U.mass / den) * (U.spikyGrad * hr * hr / r);
        t.nx = t.nx + g * rx; t.ny = t.ny + g * ry; t.nz = t.nz + g * rz;
      }
    } else {
      const float inv = __builtin_amdgcn_rcpf(den);
      const float tt = U.h2 - r2;
      const float f = (U.mass * U.poly6) * (tt * tt * tt) * inv;
      t.ax = __fmaf_rn(f, v[1].x - t.vx, t.ax);
      t.ay = __fmaf_rn(f, v[1].y - t.vy, t.ay);
      t.az = __fmaf_rn(f, v[1].z - t.vz, t.az);
      if (NORMALS && r2 >= 1.0e-12f) {
        const float ir = __builtin_amdgcn_rsqf(r2);
        const float hr = __fmaf_rn(-r2, ir, U.h);
        const float g = (U.mass * U.spikyGrad) * inv * hr * hr * ir;
        t.nx = __fmaf_rn(g, rx, t.nx);
        t.ny = __fmaf_rn(g, ry, t.ny);
        t.nz = __fmaf_rn(g, rz, t.nz);
      }
    }
  }
  __device__ void bpair(const Uniform&, Target&, float4) const {}
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    corr[i] = make_float4(U.xsphC * t.ax, U.xsphC * t.ay, U.xsphC * t.az, 0.0f);
    if (NORMALS) normal[i] = make_float4(U.h * t.nx, U.h * t.ny, U.h * t.nz, 0.0f);
  }
};

// --- tension (wgsl.js:290-335) + velFromPos/finalize (72-83, 337-350) -------
// Records: (pred.xyz, rho_j), (normal_j.xyz, 0).
template <bool EXACT>
struct TensionOp : OpBase {
  static constexpr bool FAST = false;
  static constexpr bool BOUNDARY = false;
  static constexpr int NREC = 2;
  const float4* __restrict__ pred;
  float4* __restrict__ pos;
  const float* __restrict__ density;
  const float4* __restrict__ normal;
  float4* __restrict__ corr;
  float4* __restrict__ vel;
  struct Target {
    float px, py, pz, nx, ny, nz, rho;
    float ax, ay, az;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const {
    float4 p = pred[j];
    v[0] = make_float4(p.x, p.y, p.z, density[j]);
    v[1] = normal[j];
  }
  __device__ void target(const Uniform&, uint32_t i, Target& t) const {
    float4 p = pred[i], nn = normal[i];
    t.px = p.x; t.py = p.y; t.pz = p.z;
    t.nx = nn.x; t.ny = nn.y; t.nz = nn.z;
    t.rho = density[i];
    t.ax = t.ay = t.az = 0.0f;
  }
  __device__ void pair(const Uniform& U, Target& t, uint32_t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2 || r2 < 1.0e-12f) return;
    if (EXACT) {
      const float r = sqrtf(r2);
      const float hrr = U.h - r;
      const float k = hrr * hrr * hrr * r * r * r;
      float cw = U.cohes * (2.0f * k - U.cohesTerm);
      if (2.0f * r > U.h) cw = U.cohes * k;
      const float fc = -U.tension * U.mass * cw / r;
      const float fx = fc * rx + (-U.tension * (t.nx - v[1].x));
      const float fy = fc * ry + (-U.tension * (t.ny - v[1].y));
      const float fz = fc * rz + (-U.tension * (t.nz - v[1].z));
      const float K = 2.0f * U.rho0 / (t.rho + v[0].w);
      t.ax = t.ax + K * fx; t.ay = t.ay + K * fy; t.az = t.az + K * fz;
    } else {
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * ir;
      const float hrr = U.h - r;
      const float k = hrr * hrr * hrr * r2 * r;
      const float cw = (2.0f * r > U.h) ? U.cohes * k : U.cohes * (2.0f * k - U.cohesTerm);
      const float fc = (-U.tension * U.mass) * cw * ir;
      const float mg = -U.tension;
      const float fx = __fmaf_rn(fc, rx, mg * (t.nx - v[1].x));
      const float fy = __fmaf_rn(fc, ry, mg * (t.ny - v[1].y));
      const float fz = __fmaf_rn(fc, rz, mg * (t.nz - v[1].z));
      const float K = (2.0f * U.rho0) * __builtin_amdgcn_rcpf(t.rho + v[0].w);
      t.ax = __fmaf_rn(K, fx, t.ax);
      t.ay = __fmaf_rn(K, fy, t.ay);
      t.az = __fmaf_rn(K, fz, t.az);
    }
  }
  __device__ void bpair(const Uniform&, Target&, float4) const {}
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    const float4 c = corr[i];
    const float cx = c.x + U.dt * t.ax, cy = c.y + U.dt * t.ay, cz = c.z + U.dt * t.az;
    const float4 q = pos[i];
    const float vx = (t.px - q.x) * U.invDt, vy = (t.py - q.y) * U.invDt,
                vz = (t.pz - q.z) * U.invDt;
    corr[i] = make_float4(cx, cy, cz, 0.0f);
    vel[i] = make_float4(vx + cx, vy + cy, vz + cz, 0.0f);
    pos[i] = pred[i];
  }
};

// --- fast-math forms ---------------------------------------------------------
// The same sums as the exact kernels, reorganised for throughput: every
// candidate is evaluated without branches (weights vanish outside the
// support through clamps and selects), constant factors (poly6, spiky,
// volume, mass) are applied once per particle, sqrt/division pairs use the
// hardware reciprocal square root, and the self term of the density is
// taken inside the sum (r = 0 contributes exactly W(0)).
constexpr float kFar = 1.0e18f;  // tail padding: r^2 ~ 1e36 stays finite

struct FastBase : OpBase {
  static constexpr bool FAST = true;
  __device__ static float4 sentinel(int k) {
    return k == 0 ? make_float4(kFar, kFar, kFar, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
};

// lambda is also stored in pred.w (unused by every reader of pred), so the
// following delta reads position and lambda of a neighbour in one load.
struct LambdaFast : FastBase {
  static constexpr bool CLEARS_MOVED = true;
  static constexpr bool BOUNDARY = true;
  static constexpr int NREC = 1;
  const float4* pred;  // not restrict: .w is written through predW
  float* __restrict__ density;
  float* __restrict__ lambda;
  float* predW;
  struct Target {
    float px, py, pz;
    float st3, gx, gy, gz, sg2;   // fluid: sum t^3, sum s r, sum s^2 r^2
    float rb, bx, by, bz;         // walls: sum psi t^3, sum psi s r
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const { v[0] = pred[j]; }
  __device__ void target(const Uniform&, uint32_t i, Target& t) const {
    const float4 p = pred[i];
    t.px = p.x; t.py = p.y; t.pz = p.z;
    t.st3 = t.gx = t.gy = t.gz = t.sg2 = t.rb = t.bx = t.by = t.bz = 0.0f;
  }
  __device__ void fpair(const Uniform& U, Target& t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    const float tt = fmaxf(U.h2 - r2, 0.0f);
    t.st3 = __fmaf_rn(tt * tt, tt, t.st3);
    // the gradient only for r^2 > 1e-12 (wgsl.js:512): the self pair,
    // coincident and near-coincident pairs add nothing (a select, so the
    // infinite rsq of r = 0 never reaches the sums)
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float hr = fmaxf(__fmaf_rn(-r2, ir, U.h), 0.0f);
    const float s = r2 > 1.0e-12f ? hr * hr * ir : 0.0f;
    t.gx = __fmaf_rn(s, rx, t.gx);
    t.gy = __fmaf_rn(s, ry, t.gy);
    t.gz = __fmaf_rn(s, rz, t.gz);
    t.sg2 = __fmaf_rn(s * s, r2, t.sg2);
  }
  __device__ void bpair(const Uniform& U, Target& t, float4 b) const {
    const float rx = t.px - b.x, ry = t.py - b.y, rz = t.pz - b.z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2 || r2 < 1.0e-12f) return;
    const float tt = U.h2 - r2;
    t.rb = __fmaf_rn(b.w, tt * tt * tt, t.rb);
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float hr = __fmaf_rn(-r2, ir, U.h);
    const float s = b.w * hr * hr * ir;
    t.bx = __fmaf_rn(s, rx, t.bx);
    t.by = __fmaf_rn(s, ry, t.by);
    t.bz = __fmaf_rn(s, rz, t.bz);
  }
  // the self term W(0) the walk adds when it meets the target itself
  __device__ void self_outside(const Uniform& U, Target& t) const {
    t.st3 = __fmaf_rn(U.h2 * U.h2, U.h2, t.st3);
  }
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    const float rho = U.poly6 * __fmaf_rn(U.mass, t.st3, t.rb);
    density[i] = rho;
    const float C = rho * U.invRho0 - 1.0f;
    float lam = 0.0f;
    if (C > 0.0f) {
      const float cf = U.volume * U.spikyGrad, cb = U.invRho0 * U.spikyGrad;
      const float gx = cf * t.gx + cb * t.bx, gy = cf * t.gy + cb * t.by, gz = cf * t.gz + cb * t.bz;
      const float gg = gx * gx + gy * gy + gz * gz;
      lam = -C / (gg + cf * cf * t.sg2 + U.cfmEps + 1.0e-12f);
    }
    lambda[i] = lam;
    predW[i * 4 + 3] = lam;
  }
};

// SOR = false (sorAverage off, the default): the neighbour count delta-p
// would be divided by is not formed at all.
template <bool SOR>
struct DeltaFastT : FastBase {
  static constexpr bool BOUNDARY = true;
  static constexpr int NREC = 1;
  const float4* __restrict__ pred;
  const float* __restrict__ lambda;
  float4* __restrict__ out;
  struct Target {
    float px, py, pz, li;
    float dx, dy, dz, bx, by, bz, nc;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const { v[0] = pred[j]; }
  __device__ void target(const Uniform&, uint32_t i, Target& t) const {
    const float4 p = pred[i];
    t.px = p.x; t.py = p.y; t.pz = p.z; t.li = p.w;
    t.dx = t.dy = t.dz = t.bx = t.by = t.bz = 0.0f;
    t.nc = 1.0f;
  }
  __device__ void fpair(const Uniform& U, Target& t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    const bool ok = r2 < U.h2 && r2 >= 1.0e-12f;  // [1e-12, h^2)
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float hr = __fmaf_rn(-r2, ir, U.h);
    const float tt = U.h2 - r2;
    // s_corr = -k (W(r) / W(dq))^4 as one constant times tt^12
    const float t2 = tt * tt, t4 = t2 * t2;
    const float sc = U.sCorrFast * (t4 * t4) * t4;
    float f = (t.li + v[0].w + sc) * (hr * hr * ir);
    f = ok ? f : 0.0f;
    t.dx = __fmaf_rn(f, rx, t.dx);
    t.dy = __fmaf_rn(f, ry, t.dy);
    t.dz = __fmaf_rn(f, rz, t.dz);
    if constexpr (SOR) t.nc += ok ? 1.0f : 0.0f;
  }
  __device__ void bpair(const Uniform& U, Target& t, float4 b) const {
    const float rx = t.px - b.x, ry = t.py - b.y, rz = t.pz - b.z;
    const float r2 = r2_of(rx, ry, rz);
    if (r2 >= U.h2 || r2 < 1.0e-12f) return;
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float hr = __fmaf_rn(-r2, ir, U.h);
    const float f = (t.li * b.w) * (hr * hr * ir);
    t.bx = __fmaf_rn(f, rx, t.bx);
    t.by = __fmaf_rn(f, ry, t.by);
    t.bz = __fmaf_rn(f, rz, t.bz);
  }
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    const float cf = U.spikyGrad * U.volume, cb = U.spikyGrad * U.invRho0;
    const float dx = cf * t.dx + cb * t.bx, dy = cf * t.dy + cb * t.by, dz = cf * t.dz + cb * t.bz;
    float scale = U.omega;
    if (SOR && U.sorAverage == 1u) scale = U.omega / fmaxf(t.nc, 1.0f);
    float ox = t.px + dx * scale, oy = t.py + dy * scale, oz = t.pz + dz * scale;
    ox = fminf(fmaxf(ox, U.clampMin[0]), U.clampMax[0]);
    oy = fminf(fmaxf(oy, U.clampMin[1]), U.clampMax[1]);
    oz = fminf(fmaxf(oz, U.clampMin[2]), U.clampMax[2]);
    out[i] = make_float4(ox, oy, oz, 1.0f);
    record_drift(driftOut, t.px, t.py, t.pz, ox, oy, oz);
    record_nb(U, i, ox, oy, oz);
  }
};

// Records: (pred.xyz, -), (vel_j.xyz, q_j) with vel = (pred - pos)/dt and
// q = m / max(rho_j, rho0/2) precomputed once per particle (k_prep_finish).
// normal.w carries rho_i to the tension kernel.
template <bool NORMALS>
struct XsphFast : FastBase {
  static constexpr bool BOUNDARY = false;
  static constexpr int NREC = 2;
  const float4* __restrict__ pred;
  const float4* __restrict__ velq;
  const float* __restrict__ density;
  float4* __restrict__ corr;
  float4* __restrict__ normal;
  struct Target {
    float px, py, pz, vx, vy, vz;
    float ax, ay, az, nx, ny, nz;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const {
    v[0] = pred[j];
    v[1] = velq[j];
  }
  __device__ void target(const Uniform&, uint32_t i, Target& t) const {
    const float4 p = pred[i], w = velq[i];
    t.px = p.x; t.py = p.y; t.pz = p.z;
    t.vx = w.x; t.vy = w.y; t.vz = w.z;
    t.ax = t.ay = t.az = t.nx = t.ny = t.nz = 0.0f;
  }
  __device__ void fpair(const Uniform& U, Target& t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    const float tt = fmaxf(U.h2 - r2, 0.0f);
    const float f = tt * tt * tt * v[1].w;
    t.ax = __fmaf_rn(f, v[1].x - t.vx, t.ax);
    t.ay = __fmaf_rn(f, v[1].y - t.vy, t.ay);
    t.az = __fmaf_rn(f, v[1].z - t.vz, t.az);
    if (NORMALS) {
      // the normal only for r^2 >= 1e-12 (wgsl.js:278), by a select
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float hr = fmaxf(__fmaf_rn(-r2, ir, U.h), 0.0f);
      const float g = r2 >= 1.0e-12f ? hr * hr * ir * v[1].w : 0.0f;
      t.nx = __fmaf_rn(g, rx, t.nx);
      t.ny = __fmaf_rn(g, ry, t.ny);
      t.nz = __fmaf_rn(g, rz, t.nz);
    }
  }
  __device__ void bpair(const Uniform&, Target&, float4) const {}
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    const float cx = U.xsphC * U.poly6;
    corr[i] = make_float4(cx * t.ax, cx * t.ay, cx * t.az, 0.0f);
    if (NORMALS) {
      const float cn = U.h * U.spikyGrad;
      normal[i] = make_float4(cn * t.nx, cn * t.ny, cn * t.nz, density[i]);
    }
  }
};

// Records: (pred.xyz, -), (-gamma normal_j.xyz, rho_j): the normals are
// scaled by -gamma where the record is read, so each pair takes one
// difference of scaled normals.
struct TensionFast : FastBase {
  static constexpr bool BOUNDARY = false;
  static constexpr int NREC = 2;
  const float4* __restrict__ pred;
  float4* __restrict__ pos;
  const float4* __restrict__ velq;
  const float4* __restrict__ normal;
  float4* __restrict__ corr;
  float4* __restrict__ vel;
  struct Target {
    float px, py, pz, nx, ny, nz, rho;
    float ax, ay, az;
  };
  __device__ static float4 sentinel(int k) {
    return k == 0 ? make_float4(kFar, kFar, kFar, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  }
  __device__ void load(const Uniform& U, size_t j, float4* v) const {
    const float4 nn = normal[j];
    const float mg = -U.tension;
    v[0] = pred[j];
    v[1] = make_float4(mg * nn.x, mg * nn.y, mg * nn.z, nn.w);
  }
  __device__ void target(const Uniform& U, uint32_t i, Target& t) const {
    const float4 p = pred[i], nn = normal[i];
    const float mg = -U.tension;
    t.px = p.x; t.py = p.y; t.pz = p.z;
    t.nx = mg * nn.x; t.ny = mg * nn.y; t.nz = mg * nn.z;
    t.rho = nn.w;
    t.ax = t.ay = t.az = 0.0f;
  }
  __device__ void fpair(const Uniform& U, Target& t, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    const bool ok = r2 < U.h2 && r2 >= 1.0e-12f;  // [1e-12, h^2)
    // outside the support ir = 0 keeps every term below finite
    const float ir = ok ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    const float r = r2 * ir;
    const float u = (U.h - r) * r;
    const float k = u * u * u;
    // 2 r > h exactly when r > h / 2 (both sides scale by a power of two)
    const float cw = (r > 0.5f * U.h) ? k : __fmaf_rn(2.0f, k, -U.cohesTerm);
    const float fc = (-U.tension * U.mass * U.cohes) * cw * ir;
    const float fx = __fmaf_rn(fc, rx, t.nx - v[1].x);
    const float fy = __fmaf_rn(fc, ry, t.ny - v[1].y);
    const float fz = __fmaf_rn(fc, rz, t.nz - v[1].z);
    float K = __builtin_amdgcn_rcpf(t.rho + v[1].w);
    K = ok ? K : 0.0f;
    t.ax = __fmaf_rn(K, fx, t.ax);
    t.ay = __fmaf_rn(K, fy, t.ay);
    t.az = __fmaf_rn(K, fz, t.az);
  }
  __device__ void bpair(const Uniform&, Target&, float4) const {}
  __device__ void finish(const Uniform& U, uint32_t i, Target& t) const {
    const float4 c = corr[i];
    const float w = U.dt * (2.0f * U.rho0);
    const float cx = __fmaf_rn(w, t.ax, c.x), cy = __fmaf_rn(w, t.ay, c.y),
                cz = __fmaf_rn(w, t.az, c.z);
    const float4 vq = velq[i];
    corr[i] = make_float4(cx, cy, cz, 0.0f);
    vel[i] = make_float4(vq.x + cx, vq.y + cy, vq.z + cz, 0.0f);
    pos[i] = make_float4(t.px, t.py, t.pz, 1.0f);
  }
};

// velFromPos (wgsl.js:72-83) and the xsph weight factor, once per
// particle: velq = ((pred - pos) / dt, m / max(rho, rho0/2)).
__device__ __forceinline__ void prep_finish_one(const Uniform& U, uint32_t i,
                                                const float4* __restrict__ pred,
                                                const float4* __restrict__ pos,
                                                const float* __restrict__ density,
                                                float4* __restrict__ velq) {
  const float4 p = pred[i], q = pos[i];
  velq[i] = make_float4((p.x - q.x) * U.invDt, (p.y - q.y) * U.invDt, (p.z - q.z) * U.invDt,
                        U.mass / fmaxf(density[i], 0.5f * U.rho0));
}
__global__ void __launch_bounds__(WG)
k_prep_finish(Uniform U, const float4* __restrict__ pred, const float4* __restrict__ pos,
              const float* __restrict__ density, float4* __restrict__ velq) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= U.n) return;
  prep_finish_one(U, i, pred, pos, density, velq);
}
__global__ void __launch_bounds__(WG)
k_prep_finish_batch(BatchTab B, const float4* __restrict__ pred, const float4* __restrict__ pos,
                    const float* __restrict__ density, float4* __restrict__ velq) {
  const Uniform& U = B.U[B.chunkScene[blockIdx.x]];
  const uint32_t local = blockIdx.x * WG - U.vBase + threadIdx.x;
  if (local >= U.n) return;
  prep_finish_one(U, U.pBase + local, pred, pos, density, velq);
}

// --- FP16-offset form (solver iterations) -----------------------------------
// Neighbours are read as 8-byte records (x as float, y and z as half
// offsets from the centre of the cell the particle was binned into at the
// last grid build) instead of 16-byte float4 positions.  The walk goes row
// by row as the FP32 form does: within one (y, z) row of cells every
// candidate shares the row's y and z centres, so the target's quantized y
// and z are taken relative to that row once and each candidate costs one
// fused half-to-float FMA per axis.  The target uses its own quantized
// position (self pair exactly zero, pair distances symmetric) for fluid
// pairs and, in absolute float form, for wall samples: the form is PBF
// evaluated at the quantized positions.  The master positions stay FP32 in
// pred (delta moves them by the correction) and the records are re-encoded
// after every position update (grid build, delta, shape matching).  Lambda
// stays FP32 (delta reads it from the lambda array); sums accumulate in
// float.

// a - o as one fused op on the half operand (gfx950 v_fma_mix_f32); the
// product o * -1 is exact, so this equals the rounded difference.  The -1
// comes from the uniform block so the FMA is not folded into a subtraction.
__device__ __forceinline__ float dh(float a, __half o, float minusOne) {
  return __fmaf_rn(__half2float(o), minusOne, a);
}

// r^2 between a target and an FP16-offset record of a stencil row whose
// y and z cell-centre offsets from the target are ay, az (the pair terms'
// own arithmetic)
template <class T>
__device__ __forceinline__ float r2_16(const Uniform& U, const T& t, float ay, float az,
                                       HRec o) {
  return r2_of(t.px - o.x, dh(ay, o.y, U.minusOne), dh(az, o.z, U.minusOne));
}

// The 27-cell walk of target i over FP16-offset records.  LB (the first
// lambda after a grid build): also marks each row's candidates within reach
// of the target's quantized position and stores its position and cell, as
// the FP32 mask build does.
template <class Op, bool LB>
__device__ __forceinline__ void
direct16_walk(const Uniform& U, const Op& op, uint32_t i, const uint32_t* __restrict__ cellStart,
              const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
              const HRec* __restrict__ ofs, const uint32_t* __restrict__ cellOfSlot,
              const NbArgs* nb) {
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  typename Op::Target t;
  op.target(U, i, t);
  int bx, by, bz;
  cell_xyz(U, cellOfSlot[i], bx, by, bz);
  const HRec ri = ofs[i];
  const float oy = __half2float(ri.y), oz = __half2float(ri.z);
  const Cell c = cell_of(U, t.px, t.py, t.pz);
  const int xlo = max(c.x - 1, 0), xhi = min(c.x + 1, gx - 1);
  const bool nearWall = c.x <= 1 || c.x >= gx - 2 || c.y <= 1 || c.y >= gy - 2 || c.z <= 1 ||
                        c.z >= gz - 2;
  const bool doB = U.nBoundary > 0 && nearWall;
  bool seen = false;  // the target's own record lay in a visited row
  bool unlisted = false;
  for (int dz = -1; dz <= 1; dz++) {
    const int z = c.z + dz;
    if (z < 0 || z >= gz) continue;
    const float az = __fmaf_rn((float)(bz - z), U.h, oz);
    for (int dy = -1; dy <= 1; dy++) {
      const int y = c.y + dy;
      if (y < 0 || y >= gy) continue;
      const float ay = __fmaf_rn((float)(by - y), U.h, oy);
      const uint32_t rb = (uint32_t)((z * gy + y) * gx);
      const uint32_t b = cellStart[rb + xlo], e = cellStart[rb + xhi + 1];
      seen = seen || (i >= b && i < e);
      uint64_t rowMask = 0;
      if constexpr (LB) unlisted = unlisted || (e - b > 64u);
      uint32_t j = b;
      for (; j + 3 < e; j += 4) {
        HRec r[4];
        float l[4];
        // one 64-bit base per group, the others by immediate offsets
#pragma unroll
        for (int u = 0; u < 4; u++) op.load16(ofs, (size_t)j + u, r[u], l[u]);
        if constexpr (LB) {
#pragma unroll
          for (int u = 0; u < 4; u++) rowMask = reach_in(rowMask, r2_16(U, t, ay, az, r[u]),
                                                         nb->reach2);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) op.fpair16(U, t, ay, az, r[u], l[u]);
      }
      for (; j < e; j++) {
        HRec r;
        float l;
        op.load16(ofs, (size_t)j, r, l);
        if constexpr (LB) rowMask = reach_in(rowMask, r2_16(U, t, ay, az, r), nb->reach2);
        op.fpair16(U, t, ay, az, r, l);
      }
      if constexpr (LB)
        nb->mask[(size_t)((dz + 1) * 3 + (dy + 1)) * nb->stride + i] = reach_mask(rowMask, e - b);
      if (doB) {
        const float sy = t.py, sz = t.pz;
        t.py = cell_centre(U, 1, by) + oy;
        t.pz = cell_centre(U, 2, bz) + oz;
        const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
        for (uint32_t k = bb; k < be; k++) op.bpair(U, t, bpos[k]);
        t.py = sy; t.pz = sz;
      }
    }
  }
  if (!seen) op.self_outside(U, t);
  if constexpr (LB)
    nb->ref[i] = make_float4(
        t.px, t.py, t.pz,
        __uint_as_float(cell_index(U, c) | ((unlisted || !seen) ? NB_UNLISTED : 0u)));
  op.finish16(U, i, t);
}

template <class Op, int NT>
__global__ void __launch_bounds__(NT)
k_direct16(Uniform U, Op op, const uint32_t* __restrict__ cellStart,
           const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
           const HRec* __restrict__ ofs, const uint32_t* __restrict__ cellOfSlot) {
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= U.n) return;
  direct16_walk<Op>(U, op, i, cellStart, bcellStart, bpos, ofs, cellOfSlot);
}

// The first FP16 lambda after a grid build: the full walk, building the
// masks (and restarting the travel bound and the queues).
template <class Op, int NT>
__global__ void __launch_bounds__(NT)
k_direct16_build(Uniform U, Op op, NbArgs nb, const uint32_t* __restrict__ cellStart,
                 const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
                 const HRec* __restrict__ ofs, const uint32_t* __restrict__ cellOfSlot) {
  if (blockIdx.x == 0)
    for (uint32_t k = threadIdx.x; k < NB_CTL_WORDS; k += NT) nb.ctl[k] = 0u;
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (i >= U.n) return;
  direct16_walk<Op, true>(U, op, i, cellStart, bcellStart, bpos, ofs, cellOfSlot, &nb);
}

// lambda over FP16-offset records
struct Lambda16 : LambdaFast {
  static constexpr bool USES_L = false;  // no neighbour lambda in the records
  __device__ void load16(const HRec* ofs, size_t j, HRec& r, float& l) const {
    r = ofs[j];
    l = 0.0f;
  }
  __device__ void fpair16(const Uniform& U, Target& t, float ay, float az, HRec o,
                          float) const {
    const float rx = t.px - o.x, ry = dh(ay, o.y, U.minusOne), rz = dh(az, o.z, U.minusOne);
    const float r2 = r2_of(rx, ry, rz);
    const float tt = fmaxf(U.h2 - r2, 0.0f);
    t.st3 = __fmaf_rn(tt * tt, tt, t.st3);
    // the gradient only for r^2 > 1e-12 (wgsl.js:512): the self pair,
    // coincident and near-coincident pairs add nothing (a select, so the
    // infinite rsq of r = 0 never reaches the sums)
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float hr = fmaxf(__fmaf_rn(-r2, ir, U.h), 0.0f);
    const float s = r2 > 1.0e-12f ? hr * hr * ir : 0.0f;
    t.gx = __fmaf_rn(s, rx, t.gx);
    t.gy = __fmaf_rn(s, ry, t.gy);
    t.gz = __fmaf_rn(s, rz, t.gz);
    t.sg2 = __fmaf_rn(s * s, r2, t.sg2);
  }
  __device__ void finish16(const Uniform& U, uint32_t i, Target& t) const {
    LambdaFast::finish(U, i, t);
  }
};

// delta over FP16-offset records (lambda_j from the lambda array); writes
// the moved FP32 position and its re-encoded record
using DeltaFast = DeltaFastT<true>;

template <bool SOR>
struct Delta16T : DeltaFastT<SOR> {
  static constexpr bool USES_L = true;
  using Target = typename DeltaFastT<SOR>::Target;
  using DeltaFastT<SOR>::lambda;
  using DeltaFastT<SOR>::out;
  HRec* ofsOut;
  const uint32_t* cellOfSlot;
  __device__ void load16(const HRec* ofs, size_t j, HRec& r, float& l) const {
    r = ofs[j];
    l = lambda[j];
  }
  __device__ void fpair16(const Uniform& U, Target& t, float ay, float az, HRec o,
                          float lj) const {
    const float rx = t.px - o.x, ry = dh(ay, o.y, U.minusOne), rz = dh(az, o.z, U.minusOne);
    const float r2 = r2_of(rx, ry, rz);
    const bool ok = r2 < U.h2 && r2 >= 1.0e-12f;  // [1e-12, h^2)
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float hr = __fmaf_rn(-r2, ir, U.h);
    const float tt = U.h2 - r2;
    const float t2 = tt * tt, t4 = t2 * t2;
    const float sc = U.sCorrFast * (t4 * t4) * t4;
    float f = (t.li + lj + sc) * (hr * hr * ir);
    f = ok ? f : 0.0f;
    t.dx = __fmaf_rn(f, rx, t.dx);
    t.dy = __fmaf_rn(f, ry, t.dy);
    t.dz = __fmaf_rn(f, rz, t.dz);
    if constexpr (SOR) t.nc += ok ? 1.0f : 0.0f;
  }
  __device__ void finish16(const Uniform& U, uint32_t i, Target& t) const {
    DeltaFastT<SOR>::finish(U, i, t);
    const float4 np = out[i];
    ofsOut[i] = encode_ofs(U, cellOfSlot[i], np.x, np.y, np.z);
  }
};

// A list launch over FP16-offset records: k_listed's walk (the same mask
// test, moved targets skipped, dense failures walked in full, sparse ones
// queued), with each row's 8-byte records (and, for delta, the neighbours'
// lambdas) staged in LDS and the pair terms taken at the target's quantized
// offsets from the row's cell centres, as the full FP16 walk takes them.
template <class Op>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_listed16(Uniform U, Op op, NbArgs nb, const uint32_t* __restrict__ cellStart,
           const uint32_t* __restrict__ bcellStart, const float4* __restrict__ bpos,
           const HRec* __restrict__ ofs, const uint32_t* __restrict__ cellOfSlot) {
  // one LDS slot per candidate: the record (8 bytes) or, with the
  // neighbour's lambda, record and lambda in one 16-byte slot (one read)
  using Slot = typename std::conditional<Op::USES_L, float4, HRec>::type;
  __shared__ Slot SR[NB_STAGE];
  const auto put = [](Slot& d, HRec r, float l) {
    if constexpr (Op::USES_L) {
      uint32_t yz;
      __builtin_memcpy(&yz, &r.y, 4);
      d = make_float4(r.x, __uint_as_float(yz), l, 0.0f);
    } else {
      d = r;
    }
  };
  const auto get = [](const Slot& v, HRec& r, float& l) {
    if constexpr (Op::USES_L) {
      r.x = v.x;
      const uint32_t yz = __float_as_uint(v.y);
      __builtin_memcpy(&r.y, &yz, 4);
      l = v.z;
    } else {
      r = v;
      l = 0.0f;
    }
  };
  if (blockIdx.x == 0) {
    for (uint32_t k = threadIdx.x; k < NB_QSEG; k += 64u)
      nb.ctl[nb_queue_counts(nb.slot ^ 1) + k] = 0u;
    if (Op::CLEARS_MOVED)
      for (uint32_t k = threadIdx.x; k < NB_QSEG; k += 64u)
        nb.ctl[nb_moved_counts(nb.mslot ^ 1) + k] = 0u;
  }
  const uint32_t i = (uint32_t)xcd_remap(blockIdx.x, gridDim.x) * 64u + threadIdx.x;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  const int lane = __lane_id();
  typename Op::Target t;
  Cell c{0, 0, 0};
  bool valid = false, fails = false;
  const auto enqueue = [&](bool q) {
    const uint64_t qm = __ballot(q);
    if (!qm) return;
    const int leader = __builtin_ctzll(qm);
    uint32_t base = 0;
    const uint32_t seg = blockIdx.x % NB_QSEG;
    if (lane == leader)
      base = atomicAdd(&nb.ctl[nb_queue_counts(nb.slot) + seg], (uint32_t)__popcll(qm));
    base = (uint32_t)__shfl((int)base, leader, 64);
    const uint32_t at = base + (uint32_t)__popcll(qm & ((1ull << lane) - 1ull));
    if (q && at < nb.segCap) nb.queue[nb.slot][seg * nb.segCap + at] = i;
  };
  if (i < U.n) {
    op.target(U, i, t);
    c = cell_of(U, t.px, t.py, t.pz);
    const float4 ref = nb.ref[i];
    const uint32_t built = __float_as_uint(ref.w);
    const bool stays = (built & ~NB_UNLISTED) == cell_index(U, c);
    const float travel = sqrtf(r2_of(t.px - ref.x, t.py - ref.y, t.pz - ref.z));
    const float most = __uint_as_float(*(volatile const uint32_t*)nb.ctl);
    valid = stays && built == cell_index(U, c) && travel + most <= nb.lim;
    fails = stays && !valid;
  }
  const bool dense = __popcll(__ballot(fails)) >= 16;
  enqueue(fails && !dense);
  if (__ballot(valid)) {
    int bx = 0, by = 0, bz = 0;
    float oy = 0.0f, oz = 0.0f;
    if (valid) {
      cell_xyz(U, cellOfSlot[i], bx, by, bz);
      const HRec ri = ofs[i];
      oy = __half2float(ri.y);
      oz = __half2float(ri.z);
    }
    const int xlo = max(c.x - 1, 0);
    const int xhi = min(c.x + 1, gx - 1);
    const bool nearWall = c.x <= 1 || c.x >= gx - 2 || c.y <= 1 || c.y >= gy - 2 ||
                          c.z <= 1 || c.z >= gz - 2;
    const bool doB = U.nBoundary > 0 && nearWall && valid;
    const bool mono = one_grid_row(valid, c);
    for (int dz = -1; dz <= 1; dz++) {
      for (int dy = -1; dy <= 1; dy++) {
        const int z = c.z + dz, y = c.y + dy;
        const bool rowOk = valid && z >= 0 && z < gz && y >= 0 && y < gy;
        const uint32_t rb = rowOk ? (uint32_t)((z * gy + y) * gx) : 0u;
        const float az = __fmaf_rn((float)(bz - z), U.h, oz);
        const float ay = __fmaf_rn((float)(by - y), U.h, oy);
        uint32_t b = 0, e = 0;
        uint64_t m64 = 0;
        if (rowOk) {
          b = cellStart[rb + xlo];
          e = cellStart[rb + xhi + 1];
          m64 = nb.mask[(size_t)((dz + 1) * 3 + (dy + 1)) * nb.stride + i];
        }
        const bool has = m64 != 0ull;
        uint32_t ub, ue;
        row_union(has, b, e, mono, ub, ue);
        if (ub < ue && ue - ub <= NB_STAGE) {
          const uint32_t span = ue - ub;
          for (uint32_t q = (uint32_t)lane; q < span; q += 64u) {
            HRec r;
            float l;
            op.load16(ofs, (size_t)ub + q, r, l);
            put(SR[q], r, l);
          }
          __syncthreads();
          for (int half = 0; half < 2; half++) {
            uint32_t m = (uint32_t)(m64 >> (32 * half));
            if (half && !__ballot(m != 0u)) break;
            const uint32_t base = b + 32u * (uint32_t)half - ub;
            while (m) {
              const uint32_t k = base + (uint32_t)__builtin_ctz(m);
              m &= m - 1u;
              HRec r;
              float l;
              get(SR[k], r, l);
              op.fpair16(U, t, ay, az, r, l);
            }
          }
          __syncthreads();
        } else if (has) {
          for (int half = 0; half < 2; half++) {
            uint32_t m = (uint32_t)(m64 >> (32 * half));
            while (m) {
              const uint32_t j = b + 32u * (uint32_t)half + (uint32_t)__builtin_ctz(m);
              m &= m - 1u;
              HRec r;
              float l;
              op.load16(ofs, (size_t)j, r, l);
              op.fpair16(U, t, ay, az, r, l);
            }
          }
        }
        if (doB && rowOk) {
          const float sy = t.py, sz = t.pz;
          t.py = cell_centre(U, 1, by) + oy;
          t.pz = cell_centre(U, 2, bz) + oz;
          const uint32_t bb = bcellStart[rb + xlo], be = bcellStart[rb + xhi + 1];
          for (uint32_t k = bb; k < be; k++) op.bpair(U, t, bpos[k]);
          t.py = sy; t.pz = sz;
        }
      }
    }
    if (valid) op.finish16(U, i, t);
  }
  if (dense && fails)
    direct16_walk<Op>(U, op, i, cellStart, bcellStart, bpos, ofs, cellOfSlot);
}

// --- neighbour census: |N(i)| and index checksums -------------------------
struct CensusOp : OpBase {
  static constexpr bool FAST = false;
  static constexpr bool BOUNDARY = false;
  static constexpr int NREC = 1;
  const float4* __restrict__ pred;
  uint32_t* __restrict__ cnt;
  uint32_t* __restrict__ s1;
  uint32_t* __restrict__ s2;
  struct Target {
    float px, py, pz;
    uint32_t i, c, a, b;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const { v[0] = pred[j]; }
  __device__ void target(const Uniform&, uint32_t i, Target& t) const {
    float4 p = pred[i];
    t.px = p.x; t.py = p.y; t.pz = p.z; t.i = i; t.c = t.a = t.b = 0u;
  }
  __device__ void pair(const Uniform& U, Target& t, uint32_t j, const float4* v) const {
    const float rx = t.px - v[0].x, ry = t.py - v[0].y, rz = t.pz - v[0].z;
    const float r2 = r2_of(rx, ry, rz);
    if (j == t.i || r2 >= U.h2) return;
    t.c += 1u; t.a += j; t.b += j * j;
  }
  __device__ void bpair(const Uniform&, Target&, float4) const {}
  __device__ void finish(const Uniform&, uint32_t i, Target& t) const {
    cnt[i] = t.c; s1[i] = t.a; s2[i] = t.b;
  }
};

// --- anisotropic kernel (aniso_wgsl.js:3-159) -----------------------------
// Symmetric 3x3 eigensolve by cyclic Jacobi rotations (aniso_wgsl.js:30-55):
// a <- r^T a r, v <- v r for the pairs (0,1), (0,2), (1,2).  Matrices are
// row-major here (a[row][col]); the WGSL's a[q][p] (column q, row p) is
// a[p][q].
template <class R>
__device__ void jacobi3(R a[3][3], R v[3][3], int sweeps, R tiny) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) v[i][j] = (i == j) ? (R)1 : (R)0;
  for (int sweep = 0; sweep < sweeps; sweep++) {
    for (int p = 0; p < 2; p++) {
      for (int q = p + 1; q < 3; q++) {
        const R apq = a[p][q];
        if (fabs(apq) < tiny) continue;
        const R theta = (a[q][q] - a[p][p]) / ((R)2 * apq);
        const R sg = theta > (R)0 ? (R)1 : (theta < (R)0 ? (R)-1 : (R)0);
        R t = sg / (fabs(theta) + sqrt(theta * theta + (R)1));
        if (theta == (R)0) t = (R)1;
        const R c = (R)1 / sqrt(t * t + (R)1);
        const R s = t * c;
        R r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        r[p][p] = c; r[q][q] = c;
        r[p][q] = s; r[q][p] = -s;
        R m[3][3], n[3][3];
        for (int i = 0; i < 3; i++)  // m = r^T a
          for (int j = 0; j < 3; j++) m[i][j] = r[0][i] * a[0][j] + r[1][i] * a[1][j] + r[2][i] * a[2][j];
        for (int i = 0; i < 3; i++)  // a = m r
          for (int j = 0; j < 3; j++) a[i][j] = m[i][0] * r[0][j] + m[i][1] * r[1][j] + m[i][2] * r[2][j];
        for (int i = 0; i < 3; i++)  // v = v r
          for (int j = 0; j < 3; j++) n[i][j] = v[i][0] * r[0][j] + v[i][1] * r[1][j] + v[i][2] * r[2][j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) v[i][j] = n[i][j];
      }
    }
  }
}

template <class R>
__device__ R det3(const R m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// G from eigenvectors v (columns) and eigenvalues sig (aniso_wgsl.js:133-153).
template <class R>
__device__ void aniso_g(const AnisoArgs& A, const R v[3][3], const R sig[3], int n, R G[3][3]) {
  const R s1 = fmax(sig[0], fmax(sig[1], sig[2]));
  R ext[3];
  for (int k = 0; k < 3; k++) ext[k] = (R)A.ks * fmax(sig[k], s1 / (R)A.kr);
  const R mx = fmax(ext[0], fmax(ext[1], ext[2]));
  const R dv = fmax((R)1, mx / (R)A.stretch);
  R inv[3];
  for (int k = 0; k < 3; k++) {
    ext[k] = fmax(ext[k] / dv, (R)A.minExtent);
    inv[k] = (R)1 / ((R)A.support * ext[k]);
  }
  const R e0 = (R)A.neps * (R)0.4, e1 = (R)A.neps * (R)1.6;
  R w = fmin(fmax(((R)n - e0) / (e1 - e0), (R)0), (R)1);
  w = w * w * ((R)3 - (R)2 * w);
  const R iso = (R)A.kn / (R)A.support;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const R an = v[i][0] * inv[0] * v[j][0] + v[i][1] * inv[1] * v[j][1] + v[i][2] * inv[2] * v[j][2];
      const R is = (i == j) ? iso : (R)0;
      G[i][j] = is * ((R)1 - w) + an * w;
    }
}

template <bool EXACT>
struct AnisoOp : OpBase {
  static constexpr bool FAST = false;
  static constexpr bool BOUNDARY = false;
  static constexpr int NREC = 1;
  const float4* __restrict__ pos;
  float4* __restrict__ smoothPos;
  float4* __restrict__ g;
  uint32_t* __restrict__ fallbacks;
  AnisoArgs a;
  struct Target {
    float px, py, pz, rx, ry, rz;
    float sx, sy, sz, m00, m10, m20, m11, m21, m22, wsum;
    int n;
  };
  __device__ void load(const Uniform&, size_t j, float4* v) const { v[0] = pos[j]; }
  __device__ void target(const Uniform& U, uint32_t i, Target& t) const {
    const float4 p = pos[i];
    t.px = p.x; t.py = p.y; t.pz = p.z;
    t.rx = p.x - U.boxMin[0]; t.ry = p.y - U.boxMin[1]; t.rz = p.z - U.boxMin[2];
    t.sx = t.sy = t.sz = t.m00 = t.m10 = t.m20 = t.m11 = t.m21 = t.m22 = t.wsum = 0.0f;
    t.n = 0;
  }
  // floor((x - boxMin) * invSupport), clamped (aniso_wgsl.js:72-73)
  __device__ Cell cell(const Uniform& U, const Target& t) const {
    float f[3] = {floorf(t.rx * a.invSupport), floorf(t.ry * a.invSupport),
                  floorf(t.rz * a.invSupport)};
    int c[3];
    for (int k = 0; k < 3; k++)
      c[k] = (int)fminf(fmaxf(f[k], 0.0f), (float)(U.gridDim[k] - 1));
    return Cell{c[0], c[1], c[2]};
  }
  // rows whose slab lies beyond the radius (aniso_wgsl.js:88-97)
  __device__ bool skip_row(const Uniform&, const Target& t, int z, int y) const {
    const float dzm = fmaxf(fmaxf((float)z * a.support - t.rz, t.rz - (float)(z + 1) * a.support), 0.0f);
    const float dym = fmaxf(fmaxf((float)y * a.support - t.ry, t.ry - (float)(y + 1) * a.support), 0.0f);
    const float r2max = a.radius * a.radius;
    const float dz2 = dzm * dzm;
    return dz2 >= r2max || dz2 + dym * dym >= r2max;
  }
  __device__ void pair(const Uniform&, Target& t, uint32_t, const float4* v) const {
    const float dx = v[0].x - t.px, dy = v[0].y - t.py, dz = v[0].z - t.pz;
    const float r2 = r2_of(dx, dy, dz);
    const float r2max = a.radius * a.radius;
    if (r2 >= r2max) return;
    const float invR = 1.0f / a.radius;
    const float tt = (EXACT ? sqrtf(r2) : __builtin_amdgcn_sqrtf(r2)) * invR;
    const float w = 1.0f - tt * tt * tt;
    t.sx = t.sx + dx * w; t.sy = t.sy + dy * w; t.sz = t.sz + dz * w;
    const float wx = dx * w, wy = dy * w, wz = dz * w;
    t.m00 = t.m00 + dx * wx; t.m10 = t.m10 + dy * wx; t.m20 = t.m20 + dz * wx;
    t.m11 = t.m11 + dy * wy; t.m21 = t.m21 + dz * wy;
    t.m22 = t.m22 + dz * wz;
    t.wsum = t.wsum + w;
    t.n += 1;
  }
  __device__ void bpair(const Uniform&, Target&, float4) const {}
  __device__ void finish(const Uniform&, uint32_t i, Target& t) const {
    if (t.wsum <= 0.0f) {
      smoothPos[i] = make_float4(t.px, t.py, t.pz, 0.0f);
      const float iso = a.kn / a.support;
      g[i * 2] = make_float4(iso, 0.0f, 0.0f, iso);
      g[i * 2 + 1] = make_float4(0.0f, iso, iso * iso * iso, 0.0f);
      return;
    }
    const float bx = t.sx / t.wsum, by = t.sy / t.wsum, bz = t.sz / t.wsum;
    smoothPos[i] = make_float4(t.px + bx * a.lambda, t.py + by * a.lambda, t.pz + bz * a.lambda, 0.0f);
    float C[3][3];
    C[0][0] = t.m00 / t.wsum - bx * bx;
    C[1][0] = t.m10 / t.wsum - by * bx;
    C[2][0] = t.m20 / t.wsum - bz * bx;
    C[1][1] = t.m11 / t.wsum - by * by;
    C[2][1] = t.m21 / t.wsum - bz * by;
    C[2][2] = t.m22 / t.wsum - bz * bz;
    C[0][1] = C[1][0]; C[0][2] = C[2][0]; C[1][2] = C[2][1];
    float af[3][3], vf[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) af[r][c] = C[r][c];
    jacobi3<float>(af, vf, 6, 1.0e-12f);
    const float sf[3] = {af[0][0], af[1][1], af[2][2]};
    const float scale = fmaxf(fabsf(sf[0]), fmaxf(fabsf(sf[1]), fabsf(sf[2])));
    const float gap = fminf(fabsf(sf[0] - sf[1]), fminf(fabsf(sf[0] - sf[2]), fabsf(sf[1] - sf[2])));
    float G[3][3];
    if (a.fp64Fallback && scale > 0.0f && gap < a.gapTol * scale) {
      // near-degenerate spectrum: solve again in FP64 to convergence
      double ad[3][3], vd[3][3], Gd[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) ad[r][c] = (double)C[r][c];
      jacobi3<double>(ad, vd, 12, 1.0e-300);
      const double sd[3] = {ad[0][0], ad[1][1], ad[2][2]};
      aniso_g<double>(a, vd, sd, t.n, Gd);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) G[r][c] = (float)Gd[r][c];
      atomicAdd(fallbacks, 1u);
    } else {
      aniso_g<float>(a, vf, sf, t.n, G);
    }
    g[i * 2] = make_float4(G[0][0], G[0][1], G[0][2], G[1][1]);
    g[i * 2 + 1] = make_float4(G[1][2], G[2][2], det3<float>(G), (float)t.n);
  }
};

// finalize (wgsl.js:337-350) with velFromPos (72-83) folded in, for the
// no-tension path.
__device__ __forceinline__ void finalize_one(const Uniform& U, uint32_t i,
                                             const float4* __restrict__ pred,
                                             float4* __restrict__ pos,
                                             const float4* __restrict__ corr,
                                             float4* __restrict__ vel) {
  const float4 p = pred[i], q = pos[i], c = corr[i];
  const float vx = (p.x - q.x) * U.invDt, vy = (p.y - q.y) * U.invDt, vz = (p.z - q.z) * U.invDt;
  vel[i] = make_float4(vx + c.x, vy + c.y, vz + c.z, 0.0f);
  pos[i] = p;
}
__global__ void __launch_bounds__(WG)
k_finalize(Uniform U, const float4* __restrict__ pred, float4* __restrict__ pos,
           const float4* __restrict__ corr, float4* __restrict__ vel) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= U.n) return;
  finalize_one(U, i, pred, pos, corr, vel);
}
__global__ void __launch_bounds__(WG)
k_finalize_batch(BatchTab B, const float4* __restrict__ pred, float4* __restrict__ pos,
                 const float4* __restrict__ corr, float4* __restrict__ vel) {
  const Uniform& U = B.U[B.chunkScene[blockIdx.x]];
  const uint32_t local = blockIdx.x * WG - U.vBase + threadIdx.x;
  if (local >= U.n) return;
  finalize_one(U, U.pBase + local, pred, pos, corr, vel);
}

// ---------------------------------------------------------------------------
// Rigid bodies, FP64.  One workgroup per body; the body's particles are the
// static id list bodyPids[bodyFirst[b] .. bodyFirst[b+1]), located through
// slotOf.

constexpr int BODY_UNROLL = 8;

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K>
__device__ void block_sum_d(double (&v)[K], double (*sh)[K]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wave_sum_d(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) sh[wave][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    double t = 0;
    for (int w = 0; w < WG / 64; w++) t += sh[w][k];
    v[k] = t;
  }
}

// bodyCentre + bodySeed + bodyRefLoad before the solve (wgsl.js:628-668,
// 174-185; sim.js:496-516): bodyCentre becomes the centroid of pred.
__global__ void __launch_bounds__(WG)
k_body_pre(const float4* __restrict__ pred, const uint32_t* __restrict__ slotOf,
           const uint32_t* __restrict__ bodyPids, const uint32_t* __restrict__ bodyFirst,
           double* __restrict__ centreD, float4* __restrict__ centreF) {
  __shared__ double sh[WG / 64][3];
  const int b = blockIdx.x;
  const uint32_t f = bodyFirst[b], l = bodyFirst[b + 1];
  const double cx = centreD[b * 4], cy = centreD[b * 4 + 1], cz = centreD[b * 4 + 2];
  double v[3] = {0, 0, 0};
  // BODY_UNROLL particles per thread and trip with their loads in flight
  // together (each sums its particles in the same order as one at a time)
  for (uint32_t k0 = f + threadIdx.x; k0 < l; k0 += WG * BODY_UNROLL) {
    uint32_t q[BODY_UNROLL];
    float4 p[BODY_UNROLL];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) q[u] = bodyPids[min(k0 + (uint32_t)u * WG, l - 1u)];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) q[u] = slotOf[q[u]];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) p[u] = pred[q[u]];
    // a slot past the body adds +0 (a sum that starts at +0 keeps its bits)
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) {
      const bool in = k0 + (uint32_t)u * WG < l;
      v[0] += in ? (double)p[u].x - cx : 0.0;
      v[1] += in ? (double)p[u].y - cy : 0.0;
      v[2] += in ? (double)p[u].z - cz : 0.0;
    }
  }
  block_sum_d<3>(v, sh);
  if (threadIdx.x == 0 && l > f) {
    const double cnt = (double)(l - f);
    const double nx = cx + v[0] / cnt, ny = cy + v[1] / cnt, nz = cz + v[2] / cnt;
    centreD[b * 4] = nx; centreD[b * 4 + 1] = ny; centreD[b * 4 + 2] = nz;
    centreF[b] = make_float4((float)nx, (float)ny, (float)nz, 0.0f);
  }
}

// One shape-matching round (wgsl.js:628-783, sim.js:544-563): centroid and
// A = sum (p - centroid) q^T in one pass, using
// sum (p - c0 - m) q^T = sum (p - c0) q^T - m (sum q)^T with m the mean of
// p - c0; centre moves by m * invRatio (buoyancy, wgsl.js:715-721); R is
// the orthogonal polar factor of A / cbrt|det A| by at most 8 Newton steps
// R <- (R + R^-T)/2 (wgsl.js:727-742); pred = centre + R q in float.
__global__ void __launch_bounds__(WG)
k_body_solve(float4* __restrict__ pred, const uint32_t* __restrict__ slotOf,
             const uint32_t* __restrict__ bodyPids, const uint32_t* __restrict__ bodyFirst,
             const float4* __restrict__ restOfPid, const float* __restrict__ invRatio,
             double* __restrict__ centreD, double* __restrict__ rotD,
             float4* __restrict__ centreF, float4* __restrict__ rotF, HRec* __restrict__ ofs,
             const uint32_t* __restrict__ cellOfSlot, Uniform U, uint32_t* __restrict__ driftOut,
             const float4* __restrict__ nbRef, uint32_t* __restrict__ nbTravel,
             uint32_t* __restrict__ movedQ, uint32_t* movedCnt, uint32_t segCap) {
  __shared__ double sh[WG / 64][15];
  __shared__ float sc[12];
  const int b = blockIdx.x;
  const uint32_t f = bodyFirst[b], l = bodyFirst[b + 1];
  if (l <= f) return;
  const double cx = centreD[b * 4], cy = centreD[b * 4 + 1], cz = centreD[b * 4 + 2];
  double v[15];
#pragma unroll
  for (int k = 0; k < 15; k++) v[k] = 0;
  for (uint32_t k0 = f + threadIdx.x; k0 < l; k0 += WG * BODY_UNROLL) {
    uint32_t q[BODY_UNROLL], sl[BODY_UNROLL];
    float4 p[BODY_UNROLL], r[BODY_UNROLL];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) q[u] = bodyPids[min(k0 + (uint32_t)u * WG, l - 1u)];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) {
      sl[u] = slotOf[q[u]];
      r[u] = restOfPid[q[u]];
    }
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) p[u] = pred[sl[u]];
    // a slot past the body adds +-0 to every sum (which keeps its bits)
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) {
      const bool in = k0 + (uint32_t)u * WG < l;
      const double px = in ? (double)p[u].x - cx : 0.0, py = in ? (double)p[u].y - cy : 0.0,
                   pz = in ? (double)p[u].z - cz : 0.0;
      const double qx = in ? (double)r[u].x : 0.0, qy = in ? (double)r[u].y : 0.0,
                   qz = in ? (double)r[u].z : 0.0;
      v[0] += px; v[1] += py; v[2] += pz;
      v[3] += qx; v[4] += qy; v[5] += qz;
      v[6] += px * qx; v[7] += px * qy; v[8] += px * qz;
      v[9] += py * qx; v[10] += py * qy; v[11] += py * qz;
      v[12] += pz * qx; v[13] += pz * qy; v[14] += pz * qz;
    }
  }
  block_sum_d<15>(v, sh);
  if (threadIdx.x == 0) {
    const double cnt = (double)(l - f);
    const double m[3] = {v[0] / cnt, v[1] / cnt, v[2] / cnt};
    double A[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) A[r][c] = v[6 + r * 3 + c] - m[r] * v[3 + c];
    double ir = (double)invRatio[b];
    double dm[3] = {m[0], m[1], m[2]};
    if (ir > 0.0)
      for (int k = 0; k < 3; k++) dm[k] *= ir;
    const double nc[3] = {cx + dm[0], cy + dm[1], cz + dm[2]};
    double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                 A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                 A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (fabs(det) < 1.0e-12) {
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) A[r][c] = (r == c) ? 1.0 : 0.0;
      det = 1.0;
    }
    const double s = cbrt(fabs(det));
    if (s > 1.0e-6)
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) A[r][c] *= 1.0 / s;
    double R[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) R[r][c] = A[r][c];
    for (int it = 0; it < 8; it++) {
      // R^-T = cofactor(R) / det(R): column c of R^-T is the cross product
      // of the other two columns of R.
      double C[3][3];
      for (int c = 0; c < 3; c++) {
        const int a = (c + 1) % 3, e = (c + 2) % 3;
        C[0][c] = R[1][a] * R[2][e] - R[2][a] * R[1][e];
        C[1][c] = R[2][a] * R[0][e] - R[0][a] * R[2][e];
        C[2][c] = R[0][a] * R[1][e] - R[1][a] * R[0][e];
      }
      const double d = R[0][0] * C[0][0] + R[1][0] * C[1][0] + R[2][0] * C[2][0];
      const double invd = fabs(d) > 1.0e-20 ? 1.0 / d : 0.0;
      double diff = 0;
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          const double nx = (R[r][c] + C[r][c] * invd) * 0.5;
          diff += (nx - R[r][c]) * (nx - R[r][c]);
          R[r][c] = nx;
        }
      if (diff < 1.0e-12) break;
    }
    centreD[b * 4] = nc[0]; centreD[b * 4 + 1] = nc[1]; centreD[b * 4 + 2] = nc[2];
    for (int k = 0; k < 9; k++) rotD[b * 9 + k] = R[k / 3][k % 3];
    centreF[b] = make_float4((float)nc[0], (float)nc[1], (float)nc[2], 0.0f);
    for (int r = 0; r < 3; r++)
      rotF[b * 3 + r] = make_float4((float)R[r][0], (float)R[r][1], (float)R[r][2], 0.0f);
    sc[0] = (float)nc[0]; sc[1] = (float)nc[1]; sc[2] = (float)nc[2];
    for (int k = 0; k < 9; k++) sc[3 + k] = (float)R[k / 3][k % 3];
  }
  __syncthreads();
  // bodyProject (wgsl.js:763-783)
  const float ox = sc[0], oy = sc[1], oz = sc[2];
  float dmax = 0.0f, tmax = 0.0f;
  for (uint32_t k0 = f + threadIdx.x; k0 < l; k0 += WG * BODY_UNROLL) {
    uint32_t q[BODY_UNROLL], sl[BODY_UNROLL];
    float4 r[BODY_UNROLL], op[BODY_UNROLL], ref[BODY_UNROLL];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) q[u] = bodyPids[min(k0 + (uint32_t)u * WG, l - 1u)];
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) {
      sl[u] = slotOf[q[u]];
      r[u] = restOfPid[q[u]];
    }
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) {
      op[u] = pred[sl[u]];
      if (nbTravel) ref[u] = nbRef[sl[u]];
    }
#pragma unroll
    for (int u = 0; u < BODY_UNROLL; u++) {
      const bool in = k0 + (uint32_t)u * WG < l;
      const float px = sc[3] * r[u].x + sc[4] * r[u].y + sc[5] * r[u].z;
      const float py = sc[6] * r[u].x + sc[7] * r[u].y + sc[8] * r[u].z;
      const float pz = sc[9] * r[u].x + sc[10] * r[u].y + sc[11] * r[u].z;
      const float4 np = make_float4(ox + px, oy + py, oz + pz, 1.0f);
      if (in)
        dmax = fmaxf(dmax, fmaxf(fabsf(np.x - op[u].x),
                                 fmaxf(fabsf(np.y - op[u].y), fabsf(np.z - op[u].z))));
      if (nbTravel) {
        // a NaN distance must survive the maximum (fmaxf would drop it)
        const float d = sqrtf(r2_of(np.x - ref[u].x, np.y - ref[u].y, np.z - ref[u].z));
        if (in) tmax = (d > tmax || d != d) ? d : tmax;
        // queued if the projection takes it out of its build cell (the
        // delta before queued it if it was out already)
        const uint32_t built = __float_as_uint(ref[u].w) & ~NB_UNLISTED;
        const bool was = cell_index(U, cell_of(U, op[u].x, op[u].y, op[u].z)) != built;
        const bool now = cell_index(U, cell_of(U, np.x, np.y, np.z)) != built;
        enqueue_moved(movedQ, movedCnt, segCap, in && now && !was, sl[u]);
      }
      if (in) {
        pred[sl[u]] = np;
        if (ofs) ofs[sl[u]] = encode_ofs(U, cellOfSlot[sl[u]], np.x, np.y, np.z);
      }
    }
  }
  if (driftOut) {
    // largest displacement of the body's particles into the drift slot
    __shared__ uint32_t wmax[WG / 64];
    const uint32_t m = __ockl_wfred_max_u32(__float_as_uint(dmax));
    if (__lane_id() == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t v = wmax[0];
      for (int k = 1; k < WG / 64; k++) v = max(v, wmax[k]);
      atomicMax(driftOut, v);
    }
  }
  if (nbTravel) {
    // largest distance of the body's particles from their build positions
    __shared__ uint32_t tw[WG / 64];
    const uint32_t m = __ockl_wfred_max_u32(__float_as_uint(tmax));
    if (__lane_id() == 0) tw[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t v = tw[0];
      for (int k = 1; k < WG / 64; k++) v = max(v, tw[k]);
      atomicMax(nbTravel, v);
    }
  }
}

// bodyDrag (wgsl.js:187-214) over the held body's particles.
__global__ void __launch_bounds__(WG)
k_body_drag(DragArgs D, float4* __restrict__ vel, const uint32_t* __restrict__ slotOf,
            const uint32_t* __restrict__ bodyPids, const uint32_t* __restrict__ bodyFirst,
            const float4* __restrict__ centreF) {
  const uint32_t b = D.held - 1u;
  const uint32_t f = bodyFirst[b], l = bodyFirst[b + 1];
  const uint32_t k = f + blockIdx.x * WG + threadIdx.x;
  if (k >= l) return;
  const uint32_t i = slotOf[bodyPids[k]];
  const float4 c = centreF[b];
  float wx = (D.goal[0] - c.x) * D.rate, wy = (D.goal[1] - c.y) * D.rate,
        wz = (D.goal[2] - c.z) * D.rate;
  const float sp = sqrtf(wx * wx + wy * wy + wz * wz);
  if (sp > D.limit) {
    const float s = D.limit / sp;
    wx *= s; wy *= s; wz *= s;
  }
  const float4 v = vel[i];
  vel[i] = make_float4(v.x + (wx - v.x) * D.blend, v.y + (wy - v.y) * D.blend,
                       v.z + (wz - v.z) * D.blend, v.w);
}

// impulse (wgsl.js:216-248)
__global__ void __launch_bounds__(WG)
k_impulse(RayArgs R, const float4* __restrict__ pos, float4* __restrict__ vel) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= R.n) return;
  const float4 p = pos[i];
  const float tx = p.x - R.origin[0], ty = p.y - R.origin[1], tz = p.z - R.origin[2];
  const float along = tx * R.dir[0] + ty * R.dir[1] + tz * R.dir[2];
  if (along <= 0.0f) return;
  const float ax = tx - along * R.dir[0], ay = ty - along * R.dir[1], az = tz - along * R.dir[2];
  const float d2 = ax * ax + ay * ay + az * az;
  if (d2 >= R.radius * R.radius) return;
  const float x = 1.0f - sqrtf(d2) / R.radius;
  const float4 v = vel[i];
  const float old = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  float nx = v.x + (x * x) * R.impulse[0], ny = v.y + (x * x) * R.impulse[1],
        nz = v.z + (x * x) * R.impulse[2];
  const float ns = sqrtf(nx * nx + ny * ny + nz * nz);
  const float allowed = fmaxf(old, R.speedLimit);
  if (ns > allowed) {
    const float s = allowed / ns;
    nx *= s; ny *= s; nz *= s;
  }
  vel[i] = make_float4(nx, ny, nz, 0.0f);
}

// resize (wgsl.js:125-157)
__global__ void __launch_bounds__(WG)
k_resize(ResizeArgs B, float4* __restrict__ pos) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= B.n) return;
  float4 q = pos[i];
  float p[3] = {q.x, q.y, q.z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (B.newSize[a] < B.oldSize[a]) {
      const float t = fminf(fmaxf((p[a] - B.oldMin[a]) / B.oldSize[a], 0.0f), 1.0f);
      p[a] = B.newMin[a] + t * B.newSize[a];
    }
    p[a] = fminf(fmaxf(p[a], B.clampMin[a]), B.clampMax[a]);
  }
  pos[i] = make_float4(p[0], p[1], p[2], q.w);
}

// stats (wgsl.js:785-832), reduced exactly: per-block FP64 partials, then
// one block folds the partials in a fixed order.
__global__ void __launch_bounds__(WG)
k_stats_partial(Uniform U, const float* __restrict__ density, const float4* __restrict__ vel,
                double* __restrict__ part) {
  __shared__ double sh[WG / 64][4];
  const uint32_t i0 = blockIdx.x * (WG * 8) + threadIdx.x;
  double sr = 0, mr = 0, mv = 0, ke = 0;
  for (int k = 0; k < 8; k++) {
    const uint32_t i = i0 + k * WG;
    if (i < U.n) {
      const float rho = density[i] * U.invRho0;
      const float4 v = vel[i];
      const float v2 = v.x * v.x + v.y * v.y + v.z * v.z;
      sr += rho;
      mr = fmax(mr, (double)rho);
      mv = fmax(mv, (double)sqrtf(v2));
      ke += (double)(0.5f * U.mass * v2);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    ke += __shfl_xor(ke, off, 64);
    mr = fmax(mr, __shfl_xor(mr, off, 64));
    mv = fmax(mv, __shfl_xor(mv, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sh[wave][0] = sr; sh[wave][1] = mr; sh[wave][2] = mv; sh[wave][3] = ke; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0, d = 0;
    for (int w = 0; w < WG / 64; w++) {
      a += sh[w][0]; b = fmax(b, sh[w][1]); c = fmax(c, sh[w][2]); d += sh[w][3];
    }
    part[blockIdx.x * 4 + 0] = a; part[blockIdx.x * 4 + 1] = b;
    part[blockIdx.x * 4 + 2] = c; part[blockIdx.x * 4 + 3] = d;
  }
}

__global__ void __launch_bounds__(WG)
k_stats_final(const double* __restrict__ part, int nb, uint32_t n, double* __restrict__ out) {
  __shared__ double sh[WG / 64][4];
  double sr = 0, mr = 0, mv = 0, ke = 0;
  for (int b = threadIdx.x; b < nb; b += WG) {
    sr += part[b * 4]; mr = fmax(mr, part[b * 4 + 1]);
    mv = fmax(mv, part[b * 4 + 2]); ke += part[b * 4 + 3];
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    ke += __shfl_xor(ke, off, 64);
    mr = fmax(mr, __shfl_xor(mr, off, 64));
    mv = fmax(mv, __shfl_xor(mv, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sh[wave][0] = sr; sh[wave][1] = mr; sh[wave][2] = mv; sh[wave][3] = ke; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0, d = 0;
    for (int w = 0; w < WG / 64; w++) {
      a += sh[w][0]; b = fmax(b, sh[w][1]); c = fmax(c, sh[w][2]); d += sh[w][3];
    }
    out[0] = a / (double)(n > 0 ? n : 1u);
    out[1] = b; out[2] = c; out[3] = d;
  }
}

// Batched stats: one 256-chunk per workgroup into a partial, then one
// workgroup per scene folds its chunks' partials in chunk order.
__global__ void __launch_bounds__(WG)
k_stats_partial_batch(BatchTab B, const float* __restrict__ density,
                      const float4* __restrict__ vel, double* __restrict__ part) {
  __shared__ double sh[WG / 64][4];
  const Uniform& U = B.U[B.chunkScene[blockIdx.x]];
  const uint32_t local = blockIdx.x * WG - U.vBase + threadIdx.x;
  double sr = 0, mr = 0, mv = 0, ke = 0;
  if (local < U.n) {
    const uint32_t i = U.pBase + local;
    const float rho = density[i] * U.invRho0;
    const float4 v = vel[i];
    const float v2 = v.x * v.x + v.y * v.y + v.z * v.z;
    sr = rho;
    mr = (double)rho;
    mv = (double)sqrtf(v2);
    ke = (double)(0.5f * U.mass * v2);
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    ke += __shfl_xor(ke, off, 64);
    mr = fmax(mr, __shfl_xor(mr, off, 64));
    mv = fmax(mv, __shfl_xor(mv, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sh[wave][0] = sr; sh[wave][1] = mr; sh[wave][2] = mv; sh[wave][3] = ke; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0, d = 0;
    for (int w = 0; w < WG / 64; w++) {
      a += sh[w][0]; b = fmax(b, sh[w][1]); c = fmax(c, sh[w][2]); d += sh[w][3];
    }
    part[blockIdx.x * 4] = a; part[blockIdx.x * 4 + 1] = b;
    part[blockIdx.x * 4 + 2] = c; part[blockIdx.x * 4 + 3] = d;
  }
}

__global__ void __launch_bounds__(WG)
k_stats_final_batch(BatchTab B, const double* __restrict__ part, double* __restrict__ out) {
  const Uniform& U = B.U[blockIdx.x];
  const uint32_t c0 = U.vBase >> 8, nc = (U.n + 255u) >> 8;
  __shared__ double sh[WG / 64][4];
  double sr = 0, mr = 0, mv = 0, ke = 0;
  for (uint32_t c = threadIdx.x; c < nc; c += WG) {
    const double* q = part + (size_t)(c0 + c) * 4;
    sr += q[0]; mr = fmax(mr, q[1]); mv = fmax(mv, q[2]); ke += q[3];
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    ke += __shfl_xor(ke, off, 64);
    mr = fmax(mr, __shfl_xor(mr, off, 64));
    mv = fmax(mv, __shfl_xor(mv, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sh[wave][0] = sr; sh[wave][1] = mr; sh[wave][2] = mv; sh[wave][3] = ke; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0, d = 0;
    for (int w = 0; w < WG / 64; w++) {
      a += sh[w][0]; b = fmax(b, sh[w][1]); c = fmax(c, sh[w][2]); d += sh[w][3];
    }
    double* o = out + (size_t)blockIdx.x * 4;
    o[0] = a / (double)(U.n > 0 ? U.n : 1u);
    o[1] = b; o[2] = c; o[3] = d;
  }
}

// One record per thread: on MI355X this streams at ~6.2 TB/s on 1 GiB, where
// a grid-stride loop over 8192 blocks reaches ~4.6 TB/s (one load in flight
// per thread for several rounds) and unrolled or non-temporal forms no more.
__global__ void __launch_bounds__(WG) k_copy4(const float4* __restrict__ a, float4* __restrict__ b,
                                               size_t n) {
  const size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + WG - 1) / WG); }

template <class Op, int NT, int CR = 1, int PH = 0>
hipError_t run_direct(const Uniform& U, const Arrays& A, const Op& op, hipStream_t s) {
  if (U.n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_direct<Op, NT, CR, PH>), dim3((U.n + NT - 1) / NT), dim3(NT), 0, s, U, op,
                     A.cellStart, A.bcellStart, A.bpos);
  return hipGetLastError();
}

template <class Op, int NT, int PH = 0>
hipError_t run_direct_batch(const BatchTab& B, const Arrays& A, const Op& op, hipStream_t s) {
  if (B.chunks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_direct_batch<Op, NT, PH>), dim3(B.chunks * (256u / NT)), dim3(NT), 0, s, B,
                     op, A.cellStart, A.bcellStart, A.bpos);
  return hipGetLastError();
}

template <class Op>
hipError_t run_staged(const Uniform& U, const Arrays& A, const Op& op, hipStream_t s) {
  if (U.n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_staged<Op>), dim3((U.n + 63) / 64), dim3(64), 0, s, U, op, A.cellStart,
                     A.bcellStart, A.bpos);
  return hipGetLastError();
}

template <class Op>
hipError_t run_staged_batch(const BatchTab& B, const Arrays& A, const Op& op, hipStream_t s) {
  if (B.chunks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_staged_batch<Op>), dim3(B.chunks * 4u), dim3(64), 0, s, B, op,
                     A.cellStart, A.bcellStart, A.bpos);
  return hipGetLastError();
}

template <class Op, int TX, int NT, int MAXREC, int CR = 1>
hipError_t run_tile(const Uniform& U, const Arrays& A, const Op& op, hipStream_t s) {
  const int gx = U.gridDim[0];
  const int tpr = (gx + TX - 1) / TX;
  const long nt = (long)tpr * U.gridDim[1] * U.gridDim[2];
  if (nt <= 0 || U.n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_tile<Op, TX, NT, MAXREC, CR>), dim3((unsigned)nt), dim3(NT), 0, s, U, op,
                     A.cellStart, A.bcellStart, A.bpos, tpr);
  return hipGetLastError();
}

// Tile shapes (cells per tile along x, threads, LDS record capacity).  A
// tile whose halo holds more records than the capacity walks global memory
// for all its particles; capacity 0 makes every tile do so.  p4a_tune times
// the variants on the live scene and keeps the fastest.
template <class Op>
hipError_t run_tile1(const Uniform& U, const Arrays& A, const Op& op, const TileCfg& t,
                     hipStream_t s, const BatchTab* B = nullptr) {
  if (B) return t.variant == 0   ? run_direct_batch<Op, 256>(*B, A, op, s)
                : t.variant == 1 ? run_direct_batch<Op, 128>(*B, A, op, s)
                : t.variant == 5 ? run_direct_batch<Op, 64, 1>(*B, A, op, s)
                : t.variant == 6 ? run_direct_batch<Op, 128, 1>(*B, A, op, s)
                : t.variant == 7 ? run_direct_batch<Op, 64, 2>(*B, A, op, s)
                : t.variant == 8 ? run_direct_batch<Op, 128, 2>(*B, A, op, s)
                : t.variant == 9 ? run_staged_batch<Op>(*B, A, op, s)
                                 : run_direct_batch<Op, 64>(*B, A, op, s);
  switch (t.variant) {
    case 1: return run_direct<Op, 128>(U, A, op, s);
    case 2: return run_direct<Op, 64>(U, A, op, s);
    case 3: return run_tile<Op, 32, 256, 3072>(U, A, op, s);
    case 4: return run_tile<Op, 16, 128, 2048>(U, A, op, s);
    case 5: return run_direct<Op, 64, 1, 1>(U, A, op, s);
    case 6: return run_direct<Op, 128, 1, 1>(U, A, op, s);
    case 7: return run_direct<Op, 64, 1, 2>(U, A, op, s);
    case 8: return run_direct<Op, 128, 1, 2>(U, A, op, s);
    case 9: return run_staged<Op>(U, A, op, s);
    default: return run_direct<Op, 256>(U, A, op, s);
  }
}
template <class Op>
hipError_t run_tile2(const Uniform& U, const Arrays& A, const Op& op, const TileCfg& t,
                     hipStream_t s, const BatchTab* B = nullptr) {
  if (B) return t.variant == 0   ? run_direct_batch<Op, 256>(*B, A, op, s)
                : t.variant == 1 ? run_direct_batch<Op, 128>(*B, A, op, s)
                : t.variant == 5 ? run_direct_batch<Op, 64, 1>(*B, A, op, s)
                : t.variant == 6 ? run_direct_batch<Op, 128, 1>(*B, A, op, s)
                : t.variant == 7 ? run_direct_batch<Op, 64, 2>(*B, A, op, s)
                : t.variant == 8 ? run_direct_batch<Op, 128, 2>(*B, A, op, s)
                : t.variant == 9 ? run_staged_batch<Op>(*B, A, op, s)
                                 : run_direct_batch<Op, 64>(*B, A, op, s);
  switch (t.variant) {
    case 1: return run_direct<Op, 128>(U, A, op, s);
    case 2: return run_direct<Op, 64>(U, A, op, s);
    case 3: return run_tile<Op, 16, 128, 1536>(U, A, op, s);
    case 4: return run_tile<Op, 16, 256, 1536>(U, A, op, s);
    case 5: return run_direct<Op, 64, 1, 1>(U, A, op, s);
    case 6: return run_direct<Op, 128, 1, 1>(U, A, op, s);
    case 7: return run_direct<Op, 64, 1, 2>(U, A, op, s);
    case 8: return run_direct<Op, 128, 1, 2>(U, A, op, s);
    case 9: return run_staged<Op>(U, A, op, s);
    default: return run_direct<Op, 256>(U, A, op, s);
  }
}

// --- density field (mesh_wgsl.js:23-69, mesh.js:96-128) ---------------------
// Per-slot record of the isotropic field: live position and
// w = (m / max(rho, rho0/2)) * poly6, the first two operations of the
// reference's term vol * poly6 * t * t * t; w = 0 for a body particle when
// bodies are skipped (the reference's `continue` adds nothing, so adding
// +0 to the non-negative sum leaves the same bits).
__global__ void __launch_bounds__(WG)
k_field_records(FieldArgs F, uint32_t n, const float4* __restrict__ pos,
                const float* __restrict__ density, const uint32_t* __restrict__ pid,
                const uint32_t* __restrict__ bodyOfPid, float4* __restrict__ rec) {
  const uint32_t j = blockIdx.x * WG + threadIdx.x;
  if (j >= n) return;
  const float4 p = pos[j];
  const float vol = F.mass / fmaxf(density[j], 0.5f * F.rho0);
  float w = vol * F.poly6;
  if (F.skipBodies && bodyOfPid[pid[j]] != 0u) w = 0.0f;
  rec[j] = make_float4(p.x, p.y, p.z, w);
}

// One thread per voxel vertex p = origin + voxel * v: the 27 cells around
// p's cell of the simulation grid (stale, as the reference reads it),
// acc += ((w * t) * t) * t for r^2 < h^2 in the cell-sorted order.
__global__ void __launch_bounds__(WG)
k_field_iso(Uniform U, FieldArgs F, const uint32_t* __restrict__ cellStart,
            const float4* __restrict__ rec, float* __restrict__ field) {
  const uint32_t idx = blockIdx.x * WG + threadIdx.x;
  if (idx >= F.nVerts) return;
  const int vx = (int)(idx % (uint32_t)F.vertDim[0]);
  const int vy = (int)((idx / (uint32_t)F.vertDim[0]) % (uint32_t)F.vertDim[1]);
  const int vz = (int)(idx / ((uint32_t)F.vertDim[0] * (uint32_t)F.vertDim[1]));
  const float px = F.origin[0] + F.voxel * (float)vx;
  const float py = F.origin[1] + F.voxel * (float)vy;
  const float pz = F.origin[2] + F.voxel * (float)vz;
  const int gx = U.gridDim[0], gy = U.gridDim[1], gz = U.gridDim[2];
  const int cx = min(max((int)floorf((px - U.boxMin[0]) * F.invSupport), 0), gx - 1);
  const int cy = min(max((int)floorf((py - U.boxMin[1]) * F.invSupport), 0), gy - 1);
  const int cz = min(max((int)floorf((pz - U.boxMin[2]) * F.invSupport), 0), gz - 1);
  const int x0 = max(cx - 1, 0), x1 = min(cx + 1, gx - 1);
  float acc = 0.0f;
  for (int dz = -1; dz <= 1; dz++) {
    const int z = cz + dz;
    if (z < 0 || z >= gz) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int y = cy + dy;
      if (y < 0 || y >= gy) continue;
      const uint32_t rb = (uint32_t)((z * gy + y) * gx);
      const uint32_t b = cellStart[rb + x0], e = cellStart[rb + x1 + 1];
      for (uint32_t j = b; j < e; j++) {
        const float4 q = rec[j];
        const float rx = px - q.x, ry = py - q.y, rz = pz - q.z;
        const float r2 = r2_of(rx, ry, rz);
        const float t = F.support2 - r2;
        const float term = ((q.w * t) * t) * t;
        acc = (r2 < F.support2) ? acc + term : acc;
      }
    }
  }
  field[idx] = acc;
}

// Anisotropic field (Yu & Turk): phi(x) = sum_j V_j sigma det(G_j)
// (1 - |G_j (x - xbar_j)|^2)^3 inside |G_j (x - xbar_j)| < 1, with
// xbar = smoothPos, G from the anisotropy pass, sigma = 315/(64 pi) (with
// G = I/h and xbar = x this is the isotropic poly6 field above).  Per
// particle: M = G G (so |G r|^2 = r^T M r), w = V sigma det G, and the
// bounding box of the ellipsoid, half-extents sqrt((M^-1)_kk) (FP64).
// Records: A = (xbar, w), B = (M00, M11, M22, M01), and in A2 (M02, M12,
// half-extent x, y, z) packed as B2 below.
__global__ void __launch_bounds__(WG)
k_field_aniso_prep(FieldArgs F, uint32_t n, const float4* __restrict__ smoothPos,
                   const float4* __restrict__ g, const float* __restrict__ density,
                   const uint32_t* __restrict__ pid, const uint32_t* __restrict__ bodyOfPid,
                   float4* __restrict__ recA, float4* __restrict__ recB) {
  const uint32_t j = blockIdx.x * WG + threadIdx.x;
  if (j >= n) return;
  const float4 g0 = g[2 * j], g1 = g[2 * j + 1];
  // G symmetric: (G00, G01, G02, G11), (G12, G22, det G, count)
  const float a = g0.x, b = g0.y, c = g0.z, d = g0.w, e = g1.x, f = g1.y;
  const float m00 = a * a + b * b + c * c;
  const float m11 = b * b + d * d + e * e;
  const float m22 = c * c + e * e + f * f;
  const float m01 = a * b + b * d + c * e;
  const float m02 = a * c + b * e + c * f;
  const float m12 = b * c + d * e + e * f;
  const float vol = F.mass / fmaxf(density[j], 0.5f * F.rho0);
  float w = vol * (1.5666814710608448f * g1.z);  // 315/(64 pi) * det G
  if (F.skipBodies && bodyOfPid[pid[j]] != 0u) w = 0.0f;
  // inverse of M in FP64 for the bounding box (cofactors / det)
  const double A00 = m00, A11 = m11, A22 = m22, A01 = m01, A02 = m02, A12 = m12;
  const double c00 = A11 * A22 - A12 * A12, c11 = A00 * A22 - A02 * A02,
               c22 = A00 * A11 - A01 * A01;
  const double det = A00 * c00 - A01 * (A01 * A22 - A12 * A02) + A02 * (A01 * A12 - A11 * A02);
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
  if (det > 0.0 && w > 0.0f) {
    // 1e-3 relative margin: the voxel test below decides membership
    ex = (float)(sqrt(fmax(c00 / det, 0.0)) * 1.001 + 1e-6);
    ey = (float)(sqrt(fmax(c11 / det, 0.0)) * 1.001 + 1e-6);
    ez = (float)(sqrt(fmax(c22 / det, 0.0)) * 1.001 + 1e-6);
  } else {
    w = 0.0f;
  }
  const float4 sp = smoothPos[j];
  recA[j] = make_float4(sp.x, sp.y, sp.z, w);
  recB[3 * j] = make_float4(m00, m11, m22, m01);
  recB[3 * j + 1] = make_float4(m02, m12, 0.0f, 0.0f);
  recB[3 * j + 2] = make_float4(ex, ey, ez, 0.0f);
}

// Vertex index box of particle j's ellipsoid (empty: lo > hi), and whether
// the particle contributes at all.  A non-finite centre or extent would make
// the box span the whole field, so such a particle contributes nothing.
struct SplatBox {
  int lo[3], hi[3];
};
__device__ __forceinline__ bool splat_box(const FieldArgs& F, const float4& A, const float4& E,
                                          SplatBox& bx) {
  if (!(A.w > 0.0f) || !isfinite(A.x) || !isfinite(A.y) || !isfinite(A.z) || !isfinite(E.x) ||
      !isfinite(E.y) || !isfinite(E.z))
    return false;
  const float inv = 1.0f / F.voxel;
  const float c[3] = {A.x, A.y, A.z}, ext[3] = {E.x, E.y, E.z};
  for (int k = 0; k < 3; k++) {
    const float a = (c[k] - ext[k] - F.origin[k]) * inv, b = (c[k] + ext[k] - F.origin[k]) * inv;
    bx.lo[k] = max((int)ceilf(fmaxf(a, -1.0f)), 0);
    bx.hi[k] = min((int)floorf(fminf(b, (float)F.vertDim[k])), F.vertDim[k] - 1);
  }
  return bx.lo[0] <= bx.hi[0] && bx.lo[1] <= bx.hi[1] && bx.lo[2] <= bx.hi[2];
}

// The term of vertex (x, y, z) for particle record (A, B0, B1), as a
// multiple of 2^-40, or -1 outside the ellipsoid.
__device__ __forceinline__ long long splat_term(const FieldArgs& F, const float4& A,
                                                const float4& B0, const float4& B1, int x,
                                                float ry, float rz) {
  const float rx = (F.origin[0] + F.voxel * (float)x) - A.x;
  // s^2 = M00 x^2 + M11 y^2 + M22 z^2 + 2 (M01 xy + M02 xz + M12 yz)
  const float s2 = B0.x * (rx * rx) + B0.y * (ry * ry) + B0.z * (rz * rz) +
                   2.0f * (B0.w * (rx * ry) + B1.x * (rx * rz) + B1.y * (ry * rz));
  if (!(s2 < 1.0f)) return -1;
  const float u = 1.0f - s2;
  const float term = ((A.w * u) * u) * u;
  return __double2ll_rn((double)term * FIELD_FIXED_SCALE);
}

// One workgroup per NT consecutive particles of the cell-sorted order (a
// spatially compact chunk: 64 particles are about 8 cells of one row).  The terms are first summed in an LDS tile
// covering the union of the chunk's vertex boxes, and each non-zero tile
// entry is then added to the field with one 64-bit integer atomic: the
// global atomics drop by the overlap of the chunk's ellipsoids.  A chunk
// whose union does not fit the tile adds its terms to the field directly.
// Integer sums, so the result does not depend on the order of either kind
// of atomic.
template <int NT, int CAP>
__global__ void __launch_bounds__(NT)
k_field_splat(FieldArgs F, uint32_t n, const float4* __restrict__ recA,
              const float4* __restrict__ recB, unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long tile[CAP];
  __shared__ int red[NT / 64][6];
  __shared__ int box[6];
  const uint32_t j = blockIdx.x * NT + threadIdx.x;
  float4 A = make_float4(0.0f, 0.0f, 0.0f, 0.0f), B0 = A, B1 = A, E = A;
  SplatBox bx;
  bool act = false;
  if (j < n) {
    A = recA[j];
    B0 = recB[3 * j];
    B1 = recB[3 * j + 1];
    E = recB[3 * j + 2];
    act = splat_box(F, A, E, bx);
  }
  // union of the boxes (empty boxes contribute nothing)
  int v[6];
  for (int k = 0; k < 3; k++) {
    v[k] = act ? bx.lo[k] : INT_MAX;
    v[3 + k] = act ? -bx.hi[k] : INT_MAX;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int m = __ockl_wfred_min_i32(v[k]);
    if (lane == 0) red[wave][k] = m;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int m = red[0][threadIdx.x];
    for (int w = 1; w < NT / 64; w++) m = min(m, red[w][threadIdx.x]);
    box[threadIdx.x] = m;
  }
  __syncthreads();
  const int ox = box[0], oy = box[1], oz = box[2];
  const long long sx = (long long)-box[3] - ox + 1, sy = (long long)-box[4] - oy + 1,
                  sz = (long long)-box[5] - oz + 1;
  const bool any = box[0] != INT_MAX;
  const bool useTile = any && sx * sy * sz <= (long long)CAP;
  const int tx = (int)sx, ty = (int)sy;
  if (useTile) {
    const int cnt = (int)(sx * sy * sz);
    for (int k = threadIdx.x; k < cnt; k += NT) tile[k] = 0ull;
  }
  __syncthreads();
  if (act) {
    for (int z = bx.lo[2]; z <= bx.hi[2]; z++) {
      const float rz = (F.origin[2] + F.voxel * (float)z) - A.z;
      for (int y = bx.lo[1]; y <= bx.hi[1]; y++) {
        const float ry = (F.origin[1] + F.voxel * (float)y) - A.y;
        const uint32_t rowBase = ((uint32_t)z * (uint32_t)F.vertDim[1] + (uint32_t)y) *
                                 (uint32_t)F.vertDim[0];
        const int tileRow = ((z - oz) * ty + (y - oy)) * tx - ox;
        for (int x = bx.lo[0]; x <= bx.hi[0]; x++) {
          const long long q = splat_term(F, A, B0, B1, x, ry, rz);
          if (q < 0) continue;
          if (useTile)
            atomicAdd(&tile[tileRow + x], (unsigned long long)q);
          else
            atomicAdd(&acc[rowBase + (uint32_t)x], (unsigned long long)q);
        }
      }
    }
  }
  if (!useTile) return;
  __syncthreads();
  const int cnt = (int)(sx * sy * sz);
  for (int k = threadIdx.x; k < cnt; k += NT) {
    const unsigned long long t = tile[k];
    if (t == 0ull) continue;
    const int x = ox + k % tx, y = oy + (k / tx) % ty, z = oz + k / (tx * ty);
    atomicAdd(&acc[((uint32_t)z * (uint32_t)F.vertDim[1] + (uint32_t)y) *
                       (uint32_t)F.vertDim[0] + (uint32_t)x], t);
  }
}

__global__ void __launch_bounds__(WG)
k_field_fixed_to_float(uint32_t nv, unsigned long long* __restrict__ acc,
                       float* __restrict__ field) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= nv) return;
  field[i] = (float)((double)(long long)acc[i] * (1.0 / FIELD_FIXED_SCALE));
}

// Per-slot phase in the live order (the reference's body buffer, x = body
// id, 0 for fluid; scatterPhase keeps it in slot order, wgsl.js:460-475).
__global__ void __launch_bounds__(WG)
k_export_phase(uint32_t n, const uint32_t* __restrict__ pid, const uint32_t* __restrict__ bodyOfPid,
               uint4* __restrict__ out) {
  const uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= n) return;
  out[i] = make_uint4(bodyOfPid[pid[i]], 0u, 0u, 0u);
}

// bodyPackWGSL (solid_wgsl.js:3-26): per body 6 float4 = (centre, shape id),
// (half extents, 0), the columns of R, (colour, 0).
__global__ void __launch_bounds__(64)
k_body_pack(uint32_t count, const float4* __restrict__ centre, const float4* __restrict__ rot,
            const float4* __restrict__ stat, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= count) return;
  const float4 s = stat[2 * i], c = stat[2 * i + 1];
  const float4 r0 = rot[3 * i], r1 = rot[3 * i + 1], r2 = rot[3 * i + 2];
  const float4 ce = centre[i];
  float4* o = out + 6 * (size_t)i;
  o[0] = make_float4(ce.x, ce.y, ce.z, s.w);
  o[1] = make_float4(s.x, s.y, s.z, 0.0f);
  o[2] = make_float4(r0.x, r1.x, r2.x, 0.0f);
  o[3] = make_float4(r0.y, r1.y, r2.y, 0.0f);
  o[4] = make_float4(r0.z, r1.z, r2.z, 0.0f);
  o[5] = make_float4(c.x, c.y, c.z, 0.0f);
}

__global__ void __launch_bounds__(WG)
k_fill_u64(unsigned long long* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

int stats_blocks(uint32_t n) { return (int)((n + WG * 8 - 1) / (WG * 8)); }

int tile_variants(int) { return 10; }

const char* tile_variant_name(int group, int v) {
  static const char* g1[] = {"direct/t256", "direct/t128", "direct/t64", "tile/tx32/t256/lds3072",
                             "tile/tx16/t128/lds2048", "direct2p/t64", "direct2p/t128",
                             "direct2p4/t64", "direct2p4/t128", "staged/t64"};
  static const char* g2[] = {"direct/t256", "direct/t128", "direct/t64", "tile/tx16/t128/lds1536",
                             "tile/tx16/t256/lds1536", "direct2p/t64", "direct2p/t128",
                             "direct2p4/t64", "direct2p4/t128", "staged/t64"};
  if (v < 0 || v >= 10) return "?";
  return group == 1 ? g1[v] : g2[v];
}

hipError_t launch_predict_count(const Uniform& U, const Arrays& A, int par, hipStream_t s,
                                const BatchTab* B) {
  if (B) {
    if (B->chunks > 0)
      hipLaunchKernelGGL(k_predict_count_batch, dim3(B->chunks), dim3(WG), 0, s, *B, A.pos[par],
                         A.vel[par], A.pred[par], A.key, A.cellCount, A.drift);
    return hipGetLastError();
  }
  if (U.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predict_count, dim3(blocks(U.n)), dim3(WG), 0, s, U, A.pos[par],
                     A.vel[par], A.pred[par], A.key, A.cellCount, A.drift);
  return hipGetLastError();
}

hipError_t launch_scan(const Uniform& U, const Arrays& A, hipStream_t s) {
  // nCells+1 entries are written (cellStart[nCells] = total), so the blocks
  // cover indices 0..nCells inclusive.
  const unsigned nb = (unsigned)((U.nCells + SCAN_CH) / SCAN_CH);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(WG), 0, s, A.cellCount, U.nCells, A.blockSum);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(WG), 0, s, A.cellCount, U.nCells, A.blockSum,
                     A.cellStart);
  return hipGetLastError();
}

int scan_blocks(uint32_t n) { return (int)((n + SCAN_CH) / SCAN_CH); }

hipError_t launch_exclusive_scan(uint32_t* count, uint32_t n, uint32_t* blockSum, uint32_t* out,
                                 hipStream_t s) {
  const unsigned nb = (unsigned)scan_blocks(n);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(WG), 0, s, count, n, blockSum);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(WG), 0, s, count, n, blockSum, out);
  return hipGetLastError();
}

hipError_t launch_scatter(const Uniform& U, const Arrays& A, int par, bool moveVel,
                          bool fp16, hipStream_t s) {
  if (U.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_tmp, dim3(blocks(U.n)), dim3(WG), 0, s, U.n, A.key, A.cellStart,
                     A.cellCount, A.tmp, A.fatList);
  MoveArgs M{A.pos[par], A.pred[par], A.pid[par], A.pos[par ^ 1], A.pred[par ^ 1],
             A.pid[par ^ 1], A.order, A.slot, A.slotOf,
             moveVel ? A.vel[par] : nullptr, moveVel ? A.vel[par ^ 1] : nullptr,
             fp16 ? A.ofs[0] : nullptr, A.cellOfSlot};
  // at most n / (FAT_CELL + 1) segments are longer than FAT_CELL
  const uint32_t fatCap = U.n / (FAT_CELL + 1);
  hipLaunchKernelGGL(k_sort_move, dim3(blocks(U.n)), dim3(WG), 0, s, U.n, A.tmp, A.key,
                     A.cellStart, A.cellCount, A.fatList, fatCap, M, U);
  // Such segments exist only when the state is far from a fluid at rest
  // density (a collapsed or blown-up scene); the workgroups leave at once
  // when none is listed.
  if (fatCap > 0) {
    const unsigned g = fatCap < 64u ? fatCap : 64u;
    hipLaunchKernelGGL(k_sort_fat, dim3(g), dim3(WG), 0, s, U.n, A.tmp, A.cellStart,
                       A.fatList, fatCap, M, U);
  }
  return hipGetLastError();
}

template <class Op>
Op drifted(Op op, const Drift& d) {
  op.driftOut = d.out;
  return op;
}

// Workgroup size of the mask build and the list launches: the direct walk's
// of the chosen shape, 64 for the LDS-tile shapes.
int nb_threads(const TileCfg& t) { return t.variant == 0 ? 256 : t.variant == 1 ? 128 : 64; }

template <class Op, int NT>
hipError_t run_nb(const Uniform& U, const Arrays& A, const Op& op, const NbLaunch& nb,
                  hipStream_t s) {
  if (nb.mode == Nb::BUILD) {
    hipLaunchKernelGGL((k_staged_build<Op>), dim3((U.n + 63) / 64), dim3(64), 0, s, U, op,
                       nb.args, A.cellStart, A.bcellStart, A.bpos);
    return hipGetLastError();
  }
  // the moved targets' full walks run on the side stream beside the list
  // launch (in order before it when there is no side stream)
  hipStream_t fs = nb.side ? nb.side : s;
  if (nb.side) {
    hipError_t e = hipEventRecord(nb.fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(nb.side, nb.fork, 0);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_fallback<Op, NT, true>), dim3(NB_QSEG * NB_FALLBACK_PARTS), dim3(NT), 0,
                     fs, U, op, nb.args, A.cellStart, A.bcellStart, A.bpos);
  hipLaunchKernelGGL((k_listed<Op>), dim3((U.n + 63) / 64), dim3(64), 0, s, U, op, nb.args,
                     A.cellStart, A.bcellStart, A.bpos);
  hipLaunchKernelGGL((k_fallback<Op, NT, false>), dim3(NB_QSEG * NB_FALLBACK_PARTS), dim3(NT), 0,
                     s, U, op, nb.args, A.cellStart, A.bcellStart, A.bpos);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && nb.side) {
    e = hipEventRecord(nb.join, nb.side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, nb.join, 0);
  }
  return e;
}

// The launch of a solver op under a neighbour-mask mode (OFF: the tile
// shape's walk).
template <class Op>
hipError_t run_solver(const Uniform& U, const Arrays& A, Op op, const TileCfg& t,
                      const NbLaunch& nb, hipStream_t s, const BatchTab* B) {
  if (B || nb.mode == Nb::OFF || U.n == 0) return run_tile1(U, A, op, t, s, B);
  if (nb.mode == Nb::USE) {
    // delta records travels and queues the targets it moves out of their
    // build cells for the next iteration
    op.nbRef = nb.args.ref;
    op.nbTravel = nb.args.ctl;
    op.nbMovedQ = nb.args.moved[nb.args.mslot ^ 1];
    op.nbMovedCnt = nb.args.ctl + nb_moved_counts(nb.args.mslot ^ 1);
    op.nbSegCap = nb.args.segCap;
  }
  switch (nb_threads(t)) {
    case 256: return run_nb<Op, 256>(U, A, op, nb, s);
    case 128: return run_nb<Op, 128>(U, A, op, nb, s);
    default: return run_nb<Op, 64>(U, A, op, nb, s);
  }
}

hipError_t launch_lambda(const Uniform& U, const Arrays& A, const float4* pred, bool exact,
                         const TileCfg& t, const Drift& d, hipStream_t s, const BatchTab* B,
                         const NbLaunch& nb) {
  if (exact)
    return run_solver(U, A, drifted(LambdaOp<true>{{}, pred, A.density, A.lambda}, d), t, nb,
                      s, B);
  return run_solver(U, A, drifted(LambdaFast{{}, pred, A.density, A.lambda, (float*)pred}, d), t,
                    nb, s, B);
}

hipError_t launch_delta(const Uniform& U, const Arrays& A, const float4* pred, float4* out,
                        bool exact, const TileCfg& t, const Drift& d, hipStream_t s,
                        const BatchTab* B, const NbLaunch& nb) {
  if (exact)
    return run_solver(U, A, drifted(DeltaOp<true>{{}, pred, A.lambda, out}, d), t, nb, s, B);
  if (U.sorAverage == 1u)
    return run_solver(U, A, drifted(DeltaFastT<true>{{}, pred, A.lambda, out}, d), t, nb, s, B);
  return run_solver(U, A, drifted(DeltaFastT<false>{{}, pred, A.lambda, out}, d), t, nb, s, B);
}

template <class Op, int NT>
hipError_t run_nb16(const Uniform& U, const Arrays& A, const Op& op, const HRec* ofs,
                    const NbLaunch& nb, hipStream_t s) {
  if (nb.mode == Nb::BUILD) {
    hipLaunchKernelGGL((k_direct16_build<Op, NT>), dim3((U.n + NT - 1) / NT), dim3(NT), 0, s, U,
                       op, nb.args, A.cellStart, A.bcellStart, A.bpos, ofs, A.cellOfSlot);
    return hipGetLastError();
  }
  hipStream_t fs = nb.side ? nb.side : s;
  if (nb.side) {
    hipError_t e = hipEventRecord(nb.fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(nb.side, nb.fork, 0);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_fallback<Op, NT, true, true>), dim3(NB_QSEG * NB_FALLBACK_PARTS),
                     dim3(NT), 0, fs, U, op, nb.args, A.cellStart, A.bcellStart, A.bpos, ofs,
                     A.cellOfSlot);
  hipLaunchKernelGGL((k_listed16<Op>), dim3((U.n + 63) / 64), dim3(64), 0, s, U, op, nb.args,
                     A.cellStart, A.bcellStart, A.bpos, ofs, A.cellOfSlot);
  hipLaunchKernelGGL((k_fallback<Op, NT, false, true>), dim3(NB_QSEG * NB_FALLBACK_PARTS),
                     dim3(NT), 0, s, U, op, nb.args, A.cellStart, A.bcellStart, A.bpos, ofs,
                     A.cellOfSlot);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && nb.side) {
    e = hipEventRecord(nb.join, nb.side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, nb.join, 0);
  }
  return e;
}

template <class Op>
hipError_t run_direct16(const Uniform& U, const Arrays& A, Op op, const HRec* ofs,
                        int threads, hipStream_t s, const NbLaunch& nb) {
  if (U.n == 0) return hipSuccess;
  if (nb.mode != Nb::OFF) {
    if (nb.mode == Nb::USE) {
      op.nbRef = nb.args.ref;
      op.nbTravel = nb.args.ctl;
      op.nbMovedQ = nb.args.moved[nb.args.mslot ^ 1];
      op.nbMovedCnt = nb.args.ctl + nb_moved_counts(nb.args.mslot ^ 1);
      op.nbSegCap = nb.args.segCap;
    }
    if (threads == 256) return run_nb16<Op, 256>(U, A, op, ofs, nb, s);
    if (threads == 128) return run_nb16<Op, 128>(U, A, op, ofs, nb, s);
    return run_nb16<Op, 64>(U, A, op, ofs, nb, s);
  }
  if (threads == 256)
    hipLaunchKernelGGL((k_direct16<Op, 256>), dim3((U.n + 255) / 256), dim3(256), 0, s, U, op,
                       A.cellStart, A.bcellStart, A.bpos, ofs, A.cellOfSlot);
  else if (threads == 128)
    hipLaunchKernelGGL((k_direct16<Op, 128>), dim3((U.n + 127) / 128), dim3(128), 0, s, U, op,
                       A.cellStart, A.bcellStart, A.bpos, ofs, A.cellOfSlot);
  else
    hipLaunchKernelGGL((k_direct16<Op, 64>), dim3((U.n + 63) / 64), dim3(64), 0, s, U, op,
                       A.cellStart, A.bcellStart, A.bpos, ofs, A.cellOfSlot);
  return hipGetLastError();
}

hipError_t launch_lambda16(const Uniform& U, const Arrays& A, const float4* pred,
                           const HRec* ofs, int threads, const Drift& d, hipStream_t s,
                           const NbLaunch& nb) {
  Lambda16 op;
  op.pred = pred;
  op.density = A.density;
  op.lambda = A.lambda;
  op.predW = (float*)pred;
  return run_direct16(U, A, drifted(op, d), ofs, threads, s, nb);
}

hipError_t launch_delta16(const Uniform& U, const Arrays& A, const float4* pred,
                          const HRec* ofs, float4* out, HRec* ofsOut, int threads,
                          const Drift& d, hipStream_t s, const NbLaunch& nb) {
  // the neighbour count only when sorAverage divides by it (as the FP32 form)
  auto go = [&](auto op) {
    op.pred = pred;
    op.lambda = A.lambda;
    op.out = out;
    op.ofsOut = ofsOut;
    op.cellOfSlot = A.cellOfSlot;
    return run_direct16(U, A, drifted(op, d), ofs, threads, s, nb);
  };
  if (U.sorAverage == 1u) return go(Delta16T<true>{});
  return go(Delta16T<false>{});
}

hipError_t launch_body_pre(const Uniform& U, const Arrays& A, const float4* pred,
                           hipStream_t s) {
  if (U.nBodies == 0) return hipSuccess;
  hipLaunchKernelGGL(k_body_pre, dim3(U.nBodies), dim3(WG), 0, s, pred, A.slotOf, A.bodyPids,
                     A.bodyFirst, A.bodyCentreD, A.bodyCentre);
  return hipGetLastError();
}

hipError_t launch_body_solve(const Uniform& U, const Arrays& A, float4* pred, HRec* ofs,
                             uint32_t* driftOut, hipStream_t s, int movedSlot) {
  if (U.nBodies == 0) return hipSuccess;
  const bool nb = movedSlot >= 0;
  hipLaunchKernelGGL(k_body_solve, dim3(U.nBodies), dim3(WG), 0, s, pred, A.slotOf, A.bodyPids,
                     A.bodyFirst, A.restOfPid, A.bodyInvRatio, A.bodyCentreD, A.bodyRotD,
                     A.bodyCentre, A.bodyRot, ofs, A.cellOfSlot, U, driftOut,
                     nb ? A.nbRef : nullptr, nb ? A.nbCtl : nullptr,
                     nb ? A.nbMoved[movedSlot & 1] : nullptr,
                     nb ? A.nbCtl + nb_moved_counts(movedSlot & 1) : nullptr,
                     nb_seg_cap((uint32_t)U.n));
  return hipGetLastError();
}

// xsph+normals and tension walk the same rows of the same positions: with
// a two-phase direct shape, tension takes xsph's first-chunk masks instead
// of testing those candidates again (a batch indexes them by its global
// particle index).
bool reuse_masks(const Arrays& A, const TileCfg& t, const BatchTab*) {
  return A.hitMask && t.variant >= 5 && t.variant <= 8;
}

// The finish kernels under the substep's neighbour masks (list launch and
// fallback, as the solver stages after the build), else the tile shape's
// walk.
template <class Op>
hipError_t run_finish_op(const Uniform& U, const Arrays& A, const Op& op, const TileCfg& t,
                         const NbLaunch& nb, hipStream_t s, const BatchTab* B) {
  if (B || nb.mode != Nb::USE || U.n == 0) return run_tile2(U, A, op, t, s, B);
  switch (nb_threads(t)) {
    case 256: return run_nb<Op, 256>(U, A, op, nb, s);
    case 128: return run_nb<Op, 128>(U, A, op, nb, s);
    default: return run_nb<Op, 64>(U, A, op, nb, s);
  }
}

hipError_t launch_xsph(const Uniform& U, const Arrays& A, const float4* pred, const float4* pos,
                       bool exact, bool normals, const TileCfg& t, const Drift& d,
                       hipStream_t s, const BatchTab* B, const NbLaunch& nb) {
  float4* p = const_cast<float4*>(pos);
  if (normals) {
    // the first chunk's hit masks go to the tension kernel that follows
    const auto keep = [&](auto op) {
      if (nb.mode == Nb::USE && !B) return run_finish_op(U, A, drifted(op, d), t, nb, s, B);
      if (reuse_masks(A, t, B)) {
        op.maskOut = A.hitMask;
        op.maskStride = A.hitStride;
      }
      return run_tile2(U, A, drifted(op, d), t, s, B);
    };
    if (exact) return keep(XsphOp<true, true>{{}, pred, p, A.density, A.corr, A.normal});
    return keep(XsphFast<true>{{}, pred, A.velq, A.density, A.corr, A.normal});
  }
  if (exact)
    return run_finish_op(U, A, drifted(XsphOp<true, false>{{}, pred, p, A.density, A.corr, A.normal}, d), t, nb, s, B);
  return run_finish_op(U, A, drifted(XsphFast<false>{{}, pred, A.velq, A.density, A.corr, A.normal}, d), t, nb, s, B);
}

hipError_t launch_tension(const Uniform& U, const Arrays& A, const float4* pred, float4* pos,
                          float4* vel, bool exact, const TileCfg& t, const Drift& d,
                          hipStream_t s, const BatchTab* B, const NbLaunch& nb) {
  // the masks written by xsph+normals over the same positions
  const auto take = [&](auto op) {
    if (nb.mode == Nb::USE && !B) return run_finish_op(U, A, drifted(op, d), t, nb, s, B);
    if (reuse_masks(A, t, B)) {
      op.maskIn = A.hitMask;
      op.maskStride = A.hitStride;
    }
    return run_tile2(U, A, drifted(op, d), t, s, B);
  };
  if (exact) return take(TensionOp<true>{{}, pred, pos, A.density, A.normal, A.corr, vel});
  return take(TensionFast{{}, pred, pos, A.velq, A.normal, A.corr, vel});
}

hipError_t launch_finalize(const Uniform& U, const Arrays& A, const float4* pred, float4* pos,
                           float4* vel, hipStream_t s, const BatchTab* B) {
  if (B) {
    if (B->chunks > 0)
      hipLaunchKernelGGL(k_finalize_batch, dim3(B->chunks), dim3(WG), 0, s, *B, pred, pos,
                         A.corr, vel);
    return hipGetLastError();
  }
  if (U.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3(blocks(U.n)), dim3(WG), 0, s, U, pred, pos, A.corr, vel);
  return hipGetLastError();
}

hipError_t launch_prep_finish(const Uniform& U, const Arrays& A, const float4* pred,
                              const float4* pos, hipStream_t s, const BatchTab* B) {
  if (B) {
    if (B->chunks > 0)
      hipLaunchKernelGGL(k_prep_finish_batch, dim3(B->chunks), dim3(WG), 0, s, *B, pred, pos,
                         A.density, A.velq);
    return hipGetLastError();
  }
  if (U.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_finish, dim3(blocks(U.n)), dim3(WG), 0, s, U, pred, pos, A.density,
                     A.velq);
  return hipGetLastError();
}

hipError_t launch_stats(const Uniform& U, const Arrays& A, const float4* vel, hipStream_t s,
                        const BatchTab* B) {
  if (B) {
    if (B->chunks > 0)
      hipLaunchKernelGGL(k_stats_partial_batch, dim3(B->chunks), dim3(WG), 0, s, *B, A.density,
                         vel, A.statsPart);
    if (B->nScenes > 0)
      hipLaunchKernelGGL(k_stats_final_batch, dim3(B->nScenes), dim3(WG), 0, s, *B, A.statsPart,
                         A.statsOut);
    return hipGetLastError();
  }
  const int nb = stats_blocks(U.n);
  if (nb > 0)
    hipLaunchKernelGGL(k_stats_partial, dim3(nb), dim3(WG), 0, s, U, A.density, vel,
                       A.statsPart);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(WG), 0, s, A.statsPart, nb, U.n, A.statsOut);
  return hipGetLastError();
}

hipError_t launch_stats_into(const Uniform& U, const float* density, const float4* vel,
                             double* part, double* out, hipStream_t s) {
  const int nb = stats_blocks(U.n);
  if (nb > 0)
    hipLaunchKernelGGL(k_stats_partial, dim3(nb), dim3(WG), 0, s, U, density, vel, part);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(WG), 0, s, part, nb, U.n, out);
  return hipGetLastError();
}

hipError_t launch_body_drag(const Uniform& U, const Arrays& A, float4* vel, const DragArgs& d,
                            hipStream_t s) {
  if (U.nBodies == 0 || d.held < 1 || d.held > U.nBodies) return hipSuccess;
  // the grid covers the largest body; extra threads exit
  hipLaunchKernelGGL(k_body_drag, dim3(blocks(U.nBodyParts)), dim3(WG), 0, s, d, vel,
                     A.slotOf, A.bodyPids, A.bodyFirst, A.bodyCentre);
  return hipGetLastError();
}

hipError_t launch_impulse(const float4* pos, float4* vel, const RayArgs& r, hipStream_t s) {
  if (r.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_impulse, dim3(blocks(r.n)), dim3(WG), 0, s, r, pos, vel);
  return hipGetLastError();
}

hipError_t launch_resize(const ResizeArgs& r, float4* pos, hipStream_t s) {
  if (r.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_resize, dim3(blocks(r.n)), dim3(WG), 0, s, r, pos);
  return hipGetLastError();
}

hipError_t launch_census(const Uniform& U, const Arrays& A, const float4* pred,
                         const TileCfg& t, hipStream_t s) {
  CensusOp op{{}, pred, A.census[0], A.census[1], A.census[2]};
  return run_tile1(U, A, op, t, s);
}

hipError_t launch_aniso(const Uniform& U, const Arrays& A, const float4* pos,
                        const AnisoArgs& a, float4* smoothPos, float4* g, uint32_t* fallbacks,
                        bool exact, hipStream_t s) {
  if (exact) {
    AnisoOp<true> op{{}, pos, smoothPos, g, fallbacks, a};
    return run_direct<AnisoOp<true>, 256, 2>(U, A, op, s);
  }
  AnisoOp<false> op{{}, pos, smoothPos, g, fallbacks, a};
  return run_direct<AnisoOp<false>, 256, 2>(U, A, op, s);
}

hipError_t launch_field_iso(const Uniform& U, const Arrays& A, const FieldArgs& F,
                            const float4* pos, const uint32_t* pid, float4* rec, float* field,
                            hipStream_t s) {
  if (U.n > 0)
    hipLaunchKernelGGL(k_field_records, dim3(blocks(U.n)), dim3(WG), 0, s, F, U.n, pos,
                       A.density, pid, A.bodyOfPid, rec);
  if (F.nVerts > 0)
    hipLaunchKernelGGL(k_field_iso, dim3(blocks(F.nVerts)), dim3(WG), 0, s, U, F, A.cellStart,
                       rec, field);
  return hipGetLastError();
}

hipError_t launch_field_aniso(const Uniform& U, const Arrays& A, const FieldArgs& F,
                              const float4* smoothPos, const float4* g, const uint32_t* pid,
                              float4* recA, float4* recB, unsigned long long* acc,
                              float* field, hipStream_t s) {
  if (F.nVerts == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks(F.nVerts)), dim3(WG), 0, s, acc,
                     (size_t)F.nVerts);
  if (U.n > 0) {
    hipLaunchKernelGGL(k_field_aniso_prep, dim3(blocks(U.n)), dim3(WG), 0, s, F, U.n, smoothPos,
                       g, A.density, pid, A.bodyOfPid, recA, recB);
    constexpr int NT = 64, CAP = 4096;  // 8 cells of particles, 32 KB of LDS
    hipLaunchKernelGGL((k_field_splat<NT, CAP>), dim3((U.n + NT - 1) / NT), dim3(NT), 0, s, F,
                       U.n, recA, recB, acc);
  }
  hipLaunchKernelGGL(k_field_fixed_to_float, dim3(blocks(F.nVerts)), dim3(WG), 0, s, F.nVerts,
                     acc, field);
  return hipGetLastError();
}

hipError_t launch_copy4(const float4* a, float4* b, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t g = (n + WG - 1) / WG;
  if (g > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy4, dim3((unsigned)g), dim3(WG), 0, s, a, b, n);
  return hipGetLastError();
}

hipError_t launch_export_phase(uint32_t n, const uint32_t* pid, const uint32_t* bodyOfPid,
                               uint4* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_export_phase, dim3(blocks(n)), dim3(WG), 0, s, n, pid, bodyOfPid, out);
  return hipGetLastError();
}

hipError_t launch_body_pack(uint32_t count, const float4* centre, const float4* rot,
                            const float4* stat, float4* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_body_pack, dim3((count + 63) / 64), dim3(64), 0, s, count, centre, rot,
                     stat, out);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n)), dim3(WG), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace p4a
