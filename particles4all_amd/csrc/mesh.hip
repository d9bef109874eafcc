// gfx950 kernels of the surface extraction that follows the density field
// (SurfaceMesh.smoothField, buildNormalField and the march of generate,
// src/mesh.js:273-331):
//   meshSmoothWGSL (mesh_wgsl.js:70-91)    -> k_mesh_blur (x, y, z fused)
//   meshMarchWGSL (mesh_wgsl.js:93-233)    -> k_mesh_count + scan + k_mesh_emit
//   meshIndirectWGSL (mesh_wgsl.js:235-244) -> k_mesh_finish
//
// Blur: the reference runs one pass per axis; k_mesh_blur stages a 64x4x4
// tile of vertices with a one-vertex halo in LDS and applies the x, y and
// z steps there, so a full round reads and writes the field once.  Every
// intermediate value is the one the per-axis passes produce (the halo
// holds the x and y steps of the clamped vertices, which is what the next
// axis step of the reference reads).
//
// March: the reference appends triangles with an atomic counter, so their
// order in the vertex buffer is whatever order the atomics resolved in.
// Here k_mesh_count gives every workgroup of 256 consecutive voxels its
// triangle count, an exclusive scan turns the counts into offsets, and
// k_mesh_emit writes the workgroup's triangles from its offset on.
// Triangles are numbered in voxel order (voxel, then tetrahedron, then
// triangle), the same on every run; the budget keeps the triangles
// numbered below maxTris, and the counter holds the total including the
// dropped ones, as the reference's counter does.
//
// Arithmetic: one IEEE operation per WGSL operation in the WGSL order (the
// file is compiled with -ffp-contract=off); mix(a, b, t) = a * (1 - t) +
// b * t, clamp(x, lo, hi) = min(max(x, lo), hi).
#include "p4a_internal.h"

namespace p4a {
namespace {

constexpr int BX = 64, BY = 4, BZ = 4;  // blur output tile (256 threads x 4)

__device__ __forceinline__ float blur3(float lo, float mid, float hi) {
  const float a = 0.25f * lo;
  const float b = 0.5f * mid;
  const float d = 0.25f * hi;
  return (a + b) + d;
}

// One round of meshSmoothWGSL's x, y and z steps: dst(v) = 0.25 S(v - a) +
// 0.5 S(v) + 0.25 S(v + a) along a = x, then y, then z, S clamping the
// vertex into the lattice.
__global__ void __launch_bounds__(MESH_WG)
k_mesh_blur(MeshArgs M, const float* __restrict__ src, float* __restrict__ dst) {
  __shared__ float A[(BZ + 2) * (BY + 2) * (BX + 2)];  // source with halo; later the y step
  __shared__ float X[(BZ + 2) * (BY + 2) * BX];        // x step
  const int vx = M.vertDim[0], vy = M.vertDim[1], vz = M.vertDim[2];
  // bricks in x, y, z order, handed to the XCDs in contiguous ranges so the
  // halo rows and planes two bricks share are read once into one L2
  const int tx = (vx + BX - 1) / BX, ty = (vy + BY - 1) / BY;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int x0 = (tile % tx) * BX, y0 = ((tile / tx) % ty) * BY, z0 = (tile / (tx * ty)) * BZ;
  constexpr int AX = BX + 2, AY = BY + 2, AZ = BZ + 2;
  for (int k = threadIdx.x; k < AX * AY * AZ; k += MESH_WG) {
    const int lx = k % AX, ly = (k / AX) % AY, lz = k / (AX * AY);
    const int gx = min(max(x0 - 1 + lx, 0), vx - 1);
    const int gy = min(max(y0 - 1 + ly, 0), vy - 1);
    const int gz = min(max(z0 - 1 + lz, 0), vz - 1);
    A[k] = src[((uint32_t)gz * (uint32_t)vy + (uint32_t)gy) * (uint32_t)vx + (uint32_t)gx];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < BX * AY * AZ; k += MESH_WG) {
    const int lx = k % BX, r = k / BX;  // r = lz * AY + ly
    const float* a = &A[r * AX + lx];
    X[k] = blur3(a[0], a[1], a[2]);
  }
  __syncthreads();
  float* Y = A;  // (BZ + 2) planes of BY rows
  for (int k = threadIdx.x; k < BX * BY * AZ; k += MESH_WG) {
    const int lx = k % BX, ly = (k / BX) % BY, lz = k / (BX * BY);
    const float* a = &X[(lz * AY + ly) * BX + lx];
    Y[k] = blur3(a[0], a[BX], a[2 * BX]);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < BX * BY * BZ; k += MESH_WG) {
    const int lx = k % BX, ly = (k / BX) % BY, lz = k / (BX * BY);
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    if (gx >= vx || gy >= vy || gz >= vz) continue;
    const float* a = &Y[(lz * BY + ly) * BX + lx];
    dst[((uint32_t)gz * (uint32_t)vy + (uint32_t)gy) * (uint32_t)vx + (uint32_t)gx] =
        blur3(a[0], a[BX * BY], a[2 * BX * BY]);
  }
}

// Corner k of voxel (x, y, z) is vertex (x + (k & 1), y + (k >> 1 & 1),
// z + (k >> 2)); the six tetrahedra all contain the 0-7 diagonal:
// (0,1,3,7) (0,1,5,7) (0,2,3,7) (0,2,6,7) (0,4,5,7) (0,4,6,7).
__device__ __forceinline__ uint32_t voxel_vertex(const MeshArgs& M, uint32_t v, uint32_t& x,
                                                 uint32_t& y, uint32_t& z) {
  const uint32_t row = M.divNx.div(v);
  x = v - row * (uint32_t)(M.vertDim[0] - 1);
  z = M.divNy.div(row);
  y = row - z * (uint32_t)(M.vertDim[1] - 1);
  return (z * (uint32_t)M.vertDim[1] + y) * (uint32_t)M.vertDim[0] + x;
}

__device__ __forceinline__ uint32_t corner_mask(const MeshArgs& M, const float* __restrict__ f,
                                                uint32_t b) {
  const uint32_t sy = (uint32_t)M.vertDim[0];
  const uint32_t sz = sy * (uint32_t)M.vertDim[1];
  const float d[8] = {f[b],      f[b + 1],      f[b + sy],      f[b + sy + 1],
                      f[b + sz], f[b + sz + 1], f[b + sz + sy], f[b + sz + sy + 1]};
  uint32_t mask = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) mask |= (d[k] > M.iso ? 1u : 0u) << k;
  return mask;
}

constexpr uint32_t TET_MASK[6] = {0x8Bu, 0xA3u, 0x8Du, 0xC5u, 0xB1u, 0xD1u};

// Triangles of one tetrahedron: 1 when 1 or 3 corners are inside, 2 when 2
// are (emitTet, mesh_wgsl.js:166-188).
__device__ __forceinline__ uint32_t tet_tris(uint32_t mask, int t) {
  const uint32_t n = (uint32_t)__popc(mask & TET_MASK[t]);
  return n == 2u ? 2u : (n & 1u);
}

__device__ __forceinline__ uint32_t voxel_tris(uint32_t mask) {
  uint32_t c = 0;
#pragma unroll
  for (int t = 0; t < 6; t++) c += tet_tris(mask, t);
  return c;
}

__global__ void __launch_bounds__(MESH_WG)
k_mesh_count(MeshArgs M, const float* __restrict__ field, uint32_t* __restrict__ blockCount) {
  __shared__ uint32_t red[MESH_WG / 64];
  // voxel chunk `tile`: contiguous chunks (shared corner rows and planes)
  // stay on one XCD
  const uint32_t tile = (uint32_t)xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t v = tile * MESH_WG + threadIdx.x;
  uint32_t c = 0;
  if (v < M.nVox) {
    uint32_t x, y, z;
    c = voxel_tris(corner_mask(M, field, voxel_vertex(M, v, x, y, z)));
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < MESH_WG / 64; w++) t += red[w];
    blockCount[tile] = t;
  }
}

struct MCorner {
  float d;
  float p[3];
  float g[3];
};

__device__ __forceinline__ float at_clamped(const MeshArgs& M, const float* __restrict__ f,
                                            int x, int y, int z) {
  x = min(max(x, 0), M.vertDim[0] - 1);
  y = min(max(y, 0), M.vertDim[1] - 1);
  z = min(max(z, 0), M.vertDim[2] - 1);
  return f[((uint32_t)z * (uint32_t)M.vertDim[1] + (uint32_t)y) * (uint32_t)M.vertDim[0] +
           (uint32_t)x];
}

// Corner k of the voxel at lattice vertex (x, y, z): field value, lattice
// position, and the gradient of the normal field by central differences
// (gradAt, mesh_wgsl.js:106-110, clamped at the lattice faces).
__device__ __forceinline__ MCorner load_corner(const MeshArgs& M, const float* __restrict__ field,
                                               const float* __restrict__ nfield, uint32_t x,
                                               uint32_t y, uint32_t z, int k) {
  const int X = (int)x + (k & 1), Y = (int)y + ((k >> 1) & 1), Z = (int)z + (k >> 2);
  MCorner c;
  c.d = field[((uint32_t)Z * (uint32_t)M.vertDim[1] + (uint32_t)Y) * (uint32_t)M.vertDim[0] +
              (uint32_t)X];
  c.p[0] = (float)X;
  c.p[1] = (float)Y;
  c.p[2] = (float)Z;
  c.g[0] = at_clamped(M, nfield, X + 1, Y, Z) - at_clamped(M, nfield, X - 1, Y, Z);
  c.g[1] = at_clamped(M, nfield, X, Y + 1, Z) - at_clamped(M, nfield, X, Y - 1, Z);
  c.g[2] = at_clamped(M, nfield, X, Y, Z + 1) - at_clamped(M, nfield, X, Y, Z - 1);
  return c;
}

// packVertex (mesh_wgsl.js:112-124).
__device__ __forceinline__ uint2 pack_vertex(const MeshArgs& M, const float p[3],
                                             const float n[3]) {
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float pn = fminf(fmaxf(p[k] / (float)(M.vertDim[k] - 1), 0.0f), 1.0f);
    q[k] = (uint32_t)(pn * 65535.0f + 0.5f);
  }
  const float l1 = fmaxf(fabsf(n[0]) + fabsf(n[1]) + fabsf(n[2]), 1.0e-12f);
  const float nx = n[0] / l1, ny = n[1] / l1, nz = n[2] / l1;
  float e0 = nx, e1 = ny;
  if (nz < 0.0f) {
    e0 = (1.0f - fabsf(ny)) * (nx >= 0.0f ? 1.0f : -1.0f);
    e1 = (1.0f - fabsf(nx)) * (ny >= 0.0f ? 1.0f : -1.0f);
  }
  const uint32_t o0 = (uint32_t)(fminf(fmaxf(e0 * 0.5f + 0.5f, 0.0f), 1.0f) * 255.0f + 0.5f);
  const uint32_t o1 = (uint32_t)(fminf(fmaxf(e1 * 0.5f + 0.5f, 0.0f), 1.0f) * 255.0f + 0.5f);
  return make_uint2(q[0] | (q[1] << 16), q[2] | ((o0 | (o1 << 8)) << 16));
}

// edgeVertex (mesh_wgsl.js:126-132): the crossing on edge (a, b).
__device__ __forceinline__ uint2 edge_vertex(const MeshArgs& M, const MCorner& a,
                                             const MCorner& b) {
  const float t = (M.iso - a.d) / (b.d - a.d);
  const float u = 1.0f - t;
  float p[3], n[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    p[k] = a.p[k] * u + b.p[k] * t;
    n[k] = -(a.g[k] * u + b.g[k] * t);
  }
  return pack_vertex(M, p, n);
}

// The three edges (corner pairs, first corner first) of triangle r of a
// voxel with inside mask `mask`, numbered as main / emitTet / emitLone /
// emitQuad emit them (mesh_wgsl.js:143-233).
__device__ __forceinline__ void triangle_edges(uint32_t mask, uint32_t r, int e[6]) {
  constexpr int TB[6] = {1, 1, 2, 2, 4, 4};
  constexpr int TC[6] = {3, 5, 3, 6, 5, 6};
  int k = 0;
  uint32_t sub = 0, acc = 0;
#pragma unroll
  for (int t = 0; t < 6; t++) {
    const uint32_t c = tet_tris(mask, t);
    if (r >= acc && r < acc + c) {
      k = t;
      sub = r - acc;
    }
    acc += c;
  }
  // tetrahedron (a, b, c, d) = (0, TB[k], TC[k], 7)
  const int a = 0, b = TB[k], c = TC[k], d = 7;
  const bool ia = (mask >> a) & 1u, ib = (mask >> b) & 1u, ic = (mask >> c) & 1u,
             id = (mask >> d) & 1u;
  const int nin = (int)ia + (int)ib + (int)ic + (int)id;
  if (nin != 2) {
    // emitLone: the corner alone on its side, then the others in order
    const bool look = nin == 1;
    int l, p1, p2, p3;
    if (ia == look) { l = a; p1 = b; p2 = c; p3 = d; }
    else if (ib == look) { l = b; p1 = a; p2 = c; p3 = d; }
    else if (ic == look) { l = c; p1 = a; p2 = b; p3 = d; }
    else { l = d; p1 = a; p2 = b; p3 = c; }
    e[0] = l; e[1] = p1;
    e[2] = l; e[3] = p2;
    e[4] = l; e[5] = p3;
    return;
  }
  // emitQuad(qa, qb, qc, qd): q0 = (qa,qc), q1 = (qa,qd), q2 = (qb,qd),
  // q3 = (qb,qc); triangle 0 = (q0, q1, q2), triangle 1 = (q0, q2, q3)
  int qa, qb, qc, qd;
  if (ia && ib) { qa = a; qb = b; qc = c; qd = d; }
  else if (ia && ic) { qa = a; qb = c; qc = b; qd = d; }
  else if (ia && id) { qa = a; qb = d; qc = b; qd = c; }
  else if (ib && ic) { qa = b; qb = c; qc = a; qd = d; }
  else if (ib && id) { qa = b; qb = d; qc = a; qd = c; }
  else { qa = c; qb = d; qc = a; qd = b; }
  e[0] = qa; e[1] = qc;
  e[2] = sub == 0u ? qa : qb; e[3] = qd;
  e[4] = qb; e[5] = sub == 0u ? qd : qc;
}

// Emission, one thread per triangle: a workgroup takes the voxels of one
// count workgroup, scans their triangle counts in LDS, and then thread i
// takes triangles i, i + 256, ... of the workgroup: the voxel owning it by
// binary search over the scanned counts, its tetrahedron and edges from the
// voxel's inside mask, and its three edge vertices.  All lanes run the same
// instructions whatever configuration their voxel has.
__global__ void __launch_bounds__(MESH_WG)
k_mesh_emit(MeshArgs M, const float* __restrict__ field, const float* __restrict__ nfield,
            const uint32_t* __restrict__ blockStart, uint2* __restrict__ verts) {
  __shared__ uint32_t pref[MESH_WG];
  __shared__ uint32_t masks[MESH_WG];
  __shared__ uint32_t wsum[MESH_WG / 64];
  // chunks in dispatch order (round-robin over the XCDs): the triangles sit
  // in the few chunks the surface crosses, and contiguous ranges per XCD
  // would leave most XCDs idle (243 against 145 us measured)
  const uint32_t tile = blockIdx.x;
  const uint32_t first = blockStart[tile];
  const uint32_t total = blockStart[tile + 1] - first;
  if (total == 0u) return;  // no triangle in this workgroup
  const uint32_t v = tile * MESH_WG + threadIdx.x;
  uint32_t mask = 0, cnt = 0;
  if (v < M.nVox) {
    uint32_t x, y, z;
    mask = corner_mask(M, field, voxel_vertex(M, v, x, y, z));
    cnt = voxel_tris(mask);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = cnt;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int w = 0; w < wave; w++) pre += wsum[w];
  pref[threadIdx.x] = pre + inc - cnt;
  masks[threadIdx.x] = mask;
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < total; j += MESH_WG) {
    const uint32_t k = first + j;
    if (k >= M.maxTris) break;
    // the last voxel l with pref[l] <= j owns triangle j
    uint32_t lo = 0, hi = MESH_WG;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pref[mid] <= j) lo = mid;
      else hi = mid;
    }
    uint32_t x, y, z;
    voxel_vertex(M, tile * MESH_WG + lo, x, y, z);
    int e[6];
    triangle_edges(masks[lo], j - pref[lo], e);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const MCorner ca = load_corner(M, field, nfield, x, y, z, e[2 * q]);
      const MCorner cb = load_corner(M, field, nfield, x, y, z, e[2 * q + 1]);
      verts[3 * (size_t)k + q] = edge_vertex(M, ca, cb);
    }
  }
}

// The counter (total triangles, dropped ones included) and the draw-indirect
// arguments (vertex count, instance count 1, first vertex 0, first
// instance 0) of meshIndirectWGSL.
__global__ void k_mesh_finish(const uint32_t* __restrict__ blockStart, uint32_t nBlocks,
                              uint32_t maxTris, uint32_t* __restrict__ counter,
                              uint32_t* __restrict__ indirect) {
  if (threadIdx.x != 0) return;
  const uint32_t total = blockStart[nBlocks];
  counter[0] = total;
  indirect[0] = min(total, maxTris) * 3u;
  indirect[1] = 1u;
  indirect[2] = 0u;
  indirect[3] = 0u;
}

}  // namespace

hipError_t launch_mesh_blur(const MeshArgs& M, int passes, const float* src, float* tmp,
                            float* dst, hipStream_t s) {
  if (passes <= 0 || M.nVerts == 0) return hipSuccess;
  const dim3 grid((uint32_t)((M.vertDim[0] + BX - 1) / BX) * (uint32_t)((M.vertDim[1] + BY - 1) / BY) *
                  (uint32_t)((M.vertDim[2] + BZ - 1) / BZ));
  const float* in = src;
  for (int k = 0; k < passes; k++) {
    // a round reads a halo, so it never writes its input: in place (src ==
    // dst) rounds alternate tmp, dst, ... with a copy back after an odd
    // count; otherwise the last round lands in dst
    float* out = src == dst ? ((k & 1) ? dst : tmp) : (((passes - 1 - k) & 1) ? tmp : dst);
    hipLaunchKernelGGL(k_mesh_blur, grid, dim3(MESH_WG), 0, s, M, in, out);
    in = out;
  }
  if (in != dst) {
    const hipError_t e = hipMemcpyAsync(dst, in, (size_t)M.nVerts * 4, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t launch_mesh_march(const MeshArgs& M, const float* field, const float* nfield,
                             uint32_t* blockCount, uint32_t* blockSum, uint32_t* blockStart,
                             uint2* verts, uint32_t* counter, uint32_t* indirect, hipStream_t s) {
  if (M.nBlocks > 0) {
    hipLaunchKernelGGL(k_mesh_count, dim3(M.nBlocks), dim3(MESH_WG), 0, s, M, field, blockCount);
    const hipError_t e = launch_exclusive_scan(blockCount, M.nBlocks, blockSum, blockStart, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mesh_emit, dim3(M.nBlocks), dim3(MESH_WG), 0, s, M, field, nfield,
                       blockStart, verts);
  }
  hipLaunchKernelGGL(k_mesh_finish, dim3(1), dim3(64), 0, s, blockStart, M.nBlocks, M.maxTris,
                     counter, indirect);
  return hipGetLastError();
}

}  // namespace p4a
