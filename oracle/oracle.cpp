// ORACLE - test infrastructure only.  Nothing in the product links,
// imports or calls this file; tests/, __graft_entry__.smoke() and the CPU
// baseline leg of bench.py use it as the checker.
//
// A CPU restatement of the Particles4All simulation step, following the
// reference's WGSL kernels (src/wgsl.js) and Sim orchestration (src/sim.js)
// operation by operation.  Every function runs in one of two modes:
//   mode 0 (R = float):  one IEEE float operation per WGSL operation, in the
//                        WGSL's order (no contraction; built with
//                        -ffp-contract=off), constants rounded to float as
//                        the reference's uniform buffer stores them;
//   mode 1 (R = double): the same sequence in FP64 with FP64 constants -
//                        the FP64 evaluation the GPU results are measured
//                        against;
//   mode 2 (R = float):  mode 0 with every sqrt and division of the
//                        neighbour kernels (lambda, delta, xsph, tension)
//                        one ulp above the correctly rounded result.  WGSL
//                        gives its f32 division 2.5 ULP and its sqrt (as
//                        1/inverseSqrt) more, so this is one conforming
//                        implementation of the reference's arithmetic: its
//                        distance from mode 0 is a floor for how far two
//                        conforming runs of the reference may differ.
// r^2 is fma(z,z, fma(y,y, x*x)) in both modes (the WGSL `dot` leaves the
// contraction to the WebGPU compiler; this fixes one choice).
// Rigid-body sums and the polar decomposition are FP64 in both modes; the
// reference quantises those sums to i32 fixed point (wgsl.js:628-698), and
// the difference is bounded by 0.5/bodyScale per summed term.
//
// Array I/O is double throughout (float values are exact in double).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// Constants of the reference's Params uniform (wgsl.js:4-40), computed by
// the caller from the scene and parameters (the oracle's restatement of
// Sim.uploadParams, sim.js:371-413, lives in oracle.py).
struct OrcParams {
  double dt, h, h2, poly6, spiky, gravity, mass, rho0, invRho0;
  double cfmEps, sCorrK, sCorrWq, xsphC, omega, invDt, volume, tension, cohes, cohesTerm;
  double boxMin[3], clampMin[3], clampMax[3];
  int32_t gridDim[3];
  int32_t sorAverage;
  int32_t nBoundary;
  int32_t pad;
};
}

namespace {

template <class R>
struct K {  // constants in mode R
  R dt, h, h2, poly6, spiky, gravity, mass, rho0, invRho0, cfmEps, sCorrK, sCorrWq, xsphC,
      omega, invDt, volume, tension, cohes, cohesTerm;
  R boxMin[3], clampMin[3], clampMax[3];
  int g[3];
  int sor;
  int nB;
  explicit K(const OrcParams& p) {
    dt = (R)p.dt; h = (R)p.h; h2 = (R)p.h2; poly6 = (R)p.poly6; spiky = (R)p.spiky;
    gravity = (R)p.gravity; mass = (R)p.mass; rho0 = (R)p.rho0; invRho0 = (R)p.invRho0;
    cfmEps = (R)p.cfmEps; sCorrK = (R)p.sCorrK; sCorrWq = (R)p.sCorrWq; xsphC = (R)p.xsphC;
    omega = (R)p.omega; invDt = (R)p.invDt; volume = (R)p.volume; tension = (R)p.tension;
    cohes = (R)p.cohes; cohesTerm = (R)p.cohesTerm;
    for (int k = 0; k < 3; k++) {
      boxMin[k] = (R)p.boxMin[k]; clampMin[k] = (R)p.clampMin[k]; clampMax[k] = (R)p.clampMax[k];
      g[k] = p.gridDim[k];
    }
    sor = p.sorAverage;
    nB = p.nBoundary;
  }
};

inline float fma_r(float a, float b, float c) { return fmaf(a, b, c); }
inline double fma_r(double a, double b, double c) { return fma(a, b, c); }
// mode 2's shift (read-only while a kernel runs)
int g_ulpUp = 0;
inline float sqrt_r(float a) {
  const float r = sqrtf(a);
  return g_ulpUp ? nextafterf(r, INFINITY) : r;
}
inline double sqrt_r(double a) { return sqrt(a); }
inline float div_r(float a, float b) {
  const float r = a / b;
  return g_ulpUp ? nextafterf(r, INFINITY) : r;
}
inline double div_r(double a, double b) { return a / b; }
// float for modes 0 and 2 (setting mode 2's shift), double for mode 1
bool float_mode(int mode) {
  g_ulpUp = mode == 2;
  return mode != 1;
}

template <class R>
inline R r2_of(R x, R y, R z) { return fma_r(z, z, fma_r(y, y, x * x)); }

// cellOf (wgsl.js:43-46): floor((p - boxMin)/h) clamped to the grid.
template <class R>
inline void cell_of(const K<R>& k, R x, R y, R z, int c[3]) {
  const R p[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    R f = std::floor((p[a] - k.boxMin[a]) / k.h);
    f = std::fmin(std::fmax(f, (R)0), (R)(k.g[a] - 1));  // NaN -> 0
    c[a] = (int)f;
  }
}
template <class R>
inline uint32_t cell_index(const K<R>& k, const int c[3]) {
  return (uint32_t)((c[2] * k.g[1] + c[1]) * k.g[0] + c[0]);
}

// poly6 (wgsl.js:50-53) for r2 < h2
template <class R>
inline R poly6(const K<R>& k, R r2) {
  R t = k.h2 - r2;
  return k.poly6 * t * t * t;
}

// The reference's 27-cell stencil walk (e.g. wgsl.js:497-534): rows dz, dy;
// per row the fluid span [cellStart[row+x0], cellStart[row+x1+1]) and then
// the boundary span of the same cells.
template <class R, class F, class B>
inline void stencil(const K<R>& k, R px, R py, R pz, const uint32_t* cs, const uint32_t* bcs,
                    F&& fluid, B&& bnd) {
  int c[3];
  cell_of(k, px, py, pz, c);
  const int x0 = std::max(c[0] - 1, 0), x1 = std::min(c[0] + 1, k.g[0] - 1);
  for (int dz = -1; dz <= 1; dz++) {
    const int z = c[2] + dz;
    if (z < 0 || z >= k.g[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int y = c[1] + dy;
      if (y < 0 || y >= k.g[1]) continue;
      const uint32_t rb = (uint32_t)((z * k.g[1] + y) * k.g[0]);
      for (uint32_t j = cs[rb + x0]; j < cs[rb + x1 + 1]; j++) fluid(j);
      if (bcs && k.nB > 0)
        for (uint32_t q = bcs[rb + x0]; q < bcs[rb + x1 + 1]; q++) bnd(q);
    }
  }
}

template <class R>
void predict(const OrcParams& P, int n, const double* pos, const double* vel, double* pred,
             double* velOut) {
  K<R> k(P);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    // wgsl.js:66-68
    const R gx = k.dt * (R)0, gy = k.dt * (-k.gravity), gz = k.dt * (R)0;
    const R vx = (R)vel[i * 4] + gx, vy = (R)vel[i * 4 + 1] + gy, vz = (R)vel[i * 4 + 2] + gz;
    pred[i * 4] = (R)pos[i * 4] + k.dt * vx;
    pred[i * 4 + 1] = (R)pos[i * 4 + 1] + k.dt * vy;
    pred[i * 4 + 2] = (R)pos[i * 4 + 2] + k.dt * vz;
    pred[i * 4 + 3] = 1.0;
    if (velOut) {
      velOut[i * 4] = vx; velOut[i * 4 + 1] = vy; velOut[i * 4 + 2] = vz; velOut[i * 4 + 3] = 0;
    }
  }
}

template <class R>
void lambda_k(const OrcParams& P, int n, const double* pred, const uint32_t* cs,
              const double* bpos, const double* bpsi, const uint32_t* bcs, double* density,
              double* lambda) {
  K<R> k(P);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    const R px = (R)pred[i * 4], py = (R)pred[i * 4 + 1], pz = (R)pred[i * 4 + 2];
    R rho = k.mass * (k.poly6 * k.h2 * k.h2 * k.h2);  // mass * poly6(0)
    R gx = 0, gy = 0, gz = 0, sg2 = 0;
    stencil(k, px, py, pz, cs, bcs,
      [&](uint32_t j) {
        if (j == (uint32_t)i) return;
        const R rx = px - (R)pred[j * 4], ry = py - (R)pred[j * 4 + 1], rz = pz - (R)pred[j * 4 + 2];
        const R r2 = r2_of(rx, ry, rz);
        if (r2 >= k.h2) return;
        rho = rho + k.mass * poly6(k, r2);
        if (r2 > (R)1.0e-12) {
          const R r = sqrt_r(r2);
          const R hr = k.h - r;
          const R s = div_r(k.volume * k.spiky * hr * hr, r);
          const R ax = s * rx, ay = s * ry, az = s * rz;
          gx = gx + ax; gy = gy + ay; gz = gz + az;
          sg2 = sg2 + (ax * ax + ay * ay + az * az);
        }
      },
      [&](uint32_t q) {
        const R rx = px - (R)bpos[q * 4], ry = py - (R)bpos[q * 4 + 1], rz = pz - (R)bpos[q * 4 + 2];
        const R r2 = r2_of(rx, ry, rz);
        if (r2 >= k.h2 || r2 < (R)1.0e-12) return;
        const R psi = (R)bpsi[q];
        rho = rho + psi * poly6(k, r2);
        const R r = sqrt_r(r2);
        const R hr = k.h - r;
        const R s = div_r(psi * k.invRho0 * k.spiky * hr * hr, r);
        gx = gx + s * rx; gy = gy + s * ry; gz = gz + s * rz;
      });
    density[i] = rho;
    const R C = rho * k.invRho0 - (R)1;
    R lam = 0;
    if (C > (R)0) lam = div_r(-C, (gx * gx + gy * gy + gz * gz) + sg2 + k.cfmEps + (R)1.0e-12);
    lambda[i] = lam;
  }
}

template <class R>
void delta_k(const OrcParams& P, int n, const double* pred, const double* lambda,
             const uint32_t* cs, const double* bpos, const double* bpsi, const uint32_t* bcs,
             double* out, const double* base) {
  K<R> k(P);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    const R px = (R)pred[i * 4], py = (R)pred[i * 4 + 1], pz = (R)pred[i * 4 + 2];
    const R li = (R)lambda[i];
    R dx = 0, dy = 0, dz = 0, bx = 0, by = 0, bz = 0, nc = 1;
    stencil(k, px, py, pz, cs, bcs,
      [&](uint32_t j) {
        if (j == (uint32_t)i) return;
        const R rx = px - (R)pred[j * 4], ry = py - (R)pred[j * 4 + 1], rz = pz - (R)pred[j * 4 + 2];
        const R r2 = r2_of(rx, ry, rz);
        if (r2 >= k.h2 || r2 < (R)1.0e-12) return;
        const R r = sqrt_r(r2);
        const R hr = k.h - r;
        const R s = div_r(k.spiky * hr * hr, r);
        const R t = k.h2 - r2;
        const R ratio = k.poly6 * t * t * t * k.sCorrWq;
        const R r2r = ratio * ratio;
        const R sc = -k.sCorrK * r2r * r2r;
        const R f = (li + (R)lambda[j] + sc) * s;
        dx = dx + f * rx; dy = dy + f * ry; dz = dz + f * rz;
        nc = nc + (R)1;
      },
      [&](uint32_t q) {
        const R rx = px - (R)bpos[q * 4], ry = py - (R)bpos[q * 4 + 1], rz = pz - (R)bpos[q * 4 + 2];
        const R r2 = r2_of(rx, ry, rz);
        if (r2 >= k.h2 || r2 < (R)1.0e-12) return;
        const R r = sqrt_r(r2);
        const R hr = k.h - r;
        const R f = (li * (R)bpsi[q]) * div_r(k.spiky * hr * hr, r);
        bx = bx + f * rx; by = by + f * ry; bz = bz + f * rz;
      });
    const R ex = dx * k.volume + bx * k.invRho0;
    const R ey = dy * k.volume + by * k.invRho0;
    const R ez = dz * k.volume + bz * k.invRho0;
    R scale = k.omega;
    if (k.sor == 1) scale = div_r(k.omega, std::max(nc, (R)1));
    R bxp = px, byp = py, bzp = pz;
    if (base) { bxp = (R)base[i * 4]; byp = (R)base[i * 4 + 1]; bzp = (R)base[i * 4 + 2]; }
    R o[3] = {bxp + ex * scale, byp + ey * scale, bzp + ez * scale};
    for (int a = 0; a < 3; a++) o[a] = std::min(std::max(o[a], k.clampMin[a]), k.clampMax[a]);
    out[i * 4] = o[0]; out[i * 4 + 1] = o[1]; out[i * 4 + 2] = o[2]; out[i * 4 + 3] = 1.0;
  }
}

// velFromPos (wgsl.js:72-83), xsph (85-123) and normals (250-288)
template <class R>
void xsph_k(const OrcParams& P, int n, const double* pred, const double* pos,
            const double* density, const uint32_t* cs, int normals, double* vel, double* corr,
            double* normal) {
  K<R> k(P);
  std::vector<R> v((size_t)n * 3);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++)
      v[(size_t)i * 3 + a] = ((R)pred[i * 4 + a] - (R)pos[i * 4 + a]) * k.invDt;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    const R px = (R)pred[i * 4], py = (R)pred[i * 4 + 1], pz = (R)pred[i * 4 + 2];
    const R vx = v[(size_t)i * 3], vy = v[(size_t)i * 3 + 1], vz = v[(size_t)i * 3 + 2];
    R ax = 0, ay = 0, az = 0, nx = 0, ny = 0, nz = 0;
    stencil(k, px, py, pz, cs, (const uint32_t*)nullptr,
      [&](uint32_t j) {
        if (j == (uint32_t)i) return;
        const R rx = px - (R)pred[j * 4], ry = py - (R)pred[j * 4 + 1], rz = pz - (R)pred[j * 4 + 2];
        const R r2 = r2_of(rx, ry, rz);
        if (r2 >= k.h2) return;
        const R den = std::max((R)density[j], (R)0.5 * k.rho0);
        const R f = div_r(k.mass * poly6(k, r2), den);
        ax = ax + f * (v[(size_t)j * 3] - vx);
        ay = ay + f * (v[(size_t)j * 3 + 1] - vy);
        az = az + f * (v[(size_t)j * 3 + 2] - vz);
        if (normals && r2 >= (R)1.0e-12) {
          const R r = sqrt_r(r2);
          const R hr = k.h - r;
          const R g = div_r(k.mass, den) * div_r(k.spiky * hr * hr, r);
          nx = nx + g * rx; ny = ny + g * ry; nz = nz + g * rz;
        }
      },
      [&](uint32_t) {});
    if (vel) {
      vel[i * 4] = vx; vel[i * 4 + 1] = vy; vel[i * 4 + 2] = vz; vel[i * 4 + 3] = 0;
    }
    corr[i * 4] = k.xsphC * ax; corr[i * 4 + 1] = k.xsphC * ay; corr[i * 4 + 2] = k.xsphC * az;
    corr[i * 4 + 3] = 0;
    if (normals) {
      normal[i * 4] = k.h * nx; normal[i * 4 + 1] = k.h * ny; normal[i * 4 + 2] = k.h * nz;
      normal[i * 4 + 3] = 0;
    }
  }
}

// tension (wgsl.js:290-335); corr is updated in place
template <class R>
void tension_k(const OrcParams& P, int n, const double* pred, const double* density,
               const double* normal, const uint32_t* cs, double* corr) {
  K<R> k(P);
  std::vector<R> add((size_t)n * 3);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    const R px = (R)pred[i * 4], py = (R)pred[i * 4 + 1], pz = (R)pred[i * 4 + 2];
    const R nix = (R)normal[i * 4], niy = (R)normal[i * 4 + 1], niz = (R)normal[i * 4 + 2];
    const R rhoi = (R)density[i];
    R ax = 0, ay = 0, az = 0;
    stencil(k, px, py, pz, cs, (const uint32_t*)nullptr,
      [&](uint32_t j) {
        if (j == (uint32_t)i) return;
        const R rx = px - (R)pred[j * 4], ry = py - (R)pred[j * 4 + 1], rz = pz - (R)pred[j * 4 + 2];
        const R r2 = r2_of(rx, ry, rz);
        if (r2 >= k.h2 || r2 < (R)1.0e-12) return;
        const R r = sqrt_r(r2);
        const R hrr = k.h - r;
        const R kk = hrr * hrr * hrr * r * r * r;
        R cw = k.cohes * ((R)2 * kk - k.cohesTerm);
        if ((R)2 * r > k.h) cw = k.cohes * kk;
        const R fc = div_r(-k.tension * k.mass * cw, r);
        const R fx = fc * rx + (-k.tension * (nix - (R)normal[j * 4]));
        const R fy = fc * ry + (-k.tension * (niy - (R)normal[j * 4 + 1]));
        const R fz = fc * rz + (-k.tension * (niz - (R)normal[j * 4 + 2]));
        const R KK = div_r((R)2 * k.rho0, rhoi + (R)density[j]);
        ax = ax + KK * fx; ay = ay + KK * fy; az = az + KK * fz;
      },
      [&](uint32_t) {});
    add[(size_t)i * 3] = ax; add[(size_t)i * 3 + 1] = ay; add[(size_t)i * 3 + 2] = az;
  }
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) corr[i * 4 + a] = (R)corr[i * 4 + a] + k.dt * add[(size_t)i * 3 + a];
}

// finalize (wgsl.js:337-350): vel += corr; pos = pred
template <class R>
void finalize_k(int n, const double* pred, const double* corr, double* vel, double* pos) {
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) vel[i * 4 + a] = (R)vel[i * 4 + a] + (R)corr[i * 4 + a];
    vel[i * 4 + 3] = 0;
    for (int a = 0; a < 4; a++) pos[i * 4 + a] = pred[i * 4 + a];
  }
}

// FP64 polar factor of A (row-major) by the reference's iteration
// (wgsl.js:727-760).
void polar(double A[3][3], double R[3][3]) {
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  if (std::fabs(det) < 1.0e-12) {
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) A[r][c] = r == c ? 1.0 : 0.0;
    det = 1.0;
  }
  const double s = std::cbrt(std::fabs(det));
  if (s > 1.0e-6)
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) A[r][c] *= 1.0 / s;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) R[r][c] = A[r][c];
  for (int it = 0; it < 8; it++) {
    // inverse by cross products of the columns (mat3x3fInverse)
    double a[3] = {R[0][0], R[1][0], R[2][0]}, b[3] = {R[0][1], R[1][1], R[2][1]},
           c[3] = {R[0][2], R[1][2], R[2][2]};
    auto cross = [](const double* u, const double* w, double* o) {
      o[0] = u[1] * w[2] - u[2] * w[1];
      o[1] = u[2] * w[0] - u[0] * w[2];
      o[2] = u[0] * w[1] - u[1] * w[0];
    };
    double r0[3], r1[3], r2[3];
    cross(b, c, r0); cross(c, a, r1); cross(a, b, r2);
    const double d = a[0] * r0[0] + a[1] * r0[1] + a[2] * r0[2];
    const double invd = std::fabs(d) > 1.0e-20 ? 1.0 / d : 0.0;
    // inverse rows are r0, r1, r2; so inverse^T has columns r0, r1, r2
    double IT[3][3];
    for (int q = 0; q < 3; q++) {
      IT[q][0] = r0[q] * invd; IT[q][1] = r1[q] * invd; IT[q][2] = r2[q] * invd;
    }
    double diff = 0;
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++) {
        const double nx = (R[r][cc] + IT[r][cc]) * 0.5;
        diff += (nx - R[r][cc]) * (nx - R[r][cc]);
        R[r][cc] = nx;
      }
    if (diff < 1.0e-12) break;
  }
}

// Anisotropic kernel (aniso_wgsl.js:3-159).  ap = {lambda, support, radius,
// cells, kr, kn, ks (already divided by 0.15 R^2), stretch, minExtent, neps}.
template <class R>
void jacobi_ref(R a[3][3], R v[3][3]) {  // aniso_wgsl.js:30-55, row-major
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) v[i][j] = i == j ? (R)1 : (R)0;
  for (int sweep = 0; sweep < 6; sweep++)
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        const R apq = a[p][q];
        if (std::fabs(apq) < (R)1.0e-12) continue;
        const R theta = (a[q][q] - a[p][p]) / ((R)2 * apq);
        const R sg = theta > 0 ? (R)1 : (theta < 0 ? (R)-1 : (R)0);
        R t = sg / (std::fabs(theta) + sqrt_r(theta * theta + (R)1));
        if (theta == (R)0) t = (R)1;
        const R c = (R)1 / sqrt_r(t * t + (R)1), s = t * c;
        R r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        r[p][p] = c; r[q][q] = c; r[p][q] = s; r[q][p] = -s;
        R m[3][3], nn[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) m[i][j] = r[0][i] * a[0][j] + r[1][i] * a[1][j] + r[2][i] * a[2][j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) a[i][j] = m[i][0] * r[0][j] + m[i][1] * r[1][j] + m[i][2] * r[2][j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) nn[i][j] = v[i][0] * r[0][j] + v[i][1] * r[1][j] + v[i][2] * r[2][j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) v[i][j] = nn[i][j];
      }
}

// Targets sel[0..nsel) (all n when sel is null); outputs are indexed like
// the targets, neighbours come from all n particles.
template <class R>
void aniso_k(const OrcParams& P, int n, const double* pos, const uint32_t* cs, const double* ap,
             double* smooth, double* g, const int* sel = nullptr, int nsel = 0) {
  K<R> k(P);
  const R lambda = (R)ap[0], support = (R)ap[1], radius = (R)ap[2];
  const int cells = (int)ap[3];
  const R kr = (R)ap[4], kn = (R)ap[5], ks = (R)ap[6], stretch = (R)ap[7], minExt = (R)ap[8];
  const int neps = (int)ap[9];
  const R invSupport = (R)(1.0 / ap[1]);
  const int nt = sel ? nsel : n;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int o = 0; o < nt; o++) {
    const int src = sel ? sel[o] : o;
    const int i = o;  // output slot
    const R xi[3] = {(R)pos[src * 4], (R)pos[src * 4 + 1], (R)pos[src * 4 + 2]};
    int c[3];
    for (int a = 0; a < 3; a++) {
      R f = std::floor((xi[a] - k.boxMin[a]) * invSupport);
      c[a] = (int)std::fmin(std::fmax(f, (R)0), (R)(k.g[a] - 1));
    }
    const int x0 = std::max(c[0] - cells, 0), x1 = std::min(c[0] + cells, k.g[0] - 1);
    const R r2max = radius * radius, invR = (R)1 / radius;
    const R rel[3] = {xi[0] - k.boxMin[0], xi[1] - k.boxMin[1], xi[2] - k.boxMin[2]};
    R sd[3] = {0, 0, 0}, m2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, wsum = 0;
    int cnt = 0;
    for (int dz = -cells; dz <= cells; dz++) {
      const int z = c[2] + dz;
      if (z < 0 || z >= k.g[2]) continue;
      const R dzm = std::fmax(std::fmax((R)z * support - rel[2], rel[2] - (R)(z + 1) * support), (R)0);
      const R dz2 = dzm * dzm;
      if (dz2 >= r2max) continue;
      for (int dy = -cells; dy <= cells; dy++) {
        const int y = c[1] + dy;
        if (y < 0 || y >= k.g[1]) continue;
        const R dym = std::fmax(std::fmax((R)y * support - rel[1], rel[1] - (R)(y + 1) * support), (R)0);
        if (dz2 + dym * dym >= r2max) continue;
        const uint32_t rb = (uint32_t)((z * k.g[1] + y) * k.g[0]);
        for (uint32_t j = cs[rb + x0]; j < cs[rb + x1 + 1]; j++) {
          const R d[3] = {(R)pos[j * 4] - xi[0], (R)pos[j * 4 + 1] - xi[1], (R)pos[j * 4 + 2] - xi[2]};
          const R r2 = r2_of(d[0], d[1], d[2]);
          if (r2 >= r2max) continue;
          const R t = sqrt_r(r2) * invR;
          const R w = (R)1 - t * t * t;
          for (int a = 0; a < 3; a++) sd[a] = sd[a] + d[a] * w;
          for (int col = 0; col < 3; col++)
            for (int row = 0; row < 3; row++) m2[row][col] = m2[row][col] + d[row] * (d[col] * w);
          wsum = wsum + w;
          cnt++;
        }
      }
    }
    if (wsum <= (R)0) {
      const R iso = kn / support;
      for (int a = 0; a < 3; a++) smooth[i * 4 + a] = xi[a];
      smooth[i * 4 + 3] = 0;
      g[i * 8 + 0] = iso; g[i * 8 + 1] = 0; g[i * 8 + 2] = 0; g[i * 8 + 3] = iso;
      g[i * 8 + 4] = 0; g[i * 8 + 5] = iso; g[i * 8 + 6] = iso * iso * iso; g[i * 8 + 7] = 0;
      continue;
    }
    R db[3];
    for (int a = 0; a < 3; a++) {
      db[a] = sd[a] / wsum;
      smooth[i * 4 + a] = xi[a] + db[a] * lambda;
    }
    smooth[i * 4 + 3] = 0;
    R A[3][3], V[3][3];
    for (int row = 0; row < 3; row++)
      for (int col = 0; col < 3; col++) A[row][col] = m2[row][col] / wsum - db[row] * db[col];
    jacobi_ref(A, V);
    const R sig[3] = {A[0][0], A[1][1], A[2][2]};
    const R s1 = std::fmax(sig[0], std::fmax(sig[1], sig[2]));
    R ext[3], inv[3];
    for (int a = 0; a < 3; a++) ext[a] = ks * std::fmax(sig[a], s1 / kr);
    const R dv = std::fmax((R)1, std::fmax(ext[0], std::fmax(ext[1], ext[2])) / stretch);
    for (int a = 0; a < 3; a++) {
      ext[a] = std::fmax(ext[a] / dv, minExt);
      inv[a] = (R)1 / (support * ext[a]);
    }
    const R e0 = (R)neps * (R)0.4, e1 = (R)neps * (R)1.6;
    R w = std::fmin(std::fmax(((R)cnt - e0) / (e1 - e0), (R)0), (R)1);
    w = w * w * ((R)3 - (R)2 * w);
    const R iso = kn / support;
    R G[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        const R an = V[a][0] * inv[0] * V[b][0] + V[a][1] * inv[1] * V[b][1] + V[a][2] * inv[2] * V[b][2];
        G[a][b] = (a == b ? iso : (R)0) * ((R)1 - w) + an * w;
      }
    const R det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                  G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                  G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    g[i * 8 + 0] = G[0][0]; g[i * 8 + 1] = G[0][1]; g[i * 8 + 2] = G[0][2]; g[i * 8 + 3] = G[1][1];
    g[i * 8 + 4] = G[1][2]; g[i * 8 + 5] = G[2][2]; g[i * 8 + 6] = det; g[i * 8 + 7] = cnt;
  }
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

void orc_aniso(int mode, const OrcParams* P, int n, const double* pos, const uint32_t* cs,
               const double* ap, double* smooth, double* g) {
  if (float_mode(mode)) aniso_k<float>(*P, n, pos, cs, ap, smooth, g);
  else aniso_k<double>(*P, n, pos, cs, ap, smooth, g);
}

// The same for the targets sel[0..nsel) only (outputs nsel records).
void orc_aniso_sel(int mode, const OrcParams* P, int n, const double* pos, const uint32_t* cs,
                   const double* ap, int nsel, const int* sel, double* smooth, double* g) {
  if (float_mode(mode)) aniso_k<float>(*P, n, pos, cs, ap, smooth, g, sel, nsel);
  else aniso_k<double>(*P, n, pos, cs, ap, smooth, g, sel, nsel);
}

// The Jacobi eigensolve of the anisotropy pass on one symmetric matrix
// (row-major a[9]); eigenvalues to w[3], eigenvectors as columns of v[9].
void orc_jacobi(int mode, const double* a9, double* w, double* v9) {
  if (float_mode(mode)) {
    float a[3][3], v[3][3];
    for (int k = 0; k < 9; k++) a[k / 3][k % 3] = (float)a9[k];
    jacobi_ref(a, v);
    for (int k = 0; k < 3; k++) w[k] = a[k][k];
    for (int k = 0; k < 9; k++) v9[k] = v[k / 3][k % 3];
  } else {
    double a[3][3], v[3][3];
    for (int k = 0; k < 9; k++) a[k / 3][k % 3] = a9[k];
    jacobi_ref(a, v);
    for (int k = 0; k < 3; k++) w[k] = a[k][k];
    for (int k = 0; k < 9; k++) v9[k] = v[k / 3][k % 3];
  }
}

// cell id of every particle (wgsl.js:43-49, as count/scatterSlot use it)
void orc_cell_ids(int mode, const OrcParams* P, int n, const double* pred, uint32_t* key) {
  for (int i = 0; i < n; i++) {
    int c[3];
    if (float_mode(mode)) {
      K<float> k(*P);
      cell_of(k, (float)pred[i * 4], (float)pred[i * 4 + 1], (float)pred[i * 4 + 2], c);
      key[i] = cell_index(k, c);
    } else {
      K<double> k(*P);
      cell_of(k, pred[i * 4], pred[i * 4 + 1], pred[i * 4 + 2], c);
      key[i] = cell_index(k, c);
    }
  }
}

// Stable counting sort by cell (count, scan, scatter; wgsl.js:352-438 with
// the within-cell order fixed to the input order).  cellStart has
// nCells+1 entries; slot[i] = sorted position of input i; order[s] = input
// index at sorted position s.
void orc_grid(int n, int nCells, const uint32_t* key, uint32_t* cellStart, uint32_t* slot,
              uint32_t* order) {
  std::vector<uint32_t> cnt((size_t)nCells + 1, 0);
  for (int i = 0; i < n; i++) cnt[key[i]]++;
  uint32_t run = 0;
  for (int c = 0; c < nCells; c++) {
    cellStart[c] = run;
    run += cnt[c];
  }
  cellStart[nCells] = run;
  std::vector<uint32_t> cur(cellStart, cellStart + nCells);
  for (int i = 0; i < n; i++) {
    const uint32_t s = cur[key[i]]++;
    slot[i] = s;
    order[s] = (uint32_t)i;
  }
}

void orc_predict(int mode, const OrcParams* P, int n, const double* pos, const double* vel,
                 double* pred, double* velOut) {
  if (float_mode(mode)) predict<float>(*P, n, pos, vel, pred, velOut);
  else predict<double>(*P, n, pos, vel, pred, velOut);
}

void orc_lambda(int mode, const OrcParams* P, int n, const double* pred, const uint32_t* cs,
                const double* bpos, const double* bpsi, const uint32_t* bcs, double* density,
                double* lambda) {
  if (float_mode(mode)) lambda_k<float>(*P, n, pred, cs, bpos, bpsi, bcs, density, lambda);
  else lambda_k<double>(*P, n, pred, cs, bpos, bpsi, bcs, density, lambda);
}

void orc_delta(int mode, const OrcParams* P, int n, const double* pred, const double* lambda,
               const uint32_t* cs, const double* bpos, const double* bpsi, const uint32_t* bcs,
               double* out) {
  if (float_mode(mode)) delta_k<float>(*P, n, pred, lambda, cs, bpos, bpsi, bcs, out, nullptr);
  else delta_k<double>(*P, n, pred, lambda, cs, bpos, bpsi, bcs, out, nullptr);
}

// delta evaluated at positions `pred` and applied to positions `base`:
// out = clamp(base + correction(pred)).  Restates the FP16-offset form,
// which evaluates the correction at quantized positions and moves the FP32
// master positions.
void orc_delta_at(int mode, const OrcParams* P, int n, const double* pred, const double* base,
                  const double* lambda, const uint32_t* cs, const double* bpos,
                  const double* bpsi, const uint32_t* bcs, double* out) {
  if (float_mode(mode)) delta_k<float>(*P, n, pred, lambda, cs, bpos, bpsi, bcs, out, base);
  else delta_k<double>(*P, n, pred, lambda, cs, bpos, bpsi, bcs, out, base);
}

void orc_xsph(int mode, const OrcParams* P, int n, const double* pred, const double* pos,
              const double* density, const uint32_t* cs, int normals, double* vel, double* corr,
              double* normal) {
  if (float_mode(mode)) xsph_k<float>(*P, n, pred, pos, density, cs, normals, vel, corr, normal);
  else xsph_k<double>(*P, n, pred, pos, density, cs, normals, vel, corr, normal);
}

void orc_tension(int mode, const OrcParams* P, int n, const double* pred, const double* density,
                 const double* normal, const uint32_t* cs, double* corr) {
  if (float_mode(mode)) tension_k<float>(*P, n, pred, density, normal, cs, corr);
  else tension_k<double>(*P, n, pred, density, normal, cs, corr);
}

void orc_finalize(int mode, int n, const double* pred, const double* corr, double* vel,
                  double* pos) {
  if (float_mode(mode)) finalize_k<float>(n, pred, corr, vel, pos);
  else finalize_k<double>(n, pred, corr, vel, pos);
}

// Neighbour census: fluid j != i with r^2 < h^2 in the 27-cell stencil,
// count and index sums mod 2^32 (mode 0 arithmetic).
void orc_census(const OrcParams* P, int n, const double* pred, const uint32_t* cs,
                uint32_t* cnt, uint32_t* s1, uint32_t* s2) {
  K<float> k(*P);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    const float px = (float)pred[i * 4], py = (float)pred[i * 4 + 1], pz = (float)pred[i * 4 + 2];
    uint32_t c = 0, a = 0, b = 0;
    stencil(k, px, py, pz, cs, (const uint32_t*)nullptr,
      [&](uint32_t j) {
        if (j == (uint32_t)i) return;
        const float rx = px - (float)pred[j * 4], ry = py - (float)pred[j * 4 + 1],
                    rz = pz - (float)pred[j * 4 + 2];
        if (r2_of(rx, ry, rz) >= k.h2) return;
        c += 1u; a += j; b += j * j;
      },
      [&](uint32_t) {});
    cnt[i] = c; s1[i] = a; s2[i] = b;
  }
}

// Body centroids before the solve (wgsl.js:628-668, 174-185): centre_b =
// mean of pred over the body's particles.  bodyOf[i] is 1-based, 0 =
// fluid.
void orc_body_pre(int n, const double* pred, const uint32_t* bodyOf, int nBodies,
                  double* centre) {
  std::vector<double> s((size_t)nBodies * 3, 0.0), c(nBodies, 0.0);
  for (int i = 0; i < n; i++) {
    const uint32_t b = bodyOf[i];
    if (!b) continue;
    for (int a = 0; a < 3; a++) s[(b - 1) * 3 + a] += pred[i * 4 + a] - centre[(b - 1) * 3 + a];
    c[b - 1] += 1;
  }
  for (int b = 0; b < nBodies; b++)
    if (c[b] > 0)
      for (int a = 0; a < 3; a++) centre[b * 3 + a] += s[b * 3 + a] / c[b];
}

// One shape-matching round (wgsl.js:628-783) in FP64: bodyCentre pass,
// bodySeed (centroid), bodyCov about the centroid, bodyResolve (centre +=
// mean shift * invRatio; polar factor of A), bodyProject in float:
// pred = centre + R q (rows of R and centre rounded to float).
// centre (3/body) and rot (9/body, row-major) are updated in place.
void orc_body_solve(int n, double* pred, const uint32_t* bodyOf, const double* rest,
                    int nBodies, const double* invRatio, double* centre, double* rot) {
  std::vector<double> s((size_t)nBodies * 3, 0.0), cnt(nBodies, 0.0);
  for (int i = 0; i < n; i++) {
    const uint32_t b = bodyOf[i];
    if (!b) continue;
    for (int a = 0; a < 3; a++) s[(b - 1) * 3 + a] += pred[i * 4 + a] - centre[(b - 1) * 3 + a];
    cnt[b - 1] += 1;
  }
  std::vector<double> ref((size_t)nBodies * 3), A((size_t)nBodies * 9, 0.0);
  for (int b = 0; b < nBodies; b++)
    for (int a = 0; a < 3; a++)
      ref[b * 3 + a] = cnt[b] > 0 ? centre[b * 3 + a] + s[b * 3 + a] / cnt[b] : centre[b * 3 + a];
  for (int i = 0; i < n; i++) {
    const uint32_t b = bodyOf[i];
    if (!b) continue;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++)
        A[(b - 1) * 9 + r * 3 + c] += (pred[i * 4 + r] - ref[(b - 1) * 3 + r]) * rest[i * 4 + c];
  }
  std::vector<float> cf((size_t)nBodies * 3), rf((size_t)nBodies * 9);
  for (int b = 0; b < nBodies; b++) {
    if (cnt[b] <= 0) continue;
    for (int a = 0; a < 3; a++) {
      double d = s[b * 3 + a] / cnt[b];
      if (invRatio[b] > 0) d *= invRatio[b];
      centre[b * 3 + a] += d;
    }
    double M[3][3], R[3][3];
    for (int k = 0; k < 9; k++) M[k / 3][k % 3] = A[b * 9 + k];
    polar(M, R);
    for (int k = 0; k < 9; k++) rot[b * 9 + k] = R[k / 3][k % 3];
    for (int a = 0; a < 3; a++) cf[b * 3 + a] = (float)centre[b * 3 + a];
    for (int k = 0; k < 9; k++) rf[b * 9 + k] = (float)R[k / 3][k % 3];
  }
  for (int i = 0; i < n; i++) {
    const uint32_t b = bodyOf[i];
    if (!b || cnt[b - 1] <= 0) continue;
    const float* c = &cf[(b - 1) * 3];
    const float* R = &rf[(b - 1) * 9];
    const float qx = (float)rest[i * 4], qy = (float)rest[i * 4 + 1], qz = (float)rest[i * 4 + 2];
    const float px = R[0] * qx + R[1] * qy + R[2] * qz;
    const float py = R[3] * qx + R[4] * qy + R[5] * qz;
    const float pz = R[6] * qx + R[7] * qy + R[8] * qz;
    pred[i * 4] = c[0] + px; pred[i * 4 + 1] = c[1] + py; pred[i * 4 + 2] = c[2] + pz;
    pred[i * 4 + 3] = 1.0;
  }
}

// stats (wgsl.js:785-832) reduced exactly: sum rho/rho0 / n, max rho/rho0,
// max |v|, sum 0.5 m |v|^2, per-particle terms in float.
void orc_stats(int n, const double* density, const double* vel, double invRho0, double mass,
               double* out) {
  const float ir = (float)invRho0, m = (float)mass;
  double sr = 0, mr = 0, mv = 0, ke = 0;
  for (int i = 0; i < n; i++) {
    const float rho = (float)density[i] * ir;
    const float vx = (float)vel[i * 4], vy = (float)vel[i * 4 + 1], vz = (float)vel[i * 4 + 2];
    const float v2 = vx * vx + vy * vy + vz * vz;
    sr += rho;
    mr = std::max(mr, (double)rho);
    mv = std::max(mv, (double)sqrtf(v2));
    ke += (double)(0.5f * m * v2);
  }
  out[0] = sr / std::max(1, n);
  out[1] = mr; out[2] = mv; out[3] = ke;
}

// Density field of SurfaceMesh.generateField: meshDensityWGSL
// (mesh_wgsl.js:23-69) in float with its operation order.  Vertex v sits at
// origin + voxel * v; the stencil is the 27 cells of the simulation grid
// around floor(p * invSupport) (boxMin = 0 as setParams writes it,
// mesh.js:137); per candidate j: skipped if a body particle and skipBodies;
// vol = mass / max(rho_j, 0.5 rho0); r2 = |p - pos_j|^2; if r2 < h^2:
// acc += vol * poly6 * t * t * t with t = h^2 - r2.
void orc_field_iso(const OrcParams* P, const int32_t* vertDim, double voxel,
                   const double* origin, double mass, double rho0, int n, const double* pos,
                   const double* density, const uint32_t* body, int skipBodies,
                   const uint32_t* cs, float* field) {
  const int gx = P->gridDim[0], gy = P->gridDim[1], gz = P->gridDim[2];
  const float vox = (float)voxel, ox = (float)origin[0], oy = (float)origin[1],
              oz = (float)origin[2];
  const float invS = (float)(1.0 / P->h), s2 = (float)(P->h * P->h), poly6 = (float)P->poly6;
  const float m = (float)mass, r0 = (float)rho0;
  const long nv = (long)vertDim[0] * vertDim[1] * vertDim[2];
  (void)n;
#pragma omp parallel for schedule(dynamic, 4096)
  for (long idx = 0; idx < nv; idx++) {
    const int vx = (int)(idx % vertDim[0]);
    const int vy = (int)((idx / vertDim[0]) % vertDim[1]);
    const int vz = (int)(idx / ((long)vertDim[0] * vertDim[1]));
    const float px = ox + vox * (float)vx, py = oy + vox * (float)vy, pz = oz + vox * (float)vz;
    const int cx = std::min(std::max((int)std::floor((px - (float)P->boxMin[0]) * invS), 0), gx - 1);
    const int cy = std::min(std::max((int)std::floor((py - (float)P->boxMin[1]) * invS), 0), gy - 1);
    const int cz = std::min(std::max((int)std::floor((pz - (float)P->boxMin[2]) * invS), 0), gz - 1);
    const int x0 = std::max(cx - 1, 0), x1 = std::min(cx + 1, gx - 1);
    float acc = 0.0f;
    for (int dz = -1; dz <= 1; dz++) {
      const int z = cz + dz;
      if (z < 0 || z >= gz) continue;
      for (int dy = -1; dy <= 1; dy++) {
        const int y = cy + dy;
        if (y < 0 || y >= gy) continue;
        const uint32_t rb = (uint32_t)((z * gy + y) * gx);
        for (uint32_t j = cs[rb + x0]; j < cs[rb + x1 + 1]; j++) {
          if (skipBodies && body[j] != 0u) continue;
          const float vol = m / std::max((float)density[j], 0.5f * r0);
          const float rx = px - (float)pos[j * 4], ry = py - (float)pos[j * 4 + 1],
                      rz = pz - (float)pos[j * 4 + 2];
          const float r2 = std::fmaf(rz, rz, std::fmaf(ry, ry, rx * rx));
          if (r2 >= s2) continue;
          const float t = s2 - r2;
          acc += vol * poly6 * t * t * t;
        }
      }
    }
    field[idx] = acc;
  }
}

// Anisotropic (Yu-Turk) field of this build, restated from its definition
// with the kernel's operation order: per particle M = G G, w = vol *
// (315/(64 pi)) * det G (0 for a skipped body particle), bounding box
// half-extents sqrt((M^-1)_kk) * 1.001 + 1e-6 (FP64); per vertex in the box
// with s2 = r^T M r < 1 the term w (1 - s2)^3 rounded to a multiple of
// 2^-40, summed exactly in 64-bit integers.
void orc_field_aniso(const int32_t* vertDim, double voxel, const double* origin, double mass,
                     double rho0, int n, const double* smoothPos, const double* g,
                     const double* density, const uint32_t* body, int skipBodies,
                     float* field) {
  const long nv = (long)vertDim[0] * vertDim[1] * vertDim[2];
  std::vector<long long> acc(nv, 0);
  const float vox = (float)voxel, m = (float)mass, r0 = (float)rho0;
  const float org[3] = {(float)origin[0], (float)origin[1], (float)origin[2]};
  const double scale = 1099511627776.0;  // 2^40
#pragma omp parallel for schedule(dynamic, 256)
  for (int j = 0; j < n; j++) {
    const float a = (float)g[8 * j], b = (float)g[8 * j + 1], c = (float)g[8 * j + 2],
                d = (float)g[8 * j + 3], e = (float)g[8 * j + 4], f = (float)g[8 * j + 5],
                detG = (float)g[8 * j + 6];
    const float m00 = a * a + b * b + c * c;
    const float m11 = b * b + d * d + e * e;
    const float m22 = c * c + e * e + f * f;
    const float m01 = a * b + b * d + c * e;
    const float m02 = a * c + b * e + c * f;
    const float m12 = b * c + d * e + e * f;
    const float vol = m / std::max((float)density[j], 0.5f * r0);
    float w = vol * (1.5666814710608448f * detG);
    if (skipBodies && body[j] != 0u) w = 0.0f;
    const double A00 = m00, A11 = m11, A22 = m22, A01 = m01, A02 = m02, A12 = m12;
    const double c00 = A11 * A22 - A12 * A12, c11 = A00 * A22 - A02 * A02,
                 c22 = A00 * A11 - A01 * A01;
    const double det = A00 * c00 - A01 * (A01 * A22 - A12 * A02) + A02 * (A01 * A12 - A11 * A02);
    float ext[3] = {0.0f, 0.0f, 0.0f};
    if (det > 0.0 && w > 0.0f) {
      ext[0] = (float)(std::sqrt(std::max(c00 / det, 0.0)) * 1.001 + 1e-6);
      ext[1] = (float)(std::sqrt(std::max(c11 / det, 0.0)) * 1.001 + 1e-6);
      ext[2] = (float)(std::sqrt(std::max(c22 / det, 0.0)) * 1.001 + 1e-6);
    } else {
      w = 0.0f;
    }
    if (!(w > 0.0f)) continue;
    const float ctr[3] = {(float)smoothPos[4 * j], (float)smoothPos[4 * j + 1],
                          (float)smoothPos[4 * j + 2]};
    bool finite = true;
    for (int k = 0; k < 3; k++) finite = finite && std::isfinite(ctr[k]) && std::isfinite(ext[k]);
    if (!finite) continue;
    const float inv = 1.0f / vox;
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      const float aa = (ctr[k] - ext[k] - org[k]) * inv, bb = (ctr[k] + ext[k] - org[k]) * inv;
      lo[k] = std::max((int)std::ceil(std::max(aa, -1.0f)), 0);
      hi[k] = std::min((int)std::floor(std::min(bb, (float)vertDim[k])), vertDim[k] - 1);
    }
    for (int z = lo[2]; z <= hi[2]; z++) {
      const float rz = (org[2] + vox * (float)z) - ctr[2];
      for (int y = lo[1]; y <= hi[1]; y++) {
        const float ry = (org[1] + vox * (float)y) - ctr[1];
        const long rowBase = ((long)z * vertDim[1] + y) * vertDim[0];
        for (int x = lo[0]; x <= hi[0]; x++) {
          const float rx = (org[0] + vox * (float)x) - ctr[0];
          const float s2 = m00 * (rx * rx) + m11 * (ry * ry) + m22 * (rz * rz) +
                           2.0f * (m01 * (rx * ry) + m02 * (rx * rz) + m12 * (ry * rz));
          if (!(s2 < 1.0f)) continue;
          const float u = 1.0f - s2;
          const float term = ((w * u) * u) * u;
          const long long q = std::llrint((double)term * scale);
#pragma omp atomic
          acc[rowBase + x] += q;
        }
      }
    }
  }
  for (long i = 0; i < nv; i++) field[i] = (float)((double)acc[i] * (1.0 / scale));
}

// One axis step of meshSmoothWGSL (mesh_wgsl.js:70-91): dst(v) = 0.25 S(v -
// axis) + 0.5 S(v) + 0.25 S(v + axis), S clamping the vertex into the
// lattice, evaluated left to right in float.
static void mesh_blur_axis(const int32_t* vd, int axis, const float* src, float* dst) {
  const long sx = 1, sy = vd[0], sz = (long)vd[0] * vd[1];
  const long stride = axis == 0 ? sx : axis == 1 ? sy : sz;
  for (int z = 0; z < vd[2]; z++)
    for (int y = 0; y < vd[1]; y++)
      for (int x = 0; x < vd[0]; x++) {
        const int c = axis == 0 ? x : axis == 1 ? y : z;
        const long i = x * sx + y * sy + z * sz;
        const long lo = c > 0 ? i - stride : i;
        const long hi = c < vd[axis] - 1 ? i + stride : i;
        const float a = 0.25f * src[lo];
        const float b = 0.5f * src[i];
        const float t = a + b;
        const float d = 0.25f * src[hi];
        dst[i] = t + d;
      }
}

// SurfaceMesh.smoothField / buildNormalField (mesh.js:273-313): `passes`
// rounds of the x, y and z steps applied to src; the result in dst.
void orc_mesh_smooth(const int32_t* vertDim, int passes, const float* src, float* dst) {
  const long nv = (long)vertDim[0] * vertDim[1] * vertDim[2];
  std::vector<float> a(src, src + nv), b(nv);
  for (int p = 0; p < passes; p++)
    for (int axis = 0; axis < 3; axis++) {
      mesh_blur_axis(vertDim, axis, a.data(), b.data());
      a.swap(b);
    }
  std::memcpy(dst, a.data(), nv * sizeof(float));
}

namespace {

struct MCorner {
  float d;
  float p[3];
  float g[3];
};

// packVertex (mesh_wgsl.js:112-124): grid position quantised to 16 bits per
// axis over [0, vertDim-1]; normal octahedral-encoded to 8+8 bits.
void mesh_pack(const int32_t* vd, const float* pg, const float* nIn, uint32_t* out) {
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    const float pn = std::fmin(std::fmax(pg[k] / (float)(vd[k] - 1), 0.0f), 1.0f);
    q[k] = (uint32_t)(pn * 65535.0f + 0.5f);
  }
  const float l1 = std::fmax(std::fabs(nIn[0]) + std::fabs(nIn[1]) + std::fabs(nIn[2]), 1.0e-12f);
  const float nx = nIn[0] / l1, ny = nIn[1] / l1, nz = nIn[2] / l1;
  float e[2] = {nx, ny};
  if (nz < 0.0f) {
    e[0] = (1.0f - std::fabs(ny)) * (nx >= 0.0f ? 1.0f : -1.0f);
    e[1] = (1.0f - std::fabs(nx)) * (ny >= 0.0f ? 1.0f : -1.0f);
  }
  uint32_t eo[2];
  for (int k = 0; k < 2; k++)
    eo[k] = (uint32_t)(std::fmin(std::fmax(e[k] * 0.5f + 0.5f, 0.0f), 1.0f) * 255.0f + 0.5f);
  out[0] = q[0] | (q[1] << 16);
  out[1] = q[2] | ((eo[0] | (eo[1] << 8)) << 16);
}

// edgeVertex (mesh_wgsl.js:126-132): t = (iso - da) / (db - da); position
// and gradient by mix(a, b, t) = a * (1 - t) + b * t; the normal is -g.
void mesh_edge(const int32_t* vd, float iso, const MCorner& a, const MCorner& b, uint32_t* out) {
  const float t = (iso - a.d) / (b.d - a.d);
  const float u = 1.0f - t;
  float p[3], n[3];
  for (int k = 0; k < 3; k++) {
    p[k] = a.p[k] * u + b.p[k] * t;
    n[k] = -(a.g[k] * u + b.g[k] * t);
  }
  mesh_pack(vd, p, n, out);
}

struct MeshOut {
  const int32_t* vd;
  float iso;
  uint32_t maxTris;
  uint32_t* verts;  // 6 words per triangle
  uint32_t count;
  void tri(const MCorner& a0, const MCorner& a1, const MCorner& b0, const MCorner& b1,
           const MCorner& c0, const MCorner& c1) {
    const uint32_t k = count++;
    if (k >= maxTris) return;
    mesh_edge(vd, iso, a0, a1, verts + 6 * (size_t)k);
    mesh_edge(vd, iso, b0, b1, verts + 6 * (size_t)k + 2);
    mesh_edge(vd, iso, c0, c1, verts + 6 * (size_t)k + 4);
  }
};

// emitTet (mesh_wgsl.js:166-188) with emitLone / emitQuad (143-164).
void mesh_tet(MeshOut& o, const MCorner& a, const MCorner& b, const MCorner& c,
              const MCorner& d) {
  const bool ia = a.d > o.iso, ib = b.d > o.iso, ic = c.d > o.iso, id = d.d > o.iso;
  const int nin = (int)ia + (int)ib + (int)ic + (int)id;
  if (nin == 0 || nin == 4) return;
  if (nin == 1 || nin == 3) {
    const bool look = nin == 1;
    const MCorner *l, *p1, *p2, *p3;
    if (ia == look) { l = &a; p1 = &b; p2 = &c; p3 = &d; }
    else if (ib == look) { l = &b; p1 = &a; p2 = &c; p3 = &d; }
    else if (ic == look) { l = &c; p1 = &a; p2 = &b; p3 = &d; }
    else { l = &d; p1 = &a; p2 = &b; p3 = &c; }
    o.tri(*l, *p1, *l, *p2, *l, *p3);
    return;
  }
  const MCorner *qa, *qb, *qc, *qd;
  if (ia && ib) { qa = &a; qb = &b; qc = &c; qd = &d; }
  else if (ia && ic) { qa = &a; qb = &c; qc = &b; qd = &d; }
  else if (ia && id) { qa = &a; qb = &d; qc = &b; qd = &c; }
  else if (ib && ic) { qa = &b; qb = &c; qc = &a; qd = &d; }
  else if (ib && id) { qa = &b; qb = &d; qc = &a; qd = &c; }
  else { qa = &c; qb = &d; qc = &a; qd = &b; }
  // q0 = (a,c), q1 = (a,d), q2 = (b,d), q3 = (b,c); triangles (q0,q1,q2), (q0,q2,q3)
  o.tri(*qa, *qc, *qa, *qd, *qb, *qd);
  o.tri(*qa, *qc, *qb, *qd, *qb, *qc);
}

}  // namespace

// meshMarchWGSL (mesh_wgsl.js:190-233): marching tetrahedra over the voxels
// of the lattice, six tetrahedra per voxel around the 0-7 diagonal, corners
// inside when d > iso, gradients by central differences of nfield (clamped
// at the lattice faces).  Triangles are numbered in voxel order (flat voxel
// index, then tetrahedron, then triangle); those numbered >= maxTris are
// dropped (emitTri, mesh_wgsl.js:134-141).  verts receives 6 words per kept
// triangle; returns the total count including the dropped ones.
uint32_t orc_mesh_march(const int32_t* vd, float iso, uint32_t maxTris, const float* field,
                        const float* nfield, uint32_t* verts) {
  const int nx = vd[0] - 1, ny = vd[1] - 1, nz = vd[2] - 1;
  auto at = [&](const float* f, int x, int y, int z) {
    x = std::min(std::max(x, 0), vd[0] - 1);
    y = std::min(std::max(y, 0), vd[1] - 1);
    z = std::min(std::max(z, 0), vd[2] - 1);
    return f[((long)z * vd[1] + y) * vd[0] + x];
  };
  MeshOut o{vd, iso, maxTris, verts, 0u};
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        MCorner c[8];
        int mask = 0;
        for (int k = 0; k < 8; k++) {
          c[k].d = at(field, x + off[k][0], y + off[k][1], z + off[k][2]);
          if (c[k].d > iso) mask |= 1 << k;
        }
        if (mask == 0 || mask == 255) continue;
        for (int k = 0; k < 8; k++) {
          const int X = x + off[k][0], Y = y + off[k][1], Z = z + off[k][2];
          c[k].p[0] = (float)X;
          c[k].p[1] = (float)Y;
          c[k].p[2] = (float)Z;
          c[k].g[0] = at(nfield, X + 1, Y, Z) - at(nfield, X - 1, Y, Z);
          c[k].g[1] = at(nfield, X, Y + 1, Z) - at(nfield, X, Y - 1, Z);
          c[k].g[2] = at(nfield, X, Y, Z + 1) - at(nfield, X, Y, Z - 1);
        }
        mesh_tet(o, c[0], c[1], c[3], c[7]);
        mesh_tet(o, c[0], c[1], c[5], c[7]);
        mesh_tet(o, c[0], c[2], c[3], c[7]);
        mesh_tet(o, c[0], c[2], c[6], c[7]);
        mesh_tet(o, c[0], c[4], c[5], c[7]);
        mesh_tet(o, c[0], c[4], c[6], c[7]);
      }
  return o.count;
}

}  // extern "C"
