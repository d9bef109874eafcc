// Host-side scene construction for libp4a.
//
// Restates the reference's CPU scene builder, src/scene.js:
//   defaultParams   scene.js:11-39
//   fluidBlock      scene.js:43-64   (LCG seed 12345, +-0.001d jitter)
//   boundaryParticles scene.js:76-135 (shell lattice, Akinci psi)
//   sampleBody      scene.js:137-157
//   buildScene      scene.js:159-315 (body placement, mass calibration,
//                                      denomRest)
// in FP64, the precision JavaScript evaluates it in, plus jittered-lattice
// scenes with explicit bodies for the synthetic workloads (dam breaks in
// lattice units, mass calibrated on the unjittered lattice).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/p4a.h"

struct p4a_scene_data {
  int32_t n = 0, n_fluid = 0, n_body = 0;
  std::vector<float> pos, vel, rest;
  std::vector<uint32_t> phase;
  std::vector<float> bpts, bpsi;
  std::vector<double> body_centre, body_density;
  std::vector<int32_t> body_count, body_spec;
  double mass = 0, h = 0, d = 0, denom_rest = 0;
};

namespace {

const double kPi = 3.14159265358979323846;  // Math.PI

double poly6_coef(double h) { return 315.0 / (64.0 * kPi * std::pow(h, 9)); }

// The reference's LCG, evaluated the way JavaScript does it: the product
// seed * 1103515245 is a double (rounded once past 2^53), `+ 12345` is a
// double add, and `& 0x7fffffff` applies ToInt32 (the integer value modulo
// 2^32) before masking (scene.js:52-56, main.js:100-104).
struct JsLcg {
  uint32_t seed;
  explicit JsLcg(uint32_t s) : seed(s) {}
  uint32_t next() {
    double v = (double)seed * 1103515245.0;
    v = v + 12345.0;
    uint64_t u = (uint64_t)v;  // v is a non-negative integer < 2^63
    seed = (uint32_t)(u & 0x7fffffffu);
    return seed;
  }
  // (seed / 0x7fffffff - 0.5) * 0.002 * d  (scene.js:55); amp2 = 0.002
  double jitter(double amp2, double d) {
    next();
    return ((double)seed / 2147483647.0 - 0.5) * amp2 * d;
  }
};

// Stable bucket lists of points in a dense grid: cell -> point indices in
// ascending order, which is the order of the reference's Map buckets.
struct Buckets {
  int dim[3];
  std::vector<uint32_t> start, items;
  void build(const std::vector<double>& xyz, size_t n, size_t stride,
             const int d3[3], double h) {
    dim[0] = d3[0]; dim[1] = d3[1]; dim[2] = d3[2];
    size_t cells = (size_t)dim[0] * dim[1] * dim[2];
    std::vector<uint32_t> cell(n);
    start.assign(cells + 1, 0);
    for (size_t i = 0; i < n; i++) {
      int c[3];
      for (int k = 0; k < 3; k++) {
        double f = std::floor(xyz[i * stride + k] / h);
        c[k] = (int)std::min((double)dim[k] - 1, std::max(0.0, f));
      }
      cell[i] = (uint32_t)(((size_t)c[2] * dim[1] + c[1]) * dim[0] + c[0]);
      start[cell[i] + 1]++;
    }
    for (size_t c = 0; c < cells; c++) start[c + 1] += start[c];
    items.resize(n);
    std::vector<uint32_t> cur(start.begin(), start.end() - 1);
    for (size_t i = 0; i < n; i++) items[cur[cell[i]]++] = (uint32_t)i;
  }
  void cell_of(const double* p, double h, int c[3]) const {
    for (int k = 0; k < 3; k++) {
      double f = std::floor(p[k] / h);
      c[k] = (int)std::min((double)dim[k] - 1, std::max(0.0, f));
    }
  }
};

// boundaryParticles (scene.js:76-135).
void boundary_particles(const double size[3], double d, double h, double rho0,
                        std::vector<double>& pts, std::vector<float>& psi) {
  int nb[3];
  double step[3];
  for (int k = 0; k < 3; k++) {
    nb[k] = std::max(2, (int)std::ceil(size[k] / d) + 1);
    step[k] = size[k] / (nb[k] - 1);
  }
  pts.clear();
  for (int ix = 0; ix < nb[0]; ix++)
    for (int iy = 0; iy < nb[1]; iy++)
      for (int iz = 0; iz < nb[2]; iz++) {
        bool shell = ix == 0 || ix == nb[0] - 1 || iy == 0 || iy == nb[1] - 1 ||
                     iz == 0 || iz == nb[2] - 1;
        if (shell) {
          pts.push_back(ix * step[0]);
          pts.push_back(iy * step[1]);
          pts.push_back(iz * step[2]);
        }
      }
  size_t n = pts.size() / 3;
  int dim[3];
  for (int k = 0; k < 3; k++) dim[k] = std::max(1, (int)std::ceil(size[k] / h));
  Buckets B;
  B.build(pts, n, 3, dim, h);
  const double coef = poly6_coef(h), h2 = h * h;
  psi.assign(std::max<size_t>(1, n), 0.0f);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
    int c[3];
    B.cell_of(&pts[i * 3], h, c);
    double sum = 0;
    for (int dc = -1; dc <= 1; dc++) {
      if (c[2] + dc < 0 || c[2] + dc >= dim[2]) continue;
      for (int db = -1; db <= 1; db++) {
        if (c[1] + db < 0 || c[1] + db >= dim[1]) continue;
        for (int da = -1; da <= 1; da++) {
          if (c[0] + da < 0 || c[0] + da >= dim[0]) continue;
          size_t k = ((size_t)(c[2] + dc) * dim[1] + (c[1] + db)) * dim[0] + (c[0] + da);
          for (uint32_t t = B.start[k]; t < B.start[k + 1]; t++) {
            uint32_t j = B.items[t];
            double dx = pts[i * 3] - pts[j * 3];
            double dy = pts[i * 3 + 1] - pts[j * 3 + 1];
            double dz = pts[i * 3 + 2] - pts[j * 3 + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= h2) continue;
            double tt = h2 - r2;
            sum += tt * tt * tt;
          }
        }
      }
    }
    sum *= coef;
    psi[i] = sum > 0 ? (float)(rho0 / sum) : 0.0f;
  }
}

// sampleBody (scene.js:137-157) for the reference's three shapes, and the
// two lattice shapes of the synthetic workloads.
void sample_body(int shape, double size, double d, std::vector<double>& local) {
  local.clear();
  if (shape == P4A_SHAPE_LATTICE_CUBE) {
    int e = std::max(1, (int)std::lround(size));
    double c = 0.5 * (e - 1);
    for (int ix = 0; ix < e; ix++)
      for (int iy = 0; iy < e; iy++)
        for (int iz = 0; iz < e; iz++) {
          local.push_back((ix - c) * d);
          local.push_back((iy - c) * d);
          local.push_back((iz - c) * d);
        }
    return;
  }
  if (shape == P4A_SHAPE_LATTICE_SPHERE) {
    int r = std::max(0, (int)std::lround(size));
    for (int ix = -r; ix <= r; ix++)
      for (int iy = -r; iy <= r; iy++)
        for (int iz = -r; iz <= r; iz++)
          if (ix * ix + iy * iy + iz * iz <= r * r) {
            local.push_back(ix * d);
            local.push_back(iy * d);
            local.push_back(iz * d);
          }
    return;
  }
  double ext[3];
  if (shape == P4A_SHAPE_SPHERE) { ext[0] = ext[1] = ext[2] = size; }
  else if (shape == P4A_SHAPE_TORUS) {
    ext[0] = size + 0.4 * size; ext[1] = 0.4 * size; ext[2] = size + 0.4 * size;
  } else { ext[0] = size * 0.9; ext[1] = size * 0.7; ext[2] = size * 0.9; }
  int steps[3];
  for (int k = 0; k < 3; k++) steps[k] = std::max(1, (int)std::ceil(ext[k] / d));
  for (int ix = -steps[0]; ix <= steps[0]; ix++)
    for (int iy = -steps[1]; iy <= steps[1]; iy++)
      for (int iz = -steps[2]; iz <= steps[2]; iz++) {
        double x = ix * d, y = iy * d, z = iz * d;
        bool inside;
        if (shape == P4A_SHAPE_SPHERE) inside = x * x + y * y + z * z <= size * size;
        else if (shape == P4A_SHAPE_TORUS) {
          double planar = std::sqrt(x * x + z * z) - size;
          inside = planar * planar + y * y <= (0.4 * size) * (0.4 * size);
        } else {
          inside = std::fabs(x) <= size * 0.9 && std::fabs(y) <= size * 0.7 &&
                   std::fabs(z) <= size * 0.9;
        }
        if (inside) { local.push_back(x); local.push_back(y); local.push_back(z); }
      }
}

// Mass calibration and denomRest of buildScene (scene.js:232-311), over the
// fluid particles of `pos` (float4 records read as doubles).
void calibrate_jittered(p4a_scene_data& S, const double box[3], double rho0) {
  const double d = S.d, h = S.h;
  const double coef = poly6_coef(h), h2 = h * h;
  const double spiky = -45.0 / (kPi * std::pow(h, 6));
  int dim[3];
  for (int k = 0; k < 3; k++) dim[k] = std::max(1, (int)std::floor(box[k] / h));
  const size_t nF = (size_t)S.n_fluid;
  std::vector<double> xyz(nF * 3);
  for (size_t i = 0; i < nF; i++)
    for (int k = 0; k < 3; k++) xyz[i * 3 + k] = (double)S.pos[i * 4 + k];
  Buckets B;
  B.build(xyz, nF, 3, dim, h);
  auto for_each_nb = [&](size_t i, auto&& f) {
    int c[3];
    B.cell_of(&xyz[i * 3], h, c);
    for (int dc = -1; dc <= 1; dc++) {
      if (c[2] + dc < 0 || c[2] + dc >= dim[2]) continue;
      for (int db = -1; db <= 1; db++) {
        if (c[1] + db < 0 || c[1] + db >= dim[1]) continue;
        for (int da = -1; da <= 1; da++) {
          if (c[0] + da < 0 || c[0] + da >= dim[0]) continue;
          size_t k = ((size_t)(c[2] + dc) * dim[1] + (c[1] + db)) * dim[0] + (c[0] + da);
          for (uint32_t t = B.start[k]; t < B.start[k + 1]; t++)
            if (B.items[t] != i) f(B.items[t]);
        }
      }
    }
  };
  auto density_at = [&](size_t i, double m) {
    double rho = m * coef * h2 * h2 * h2;
    for_each_nb(i, [&](uint32_t j) {
      double dx = xyz[i * 3] - xyz[j * 3], dy = xyz[i * 3 + 1] - xyz[j * 3 + 1],
             dz = xyz[i * 3 + 2] - xyz[j * 3 + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= h2) return;
      double t = h2 - r2;
      rho += m * coef * t * t * t;
    });
    return rho;
  };
  double mass = rho0 * d * d * d;
  double maxRho = 0;
#pragma omp parallel for reduction(max : maxRho) schedule(static)
  for (long i = 0; i < (long)nF; i++) maxRho = std::max(maxRho, density_at((size_t)i, mass));
  if (maxRho > 0.5 * rho0) mass *= rho0 / maxRho;
  double denom = 0;
  const double volume = mass / rho0;
  for (size_t i = 0; i < nF; i++) {
    if (density_at(i, mass) < 0.99 * rho0) continue;
    double gx = 0, gy = 0, gz = 0, sg2 = 0;
    for_each_nb(i, [&](uint32_t j) {
      double rx = xyz[i * 3] - xyz[j * 3], ry = xyz[i * 3 + 1] - xyz[j * 3 + 1],
             rz = xyz[i * 3 + 2] - xyz[j * 3 + 2];
      double r2 = rx * rx + ry * ry + rz * rz;
      if (r2 < 1e-12 || r2 >= h2) return;
      double r = std::sqrt(r2), hr = h - r;
      double s = volume * spiky * hr * hr / r;
      gx += s * rx; gy += s * ry; gz += s * rz;
      sg2 += s * s * r2;
    });
    denom = std::max(denom, gx * gx + gy * gy + gz * gz + sg2);
    if (i > 2000 && denom > 0) break;
  }
  S.mass = mass;
  S.denom_rest = denom;
}

// Mass and denomRest from an interior site of the unjittered lattice of
// spacing d: rho0 = m * sum_o W(|o|d) over lattice offsets o with
// |o|d < h (self included); denomRest = |sum g|^2 + sum |g|^2 with the
// first term zero by the lattice's point symmetry.
void calibrate_lattice(p4a_scene_data& S, double rho0) {
  const double d = S.d, h = S.h, h2 = h * h;
  const double coef = poly6_coef(h);
  const double spiky = -45.0 / (kPi * std::pow(h, 6));
  int R = (int)std::ceil(h / d) + 1;
  double wsum = 0;
  std::vector<double> r2s;
  for (int a = -R; a <= R; a++)
    for (int b = -R; b <= R; b++)
      for (int c = -R; c <= R; c++) {
        double r2 = (a * d) * (a * d) + (b * d) * (b * d) + (c * d) * (c * d);
        if (r2 >= h2) continue;
        double t = h2 - r2;
        wsum += coef * t * t * t;
        if (r2 >= 1e-12) r2s.push_back(r2);
      }
  S.mass = rho0 / wsum;
  const double volume = S.mass / rho0;
  double sg2 = 0;
  for (double r2 : r2s) {
    double r = std::sqrt(r2), hr = h - r;
    double s = volume * spiky * hr * hr / r;
    sg2 += s * s * r2;
  }
  S.denom_rest = sg2;
}

void set_err(std::string* e, const char* m) { if (e) *e = m; }

}  // namespace

extern "C" {

P4A_API int p4a_params_default(p4a_params* p) {
  if (!p) return P4A_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->spacing = 0.02;
  p->rest_density = 1000;
  p->gravity = 9.81;
  p->substeps = 4;
  p->iterations = 4;
  p->cfm_epsilon_rel = 0.01;
  p->scorr_k = 0.1;
  p->scorr_dq = 0.3;
  p->xsph_c = 0.05;
  p->omega = 1.0;
  p->sor_average = 0;
  p->surface_tension_k = 0;
  p->box[0] = 1.5; p->box[1] = 1.5; p->box[2] = 1.0;
  p->no_boundary = 0;
  p->pour_budget = -1;
  p->precision = P4A_PREC_FP32;
  p->math = P4A_MATH_FAST;
  p->use_graph = 1;
  p->neighbour_skin = 0.1;
  p->neighbour_min_particles = 1 << 19;
  return P4A_OK;
}

P4A_API int p4a_boundary_build(const double box[3], double d, double h, double rho0,
                               float* pts, float* psi, int32_t capacity,
                               int32_t* count) {
  if (!box || !count || d <= 0 || h <= 0) return P4A_ERR_ARG;
  std::vector<double> P;
  std::vector<float> Psi;
  boundary_particles(box, d, h, rho0, P, Psi);
  int32_t n = (int32_t)(P.size() / 3);
  *count = n;
  if (pts || psi) {
    if (capacity < n) return P4A_ERR_RANGE;
    for (int32_t i = 0; i < n; i++) {
      if (pts) for (int k = 0; k < 3; k++) pts[i * 3 + k] = (float)P[i * 3 + k];
      if (psi) psi[i] = Psi[i];
    }
  }
  return P4A_OK;
}

P4A_API int p4a_scene_build(const p4a_params* p, const p4a_scene_spec* spec,
                            p4a_scene_data** out) {
  if (!p || !spec || !out) return P4A_ERR_ARG;
  if (p->spacing <= 0 || p->box[0] <= 0 || p->box[1] <= 0 || p->box[2] <= 0)
    return P4A_ERR_ARG;
  if (spec->n_bodies < 0 || (spec->n_bodies > 0 && !spec->bodies)) return P4A_ERR_ARG;
  p4a_scene_data* S = new (std::nothrow) p4a_scene_data();
  if (!S) return P4A_ERR_NOMEM;
  const double d = p->spacing, h = 2 * d;
  const double bx = p->box[0], by = p->box[1], bz = p->box[2];
  S->d = d; S->h = h;

  // Fluid.
  std::vector<double> fluid;
  if (spec->kind == P4A_SCENE_PRESET) {
    // fluidBlock (scene.js:43-64)
    const double margin = d;
    const long nx = std::max(1L, (long)std::floor(bx * 0.35 / d));
    const long nz = std::max(1L, (long)std::floor((bz - 2 * margin) / d));
    const long maxLayers = std::max(1L, (long)std::floor((by - 2 * margin) / d));
    const long target = std::max(0, spec->target_particle_count);
    const long layers = std::min((long)std::ceil((double)target / (double)(nx * nz)), maxLayers);
    JsLcg rng(12345);
    fluid.reserve((size_t)target * 3);
    long cnt = 0;
    for (long iy = 0; iy < layers && cnt < target; iy++)
      for (long ix = 0; ix < nx && cnt < target; ix++)
        for (long iz = 0; iz < nz && cnt < target; iz++) {
          double x = margin + (ix + 0.5) * d + rng.jitter(0.002, d);
          double y = margin + (iy + 0.5) * d + rng.jitter(0.002, d);
          double z = margin + (iz + 0.5) * d + rng.jitter(0.002, d);
          fluid.push_back(x); fluid.push_back(y); fluid.push_back(z);
          cnt++;
        }
  } else if (spec->kind == P4A_SCENE_LATTICE) {
    const long nx = spec->lattice[0], ny = spec->lattice[1], nz = spec->lattice[2];
    if (nx < 0 || ny < 0 || nz < 0) { delete S; return P4A_ERR_ARG; }
    JsLcg rng(spec->seed);
    const double amp2 = 2.0 * spec->jitter;  // (u - .5) * 2j * d spans +-j*d
    fluid.reserve((size_t)(nx * ny * nz) * 3);
    for (long iy = 0; iy < ny; iy++)
      for (long ix = 0; ix < nx; ix++)
        for (long iz = 0; iz < nz; iz++) {
          double x = spec->origin[0] + (ix + 0.5) * d + rng.jitter(amp2, d);
          double y = spec->origin[1] + (iy + 0.5) * d + rng.jitter(amp2, d);
          double z = spec->origin[2] + (iz + 0.5) * d + rng.jitter(amp2, d);
          fluid.push_back(x); fluid.push_back(y); fluid.push_back(z);
        }
  } else {
    delete S;
    return P4A_ERR_ARG;
  }
  const size_t nFluid = fluid.size() / 3;

  // Bodies (scene.js:167-197).
  std::vector<double> bodyPts, bodyRest;
  std::vector<uint32_t> bodyPhase;
  const int nspec = spec->n_bodies;
  const double sz = spec->body_size > 0 ? spec->body_size : 0.12 * std::min(bx, std::min(by, bz));
  const double step = std::min(2.5 * sz, bx * 0.25);
  const double span = step * (nspec - 1);
  std::vector<double> local;
  for (int r = 0; r < nspec; r++) {
    const p4a_body_spec& bs = spec->bodies[r];
    double centre[3];
    if (spec->kind == P4A_SCENE_PRESET) {
      sample_body(bs.shape, sz, d, local);
      double startY = (bs.start_y == 0 || std::isnan(bs.start_y)) ? 0.75 : bs.start_y;
      centre[0] = 0.5 * bx - 0.5 * span + step * r;
      centre[1] = by * startY;
      centre[2] = bz * 0.5;
    } else {
      sample_body(bs.shape, bs.size, d, local);
      for (int k = 0; k < 3; k++) centre[k] = bs.centre[k];
    }
    const size_t count = local.size() / 3;
    if (!count) continue;
    double cx = 0, cy = 0, cz = 0;
    for (size_t i = 0; i < count; i++) {
      cx += local[i * 3]; cy += local[i * 3 + 1]; cz += local[i * 3 + 2];
    }
    cx /= count; cy /= count; cz /= count;
    const uint32_t id = (uint32_t)S->body_count.size() + 1;
    for (size_t i = 0; i < count; i++) {
      double lx = local[i * 3], ly = local[i * 3 + 1], lz = local[i * 3 + 2];
      bodyPts.push_back(centre[0] + lx);
      bodyPts.push_back(centre[1] + ly);
      bodyPts.push_back(centre[2] + lz);
      bodyRest.push_back(lx - cx);
      bodyRest.push_back(ly - cy);
      bodyRest.push_back(lz - cz);
      bodyPhase.push_back(id);
    }
    double dens = bs.density;
    if ((dens == 0 || std::isnan(dens)) && spec->kind == P4A_SCENE_PRESET) dens = 0.5;  // `|| 0.5`
    S->body_centre.push_back(centre[0]);
    S->body_centre.push_back(centre[1]);
    S->body_centre.push_back(centre[2]);
    S->body_density.push_back(dens);
    S->body_count.push_back((int32_t)count);
    S->body_spec.push_back(r);
  }
  const size_t nBody = bodyPts.size() / 3;
  const size_t n = nFluid + nBody;
  if (n > (size_t)INT32_MAX / 16) { delete S; return P4A_ERR_RANGE; }
  S->n = (int32_t)n;
  S->n_fluid = (int32_t)nFluid;
  S->n_body = (int32_t)nBody;
  S->pos.assign(n * 4, 0.0f);
  S->vel.assign(n * 4, 0.0f);
  S->rest.assign(n * 4, 0.0f);
  S->phase.assign(n * 4, 0u);
  for (size_t i = 0; i < nFluid; i++)
    for (int k = 0; k < 3; k++) S->pos[i * 4 + k] = (float)fluid[i * 3 + k];
  for (size_t i = 0; i < nBody; i++) {
    size_t o = (nFluid + i) * 4;
    for (int k = 0; k < 3; k++) {
      S->pos[o + k] = (float)bodyPts[i * 3 + k];
      S->rest[o + k] = (float)bodyRest[i * 3 + k];
    }
    S->phase[o] = bodyPhase[i];
  }

  // Walls.
  std::vector<double> bp;
  boundary_particles(p->box, d, h, p->rest_density, bp, S->bpsi);
  S->bpts.resize(bp.size());
  for (size_t i = 0; i < bp.size(); i++) S->bpts[i] = (float)bp[i];
  if (bp.empty()) S->bpsi.assign(1, 0.0f);

  // Mass and denomRest.
  if (spec->kind == P4A_SCENE_LATTICE && spec->calibrate_unjittered) {
    calibrate_lattice(*S, p->rest_density);
  } else {
    calibrate_jittered(*S, p->box, p->rest_density);
  }
  *out = S;
  return P4A_OK;
}

P4A_API int p4a_scene_view(const p4a_scene_data* S, p4a_scene* v) {
  if (!S || !v) return P4A_ERR_ARG;
  std::memset(v, 0, sizeof(*v));
  v->n = S->n; v->n_fluid = S->n_fluid; v->n_body = S->n_body;
  v->pos = S->pos.data(); v->vel = S->vel.data(); v->rest = S->rest.data();
  v->phase = S->phase.data();
  v->n_boundary = (int32_t)(S->bpts.size() / 3);
  v->boundary_pts = S->bpts.data();
  v->boundary_psi = S->bpsi.data();
  v->n_bodies = (int32_t)S->body_count.size();
  v->body_centre = S->body_centre.data();
  v->body_density = S->body_density.data();
  v->body_count = S->body_count.data();
  v->body_spec = S->body_spec.data();
  v->mass = S->mass; v->h = S->h; v->d = S->d; v->denom_rest = S->denom_rest;
  return P4A_OK;
}

P4A_API void p4a_scene_free(p4a_scene_data* S) { delete S; }

}  // extern "C"
