// Node.js N-API addon over libp4a.so (include/p4a.h).  Each export is a
// thin marshalling wrapper; node/sim.js builds the reference's `Sim`
// surface (src/sim.js) on top of them.  Errors from the library are thrown
// as JS exceptions carrying p4a_last_error().
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "p4a.h"

namespace {

#define NAPI_OK(call)                                                   \
  do {                                                                  \
    if ((call) != napi_ok) {                                            \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);      \
      return nullptr;                                                   \
    }                                                                   \
  } while (0)

napi_value throw_p4a(napi_env env, p4a_sim* h, int rc, const char* what) {
  std::string m = std::string(what) + " failed (status " + std::to_string(rc) + ")";
  if (h) m += std::string(": ") + p4a_last_error(h);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

#define P4A_CALL(h, call, what)                                         \
  do {                                                                  \
    int rc_ = (call);                                                   \
    if (rc_) return throw_p4a(env, (h), rc_, (what));                   \
  } while (0)

std::vector<napi_value> args(napi_env env, napi_callback_info info, size_t n) {
  std::vector<napi_value> a(n);
  size_t got = n;
  napi_get_cb_info(env, info, &got, a.data(), nullptr, nullptr);
  for (size_t i = got; i < n; i++) napi_get_undefined(env, &a[i]);
  return a;
}

double num(napi_env env, napi_value v, double def = 0) {
  double d = def;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_number) napi_get_value_double(env, v, &d);
  else if (t == napi_boolean) {
    bool b = false;
    napi_get_value_bool(env, v, &b);
    d = b ? 1 : 0;
  }
  return d;
}

napi_value prop(napi_env env, napi_value o, const char* k) {
  napi_value v = nullptr;
  bool has = false;
  napi_valuetype t;
  napi_typeof(env, o, &t);
  if (t == napi_object && napi_has_named_property(env, o, k, &has) == napi_ok && has)
    napi_get_named_property(env, o, k, &v);
  else
    napi_get_undefined(env, &v);
  return v;
}

double pnum(napi_env env, napi_value o, const char* k, double def) {
  return num(env, prop(env, o, k), def);
}

void vec3(napi_env env, napi_value a, double out[3], double def = 0) {
  for (int i = 0; i < 3; i++) {
    napi_value e;
    out[i] = def;
    bool isArr = false;
    napi_is_array(env, a, &isArr);
    if (isArr && napi_get_element(env, a, i, &e) == napi_ok) out[i] = num(env, e, def);
  }
}

p4a_sim* handle(napi_env env, napi_value v) {
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  return (p4a_sim*)p;
}

void set_num(napi_env env, napi_value o, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, k, x);
}

// p4a_params from a JS object in the reference's parameter names
// (defaultParams, scene.js:14-31) plus the build's own switches.
void params_from(napi_env env, napi_value o, p4a_params* p) {
  p4a_params_default(p);
  p->spacing = pnum(env, o, "spacing", p->spacing);
  p->rest_density = pnum(env, o, "restDensity", p->rest_density);
  p->gravity = pnum(env, o, "gravity", p->gravity);
  p->substeps = (int32_t)pnum(env, o, "substeps", p->substeps);
  p->iterations = (int32_t)pnum(env, o, "iterations", p->iterations);
  p->cfm_epsilon_rel = pnum(env, o, "cfmEpsilonRel", p->cfm_epsilon_rel);
  p->scorr_k = pnum(env, o, "sCorrK", p->scorr_k);
  p->scorr_dq = pnum(env, o, "sCorrDq", p->scorr_dq);
  p->xsph_c = pnum(env, o, "xsphC", p->xsph_c);
  p->omega = pnum(env, o, "omega", p->omega);
  p->sor_average = pnum(env, o, "sorAverage", 0) != 0;
  p->surface_tension_k = pnum(env, o, "surfaceTensionK", p->surface_tension_k);
  napi_value b = prop(env, o, "box");
  double box[3] = {p->box[0], p->box[1], p->box[2]};
  napi_valuetype t;
  napi_typeof(env, b, &t);
  if (t == napi_object) vec3(env, b, box, 0);
  for (int k = 0; k < 3; k++) p->box[k] = box[k];
  p->no_boundary = pnum(env, o, "noBoundary", 0) != 0;
  // `pour === false` disables the pour budget (sim.js:94)
  napi_value pour = prop(env, o, "pour");
  napi_typeof(env, pour, &t);
  bool pourOn = true;
  if (t == napi_boolean) napi_get_value_bool(env, pour, &pourOn);
  p->pour_budget = pourOn ? (int32_t)pnum(env, o, "pourBudget", -1) : 0;
  p->precision = (int32_t)pnum(env, o, "precision", 0);
  p->math = (int32_t)pnum(env, o, "math", 0);
  p->use_graph = (int32_t)pnum(env, o, "useGraph", 1);
  p->neighbour_skin = pnum(env, o, "neighbourSkin", p->neighbour_skin);
  p->neighbour_min_particles =
      (int32_t)pnum(env, o, "neighbourMinParticles", p->neighbour_min_particles);
}

void finalize_scene(napi_env, void* data, void*) { p4a_scene_free((p4a_scene_data*)data); }

napi_value f32array(napi_env env, const float* src, size_t count) {
  napi_value ab, ta;
  void* dst = nullptr;
  napi_create_arraybuffer(env, count * 4, &dst, &ab);
  if (count) std::memcpy(dst, src, count * 4);
  napi_create_typedarray(env, napi_float32_array, count, ab, 0, &ta);
  return ta;
}

napi_value u32array(napi_env env, const uint32_t* src, size_t count) {
  napi_value ab, ta;
  void* dst = nullptr;
  napi_create_arraybuffer(env, count * 4, &dst, &ab);
  if (count) std::memcpy(dst, src, count * 4);
  napi_create_typedarray(env, napi_uint32_array, count, ab, 0, &ta);
  return ta;
}

bool typed_f32(napi_env env, napi_value v, const float** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  void* d = nullptr;
  napi_value ab;
  size_t off = 0;
  napi_get_typedarray_info(env, v, &t, len, &d, &ab, &off);
  if (t != napi_float32_array) return false;
  *data = (const float*)d;
  return true;
}

// ---- exports ---------------------------------------------------------------

napi_value Create(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = nullptr;
  P4A_CALL(nullptr, p4a_create((int)num(env, a[0], 0), &h), "p4a_create");
  napi_value ext;
  NAPI_OK(napi_create_external(env, h, nullptr, nullptr, &ext));
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  if (h) p4a_destroy(h);
  return nullptr;
}

// buildScene(p) with p in the reference's names; bodies as the
// 'shape:density:startY' strings of defaultParams.
napi_value SceneBuild(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_params p;
  params_from(env, a[0], &p);
  p4a_scene_spec spec;
  std::memset(&spec, 0, sizeof(spec));
  spec.kind = (int32_t)pnum(env, a[1], "kind", P4A_SCENE_PRESET);
  spec.target_particle_count = (int32_t)pnum(env, a[0], "targetParticleCount", 60000);
  spec.body_size = pnum(env, a[0], "bodySize", 0);
  double lat[3] = {0, 0, 0}, org[3] = {0, 0, 0};
  vec3(env, prop(env, a[1], "lattice"), lat);
  vec3(env, prop(env, a[1], "origin"), org);
  for (int k = 0; k < 3; k++) {
    spec.lattice[k] = (int32_t)lat[k];
    spec.origin[k] = org[k];
  }
  spec.jitter = pnum(env, a[1], "jitter", 0.05);
  spec.seed = (uint32_t)pnum(env, a[1], "seed", 1);
  spec.calibrate_unjittered = pnum(env, a[1], "calibrateUnjittered", 1) != 0;
  std::vector<p4a_body_spec> bodies;
  napi_value list = prop(env, a[1], "bodies");
  bool isArr = false;
  napi_is_array(env, list, &isArr);
  if (isArr) {
    uint32_t len = 0;
    napi_get_array_length(env, list, &len);
    for (uint32_t i = 0; i < len; i++) {
      napi_value e;
      napi_get_element(env, list, i, &e);
      p4a_body_spec b;
      std::memset(&b, 0, sizeof(b));
      b.shape = (int32_t)pnum(env, e, "shape", P4A_SHAPE_BOX);
      b.size = pnum(env, e, "size", 0);
      b.density = pnum(env, e, "density", 0.5);
      b.start_y = pnum(env, e, "startY", 0.75);
      vec3(env, prop(env, e, "centre"), b.centre);
      bodies.push_back(b);
    }
  }
  spec.n_bodies = (int32_t)bodies.size();
  spec.bodies = bodies.empty() ? nullptr : bodies.data();
  p4a_scene_data* sd = nullptr;
  P4A_CALL(nullptr, p4a_scene_build(&p, &spec, &sd), "p4a_scene_build");
  p4a_scene v;
  p4a_scene_view(sd, &v);
  napi_value out, ext;
  NAPI_OK(napi_create_object(env, &out));
  NAPI_OK(napi_create_external(env, sd, finalize_scene, nullptr, &ext));
  napi_set_named_property(env, out, "handle", ext);
  set_num(env, out, "n", v.n);
  set_num(env, out, "nFluid", v.n_fluid);
  set_num(env, out, "nBody", v.n_body);
  set_num(env, out, "mass", v.mass);
  set_num(env, out, "h", v.h);
  set_num(env, out, "d", v.d);
  set_num(env, out, "denomRest", v.denom_rest);
  napi_set_named_property(env, out, "pos", f32array(env, v.pos, (size_t)v.n * 4));
  napi_set_named_property(env, out, "vel", f32array(env, v.vel, (size_t)v.n * 4));
  napi_set_named_property(env, out, "rest", f32array(env, v.rest, (size_t)v.n * 4));
  napi_set_named_property(env, out, "phase", u32array(env, v.phase, (size_t)v.n * 4));
  napi_value bnd;
  napi_create_object(env, &bnd);
  napi_set_named_property(env, bnd, "pts", f32array(env, v.boundary_pts, (size_t)v.n_boundary * 3));
  napi_set_named_property(env, bnd, "psi", f32array(env, v.boundary_psi, (size_t)v.n_boundary));
  set_num(env, bnd, "count", v.n_boundary);
  napi_set_named_property(env, out, "boundary", bnd);
  napi_value arr;
  napi_create_array_with_length(env, v.n_bodies, &arr);
  for (int b = 0; b < v.n_bodies; b++) {
    napi_value o, c;
    napi_create_object(env, &o);
    set_num(env, o, "id", b + 1);
    set_num(env, o, "density", v.body_density[b]);
    set_num(env, o, "count", v.body_count[b]);
    set_num(env, o, "spec", v.body_spec[b]);
    napi_create_array_with_length(env, 3, &c);
    for (int k = 0; k < 3; k++) {
      napi_value x;
      napi_create_double(env, v.body_centre[b * 3 + k], &x);
      napi_set_element(env, c, k, x);
    }
    napi_set_named_property(env, o, "centre", c);
    napi_set_element(env, arr, b, o);
  }
  napi_set_named_property(env, out, "bodies", arr);
  return out;
}

napi_value BoundaryBuild(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 4);
  double box[3];
  vec3(env, a[0], box);
  const double d = num(env, a[1]), h = num(env, a[2]), rho0 = num(env, a[3], 1000);
  int32_t nb = 0;
  P4A_CALL(nullptr, p4a_boundary_build(box, d, h, rho0, nullptr, nullptr, 0, &nb), "p4a_boundary_build");
  std::vector<float> pts((size_t)nb * 3), psi(nb);
  P4A_CALL(nullptr, p4a_boundary_build(box, d, h, rho0, pts.data(), psi.data(), nb, &nb),
           "p4a_boundary_build");
  napi_value out;
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "pts", f32array(env, pts.data(), pts.size()));
  napi_set_named_property(env, out, "psi", f32array(env, psi.data(), psi.size()));
  set_num(env, out, "count", nb);
  return out;
}

napi_value Reset(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  p4a_sim* h = handle(env, a[0]);
  p4a_params p;
  params_from(env, a[1], &p);
  void* sd = nullptr;
  napi_get_value_external(env, prop(env, a[2], "handle"), &sd);
  if (!sd) {
    napi_throw_error(env, nullptr, "reset: scene has no native handle");
    return nullptr;
  }
  p4a_scene v;
  p4a_scene_view((p4a_scene_data*)sd, &v);
  P4A_CALL(h, p4a_reset(h, &p, &v), "p4a_reset");
  return nullptr;
}

napi_value SetParams(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  p4a_params p;
  params_from(env, a[1], &p);
  P4A_CALL(h, p4a_set_params(h, &p), "p4a_set_params");
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  int32_t sub = 0;
  P4A_CALL(h, p4a_step(h, num(env, a[1]), &sub), "p4a_step");
  napi_value r;
  napi_create_int32(env, sub, &r);
  return r;
}

napi_value Substeps(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_substeps(h, num(env, a[1]), (int32_t)num(env, a[2])), "p4a_substeps");
  return nullptr;
}

napi_value ResizeBox(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  double box[3];
  vec3(env, a[1], box);
  P4A_CALL(h, p4a_resize_box(h, box), "p4a_resize_box");
  return nullptr;
}

napi_value AppendFluid(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  p4a_sim* h = handle(env, a[0]);
  // plain arrays or Float32Arrays of x,y,z triples
  std::vector<float> P, V;
  for (int w = 0; w < 2; w++) {
    std::vector<float>& dst = w == 0 ? P : V;
    const float* f = nullptr;
    size_t len = 0;
    if (typed_f32(env, a[1 + w], &f, &len)) {
      dst.assign(f, f + len);
    } else {
      uint32_t n = 0;
      napi_get_array_length(env, a[1 + w], &n);
      dst.resize(n);
      for (uint32_t i = 0; i < n; i++) {
        napi_value e;
        napi_get_element(env, a[1 + w], i, &e);
        dst[i] = (float)num(env, e);
      }
    }
  }
  const int32_t count = (int32_t)(std::min(P.size(), V.size()) / 3);
  int32_t got = 0;
  P4A_CALL(h, p4a_append_fluid(h, P.data(), V.data(), count, &got), "p4a_append_fluid");
  napi_value r;
  napi_create_int32(env, got, &r);
  return r;
}

napi_value ApplyRayImpulse(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 6);
  p4a_sim* h = handle(env, a[0]);
  double o[3], d[3], imp[3];
  vec3(env, a[1], o);
  vec3(env, a[2], d);
  vec3(env, a[3], imp);
  P4A_CALL(h, p4a_apply_ray_impulse(h, o, d, imp, num(env, a[4]), num(env, a[5])),
           "p4a_apply_ray_impulse");
  return nullptr;
}

napi_value HoldBody(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 5);
  p4a_sim* h = handle(env, a[0]);
  double t[3];
  vec3(env, a[2], t);
  P4A_CALL(h, p4a_hold_body(h, (int32_t)num(env, a[1]), t, num(env, a[3], 10), num(env, a[4], 4)),
           "p4a_hold_body");
  return nullptr;
}

napi_value ReleaseBody(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_release_body(h), "p4a_release_body");
  return nullptr;
}

napi_value Info(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  p4a_info i;
  P4A_CALL(h, p4a_get_info(h, &i), "p4a_get_info");
  napi_value o, g;
  napi_create_object(env, &o);
  set_num(env, o, "n", i.n);
  set_num(env, o, "cap", i.cap);
  set_num(env, o, "nFluid", i.n_fluid);
  set_num(env, o, "nBodyParts", i.n_body_parts);
  set_num(env, o, "nBodies", i.n_bodies);
  set_num(env, o, "nBoundary", i.n_boundary);
  set_num(env, o, "nCells", i.n_cells);
  napi_create_array_with_length(env, 3, &g);
  for (int k = 0; k < 3; k++) {
    napi_value x;
    napi_create_int32(env, i.grid_dim[k], &x);
    napi_set_element(env, g, k, x);
  }
  napi_set_named_property(env, o, "gridDim", g);
  set_num(env, o, "parity", i.parity);
  set_num(env, o, "gen", i.gen);
  set_num(env, o, "lastSubsteps", i.last_substeps);
  set_num(env, o, "pourBudget", i.pour_budget);
  set_num(env, o, "h", i.h);
  set_num(env, o, "mass", i.mass);
  set_num(env, o, "denomRest", i.denom_rest);
  set_num(env, o, "simTime", i.sim_time);
  set_num(env, o, "lastAdvanced", i.last_advanced);
  set_num(env, o, "timeBank", i.time_bank);
  set_num(env, o, "bodyScale", i.body_scale);
  set_num(env, o, "covScale", i.cov_scale);
  return o;
}

// read(h, bufId, elemType 'f32'|'u32', count) -> typed array
napi_value Read(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 4);
  p4a_sim* h = handle(env, a[0]);
  const int32_t id = (int32_t)num(env, a[1]);
  const size_t count = (size_t)num(env, a[3]);
  char kind[8] = {0};
  size_t kl = 0;
  napi_get_value_string_utf8(env, a[2], kind, sizeof(kind), &kl);
  napi_value ab, ta;
  void* dst = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, count * 4, &dst, &ab));
  P4A_CALL(h, p4a_read(h, id, dst, 0, count * 4), "p4a_read");
  NAPI_OK(napi_create_typedarray(env, std::strcmp(kind, "u32") == 0 ? napi_uint32_array
                                                                   : napi_float32_array,
                                 count, ab, 0, &ta));
  return ta;
}

napi_value DevicePtr(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  void* p = nullptr;
  size_t bytes = 0;
  P4A_CALL(h, p4a_device_ptr(h, (int32_t)num(env, a[1]), &p, &bytes), "p4a_device_ptr");
  napi_value o, big;
  napi_create_object(env, &o);
  napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)p, &big);
  napi_set_named_property(env, o, "ptr", big);
  set_num(env, o, "bytes", (double)bytes);
  return o;
}

napi_value Stats(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  double s[4];
  P4A_CALL(h, p4a_stats(h, s), "p4a_stats");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "avgRho", s[0]);
  set_num(env, o, "maxRho", s[1]);
  set_num(env, o, "maxSpeed", s[2]);
  set_num(env, o, "ke", s[3]);
  return o;
}

napi_value BodyPose(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  const size_t nB = (size_t)num(env, a[1]);
  std::vector<float> c(nB * 3 + 1), r(nB * 9 + 1);
  P4A_CALL(h, p4a_body_pose(h, c.data(), r.data()), "p4a_body_pose");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "centre", f32array(env, c.data(), nB * 3));
  napi_set_named_property(env, o, "rot", f32array(env, r.data(), nB * 9));
  return o;
}

napi_value RunStage(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_run_stage(h, (int32_t)num(env, a[1]), num(env, a[2])), "p4a_run_stage");
  return nullptr;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_sync(h), "p4a_sync");
  return nullptr;
}

// aniso(h, {radiusScale, lambda, kr, kn, ks, stretch, minExtent, neps,
// fp64Fallback, gapTol}) -> number of FP64 re-solves (missing keys keep
// p4a_aniso_params_default)
napi_value Aniso(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  p4a_aniso_params q;
  p4a_aniso_params_default(&q);
  q.radius_scale = pnum(env, a[1], "radiusScale", q.radius_scale);
  q.lambda = pnum(env, a[1], "lambda", q.lambda);
  q.kr = pnum(env, a[1], "kr", q.kr);
  q.kn = pnum(env, a[1], "kn", q.kn);
  q.ks = pnum(env, a[1], "ks", q.ks);
  q.stretch = pnum(env, a[1], "stretch", q.stretch);
  q.min_extent = pnum(env, a[1], "minExtent", q.min_extent);
  q.neps = (int32_t)pnum(env, a[1], "neps", q.neps);
  q.fp64_fallback = (int32_t)pnum(env, a[1], "fp64Fallback", q.fp64_fallback);
  q.gap_tol = pnum(env, a[1], "gapTol", q.gap_tol);
  int32_t solves = 0;
  P4A_CALL(h, p4a_aniso(h, &q, &solves), "p4a_aniso");
  napi_value v;
  napi_create_int32(env, solves, &v);
  return v;
}

// fieldConfigure(h, boxMin[3], boxMax[3], resolution, maxTriangles)
napi_value FieldConfigure(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 5);
  p4a_sim* h = handle(env, a[0]);
  double lo[3], hi[3];
  vec3(env, a[1], lo);
  vec3(env, a[2], hi);
  P4A_CALL(h, p4a_field_configure(h, lo, hi, (int32_t)num(env, a[3]),
                                  (int64_t)num(env, a[4], 0)),
           "p4a_field_configure");
  return nullptr;
}

// snapshotSave(h) -> Uint8Array
napi_value SnapshotSave(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  size_t bytes = 0;
  P4A_CALL(h, p4a_snapshot_size(h, &bytes), "p4a_snapshot_size");
  void* data = nullptr;
  napi_value ab, u8;
  NAPI_OK(napi_create_arraybuffer(env, bytes, &data, &ab));
  P4A_CALL(h, p4a_snapshot_save(h, data, bytes), "p4a_snapshot_save");
  NAPI_OK(napi_create_typedarray(env, napi_uint8_array, bytes, ab, 0, &u8));
  return u8;
}

// snapshotLoad(h, Uint8Array | Buffer)
napi_value SnapshotLoad(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  bool is = false;
  napi_is_typedarray(env, a[1], &is);
  if (!is) {
    napi_throw_type_error(env, nullptr, "snapshotLoad: expected a Uint8Array");
    return nullptr;
  }
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  void* d = nullptr;
  napi_value ab;
  NAPI_OK(napi_get_typedarray_info(env, a[1], &t, &len, &d, &ab, &off));
  if (t != napi_uint8_array) {
    napi_throw_type_error(env, nullptr, "snapshotLoad: expected a Uint8Array");
    return nullptr;
  }
  P4A_CALL(h, p4a_snapshot_load(h, d, len), "p4a_snapshot_load");
  return nullptr;
}

// bodyPack(h, stat Float32Array | null, count)
napi_value BodyPack(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  p4a_sim* h = handle(env, a[0]);
  const float* f = nullptr;
  size_t len = 0;
  const int32_t count = (int32_t)num(env, a[2]);
  if (typed_f32(env, a[1], &f, &len) && len < (size_t)count * 8) {
    napi_throw_error(env, nullptr, "bodyPack: stat needs 8 floats per body");
    return nullptr;
  }
  P4A_CALL(h, p4a_body_pack(h, f, count), "p4a_body_pack");
  return nullptr;
}

// fieldSmooth(h, passes)
napi_value FieldSmooth(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_field_smooth(h, (int32_t)num(env, a[1], 0)), "p4a_field_smooth");
  return nullptr;
}

// meshNormalField(h, passes)
napi_value MeshNormalField(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_mesh_normal_field(h, (int32_t)num(env, a[1], 0)), "p4a_mesh_normal_field");
  return nullptr;
}

// meshMarch(h, iso)
napi_value MeshMarch(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 2);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_mesh_march(h, (float)num(env, a[1])), "p4a_mesh_march");
  return nullptr;
}

// meshTriangles(h) -> {total, emitted}
napi_value MeshTriangles(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  int64_t total = 0, emitted = 0;
  P4A_CALL(h, p4a_mesh_triangles(h, &total, &emitted), "p4a_mesh_triangles");
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_double(env, (double)total, &v);
  napi_set_named_property(env, o, "total", v);
  napi_create_double(env, (double)emitted, &v);
  napi_set_named_property(env, o, "emitted", v);
  return o;
}

// tune(h) -> {lambda: shape name, xsph: shape name} (p4a_tune)
napi_value Tune(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  int32_t chosen[2] = {0, 0};
  P4A_CALL(h, p4a_tune(h, chosen, nullptr, 0), "p4a_tune");
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_string_utf8(env, p4a_tile_variant_name(1, chosen[0]), NAPI_AUTO_LENGTH, &v);
  napi_set_named_property(env, o, "lambda", v);
  napi_create_string_utf8(env, p4a_tile_variant_name(2, chosen[1]), NAPI_AUTO_LENGTH, &v);
  napi_set_named_property(env, o, "xsph", v);
  return o;
}

// fieldGenerate(h, anisotropic, skipBodies)
napi_value FieldGenerate(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 3);
  p4a_sim* h = handle(env, a[0]);
  P4A_CALL(h, p4a_field_generate(h, num(env, a[1]) != 0 ? P4A_FIELD_ANISOTROPIC
                                                          : P4A_FIELD_ISOTROPIC,
                                 (int32_t)num(env, a[2], -1)),
           "p4a_field_generate");
  return nullptr;
}

// fieldInfo(h) -> {vertDim, voxel, origin, nVerts, maxTriangles}
napi_value FieldInfo(napi_env env, napi_callback_info info) {
  auto a = args(env, info, 1);
  p4a_sim* h = handle(env, a[0]);
  p4a_field_info f;
  P4A_CALL(h, p4a_field_get_info(h, &f), "p4a_field_get_info");
  napi_value o, vd, org, v;
  napi_create_object(env, &o);
  napi_create_array_with_length(env, 3, &vd);
  napi_create_array_with_length(env, 3, &org);
  for (int k = 0; k < 3; k++) {
    napi_create_int32(env, f.vert_dim[k], &v);
    napi_set_element(env, vd, k, v);
    napi_create_double(env, f.origin[k], &v);
    napi_set_element(env, org, k, v);
  }
  napi_set_named_property(env, o, "vertDim", vd);
  napi_set_named_property(env, o, "origin", org);
  napi_create_double(env, f.voxel, &v);
  napi_set_named_property(env, o, "voxel", v);
  napi_create_double(env, (double)f.n_verts, &v);
  napi_set_named_property(env, o, "nVerts", v);
  napi_create_double(env, (double)f.max_triangles, &v);
  napi_set_named_property(env, o, "maxTriangles", v);
  return o;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  napi_create_string_utf8(env, p4a_version(), NAPI_AUTO_LENGTH, &s);
  return s;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback fn;
  } fns[] = {{"create", Create},         {"destroy", Destroy},
             {"sceneBuild", SceneBuild}, {"boundaryBuild", BoundaryBuild},
             {"reset", Reset},           {"setParams", SetParams},
             {"step", Step},             {"substeps", Substeps},
             {"resizeBox", ResizeBox},   {"appendFluid", AppendFluid},
             {"applyRayImpulse", ApplyRayImpulse},
             {"holdBody", HoldBody},     {"releaseBody", ReleaseBody},
             {"info", Info},             {"read", Read},
             {"devicePtr", DevicePtr},   {"stats", Stats},
             {"bodyPose", BodyPose},     {"runStage", RunStage},
             {"sync", Sync},             {"version", Version},
             {"aniso", Aniso},           {"fieldConfigure", FieldConfigure},
             {"fieldGenerate", FieldGenerate}, {"fieldInfo", FieldInfo},
             {"fieldSmooth", FieldSmooth}, {"meshNormalField", MeshNormalField},
             {"meshMarch", MeshMarch},   {"meshTriangles", MeshTriangles},
             {"bodyPack", BodyPack},     {"snapshotSave", SnapshotSave},
             {"snapshotLoad", SnapshotLoad}, {"tune", Tune}};
  for (auto& f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
