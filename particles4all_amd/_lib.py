"""ctypes binding of libp4a.so (include/p4a.h) for Python callers: the
tests and bench.py.  The JavaScript `Sim` (node/sim.js) is the drop-in
surface; this module is plumbing over the same C ABI.

Loading fails loudly when the library has not been built: there is no
CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libp4a.so")

P4A_OK = 0
ERRORS = {1: "P4A_ERR_ARG", 2: "P4A_ERR_HIP", 3: "P4A_ERR_STATE", 4: "P4A_ERR_NOMEM",
          5: "P4A_ERR_RANGE", 6: "P4A_ERR_NODEV"}

PREC_FP32, PREC_FP16_OFFSET = 0, 1
MATH_FAST, MATH_EXACT = 0, 1
SHAPE_SPHERE, SHAPE_TORUS, SHAPE_BOX, SHAPE_LATTICE_CUBE, SHAPE_LATTICE_SPHERE = range(5)
SCENE_PRESET, SCENE_LATTICE = 0, 1

BUF = dict(POS=0, VEL=1, PRED=2, BODY=3, REST=4, LAMBDA=5, DENSITY=6, SLOT=7, CORR=8,
           NORMAL=9, CELL_START=10, BPOS=11, BCELL_START=12, BODY_CENTRE=13, BODY_ROT=14,
           PID=15, KEY=16, ORDER=17, BPSI=18, SMOOTH_POS=19, ANISO_G=20,
           OFS=21, CELL_OF_SLOT=22, DRIFT=23, FIELD=24, FIELD_NORMAL=25, MESH_VERTS=26,
           MESH_INDIRECT=27, BODY_PACKED=28, NEIGHBOUR_STATE=29)
STAGE = dict(PREDICT=0, GRID=1, BODY_PRE=2, LAMBDA=3, DELTA=4, BODY_SOLVE=5, FINISH=6, STATS=7)


class Params(C.Structure):
    _fields_ = [("spacing", C.c_double), ("rest_density", C.c_double), ("gravity", C.c_double),
                ("substeps", C.c_int32), ("iterations", C.c_int32),
                ("cfm_epsilon_rel", C.c_double), ("scorr_k", C.c_double),
                ("scorr_dq", C.c_double), ("xsph_c", C.c_double), ("omega", C.c_double),
                ("sor_average", C.c_int32), ("surface_tension_k", C.c_double),
                ("box", C.c_double * 3), ("no_boundary", C.c_int32),
                ("pour_budget", C.c_int32), ("precision", C.c_int32), ("math", C.c_int32),
                ("use_graph", C.c_int32), ("neighbour_skin", C.c_double),
                ("neighbour_min_particles", C.c_int32)]


class BodySpec(C.Structure):
    _fields_ = [("shape", C.c_int32), ("size", C.c_double), ("density", C.c_double),
                ("start_y", C.c_double), ("centre", C.c_double * 3)]


class SceneSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("target_particle_count", C.c_int32),
                ("body_size", C.c_double), ("lattice", C.c_int32 * 3),
                ("origin", C.c_double * 3), ("jitter", C.c_double), ("seed", C.c_uint32),
                ("calibrate_unjittered", C.c_int32), ("n_bodies", C.c_int32),
                ("bodies", C.POINTER(BodySpec))]


class SceneView(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_fluid", C.c_int32), ("n_body", C.c_int32),
                ("pos", C.POINTER(C.c_float)), ("vel", C.POINTER(C.c_float)),
                ("rest", C.POINTER(C.c_float)), ("phase", C.POINTER(C.c_uint32)),
                ("n_boundary", C.c_int32), ("boundary_pts", C.POINTER(C.c_float)),
                ("boundary_psi", C.POINTER(C.c_float)), ("n_bodies", C.c_int32),
                ("body_centre", C.POINTER(C.c_double)), ("body_density", C.POINTER(C.c_double)),
                ("body_count", C.POINTER(C.c_int32)), ("body_spec", C.POINTER(C.c_int32)),
                ("mass", C.c_double), ("h", C.c_double),
                ("d", C.c_double), ("denom_rest", C.c_double)]


class AnisoParams(C.Structure):
    _fields_ = [("radius_scale", C.c_double), ("lambda_", C.c_double), ("kr", C.c_double),
                ("kn", C.c_double), ("ks", C.c_double), ("stretch", C.c_double),
                ("min_extent", C.c_double), ("neps", C.c_int32), ("fp64_fallback", C.c_int32),
                ("gap_tol", C.c_double)]


class FieldInfo(C.Structure):
    _fields_ = [("vert_dim", C.c_int32 * 3), ("pad_", C.c_int32), ("voxel", C.c_double),
                ("origin", C.c_double * 3), ("n_verts", C.c_int64),
                ("max_triangles", C.c_int64)]


class Info(C.Structure):
    _fields_ = [("n", C.c_int32), ("cap", C.c_int32), ("n_fluid", C.c_int32),
                ("n_body_parts", C.c_int32), ("n_bodies", C.c_int32), ("n_boundary", C.c_int32),
                ("n_cells", C.c_int32), ("grid_dim", C.c_int32 * 3), ("parity", C.c_int32),
                ("gen", C.c_int32), ("last_substeps", C.c_int32), ("pour_budget", C.c_int32),
                ("h", C.c_double), ("mass", C.c_double), ("denom_rest", C.c_double),
                ("sim_time", C.c_double), ("last_advanced", C.c_double),
                ("time_bank", C.c_double), ("body_scale", C.c_double), ("cov_scale", C.c_double)]


_lib = None

# Every symbol include/p4a.h declares.
EXPORTS = ["p4a_params_default", "p4a_scene_build", "p4a_scene_view", "p4a_scene_free",
           "p4a_boundary_build", "p4a_create", "p4a_destroy", "p4a_last_error", "p4a_reset",
           "p4a_set_params", "p4a_step", "p4a_substeps", "p4a_resize_box", "p4a_append_fluid",
           "p4a_apply_ray_impulse", "p4a_hold_body", "p4a_release_body", "p4a_read",
           "p4a_write", "p4a_device_ptr", "p4a_get_info", "p4a_stats", "p4a_body_pose",
           "p4a_sync", "p4a_run_stage", "p4a_neighbour_census", "p4a_set_timing",
           "p4a_timing", "p4a_version", "p4a_aniso_params_default", "p4a_aniso",
           "p4a_copy_probe", "p4a_tune", "p4a_tile_variant_name", "p4a_field_configure",
           "p4a_field_generate", "p4a_field_get_info", "p4a_field_smooth",
           "p4a_mesh_normal_field", "p4a_mesh_march", "p4a_mesh_triangles", "p4a_body_pack",
           "p4a_snapshot_size", "p4a_snapshot_save", "p4a_snapshot_load",
           "p4a_batch_create", "p4a_batch_destroy",
           "p4a_batch_last_error", "p4a_batch_substeps", "p4a_batch_tune", "p4a_batch_sync",
           "p4a_batch_scene_size", "p4a_batch_read", "p4a_batch_stats", "p4a_stats_noalloc"]


def lib():
    """The loaded libp4a.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                           "(there is no fallback path)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.p4a_params_default.argtypes = [C.POINTER(Params)]
    L.p4a_scene_build.argtypes = [C.POINTER(Params), C.POINTER(SceneSpec), C.POINTER(vp)]
    L.p4a_scene_view.argtypes = [vp, C.POINTER(SceneView)]
    L.p4a_scene_free.argtypes = [vp]
    L.p4a_scene_free.restype = None
    L.p4a_boundary_build.argtypes = [C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double,
                                     vp, vp, C.c_int32, C.POINTER(C.c_int32)]
    L.p4a_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.p4a_destroy.argtypes = [vp]
    L.p4a_last_error.argtypes = [vp]
    L.p4a_last_error.restype = C.c_char_p
    L.p4a_reset.argtypes = [vp, C.POINTER(Params), C.POINTER(SceneView)]
    L.p4a_set_params.argtypes = [vp, C.POINTER(Params)]
    L.p4a_step.argtypes = [vp, C.c_double, C.POINTER(C.c_int32)]
    L.p4a_substeps.argtypes = [vp, C.c_double, C.c_int32]
    L.p4a_resize_box.argtypes = [vp, C.POINTER(C.c_double)]
    L.p4a_append_fluid.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(C.c_int32)]
    L.p4a_apply_ray_impulse.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.c_double, C.c_double]
    L.p4a_hold_body.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.c_double, C.c_double]
    L.p4a_stats_noalloc.argtypes = [vp, vp]
    L.p4a_release_body.argtypes = [vp]
    L.p4a_read.argtypes = [vp, C.c_int32, vp, C.c_size_t, C.c_size_t]
    L.p4a_field_configure.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.c_int32, C.c_int64]
    L.p4a_field_smooth.argtypes = [vp, C.c_int32]
    L.p4a_body_pack.argtypes = [vp, vp, C.c_int32]
    L.p4a_snapshot_size.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.p4a_snapshot_save.argtypes = [vp, vp, C.c_size_t]
    L.p4a_snapshot_load.argtypes = [vp, vp, C.c_size_t]
    L.p4a_mesh_normal_field.argtypes = [vp, C.c_int32]
    L.p4a_mesh_march.argtypes = [vp, C.c_float]
    L.p4a_mesh_triangles.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.p4a_field_generate.argtypes = [vp, C.c_int32, C.c_int32]
    L.p4a_field_get_info.argtypes = [vp, C.POINTER(FieldInfo)]
    L.p4a_batch_create.argtypes = [C.c_int32, C.c_int32, C.POINTER(Params),
                                   C.POINTER(SceneView), C.POINTER(vp)]
    L.p4a_batch_destroy.argtypes = [vp]
    L.p4a_batch_last_error.argtypes = [vp]
    L.p4a_batch_last_error.restype = C.c_char_p
    L.p4a_batch_substeps.argtypes = [vp, C.c_double, C.c_int32]
    L.p4a_batch_tune.argtypes = [vp, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                 C.c_int32]
    L.p4a_batch_sync.argtypes = [vp]
    L.p4a_batch_scene_size.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.p4a_batch_read.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_size_t, C.c_size_t]
    L.p4a_batch_stats.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.p4a_write.argtypes = [vp, C.c_int32, vp, C.c_size_t, C.c_size_t]
    L.p4a_device_ptr.argtypes = [vp, C.c_int32, C.POINTER(vp), C.POINTER(C.c_size_t)]
    L.p4a_get_info.argtypes = [vp, C.POINTER(Info)]
    L.p4a_stats.argtypes = [vp, C.POINTER(C.c_double)]
    L.p4a_body_pose.argtypes = [vp, vp, vp]
    L.p4a_sync.argtypes = [vp]
    L.p4a_run_stage.argtypes = [vp, C.c_int32, C.c_double]
    L.p4a_neighbour_census.argtypes = [vp, vp, vp, vp]
    L.p4a_set_timing.argtypes = [vp, C.c_int32]
    L.p4a_timing.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_double),
                             C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]
    L.p4a_version.restype = C.c_char_p
    L.p4a_tune.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int32]
    L.p4a_tile_variant_name.argtypes = [C.c_int32, C.c_int32]
    L.p4a_tile_variant_name.restype = C.c_char_p
    L.p4a_copy_probe.argtypes = [C.c_int, C.c_size_t, C.c_int32, C.POINTER(C.c_double)]
    L.p4a_aniso_params_default.argtypes = [C.POINTER(AnisoParams)]
    L.p4a_aniso.argtypes = [vp, C.POINTER(AnisoParams), C.POINTER(C.c_int32)]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _dvec(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def copy_probe(nbytes, reps=5, device=0):
    """Mean ms of one device-to-device copy of `nbytes` (float4 kernel)."""
    ms = C.c_double()
    rc = lib().p4a_copy_probe(device, nbytes, reps, C.byref(ms))
    if rc:
        raise RuntimeError(f"p4a_copy_probe failed: {ERRORS.get(rc, rc)}")
    return ms.value


def default_params(**kw):
    p = Params()
    lib().p4a_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "box":
            p.box = (C.c_double * 3)(*v)
        else:
            setattr(p, k, v)
    return p


class Scene:
    """Scene arrays built by the library (buildScene or a lattice scene)."""

    def __init__(self, params, spec, bodies=()):
        L = lib()
        self._bodies = (BodySpec * max(1, len(bodies)))(*bodies)
        spec.n_bodies = len(bodies)
        spec.bodies = C.cast(self._bodies, C.POINTER(BodySpec)) if bodies else None
        h = C.c_void_p()
        rc = L.p4a_scene_build(C.byref(params), C.byref(spec), C.byref(h))
        if rc:
            raise RuntimeError(f"p4a_scene_build failed: {ERRORS.get(rc, rc)}")
        self.handle = h
        self.view = SceneView()
        L.p4a_scene_view(h, C.byref(self.view))
        v = self.view
        n, nb, nB = v.n, v.n_boundary, v.n_bodies
        self.n, self.n_fluid, self.n_body = n, v.n_fluid, v.n_body
        self.pos = np.ctypeslib.as_array(v.pos, (max(n, 1) * 4,))[: n * 4].reshape(n, 4)
        self.vel = np.ctypeslib.as_array(v.vel, (max(n, 1) * 4,))[: n * 4].reshape(n, 4)
        self.rest = np.ctypeslib.as_array(v.rest, (max(n, 1) * 4,))[: n * 4].reshape(n, 4)
        self.phase = np.ctypeslib.as_array(v.phase, (max(n, 1) * 4,))[: n * 4].reshape(n, 4)
        self.boundary_pts = np.ctypeslib.as_array(v.boundary_pts, (max(nb, 1) * 3,))[: nb * 3].reshape(nb, 3)
        self.boundary_psi = np.ctypeslib.as_array(v.boundary_psi, (max(nb, 1),))[:nb]
        self.n_bodies = nB
        if nB:
            self.body_centre = np.ctypeslib.as_array(v.body_centre, (nB * 3,)).reshape(nB, 3)
            self.body_density = np.ctypeslib.as_array(v.body_density, (nB,))
            self.body_count = np.ctypeslib.as_array(v.body_count, (nB,))
        else:
            self.body_centre = np.zeros((0, 3))
            self.body_density = np.zeros(0)
            self.body_count = np.zeros(0, np.int32)
        self.mass, self.h, self.d, self.denom_rest = v.mass, v.h, v.d, v.denom_rest

    def free(self):
        if self.handle:
            lib().p4a_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def preset_scene(params, target, bodies=(), body_size=0.0):
    """buildScene(p) for a preset-style parameter set.  `bodies` is a list
    of (shape, density, startY) as the reference's 'shape:density:startY'
    strings give them."""
    shapes = {"sphere": SHAPE_SPHERE, "torus": SHAPE_TORUS}
    spec = SceneSpec()
    spec.kind = SCENE_PRESET
    spec.target_particle_count = int(target)
    spec.body_size = float(body_size)
    bs = [BodySpec(shape=shapes.get(s, SHAPE_BOX), size=0.0, density=float(dn),
                   start_y=float(sy), centre=(C.c_double * 3)(0, 0, 0))
          for (s, dn, sy) in bodies]
    return Scene(params, spec, bs)


def lattice_scene(params, lattice, origin=(0.0, 0.0, 0.0), jitter=0.05, seed=1, bodies=(),
                  calibrate_unjittered=True):
    """Jittered lattice block (sites origin + (i+0.5)d, +-jitter*d) with
    explicit bodies [(shape, size, density, (cx, cy, cz))]."""
    spec = SceneSpec()
    spec.kind = SCENE_LATTICE
    spec.lattice = (C.c_int32 * 3)(*lattice)
    spec.origin = _dvec(origin)
    spec.jitter = float(jitter)
    spec.seed = int(seed) & 0xFFFFFFFF
    spec.calibrate_unjittered = 1 if calibrate_unjittered else 0
    bs = [BodySpec(shape=sh, size=float(sz), density=float(dn), start_y=0.0, centre=_dvec(c))
          for (sh, sz, dn, c) in bodies]
    return Scene(params, spec, bs)


class Sim:
    """One scene on one GPU through the C ABI."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.p4a_create(device, C.byref(h))
        if rc:
            raise RuntimeError(f"p4a_create({device}) failed: {ERRORS.get(rc, rc)}")
        self.h = h

    def _check(self, rc, what):
        if rc:
            msg = lib().p4a_last_error(self.h).decode()
            raise RuntimeError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")

    def close(self):
        if self.h:
            lib().p4a_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, params, scene):
        self._check(lib().p4a_reset(self.h, C.byref(params), C.byref(scene.view)), "reset")
        self.params = params

    def set_params(self, params):
        self._check(lib().p4a_set_params(self.h, C.byref(params)), "set_params")
        self.params = params

    def step(self, frame_dt):
        sub = C.c_int32()
        self._check(lib().p4a_step(self.h, frame_dt, C.byref(sub)), "step")
        return sub.value

    def substeps(self, dt, count):
        self._check(lib().p4a_substeps(self.h, dt, count), "substeps")

    def run_stage(self, stage, dt=0.0):
        self._check(lib().p4a_run_stage(self.h, STAGE[stage], dt), f"stage {stage}")

    def info(self):
        i = Info()
        self._check(lib().p4a_get_info(self.h, C.byref(i)), "info")
        return i

    def read(self, name, dtype=None, cols=None):
        i = self.info()
        n, nc = i.n, i.n_cells
        spec = {"POS": (np.float32, n, 4), "VEL": (np.float32, n, 4), "PRED": (np.float32, n, 4),
                "BODY": (np.uint32, n, 4), "REST": (np.float32, n, 4),
                "LAMBDA": (np.float32, n, 1), "DENSITY": (np.float32, n, 1),
                "SLOT": (np.uint32, n, 1), "CORR": (np.float32, n, 4),
                "NORMAL": (np.float32, n, 4), "CELL_START": (np.uint32, nc + 2, 1),
                "BPOS": (np.float32, i.n_boundary, 4),
                "BCELL_START": (np.uint32, nc + 2, 1),
                "BODY_CENTRE": (np.float32, i.n_bodies, 4),
                "BODY_ROT": (np.float32, i.n_bodies * 3, 4), "PID": (np.uint32, n, 1),
                "KEY": (np.uint32, n, 1), "ORDER": (np.uint32, n, 1),
                "BPSI": (np.float32, i.n_boundary, 1), "SMOOTH_POS": (np.float32, n, 4),
                "ANISO_G": (np.float32, 2 * n, 4), "OFS": (np.uint32, n, 2),
                "CELL_OF_SLOT": (np.uint32, n, 1), "DRIFT": (np.float32, 256, 1),
                "FIELD": (np.float32, self.field_info().n_verts, 1),
                "FIELD_NORMAL": (np.float32, self.field_info().n_verts, 1),
                "MESH_VERTS": (np.uint32, 3 * self.field_info().max_triangles, 2),
                "MESH_INDIRECT": (np.uint32, 4, 1),
                "BODY_PACKED": (np.float32, 6 * i.n_bodies, 4),
                "NEIGHBOUR_STATE": (np.uint32, 1025, 1)}[name]
        dt, rows, c = spec
        out = np.empty((rows, c) if c > 1 else (rows,), dtype=dt)
        self._check(lib().p4a_read(self.h, BUF[name], _ptr(out), 0, out.nbytes), f"read {name}")
        return out

    def write(self, name, arr, offset=0):
        a = np.ascontiguousarray(arr)
        self._check(lib().p4a_write(self.h, BUF[name], _ptr(a), offset, a.nbytes), f"write {name}")

    def stats(self):
        out = (C.c_double * 4)()
        self._check(lib().p4a_stats(self.h, out), "stats")
        return list(out)

    def body_pose(self):
        nB = self.info().n_bodies
        c = np.zeros((nB, 3), np.float32)
        r = np.zeros((nB, 9), np.float32)
        self._check(lib().p4a_body_pose(self.h, _ptr(c), _ptr(r)), "body_pose")
        return c, r

    def census(self):
        n = self.info().n
        a = np.zeros(n, np.uint32)
        b = np.zeros(n, np.uint32)
        c = np.zeros(n, np.uint32)
        self._check(lib().p4a_neighbour_census(self.h, _ptr(a), _ptr(b), _ptr(c)), "census")
        return a, b, c

    def resize_box(self, box):
        self._check(lib().p4a_resize_box(self.h, _dvec(box)), "resize_box")

    def append_fluid(self, pos3, vel3):
        p = np.ascontiguousarray(pos3, np.float32).reshape(-1)
        v = np.ascontiguousarray(vel3, np.float32).reshape(-1)
        k = C.c_int32()
        self._check(lib().p4a_append_fluid(self.h, _ptr(p), _ptr(v), min(p.size, v.size) // 3,
                                           C.byref(k)), "append_fluid")
        return k.value

    def apply_ray_impulse(self, origin, direction, impulse, radius, speed_limit):
        self._check(lib().p4a_apply_ray_impulse(self.h, _dvec(origin), _dvec(direction),
                                                _dvec(impulse), radius, speed_limit), "impulse")

    def hold_body(self, body, target, rate, limit):
        self._check(lib().p4a_hold_body(self.h, body, _dvec(target), rate, limit), "hold_body")

    def release_body(self):
        self._check(lib().p4a_release_body(self.h), "release_body")

    def sync(self):
        self._check(lib().p4a_sync(self.h), "sync")

    def set_timing(self, on):
        self._check(lib().p4a_set_timing(self.h, 1 if on else 0), "set_timing")

    def timing(self):
        names = C.create_string_buffer(4096)
        ms = (C.c_double * 64)()
        cnt = (C.c_int32 * 64)()
        k = C.c_int32()
        self._check(lib().p4a_timing(self.h, names, 4096, ms, cnt, 64, C.byref(k)), "timing")
        keys = names.value.decode().split("\n")[: k.value]
        return {key: (ms[i], cnt[i]) for i, key in enumerate(keys)}

    def aniso(self, **kw):
        a = AnisoParams()
        lib().p4a_aniso_params_default(C.byref(a))
        for k, v in kw.items():
            setattr(a, "lambda_" if k == "lambda" else k, v)
        cnt = C.c_int32()
        self._check(lib().p4a_aniso(self.h, C.byref(a), C.byref(cnt)), "aniso")
        return cnt.value

    def field_configure(self, box_min, box_max, resolution, max_triangles=0):
        """SurfaceMesh.configure (mesh.js:96-128); max_triangles <= 0 is the
        reference's 3,000,000."""
        self._check(lib().p4a_field_configure(self.h, (C.c_double * 3)(*box_min),
                                              (C.c_double * 3)(*box_max), int(resolution),
                                              int(max_triangles)),
                    "field_configure")

    def snapshot(self):
        """The whole simulation state as bytes (p4a_snapshot_save)."""
        size = C.c_size_t()
        self._check(lib().p4a_snapshot_size(self.h, C.byref(size)), "snapshot_size")
        buf = np.empty(size.value, np.uint8)
        self._check(lib().p4a_snapshot_save(self.h, _ptr(buf), buf.nbytes), "snapshot_save")
        return buf.tobytes()

    def load_snapshot(self, data):
        """Reset this handle from snapshot bytes (p4a_snapshot_load)."""
        buf = np.frombuffer(bytes(data), np.uint8)
        self._check(lib().p4a_snapshot_load(self.h, _ptr(buf), buf.nbytes), "snapshot_load")

    def body_state_fp64(self, snap=None):
        """The FP64 body state (centres (nB, 3), row-major rotations (nB, 9))
        as the snapshot layout holds it: after the per-particle arrays and
        the cell starts of the last grid build."""
        snap = self.snapshot() if snap is None else snap
        i = self.info()
        n, nc, nB, nb = i.n, i.n_cells, i.n_bodies, i.n_boundary
        hdr = len(snap) - (n * 60 + (nc + 2) * 4 + nB * 172 + nb * 16)
        off = hdr + n * 60 + (nc + 2) * 4
        centre = np.frombuffer(snap, np.float64, 4 * nB, off).reshape(nB, 4)[:, :3].copy()
        rot = np.frombuffer(snap, np.float64, 9 * nB, off + 32 * nB).reshape(nB, 9).copy()
        return centre, rot

    def save_snapshot(self, path):
        with open(path, "wb") as f:
            f.write(self.snapshot())

    def load_snapshot_file(self, path):
        with open(path, "rb") as f:
            self.load_snapshot(f.read())

    def body_pack(self, stat=None):
        """Solids.pack (solids.js:147-161); stat (n_bodies, 8) float32 as
        Solids.build fills it, None keeps the last upload."""
        arr = None if stat is None else np.ascontiguousarray(stat, np.float32)
        self._check(lib().p4a_body_pack(self.h, None if arr is None else _ptr(arr),
                                        self.info().n_bodies), "body_pack")

    def field_smooth(self, passes):
        """SurfaceMesh.smoothField (mesh.js:273-291), in place."""
        self._check(lib().p4a_field_smooth(self.h, int(passes)), "field_smooth")

    def mesh_normal_field(self, passes):
        """SurfaceMesh.buildNormalField (mesh.js:293-313)."""
        self._check(lib().p4a_mesh_normal_field(self.h, int(passes)), "mesh_normal_field")

    def mesh_march(self, iso):
        """The march of SurfaceMesh.generate (mesh.js:320-331); asynchronous."""
        self._check(lib().p4a_mesh_march(self.h, float(iso)), "mesh_march")

    def mesh_triangles(self):
        """(total triangles including dropped ones, triangles written)."""
        t, e = C.c_int64(), C.c_int64()
        self._check(lib().p4a_mesh_triangles(self.h, C.byref(t), C.byref(e)), "mesh_triangles")
        return t.value, e.value

    def mesh(self):
        """Written triangles of the last march as uint32 (k, 6) and the total."""
        total, emitted = self.mesh_triangles()
        out = np.empty((max(emitted, 0), 6), np.uint32)
        if emitted:
            self._check(lib().p4a_read(self.h, BUF["MESH_VERTS"], _ptr(out), 0, out.nbytes),
                        "read MESH_VERTS")
        return out, total

    def field_generate(self, anisotropic=False, skip_bodies=-1):
        """SurfaceMesh.generateField (mesh.js:263-271); anisotropic uses the
        last aniso() pass."""
        self._check(lib().p4a_field_generate(self.h, 1 if anisotropic else 0, int(skip_bodies)),
                    "field_generate")

    def field_info(self):
        f = FieldInfo()
        self._check(lib().p4a_field_get_info(self.h, C.byref(f)), "field_info")
        return f

    def field(self):
        """The last generated field as an array of shape (z, y, x)."""
        f = self.field_info()
        vd = list(f.vert_dim)
        return self.read("FIELD").reshape(vd[2], vd[1], vd[0])

    def tune(self):
        """Autotune the tile shapes; returns (chosen names, {name: ms})."""
        ch = (C.c_int32 * 2)()
        ms = (C.c_double * 32)()
        self._check(lib().p4a_tune(self.h, ch, ms, 32), "tune")
        L = lib()
        times, k = {}, 0
        for g in (1, 2):
            v = 0
            while L.p4a_tile_variant_name(g, v).decode() != "?":
                times[f"g{g}:" + L.p4a_tile_variant_name(g, v).decode()] = ms[k]
                k += 1
                v += 1
        return ([L.p4a_tile_variant_name(1, ch[0]).decode(),
                 L.p4a_tile_variant_name(2, ch[1]).decode()], times)

    def device_ptr(self, name):
        p = C.c_void_p()
        b = C.c_size_t()
        self._check(lib().p4a_device_ptr(self.h, BUF[name], C.byref(p), C.byref(b)), "device_ptr")
        return p.value, b.value


class Batch:
    """Many independent scenes stepped together (p4a_batch_*): every stage
    is one launch over all scenes and the substep is one hipGraph."""

    def __init__(self, items, device=0):
        """items: list of (Params, Scene)."""
        L = lib()
        n = len(items)
        self._scenes = [sc for _, sc in items]  # keep the arrays alive
        P = (Params * n)(*[p for p, _ in items])
        V = (SceneView * n)(*[sc.view for _, sc in items])
        h = C.c_void_p()
        rc = L.p4a_batch_create(device, n, P, V, C.byref(h))
        self.h = h
        self.n_scenes = n
        if rc:
            msg = L.p4a_batch_last_error(h).decode() if h else ""
            self.close()
            raise RuntimeError(f"p4a_batch_create: {ERRORS.get(rc, rc)}: {msg}")

    def _check(self, rc, what):
        if rc:
            msg = lib().p4a_batch_last_error(self.h).decode()
            raise RuntimeError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")

    def close(self):
        if self.h:
            lib().p4a_batch_destroy(self.h)
            self.h = None

    def substeps(self, dt, count):
        self._check(lib().p4a_batch_substeps(self.h, dt, int(count)), "batch_substeps")

    def tune(self, dt):
        """p4a_batch_tune: returns (chosen variant names, {name: ms})."""
        L = lib()
        ch = (C.c_int32 * 2)()
        ms = (C.c_double * 16)()
        self._check(L.p4a_batch_tune(self.h, dt, ch, ms, 16), "batch_tune")
        times, k = {}, 0
        for g in (1, 2):
            for v in (0, 1, 2, 5, 6, 7, 8, 9):
                times[f"g{g}:" + L.p4a_tile_variant_name(g, v).decode()] = ms[k]
                k += 1
        return [L.p4a_tile_variant_name(g, ch[g - 1]).decode() for g in (1, 2)], times

    def sync(self):
        self._check(lib().p4a_batch_sync(self.h), "batch_sync")

    def scene_size(self, s):
        n = C.c_int32()
        self._check(lib().p4a_batch_scene_size(self.h, s, C.byref(n)), "batch_scene_size")
        return n.value

    def read(self, s, name):
        n = self.scene_size(s)
        dt, c = {"POS": (np.float32, 4), "VEL": (np.float32, 4), "PRED": (np.float32, 4),
                 "DENSITY": (np.float32, 1), "LAMBDA": (np.float32, 1),
                 "PID": (np.uint32, 1)}[name]
        out = np.empty((n, c) if c > 1 else (n,), dtype=dt)
        self._check(lib().p4a_batch_read(self.h, s, BUF[name], _ptr(out), 0, out.nbytes),
                    f"batch_read {name}")
        return out

    def stats(self, s):
        out = (C.c_double * 4)()
        self._check(lib().p4a_batch_stats(self.h, s, out), "batch_stats")
        return list(out)

