"""Benchmark of the simulation step on one MI355X.

Workload (the headline configuration): block_4m_bodies - a 256x128x128
jittered fluid lattice (4,194,304 particles, spacing d = 1, h = 2d, sites
one spacing inside the walls) in a 512x192x130 box plus 64 rigid bodies
(32 lattice cubes with 6-12 particle edges, 32 lattice spheres with 4-8
particle radii, densities log-uniform in [0.3, 2.0]); 4 solver iterations,
substep 1/120 s, XSPH 0.01, surface tension 0.5, FP32.  The dam break first
flows for --settle substeps (2 s of simulated time by default, untimed), so
the timed steps see a collapsing, flowing block rather than the first
instants after release.  One step = one Sim.step(1/120) = one substep
(predict, grid build, body centroids, 4 x (lambda, delta, shape matching),
xsph + normals, tension + finalize) plus the stats reduction, all resident
in HBM.

Prints one JSON line; `value` is particle-substeps/s of the default (fast)
arithmetic.  Beside it, on the same state:
* per-kernel HIP-event times (eager launches on the simulation stream) and
  the roofline of the dominant kernel, with the committed rocprofv3
  summary's duration and HBM traffic for the same kernel;
* the exact arithmetic (bit-identical to the oracle's float chain) timed
  the same way;
* the FP16-offset form timed the same way;
* accuracy of one substep from the flowing state: the fast, exact and
  FP16-offset forms against the oracle's float chain (the reference's FP32
  arithmetic) and against its FP64 evaluation (positions, velocities,
  lambdas, densities), and the float chain with every sqrt and division one
  ulp up (within WGSL's accuracy for those operations) against the float
  chain: how far two conforming runs of the reference may differ;
* `batch`: 256 independent scenes stepped as one batch against the same
  scenes stepped one by one;
* `field`: the anisotropy pass and the anisotropic density field of the
  2,097,152-particle dam break after 600 substeps;
* `cpu_baseline`: the oracle's float restatement of the same substep
  (OpenMP threads on the host), timed in the accuracy run; the probe for a
  software WebGPU adapter that could run the reference WGSL.
"""
import argparse
import glob
import json
import os
import platform
import shutil
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per target particle and launch: compulsory per-particle
# reads/writes plus the cell-tile gather of the 27-cell algorithm without
# reuse across cells (27 records per target at 8 particles per cell).
ALGO_BYTES = {
    "lambda": 24 + 27 * 16,       # pred in; density, lambda out; pred records
    "delta": 36 + 27 * 20,        # pred, lambda in; pred out; pred+lambda records
    "xsph": 68 + 27 * 36,         # pred, pos, rho in; corr, normal out
    "tension": 96 + 27 * 32,      # pred, pos, rho, normal, corr in; pos, vel, corr out
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="block_4m_bodies")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--settle", type=int, default=240,
                    help="untimed substeps of flow before the timed steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the oracle accuracy figures (and the CPU baseline timed there)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-field", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the device copy probe on MIB mebibytes and exit")
    ap.add_argument("--field", action="store_true",
                    help="only the field measurement (dam_break_2m), as its own JSON line")
    ap.add_argument("--field-settle", type=int, default=600,
                    help="substeps of flow before the field measurement")
    ap.add_argument("--resolution", type=int, default=256)
    ap.add_argument("--tiles", default="", metavar="G1,G2",
                    help="fix the traversal shapes (variant numbers of groups 1 and 2) "
                         "instead of autotuning, e.g. for counter runs that must profile "
                         "the shapes a timed run picked")
    ap.add_argument("--batch", type=int, default=0, metavar="SCENES",
                    help="only the batch measurement over SCENES scenes, as its own JSON line")
    ap.add_argument("--neighbour-skin", type=float, default=None,
                    help="skin of the per-substep neighbour masks in units of h (0: full "
                         "27-cell walks); default: the library's")
    ap.add_argument("--no-sequential", action="store_true",
                    help="with --batch: time the batched launches only, not the same "
                         "scenes stepped one by one")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def settle(sim, dt, count, what):
    t0 = time.perf_counter()
    done = 0
    while done < count:
        k = min(60, count - done)
        sim.substeps(dt, k)
        sim.sync()
        done += k
        log(f"{what}: {done}/{count} substeps of flow, {time.perf_counter() - t0:.1f}s")


def host_cpu():
    model = platform.processor() or "?"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return model, os.cpu_count(), usable


def webgpu_probe():
    """Looks for an installed software WebGPU stack that could run the
    reference WGSL on the host cores: a Vulkan CPU driver (lavapipe or
    SwiftShader) with its ICD manifest, a Dawn or wgpu-native library, a
    Node.js WebGPU module, a Chrome/Chromium or Deno binary.  Only looks
    for files; runs nothing it finds."""
    found = []
    for d in ("/usr/share/vulkan/icd.d", "/etc/vulkan/icd.d", "/usr/local/share/vulkan/icd.d"):
        found += glob.glob(os.path.join(d, "*.json"))
    libdirs = ["/usr/lib", "/usr/lib64", "/usr/local/lib", "/usr/lib/x86_64-linux-gnu"]
    for pat in ("libvulkan_lvp.so*", "libvk_swiftshader.so*", "libwebgpu_dawn.so*",
                "libdawn*.so*", "libwgpu_native.so*"):
        for d in libdirs:
            found += glob.glob(os.path.join(d, pat)) + glob.glob(os.path.join(d, "*", pat))
    for exe in ("google-chrome", "chromium", "chromium-browser", "chrome", "deno",
                "vulkaninfo"):
        if shutil.which(exe):
            found.append(shutil.which(exe))
    for d in ("/usr/lib/node_modules", "/usr/local/lib/node_modules",
              os.path.expanduser("~/node_modules")):
        for mod in ("webgpu", "@webgpu", "dawn", "wgpu"):
            if os.path.isdir(os.path.join(d, mod)):
                found.append(os.path.join(d, mod))
    node = shutil.which("node")
    return {"available": bool(found), "found": found[:20],
            "node": node,
            "note": ("no Vulkan ICD, Dawn/wgpu library, WebGPU Node module, Chrome or Deno on "
                     "this host: the reference WGSL cannot run on its cores") if not found
            else "candidate WebGPU components present (not used by this bench)"}


def profiled(workload, kernel):
    """Mean duration (us) and HBM bytes per launch of `kernel` from the
    committed rocprofv3 summary of this workload, or (None, None)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_fp32.md")
    us = traffic = None
    if not os.path.exists(path):
        return us, traffic
    for line in open(path):
        if line.startswith("TRAFFIC ") or line.startswith("AVG_US "):
            vals = dict(kv.split("=") for kv in line.split()[1:])
            if kernel in vals:
                if line.startswith("TRAFFIC "):
                    traffic = int(vals[kernel])
                else:
                    us = float(vals[kernel])
    return us, traffic


def timed_steps(sim, frame, steps):
    sim.sync()
    t0 = time.perf_counter()
    sub = 0
    for _ in range(steps):
        sub += sim.step(frame)
    sim.sync()
    return sub, time.perf_counter() - t0


def kernel_timing(sim, frame, steps):
    sim.set_timing(True)
    sim.substeps(frame, steps)
    sim.sync()
    kt = sim.timing()
    sim.set_timing(False)
    return kt


def set_form(sim, p, math=None, precision=None):
    if math is not None:
        p.math = math
    if precision is not None:
        p.precision = precision
    sim.set_params(p)


def time_form(sim, p, frame, steps, math, precision, base_value):
    """Graph-replay throughput and per-kernel times of one arithmetic form
    on the simulation's current state (two untimed steps capture it)."""
    set_form(sim, p, math, precision)
    n = sim.info().n
    for _ in range(2):
        sim.step(frame)
    sub, el = timed_steps(sim, frame, steps)
    kt = kernel_timing(sim, frame, steps)
    value = n * sub / el
    return {"value": value, "ms_per_step": el / steps * 1e3, "vs_fast": value / base_value,
            "kernels_ms": {k: round(v[0], 4) for k, v in kt.items()}}, kt


def _err(a, b, scale):
    import numpy as np
    e = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / scale
    return {"max": float(e.max()), "rms": float(np.sqrt((e ** 2).mean()))}


def accuracy(sim, p, sc, dt):
    """One substep from the simulation's current (flowing) state in the
    fast, exact and FP16-offset forms, each against the oracle's float chain
    (the reference's FP32 arithmetic) and its FP64 evaluation started from
    the same state (positions and FP64 body poses read back).  Positions in
    units of h, velocities in units of h per substep (|dv| dt / h), lambdas
    in units of max |lambda|, densities in units of rho0.  Also times the
    float chain's substep (the CPU baseline)."""
    import numpy as np
    from oracle import oracle as O
    snap = sim.snapshot()
    centre, rot = sim.body_state_fp64(snap)
    f64 = lambda k: sim.read(k).astype(np.float64)
    pos0, vel0, body0, rest0, pid0 = f64("POS"), f64("VEL"), sim.read("BODY")[:, 0], \
        f64("REST"), sim.read("PID")
    P = O.constants_from(p, sc, dt)
    rec, start = O.sort_boundary(P, sc.boundary_pts, sc.boundary_psi)
    bnd = (rec, rec[:, 3].copy(), start)
    inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(np.float64)
    tension = p.surface_tension_k > 0
    ref = {}
    cpu_s = None
    for name, mode in (("float", O.F32), ("fp64", O.F64), ("float_ulp", O.F32_ULP)):
        st = O.State(pos0, vel0, body0, rest0)
        t0 = time.perf_counter()
        info = O.substep(mode, P, st, bnd, centre.copy(), rot.copy(), inv, p.iterations,
                         tension)
        if mode == O.F32:
            cpu_s = time.perf_counter() - t0
        ref[name] = (st, info)
        log(f"oracle {name} substep from the flowing state: {time.perf_counter() - t0:.1f}s")
    # every result is compared particle by particle: each side's slots are
    # put in particle-id order (the FP64 chain may bin a particle near a
    # cell face into the other cell, so its sorted order can differ)
    def by_id(order):
        return np.argsort(pid0[order], kind="stable")

    al = {name: by_id(ref[name][1]["order"]) for name in ref}
    lam_scale = max(float(np.abs(ref["fp64"][1]["lam"]).max()), 1e-30)

    def fields(pos, vel, lam, den, idx):
        return pos[idx], vel[idx], lam[idx], den[idx]

    R = {name: fields(ref[name][0].pos[:, :3], ref[name][0].vel[:, :3], ref[name][1]["lam"],
                      ref[name][1]["density"], al[name]) for name in ref}
    out = {}
    for form, math, prec in (("fast", 0, 0), ("exact", 1, 0), ("fp16_offset", 0, 1)):
        sim.load_snapshot(snap)
        set_form(sim, p, math, prec)
        sim.substeps(dt, 1)
        gidx = np.argsort(sim.read("PID"), kind="stable")
        same = bool(np.array_equal(sim.read("PID"), pid0[ref["float"][1]["order"]]))
        pos, vel, lam, den = fields(f64("POS")[:, :3], f64("VEL")[:, :3], f64("LAMBDA"),
                                    f64("DENSITY"), gidx)
        r = {"same_order_as_float_chain": same}
        for name in ("float", "fp64"):
            rp, rv, rl, rd = R[name]
            r[f"vs_{name}"] = {
                "pos_h": _err(pos, rp, sc.h),
                "vel_h_per_substep": _err(vel * dt, rv * dt, sc.h),
                "lambda": _err(lam, rl, lam_scale),
                "density_rho0": _err(den, rd, p.rest_density),
            }
        if form == "exact":
            rp, rv, rl, _ = R["float"]
            r["bit_identical_to_float_chain"] = bool(
                same and np.array_equal(pos, rp) and np.array_equal(vel, rv) and
                np.array_equal(lam, rl))
        pf = r["vs_fp64"]["pos_h"]
        r["pos_within_1e-5_h_of_fp64"] = float(
            (np.abs(pos - R["fp64"][0]).max(axis=1) <= 1e-5 * sc.h).mean())
        r["pos_within_1e-5_h_of_float_chain"] = float(
            (np.abs(pos - R["float"][0]).max(axis=1) <= 1e-5 * sc.h).mean())
        out[form] = r
        log(f"{form}: vs float chain pos max {r['vs_float']['pos_h']['max']:.3e} h, "
            f"vs FP64 pos max {pf['max']:.3e} h rms {pf['rms']:.3e} h")
    out["float_chain_vs_fp64"] = {
        "pos_h": _err(R["float"][0], R["fp64"][0], sc.h),
        "vel_h_per_substep": _err(R["float"][1] * dt, R["fp64"][1] * dt, sc.h),
        "lambda": _err(R["float"][2], R["fp64"][2], lam_scale),
        "pos_within_1e-5_h": float(
            (np.abs(R["float"][0] - R["fp64"][0]).max(axis=1) <= 1e-5 * sc.h).mean())}
    # a conforming implementation of the reference's FP32 arithmetic: WGSL
    # allows its f32 sqrt and division more than one ulp, so the float chain
    # with every sqrt and division of the neighbour kernels one ulp up is as
    # much the reference as the float chain is
    out["float_chain_one_ulp_vs_float_chain"] = {
        "pos_h": _err(R["float_ulp"][0], R["float"][0], sc.h),
        "vel_h_per_substep": _err(R["float_ulp"][1] * dt, R["float"][1] * dt, sc.h),
        "lambda": _err(R["float_ulp"][2], R["float"][2], lam_scale),
        "pos_within_1e-5_h": float(
            (np.abs(R["float_ulp"][0] - R["float"][0]).max(axis=1) <= 1e-5 * sc.h).mean())}
    log("float chain with sqrt and division one ulp up vs the float chain: pos max "
        f"{out['float_chain_one_ulp_vs_float_chain']['pos_h']['max']:.3e} h")
    sim.load_snapshot(snap)
    set_form(sim, p, 0, 0)
    return out, cpu_s


def density_errors(sim, p, sc, frame):
    """Densities of the FP16-offset and the FP32 form on the same predicted
    state (the simulation's current one) against the FP64 oracle; |error| /
    rho0."""
    import numpy as np
    from oracle import oracle as O
    P = O.constants_from(p, sc, frame)
    set_form(sim, p, 0, 1)
    sim.run_stage("PREDICT", frame)
    sim.run_stage("GRID", frame)
    pred = sim.read("PRED").astype(np.float64)
    cs = sim.read("CELL_START")
    bpos = sim.read("BPOS").astype(np.float64)
    bcs = sim.read("BCELL_START")
    sim.run_stage("LAMBDA", frame)
    d16 = sim.read("DENSITY").astype(np.float64)
    set_form(sim, p, 0, 0)
    sim.run_stage("LAMBDA", frame)
    d32 = sim.read("DENSITY").astype(np.float64)
    den, _ = O.lambda_(O.F64, P, pred, cs, bpos, bpos[:, 3].copy(), bcs)
    out = {}
    for name, d in (("fp16_offset", d16), ("fp32", d32)):
        e = np.abs(d - den) / P.rho0
        out[name] = {"mean": float(e.mean()), "max": float(e.max())}
    return out


def field_measure(scale=1.0, settle_substeps=600, resolution=256, reps=5):
    """Anisotropy pass + density fields on the 2,097,152-particle dam break
    after `settle_substeps` substeps of flow: HIP-event times of the
    anisotropy kernel, the anisotropic splat into the 256^3 lattice and the
    isotropic (reference) field, the normal field and the march, plus
    size-independent checks of the outputs."""
    import numpy as np
    from particles4all_amd import Sim, workloads
    p, sc, frame = workloads.dam_break_2m(scale=scale)
    sim = Sim(0)
    sim.reset(p, sc)
    n = sim.info().n
    settle(sim, 1.0 / 120.0, settle_substeps, "field scene")
    st = sim.stats()
    box = list(p.box)
    sim.field_configure([0, 0, 0], box, resolution)
    fi = sim.field_info()
    # SurfaceMesh's anisoLimitToField default: extents >= 1.25 voxel / h
    # (mesh.js:238)
    akw = dict(min_extent=1.25 * fi.voxel / sc.h)
    sim.aniso(**akw)
    sim.field_generate(anisotropic=True)
    sim.field_generate(anisotropic=False)
    sim.sync()
    sim.set_timing(True)
    solves = 0
    for _ in range(reps):
        solves = sim.aniso(**akw)
        sim.field_generate(anisotropic=True)
    kt = sim.timing()
    sim.set_timing(True)
    for _ in range(reps):
        sim.field_generate(anisotropic=False)
    kti = sim.timing()
    sim.set_timing(False)
    # size-independent checks: G positive definite (det > 0), neighbour
    # counts >= 1 (self), and the anisotropic field integrates the
    # particles' volumes (each kernel integrates to V_j)
    g = sim.read("ANISO_G").reshape(-1, 2, 4)
    det_ok = float(np.mean(g[:, 1, 2] > 0))
    counts = g[:, 1, 3]
    sim.field_generate(anisotropic=True)
    field = sim.read("FIELD").astype(np.float64)
    sim.mesh_normal_field(2)
    sim.mesh_march(0.4)
    sim.sync()
    sim.set_timing(True)
    for _ in range(reps):
        sim.mesh_normal_field(2)
        sim.mesh_march(0.4)
    ktm = sim.timing()
    sim.set_timing(False)
    tri_total, tri_written = sim.mesh_triangles()
    rho = sim.read("DENSITY").astype(np.float64)
    vol = float(sc.mass) / np.maximum(rho, 0.5 * p.rest_density)
    integral = field.sum() * fi.voxel ** 3
    sim.close()
    a_ms, f_ms, i_ms = kt["aniso"][0], kt["field"][0], kti["field"][0]
    return {
        "workload": f"dam_break_2m (scale {scale}): jittered 128^3 lattice in a 256^3 box after "
                    f"{settle_substeps} substeps of 1/120 s",
        "particles": n, "field_vert_dim": list(fi.vert_dim), "voxel": fi.voxel,
        "aniso": "R = 2h, lambda 0.9, kr 4, ks 1, stretch 2, kn 1, neps 25, min extent "
                 "1.25 voxel/h, FP64 re-solve below a 1e-4 eigen-gap",
        "particles_per_s": n / ((a_ms + f_ms) * 1e-3),
        "ms": {"aniso": a_ms, "field_anisotropic": f_ms, "field_isotropic": i_ms,
               "normal_field_2_passes": ktm["normal field"][0], "march": ktm["march"][0]},
        "surface": {"iso": 0.4, "triangles": tri_total, "written": tri_written},
        "fp64_solves": solves,
        "checks": {"det_G_positive": det_ok, "min_count": float(counts.min()),
                   "mean_count": float(counts.mean()),
                   "field_integral_over_volume": integral / float(vol.sum())},
        "stats": {"avgRho": st[0], "maxRho": st[1], "maxSpeed": st[2]},
    }


def batch_measure(n_scenes=256, steps=5, warmup=2, sequential=True):
    """n_scenes scenes of 2e3-2e5 particles: one batched graph per substep
    against each scene's own graph replayed in turn (synchronised after each
    scene, so scenes do not overlap), the same substeps of the same
    scenes."""
    from particles4all_amd import Batch, Sim, workloads
    t0 = time.time()
    items = workloads.batch_scenes(n_scenes=n_scenes)
    n_total = sum(sc.n for _, sc in items)
    log(f"{n_scenes} scenes, {n_total} particles, built in {time.time() - t0:.1f}s")
    dt = 1.0 / 120.0
    b = Batch(items)
    b.substeps(dt, warmup)
    b.sync()
    chosen, tune_ms = b.tune(dt)
    log("batch autotuner: " + ", ".join(f"{k} {v:.3f}ms" for k, v in tune_ms.items()) +
        f" -> {chosen}")
    b.substeps(dt, 1)  # captures the graphs with the chosen shapes
    b.sync()
    t0 = time.perf_counter()
    b.substeps(dt, steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    b.close()
    log(f"batched: {steps} substeps in {t_batch:.3f}s")
    if not sequential:
        return {
            "workload": f"batch_scenes: {n_scenes} scenes, particle counts log-uniform in "
                        "[2e3, 2e5], boxes U[32,128] d, 0-4 cubes, I = 4, FP32",
            "particles": n_total, "scenes": n_scenes, "steps": steps,
            "tile_shapes": chosen,
            "batched": {"particle_substeps_per_s": n_total * steps / t_batch,
                        "ms_per_substep": t_batch / steps * 1e3},
        }
    sims = []
    for p, sc in items:
        s = Sim(0)
        s.reset(p, sc)
        sims.append(s)
    for s in sims:
        s.substeps(dt, warmup)
        s.sync()
    t0 = time.perf_counter()
    for s in sims:
        s.substeps(dt, steps)
        s.sync()
    t_seq = time.perf_counter() - t0
    log(f"one by one: {steps} substeps per scene in {t_seq:.3f}s")
    for s in sims:
        s.close()
    return {
        "workload": f"batch_scenes: {n_scenes} scenes, particle counts log-uniform in "
                    "[2e3, 2e5], boxes U[32,128] d, 0-4 cubes, I = 4, FP32",
        "particles": n_total, "scenes": n_scenes, "steps": steps,
        "tile_shapes": chosen,
        "batched": {"particle_substeps_per_s": n_total * steps / t_batch,
                    "ms_per_substep": t_batch / steps * 1e3},
        "sequential": {"particle_substeps_per_s": n_total * steps / t_seq,
                       "ms_per_substep": t_seq / steps * 1e3},
        "speedup": t_seq / t_batch,
    }


def headline(args):
    from particles4all_amd import Sim, workloads
    from particles4all_amd._lib import copy_probe, lib

    if args.tiles:
        # read when a Sim is created
        os.environ["P4A_TILE1"], os.environ["P4A_TILE2"] = args.tiles.split(",")
    t0 = time.time()
    p, sc, dt = workloads.WORKLOADS[args.workload](scale=args.scale)
    if args.neighbour_skin is not None:
        p.neighbour_skin = args.neighbour_skin
    log(f"scene {args.workload}: n={sc.n} bodies={sc.n_bodies} built in {time.time() - t0:.1f}s")
    sim = Sim(0)
    sim.reset(p, sc)
    n = sim.info().n
    copy_ms = copy_probe(1 << 30, reps=5)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"device copy probe: {copy_gbs:.0f} GB/s (read+write) on 1 GiB")
    frame = 1.0 / (60.0 * max(1, p.substeps))  # one substep per step()
    settle(sim, frame, args.settle, args.workload)
    for _ in range(args.warmup):
        sim.step(frame)
    sim.sync()
    if args.tiles:
        chosen = [lib().p4a_tile_variant_name(g, int(v)).decode()
                  for g, v in zip((1, 2), args.tiles.split(","))]
        log(f"fixed shapes: {chosen}")
    else:
        chosen, tune_ms = sim.tune()
        log("autotuner: " + ", ".join(f"{k} {v:.3f}ms" for k, v in tune_ms.items()) +
            f" -> {chosen}")
    sim.step(frame)  # captures the graphs with the chosen shapes
    sim.sync()
    st0 = sim.stats()
    log(f"warmup done; stats avgRho {st0[0]:.4f} maxRho {st0[1]:.4f} maxSpeed {st0[2]:.3f}")
    sub, el = timed_steps(sim, frame, args.steps)
    value = n * sub / el
    log(f"timed: {args.steps} steps ({sub} substeps) in {el:.3f}s")

    # per-kernel timing pass (same work, eager launches with events)
    kt = kernel_timing(sim, frame, args.steps)
    dom = max((k for k in kt if k in ALGO_BYTES), key=lambda k: kt[k][0] * kt[k][1])
    dom_ms = kt[dom][0]
    achieved = ALGO_BYTES[dom] * n / (dom_ms * 1e-3) / 1e9
    st = sim.stats()
    log("kernels (mean ms/launch, launches): " +
        ", ".join(f"{k} {v[0]:.3f}x{v[1]}" for k, v in kt.items()))
    log(f"stats avgRho {st[0]:.4f} maxRho {st[1]:.4f} maxSpeed {st[2]:.3f}")
    prof_us, traffic = profiled(args.workload, dom)
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_particle": ALGO_BYTES[dom], "particles_per_launch": n,
            "mean_launch_ms": dom_ms, "duration_source": "HIP events on the simulation "
                                                         "stream, this run",
            "launch": "one stage: the mask build (first lambda after a grid build), or a "
                      "list launch with the fallback launch after it and the moved targets' "
                      "full walks on a second stream beside it (later stages); the stage's "
                      "HBM traffic (all three kernels) and wall time from profiles/" +
                      args.workload + "_fp32.md",
            "copy_probe_GBps": copy_gbs, "frac_of_copy_probe": achieved / copy_gbs}
    if prof_us:
        roof["profile_mean_launch_ms"] = prof_us / 1e3
        roof["frac_at_profile_duration"] = ALGO_BYTES[dom] * n / (prof_us * 1e-6) / 1e9 / \
            HBM_PEAK_GBS
    if traffic:
        roof["measured_GBps"] = traffic / (dom_ms * 1e-3) / 1e9
        roof["measured_frac"] = roof["measured_GBps"] / HBM_PEAK_GBS
    # the same byte model for every neighbour stage.  Every stage after the
    # mask build walks only the marked candidates (about a fifth of the 27
    # cells' records) and most gathers hit L1/L2, so a stage can finish faster
    # than the model's bytes would stream at peak: a fraction above 1 says the
    # model does not bound it (the measured traffic of the profile does)
    by_kernel = {}
    for k in ("lambda", "delta", "xsph", "tension"):
        if k in kt:
            by_kernel[k] = round(ALGO_BYTES[k] * n / (kt[k][0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    roof["frac_by_kernel"] = by_kernel
    measured = {}
    for k in by_kernel:
        us, tr = profiled(args.workload, k)
        if us and tr:
            measured[k] = round(tr / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    roof["measured_frac_by_kernel"] = measured
    out = {
        "metric": "particle-substeps/s at 4M particles, 4 iters; HBM GB/s vs peak; FP16 vs FP32",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} (scale {args.scale}): jittered lattice dam "
                               "break + 64 rigid bodies, 4 iterations, substep 1/120 s, timed "
                               f"after {args.settle} substeps of flow",
                   "particles": n, "bodies": sc.n_bodies, "iterations": p.iterations,
                   "settle_substeps": args.settle, "parallelism": "single GPU"},
        "state": {"avgRho": st0[0], "maxRho": st0[1], "maxSpeed": st0[2]},
        "roofline": roof,
        "kernels_ms": {k: round(v[0], 4) for k, v in kt.items()},
        "tile_shapes": chosen,
    }
    if not args.no_exact:
        out["exact_math"], _ = time_form(sim, p, frame, args.steps, 1, 0, value)
        log(f"exact math: {out['exact_math']['value']:.4g} particle-substeps/s "
            f"({out['exact_math']['vs_fast']:.3f}x fast)")
    if not args.no_fp16:
        fp16, kt16 = time_form(sim, p, frame, args.steps, 0, 1, value)
        solver = lambda k: sum(k[s][0] * k[s][1] for s in ("lambda", "delta") if s in k)
        fp16["speedup_substep"] = fp16["vs_fast"]
        fp16["speedup_solver_kernels"] = solver(kt) / max(solver(kt16), 1e-9)
        log(f"fp16-offset: {fp16['value']:.4g} particle-substeps/s "
            f"({fp16['speedup_substep']:.3f}x); solver kernels {fp16['speedup_solver_kernels']:.3f}x")
        if not args.no_accuracy:
            fp16["density_error_evolved"] = density_errors(sim, p, sc, frame)
            log(f"density error on the evolved state: {fp16['density_error_evolved']}")
        out["fp16_offset"] = fp16
    set_form(sim, p, 0, 0)
    cpu_s = None
    if not args.no_accuracy:
        out["accuracy_one_substep"], cpu_s = accuracy(sim, p, sc, frame)
    sim.close()
    return out, cpu_s, n


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU slab decomposition is design-only)"}))
        sys.exit(2)
    from particles4all_amd._lib import copy_probe

    if args.copy_probe:
        nbytes = args.copy_probe << 20
        ms = copy_probe(nbytes, reps=5)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms,
                          "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.field:
        f = field_measure(args.scale, args.field_settle, args.resolution, max(1, args.steps))
        print(json.dumps({"metric": "anisotropy pass + anisotropic density field splat, "
                                    "particles/s", "value": f["particles_per_s"],
                          "unit": "particles/s", "n_gpus": 1, "dtype": "f32",
                          "data": "synthetic", "field": f}), flush=True)
        return
    if args.batch:
        b = batch_measure(args.batch, args.steps, args.warmup, not args.no_sequential)
        print(json.dumps({"metric": "batched scenes: particle-substeps/s",
                          "value": b["batched"]["particle_substeps_per_s"],
                          "unit": "particle-substeps/s", "n_gpus": 1, "dtype": "f32",
                          "data": "synthetic", "batch": b}), flush=True)
        return

    out, cpu_s, n = headline(args)
    if not args.no_batch:
        out["batch"] = batch_measure(args.batch_scenes, args.batch_steps)
        log(f"batch: {out['batch']['speedup']:.2f}x over scene-by-scene launches")
    if not args.no_field:
        out["field"] = field_measure(args.scale, args.field_settle, args.resolution)
        log(f"field: aniso {out['field']['ms']['aniso']:.3f} ms, anisotropic splat "
            f"{out['field']['ms']['field_anisotropic']:.3f} ms")
    if not args.no_cpu_baseline and cpu_s:
        model, ncpu, usable = host_cpu()
        threads = int(os.environ.get("OMP_NUM_THREADS", usable or 1))
        out["cpu_baseline"] = {
            "value": n / cpu_s, "unit": "particle-substeps/s", "cores": threads,
            "kind": "port",
            "sample": f"1 substep of the same flowing {n}-particle state (the accuracy run's "
                      f"float chain): the oracle's float restatement, {threads} OpenMP "
                      f"threads, {cpu_s:.1f} s",
            "cpu_model": model, "nproc": ncpu, "usable_cpus": usable,
            "reference_wgsl_on_cpu_adapter": webgpu_probe()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
