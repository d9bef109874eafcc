"""Benchmark of the simulation step on one MI355X.

Workload (the headline configuration): block_4m_bodies - a 256x128x128
jittered fluid lattice (4,194,304 particles, spacing d = 1, h = 2d) in a
512x192x128 box plus 64 rigid bodies (32 lattice cubes with 6-12 particle
edges, 32 lattice spheres with 4-8 particle radii, densities log-uniform in
[0.3, 2.0]); 4 solver iterations, substep 1/120 s, XSPH 0.01, surface
tension 0.5, FP32.  One step = one Sim.step(1/120) = one substep (predict,
grid build, body centroids, 4 x (lambda, delta, shape matching), xsph +
normals, tension + finalize) plus the stats reduction, all resident in HBM.

Prints one JSON line; `value` is particle-substeps/s.  After the timed
region a second pass of the same steps records per-kernel HIP events on
the simulation stream (eager launches) for the roofline of the dominant
kernel.  The FP16-offset form (solver iterations read 8-byte half-offset
records) is then timed the same way on the same simulation, and its
accuracy is measured against the FP64 oracle: densities of both forms on
one shared evolved state, and positions/densities after one substep from
the initial scene.  The CPU baseline times the oracle's float restatement
of the same substep (OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64-oracle accuracy figures of the FP16-offset form")
    ap.add_argument("--cpu-substeps", type=int, default=2)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the device copy probe on MIB mebibytes and exit")
    ap.add_argument("--field", action="store_true",
                    help="instead of the headline line: settle dam_break_2m, then time the "
                         "anisotropy pass and the 256^3 density fields (one JSON line)")
    ap.add_argument("--settle", type=int, default=600,
                    help="substeps of settling before the field measurement")
    ap.add_argument("--resolution", type=int, default=256)
    ap.add_argument("--tiles", default="", metavar="G1,G2",
                    help="fix the traversal shapes (variant numbers of groups 1 and 2) "
                         "instead of autotuning, e.g. for counter runs that must profile "
                         "the shapes a timed run picked")
    ap.add_argument("--batch", type=int, default=0, metavar="SCENES",
                    help="instead of the headline line: SCENES independent scenes batched "
                         "against the same scenes stepped one by one (one JSON line)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, sc, dt, substeps):
    """The oracle's float-mode substep (OpenMP threads) on the same scene."""
    import numpy as np
    from oracle import oracle as O
    P = O.constants_from(p, sc, dt)
    rec, start = O.sort_boundary(P, sc.boundary_pts, sc.boundary_psi)
    st = O.State(sc.pos.astype(np.float64), sc.vel.astype(np.float64), sc.phase[:, 0],
                 sc.rest.astype(np.float64))
    inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(np.float64)
    centre = np.array(sc.body_centre, np.float64)
    rot = np.tile(np.eye(3).reshape(1, 9), (max(sc.n_bodies, 0), 1))
    t0 = time.perf_counter()
    for _ in range(substeps):
        info = O.substep(O.F32, P, st, (rec, rec[:, 3].copy(), start), centre, rot, inv,
                         p.iterations, p.surface_tension_k > 0)
        centre, rot = info["centre"], info["rot"]
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sc.n * substeps / el, "unit": "particle-substeps/s", "cores": threads,
            "kind": "port",
            "sample": f"{substeps} substep(s) of the full {sc.n}-particle workload, oracle float "
                      f"restatement, {threads} OpenMP threads, {el:.1f} s"}


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    this workload (FETCH_SIZE and WRITE_SIZE passes, copy-probe calibrated),
    or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_fp32.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC "):
            vals = dict(kv.split("=") for kv in line.split()[1:])
            return int(vals[kernel]) if kernel in vals else None
    return None


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


def density_errors(sim, p, sc, frame):
    """Densities of the FP16 and the FP32 form on the same predicted state
    (the simulation's current one) against the FP64 oracle; |error|/rho0."""
    import numpy as np
    from oracle import oracle as O
    P = O.constants_from(p, sc, frame)
    p.precision = 1
    sim.set_params(p)
    sim.run_stage("PREDICT", frame)
    sim.run_stage("GRID", frame)
    pred = sim.read("PRED").astype(np.float64)
    cs = sim.read("CELL_START")
    bpos = sim.read("BPOS").astype(np.float64)
    bcs = sim.read("BCELL_START")
    sim.run_stage("LAMBDA", frame)
    d16 = sim.read("DENSITY").astype(np.float64)
    p.precision = 0
    sim.set_params(p)
    sim.run_stage("LAMBDA", frame)
    d32 = sim.read("DENSITY").astype(np.float64)
    den, _ = O.lambda_(O.F64, P, pred, cs, bpos, bpos[:, 3].copy(), bcs)
    out = {}
    for name, d in (("fp16_offset", d16), ("fp32", d32)):
        e = np.abs(d - den) / P.rho0
        out[name] = {"mean": float(e.mean()), "max": float(e.max())}
    return out


def substep_errors(p, sc, dt):
    """One substep from the initial scene in both forms against the FP64
    oracle, which starts from the state read back after the reset (the live
    cell-sorted order): position error (units of h) and density error
    (units of rho0)."""
    import numpy as np
    from particles4all_amd import Sim
    from oracle import oracle as O
    P = O.constants_from(p, sc, dt)
    rec, start = O.sort_boundary(P, sc.boundary_pts, sc.boundary_psi)
    inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(np.float64)
    rot = np.tile(np.eye(3).reshape(1, 9), (max(sc.n_bodies, 0), 1))
    out, st, info, pid0 = {}, None, None, None
    for name, prec in (("fp16_offset", 1), ("fp32", 0)):
        p.precision = prec
        sim = Sim(0)
        sim.reset(p, sc)
        if st is None:
            f64 = lambda k: sim.read(k).astype(np.float64)
            st = O.State(f64("POS"), f64("VEL"), sim.read("BODY")[:, 0], f64("REST"))
            pid0 = sim.read("PID")
            info = O.substep(O.F64, P, st, (rec, rec[:, 3].copy(), start),
                             np.array(sc.body_centre, np.float64), rot, inv, p.iterations,
                             p.surface_tension_k > 0)
        sim.substeps(dt, 1)
        same = bool(np.array_equal(sim.read("PID"), pid0[info["order"]]))
        pos = sim.read("POS").astype(np.float64)[:, :3]
        den = sim.read("DENSITY").astype(np.float64)
        sim.close()
        ep = np.abs(pos - st.pos[:, :3]) / sc.h
        ed = np.abs(den - info["density"]) / P.rho0
        out[name] = {"same_order": same, "pos_max_h": float(ep.max()),
                     "pos_rms_h": float(np.sqrt((ep ** 2).mean())),
                     "density_mean": float(ed.mean()), "density_max": float(ed.max())}
    p.precision = 0
    return out


def field_bench(args):
    """Anisotropy pass + density fields on the settled 2,097,152-particle dam
    break: HIP-event times of the anisotropy kernel, the anisotropic splat
    into the 256^3 lattice and the isotropic (reference) field, plus
    size-independent checks of the outputs."""
    import numpy as np
    from particles4all_amd import Sim, workloads
    p, sc, frame = workloads.dam_break_2m(scale=args.scale)
    sim = Sim(0)
    sim.reset(p, sc)
    n = sim.info().n
    dt = 1.0 / 120.0
    t0 = time.perf_counter()
    done = 0
    while done < args.settle:
        k = min(60, args.settle - done)
        sim.substeps(dt, k)
        sim.sync()
        done += k
        log(f"settling: {done}/{args.settle} substeps, {time.perf_counter() - t0:.1f}s")
    st = sim.stats()
    box = list(p.box)
    sim.field_configure([0, 0, 0], box, args.resolution)
    fi = sim.field_info()
    # SurfaceMesh's anisoLimitToField default: extents >= 1.25 voxel / h
    # (mesh.js:238)
    akw = dict(min_extent=1.25 * fi.voxel / sc.h)
    sim.aniso(**akw)
    sim.field_generate(anisotropic=True)
    sim.field_generate(anisotropic=False)
    sim.sync()
    reps = max(1, args.steps)
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
    # size-independent checks: G symmetric positive definite (det > 0),
    # neighbour counts >= 1 (self), and the anisotropic field integrates the
    # particles' volumes (each kernel integrates to V_j)
    g = sim.read("ANISO_G").reshape(-1, 2, 4)
    det_ok = float(np.mean(g[:, 1, 2] > 0))
    counts = g[:, 1, 3]
    sim.field_generate(anisotropic=True)
    field = sim.read("FIELD").astype(np.float64)
    # the surface of generate: normal field (2 passes, the default of the
    # reference's UI) and the marching tetrahedra at meshIso 0.4 (one untimed
    # round first)
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
    out = {
        "metric": "anisotropy pass + anisotropic density field splat, particles/s",
        "value": n / ((a_ms + f_ms) * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": a_ms + f_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"dam_break_2m (scale {args.scale}): jittered 128^3 lattice in a "
                               f"256^3 box, settled for {args.settle} substeps of 1/120 s",
                   "particles": n, "field_vert_dim": list(fi.vert_dim), "voxel": fi.voxel,
                   "aniso": "R = 2h, lambda 0.9, kr 4, ks 1, stretch 2, kn 1, neps 25, "
                            "min extent 1.25 voxel/h, FP64 re-solve below a 1e-4 eigen-gap"},
        "ms": {"aniso": a_ms, "field_anisotropic": f_ms, "field_isotropic": i_ms,
               "normal_field_2_passes": ktm["normal field"][0], "march": ktm["march"][0]},
        "surface": {"iso": 0.4, "triangles": tri_total, "written": tri_written},
        "fp64_solves": solves,
        "checks": {"det_G_positive": det_ok, "min_count": float(counts.min()),
                   "mean_count": float(counts.mean()),
                   "field_integral_over_volume": integral / float(vol.sum())},
        "settled_stats": {"avgRho": st[0], "maxRho": st[1], "maxSpeed": st[2]},
    }
    print(json.dumps(out), flush=True)


def batch_bench(args):
    """SCENES scenes of 2e3-2e5 particles: one batched graph per substep
    against each scene's own graph replayed in turn (synchronised after each
    scene, so scenes do not overlap), the same substeps of the same scenes."""
    from particles4all_amd import Batch, Sim, workloads
    t0 = time.time()
    items = workloads.batch_scenes(n_scenes=args.batch)
    n_total = sum(sc.n for _, sc in items)
    log(f"{args.batch} scenes, {n_total} particles, built in {time.time() - t0:.1f}s")
    dt = 1.0 / 120.0
    K = args.steps
    b = Batch(items)
    b.substeps(dt, args.warmup)
    b.sync()
    chosen, tune_ms = b.tune(dt)
    log("batch autotuner: " + ", ".join(f"{k} {v:.3f}ms" for k, v in tune_ms.items()) +
        f" -> {chosen}")
    b.substeps(dt, 1)  # captures the graphs with the chosen shapes
    b.sync()
    t0 = time.perf_counter()
    b.substeps(dt, K)
    b.sync()
    t_batch = time.perf_counter() - t0
    b.close()
    log(f"batched: {K} substeps in {t_batch:.3f}s")
    sims = []
    for p, sc in items:
        s = Sim(0)
        s.reset(p, sc)
        sims.append(s)
    for s in sims:
        s.substeps(dt, args.warmup)
        s.sync()
    t0 = time.perf_counter()
    for s in sims:
        s.substeps(dt, K)
        s.sync()
    t_seq = time.perf_counter() - t0
    log(f"one by one: {K} substeps per scene in {t_seq:.3f}s")
    for s in sims:
        s.close()
    out = {
        "metric": "batched scenes: particle-substeps/s, speed-up over scene-by-scene launches",
        "value": n_total * K / t_batch,
        "unit": "particle-substeps/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": t_batch / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch_scenes: {args.batch} scenes, particle counts "
                               "log-uniform in [2e3, 2e5], boxes U[32,128] d, 0-4 cubes, I = 4",
                   "particles": n_total, "scenes": args.batch},
        "tile_shapes": chosen,
        "sequential": {"value": n_total * K / t_seq, "ms_per_step": t_seq / K * 1e3},
        "speedup": t_seq / t_batch,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU slab decomposition is design-only)"}))
        sys.exit(2)
    from particles4all_amd import Sim, workloads
    from particles4all_amd._lib import copy_probe, lib

    if args.field:
        field_bench(args)
        return
    if args.batch:
        batch_bench(args)
        return
    if args.copy_probe:
        nbytes = args.copy_probe << 20
        ms = copy_probe(nbytes, reps=5)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms,
                          "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return

    if args.tiles:
        # read when a Sim is created
        os.environ["P4A_TILE1"], os.environ["P4A_TILE2"] = args.tiles.split(",")
    t0 = time.time()
    p, sc, dt = workloads.WORKLOADS[args.workload](scale=args.scale)
    log(f"scene {args.workload}: n={sc.n} bodies={sc.n_bodies} built in {time.time() - t0:.1f}s")
    sim = Sim(0)
    sim.reset(p, sc)
    n = sim.info().n
    copy_ms = copy_probe(1 << 30, reps=5)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"device copy probe: {copy_gbs:.0f} GB/s (read+write) on 1 GiB")
    frame = 1.0 / (60.0 * max(1, p.substeps))  # one substep per step()
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
    log("warmup done")
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

    # FP16-offset form on the same simulation
    p.precision = 1
    sim.set_params(p)
    for _ in range(2):
        sim.step(frame)
    sub16, el16 = timed_steps(sim, frame, args.steps)
    value16 = n * sub16 / el16
    kt16 = kernel_timing(sim, frame, args.steps)
    solver = lambda k: sum(k[s][0] * k[s][1] for s in ("lambda", "delta") if s in k)
    fp16 = {"value": value16, "speedup_substep": value16 / value,
            "speedup_solver_kernels": solver(kt) / max(solver(kt16), 1e-9),
            "kernels_ms": {k: round(v[0], 4) for k, v in kt16.items() if k in ("lambda", "delta")}}
    log(f"fp16-offset: {value16:.4g} particle-substeps/s ({fp16['speedup_substep']:.3f}x); "
        f"lambda {kt16['lambda'][0]:.3f} ms, delta {kt16['delta'][0]:.3f} ms")
    if not args.no_accuracy:
        fp16["density_error_evolved"] = density_errors(sim, p, sc, frame)
        log(f"density error on the evolved state: {fp16['density_error_evolved']}")
    p.precision = 0

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
                               "break + 64 rigid bodies, 4 iterations, substep 1/120 s",
                   "particles": n, "bodies": sc.n_bodies, "iterations": p.iterations,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.workload, dom),
                     "algorithmic_bytes_per_particle": ALGO_BYTES[dom],
                     "mean_launch_ms": dom_ms, "copy_probe_GBps": copy_gbs,
                     # the same model for every neighbour kernel (tension's figure
                     # counts the row tests it takes from xsph's masks as gathered)
                     "frac_by_kernel": {k: round(ALGO_BYTES[k] * n / (kt[k][0] * 1e-3) / 1e9
                                                 / HBM_PEAK_GBS, 4)
                                        for k in ALGO_BYTES if k in kt}},
        "kernels_ms": {k: round(v[0], 4) for k, v in kt.items()},
        "tile_shapes": chosen,
        "fp16_offset": fp16,
    }
    sim.close()
    if not args.no_accuracy:
        fp16["first_substep_vs_fp64"] = substep_errors(p, sc, frame)
        log(f"first substep vs FP64 oracle: {fp16['first_substep_vs_fp64']}")
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, sc, dt, args.cpu_substeps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
