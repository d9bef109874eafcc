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
kernel.  The CPU baseline times the oracle's float restatement of the same
substep (OpenMP) on a bounded sample of the same workload.
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
    ap.add_argument("--cpu-substeps", type=int, default=2)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the device copy probe on MIB mebibytes and exit")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU slab decomposition is design-only)"}))
        sys.exit(2)
    from particles4all_amd import Sim, workloads
    from particles4all_amd._lib import copy_probe

    if args.copy_probe:
        nbytes = args.copy_probe << 20
        ms = copy_probe(nbytes, reps=5)
        print(json.dumps({"copy_bytes": nbytes, "ms": ms,
                          "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return

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
    chosen, tune_ms = sim.tune()
    log("autotuner: " + ", ".join(f"{k} {v:.3f}ms" for k, v in tune_ms.items()) +
        f" -> {chosen}")
    sim.step(frame)  # captures the graphs with the chosen shapes
    sim.sync()
    log("warmup done")
    t0 = time.perf_counter()
    sub = 0
    for _ in range(args.steps):
        sub += sim.step(frame)
    sim.sync()
    el = time.perf_counter() - t0
    value = n * sub / el
    log(f"timed: {args.steps} steps ({sub} substeps) in {el:.3f}s")

    # per-kernel timing pass (same work, eager launches with events)
    sim.set_timing(True)
    sim.substeps(frame, args.steps)
    sim.sync()
    kt = sim.timing()
    sim.set_timing(False)
    dom = max((k for k in kt if k in ALGO_BYTES), key=lambda k: kt[k][0] * kt[k][1])
    dom_ms = kt[dom][0]
    achieved = ALGO_BYTES[dom] * n / (dom_ms * 1e-3) / 1e9
    st = sim.stats()
    log("kernels (mean ms/launch, launches): " +
        ", ".join(f"{k} {v[0]:.3f}x{v[1]}" for k, v in kt.items()))
    log(f"stats avgRho {st[0]:.4f} maxRho {st[1]:.4f} maxSpeed {st[2]:.3f}")

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
                     "traffic": None,
                     "algorithmic_bytes_per_particle": ALGO_BYTES[dom],
                     "mean_launch_ms": dom_ms, "copy_probe_GBps": copy_gbs},
        "kernels_ms": {k: round(v[0], 4) for k, v in kt.items()},
        "tile_shapes": chosen,
        "fp16_offset": None,
    }
    if not args.no_cpu_baseline:
        sim.close()
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, sc, dt, args.cpu_substeps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
