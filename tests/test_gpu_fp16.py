"""GPU checks of the FP16-offset form of the solver iterations (lambda and
delta read neighbours as 8-byte records: x as float, y and z as half
offsets from the centre of the cell each particle was binned into),
through the C ABI.

Bars:
  * the records: after every position update (grid build, delta, shape
    matching) each record holds pred.x bit for bit and half(pred.yz -
    centre(binned cell).yz) rounded to nearest even, bit for bit; the
    binned cell equals the grid's cell id of the slot; lambda leaves them
    unchanged;
  * one stage at a time: density, lambda and delta equal the FP64 oracle
    evaluated at the quantized positions (delta applied to the FP32
    positions) within the
    fast-math bars for interior particles (2e-6 relative density,
    1e-5*max|lambda|, 1e-5*h plus one float ulp) and within ten times those
    (twice for delta) where wall samples enter (the absolute quantized
    position is a float there);
  * density against the FP64 oracle at the FP32 positions: mean
    |error| <= 1e-3*rho0 (the precision target), max printed;
  * a whole substep against the oracle's restatement of the form and
    against plain FP64 (see that test);
  * a captured substep replays bit-identically to stage-by-stage launches.
"""
import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float64


def rd(sim, name):
    return sim.read(name).astype(F)


def fp16_sim(scale=1.0, warm=4, bodies=True):
    """dam_break_32 after `warm` FP32 substeps, switched to the FP16 form."""
    p, sc, _ = workloads.dam_break_32(scale=scale, math=0, bodies=bodies)
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 1 / 120
    if warm:
        sim.substeps(dt, warm)
    p.precision = 1
    sim.set_params(p)
    return sim, p, sc, dt


def cell_xyz(P, c):
    gx, gy = P.gridDim[0], P.gridDim[1]
    c = c.astype(np.int64)
    return np.stack([c % gx, (c // gx) % gy, c // (gx * gy)], axis=1)


def centres32(P, cells):
    """boxMin + (x + 0.5) * h in float, as the kernels form it."""
    xyz = cell_xyz(P, cells).astype(np.float32)
    h = np.float32(P.h)
    bmin = np.array(P.boxMin[:], np.float32)
    return bmin + (xyz + np.float32(0.5)) * h


def records(sim):
    """(x float32, yz float16 pairs) of the FP16-offset records."""
    w = sim.read("OFS")
    return w[:, 0].copy().view(np.float32), w[:, 1:2].copy().view(np.float16)


def check_records(sim, P, where):
    pred = sim.read("PRED")
    cos = sim.read("CELL_OF_SLOT")
    x, yz = records(sim)
    assert np.array_equal(x.view(np.uint32), pred[:, 0].view(np.uint32)), where
    want = (pred[:, 1:3] - centres32(P, cos)[:, 1:]).astype(np.float16)
    bad = np.flatnonzero(np.any(yz.view(np.uint16) != want.view(np.uint16), axis=1))
    assert bad.size == 0, f"{where}: {bad.size} records differ, first slot {bad[:5]}"
    return pred, cos, (x, yz)


def quantized(P, cos, rec):
    x, yz = rec
    q = np.empty((len(x), 3))
    q[:, 0] = x
    q[:, 1:] = centres32(P, cos)[:, 1:].astype(F) + yz.astype(F)
    return q


def interior(P, pred, margin=3):
    c = cell_xyz(P, O.cell_ids(O.F32, P, pred.astype(F)))
    g = np.array(P.gridDim[:])
    return np.all((c >= margin) & (c <= g - 1 - margin), axis=1)


def oracle_boundary(sim):
    bpos = rd(sim, "BPOS")
    return bpos, bpos[:, 3].copy(), sim.read("BCELL_START")


def test_records_follow_every_position_update():
    sim, p, sc, dt = fp16_sim(scale=0.5)
    P = O.constants_from(p, sc, dt)
    for rep in range(2):
        sim.run_stage("PREDICT", dt)
        sim.run_stage("GRID", dt)
        check_records(sim, P, "grid")
        key, order = sim.read("KEY"), sim.read("ORDER")
        assert np.array_equal(sim.read("CELL_OF_SLOT"), key[order])
        if sc.n_bodies:
            sim.run_stage("BODY_PRE", dt)
        for it in range(p.iterations):
            before = sim.read("OFS")
            sim.run_stage("LAMBDA", dt)
            assert np.array_equal(sim.read("OFS"), before)
            sim.run_stage("DELTA", dt)
            check_records(sim, P, f"delta {it}")
            if sc.n_bodies:
                sim.run_stage("BODY_SOLVE", dt)
                check_records(sim, P, f"shape matching {it}")
        sim.run_stage("FINISH", dt)
    sim.close()


def test_stale_records_are_refused():
    p, sc, _ = workloads.dam_break_32(scale=0.25, math=0, bodies=False)
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 1 / 120
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)          # FP32 grid build: no records
    p.precision = 1
    sim.set_params(p)
    with pytest.raises(RuntimeError, match="stale"):
        sim.run_stage("LAMBDA", dt)
    sim.close()


def test_exact_math_with_fp16_is_rejected():
    p, sc, _ = workloads.dam_break_32(scale=0.25, math=1, bodies=False)
    p.precision = 1
    sim = Sim(0)
    with pytest.raises(RuntimeError):
        sim.reset(p, sc)
    sim.close()


def test_stages_against_fp64_oracle_on_quantized_positions():
    sim, p, sc, dt = fp16_sim(scale=1.0, bodies=False)
    P = O.constants_from(p, sc, dt)
    h = sc.h
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    bpos, bpsi, bcs = oracle_boundary(sim)
    cs = sim.read("CELL_START")
    pred32, cos, rec = check_records(sim, P, "grid")
    pred = pred32.astype(F)
    q = pred.copy()
    q[:, :3] = quantized(P, cos, rec)
    assert np.array_equal(q, O.fp16_quantize(P, pred32, cos))
    inner = interior(P, pred32)
    assert 0.3 < inner.mean() < 1.0
    sim.run_stage("LAMBDA", dt)
    gden, glam = rd(sim, "DENSITY"), rd(sim, "LAMBDA")
    den_q, lam_q = O.lambda_(O.F64, P, q, cs, bpos, bpsi, bcs)
    e_d = np.abs(gden - den_q) / den_q
    e_l = np.abs(glam - lam_q)
    # near rest, lambda carries the float error of rho/rho0 - 1: the bar
    # takes the float chain's own distance from FP64 on these inputs
    _, lam_q32 = O.lambda_(O.F32, P, q, cs, bpos, bpsi, bcs)
    chain = np.abs(lam_q32 - lam_q)
    lmax = np.abs(lam_q).max()
    print(f"density rel: interior max {e_d[inner].max():.3e}, all max {e_d.max():.3e}; "
          f"lambda: interior max {e_l[inner].max():.3e}, all max {e_l.max():.3e} "
          f"(max|lambda| {lmax:.3e})")
    # interior: fluid pairs only, distances exact -> fast-math bars; walls:
    # the absolute quantized position is rounded to float (<= 4e-6 here)
    assert e_d[inner].max() <= 2e-6 and e_d.max() <= 2e-5
    assert e_l[inner].max() <= max(1e-5 * lmax, 4 * chain[inner].max())
    assert e_l.max() <= max(1e-4 * lmax, 4 * chain.max())
    # density error of the form against the FP64 oracle at the FP32 positions
    den_t, _ = O.lambda_(O.F64, P, pred, cs, bpos, bpsi, bcs)
    err = np.abs(gden - den_t) / P.rho0
    print(f"density vs FP64 at FP32 positions: mean {err.mean():.3e} rho0, "
          f"max {err.max():.3e} rho0")
    assert err.mean() <= 1e-3
    # delta evaluated at the quantized positions, applied to the FP32 ones
    sim.run_stage("DELTA", dt)
    g = rd(sim, "PRED")
    out = O.delta_at(O.F64, P, q, pred, glam, cs, bpos, bpsi, bcs)
    ulp = np.spacing(np.abs(out[:, :3]).astype(np.float32)).astype(F)
    err = np.abs(g[:, :3] - out[:, :3])
    print(f"delta: interior max {err[inner].max():.3e}, all max {err.max():.3e} "
          f"(1e-5 h = {1e-5 * h:.1e})")
    assert np.all(err[inner] <= 1e-5 * h + ulp[inner])
    assert np.all(err <= 2e-5 * h + ulp)
    sim.close()


def test_one_substep_against_fp64_oracle():
    """A whole substep of the 33,792-particle dam break with its two bodies
    in the FP16 form, from the same state as the oracle.

    Against the oracle's restatement of the FP16 form (FP64 arithmetic at
    the quantized positions) the comparison is statistical: half rounding is
    discontinuous, so the ~1e-7 relative difference between a float and an
    FP64 evaluation of the same position moves some y or z offsets by one
    half ulp (up to 2.4e-4 h), and each such flip reaches the neighbours in the
    following iterations.  Bars: RMS within 5e-5 h (a fifth of that step),
    max within ten steps (2.4e-3 h).  The stage-by-stage test above pins the
    kernels on identical inputs.

    Against the plain FP64 oracle (the physical error of the form): density
    mean within 1e-3 rho0, and the position max and RMS within twice the
    error the FP64 restatement of the form itself shows (the quantization
    error of the start state, amplified by the compressed corner), plus
    1e-4 h."""
    p, sc, _ = workloads.dam_break_32(scale=1.0, math=0, precision=1)
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 1 / 120
    h = sc.h
    P = O.constants_from(p, sc, dt)
    args = (P, None, oracle_boundary(sim), np.array(sc.body_centre, F),
            np.tile(np.eye(3).reshape(1, 9), (sc.n_bodies, 1)),
            (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(F),
            p.iterations, p.surface_tension_k > 0)
    state = (rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    pid0 = sim.read("PID")
    res = {}
    for form in (True, False):
        st = O.State(*state)
        info = O.substep(O.F64, args[0], st, *args[2:], fp16=form)
        res[form] = (st.pos[:, :3].copy(), info)
    sim.substeps(dt, 1)
    assert np.array_equal(sim.read("PID"), pid0[res[True][1]["order"]])
    gp = rd(sim, "POS")[:, :3]
    gd = rd(sim, "DENSITY")
    rms = lambda e: float(np.sqrt((e ** 2).mean()))
    e_em = np.abs(gp - res[True][0])
    e_64 = np.abs(gp - res[False][0])
    e_in = np.abs(res[True][0] - res[False][0])
    ed = np.abs(gd - res[False][1]["density"]) / P.rho0
    frac = float(np.mean(e_em <= 1e-5 * h))
    print(f"vs FP16 restatement: max {e_em.max() / h:.3e} h rms {rms(e_em) / h:.3e} h, "
          f"within 1e-5 h {frac:.5f}")
    print(f"vs FP64: max {e_64.max() / h:.3e} h rms {rms(e_64) / h:.3e} h; restatement vs "
          f"FP64: max {e_in.max() / h:.3e} h rms {rms(e_in) / h:.3e} h; density mean "
          f"{ed.mean():.3e} max {ed.max():.3e} rho0")
    assert rms(e_em) <= 5e-5 * h and e_em.max() <= 2.4e-3 * h
    assert e_64.max() <= 2 * e_in.max() + 1e-4 * h
    assert rms(e_64) <= 2 * rms(e_in) + 1e-4 * h
    assert ed.mean() <= 1e-3
    sim.close()


def test_graph_replay_equals_stage_by_stage():
    simA, p, sc, dt = fp16_sim(scale=0.5, warm=0)
    simB, _, _, _ = fp16_sim(scale=0.5, warm=0)
    simA.substeps(dt, 6)
    for _ in range(6):
        simB.run_stage("PREDICT", dt)
        simB.run_stage("GRID", dt)
        if sc.n_bodies:
            simB.run_stage("BODY_PRE", dt)
        for _ in range(p.iterations):
            simB.run_stage("LAMBDA", dt)
            simB.run_stage("DELTA", dt)
            if sc.n_bodies:
                simB.run_stage("BODY_SOLVE", dt)
        simB.run_stage("FINISH", dt)
    for name in ("POS", "VEL", "PID", "DENSITY", "LAMBDA"):
        assert np.array_equal(simA.read(name), simB.read(name)), name
    simA.close()
    simB.close()
