"""Solver parameters on the GPU against the oracle's float chain, exact
math, through the C ABI: the iteration count (the predicted positions
alternate between two buffers, so odd and even counts end in different
ones), both substep parities (one captured graph each), and parameters
changed between substeps by set_params (iterations, XSPH, surface
tension, gravity, SOR weighting and walls off), which rebuild the graphs.

Bar: positions, velocities, lambdas, densities and the permutation equal
the oracle's float chain bit for bit.
"""
import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float64
DT = 1 / 120


def rd(sim, name):
    return sim.read(name).astype(F)


def checked_substep(sim, p, sc, from_reset_bodies=False):
    """One GPU substep against the oracle started from the GPU's own state.
    Bodies are only compared from the reset pose (identity rotations)."""
    P = O.constants_from(p, sc, DT)
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    pid0 = sim.read("PID")
    bpos = rd(sim, "BPOS")
    bnd = (bpos, bpos[:, 3].copy(), sim.read("BCELL_START"))
    if from_reset_bodies:
        centre = np.array(sc.body_centre, F)
        rot = np.tile(np.eye(3).reshape(1, 9), (sc.n_bodies, 1))
        inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(F)
    else:
        centre, rot, inv = np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0)
    info = O.substep(O.F32, P, st, bnd, centre, rot, inv, p.iterations, p.surface_tension_k > 0)
    sim.substeps(DT, 1)
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    assert np.array_equal(rd(sim, "POS"), st.pos)
    assert np.array_equal(rd(sim, "VEL")[:, :3], st.vel[:, :3])
    assert np.array_equal(rd(sim, "LAMBDA"), info["lam"])
    assert np.array_equal(rd(sim, "DENSITY"), info["density"])


@pytest.mark.parametrize("iterations", [1, 2, 3, 8])
@pytest.mark.parametrize("bodies", [False, True])
def test_iteration_counts_bit_exact(iterations, bodies):
    p, sc, _ = workloads.dam_break_32(scale=0.5, math=1, bodies=bodies)
    p.iterations = iterations
    p.surface_tension_k = 0.5
    sim = Sim(0)
    sim.reset(p, sc)
    print(f"iterations={iterations} bodies={bodies} n={sim.info().n}", flush=True)
    checked_substep(sim, p, sc, from_reset_bodies=bodies)
    if not bodies:
        checked_substep(sim, p, sc)   # the other parity's graph
    sim.close()


def test_set_params_between_substeps_bit_exact():
    p, sc, _ = workloads.dam_break_32(scale=0.5, math=1, bodies=False)
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(DT, 2)
    p.iterations = 3
    p.xsph_c = 0.02
    p.surface_tension_k = 0.5
    p.gravity = 4.0
    p.omega = 1.3
    p.sor_average = 0 if p.sor_average else 1
    sim.set_params(p)
    print(f"after set_params: iterations=3 omega=1.3 sor_average={p.sor_average}", flush=True)
    checked_substep(sim, p, sc)
    checked_substep(sim, p, sc)
    p.no_boundary = 1
    sim.set_params(p)
    print("after set_params: no_boundary=1", flush=True)
    checked_substep(sim, p, sc)
    p.no_boundary = 0
    p.iterations = 4
    p.surface_tension_k = 0.0
    sim.set_params(p)
    checked_substep(sim, p, sc)
    sim.close()


def _err(a, b):
    d = np.abs(a - b)
    return float(d.max()), float(np.sqrt((d ** 2).mean()))


@pytest.mark.parametrize("iterations", [1, 3])
@pytest.mark.parametrize("precision", [0, 1])
def test_fast_forms_odd_iteration_counts(iterations, precision):
    """Fast math with odd iteration counts, FP32 and FP16-offset forms,
    against the FP64 oracle (the FP16 form against the oracle's FP64
    restatement of that form, evaluated at the same quantized positions).
    Bars as for four iterations: positions within 1e-4 h at most and 1e-6 h
    RMS (FP32), 1e-3 h and 1e-5 h (FP16); velocities within the position
    bar over dt."""
    p, sc, _ = workloads.dam_break_32(scale=0.5, math=0, bodies=False, precision=precision)
    p.iterations = iterations
    p.surface_tension_k = 0.5
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(DT, 1)
    h = sc.h
    P = O.constants_from(p, sc, DT)
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    pid0 = sim.read("PID")
    bpos = rd(sim, "BPOS")
    bnd = (bpos, bpos[:, 3].copy(), sim.read("BCELL_START"))
    info = O.substep(O.F64, P, st, bnd, np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0),
                     iterations, True, fp16=bool(precision))
    sim.substeps(DT, 1)
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    ep = _err(rd(sim, "POS")[:, :3], st.pos[:, :3])
    ev = _err(rd(sim, "VEL")[:, :3], st.vel[:, :3])
    print(f"iterations={iterations} precision={precision}: pos max {ep[0] / h:.3e} h "
          f"rms {ep[1] / h:.3e} h; vel max {ev[0]:.3e}", flush=True)
    bar, rms = (1e-3, 1e-5) if precision else (1e-4, 1e-6)
    assert ep[0] <= bar * h and ep[1] <= rms * h
    assert ev[0] <= bar * h / DT
    sim.close()
