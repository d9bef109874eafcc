"""The Sim's interaction entry points on the GPU against the oracle, through
the C ABI: ray impulse (impulseWGSL, applyRayImpulse), held-body drag
(bodyDragWGSL inside the predict stage), box resize (resizeWGSL plus the
rebuilt walls and grid) and poured particles (appendFluid).

Bars (exact math, where every kernel is one IEEE operation per source
operation): velocities after an impulse and after a drag, the positions
predicted from the dragged velocities,
positions after a resize, the resized walls (sortBoundary of the new
box's samples) and the appended particles' records are bit-identical to
the oracle; a whole substep after a resize and after an append is
bit-identical to the oracle's float chain on the new state.
"""
import ctypes as C

import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from particles4all_amd._lib import lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float64


def rd(sim, name):
    return sim.read(name).astype(F)


def same(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape
    bad = np.flatnonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=-1))
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:4]}: {a[bad[:2]]} vs {b[bad[:2]]}"


def moved(bodies=True, pour=0, substeps=3):
    p, sc, _ = workloads.dam_break_32(scale=1.0, math=1, bodies=bodies, pour_budget=pour)
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(1 / 120, substeps)
    return sim, p, sc


@pytest.mark.parametrize("origin,direction,imp,radius,limit", [
    ([-5.0, 16.0, 16.0], [1.0, 0.1, 0.0], [0.0, 30.0, 0.0], 6.0, 5.0),    # caps some
    ([16.0, 40.0, 16.0], [3.0, -4.0, 0.5], [2.0, 1.0, -3.0], 4.5, 100.0),
    ([10.0, 10.0, 10.0], [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], 8.0, 10.0),     # dir 0: no hit
])
def test_impulse_bit_exact(origin, direction, imp, radius, limit):
    sim, p, sc = moved()
    pos, vel = sim.read("POS"), sim.read("VEL")
    want = O.impulse(pos, vel, origin, direction, imp, radius, limit)
    sim.apply_ray_impulse(origin, direction, imp, radius, limit)
    got = sim.read("VEL")
    hit = np.any(want != vel, axis=1).sum()
    print(f"impulse: {hit} particles changed")
    same(got, want)
    assert np.array_equal(sim.read("POS"), pos)
    if direction != [0.0, 0.0, 0.0]:
        assert hit > 10
    sim.close()


def test_held_body_drag_bit_exact():
    sim, p, sc = moved()
    dt = 1 / 120
    P = O.constants_from(p, sc, dt)
    goal, rate, limit = [20.0, 50.0, 12.0], 3.0, 9.0
    sim.hold_body(1, goal, rate, limit)
    pos, vel = rd(sim, "POS"), sim.read("VEL")
    body = sim.read("BODY")[:, 0]
    centre = sim.read("BODY_CENTRE")[0]
    members = np.flatnonzero(body == 1)
    assert len(members) > 100
    dragged = O.body_drag(vel, members, centre, goal, rate, limit, dt)
    assert not np.array_equal(dragged, vel)
    sim.run_stage("PREDICT", dt)
    # the predict kernel leaves the velocities as the drag left them
    # (velFromPos overwrites them before any read); the predicted positions
    # come from the dragged velocities
    same(sim.read("VEL"), dragged)
    sim.run_stage("GRID", dt)
    pred, _ = O.predict(O.F32, P, pos, dragged.astype(F))
    cs, slot, order = O.grid(P, O.cell_ids(O.F32, P, pred))
    same(sim.read("PRED"), pred[order])
    same(sim.read("VEL"), dragged[order])
    # a released body is not dragged
    sim.release_body()
    vel = sim.read("VEL")
    sim.run_stage("PREDICT", dt)
    same(sim.read("VEL"), vel)
    sim.close()


def boundary_of(box, d, h, rho0):
    """The walls of `box` from the library's host builder (boundaryParticles
    restated; pinned against the Python restatement by the scene tests)."""
    L = lib()
    b = (C.c_double * 3)(*box)
    nb = C.c_int32()
    assert L.p4a_boundary_build(b, d, h, rho0, None, None, 0, C.byref(nb)) == 0
    pts = np.zeros((max(1, nb.value), 3), np.float32)
    psi = np.zeros(max(1, nb.value), np.float32)
    assert L.p4a_boundary_build(b, d, h, rho0, pts.ctypes.data, psi.ctypes.data, nb.value,
                                C.byref(nb)) == 0
    return pts[: nb.value], psi[: nb.value]


def test_resize_bit_exact_and_the_next_substep():
    sim, p, sc = moved(bodies=False)
    dt = 1 / 120
    old = list(p.box)
    new = [48.0, 72.0, 64.0]  # x shrinks, y grows, z stays
    pos = sim.read("POS")
    want = O.resize(pos, old, new, p.spacing)
    sim.resize_box(new)
    same(sim.read("POS"), want)
    info = sim.info()
    assert info.n_cells == 24 * 36 * 32
    # walls of the new box, sorted into the new grid
    p.box = (C.c_double * 3)(*new)
    P = O.constants_from(p, sc, dt)
    pts, psi = boundary_of(new, p.spacing, sc.h, p.rest_density)
    P.nBoundary = len(pts)
    rec, start = O.sort_boundary(P, pts, psi)
    assert info.n_boundary == len(pts)
    assert np.array_equal(rd(sim, "BPOS"), rec)
    assert np.array_equal(sim.read("BCELL_START"), start)
    # one substep on the resized state against the float chain
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    pid0 = sim.read("PID")
    bnd = (rec, rec[:, 3].copy(), start)
    info2 = O.substep(O.F32, P, st, bnd, np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0),
                      p.iterations, p.surface_tension_k > 0)
    sim.substeps(dt, 1)
    assert np.array_equal(sim.read("PID"), pid0[info2["order"]])
    same(sim.read("POS"), st.pos)
    same(sim.read("VEL")[:, :3], st.vel[:, :3])
    sim.close()


def test_append_bit_exact_and_the_next_substep():
    sim, p, sc = moved(bodies=False, pour=300)
    dt = 1 / 120
    i0 = sim.info()
    rng = np.random.default_rng(9)
    k = 200
    pos3 = rng.uniform([4, 40, 4], [28, 60, 28], size=(k, 3)).astype(np.float32)
    vel3 = rng.uniform(-3, 3, size=(k, 3)).astype(np.float32)
    assert sim.append_fluid(pos3, vel3) == k
    n0, n1 = i0.n, i0.n + k
    assert sim.info().n == n1
    want_p = np.concatenate([pos3, np.ones((k, 1), np.float32)], 1)
    want_v = np.concatenate([vel3, np.zeros((k, 1), np.float32)], 1)
    same(sim.read("POS")[n0:], want_p)
    same(sim.read("VEL")[n0:], want_v)
    same(sim.read("PRED")[n0:], want_p)
    assert np.all(sim.read("DENSITY")[n0:] == np.float32(p.rest_density))
    assert np.all(sim.read("LAMBDA")[n0:] == 0)
    for name in ("CORR", "NORMAL", "REST"):
        assert np.all(sim.read(name)[n0:] == 0), name
    assert np.all(sim.read("BODY")[n0:] == 0)
    assert np.array_equal(sim.read("PID")[n0:], np.arange(n0, n1, dtype=np.uint32))
    # the room left is the pour budget minus what was poured
    more = np.tile(pos3[:1], (200, 1))
    assert sim.append_fluid(more, np.zeros_like(more)) == 100
    assert sim.append_fluid(more, np.zeros_like(more)) == 0
    # one substep with the poured particles against the float chain
    P = O.constants_from(p, sc, dt)
    bpos = rd(sim, "BPOS")
    bnd = (bpos, bpos[:, 3].copy(), sim.read("BCELL_START"))
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    pid0 = sim.read("PID")
    info = O.substep(O.F32, P, st, bnd, np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0),
                     p.iterations, p.surface_tension_k > 0)
    sim.substeps(dt, 1)
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    same(sim.read("POS"), st.pos)
    same(sim.read("VEL")[:, :3], st.vel[:, :3])
    sim.close()


@pytest.mark.parametrize("math", [0, 1])
def test_stats_and_body_pose(math):
    """stats (wgsl.js:785-832) after a frame against the oracle's reduction of
    the read-back densities and velocities: the maxima exactly, the sums to
    FP64 rounding (the reduction trees differ); the body poses in the
    reference's readback layout (sim.js:635-655: rot column-major from the
    stored rows)."""
    p, sc, frame = workloads.dam_break_32(scale=1.0, math=math)
    sim = Sim(0)
    sim.reset(p, sc)
    for _ in range(3):
        sim.step(frame)
    got = sim.stats()
    want = O.stats(rd(sim, "DENSITY"), rd(sim, "VEL"), 1.0 / p.rest_density, sc.mass)
    print(f"stats: {got}")
    assert got[1] == want[1] and got[2] == want[2]
    np.testing.assert_allclose([got[0], got[3]], [want[0], want[3]], rtol=1e-12, atol=0)
    assert 0.5 < got[0] < 1.5 and got[3] > 0
    c, r = sim.body_pose()
    centre, rows = sim.read("BODY_CENTRE"), sim.read("BODY_ROT").reshape(-1, 3, 4)
    assert np.array_equal(c, centre[:, :3])
    assert np.array_equal(r.reshape(-1, 3, 3), rows[:, :, :3].transpose(0, 2, 1))
    R = r.reshape(-1, 3, 3).astype(F)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
    sim.close()
