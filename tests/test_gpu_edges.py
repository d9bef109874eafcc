"""GPU edge cases, through the C ABI.

* an empty scene (n = 0) resets, steps, reports stats and takes poured
  particles, which then step like any others;
* one particle with no neighbour falls freely: after k substeps its
  velocity is -g k dt and its height y0 - g dt^2 k (k + 1) / 2, within the
  float spacing of its height (velocities come from (pred - pos) / dt),
  and the whole substep equals the oracle's float chain bit for bit in
  exact math;
* a dense cluster (200 particles within 0.02 h: rows far longer than one
  32-candidate chunk, coincident and near-coincident pairs) stays finite
  and equals the float oracle bit for bit in exact math; in fast math every
  lambda and delta stage on the GPU's own inputs is within the fast bars of
  FP64, including the iterations after the cluster has flown apart, when
  targets lie more than a cell from where they were binned and their own
  records fall outside their stencils (the fast lambda then adds its self
  term W(0) itself), for the single- and two-phase traversals;
* pairs a few float ulps apart near the box corner (0 < r^2 < 1e-12, below
  the reference's cutoffs of the gradient, normal, delta and tension terms)
  give the fast lambda, delta and finish stages within the fast bars of
  FP64 in both forms and for both traversals (this found the fast gradient
  and normal terms taking such pairs in);
* positions outside the box, on its faces, huge and NaN are predicted and
  binned as the oracle predicts and bins them (clamped cells): cell ids,
  cell starts and the stable permutation bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from particles4all_amd._lib import SceneView
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float64


class CustomScene:
    """Fluid particles at given positions in the box of a lattice scene,
    with its mass, h, d, denomRest and wall samples."""

    def __init__(self, base, pos, vel=None):
        self.base = base
        n = len(pos)
        self.n, self.n_fluid, self.n_body, self.n_bodies = n, n, 0, 0
        m = max(n, 1)
        self.pos = np.zeros((m, 4), np.float32)
        self.vel = np.zeros((m, 4), np.float32)
        self.rest = np.zeros((m, 4), np.float32)
        self.phase = np.zeros((m, 4), np.uint32)
        if n:
            self.pos[:n, :3] = pos
            if vel is not None:
                self.vel[:n, :3] = vel
        v = SceneView()
        C.memmove(C.byref(v), C.byref(base.view), C.sizeof(SceneView))
        fp = C.POINTER(C.c_float)
        v.n, v.n_fluid, v.n_body = n, n, 0
        v.pos = self.pos.ctypes.data_as(fp)
        v.vel = self.vel.ctypes.data_as(fp)
        v.rest = self.rest.ctypes.data_as(fp)
        v.phase = self.phase.ctypes.data_as(C.POINTER(C.c_uint32))
        v.n_bodies = 0
        self.view = v
        self.boundary_pts, self.boundary_psi = base.boundary_pts, base.boundary_psi
        self.body_centre, self.body_density = np.zeros((0, 3)), np.zeros(0)
        self.mass, self.h, self.d, self.denom_rest = base.mass, base.h, base.d, base.denom_rest


def rd(sim, name):
    return sim.read(name).astype(F)


def base_scene(math=1, pour=0):
    return workloads.dam_break_32(scale=0.5, math=math, bodies=False, pour_budget=pour)


def test_empty_scene_steps_and_takes_poured_particles():
    p, base, _ = base_scene(math=0, pour=64)
    sc = CustomScene(base, np.zeros((0, 3)))
    sim = Sim(0)
    sim.reset(p, sc)
    assert sim.info().n == 0 and sim.info().cap == 64
    sim.substeps(1 / 120, 3)
    assert sim.step(1 / 60) >= 1
    st = sim.stats()
    assert all(np.isfinite(st))
    pos = np.array([[10.0 + 0.5 * k, 20.0, 10.0] for k in range(8)], np.float32)
    assert sim.append_fluid(pos, np.zeros_like(pos)) == 8
    sim.substeps(1 / 120, 10)
    out = sim.read("POS")
    assert out.shape == (8, 4) and np.all(np.isfinite(out))
    assert np.all(out[:, 1] < 20.0)  # they fell
    sim.close()


@pytest.mark.parametrize("math", [0, 1])
def test_single_particle_falls_freely(math):
    p, base, _ = base_scene(math=math)
    y0 = 20.0
    sc = CustomScene(base, np.array([[16.3, y0, 15.7]]))
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 1 / 120
    P = O.constants_from(p, sc, dt)
    if math == 1:
        st = O.State(rd(sim, "POS"), rd(sim, "VEL"), np.zeros(1, np.uint32), rd(sim, "REST"))
        bnd = (rd(sim, "BPOS"), rd(sim, "BPOS")[:, 3].copy(), sim.read("BCELL_START"))
        O.substep(O.F32, P, st, bnd, np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0),
                  p.iterations, p.surface_tension_k > 0)
        sim.substeps(dt, 1)
        assert np.array_equal(rd(sim, "POS"), st.pos)
        assert np.array_equal(rd(sim, "VEL")[:, :3], st.vel[:, :3])
        k = 1
    else:
        k = 0
    steps = 12
    sim.substeps(dt, steps - k)
    g = p.gravity
    v = rd(sim, "VEL")[0]
    pos = rd(sim, "POS")[0]
    want_v = -g * dt * steps
    want_y = y0 - g * dt * dt * steps * (steps + 1) / 2
    # velFromPos forms (pred - pos) / dt in float: each substep carries up to
    # a few position ulps / dt of velocity error, and the height a few ulps
    ulp = float(np.spacing(np.float32(y0)))
    assert abs(v[1] - want_v) <= 4 * ulp / dt and abs(v[0]) < 1e-9 and abs(v[2]) < 1e-9
    assert abs(pos[1] - want_y) <= 4 * steps * ulp
    assert pos[0] == np.float32(16.3) and pos[2] == np.float32(15.7)
    assert rd(sim, "LAMBDA")[0] == 0.0
    sim.close()


def cluster(n=200, centre=(16.2, 3.1, 15.9), radius=0.04, seed=3):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.uniform(0, 1, size=(n, 1)) ** (1 / 3)
    pts = np.array(centre) + d * r
    pts[1] = pts[0]          # one exactly coincident pair
    pts[2, 0] += 2e-6        # and one a float ulp apart (r^2 just above 1e-12)
    pts[2, 1:] = pts[0, 1:]
    return pts


def near_coincident(n=48, centre=(1.3, 1.2, 1.4), seed=5):
    """A cluster near the box corner (coordinates near 1: float ulp 1.2e-7)
    with eight pairs one to three ulps apart, r^2 ~ 1e-14."""
    rng = np.random.default_rng(seed)
    pts = (np.array(centre) + rng.uniform(-0.5, 0.5, size=(n, 3))).astype(np.float32)
    for k in range(0, 16, 2):
        q = pts[k].copy()
        for a in range(1 + (k // 2) % 3):
            q[a] = np.nextafter(q[a], np.float32(10))
        pts[k + 1] = q
    return pts.astype(F)


@pytest.mark.parametrize("precision,variant", [(0, 2), (0, 6), (1, 2)])
def test_near_coincident_pairs(precision, variant, monkeypatch):
    monkeypatch.setenv("P4A_TILE1", str(variant))
    monkeypatch.setenv("P4A_TILE2", str(variant))
    dt = 1 / 120
    p, base, _ = workloads.block_1m(scale=0.125, math=0, precision=precision)
    assert p.surface_tension_k > 0
    sc = CustomScene(base, near_coincident())
    sim = Sim(0)
    sim.reset(p, sc)
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    P = O.constants_from(p, sc, dt)
    bnd = (rd(sim, "BPOS"), rd(sim, "BPOS")[:, 3].copy(), sim.read("BCELL_START"))
    cs = sim.read("CELL_START")
    pred = rd(sim, "PRED")
    if precision:
        pred = O.fp16_quantize(P, sim.read("PRED"), sim.read("CELL_OF_SLOT"))
    d = pred[:, None, :3] - pred[None, :, :3]
    r2 = (d ** 2).sum(-1)
    tiny = np.count_nonzero((r2 > 0) & (r2 < 1e-12)) // 2
    print(f"precision {precision}: {tiny} pairs with 0 < r^2 < 1e-12")
    assert tiny >= 4
    h = sc.h
    for it in range(p.iterations):
        pred = rd(sim, "PRED")
        at = O.fp16_quantize(P, sim.read("PRED"), sim.read("CELL_OF_SLOT")) if precision else pred
        # the FP16-offset form walks the stencil of a target's FP32 position;
        # the oracle takes the stencil of the positions it is given, so a
        # target whose quantized position lies across a cell face (within the
        # half rounding of its offset) is left out of the comparison
        same = O.cell_ids(O.F32, P, pred) == O.cell_ids(O.F64, P, at)
        assert np.count_nonzero(~same) <= 2
        rho64, lam64 = O.lambda_(O.F64, P, at, cs, *bnd)
        sim.run_stage("LAMBDA", dt)
        rho_err = np.abs(rd(sim, "DENSITY") - rho64)[same].max() / rho64.max()
        lam_err = np.abs(rd(sim, "LAMBDA") - lam64)[same].max() / max(np.abs(lam64).max(), 1e-30)
        print(f"iteration {it}: density {rho_err:.2e}, lambda {lam_err:.2e} "
              f"({np.count_nonzero(~same)} across a face)")
        assert rho_err <= (1e-3 if precision else 1e-5) and lam_err <= (1e-2 if precision else 1e-4)
        if precision:
            out64 = O.delta_at(O.F64, P, at, pred, rd(sim, "LAMBDA"), cs, *bnd)
        else:
            out64 = O.delta(O.F64, P, pred, rd(sim, "LAMBDA"), cs, *bnd)
        sim.run_stage("DELTA", dt)
        err = np.abs(rd(sim, "PRED")[:, :3] - out64[:, :3])[same].max()
        assert err <= (1e-3 if precision else 1e-4) * h, err
    pred, pos, den = rd(sim, "PRED"), rd(sim, "POS"), rd(sim, "DENSITY")
    _, corr, normal = O.xsph(O.F64, P, pred, pos, den, cs, True)
    corr = O.tension(O.F64, P, pred, den, normal, cs, corr)
    sim.run_stage("FINISH", dt)
    gn = rd(sim, "NORMAL")[:, :3]
    n_err = np.abs(gn - normal[:, :3]).max() / np.abs(normal[:, :3]).max()
    c_err = np.abs(rd(sim, "CORR")[:, :3] - corr[:, :3]).max() / max(np.abs(corr[:, :3]).max(), 1e-30)
    print(f"finish: normal {n_err:.2e}, corr {c_err:.2e}")
    assert n_err <= 1e-4 and c_err <= 1e-3
    sim.close()


@pytest.mark.parametrize("variant", [0, 2, 6])
def test_dense_cluster(variant, monkeypatch):
    monkeypatch.setenv("P4A_TILE1", str(variant))
    monkeypatch.setenv("P4A_TILE2", str(variant))
    dt = 1 / 120
    # exact math: the whole substep equals the oracle's float chain
    p, base, _ = base_scene(math=1)
    sc = CustomScene(base, cluster())
    sim = Sim(0)
    sim.reset(p, sc)
    P = O.constants_from(p, sc, dt)
    bnd = (rd(sim, "BPOS"), rd(sim, "BPOS")[:, 3].copy(), sim.read("BCELL_START"))
    start = (rd(sim, "POS"), rd(sim, "VEL"), rd(sim, "REST"))
    st32 = O.State(*start[:2], np.zeros(sc.n, np.uint32), start[2])
    info = O.substep(O.F32, P, st32, bnd, np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0),
                     p.iterations, p.surface_tension_k > 0)
    sim.substeps(dt, 1)
    gp, gv = rd(sim, "POS"), rd(sim, "VEL")
    assert np.all(np.isfinite(gp)) and np.all(np.isfinite(gv))
    assert np.array_equal(gp, st32.pos)
    assert np.array_equal(gv[:, :3], st32.vel[:, :3])
    assert np.array_equal(rd(sim, "DENSITY"), info["density"])
    sim.close()
    st64 = O.State(*start[:2], np.zeros(sc.n, np.uint32), start[2])
    O.substep(O.F64, P, st64, bnd, np.zeros((0, 3)), np.zeros((0, 9)), np.zeros(0),
              p.iterations, p.surface_tension_k > 0)
    float_gap = np.abs(st32.pos[:, :3] - st64.pos[:, :3]).max()
    # fast math, stage by stage on the GPU's own inputs: after the first
    # delta the cluster has flown apart by several cells, so targets sit far
    # from the cells they were binned into and their own records lie outside
    # their stencils; every lambda still counts the self term W(0) once
    p, base, _ = base_scene(math=0)
    sc = CustomScene(base, cluster())
    sim = Sim(0)
    sim.reset(p, sc)
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    P0 = O.constants_from(p, sc, dt)
    cs = sim.read("CELL_START")
    worst = 0.0
    for it in range(p.iterations):
        pred = rd(sim, "PRED")
        rho64, lam64 = O.lambda_(O.F64, P0, pred, cs, bnd[0], bnd[1], bnd[2])
        sim.run_stage("LAMBDA", dt)
        rho_err = np.abs(rd(sim, "DENSITY") - rho64).max() / rho64.max()
        lam_err = np.abs(rd(sim, "LAMBDA") - lam64).max() / max(np.abs(lam64).max(), 1e-30)
        print(f"cluster iteration {it}: lambda stage vs FP64: density {rho_err:.2e}, "
              f"lambda {lam_err:.2e}")
        assert rho_err <= 1e-5 and lam_err <= 1e-4
        out64 = O.delta(O.F64, P0, pred, rd(sim, "LAMBDA"), cs, bnd[0], bnd[1], bnd[2])
        sim.run_stage("DELTA", dt)
        worst = max(worst, np.abs(rd(sim, "PRED")[:, :3] - out64[:, :3]).max())
    print(f"cluster: delta stages vs FP64 max {worst:.3e}; float chain substep gap {float_gap:.3e}")
    assert worst <= 1e-4 * sc.h
    sim.close()
    # and the whole fast substep stays finite
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(dt, 1)
    assert np.all(np.isfinite(rd(sim, "POS"))) and np.all(np.isfinite(rd(sim, "VEL")))
    sim.close()


def test_outside_nan_and_face_positions_are_binned_as_the_oracle_bins_them():
    p, base, _ = base_scene(math=1)
    box = np.array(p.box)
    pts = np.array([[-5.0, 3.0, 4.0], [box[0] + 7.0, 3.0, 4.0], [4.0, -0.0, 4.0],
                    [box[0], box[1], box[2]], [0.0, 0.0, 0.0], [np.nan, 4.0, 4.0],
                    [4.0, np.nan, np.nan], [1e30, -1e30, 3.0], [5.5, 6.5, 7.5],
                    [box[0] - 1e-4, 2.0, 2.0]])
    sc = CustomScene(base, pts)
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 0.0
    P = O.constants_from(p, sc, dt)
    pos0, vel0, pid0 = rd(sim, "POS"), rd(sim, "VEL"), sim.read("PID")
    sim.run_stage("PREDICT", dt)
    pred, _ = O.predict(O.F32, P, pos0, vel0)
    sim.run_stage("GRID", dt)
    key = O.cell_ids(O.F32, P, pred)
    cs, slot, order = O.grid(P, key)
    assert np.array_equal(sim.read("KEY"), key)
    assert np.array_equal(sim.read("CELL_START")[: len(cs)], cs)
    assert np.array_equal(sim.read("ORDER"), order)
    assert np.array_equal(sim.read("PID"), pid0[order])
    nc = int(np.prod(P.gridDim[:]))
    assert key.max() < nc
    sim.close()


@pytest.mark.parametrize("math", [1, 0])
def test_tiny_bodies(math):
    """Bodies of one particle (zero covariance: the reference's |det| < 1e-12
    falls back to the identity), a 2x2x2 cube and a radius-1 lattice sphere:
    one substep bit-exact against the float chain in exact math (within the
    FP64 bars in fast math), then finite poses with orthonormal rotations."""
    from particles4all_amd._lib import SHAPE_LATTICE_CUBE, SHAPE_LATTICE_SPHERE, lattice_scene
    p, _, _ = workloads.dam_break_32(scale=0.5, math=math, bodies=False)
    bl = [(SHAPE_LATTICE_CUBE, 1, 0.5, (8.3, 20.2, 8.1)), (SHAPE_LATTICE_CUBE, 2, 1.5, (16.2, 22.4, 8.3)),
          (SHAPE_LATTICE_SPHERE, 1, 0.8, (24.1, 21.7, 16.6))]
    sc = lattice_scene(p, (16, 16, 16), jitter=0.05, seed=5, bodies=bl)
    sim = Sim(0)
    sim.reset(p, sc)
    counts = list(sc.body_count)
    assert sc.n_bodies == 3 and counts[0] == 1
    dt = 1 / 120
    P = O.constants_from(p, sc, dt)
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    pid0 = sim.read("PID")
    inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(F)
    bnd = (rd(sim, "BPOS"), rd(sim, "BPOS")[:, 3].copy(), sim.read("BCELL_START"))
    rot = np.tile(np.eye(3).reshape(1, 9), (sc.n_bodies, 1))
    info = O.substep(O.F32 if math == 1 else O.F64, P, st, bnd, np.array(sc.body_centre, F),
                     rot, inv, p.iterations, p.surface_tension_k > 0)
    sim.substeps(dt, 1)
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    if math == 1:
        assert np.array_equal(rd(sim, "POS"), st.pos)
        assert np.array_equal(rd(sim, "VEL")[:, :3], st.vel[:, :3])
    else:
        assert np.abs(rd(sim, "POS")[:, :3] - st.pos[:, :3]).max() <= 1e-4 * sc.h
    sim.substeps(dt, 20)
    R = sim.read("BODY_ROT").reshape(-1, 3, 4)[:, :, :3].astype(F)
    c = sim.read("BODY_CENTRE")
    assert np.all(np.isfinite(R)) and np.all(np.isfinite(c)) and np.all(np.isfinite(rd(sim, "POS")))
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
    assert np.allclose(R[0], np.eye(3))  # the one-particle body never turns
    sim.close()


@pytest.mark.parametrize("math", [1, 0])
def test_batch_with_empty_and_single_particle_scenes(math):
    """A batch whose scenes include an empty one and a one-particle one
    between ordinary scenes: every scene bit-identical to its own Sim."""
    from particles4all_amd import Batch
    items = workloads.batch_scenes(n_scenes=3, seed=7, n_min=300, n_max=2000, math=math)
    p0, base0 = items[0]
    empty = (p0, CustomScene(base0, np.zeros((0, 3))))
    p1, base1 = items[1]
    lone = (p1, CustomScene(base1, np.array([[5.3, 9.1, 6.2]])))
    items = [items[0], empty, items[1], lone, items[2]]
    dt = 1 / 120
    b = Batch(items)
    b.substeps(dt, 4)
    for s, (p, sc) in enumerate(items):
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 4)
        assert b.read(s, "POS").shape == sim.read("POS").shape
        for name in ("POS", "VEL", "PID", "DENSITY", "LAMBDA"):
            a, g = b.read(s, name), sim.read(name)
            assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), (s, name)
        sim.close()
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_density_never_below_the_self_term(precision):
    """Every density includes the particle's own W(0) and only non-negative
    terms, so rho_i >= m W(0) after every lambda, also for targets far from
    where they were binned (the exploding cluster), in the FP32 and the
    FP16-offset forms."""
    p, base, _ = workloads.dam_break_32(scale=0.5, math=0, bodies=False, precision=precision)
    sc = CustomScene(base, cluster())
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 1 / 120
    P = O.constants_from(p, sc, dt)
    self_rho = sc.mass * P.poly6 * P.h2 ** 3
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    for it in range(p.iterations):
        sim.run_stage("LAMBDA", dt)
        rho = rd(sim, "DENSITY")
        assert np.all(np.isfinite(rho))
        print(f"precision {precision} iteration {it}: min rho / (m W(0)) = {rho.min() / self_rho:.6f}")
        assert rho.min() >= self_rho * (1 - 1e-5)
        sim.run_stage("DELTA", dt)
    sim.run_stage("FINISH", dt)
    assert np.all(np.isfinite(rd(sim, "POS")))
    sim.close()


def test_empty_scene_fields_and_mesh():
    p, base, _ = base_scene(math=0)
    sc = CustomScene(base, np.zeros((0, 3)))
    sim = Sim(0)
    sim.reset(p, sc)
    sim.aniso()
    sim.field_configure([0, 0, 0], list(p.box), 24)
    for aniso in (False, True):
        sim.field_generate(anisotropic=aniso)
        assert np.all(sim.read("FIELD") == 0)
    sim.mesh_normal_field(2)
    sim.mesh_march(0.4)
    verts, total = sim.mesh()
    assert total == 0 and len(verts) == 0
    assert list(sim.read("MESH_INDIRECT")) == [0, 1, 0, 0]
    sim.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_crowded_cells_sort_stably_in_linear_time(precision):
    """Cells holding 65 to 20,000 particles (the radix-sorted segments)
    beside ordinary ones, with the pre-sort indices shuffled so every
    crowded segment spans most of the index range: cell ids, cell starts,
    the stable permutation and the moved records equal the oracle's
    counting sort, on two grid builds in a row; the whole build of 37k
    particles (one cell of 20,000) finishes in well under a second."""
    import time
    p, base, _ = workloads.dam_break_32(scale=0.5, math=1 - precision, bodies=False,
                                        precision=precision)
    rng = np.random.default_rng(11)
    box = np.array(p.box)
    h = base.h
    parts = [rng.uniform(h, box - h, size=(10000, 3))]
    for size, cell in ((20000, (3, 3, 3)), (65, (5, 2, 4)), (66, (1, 6, 2)), (300, (6, 6, 6)),
                       (4000, (2, 5, 1)), (2601, (4, 1, 5))):
        lo = np.array(cell, F) * h
        parts.append(lo + rng.uniform(0.05, 0.95, size=(size, 3)) * h)
    pts = np.concatenate(parts)
    pts = pts[rng.permutation(len(pts))]
    sc = CustomScene(base, pts)
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 0.0
    P = O.constants_from(p, sc, dt)
    for rep in range(2):
        pos0, vel0, pid0 = rd(sim, "POS"), rd(sim, "VEL"), sim.read("PID")
        sim.run_stage("PREDICT", dt)
        pred, _ = O.predict(O.F32, P, pos0, vel0)
        t0 = time.perf_counter()
        sim.run_stage("GRID", dt)
        sim.sync()
        took = time.perf_counter() - t0
        key = O.cell_ids(O.F32, P, pred)
        cs, slot, order = O.grid(P, key)
        counts = np.diff(cs.astype(np.int64))
        print(f"precision {precision} build {rep}: largest cell {counts.max()}, "
              f"{int((counts > 64).sum())} cells over 64, grid build {took * 1e3:.2f} ms")
        assert counts.max() >= 20000 and (counts > 64).sum() >= 6
        assert np.array_equal(sim.read("KEY"), key)
        assert np.array_equal(sim.read("CELL_START")[: len(cs)], cs)
        assert np.array_equal(sim.read("SLOT"), slot)
        assert np.array_equal(sim.read("ORDER"), order)
        assert np.array_equal(sim.read("PID"), pid0[order])
        assert np.array_equal(rd(sim, "PRED")[:, :3], pred[order][:, :3])
        assert np.array_equal(rd(sim, "POS"), pos0[order])
        if precision == 1:
            assert np.array_equal(sim.read("CELL_OF_SLOT"), key[order])
        assert took < 1.0
    sim.close()
