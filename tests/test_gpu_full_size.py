"""Parity at the workloads' full sizes, on the GPU, through the C ABI.

The small-scene parity tests pin every kernel; these run the same bars at
the sizes the throughput figures are quoted on, against the oracle on the
same inputs (the C oracle is OpenMP-parallel: each comparison takes
seconds):

* block_1m (1,048,576 jittered particles, a 128x64x128 block in a
  256x128x128 box, tension on), exact math: one whole substep (graph replay) bit-identical
  to the oracle's float chain (stable permutation, positions, velocities,
  lambdas, densities); the neighbour census bit-exact on the binned and on
  the moved (stale-grid) state;
* block_4m_bodies (4,254,487 particles, 64 bodies): one substep in exact
  math bit-identical to the oracle's float chain; in fast math against the
  FP64 oracle within twice the float chain's own distance from FP64 (or
  the small-scene bars where larger); the FP16-offset form's densities against
  the FP64 densities at the unquantized positions: mean error <= 0.1% of
  the rest density;
* 256 batched scenes (12.6M particles): every scene's ids a permutation,
  sampled scenes bit-identical to their own Sims after a batched substep;
* dam_break_2m (2,097,152 particles): the anisotropy counts, the isotropic
  and the anisotropic field on the 259^3 lattice of resolution 256, the
  blurred field, the normal field and the marched triangles bit-identical
  to the oracle.
"""
import sys
import time

import numpy as np
import pytest

from particles4all_amd import Batch, Sim, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float64


def rd(sim, name):
    return sim.read(name).astype(F)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _err(a, b):
    d = np.abs(a - b)
    return float(d.max()), float(np.sqrt((d ** 2).mean()))


def substep_inputs(sim, p, sc, dt):
    P = O.constants_from(p, sc, dt)
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    bpos = rd(sim, "BPOS")
    bnd = (bpos, bpos[:, 3].copy(), sim.read("BCELL_START"))
    inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(F)
    rot = np.tile(np.eye(3).reshape(1, 9), (sc.n_bodies, 1))
    centre = np.array(sc.body_centre, F).reshape(-1, 3)
    return P, st, bnd, inv, rot, centre


def test_block_1m_substep_and_census_bit_exact():
    p, sc, dt = workloads.block_1m(math=1)
    assert sc.n == 1_048_576 and p.surface_tension_k > 0
    sim = Sim(0)
    sim.reset(p, sc)
    P, st, bnd, inv, rot, centre = substep_inputs(sim, p, sc, dt)
    pid0 = sim.read("PID")
    sim.substeps(dt, 1)
    t = time.time()
    info = O.substep(O.F32, P, st, bnd, centre, rot, inv, p.iterations, True)
    say(f"block_1m: oracle substep {time.time() - t:.1f} s")
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    assert np.array_equal(rd(sim, "POS"), st.pos)
    assert np.array_equal(rd(sim, "VEL")[:, :3], st.vel[:, :3])
    assert np.array_equal(rd(sim, "LAMBDA"), info["lam"])
    assert np.array_equal(rd(sim, "DENSITY"), info["density"])
    # neighbour sets: binned state, then moved by one iteration (stale grid)
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    for it in range(2):
        cnt, s1, s2 = sim.census()
        a, b, c = O.census(P, rd(sim, "PRED"), sim.read("CELL_START"))
        say(f"block_1m census {it}: mean |N| {cnt.mean():.2f}, max {cnt.max()}")
        assert np.array_equal(cnt, a) and np.array_equal(s1, b) and np.array_equal(s2, c)
        sim.run_stage("LAMBDA", dt)
        sim.run_stage("DELTA", dt)
    sim.close()


def _bodies_substep(math):
    p, sc, dt = workloads.block_4m_bodies(math=math)
    assert sc.n_bodies == 64 and sc.n > 4_194_304
    sim = Sim(0)
    sim.reset(p, sc)
    P, st, bnd, inv, rot, centre = substep_inputs(sim, p, sc, dt)
    pid0 = sim.read("PID")
    sim.substeps(dt, 1)
    return sim, p, sc, dt, P, st, bnd, inv, rot, centre, pid0


def test_block_4m_bodies_substep_bit_exact_in_exact_math():
    sim, p, sc, dt, P, st, bnd, inv, rot, centre, pid0 = _bodies_substep(1)
    t = time.time()
    info = O.substep(O.F32, P, st, bnd, centre, rot, inv, p.iterations, True)
    say(f"block_4m_bodies: float oracle substep {time.time() - t:.1f} s")
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    assert np.array_equal(rd(sim, "POS"), st.pos)
    assert np.array_equal(rd(sim, "VEL")[:, :3], st.vel[:, :3])
    assert np.array_equal(rd(sim, "LAMBDA"), info["lam"])
    assert np.array_equal(rd(sim, "DENSITY"), info["density"])
    sim.close()


def test_block_4m_bodies_fast_substep_against_fp64():
    """Fast math against the FP64 oracle.  At this size a few particles
    (colliding bodies, the free surface under tension) amplify rounding over
    the four Jacobi-style iterations: the bar is the reference's own float
    arithmetic (the oracle's float chain) against FP64, doubled, or the
    small-scene bars (1e-4 h max, 1e-6 h RMS) if larger."""
    sim, p, sc, dt, P, st, bnd, inv, rot, centre, pid0 = _bodies_substep(0)
    h = sc.h
    st32 = O.State(st.pos, st.vel, st.body_of, st.rest)
    t = time.time()
    info = O.substep(O.F64, P, st, bnd, centre, rot, inv, p.iterations, True)
    info32 = O.substep(O.F32, P, st32, bnd, centre, rot, inv, p.iterations, True)
    say(f"block_4m_bodies: FP64 and float oracle substeps {time.time() - t:.1f} s")
    assert np.array_equal(sim.read("PID"), pid0[info["order"]])
    assert np.array_equal(info32["order"], info["order"])
    gp, gv, gl = rd(sim, "POS")[:, :3], rd(sim, "VEL")[:, :3], rd(sim, "LAMBDA")
    want = st.pos[:, :3]
    ep, ev, el = _err(gp, want), _err(gv, st.vel[:, :3]), _err(gl, info["lam"])
    cp, cv = _err(st32.pos[:, :3], want), _err(st32.vel[:, :3], st.vel[:, :3])
    cl = _err(info32["lam"], info["lam"])
    body = sc.phase[:, 0][pid0][info["order"]] > 0
    eb = _err(gp[body], want[body])
    frac = float(np.mean(np.abs(gp - want) <= 1e-5 * h))
    say(f"block_4m_bodies fast vs FP64: pos max {ep[0]:.3e} rms {ep[1]:.3e} (within 1e-5 h: "
        f"{frac:.6f}; body particles max {eb[0]:.3e}); vel max {ev[0]:.3e}; lambda max "
        f"{el[0]:.3e}; float chain vs FP64: pos max {cp[0]:.3e} rms {cp[1]:.3e}, vel max "
        f"{cv[0]:.3e}, lambda max {cl[0]:.3e}")
    assert ep[0] <= max(1e-4 * h, 2 * cp[0]) and ep[1] <= max(1e-6 * h, 2 * cp[1])
    assert ev[0] <= max(1e-4 * h / dt, 2 * cv[0])
    assert el[0] <= max(1e-4 * np.abs(info["lam"]).max(), 2 * cl[0])
    gc = rd(sim, "BODY_CENTRE")[:, :3]
    assert np.abs(gc - info["centre"]).max() <= 1e-4 * h
    sim.close()


def test_block_4m_bodies_fp16_density_error():
    p, sc, dt = workloads.block_4m_bodies(math=0, precision=1)
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(dt, 2)
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    if sc.n_bodies:
        sim.run_stage("BODY_PRE", dt)
    P = O.constants_from(p, sc, dt)
    bpos = rd(sim, "BPOS")
    pred, cs = rd(sim, "PRED"), sim.read("CELL_START")
    sim.run_stage("LAMBDA", dt)
    den, _ = O.lambda_(O.F64, P, pred, cs, bpos, bpos[:, 3].copy(), sim.read("BCELL_START"))
    err = np.abs(rd(sim, "DENSITY") - den) / p.rest_density
    say(f"block_4m_bodies FP16-offset density vs FP64: mean {err.mean():.3e} max "
        f"{err.max():.3e} (of rest density)")
    assert err.mean() <= 1e-3
    sim.close()


def test_batch_256_scenes_full_size():
    items = workloads.batch_scenes(n_scenes=256)
    n = [sc.n for _, sc in items]
    say(f"batch: {len(items)} scenes, {sum(n)} particles")
    dt = 1 / 120
    b = Batch(items)
    b.substeps(dt, 1)
    for s in range(len(items)):
        pid = b.read(s, "PID")
        assert np.array_equal(np.sort(pid), np.arange(n[s], dtype=np.uint32)), s
    rng = np.random.default_rng(7)
    picks = [int(x) for x in rng.choice(len(items), 6, replace=False)]
    sample = sorted({int(np.argmin(n)), int(np.argmax(n)), *picks})
    for s in sample:
        p, sc = items[s]
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 1)
        for name in ("POS", "VEL", "PID", "DENSITY", "LAMBDA"):
            assert np.array_equal(b.read(s, name).view(np.uint32),
                                  sim.read(name).view(np.uint32)), (s, name)
        sim.close()
    say(f"batch: scenes {sample} bit-identical to their own Sims")
    b.close()


def test_dam_break_2m_fields_and_mesh_bit_exact():
    p, sc, dt = workloads.dam_break_2m(math=0)
    assert sc.n == 128 ** 3
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(1 / 120, 4)
    P = O.constants_from(p, sc, 1 / 120)
    pos, cs = rd(sim, "POS"), sim.read("CELL_START")
    den, body = rd(sim, "DENSITY"), sim.read("BODY")[:, 0]
    sim.aniso()
    _, g32 = O.aniso(O.F32, P, pos, cs)
    gg = sim.read("ANISO_G").reshape(-1, 8)
    assert np.array_equal(gg[:, 7], g32.reshape(-1, 8)[:, 7].astype(np.float32))
    box = list(p.box)
    sim.field_configure([0, 0, 0], box, 256)
    lat = O.field_lattice([0, 0, 0], box, 256)
    vd = lat[0]
    assert list(sim.field_info().vert_dim) == vd
    sim.field_generate(anisotropic=False)
    want = O.field_iso(P, lat, sc.mass, p.rest_density, pos, den, body, 1, cs)
    assert np.array_equal(sim.read("FIELD").view(np.uint32), want.view(np.uint32))
    say(f"dam_break_2m: isotropic field over {want.size} vertices bit-exact")
    sim.field_generate(anisotropic=True)
    fa = sim.read("FIELD")
    want = O.field_aniso(lat, sc.mass, p.rest_density, rd(sim, "SMOOTH_POS"), rd(sim, "ANISO_G"),
                         den, body, 1)
    assert np.array_equal(fa.view(np.uint32), want.view(np.uint32))
    sim.field_smooth(1)
    f1 = sim.read("FIELD")
    assert np.array_equal(f1.view(np.uint32), O.mesh_smooth(vd, 1, fa).view(np.uint32))
    sim.mesh_normal_field(2)
    nf = sim.read("FIELD_NORMAL")
    assert np.array_equal(nf.view(np.uint32), O.mesh_smooth(vd, 2, f1).view(np.uint32))
    sim.mesh_march(0.4)
    verts, total = sim.mesh()
    want_v, want_total = O.mesh_march(vd, 0.4, sim.field_info().max_triangles, f1, nf)
    say(f"dam_break_2m: anisotropic field, blur, normal field and {total} triangles bit-exact")
    assert total == want_total > 100000
    assert np.array_equal(verts, want_v)
    sim.close()
