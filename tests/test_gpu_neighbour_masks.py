"""Per-substep neighbour masks against the full 27-cell walk, through the
C ABI.  The masks change which candidates a kernel visits, never a sum: a
simulation stepped with masks must equal the same simulation stepped with
full walks bit for bit, in both math forms, with bodies, with targets that
leave their cells (queued for the full walk) and with a skin so small that
most targets are queued.

* dam_break_32, a quarter-scale block_4m_bodies and a half-scale
  dam_break_2m (a block in the corner of an empty box: a wave spanning two
  grid rows of the block has unions too large to stage, and the rows above
  the block's top layer are empty and past the last occupied cell): 12 substeps
  with skins 0.1 and 0.3 against skin 0 (positions, velocities, densities,
  lambdas, particle ids, body poses), fast and exact math;
* skin 1e-4: nearly every target fails the mask test after the first delta
  (fast math queues them for the fallback launch; in exact math the waves
  walk them in full inside the list launch and only a few go to the
  fallback), and the result is still the full walk's; the block's fallback
  launches take queued targets at every skin;
* the mask state: the travel bound equals the largest distance of any
  particle from its build position after the deltas (read back stage by
  stage), and the queues hold a few percent of the targets at most on the
  flowing state;
* the moved queues: after each delta and shape matching, the queue of the
  next iteration holds exactly the targets that left the cell they were
  binned into (their full walks run beside the list launches, which skip
  them); deltas run back to back through the stage API refill the queue
  from empty and still give the full walk's bits;
* the FP16-offset form with masks equals its own full walk bit for bit;
* a build after a delta (targets outside the cells they were binned into)
  gives the full walk's bits;
* the masks are used from neighbour_min_particles on (by default only for
  scenes of 2^19 particles or more), and the threshold gives the same bits;
* skins outside [0, 1], non-finite skins and negative thresholds are
  refused.
"""
import numpy as np
import pytest

from oracle import oracle as O
from particles4all_amd import Sim, workloads

MOVED = 1 + 2 * 256  # NEIGHBOUR_STATE: first count of the moved queues

pytestmark = pytest.mark.gpu

FIELDS = ("POS", "VEL", "DENSITY", "LAMBDA", "PID")


def run(maker, skin, math, steps, precision=0, **kw):
    p, sc, _ = maker(math=math, **kw)
    p.precision = precision
    p.neighbour_skin = skin
    p.neighbour_min_particles = 0  # these scenes are below the default size
    sim = Sim(0)
    sim.reset(p, sc)
    dt = 1 / 120
    sim.substeps(dt, steps)
    out = {k: sim.read(k).copy() for k in FIELDS}
    if sc.n_bodies:
        out["BODY_CENTRE"] = sim.read("BODY_CENTRE").copy()
        out["BODY_ROT"] = sim.read("BODY_ROT").copy()
    st = sim.read("NEIGHBOUR_STATE")
    sim.close()
    return out, st


def same(a, b, what):
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (what, k)


@pytest.mark.parametrize("math", [0, 1])
@pytest.mark.parametrize("maker,kw", [(workloads.dam_break_32, {}),
                                      (workloads.block_4m_bodies, {"scale": 0.25}),
                                      (workloads.dam_break_2m, {"scale": 0.5})])
def test_masks_equal_the_full_walk(math, maker, kw):
    full, st0 = run(maker, 0.0, math, 12, **kw)
    assert st0[0] == 0 and not st0[1:].any()  # no list launch ran
    for skin in (0.1, 0.3, 1e-4):
        got, st = run(maker, skin, math, 12, **kw)
        queued = int(st[1:].sum())
        travel = float(st[:1].view(np.float32)[0])
        print(f"math {math} {maker.__name__} skin {skin}: last queues {queued} of "
              f"{len(got['PID'])}, travel {travel:.3e}")
        same(full, got, f"skin {skin}")
        if maker is workloads.block_4m_bodies:
            assert queued > 0  # the fallback launch walked some targets


@pytest.mark.parametrize("maker,kw", [(workloads.dam_break_32, {}),
                                      (workloads.block_4m_bodies, {"scale": 0.25})])
def test_fp16_offset_masks_equal_its_full_walk(maker, kw):
    """The FP16-offset form marks and walks quantized positions (its travel
    limit keeps the half rounding of both particles at build and at use)."""
    full, _ = run(maker, 0.0, 0, 12, precision=1, **kw)
    for skin in (0.1, 0.3, 1e-4):
        got, st = run(maker, skin, 0, 12, precision=1, **kw)
        print(f"fp16 {maker.__name__} skin {skin}: moved queue "
              f"{int(st[MOVED:].sum())}, list queue {int(st[1:MOVED].sum())}")
        same(full, got, f"fp16 skin {skin}")


def test_travel_bound_and_queues_on_a_flowing_state():
    p, sc, dt = workloads.block_4m_bodies(scale=0.25, math=0)
    p.neighbour_skin = 0.0
    p.neighbour_min_particles = 0
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(dt, 120)
    p.neighbour_skin = 0.1
    sim.set_params(p)
    n = sim.info().n
    sim.run_stage("PREDICT", dt)
    sim.run_stage("GRID", dt)
    sim.run_stage("BODY_PRE", dt)
    sim.run_stage("LAMBDA", dt)
    st = sim.read("NEIGHBOUR_STATE")
    assert not st.any()  # the build restarts the bound and the queues
    p0 = sim.read("PRED")[:, :3].astype(np.float64)
    most = 0.0
    for it in range(p.iterations):
        if it:
            sim.run_stage("LAMBDA", dt)
        sim.run_stage("DELTA", dt)
        q = sim.read("PRED")[:, :3].astype(np.float64)
        most = max(most, float(np.linalg.norm(q - p0, axis=1).max()))
        sim.run_stage("BODY_SOLVE", dt)  # moves the body particles again
        q = sim.read("PRED")[:, :3].astype(np.float64)
        most = max(most, float(np.linalg.norm(q - p0, axis=1).max()))
        st = sim.read("NEIGHBOUR_STATE")
        bound = float(st[:1].view(np.float32)[0])
        queued = int(st[1:].sum())
        print(f"iteration {it}: travel bound {bound:.4e}, largest travel {most:.4e}, "
              f"queued {queued} of {n}")
        # the float distance computed on the GPU, against FP64 of the same
        # float positions
        assert abs(bound - most) <= 1e-6 * max(most, 1.0)
        assert queued <= 0.05 * n
    sim.run_stage("FINISH", dt)
    sim.close()


def test_bad_skins_are_refused():
    p, sc, _ = workloads.dam_break_32(scale=0.25)
    sim = Sim(0)
    for bad in (-0.1, 1.5, float("nan"), float("inf")):
        p.neighbour_skin = bad
        with pytest.raises(RuntimeError, match="P4A_ERR_ARG"):
            sim.reset(p, sc)
    p.neighbour_skin = 0.1
    sim.reset(p, sc)
    p.neighbour_skin = -1.0
    with pytest.raises(RuntimeError, match="P4A_ERR_ARG"):
        sim.set_params(p)
    p.neighbour_skin = 0.1
    p.neighbour_min_particles = -1
    with pytest.raises(RuntimeError, match="P4A_ERR_ARG"):
        sim.set_params(p)
    sim.substeps(1 / 120, 2)
    sim.close()


def test_moved_queues_hold_the_targets_that_left_their_build_cells():
    p, sc, dt = workloads.block_4m_bodies(scale=0.25, math=0)
    P = O.constants_from(p, sc, dt)
    runs = {}
    p.neighbour_min_particles = 0
    for skin in (0.1, 0.0):
        p.neighbour_skin = skin
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 60)
        sim.run_stage("PREDICT", dt)
        sim.run_stage("GRID", dt)
        sim.run_stage("BODY_PRE", dt)
        sim.run_stage("LAMBDA", dt)
        built = O.cell_ids(O.F32, P, sim.read("PRED"))

        def moved_count(slot):
            st = sim.read("NEIGHBOUR_STATE")
            return int(st[MOVED + 256 * slot: MOVED + 256 * (slot + 1)].sum())

        it = 0
        stages = []
        for k in range(p.iterations):
            stages += (["LAMBDA"] if k else []) + ["DELTA", "BODY_SOLVE"]
        stages += ["DELTA", "DELTA", "BODY_SOLVE"]  # deltas back to back
        for st in stages:
            sim.run_stage(st, dt)
            if st == "LAMBDA":
                continue
            q = O.cell_ids(O.F32, P, sim.read("PRED")) != built
            if st == "DELTA":
                it += 1
                out = q
            else:
                out = out | q  # shape matching queues the ones it moves out
            if skin > 0:
                got = moved_count(it & 1)
                print(f"{st} (iteration {it}): queued {got}, left their cells {int(q.sum())}")
                assert got == int(out.sum())
        sim.run_stage("FINISH", dt)
        runs[skin] = {k: sim.read(k).copy() for k in FIELDS}
        sim.close()
    same(runs[0.0], runs[0.1], "back-to-back deltas")


@pytest.mark.parametrize("precision,math", [(0, 0), (0, 1), (1, 0)])
def test_build_after_a_delta_matches_the_full_walk(precision, math):
    """A mask build that finds targets outside the cell they were binned
    into (a delta ran between the grid build and the first lambda, through
    the stage API): those targets take the full walk in the build and are
    left to the fallback walks after it, and every stage still gives the
    full walk's bits (fast and exact math, FP16-offset form)."""
    p, sc, dt = workloads.block_4m_bodies(scale=0.25, math=math)
    p.precision = precision
    p.neighbour_min_particles = 0
    runs = {}
    for skin in (0.1, 0.0):
        p.neighbour_skin = skin
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 30)
        for st in ("PREDICT", "GRID", "BODY_PRE", "DELTA", "LAMBDA", "DELTA", "BODY_SOLVE",
                   "LAMBDA", "DELTA", "BODY_SOLVE", "FINISH"):
            sim.run_stage(st, dt)
        runs[skin] = {k: sim.read(k).copy() for k in FIELDS}
        if skin > 0:
            st = sim.read("NEIGHBOUR_STATE")
            print(f"queued {int(st[1:].sum())} of {len(runs[skin]['PID'])}")
        sim.close()
    same(runs[0.0], runs[0.1], "build after a delta")


def test_masks_only_from_the_particle_threshold():
    p, sc, dt = workloads.dam_break_32(math=0)
    n = sc.n
    default = p.neighbour_min_particles
    assert default == 1 << 19
    out = {}
    for lo in (n + 1, n, None):
        p.neighbour_min_particles = default if lo is None else lo
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 4)
        st = sim.read("NEIGHBOUR_STATE")
        used = bool(st.any())
        out[lo] = {k: sim.read(k).copy() for k in FIELDS}
        sim.close()
        # the default threshold (2^19) is above this scene
        assert used == (lo == n), (lo, used)
    same(out[n + 1], out[n], "threshold")
    same(out[n + 1], out[None], "default threshold")
