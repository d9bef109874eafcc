"""Batched scenes against the same scenes run one by one, through the C
ABI: after several substeps every scene's positions, velocities, particle
ids, densities and lambdas are bit-identical to a separate Sim of that
scene (same kernels, same per-scene arithmetic; only the slices of the
concatenated arrays differ), in both math modes, also after the batch's
autotuner has changed the traversal shapes between substeps; per-scene stats agree to
FP64 rounding (the reduction tree differs)."""
import numpy as np
import pytest

from particles4all_amd import Batch, Sim, workloads
from particles4all_amd._lib import lib

pytestmark = pytest.mark.gpu


def variant_ids(names):
    L = lib()
    ids = []
    for g, name in zip((1, 2), names):
        ids.append(next(v for v in range(10) if L.p4a_tile_variant_name(g, v).decode() == name))
    return ids


@pytest.mark.parametrize("math,tension", [(1, 0.0), (0, 0.0), (0, 0.5)])
def test_batch_equals_separate_sims(math, tension, monkeypatch):
    items = workloads.batch_scenes(n_scenes=7, seed=11, n_min=300, n_max=6000, math=math,
                                   tension=tension)
    assert sum(sc.n_bodies for _, sc in items) > 0
    dt = 1 / 120
    b = Batch(items)
    if math:
        # exact math: every shape gives the same bits, so the batch retunes
        # (and recaptures its graphs) between substeps and the separate Sims
        # keep their default shapes
        b.substeps(dt, 3)
        chosen, times = b.tune(dt)
        b.substeps(dt, 2)
    else:
        # fast math: the one- and two-phase walks differ in the last bits, so
        # the separate Sims run the shapes the batch's autotuner picked
        chosen, times = b.tune(dt)
        b.substeps(dt, 5)
        t1, t2 = variant_ids(chosen)
        monkeypatch.setenv("P4A_TILE1", str(t1))
        monkeypatch.setenv("P4A_TILE2", str(t2))
    assert len(times) == 16 and all(t > 0 for t in times.values())
    for s, (p, sc) in enumerate(items):
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 3)
        sim.substeps(dt, 2)
        for name in ("POS", "VEL", "PID", "DENSITY", "LAMBDA"):
            a, g = b.read(s, name), sim.read(name)
            assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), (s, name, chosen)
        np.testing.assert_allclose(b.stats(s), sim.stats(), rtol=1e-12, atol=0)
        sim.close()
    b.close()


def test_batch_rejects_mixed_iteration_counts():
    items = workloads.batch_scenes(n_scenes=2, seed=3, n_min=300, n_max=600)
    items[1][0].iterations = 3
    with pytest.raises(RuntimeError, match="iterations"):
        Batch(items)


@pytest.mark.parametrize("math", [0, 1])
def test_batch_staged_walk_equals_separate_sims(math, monkeypatch):
    """The staged shape (variant 9) in the batch, each wave inside one
    scene's slice, against separate Sims running it."""
    monkeypatch.setenv("P4A_TILE1", "9")
    monkeypatch.setenv("P4A_TILE2", "9")
    items = workloads.batch_scenes(n_scenes=7, seed=11, n_min=300, n_max=6000, math=math,
                                   tension=0.5)
    dt = 1 / 120
    b = Batch(items)
    b.substeps(dt, 5)
    for s, (p, sc) in enumerate(items):
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 5)
        for name in ("POS", "VEL", "PID", "DENSITY", "LAMBDA"):
            a, g = b.read(s, name), sim.read(name)
            assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), (s, name)
        sim.close()
    b.close()
