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

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("math", [1, 0])
def test_batch_equals_separate_sims(math):
    items = workloads.batch_scenes(n_scenes=7, seed=11, n_min=300, n_max=6000, math=math)
    assert sum(sc.n_bodies for _, sc in items) > 0
    dt = 1 / 120
    b = Batch(items)
    b.substeps(dt, 3)
    chosen, times = b.tune(dt)  # every shape gives the same bits
    assert len(times) == 14 and all(t > 0 for t in times.values())
    b.substeps(dt, 2)
    for s, (p, sc) in enumerate(items):
        sim = Sim(0)
        sim.reset(p, sc)
        sim.substeps(dt, 3)
        sim.substeps(dt, 2)
        for name in ("POS", "VEL", "PID", "DENSITY", "LAMBDA"):
            a, g = b.read(s, name), sim.read(name)
            assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), (s, name)
        np.testing.assert_allclose(b.stats(s), sim.stats(), rtol=1e-12, atol=0)
        sim.close()
    b.close()


def test_batch_rejects_mixed_iteration_counts():
    items = workloads.batch_scenes(n_scenes=2, seed=3, n_min=300, n_max=600)
    items[1][0].iterations = 3
    with pytest.raises(RuntimeError, match="iterations"):
        Batch(items)
