"""A long run on the GPU against the oracle, through the C ABI.

The 120-frame CPU case, whole: the 32^3 jittered dam break with its two
floating cubes (33,792 particles), 120 frames of 1/60 s (240 substeps through
the time bank and the graph replay) in exact math.  Every 20 frames the
positions, velocities and particle ids equal the oracle's float chain bit for
bit and the body centres agree to FP64 rounding (both sides sum in FP64, in
different orders; the reductions here are deterministic trees); every frame's
stats equal the oracle's reduction (maxima exactly, sums to FP64 rounding).
Nothing is re-synchronised: the GPU and the oracle each carry their own state
through all 240 substeps.  The run is made twice: with the library's default
(full 27-cell walks at this size) and with the per-substep neighbour masks
forced on.
"""
import time

import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float64


def rd(sim, name):
    return sim.read(name).astype(F)


@pytest.mark.parametrize("masks", [False, True])
def test_120_frames_bit_exact(masks):
    p, sc, frame = workloads.dam_break_32(scale=1.0, math=1)
    if masks:
        p.neighbour_min_particles = 0
    assert sc.n == 32 ** 3 + 2 * 8 ** 3 and sc.n_bodies == 2 and p.substeps == 2
    sim = Sim(0)
    sim.reset(p, sc)
    dt = frame / p.substeps
    P = O.constants_from(p, sc, dt)
    bpos = rd(sim, "BPOS")
    bnd = (bpos, bpos[:, 3].copy(), sim.read("BCELL_START"))
    st = O.State(rd(sim, "POS"), rd(sim, "VEL"), sim.read("BODY")[:, 0], rd(sim, "REST"))
    inv = (1.0 / np.maximum(1e-6, sc.body_density)).astype(np.float32).astype(F)
    centre = np.array(sc.body_centre, F).reshape(-1, 3)
    rot = np.tile(np.eye(3).reshape(1, 9), (sc.n_bodies, 1))
    pid = sim.read("PID")
    t0 = time.time()
    for f in range(1, 121):
        assert sim.step(frame) == 2
        info = None
        for _ in range(2):
            info = O.substep(O.F32, P, st, bnd, centre, rot, inv, p.iterations,
                             p.surface_tension_k > 0)
            centre, rot = info["centre"], info["rot"]
            pid = pid[info["order"]]
        got = sim.stats()
        want = O.stats(info["density"], st.vel, 1.0 / p.rest_density, sc.mass)
        assert got[1] == want[1] and got[2] == want[2], f
        np.testing.assert_allclose([got[0], got[3]], [want[0], want[3]], rtol=1e-12, atol=0)
        if f % 20 == 0:
            assert np.array_equal(sim.read("PID"), pid), f
            assert np.array_equal(rd(sim, "POS"), st.pos), f
            assert np.array_equal(rd(sim, "VEL")[:, :3], st.vel[:, :3]), f
            np.testing.assert_allclose(rd(sim, "BODY_CENTRE")[:, :3], centre, rtol=1e-6)
            print(f"frame {f}: bit-identical; maxRho {got[1]:.4f} maxSpeed {got[2]:.2f} "
                  f"({time.time() - t0:.1f} s)", flush=True)
    sim.close()
