"""p4a_stats_noalloc, the stats reduction that allocates nothing (its
partial sums live in the library's fixed device array), against p4a_stats
on the same state, through the C ABI: two calls in a row, the second on a
smaller scene, each bit-identical to p4a_stats of its own handle and equal
to an FP64 reduction of the read-back densities and velocities; a null
output and a call before any reset are refused."""
import ctypes as C

import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from particles4all_amd._lib import lib

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


def noalloc_stats(sim, dev):
    assert lib().p4a_stats_noalloc(sim.h, dev) == 0, lib().p4a_last_error(sim.h)
    sim.sync()
    out = np.zeros(4, np.float64)
    assert hip().hipMemcpy(out.ctypes.data, dev, 32, 2) == 0
    return out


def fp64_stats(sim, p, sc):
    """Per-particle terms in float as the kernel forms them, sums in FP64."""
    f = np.float32
    rho = sim.read("DENSITY") * f(1.0 / p.rest_density)
    v = sim.read("VEL")
    v2 = v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]
    ke = (f(0.5) * f(sc.mass)) * v2
    return np.array([rho.astype(np.float64).mean(), float(rho.max()), float(np.sqrt(v2).max()),
                     ke.astype(np.float64).sum()])


def test_stats_noalloc_equals_stats_on_two_scenes():
    dev = C.c_void_p()
    assert hip().hipMalloc(C.byref(dev), 64) == 0
    try:
        for scale in (1.0, 0.5):  # the second call on a smaller input
            p, sc, frame = workloads.dam_break_32(scale=scale, math=0)
            sim = Sim(0)
            sim.reset(p, sc)
            sim.step(frame)
            sim.step(frame)
            want = np.array(sim.stats())
            assert hip().hipMemset(dev, 0xFF, 64) == 0  # NaNs: every word is written
            got = noalloc_stats(sim, dev)
            print(f"scale {scale}: n {sim.info().n} p4a_stats {want} noalloc {got}")
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
            np.testing.assert_allclose(got, fp64_stats(sim, p, sc), rtol=1e-12)
            # again on the same state: the same bits
            assert np.array_equal(noalloc_stats(sim, dev), got)
            sim.close()
    finally:
        hip().hipFree(dev)


def test_stats_noalloc_refusals():
    p, sc, frame = workloads.dam_break_32(scale=0.25)
    sim = Sim(0)
    assert lib().p4a_stats_noalloc(sim.h, None) != 0
    sim.reset(p, sc)
    assert lib().p4a_stats_noalloc(sim.h, None) == 1  # P4A_ERR_ARG
    sim.close()
