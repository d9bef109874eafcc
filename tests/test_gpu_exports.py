"""GPU renderer exports, through the C ABI.

* per-slot phase (the reference's body buffer that the SSFR splat binds):
  once its device pointer is requested, the device buffer equals the
  host-built P4A_BUF_BODY after every substep batch, a GRID stage and an
  append of fluid;
* the SSFR format: smoothed centres with w = 0, and the packed G read back
  as the splat reads it (ssfr_wgsl.js:21-28) symmetric positive definite,
  its determinant the third word of the second float4, the count the
  fourth;
* body poses (bodyPackWGSL): bit-identical to the oracle's packing of the
  live centre and rotation rows, before and after a step.
"""
import ctypes as C

import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_hip = None


def device_read(ptr, nbytes, dtype):
    """Copy device memory at ptr into a new array (hipMemcpy, device to host)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
    assert _hip.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), nbytes, 2) == 0
    return out


def dam_break():
    p, sc, _ = workloads.dam_break_32(scale=1.0, math=0, pour_budget=64)
    sim = Sim(0)
    sim.reset(p, sc)
    return sim, p, sc


def check_phase(sim):
    ptr, nbytes = sim.device_ptr("BODY")
    n = sim.info().n
    assert nbytes == n * 16
    dev = device_read(ptr, nbytes, np.uint32).reshape(n, 4)
    host = sim.read("BODY")
    assert np.array_equal(dev, host)
    return dev


def test_phase_export_follows_the_sort():
    sim, p, sc = dam_break()
    sim.substeps(1 / 120, 2)
    d0 = check_phase(sim)
    assert sim.info().n_bodies > 0 and np.count_nonzero(d0[:, 0]) == sim.info().n_body_parts
    sim.substeps(1 / 120, 60)
    d1 = check_phase(sim)
    assert not np.array_equal(d0, d1)  # the sort moved particles
    sim.run_stage("PREDICT", 1 / 120)
    sim.run_stage("GRID", 1 / 120)
    check_phase(sim)
    k = sim.append_fluid(np.array([[5.0, 30.0, 5.0], [6.0, 30.0, 5.0]], np.float32),
                         np.zeros((2, 3), np.float32))
    assert k == 2
    d2 = check_phase(sim)
    assert np.all(d2[-2:] == 0)
    sim.close()


def test_ssfr_format():
    sim, p, sc = dam_break()
    sim.substeps(1 / 120, 4)
    sim.aniso()
    n = sim.info().n
    sp = sim.read("SMOOTH_POS")
    assert np.all(sp[:, 3] == 0) and np.all(np.isfinite(sp))
    g = sim.read("ANISO_G")
    G = O.ssfr_kernel_g(g)
    assert np.array_equal(G, G.transpose(0, 2, 1))
    ev = np.linalg.eigvalsh(G)
    assert np.all(ev > 0)
    det = np.linalg.det(G)
    packed = g.reshape(n, 2, 4)[:, 1, 2].astype(np.float64)
    assert np.all(np.abs(packed - det) <= 1e-4 * np.abs(det))
    counts = g.reshape(n, 2, 4)[:, 1, 3]
    assert np.all(counts >= 1) and np.all(counts == np.round(counts))
    # the device pointers a renderer binds
    for name, per in (("SMOOTH_POS", 16), ("ANISO_G", 32)):
        ptr, nbytes = sim.device_ptr(name)
        assert ptr and nbytes == n * per
    sim.close()


def test_body_pack_matches_oracle():
    sim, p, sc = dam_break()
    nb = sim.info().n_bodies
    assert nb > 0
    rng = np.random.default_rng(5)
    stat = rng.uniform(0.1, 2.0, size=(nb, 8)).astype(np.float32)
    stat[:, 3] = np.arange(nb) % 3
    stat[:, 7] = 0
    sim.body_pack(stat)
    want = O.body_pack(sim.read("BODY_CENTRE"), sim.read("BODY_ROT"), stat)
    assert np.array_equal(sim.read("BODY_PACKED"), want)
    sim.substeps(1 / 120, 6)
    sim.body_pack()  # keeps the static data
    got = sim.read("BODY_PACKED")
    want = O.body_pack(sim.read("BODY_CENTRE"), sim.read("BODY_ROT"), stat)
    assert np.array_equal(got, want)
    # the packed columns of a rotation are orthonormal
    R = got.reshape(nb, 6, 4)[:, 2:5, :3].astype(np.float64)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
    sim.close()
