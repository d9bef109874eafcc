"""Error paths of the C ABI on the GPU.

Bars: every refused call returns its documented code (P4A_ERR_ARG, _RANGE,
_STATE) with a message and before any kernel runs; a refused call changes
nothing: a sim that took a battery of refused calls (bad ids, reads and
writes past the end, offsets near 2^64, unknown stages and forms, resets
and resizes whose grid would not fit, phases naming missing bodies, field
and mesh calls out of order, corrupt snapshots) continues bit for bit as
its twin that took none, in the FP32 and FP16-offset forms, stage by stage;
a NaN frame runs no substep and the time bank restarts at the next frame as
the reference's `(timeBank || 0) + frameDt` does (sim.js:420); an infinite
frame runs the work cap and empties the bank (sim.js:423).
"""
import ctypes as C

import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from particles4all_amd._lib import BUF, STAGE, lib

pytestmark = pytest.mark.gpu

NAMES = ("POS", "VEL", "PID", "DENSITY", "BODY_CENTRE", "BODY_ROT")


def state(sim):
    return {k: sim.read(k).copy() for k in NAMES}


def same(a, b):
    for k in NAMES:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def refused(code, fn):
    with pytest.raises(RuntimeError, match=code):
        fn()


def raw(sim, rc, code, what):
    """A direct C call's return code, with the message kept on the handle."""
    names = {1: "P4A_ERR_ARG", 3: "P4A_ERR_STATE", 5: "P4A_ERR_RANGE"}
    assert names.get(rc) == code, f"{what}: rc {rc}"
    assert lib().p4a_last_error(sim.h).decode() != "" or code == "P4A_ERR_ARG"


def scene(precision):
    return workloads.dam_break_32(scale=0.5, math=0, precision=precision, pour_budget=16)


def test_calls_before_reset_report_state():
    sim = Sim(0)
    calls = [lambda: sim.step(1 / 60), lambda: sim.substeps(1 / 120, 1),
             lambda: sim.run_stage("PREDICT", 1 / 120), lambda: sim.read("POS"),
             lambda: sim.stats(), lambda: sim.body_pose(), lambda: sim.aniso(),
             lambda: sim.field_configure([0, 0, 0], [1, 1, 1], 8), lambda: sim.snapshot(),
             lambda: sim.body_pack(), lambda: sim.release_body(),
             lambda: sim.hold_body(1, [0, 0, 0], 1, 1), lambda: sim.resize_box([8, 8, 8]),
             lambda: sim.device_ptr("POS"), lambda: sim.mesh_triangles()]
    for c in calls:
        refused("P4A_ERR_STATE", c)
    buf = np.zeros(4, np.float32)
    raw(sim, lib().p4a_write(sim.h, BUF["POS"], buf.ctypes.data, 0, 16), "P4A_ERR_STATE", "write")
    sim.close()


def battery(sim, p, sc, snap):
    """Refused calls on a live sim; each names its code."""
    L = lib()
    i = sim.info()
    n, nB = i.n, i.n_bodies
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    big = (1 << 64) - 8
    raw(sim, L.p4a_read(sim.h, 999, ptr, 0, 16), "P4A_ERR_ARG", "read unknown id")
    raw(sim, L.p4a_read(sim.h, BUF["POS"], ptr, n * 16, 16), "P4A_ERR_RANGE", "read past end")
    raw(sim, L.p4a_read(sim.h, BUF["POS"], ptr, big, 16), "P4A_ERR_RANGE", "read offset 2^64")
    raw(sim, L.p4a_read(sim.h, BUF["BODY"], ptr, n * 16 - 8, 16), "P4A_ERR_RANGE", "read body")
    raw(sim, L.p4a_read(sim.h, BUF["POS"], ptr, 0, big), "P4A_ERR_RANGE", "read 2^64 bytes")
    raw(sim, L.p4a_write(sim.h, BUF["POS"], ptr, n * 16 - 8, 16), "P4A_ERR_RANGE", "write past end")
    raw(sim, L.p4a_write(sim.h, BUF["POS"], ptr, big, 16), "P4A_ERR_RANGE", "write offset 2^64")
    raw(sim, L.p4a_write(sim.h, BUF["PID"], ptr, 0, 4), "P4A_ERR_ARG", "write ids")
    raw(sim, L.p4a_write(sim.h, BUF["SLOT"], ptr, 0, 4), "P4A_ERR_ARG", "write slots")
    raw(sim, L.p4a_write(sim.h, 999, ptr, 0, 4), "P4A_ERR_ARG", "write unknown id")
    # index buffers the walks and the sort index through: not writable
    for name in ("CELL_START", "BCELL_START", "CELL_OF_SLOT", "ORDER", "KEY"):
        raw(sim, L.p4a_write(sim.h, BUF[name], ptr, 0, 4), "P4A_ERR_ARG", f"write {name}")
    # positions and velocities: whole finite floats only
    bad = np.zeros(8, np.float32)
    for k, v in ((1, np.nan), (6, np.inf), (3, -np.inf)):
        bad[:] = 1.0
        bad[k] = v
        for name in ("POS", "VEL", "PRED"):
            raw(sim, L.p4a_write(sim.h, BUF[name], bad.ctypes.data, 0, 32), "P4A_ERR_ARG",
                f"write non-finite {name}")
    raw(sim, L.p4a_write(sim.h, BUF["POS"], ptr, 2, 16), "P4A_ERR_ARG", "write part of a float")
    # non-finite interaction arguments
    nan = float("nan")
    refused("P4A_ERR_ARG", lambda: sim.apply_ray_impulse([0, 0, 0], [1, 0, 0], [nan, 1, 1], 4.0, 10.0))
    refused("P4A_ERR_ARG", lambda: sim.apply_ray_impulse([0, nan, 0], [1, 0, 0], [1, 1, 1], 4.0, 10.0))
    refused("P4A_ERR_ARG", lambda: sim.apply_ray_impulse([0, 0, 0], [1, 0, 0], [1, 1, 1], nan, 10.0))
    refused("P4A_ERR_ARG", lambda: sim.apply_ray_impulse([0, 0, 0], [1, 0, 0], [1, 1, 1], 4.0, float("inf")))
    refused("P4A_ERR_ARG", lambda: sim.hold_body(1, [nan, 2.0, 3.0], 10.0, 4.0))
    refused("P4A_ERR_ARG", lambda: sim.hold_body(1, [1.0, 2.0, 3.0], float("inf"), 4.0))
    refused("P4A_ERR_ARG", lambda: sim.hold_body(1, [1.0, 2.0, 3.0], 10.0, nan))
    pts = np.full((4, 3), 5.0, np.float32)
    pts[2, 1] = np.nan
    refused("P4A_ERR_ARG", lambda: sim.append_fluid(pts, np.zeros_like(pts)))
    pts[2, 1] = 5.0
    vel = np.zeros_like(pts)
    vel[3, 0] = np.inf
    refused("P4A_ERR_ARG", lambda: sim.append_fluid(pts, vel))
    d, nbytes = C.c_void_p(), C.c_size_t()
    raw(sim, L.p4a_device_ptr(sim.h, 999, C.byref(d), C.byref(nbytes)), "P4A_ERR_ARG", "ptr")
    raw(sim, L.p4a_run_stage(sim.h, 99, 1 / 120), "P4A_ERR_ARG", "unknown stage")
    refused("P4A_ERR_ARG", lambda: sim.substeps(1 / 120, -1))
    # forms: unknown math and precision, FP16 with exact math
    for kw in (dict(math=7), dict(precision=9), dict(math=1, precision=1)):
        q = workloads.dam_break_32(scale=0.5, **{**dict(math=0, precision=0), **kw})[0]
        refused("P4A_ERR_ARG", lambda: sim.set_params(q))
        refused("P4A_ERR_ARG", lambda: sim.reset(q, sc))
    # resets that cannot be built are refused before the old scene goes
    q = workloads.dam_break_32(scale=0.5, math=p.math, precision=p.precision)[0]
    q.box[0] = 1e12
    refused("P4A_ERR_RANGE", lambda: sim.reset(q, sc))
    q.box[0] = float("nan")
    refused("P4A_ERR_RANGE", lambda: sim.reset(q, sc))
    old = int(sc.phase[0, 0])
    for bad in (nB + 1, 0x80000001):
        sc.phase[0, 0] = bad
        refused("P4A_ERR_ARG", lambda: sim.reset(p, sc))
    sc.phase[0, 0] = old
    q = workloads.dam_break_32(scale=0.5, math=p.math, precision=p.precision, pour_budget=1 << 28)[0]
    refused("P4A_ERR_RANGE", lambda: sim.reset(q, sc))
    # resizes
    refused("P4A_ERR_RANGE", lambda: sim.resize_box([1e12, p.box[1], p.box[2]]))
    refused("P4A_ERR_ARG", lambda: sim.resize_box([0.0, p.box[1], p.box[2]]))
    sim.resize_box(list(p.box))  # the same box: nothing to do
    # field and mesh out of order, bad field boxes and sizes
    refused("P4A_ERR_ARG", lambda: sim.field_generate())
    refused("P4A_ERR_ARG", lambda: sim.field_smooth(1))
    refused("P4A_ERR_ARG", lambda: sim.mesh_normal_field(1))
    refused("P4A_ERR_ARG", lambda: sim.mesh_march(0.5))
    refused("P4A_ERR_ARG", lambda: sim.mesh_triangles())
    refused("P4A_ERR_ARG", lambda: sim.field_configure([0, 0, 0], [8, 8, 8], 0))
    refused("P4A_ERR_ARG", lambda: sim.field_configure([0, 0, 0], [8, 0, 8], 8))
    refused("P4A_ERR_RANGE", lambda: sim.field_configure([0, 0, 0], [8, 8, 8], 8, 1 << 31))
    refused("P4A_ERR_RANGE", lambda: sim.field_configure([0, 0, 0], [8, 8, 8], 1 << 20))
    refused("P4A_ERR_RANGE", lambda: sim.field_configure([0, 0, 0], [float("inf"), 8, 8], 8))
    # a refused configure keeps the field that was configured before it
    sim.field_configure([0, 0, 0], list(p.box), 12)
    sim.field_generate()
    f0 = sim.read("FIELD").copy()
    refused("P4A_ERR_RANGE", lambda: sim.field_configure([0, 0, 0], [8, 8, 8], 1 << 20))
    refused("P4A_ERR_ARG", lambda: sim.field_configure([0, 0, 0], [8, 8, 8], -3))
    sim.field_generate()
    assert np.array_equal(sim.read("FIELD").view(np.uint32), f0.view(np.uint32))
    refused("P4A_ERR_ARG", lambda: sim.aniso(radius_scale=3.0))
    # bodies: no static data yet, a count that is not the number of bodies
    refused("P4A_ERR_ARG", lambda: sim.body_pack())
    stat = np.ones((nB + 1, 8), np.float32)
    raw(sim, L.p4a_body_pack(sim.h, stat.ctypes.data, nB + 1), "P4A_ERR_ARG", "body_pack count")
    # snapshots: foreign, truncated, rows out of order; all before any reset
    refused("P4A_ERR_RANGE", lambda: sim.load_snapshot(snap[:64]))
    refused("P4A_ERR_ARG", lambda: sim.load_snapshot(b"X" + snap[1:]))
    refused("P4A_ERR_RANGE", lambda: sim.load_snapshot(snap[:-4]))
    nc = i.n_cells
    hdr = len(snap) - (n * 60 + (nc + 2) * 4 + nB * 172 + i.n_boundary * 16)
    cs0 = hdr + n * 60
    cs = np.frombuffer(snap, np.uint32, nc + 2, cs0).copy()
    rising = np.nonzero(np.diff(cs[: nc + 1]) > 0)[0]
    k = int(rising[len(rising) // 2])
    bad = cs.copy()
    bad[k] = bad[k + 1] + 1
    refused("P4A_ERR_ARG", lambda: sim.load_snapshot(snap[:cs0] + bad.tobytes() + snap[cs0 + bad.nbytes:]))
    bad = cs.copy()
    bad[0] = 1
    refused("P4A_ERR_ARG", lambda: sim.load_snapshot(snap[:cs0] + bad.tobytes() + snap[cs0 + bad.nbytes:]))
    # accepted calls that do nothing here: a hold on a body that does not
    # exist, an impulse of zero radius
    sim.hold_body(nB + 7, [1.0, 2.0, 3.0], 10.0, 4.0)
    sim.apply_ray_impulse([0, 0, 0], [1, 0, 0], [1, 1, 1], 0.0, 10.0)


@pytest.mark.parametrize("precision", [0, 1])
def test_refused_calls_change_nothing(precision):
    p, sc, _ = scene(precision)
    a, b = Sim(0), Sim(0)
    for s in (a, b):
        s.reset(p, sc)
        s.substeps(1 / 120, 3)
        s.run_stage("PREDICT", 1 / 120)
        s.run_stage("GRID", 1 / 120)
    snap = b.snapshot()
    battery(b, p, sc, snap)
    # stage by stage: the FP16 records built by GRID are still valid on b
    for s in (a, b):
        s.run_stage("BODY_PRE", 1 / 120)
        for _ in range(p.iterations):
            s.run_stage("LAMBDA", 1 / 120)
            s.run_stage("DELTA", 1 / 120)
            s.run_stage("BODY_SOLVE", 1 / 120)
        s.run_stage("FINISH", 1 / 120)
        s.run_stage("STATS", 1 / 120)
    same(state(b), state(a))
    for s in (a, b):
        s.substeps(1 / 120, 2)
        s.step(1 / 60)
    same(state(b), state(a))
    assert np.array_equal(np.array(a.stats()), np.array(b.stats()))
    ia, ib = a.info(), b.info()
    assert (ia.n, ia.cap, ia.n_cells, ia.time_bank) == (ib.n, ib.cap, ib.n_cells, ib.time_bank)
    # the snapshot taken before the battery still loads on the handle
    b.load_snapshot(snap)
    a.close()
    b.close()


def test_nan_and_infinite_frames():
    p, sc, _ = scene(0)
    a, b = Sim(0), Sim(0)
    for s in (a, b):
        s.reset(p, sc)
    assert b.step(float("nan")) == 0
    assert np.isnan(b.info().time_bank)
    # the bank restarts at zero: the next frame runs as on a fresh bank
    assert a.step(1 / 60) == 2 and b.step(1 / 60) == 2
    same(state(b), state(a))
    assert a.info().time_bank == b.info().time_bank
    # an infinite frame runs the work cap, 8 frames' worth, and empties the bank
    assert b.step(float("inf")) == 16
    assert b.info().time_bank == 0.0
    a.substeps(1 / 120, 16)
    same(state(b), state(a))
    a.close()
    b.close()


def test_batch_refuses_bad_arguments():
    """p4a_batch_* refuse unknown scenes, unreadable buffers, reads past a
    scene's slice (also at offsets near 2^64) and ids read in part of a
    word; the batch still steps afterwards."""
    from particles4all_amd import Batch
    items = workloads.batch_scenes(n_scenes=3, seed=5, n_min=300, n_max=900)
    b = Batch(items)
    L = lib()
    n0 = b.scene_size(0)
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    assert L.p4a_batch_read(b.h, 3, BUF["POS"], ptr, 0, 16) == 1        # no scene 3
    assert L.p4a_batch_read(b.h, -1, BUF["POS"], ptr, 0, 16) == 1
    assert L.p4a_batch_read(b.h, 0, BUF["CORR"], ptr, 0, 16) == 1      # not per scene
    assert L.p4a_batch_read(b.h, 0, BUF["POS"], ptr, n0 * 16, 16) == 5
    assert L.p4a_batch_read(b.h, 0, BUF["POS"], ptr, (1 << 64) - 8, 16) == 5
    assert L.p4a_batch_read(b.h, 0, BUF["PID"], ptr, 2, 4) == 1        # part of a word
    assert L.p4a_batch_substeps(b.h, -1.0, 1) == 1
    assert L.p4a_batch_substeps(b.h, float("nan"), 1) == 1
    assert L.p4a_batch_substeps(b.h, 1 / 120, -1) == 1
    b.substeps(1 / 120, 2)
    for s in range(3):
        pid = b.read(s, "PID")
        assert np.array_equal(np.sort(pid), np.arange(len(pid), dtype=np.uint32))
    b.close()
