"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/p4a.h declares; the native scene builder equals a
pure-Python restatement of the reference's buildScene on small presets and
matches closed forms on the lattice workloads; the Node.js host layer
(settings reader, presets, pour hose) behaves as the reference's."""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from particles4all_amd import _lib, workloads
from oracle import scene_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    header = open(os.path.join(ROOT, "include", "p4a.h")).read()
    declared = set(re.findall(r"P4A_API\s+[\w\s\*]+?\b(p4a_\w+)\s*\(", header))
    assert len(declared) >= 25
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    assert _lib.lib().p4a_version().startswith(b"p4a")


def test_create_without_gpu_reports_no_device():
    import ctypes
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.p4a_create(99, ctypes.byref(h))
    assert rc == 6  # P4A_ERR_NODEV
    assert not h.value


@pytest.mark.parametrize("target,bodies,box,size", [
    (300, [("sphere", 0.5, 0.75), ("torus", 0.8, 0.75), ("box", 1.4, 0.75)], (0.3, 0.24, 0.2), 0.04),
    (500, [("box", 0.5, 0.6)], (0.4, 0.3, 0.2), 0.0),
    (150, [], (0.2, 0.2, 0.2), 0.0),
])
def test_preset_scene_matches_python_restatement(target, bodies, box, size):
    p = _lib.default_params(spacing=0.02, box=list(box))
    sc = _lib.preset_scene(p, target, bodies, size)
    ref = scene_ref.build_scene(list(box), 0.02, 1000.0, target, bodies, size)
    assert sc.n == ref["n"] and sc.n_fluid == ref["n_fluid"]
    assert np.array_equal(sc.pos, ref["pos"])
    assert np.array_equal(sc.rest, ref["rest"])
    assert np.array_equal(sc.phase, ref["phase"])
    assert np.array_equal(sc.boundary_pts, ref["boundary_pts"])
    np.testing.assert_allclose(sc.boundary_psi, ref["boundary_psi"], rtol=1e-6)
    assert sc.mass == pytest.approx(ref["mass"], rel=1e-12)
    assert sc.denom_rest == pytest.approx(ref["denom_rest"], rel=1e-9)
    assert sc.n_bodies == len(ref["bodies"])
    for b, rb in zip(range(sc.n_bodies), ref["bodies"]):
        np.testing.assert_allclose(sc.body_centre[b], rb["centre"], rtol=0, atol=0)
        assert sc.body_count[b] == rb["count"]


def test_lcg_matches_javascript_double_arithmetic():
    # first values of the fluidBlock LCG (seed 12345), computed with exact
    # integers where JavaScript's double product is exact (< 2^53)
    s = 12345
    v = scene_ref.js_lcg(s)
    assert v == (12345 * 1103515245 + 12345) % (1 << 32) & 0x7FFFFFFF
    # past 2^53 the double product rounds; the restatement must follow it
    big = 0x7FFFFFF0
    exact = (big * 1103515245 + 12345) % (1 << 32) & 0x7FFFFFFF
    rounded = scene_ref.js_lcg(big)
    assert rounded == int(float(big) * 1103515245.0 + 12345.0) % (1 << 32) & 0x7FFFFFFF
    assert isinstance(exact, int)


def test_lattice_workloads_closed_forms():
    p, sc, dt = workloads.dam_break_32()
    assert sc.n_fluid == 32 ** 3 and sc.n_body == 2 * 8 ** 3 and sc.n == 33792
    assert len(sc.boundary_psi) == 65 ** 3 - 63 ** 3
    # mass from the unjittered lattice: rho0 = m * sum_o W(|o|)
    h = 2.0
    coef = 315 / (64 * math.pi * h ** 9)
    s = sum(coef * (h * h - (a * a + b * b + c * c)) ** 3
            for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)
            if a * a + b * b + c * c < 4)
    assert sc.mass == pytest.approx(1000.0 / s, rel=1e-14)
    # jitter within +-0.05d of the lattice sites
    site = np.floor(sc.pos[: sc.n_fluid, :3]) + 0.5
    assert np.abs(sc.pos[: sc.n_fluid, :3] - site).max() <= 0.05 + 1e-6
    # body rest offsets are centred
    for b in range(2):
        sel = sc.phase[:, 0] == b + 1
        assert sel.sum() == 512
        assert np.abs(sc.rest[sel, :3].mean(axis=0)).max() < 1e-6
    p, sc, dt = workloads.block_1m(scale=0.125)
    assert sc.n == 16 * 8 * 16 and len(sc.boundary_psi) == 33 * 17 * 17 - 31 * 15 * 15
    p, sc, dt = workloads.block_1m()
    assert sc.n == 1_048_576 and list(p.box) == [256.0, 128.0, 128.0]


def _node(script):
    r = subprocess.run(["node", "-e", script], cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_js_presets_and_settings_reader():
    out = _node("""
const S = require('./particles4all_amd/node/settingsini.js');
const sc = require('./particles4all_amd/node/scene.js');
const res = {};
for (const name of ['small', 'medium', 'large', 'extreme']) res[name] = S.loadPreset(name).params;
res.kv = S.parseIni('# c\\n; d\\nA = 1\\nb=two = 2\\n =x\\nnoeq\\n');
res.q = S.iniToQuery('substeps=3\\ndisplay=4\\nbody=\\nfoo=1\\ncubemap="a/b/c.hdr"\\n');
res.def = sc.defaultParams();
res.cap = sc.blockCapacity({box: [1.5, 1, 1]}, 0.02);
res.over = S.loadPreset('small', 'particles=1234&box=1,2,3&boundary=0').params;
console.log(JSON.stringify(res));
""")
    r = json.loads(out)
    assert r["small"]["substeps"] == 2 and r["small"]["targetParticleCount"] == 30000
    assert r["small"]["box"] == [1.5, 1, 1] and r["small"]["xsphC"] == 0.066
    assert r["small"]["bodies"] == ["sphere:0.5", "torus:0.5", "box:0.5"]
    assert r["small"]["sorAverage"] is False  # soraverage is unsupported
    assert r["medium"]["targetParticleCount"] == 100000 and r["medium"]["box"] == [2, 1.5, 1.5]
    assert r["large"]["surfaceTensionK"] == 0.267 and r["large"]["substeps"] == 4
    assert r["extreme"]["targetParticleCount"] == 1000000 and r["extreme"]["box"] == [3, 3, 3]
    assert r["kv"] == {"a": "1", "b": "two = 2"}
    assert r["q"]["applied"] == ["substeps", "display", "body", "cubemap"]
    assert r["q"]["skipped"] == ["foo"] and r["q"]["missingPanorama"] == "c.hdr"
    assert "bodies=0" in r["q"]["query"] and "view=ssfr" in r["q"]["query"]
    assert r["def"]["xsphC"] == 0.05 and r["def"]["box"] == [1.5, 1.5, 1.0]
    assert r["cap"] == 26 * 48 * 48
    assert r["over"]["targetParticleCount"] == 1234 and r["over"]["box"] == [1, 2, 3]
    assert r["over"]["noBoundary"] is True


def test_js_pour_hose_math():
    out = _node("""
const D = require('./particles4all_amd/node/driver.js');
const added = [];
const sim = { params: { spacing: 0.02, box: [1.5, 1, 1], gravity: 9.81, substeps: 2 }, h: 0.04,
              pourBudget: 1000,
              appendFluid(p, v) { added.push(p.length / 3); return p.length / 3; } };
const ui = D.makeUi(sim, { pouring: true, pourSpeed: 4, pourWidth: 0.13, pourTilt: 25 });
let n = 0;
for (let f = 0; f < 10; f++) n += D.pourStep(sim, ui, 1 / 120);
console.log(JSON.stringify({ n, left: ui.pourLeft, across: D.pourLatticeAcross(0.13, 0.02),
  rate: D.pourRatePerSecond(ui, 0.02), layers: ui.pourNextLayer, added }));
""")
    r = json.loads(out)
    k = math.floor(0.13 / 0.02 + 0.5)  # Math.round
    assert r["across"] == k and r["rate"] == math.floor(k * k * 4 / (0.02 * 1.25))
    assert r["n"] + r["left"] == 1000 and r["n"] == sum(r["added"])
    # extruded 10 * 4/120 m = 0.333 m over layers of 1.25 d = 0.025 m
    assert r["layers"] == math.floor((10 * 4 / 120) / 0.025) + 1
