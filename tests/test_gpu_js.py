"""The JavaScript Sim surface on the GPU (Node.js addon over the C ABI):
the reference's verify harness on the small preset, then pour, box-wall
move, ray impulse and body drag through the JS calls."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_js_sim_verify_harness_and_interactions():
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "verify_sim.js")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(out)
    assert out["nBodies"] == 3 and out["bodyShapes"] == ["sphere", "torus", "box"]
    assert out["cap"] > out["n0"] and out["sub"] == 2
    assert out["tuneKeepsState"] and out["subAfterTune"] == 2
    assert set(out["tuned"]) == {"lambda", "xsph"} and all(out["tuned"].values())
    assert out["scanMonotonic"] and out["scanTotal"] == out["n0"]
    assert out["worstDensityRelError"] < 1e-4
    assert out["nonFinite"] == 0 and out["lambdaMax"] <= 0
    assert out["outsideBox"] == 0
    assert out["poured"] > 0 and out["nAfterPour"] == out["n0"] + out["poured"]
    assert abs(out["boxX"] - 1.5) > 0.05 and out["gen"] >= 2
    assert out["impulsed"] > 0
    assert 0.9 < out["stats"]["avgRho"] < 1.1
    assert out["masksSameBits"] and out["neighbourStateLen"] == 1025


def test_js_surface_mesh_field():
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "verify_field.js")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(out)
    assert out["nVerts"] == out["vertDim"][0] * out["vertDim"][1] * out["vertDim"][2]
    assert out["checked"] >= 30 and out["isoWorstRelError"] < 1e-4
    assert out["nonFinite"] == 0 and out["anisoMax"] > 0
    # both fields integrate each particle's volume m / max(rho, rho0/2)
    assert abs(out["anisoIntegral"] / out["isoIntegral"] - 1) < 0.05
    # generate: a surface within the budget, its counter, packed vertices
    # inside the lattice
    assert out["triangles"] > 1000 and out["emitted"] == out["triangles"]
    assert out["lastTriangles"] == out["triangles"]
    assert out["vertsPtrNonZero"] and out["vertsBytes"] == 3000000 * 24
    assert out["vertWords"] == 6 * out["emitted"]
    assert out["vertsOutside"] == 0 and out["badNormal"] == 0
    # renderer exports: device phase buffer, packed body poses
    assert out["liveBodyPtrNonZero"] and out["liveBodyBytes"] > 0
    assert out["solidsCount"] == 3 and out["packedBytes"] == 3 * 96
    assert out["packErr"] == 0
    # snapshot through JS: the loaded sim continues bit for bit
    assert out["snapshotBytes"] > 0 and out["snapshotDiff"] == 0 and out["snapshotBodyPose"]
