"""The anisotropy pass on the GPU (aniso_wgsl.js:3-159 with the parameters
of mesh.js:221-242) against closed forms and the FP64 eigensolve, through
the C ABI:

* unjittered lattices and sheets, where the covariance is sigma I (or
  diag(sigma, 0, sigma)) with sigma a lattice sum: G equals the size rule's
  closed form with no cap binding, with the stretch cap, the minimum extent,
  the kr clamp and the neighbour-count blend each binding, in both math
  forms;
* a lattice stretched by Q diag(1, 1 + e, 1 + 2e) Q^T (e = 1e-5, a random
  frame): every eigen-gap is below 1e-4 of the spectrum, so each interior
  particle is re-solved in FP64; G equals the FP64 oracle and the size rule
  applied to numpy.linalg.eigh of the covariance summed in FP64, and with
  the re-solve switched off no particle is re-solved;
* the dam break of 2,097,152 particles after 600 substeps (5 s of flow),
  with the reference's minimum extent for a 256 field: neighbour counts
  equal the float oracle's and G is within 2e-4 |G| of the FP64 oracle on
  20,000 sampled particles.
"""
import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from oracle import oracle as O
from test_gpu_edges import CustomScene
from test_scene_pins import block, inner, offsets, sheet, size_rule, unpack, variance

pytestmark = pytest.mark.gpu


def gpu_aniso(pts, math, **kw):
    p, base, _ = workloads.dam_break_32(scale=0.5, math=math, bodies=False)
    assert list(p.box) == [32.0, 32.0, 32.0]
    sc = CustomScene(base, pts)
    sim = Sim(0)
    sim.reset(p, sc)
    fb = sim.aniso(**kw)
    pos = sim.read("POS").astype(np.float64)
    g = sim.read("ANISO_G").astype(np.float64).reshape(-1, 8)
    cs = sim.read("CELL_START")
    P = O.constants_from(p, sc, 1 / 120)
    sim.close()
    return pos, g, fb, cs, P


@pytest.mark.parametrize("math", [1, 0])
@pytest.mark.parametrize("kw,binds", [(dict(), "none"), (dict(ks=10.0), "stretch"),
                                      (dict(ks=0.1, min_extent=0.5), "min_extent")])
def test_lattice_sizes_closed_form(math, kw, binds):
    R = 4.0
    o = offsets(3, R)
    want = size_rule(variance(o, R), len(o), **kw)
    pos, g, fb, _, _ = gpu_aniso(block(18), math, **kw)
    sel = inner(pos, 6.5 + R, 6.5 + 17 - R)
    assert sel.sum() >= 27
    assert np.all(g[sel, 7] == len(o))
    G = unpack(g[sel])
    err = np.abs(G - np.diag(want)).max() / want.max()
    print(f"math {math} {binds}: G {want} max rel error {err:.2e}, FP64 re-solves {fb}")
    assert err <= 5e-6
    np.testing.assert_allclose(g[sel, 6], np.prod(want), rtol=1e-5)
    assert fb >= sel.sum()  # a triple eigenvalue: every interior particle re-solves


@pytest.mark.parametrize("math", [1, 0])
def test_sheet_kr_clamp_and_blend(math):
    R = 4.0
    for spacing, count in ((1.0, 45), (1.5, 21)):
        o = offsets(2, R, spacing)[:, [0, 2, 1]]
        want = size_rule(variance(o, R), count)
        pos, g, fb, _, _ = gpu_aniso(sheet(int(20 / spacing), spacing), math)
        sel = (np.abs(pos[:, 0] - 16.5) < 2.5) & (np.abs(pos[:, 2] - 16.5) < 2.5)
        assert sel.sum() >= 4 and np.all(g[sel, 7] == count)
        G = unpack(g[sel])
        err = np.abs(G - np.diag(want)).max() / want.max()
        print(f"math {math} sheet spacing {spacing}: G {want} max rel error {err:.2e}")
        assert err <= 5e-6


def eigh_g(pos, i, R=4.0, h=2.0):
    """The size rule on numpy.linalg.eigh of target i's covariance, summed
    in FP64 over all particles within R."""
    d = pos[:, :3] - pos[i, :3]
    r2 = (d * d).sum(1)
    m = r2 < R * R
    d, r = d[m], np.sqrt(r2[m])
    w = 1.0 - (r / R) ** 3
    db = (w[:, None] * d).sum(0) / w.sum()
    C = (w[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0) / w.sum() - np.outer(db, db)
    sig, V = np.linalg.eigh(C)
    return V @ np.diag(size_rule(sig, m.sum())) @ V.T, sig


@pytest.mark.parametrize("math", [1, 0])
def test_near_degenerate_spectra_take_the_fp64_solve(math):
    rng = np.random.default_rng(12)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    e = 1.0e-5
    A = q @ np.diag([1.0, 1.0 + e, 1.0 + 2 * e]) @ q.T
    c = np.array([16.0, 16.0, 16.0])
    pts = c + (block(16, origin=(8.0, 8.0, 8.0)) - c) @ A.T
    pos, g, fb, cs, P = gpu_aniso(pts, math)
    sel = np.nonzero(np.linalg.norm(pos[:, :3] - c, axis=1) < 3.0)[0]
    assert len(sel) >= 27
    _, g64 = O.aniso(O.F64, P, pos, cs, sel=sel)
    g64 = g64.reshape(-1, 8)
    G, G64 = unpack(g[sel]), unpack(g64)
    scale = np.abs(G64).max()
    worst_gap = 0.0
    for k, i in enumerate(sel):
        Ge, sig = eigh_g(pos, i)
        worst_gap = max(worst_gap, (np.diff(sig).min()) / sig.max())
        np.testing.assert_allclose(G64[k], Ge, atol=1e-9 * scale)
    err = np.abs(G - G64).max() / scale
    print(f"math {math}: largest relative eigen-gap {worst_gap:.2e}, FP64 re-solves {fb}, "
          f"G max rel error against FP64 {err:.2e}")
    assert worst_gap < 1e-4
    assert fb >= len(sel)
    assert err <= 1e-5
    assert np.array_equal(g[sel, 7], g64[:, 7])
    _, _, fb_off, _, _ = gpu_aniso(pts, math, fp64_fallback=0)
    assert fb_off == 0


def test_dam_break_2m_after_600_substeps_against_fp64():
    p, sc, frame = workloads.dam_break_2m(math=0)
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(1 / 120, 600)
    P = O.constants_from(p, sc, 1 / 120)
    pos, cs = sim.read("POS").astype(np.float64), sim.read("CELL_START")
    st = sim.stats()
    # the reference's minimum extent for a field of resolution 256 over the
    # 256^3 box (anisoLimitToField, mesh.js:89, 238): 1.25 voxel / h
    me = 1.25 * 1.0 / sc.h
    fb = sim.aniso(min_extent=me)
    g = sim.read("ANISO_G").astype(np.float64).reshape(-1, 8)
    sim.close()
    rng = np.random.default_rng(21)
    sel = np.sort(rng.choice(len(pos), 20000, replace=False))
    _, g64 = O.aniso(O.F64, P, pos, cs, sel=sel, min_extent=me)
    _, g32 = O.aniso(O.F32, P, pos, cs, sel=sel, min_extent=me)
    g64, g32 = g64.reshape(-1, 8), g32.reshape(-1, 8)
    assert np.array_equal(g[sel, 7], g32[:, 7])
    scale = np.abs(g64[:, [0, 3, 5]]).max()
    err = np.abs(g[sel, :6] - g64[:, :6])
    print(f"dam_break_2m after 600 substeps (maxSpeed {st[2]:.2f}): FP64 re-solves {fb}; G on 20000 "
          f"samples max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e} (|G| {scale:.3f})")
    assert err.max() <= 2e-4 * scale
    np.testing.assert_allclose(g[sel, 6], g64[:, 6], rtol=1e-3, atol=1e-6 * scale ** 3)
