"""Pins of the oracle's density fields (CPU).

* isotropic field (meshDensityWGSL restated): one particle gives the
  closed form m/rho * poly6 * (h^2 - r^2)^3 at every vertex within h and 0
  elsewhere; the lattice follows SurfaceMesh.configure;
* anisotropic splat: a single rotated ellipsoid is non-zero exactly at the
  lattice vertices inside it, with w (1 - s^2)^3, w = V * 315/(64 pi) *
  det G; its sum times voxel^3 is V (the kernel integrates to one);
* the two different algorithms (a float gather over the stale grid, and a
  fixed-point splat over ellipsoid bounding boxes) agree on a random cloud
  when G = I/h and the centres are not shifted, which is where the
  anisotropic field reduces to the isotropic one.
"""
import math

import numpy as np

from oracle import oracle as O

F = np.float64


def make_P(box=(16.0, 12.0, 10.0)):
    return O.constants(list(box), 1.0, 2.0, 1.0, 0.5, 1000.0, 9.8, 4, 0.01, 0.1, 0.3, 0.05,
                       1.0, False, 0.5, 0, 1 / 120)


def float4(p):
    out = np.zeros((len(p), 4))
    out[:, :3] = np.asarray(p, np.float32)
    return out


def vertices(lattice):
    vd, voxel, origin = lattice
    z, y, x = np.meshgrid(np.arange(vd[2]), np.arange(vd[1]), np.arange(vd[0]), indexing="ij")
    org = np.array(origin, np.float32)
    vox = np.float32(voxel)
    return np.stack([org[0] + vox * x.astype(np.float32), org[1] + vox * y.astype(np.float32),
                     org[2] + vox * z.astype(np.float32)], -1).reshape(-1, 3).astype(F)


def test_field_lattice_follows_configure():
    vd, voxel, origin = O.field_lattice([0, 0, 0], [16, 12, 10], 32)
    assert voxel == 0.5 and vd == [35, 27, 23] and origin == [-0.5, -0.5, -0.5]
    vd, voxel, origin = O.field_lattice([1, 2, 3], [8, 5, 4.5], 10)
    assert voxel == 0.7 and vd == [13, 8, 6]
    np.testing.assert_allclose(origin, [0.3, 1.3, 2.3])


def test_isotropic_single_particle_closed_form():
    P = make_P()
    lat = O.field_lattice([0, 0, 0], [16, 12, 10], 32)
    pos = float4([[5.3, 4.7, 6.1]])
    key = O.cell_ids(O.F32, P, pos)
    cs, _, _ = O.grid(P, key)
    rho = np.array([900.0])
    f = O.field_iso(P, lat, 1.0, 1000.0, pos, rho, np.zeros(1, np.uint32), 0, cs)
    p = vertices(lat)
    r2 = ((p - pos[:, :3]) ** 2).sum(1)
    want = np.where(r2 < 4.0, (1.0 / 900.0) * P.poly6 * (4.0 - r2) ** 3, 0.0)
    assert (want > 0).sum() > 100
    np.testing.assert_allclose(f, want, rtol=2e-5, atol=1e-12)
    # a body particle is skipped only when asked
    f0 = O.field_iso(P, lat, 1.0, 1000.0, pos, rho, np.ones(1, np.uint32), 1, cs)
    assert not f0.any()


def ellipsoid_g(axes, R):
    G = R @ np.diag(1.0 / np.asarray(axes)) @ R.T
    det = np.linalg.det(G)
    return np.array([[G[0, 0], G[0, 1], G[0, 2], G[1, 1]], [G[1, 2], G[2, 2], det, 30.0]]), G


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_anisotropic_single_ellipsoid():
    lat = O.field_lattice([0, 0, 0], [16, 12, 10], 64)   # voxel 0.25
    ctr = np.array([[7.1, 5.3, 4.9, 0.0]])
    g, G = ellipsoid_g([2.6, 1.1, 0.7], rotation(1))
    rho = np.array([1000.0])
    f = O.field_aniso(lat, 1.0, 1000.0, ctr, g, rho, np.zeros(1, np.uint32), 0)
    p = vertices(lat)
    d = (p - ctr[:, :3]) @ G.T
    s2 = (d ** 2).sum(1)
    w = (1.0 / 1000.0) * 315.0 / (64.0 * math.pi) * np.linalg.det(G)
    want = np.where(s2 < 1.0, w * (1.0 - s2) ** 3, 0.0)
    sure = np.abs(s2 - 1.0) > 1e-5
    assert (want > 0).sum() > 500
    np.testing.assert_allclose(f[sure], want[sure], rtol=1e-4, atol=1e-11)
    # the kernel integrates to one: sum * voxel^3 = V
    vox = lat[1]
    assert abs(f.astype(F).sum() * vox ** 3 / (1.0 / 1000.0) - 1.0) < 2e-3


def test_anisotropic_isotropic_limit_equals_the_isotropic_gather():
    P = make_P(box=(20.0, 16.0, 16.0))
    rng = np.random.default_rng(7)
    n = 4000
    pos = float4(rng.uniform([3, 3, 3], [17, 13, 13], (n, 3)))
    key = O.cell_ids(O.F32, P, pos)
    cs, _, order = O.grid(P, key)
    pos = pos[order]
    rho = rng.uniform(700, 1300, n)
    lat = O.field_lattice([0, 0, 0], [20, 16, 16], 40)
    fi = O.field_iso(P, lat, 1.0, 1000.0, pos, rho, np.zeros(n, np.uint32), 0, cs)
    h = 2.0
    g = np.zeros((2 * n, 4))
    g[0::2] = [1 / h, 0.0, 0.0, 1 / h]
    g[1::2] = [0.0, 1 / h, 1 / h ** 3, 30.0]
    fa = O.field_aniso(lat, 1.0, 1000.0, pos, g, rho, np.zeros(n, np.uint32), 0)
    assert fi.max() > 1e-3 and (fi > 0).mean() > 0.2
    np.testing.assert_allclose(fa, fi, rtol=0, atol=2e-6 * fi.max())
