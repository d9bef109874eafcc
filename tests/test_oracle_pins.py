"""Pins the oracle (oracle/oracle.cpp) against closed forms, library
routines, invariants and brute force, so that the GPU parity tests compare
against something that is itself checked.  CPU only.

What each test pins:
  cell ids           IEEE float division + floor + clamp, written in numpy
  grid               numpy's stable argsort and bincount
  census             brute-force O(n^2) neighbour sets
  lattice density    rho = m * sum_o W(|o| d) on an unjittered lattice
  two-body forms     lambda / delta / xsph / normals / tension for a pair,
                     derived from the formulas of the reference kernels
  momentum           sum of delta-p and of the tension impulses vanish
                     (antisymmetric pair terms), uniform velocity -> xsph 0
  polar              a rigidly rotated body gives back its rotation
  modes              float mode and FP64 mode agree to float precision
  kernels            poly6 and spiky are unit-normalised (radial quadrature)
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def make_P(box=(16.0, 12.0, 10.0), d=1.0, mass=1.0, denom=0.5, dt=1 / 120, tension=0.5,
           xsph=0.05, sor=False, nb=0, iters=4):
    return O.constants(list(box), d, 2 * d, mass, denom, 1000.0, 9.8, iters, 0.01, 0.1, 0.3,
                       xsph, 1.0, sor, tension, nb, dt)


def float4(p):
    out = np.zeros((len(p), 4))
    out[:, :3] = np.asarray(p, np.float32)
    return out


def fresh_grid(P, pred, mode=O.F32):
    key = O.cell_ids(mode, P, pred)
    cs, slot, order = O.grid(P, key)
    return key, cs, slot, order


def test_kernel_normalisation():
    h = 2.0
    r = np.linspace(0, h, 200001)
    poly6 = 315 / (64 * math.pi * h ** 9) * (h * h - r * r) ** 3
    spiky = 15 / (math.pi * h ** 6) * (h - r) ** 3
    assert abs(np.trapezoid(4 * math.pi * r * r * poly6, r) - 1) < 1e-8
    assert abs(np.trapezoid(4 * math.pi * r * r * spiky, r) - 1) < 1e-8
    P = make_P()
    # spiky gradient coefficient is d/dr of the unit spiky kernel
    assert P.spiky == pytest.approx(-3 * 15 / (math.pi * h ** 6), rel=1e-15)


def test_cell_ids_closed_form():
    rng = np.random.default_rng(0)
    P = make_P()
    g = [int(x) for x in P.gridDim]
    p = rng.uniform(-3, 20, (5000, 3)).astype(np.float32)
    p[:100] = np.round(p[:100])  # exact cell boundaries
    p[100] = np.nan
    pred = float4(p)
    key = O.cell_ids(O.F32, P, pred)
    h32 = np.float32(P.h)
    with np.errstate(invalid="ignore"):
        c = np.floor(p / h32).astype(np.float64)
    c = np.nan_to_num(c, nan=0.0)
    for a in range(3):
        c[:, a] = np.clip(c[:, a], 0, g[a] - 1)
    c = c.astype(np.int64)
    want = (c[:, 2] * g[1] + c[:, 1]) * g[0] + c[:, 0]
    assert np.array_equal(key, want.astype(np.uint32))


def test_grid_is_stable_counting_sort():
    rng = np.random.default_rng(1)
    P = make_P()
    nc = O.n_cells(P)
    key = rng.integers(0, nc, 20000).astype(np.uint32)
    key[:5000] = rng.integers(0, 7, 5000)  # heavy cells
    cs, slot, order = O.grid(P, key)
    perm = np.argsort(key, kind="stable")
    assert np.array_equal(order, perm)
    assert np.array_equal(slot[perm], np.arange(len(key)))
    counts = np.bincount(key, minlength=nc)
    assert np.array_equal(cs[1:], np.cumsum(counts))
    assert cs[0] == 0 and cs[-1] == len(key)


def _r2_f32(a, b):
    """fma(z,z, fma(y,y, x*x)) of float32 differences, evaluated in FP64
    (exact except in rare double-rounding cases, which the callers treat as
    borderline)."""
    d = (a.astype(np.float32) - b.astype(np.float32)).astype(np.float64)
    xx = np.float32(d[..., 0] * d[..., 0]).astype(np.float64)
    t = np.float32(d[..., 1] * d[..., 1] + xx).astype(np.float64)
    return np.float32(d[..., 2] * d[..., 2] + t).astype(np.float64)


def test_census_matches_brute_force():
    rng = np.random.default_rng(2)
    P = make_P(box=(14.0, 10.0, 8.0))
    p = rng.uniform(0, [14, 10, 8], (1500, 3)).astype(np.float32)
    key, cs, slot, order = fresh_grid(P, float4(p))
    ps = p[order]
    cnt, s1, s2 = O.census(P, float4(ps), cs)
    r2 = _r2_f32(ps[:, None, :], ps[None, :, :])
    h2 = np.float32(P.h2)
    near = r2 < h2
    np.fill_diagonal(near, False)
    border = np.abs(r2 - h2) <= 4 * np.spacing(h2)
    ok_rows = ~border.any(axis=1)
    assert ok_rows.mean() > 0.99
    j = np.arange(len(ps), dtype=np.uint64)
    want_cnt = near.sum(axis=1)
    want_s1 = (near * j[None, :]).sum(axis=1) % (1 << 32)
    want_s2 = (near * (j * j % (1 << 32))[None, :]).sum(axis=1) % (1 << 32)
    assert np.array_equal(cnt[ok_rows], want_cnt[ok_rows])
    assert np.array_equal(s1[ok_rows], want_s1[ok_rows].astype(np.uint32))
    assert np.array_equal(s2[ok_rows], want_s2[ok_rows].astype(np.uint32))


def lattice_mass(d, h, rho0):
    coef = 315 / (64 * math.pi * h ** 9)
    s = 0.0
    R = int(math.ceil(h / d)) + 1
    for a in range(-R, R + 1):
        for b in range(-R, R + 1):
            for c in range(-R, R + 1):
                r2 = (a * a + b * b + c * c) * d * d
                if r2 < h * h:
                    s += coef * (h * h - r2) ** 3
    return rho0 / s


def test_lattice_density_equals_rest_density():
    m = lattice_mass(1.0, 2.0, 1000.0)
    P = make_P(box=(12.0, 12.0, 12.0), mass=m)
    g = np.arange(12) + 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    key, cs, slot, order = fresh_grid(P, float4(pts), O.F64)
    ps = float4(pts[order])
    den, lam = O.lambda_(O.F64, P, ps, cs)
    interior = np.all((ps[:, :3] > 2.5) & (ps[:, :3] < 9.5), axis=1)
    assert interior.sum() > 100
    assert np.max(np.abs(den[interior] / 1000.0 - 1)) < 1e-12
    assert np.all(lam <= 0)  # lambda is never positive
    # fewer neighbours at the surface: density below rest, lambda zero
    assert np.all(lam[~interior & (den < 1000.0)] == 0)


def pair_P(r, mass=3000.0, **kw):
    P = make_P(box=(20.0, 20.0, 20.0), mass=mass, **kw)
    return P


def pair_state(r, u=(0.6, 0.48, 0.64)):
    u = np.asarray(u) / np.linalg.norm(u)
    a = np.array([10.1, 10.2, 10.3])
    return float4([a, a + r * u]), u


@pytest.mark.parametrize("r", [0.3, 1.0, 1.7])
def test_pair_lambda_delta_closed_form(r):
    P = pair_P(r)
    pred, u = pair_state(r)
    key, cs, slot, order = fresh_grid(P, pred, O.F64)
    ps = pred[order]
    den, lam = O.lambda_(O.F64, P, ps, cs)
    h, h2 = P.h, P.h2
    r = float(np.linalg.norm(ps[1, :3] - ps[0, :3]))
    W = lambda q: P.poly6 * (h2 - q * q) ** 3
    rho = P.mass * (W(0.0) + W(r))
    s = P.volume * P.spiky * (h - r) ** 2 / r  # |grad| / r
    g2 = (s * r) ** 2
    C = rho / P.rho0 - 1
    want = -C / (g2 + g2 + P.cfmEps + 1e-12) if C > 0 else 0.0
    assert den == pytest.approx([rho, rho], rel=1e-13)
    assert lam == pytest.approx([want, want], rel=1e-12)
    out = O.delta(O.F64, P, ps, lam, cs)
    sc = -P.sCorrK * (W(r) * P.sCorrWq) ** 4
    rij = ps[0, :3] - ps[1, :3]
    dp = (lam[0] + lam[1] + sc) * (P.spiky * (h - r) ** 2 / r) * rij * P.volume
    np.testing.assert_allclose(out[0, :3], ps[0, :3] + dp, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[1, :3], ps[1, :3] - dp, rtol=0, atol=1e-12)


@pytest.mark.parametrize("r", [0.3, 1.0, 1.7])
def test_one_ulp_mode_shifts_only_sqrt_and_division(r):
    """The oracle's F32_ULP mode (a conforming implementation of WGSL's f32
    sqrt and division, one ulp above the correctly rounded result) against
    a float32 evaluation of the pair lambda with numpy's correctly rounded
    sqrt and division, unshifted (= F32) and shifted by np.nextafter
    (= F32_ULP), bit for bit.  The pair lies along x, so r^2 is one rounded
    product (the fma chain adds exact zeros)."""
    P = pair_P(r)
    pred, _ = pair_state(r, u=(1.0, 0.0, 0.0))
    key, cs, slot, order = fresh_grid(P, pred)
    ps = pred[order]
    f = np.float32
    h, h2, mass, poly6 = f(P.h), f(P.h2), f(P.mass), f(P.poly6)
    vol, spiky, inv_rho0, eps = f(P.volume), f(P.spiky), f(P.invRho0), f(P.cfmEps)
    up = lambda x: np.nextafter(f(x), f(np.inf))

    def lam_pair(i, shift):
        sq = (lambda x: up(np.sqrt(x))) if shift else np.sqrt
        dv = (lambda a, b: up(a / b)) if shift else (lambda a, b: f(a / b))
        rx = f(ps[i, 0]) - f(ps[1 - i, 0])
        r2 = rx * rx
        rho = mass * (poly6 * h2 * h2 * h2)
        t = h2 - r2
        rho = rho + mass * (poly6 * t * t * t)
        rr = sq(r2)
        hr = h - rr
        s = dv(vol * spiky * hr * hr, rr)
        ax, ay, az = s * rx, s * f(0), s * f(0)
        sg2 = f(0) + (ax * ax + ay * ay + az * az)
        C = rho * inv_rho0 - f(1)
        if not C > 0:
            return f(0)
        return dv(-C, (ax * ax + ay * ay + az * az) + sg2 + eps + f(1.0e-12))

    lams = {}
    for mode, shift in ((O.F32, False), (O.F32_ULP, True)):
        _, lam = O.lambda_(mode, P, ps, cs)
        lams[mode] = lam
        for i in range(2):
            assert np.float32(lam[i]) == lam_pair(i, shift), (mode, i)
    l0, l2 = lams[O.F32][0], lams[O.F32_ULP][0]
    if r < 1.0:
        assert l0 < 0  # a compressed pair: the shift reaches lambda
    if l0 != 0:
        assert l0 != l2 and abs(l2 - l0) <= 8 * abs(np.spacing(np.float32(l0)))


def test_pair_xsph_normals_tension_closed_form():
    r = 1.2
    P = pair_P(r, mass=900.0)
    pred, u = pair_state(r)
    key, cs, slot, order = fresh_grid(P, pred, O.F64)
    ps = pred[order]
    pos = ps.copy()
    v = np.array([[0.5, -1.0, 0.25], [-0.75, 0.5, 1.5]])
    pos[:, :3] = ps[:, :3] - v * P.dt  # so that (pred - pos)/dt = v
    vel_true = (ps[:, :3] - pos[:, :3]) * P.invDt
    den = np.array([1020.0, 480.0])
    vel, corr, normal = O.xsph(O.F64, P, ps, pos, den, cs, True)
    h, h2 = P.h, P.h2
    dh = np.maximum(den, 0.5 * P.rho0)
    rij = ps[0, :3] - ps[1, :3]
    rr = np.linalg.norm(rij)  # r after rounding the positions to float
    assert rr == pytest.approx(r, rel=1e-6)
    W = P.poly6 * (h2 - rr * rr) ** 3
    want0 = P.xsphC * P.mass * W / dh[1] * (vel_true[1] - vel_true[0])
    want1 = P.xsphC * P.mass * W / dh[0] * (vel_true[0] - vel_true[1])
    np.testing.assert_allclose(corr[0, :3], want0, atol=1e-12)
    np.testing.assert_allclose(corr[1, :3], want1, atol=1e-12)
    g = P.spiky * (h - rr) ** 2 / rr
    np.testing.assert_allclose(normal[0, :3], h * P.mass / dh[1] * g * rij, atol=1e-12)
    np.testing.assert_allclose(normal[1, :3], -h * P.mass / dh[0] * g * rij, atol=1e-12)
    # Akinci cohesion + curvature for the pair
    c2 = O.tension(O.F64, P, ps, den, normal, cs, corr)
    k = (h - rr) ** 3 * rr ** 3
    cw = P.cohes * k if 2 * rr > h else P.cohes * (2 * k - P.cohesTerm)
    K = 2 * P.rho0 / (den[0] + den[1])
    f0 = K * (-P.tension * P.mass * cw / rr * rij - P.tension * (normal[0, :3] - normal[1, :3]))
    np.testing.assert_allclose(c2[0, :3] - corr[0, :3], P.dt * f0, atol=1e-12)
    np.testing.assert_allclose(c2[1, :3] - corr[1, :3], -P.dt * f0, atol=1e-12)


def random_cloud(n=3000, seed=3, box=(24.0, 20.0, 20.0)):
    rng = np.random.default_rng(seed)
    lo = np.array([4.0, 4.0, 4.0])
    hi = np.array(box) - 4.0
    return rng.uniform(lo, hi, (n, 3))


def test_delta_momentum_and_tension_momentum_vanish():
    m = lattice_mass(1.0, 2.0, 1000.0) * 1.3
    P = make_P(box=(24.0, 20.0, 20.0), mass=m)
    P.clampMin = type(P.clampMin)(-1e9, -1e9, -1e9)
    P.clampMax = type(P.clampMax)(1e9, 1e9, 1e9)
    p = random_cloud()
    key, cs, slot, order = fresh_grid(P, float4(p), O.F64)
    ps = float4(p[order])
    den, lam = O.lambda_(O.F64, P, ps, cs)
    assert (lam < 0).sum() > 100
    out = O.delta(O.F64, P, ps, lam, cs)
    dp = out[:, :3] - ps[:, :3]
    assert np.abs(dp).max() > 1e-3
    assert np.abs(dp.sum(axis=0)).max() < 1e-9 * np.abs(dp).sum()
    pos = ps.copy()
    vel, corr, normal = O.xsph(O.F64, P, ps, pos, den, cs, True)
    c2 = O.tension(O.F64, P, ps, den, normal, cs, corr)
    imp = c2[:, :3] - corr[:, :3]
    assert np.abs(imp).max() > 0
    assert np.abs(imp.sum(axis=0)).max() < 1e-9 * np.abs(imp).sum()


def test_uniform_velocity_gives_no_xsph():
    P = make_P(box=(24.0, 20.0, 20.0))
    p = random_cloud()
    key, cs, slot, order = fresh_grid(P, float4(p), O.F64)
    ps = float4(p[order])
    pos = ps.copy()
    pos[:, :3] -= np.array([0.3, -0.2, 0.1]) * P.dt
    den = np.full(len(ps), 1000.0)
    vel, corr, normal = O.xsph(O.F64, P, ps, pos, den, cs, False)
    assert np.abs(corr[:, :3]).max() < 1e-10


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("inv_ratio", [1.0, 2.0])
def test_polar_recovers_rotation_and_centroid(inv_ratio):
    rng = np.random.default_rng(4)
    q = rng.uniform(-3, 3, (200, 3))
    q -= q.mean(axis=0)
    R = random_rotation(rng)
    t = np.array([30.0, 20.0, 10.0])
    pred = float4(q @ R.T + t)
    pred[:, :3] = (q @ R.T + t)  # keep FP64 positions exactly
    rest = float4(q)
    rest[:, :3] = q
    body = np.ones(len(q), np.uint32)
    c0 = np.array([[29.0, 21.0, 10.5]])
    out, centre, rot = O.body_solve(pred, body, rest, [inv_ratio], c0, np.eye(3).reshape(1, 9))
    np.testing.assert_allclose(rot.reshape(3, 3), R, atol=1e-12)
    np.testing.assert_allclose(centre[0], c0[0] + inv_ratio * (t - c0[0]), atol=1e-12)
    cen = c0 = O.body_pre(pred, body, np.array([[0.0, 0.0, 0.0]]))
    np.testing.assert_allclose(cen[0], t, atol=1e-12)


def test_float_and_fp64_modes_agree():
    m = lattice_mass(1.0, 2.0, 1000.0)
    P = make_P(box=(24.0, 20.0, 20.0), mass=m)
    p = random_cloud()
    key, cs, slot, order = fresh_grid(P, float4(p))
    ps = float4(p[order])
    d32, l32 = O.lambda_(O.F32, P, ps, cs)
    d64, l64 = O.lambda_(O.F64, P, ps, cs)
    np.testing.assert_allclose(d32, d64, rtol=2e-6)
    np.testing.assert_allclose(l32, l64, atol=2e-6 * np.abs(l64).max())
    o32 = O.delta(O.F32, P, ps, l64, cs)
    o64 = O.delta(O.F64, P, ps, l64, cs)
    np.testing.assert_allclose(o32[:, :3], o64[:, :3], atol=4e-6)


def test_stats_closed_form():
    rng = np.random.default_rng(5)
    den = rng.uniform(900, 1100, 1000)
    vel = np.zeros((1000, 4))
    vel[:, :3] = rng.normal(size=(1000, 3))
    s = O.stats(den, vel, 1 / 1000.0, 2.0)
    v = np.float32(vel[:, :3]).astype(np.float64)
    assert s[0] == pytest.approx(np.mean(den / 1000.0), rel=1e-6)
    assert s[1] == pytest.approx(den.max() / 1000.0, rel=1e-6)
    assert s[2] == pytest.approx(np.sqrt((v ** 2).sum(1)).max(), rel=1e-6)
    assert s[3] == pytest.approx(0.5 * 2.0 * (v ** 2).sum(), rel=1e-5)


def test_boundary_sort_is_stable_counting_sort():
    from oracle import scene_ref
    P = make_P(box=(8.0, 6.0, 6.0))
    pts, psi = scene_ref.boundary_particles([8.0, 6.0, 6.0], 1.0, 2.0, 1000.0)
    pts = np.array(pts, np.float32)
    rec, start = O.sort_boundary(P, pts, psi)
    assert len(pts) == 9 * 7 * 7 - 7 * 5 * 5
    assert start[-1] == len(pts) and np.all(np.diff(start.astype(np.int64)) >= 0)
    g = [int(x) for x in P.gridDim]
    for c in range(O.n_cells(P)):
        seg = rec[start[c]:start[c + 1], :3]
        cc = np.clip(np.floor(seg / P.h), 0, np.array(g) - 1).astype(int)
        assert np.all((cc[:, 2] * g[1] + cc[:, 1]) * g[0] + cc[:, 0] == c)


def test_jacobi_matches_numpy_eigh():
    rng = np.random.default_rng(6)
    for mode, tol in ((O.F64, 1e-10), (O.F32, 2e-5)):
        for _ in range(200):
            m = rng.normal(size=(3, 3))
            a = m @ m.T
            w, v = O.jacobi(mode, a)
            ref = np.linalg.eigh(a)[0]
            np.testing.assert_allclose(np.sort(w), ref, atol=tol * np.abs(ref).max())
            # the columns of v are orthonormal eigenvectors: a v = v diag(w)
            np.testing.assert_allclose(a @ v, v * w, atol=10 * tol * np.abs(ref).max())


def test_aniso_on_a_lattice_is_isotropic_with_closed_form_count():
    P = make_P(box=(16.0, 16.0, 16.0))
    g = np.arange(16) + 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    key, cs, slot, order = fresh_grid(P, float4(pts), O.F64)
    ps = float4(pts[order])
    smooth, G = O.aniso(O.F64, P, ps, cs)
    G = G.reshape(-1, 8)
    R = 2 * P.h  # radius_scale 2
    want = sum(1 for a in range(-5, 6) for b in range(-5, 6) for c in range(-5, 6)
               if a * a + b * b + c * c < R * R)
    inner = np.all((ps[:, :3] > R + 0.5) & (ps[:, :3] < 16 - R - 0.5), axis=1)
    assert inner.sum() > 50
    assert np.all(G[inner, 7] == want)             # neighbour count incl. self
    np.testing.assert_allclose(smooth[inner, :3], ps[inner, :3], atol=1e-12)  # no drift
    off = np.abs(G[inner][:, [1, 2, 4]]).max()
    diag = G[inner][:, [0, 3, 5]]
    assert off < 1e-12 * diag.max()                # G is a multiple of I
    np.testing.assert_allclose(diag, diag[:, :1] * np.ones((1, 3)), rtol=1e-12)
    np.testing.assert_allclose(G[inner, 6], diag[:, 0] ** 3, rtol=1e-12)


def test_delta_at_moves_the_base_by_the_correction():
    """delta_at(pred, base=pred) is delta; with another base the same
    correction is added to it (away from the clamp)."""
    m = lattice_mass(1.0, 2.0, 1000.0) * 1.3
    P = make_P(box=(24.0, 20.0, 20.0), mass=m)
    p = random_cloud()
    key, cs, slot, order = fresh_grid(P, float4(p), O.F64)
    ps = float4(p[order])
    den, lam = O.lambda_(O.F64, P, ps, cs)
    out = O.delta(O.F64, P, ps, lam, cs)
    assert np.array_equal(O.delta_at(O.F64, P, ps, ps, lam, cs), out)
    base = ps.copy()
    base[:, :3] += 0.125
    out2 = O.delta_at(O.F64, P, ps, base, lam, cs)
    np.testing.assert_allclose(out2[:, :3] - base[:, :3], out[:, :3] - ps[:, :3], atol=1e-12)


def test_fp16_quantize_rounds_offsets_from_cell_centres():
    """x is kept as float; y and z offsets are taken from the cell centre
    (boxMin + (index + 0.5) h) and rounded to nearest-even half: a position
    at centre + k/1024 is kept exactly; centre + 0.3 moves by half(0.3) -
    0.3 (the corner-origin offset 1.3 would round on a twice coarser grid);
    every move is within half a half-ulp of the offset."""
    P = make_P()  # h = 2, boxMin = 0
    cells = np.array([0, 1, 8 + 3, 8 * 6 * 2 + 8 * 5 + 7], np.int64)
    centres = O.cell_centres(P, cells).astype(np.float64)
    assert np.array_equal(centres[0], [1.0, 1.0, 1.0])
    assert np.array_equal(centres[3], [15.0, 11.0, 5.0])
    exact = float4(centres + np.array([17, -300, 1023]) / 1024.0)
    q = O.fp16_quantize(P, exact, cells)
    assert np.array_equal(q[:, :3], exact[:, :3])
    off = float4(centres + 0.3)
    q = O.fp16_quantize(P, off, cells)
    assert np.array_equal(q[:, 0], off[:, 0])
    o = (off[:, 1:3] - centres[:, 1:]).astype(np.float32)  # exact: |o| < 1
    step = o.astype(np.float16).astype(np.float64) - o
    np.testing.assert_array_equal(q[:, 1:3] - off[:, 1:3], step)
    o_corner = o + np.float32(1.0)
    step_corner = o_corner.astype(np.float16).astype(np.float64) - o_corner
    assert np.any(step_corner != step)
    rng = np.random.default_rng(5)
    pts = float4(centres[rng.integers(0, 4, 2000)] + rng.uniform(-1.5, 1.5, (2000, 3)))
    cid = cells[np.argmin(np.abs(pts[:, None, :3] - centres[None]).sum(-1), axis=1)]
    q = O.fp16_quantize(P, pts, cid)
    assert np.array_equal(q[:, 0], pts[:, 0])
    o = pts[:, 1:3] - O.cell_centres(P, cid)[:, 1:]
    bound = 0.5 * np.spacing(np.abs(o).astype(np.float16)).astype(np.float64)
    assert np.all(np.abs(q[:, 1:3] - pts[:, 1:3]) <= bound + 1e-9)


def test_body_pack_closed_form():
    """bodyPackWGSL restated: a rotation of 90 degrees about z (rows (0,-1,0),
    (1,0,0), (0,0,1)) packs its columns (0,1,0), (-1,0,0), (0,0,1)."""
    centre = np.array([[1.0, 2.0, 3.0, 0.0], [4.0, 5.0, 6.0, 0.0]])
    rot = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0],
                    [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float64)
    stat = np.array([[0.9, 0.7, 0.9, 0, 0.55, 0.72, 0.95, 0],
                     [2.0, 2.0, 2.0, 1, 0.60, 0.88, 0.66, 0]])
    out = O.body_pack(centre, rot, stat).reshape(2, 6, 4)
    assert np.array_equal(out[0, 0], np.float32([1, 2, 3, 0]))
    assert np.array_equal(out[1, 0], np.float32([4, 5, 6, 1]))
    assert np.array_equal(out[0, 1], np.float32([0.9, 0.7, 0.9, 0]))
    assert np.array_equal(out[0, 2:5, :3], np.float32([[0, 1, 0], [-1, 0, 0], [0, 0, 1]]))
    assert np.array_equal(out[1, 2:5, :3], np.eye(3, dtype=np.float32))
    assert np.array_equal(out[1, 5], np.float32([0.60, 0.88, 0.66, 0]))


def test_ssfr_kernel_g_layout():
    """The packed G (G00, G10, G20, G11), (G21, G22, det, n) read back as the
    SSFR splat does gives the symmetric matrix it was packed from."""
    rng = np.random.default_rng(3)
    a = rng.normal(size=(5, 3, 3))
    G = a @ a.transpose(0, 2, 1) + np.eye(3)
    g = np.zeros((5, 2, 4))
    g[:, 0] = np.stack([G[:, 0, 0], G[:, 1, 0], G[:, 2, 0], G[:, 1, 1]], -1)
    g[:, 1] = np.stack([G[:, 2, 1], G[:, 2, 2], np.linalg.det(G), np.full(5, 30.0)], -1)
    assert np.allclose(O.ssfr_kernel_g(g), G)


def test_impulse_closed_forms():
    """impulseWGSL restated: a particle on the ray gets the full impulse (x =
    1), one at radial distance r gets (1 - r/R)^2 of it; particles behind the
    origin or at r >= R keep their velocity (w too); a capped velocity keeps
    its direction with length max(|v|, limit); hit particles get w = 0."""
    pos = np.float32([[5, 1, 1, 1], [5, 1, 2.5, 1], [-1, 1, 1, 1], [5, 3, 1, 1],
                      [7, 1, 1, 1]])
    vel = np.zeros((5, 4), np.float32)
    vel[:, 3] = 7
    vel[4] = [0, 2, 0, 7]
    out = O.impulse(pos, vel, [0, 1, 1], [4, 0, 0], [0.5, 0, 0], 2.0, 10.0)
    assert np.array_equal(out[0], np.float32([0.5, 0, 0, 0]))
    assert np.array_equal(out[1], np.float32([0.5 * 0.25 ** 2, 0, 0, 0]))
    assert np.array_equal(out[2], vel[2]) and np.array_equal(out[3], vel[3])
    assert np.array_equal(out[4], np.float32([0.5, 2, 0, 0]))
    # cap: |v + J| = |(30, 2, 0)| > max(|v| = 2, limit 3) -> length 3, same direction
    out = O.impulse(pos, vel, [0, 1, 1], [1, 0, 0], [30, 0, 0], 2.0, 3.0)
    n = np.linalg.norm(out[4, :3].astype(np.float64))
    assert abs(n - 3.0) < 1e-6
    assert abs(out[4, 0] / out[4, 1] - 15.0) < 1e-5
    # no-ops: radius <= 0, zero impulse
    assert np.array_equal(O.impulse(pos, vel, [0, 1, 1], [1, 0, 0], [1, 0, 0], 0.0, 1.0), vel)
    assert np.array_equal(O.impulse(pos, vel, [0, 1, 1], [1, 0, 0], [0, 0, 0], 2.0, 1.0), vel)


def test_body_drag_closed_forms():
    """bodyDragWGSL restated: from rest a member moves to blend * w with
    w = (goal - centre) * rate, or w scaled to `limit` when longer;
    blend = 1 - exp(-dt/0.02); other particles are untouched."""
    vel = np.zeros((4, 4), np.float32)
    vel[:, 3] = 5
    centre = np.float32([1, 2, 3, 0])
    dt = 1 / 120
    blend = np.float32(1 - math.exp(-dt / 0.02))
    out = O.body_drag(vel, [1, 3], centre, [2, 2, 3], 4.0, 100.0, dt)
    assert np.allclose(out[1, :3], [4 * blend, 0, 0], rtol=1e-6, atol=0)
    assert np.array_equal(out[0], vel[0]) and np.array_equal(out[2], vel[2])
    assert out[1, 3] == 5 and out[3, 3] == 5
    out = O.body_drag(vel, [1], centre, [1, 2, 13], 4.0, 2.0, dt)  # |w| = 40 > 2
    assert np.allclose(out[1, :3], [0, 0, 2 * blend], rtol=1e-6, atol=0)


def test_resize_closed_forms():
    """resizeWGSL restated: a shrinking axis maps [0, old] onto [0, new]
    (outside the old box clamps to its face), a growing axis keeps the
    coordinate; then every axis is clamped to [d/2, box - d/2]."""
    pos = np.float32([[8, 9.8, 1, 1], [12, 0.1, 1, 1], [2, 5, 0.2, 1]])
    out = O.resize(pos, [10, 10, 10], [5, 20, 10], 1.0)
    assert np.array_equal(out[0, :3], np.float32([4, 9.8, 1]))
    assert np.array_equal(out[1, :3], np.float32([4.5, 0.5, 1]))
    assert np.array_equal(out[2, :3], np.float32([1, 5, 0.5]))
    assert np.array_equal(out[:, 3], pos[:, 3])
