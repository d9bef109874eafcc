"""ORACLE - test infrastructure only (tests/, __graft_entry__.smoke() and the
CPU baseline of bench.py).  Nothing in particles4all_amd imports it.

Python side of the CPU restatement in oracle.cpp:
  * `constants()` restates Sim.uploadParams (reference src/sim.js:371-413)
    from the scene scalars, independently of the library's own derivation;
  * `sort_boundary()` restates Sim.sortBoundary (sim.js:255-286);
  * thin ctypes wrappers of the kernels (mode 0 = op-by-op float in the
    WGSL order, mode 1 = FP64);
  * `substep()` chains them in the order of Sim.step (sim.js:458-596).
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
F32, F64 = 0, 1
# F32 with every sqrt and division of lambda, delta, xsph and tension one
# ulp above the correctly rounded result: within WGSL's accuracy for f32
# division (2.5 ULP) and sqrt (inherited from 1/inverseSqrt), so one
# conforming implementation of the reference's arithmetic.
F32_ULP = 2


class OrcParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("dt", "h", "h2", "poly6", "spiky", "gravity", "mass", "rho0", "invRho0",
                 "cfmEps", "sCorrK", "sCorrWq", "xsphC", "omega", "invDt", "volume", "tension",
                 "cohes", "cohesTerm")] + [
        ("boxMin", C.c_double * 3), ("clampMin", C.c_double * 3), ("clampMax", C.c_double * 3),
        ("gridDim", C.c_int32 * 3), ("sorAverage", C.c_int32), ("nBoundary", C.c_int32),
        ("pad", C.c_int32)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: build it with `make -C oracle`")
        _L = C.CDLL(LIB)
        _setup(_L)
    return _L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def grid_dim(box, h):
    """Sim.reset / resizeBox (sim.js:80-84): max(1, floor(box/h)) per axis."""
    return [max(1, int(math.floor(b / h))) for b in box]


def constants(box, spacing, h, mass, denom_rest, rho0, gravity, iterations, cfm, scorr_k,
              scorr_dq, xsph_c, omega, sor_average, tension, n_boundary, dt, no_boundary=False):
    """Sim.uploadParams (sim.js:371-413) as FP64 values."""
    P = OrcParams()
    P.dt = dt
    P.h = h
    P.h2 = h * h
    P.poly6 = 315.0 / (64.0 * math.pi * math.pow(h, 9))
    P.spiky = -45.0 / (math.pi * math.pow(h, 6))
    P.gravity = gravity
    P.mass = mass
    P.rho0 = rho0
    P.invRho0 = 1.0 / rho0
    P.cfmEps = max(1e-9, cfm * denom_rest)
    P.sCorrK = scorr_k / (denom_rest * max(1, iterations))
    rq = scorr_dq * h
    tq = h * h - rq * rq
    wq = P.poly6 * tq * tq * tq
    P.sCorrWq = 1.0 / wq if wq > 0 else 0.0
    P.xsphC = xsph_c
    P.omega = omega
    P.invDt = 1.0 / dt if dt > 0 else 0.0
    P.volume = mass / rho0
    P.tension = tension
    P.cohes = 32.0 / (math.pi * math.pow(h, 9))
    P.cohesTerm = math.pow(h, 6) / 64.0
    half = 0.5 * spacing
    P.boxMin = (C.c_double * 3)(0, 0, 0)
    P.clampMin = (C.c_double * 3)(half, half, half)
    P.clampMax = (C.c_double * 3)(*[b - half for b in box])
    P.gridDim = (C.c_int32 * 3)(*grid_dim(box, h))
    P.sorAverage = 1 if sor_average else 0
    P.nBoundary = 0 if no_boundary else int(n_boundary)
    return P


def constants_from(params, scene, dt):
    """constants() from a particles4all_amd Params struct and Scene."""
    return constants(list(params.box), params.spacing, scene.h, scene.mass, scene.denom_rest,
                     params.rest_density, params.gravity, params.iterations,
                     params.cfm_epsilon_rel, params.scorr_k, params.scorr_dq, params.xsph_c,
                     params.omega, params.sor_average, params.surface_tension_k,
                     len(scene.boundary_psi), dt, bool(params.no_boundary))


def n_cells(P):
    return int(P.gridDim[0]) * int(P.gridDim[1]) * int(P.gridDim[2])


def sort_boundary(P, pts, psi):
    """Sim.sortBoundary (sim.js:255-286): stable counting sort of the wall
    samples into the simulation grid, cells from float samples in FP64.
    Returns (sorted float4 records with w = psi, bcellStart[nCells+2])."""
    pts = np.asarray(pts, np.float32).astype(np.float64)
    g = [int(x) for x in P.gridDim]
    c = np.floor(pts / P.h)
    for a in range(3):
        c[:, a] = np.clip(c[:, a], 0, g[a] - 1)
    c = c.astype(np.int64)
    cell = (c[:, 2] * g[1] + c[:, 1]) * g[0] + c[:, 0]
    nc = n_cells(P)
    order = np.argsort(cell, kind="stable")
    counts = np.bincount(cell, minlength=nc + 2)
    start = np.zeros(nc + 2, np.uint32)
    start[1:nc + 1] = np.cumsum(counts[:nc])
    start[nc + 1] = start[nc] + counts[nc]
    rec = np.zeros((len(pts), 4), np.float64)
    rec[:, :3] = pts[order]
    rec[:, 3] = np.asarray(psi, np.float32)[order]
    return rec, start


def cell_ids(mode, P, pred):
    pred = np.ascontiguousarray(pred, np.float64)
    if pred.ndim != 2 or pred.shape[1] != 4:
        raise ValueError("cell_ids takes float4 records (n x 4)")
    key = np.zeros(len(pred), np.uint32)
    lib().orc_cell_ids(mode, C.byref(P), len(pred), _p(pred), _p(key))
    return key


def grid(P, key):
    """Stable counting sort by cell: (cellStart[nCells+1], slot, order)."""
    nc = n_cells(P)
    key = np.ascontiguousarray(key, np.uint32)
    n = len(key)
    cs = np.zeros(nc + 1, np.uint32)
    slot = np.zeros(n, np.uint32)
    order = np.zeros(n, np.uint32)
    lib().orc_grid(n, nc, _p(key), _p(cs), _p(slot), _p(order))
    return cs, slot, order


def predict(mode, P, pos, vel):
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    pred = np.zeros_like(pos)
    v = np.zeros_like(vel)
    lib().orc_predict(mode, C.byref(P), len(pos), _p(pos), _p(vel), _p(pred), _p(v))
    return pred, v


def _bnd(bpos, bpsi, bcs):
    if bpos is None or len(bpos) == 0:
        return (np.zeros((1, 4)), np.zeros(1), None)
    return (np.ascontiguousarray(bpos, np.float64), np.ascontiguousarray(bpsi, np.float64),
            np.ascontiguousarray(bcs, np.uint32))


def lambda_(mode, P, pred, cs, bpos=None, bpsi=None, bcs=None):
    pred = np.ascontiguousarray(pred, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    bp, bs, bc = _bnd(bpos, bpsi, bcs)
    n = len(pred)
    den = np.zeros(n)
    lam = np.zeros(n)
    lib().orc_lambda(mode, C.byref(P), n, _p(pred), _p(cs), _p(bp), _p(bs),
                     _p(bc) if bc is not None else None, _p(den), _p(lam))
    return den, lam


def delta(mode, P, pred, lam, cs, bpos=None, bpsi=None, bcs=None):
    pred = np.ascontiguousarray(pred, np.float64)
    lam = np.ascontiguousarray(lam, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    bp, bs, bc = _bnd(bpos, bpsi, bcs)
    out = np.zeros_like(pred)
    lib().orc_delta(mode, C.byref(P), len(pred), _p(pred), _p(lam), _p(cs), _p(bp), _p(bs),
                    _p(bc) if bc is not None else None, _p(out))
    return out


def delta_at(mode, P, pred, base, lam, cs, bpos=None, bpsi=None, bcs=None):
    """delta evaluated at `pred`, applied to `base`: clamp(base + correction)."""
    pred = np.ascontiguousarray(pred, np.float64)
    base = np.ascontiguousarray(base, np.float64)
    lam = np.ascontiguousarray(lam, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    bp, bs, bc = _bnd(bpos, bpsi, bcs)
    out = np.zeros_like(pred)
    lib().orc_delta_at(mode, C.byref(P), len(pred), _p(pred), _p(base), _p(lam), _p(cs), _p(bp),
                       _p(bs), _p(bc) if bc is not None else None, _p(out))
    return out


def cell_centres(P, cells):
    """Float centre boxMin + (x + 0.5) * h of flat cell indices."""
    gx, gy = P.gridDim[0], P.gridDim[1]
    c = np.asarray(cells, np.int64)
    xyz = np.stack([c % gx, (c // gx) % gy, c // (gx * gy)], axis=1).astype(np.float32)
    return np.array(P.boxMin[:], np.float32) + (xyz + np.float32(0.5)) * np.float32(P.h)


def fp16_quantize(P, pred, cells):
    """Positions of the FP16-offset form: x as float; y and z as
    centre(binned cell) + half(pred - centre), the offset formed in float
    and rounded to nearest even."""
    c32 = cell_centres(P, cells)[:, 1:]
    p32 = np.asarray(pred)[:, :3].astype(np.float32)
    o = (p32[:, 1:] - c32).astype(np.float16)
    q = np.array(pred, np.float64)
    q[:, 0] = p32[:, 0]
    q[:, 1:3] = c32.astype(np.float64) + o.astype(np.float64)
    return q


def field_lattice(box_min, box_max, resolution):
    """SurfaceMesh.configure (mesh.js:96-128): (vert_dim, voxel, origin)."""
    size = [b - a for a, b in zip(box_min, box_max)]
    voxel = max(size) / resolution
    vd = [int(math.ceil(s / voxel)) + 3 for s in size]
    return vd, voxel, [a - voxel for a in box_min]


def field_iso(P, lattice, mass, rho0, pos, density, body, skip_bodies, cs):
    """meshDensityWGSL (mesh_wgsl.js:23-69) in float; returns the flat field."""
    vd, voxel, origin = lattice
    pos = np.ascontiguousarray(pos, np.float64)
    density = np.ascontiguousarray(density, np.float64)
    body = np.ascontiguousarray(body, np.uint32)
    cs = np.ascontiguousarray(cs, np.uint32)
    out = np.zeros(vd[0] * vd[1] * vd[2], np.float32)
    vdc = np.array(vd, np.int32)
    org = np.array(origin, np.float64)
    lib().orc_field_iso(C.byref(P), _p(vdc), voxel, _p(org), mass, rho0, len(pos), _p(pos),
                        _p(density), _p(body), int(skip_bodies), _p(cs), _p(out))
    return out


def field_aniso(lattice, mass, rho0, smooth_pos, g, density, body, skip_bodies):
    """The anisotropic splat of smoothed centres and G (2 float4 per particle)."""
    vd, voxel, origin = lattice
    sp = np.ascontiguousarray(smooth_pos, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    density = np.ascontiguousarray(density, np.float64)
    body = np.ascontiguousarray(body, np.uint32)
    out = np.zeros(vd[0] * vd[1] * vd[2], np.float32)
    vdc = np.array(vd, np.int32)
    org = np.array(origin, np.float64)
    lib().orc_field_aniso(_p(vdc), voxel, _p(org), mass, rho0, len(sp), _p(sp), _p(g),
                          _p(density), _p(body), int(skip_bodies), _p(out))
    return out


def mesh_smooth(vd, passes, field):
    """`passes` rounds of meshSmoothWGSL's x, y, z steps (mesh_wgsl.js:70-91),
    as SurfaceMesh.smoothField / buildNormalField (mesh.js:273-313)."""
    src = np.ascontiguousarray(field, np.float32)
    out = np.zeros_like(src)
    lib().orc_mesh_smooth(_p(np.array(vd, np.int32)), int(passes), _p(src), _p(out))
    return out


def mesh_march(vd, iso, max_tris, field, nfield=None):
    """meshMarchWGSL (mesh_wgsl.js:190-233) with triangles numbered in voxel
    order: returns (kept triangles as uint32 (k, 6), total count)."""
    vdc = np.array(vd, np.int32)
    f = np.ascontiguousarray(field, np.float32)
    nf = f if nfield is None else np.ascontiguousarray(nfield, np.float32)
    total = lib().orc_mesh_march(_p(vdc), float(iso), 0, _p(f), _p(nf), None)
    keep = min(int(total), int(max_tris))
    verts = np.zeros((max(keep, 1), 6), np.uint32)
    lib().orc_mesh_march(_p(vdc), float(iso), keep, _p(f), _p(nf), _p(verts))
    return verts[:keep], int(total)


def body_pack(centre, rot, stat):
    """bodyPackWGSL (solid_wgsl.js:3-26): per body 6 float4 = (centre, shape
    id), (half extents, 0), the columns of R (rot holds 3 rows per body),
    (colour, 0)."""
    centre = np.asarray(centre, np.float32).reshape(-1, 4)
    rot = np.asarray(rot, np.float32).reshape(-1, 3, 4)
    stat = np.asarray(stat, np.float32).reshape(-1, 2, 4)
    nb = len(centre)
    out = np.zeros((nb, 6, 4), np.float32)
    out[:, 0, :3] = centre[:, :3]
    out[:, 0, 3] = stat[:, 0, 3]
    out[:, 1, :3] = stat[:, 0, :3]
    for k in range(3):
        out[:, 2 + k, :3] = rot[:, :, k]
    out[:, 5, :3] = stat[:, 1, :3]
    return out.reshape(nb * 6, 4)


def _len3(x, y, z):
    """WGSL length() as the kernels evaluate it: sqrt((x*x + y*y) + z*z) in
    float32, one rounding per operation."""
    return np.sqrt((x * x + y * y) + z * z)


def impulse(pos, vel, origin, direction, imp, radius, speed_limit):
    """impulseWGSL (wgsl.js:216-248) with Sim.applyRayImpulse's host side
    (sim.js:658-669): nothing happens for radius <= 0 or a zero impulse; the
    direction is divided by Math.hypot (|| 1) in double; every value then
    lands in a float32 uniform and the kernel works in float32.  Returns the
    new live velocities (float32, n x 4)."""
    f32 = np.float32
    v = np.array(vel, np.float32, copy=True).reshape(-1, 4)
    if not (len(v) and not (radius <= 0) and not (math.hypot(*imp) <= 0)):
        return v
    dl = math.hypot(*direction) or 1.0
    o = [f32(c) for c in origin]
    d = [f32(c / dl) for c in direction]
    J = [f32(c) for c in imp]
    R, L = f32(radius), f32(speed_limit)
    p = np.asarray(pos, np.float32).reshape(-1, 4)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        tx, ty, tz = p[:, 0] - o[0], p[:, 1] - o[1], p[:, 2] - o[2]
        along = (tx * d[0] + ty * d[1]) + tz * d[2]
        ax, ay, az = tx - along * d[0], ty - along * d[1], tz - along * d[2]
        d2 = (ax * ax + ay * ay) + az * az
        hit = ~(along <= 0) & ~(d2 >= R * R)
        x = f32(1.0) - np.sqrt(d2) / R
        old = _len3(v[:, 0], v[:, 1], v[:, 2])
        w = x * x
        n = [v[:, k] + w * J[k] for k in range(3)]
        ns = _len3(*n)
        allowed = np.maximum(old, L)
        over = ns > allowed
        sc = allowed / ns
        n = [np.where(over, c * sc, c) for c in n]
    for k in range(3):
        v[hit, k] = n[k][hit]
    v[hit, 3] = 0.0
    return v


def body_drag(vel, slot_of_members, centre, goal, rate, limit, dt):
    """bodyDragWGSL (wgsl.js:187-214) on the held body's particles (slots
    `slot_of_members`) with Sim.step's uniform (sim.js:444-455): blend =
    1 - exp(-max(dt, 0)/0.02) in double, stored as float32; the wanted
    velocity (goal - centre) * rate, scaled to `limit` when longer."""
    f32 = np.float32
    v = np.array(vel, np.float32, copy=True).reshape(-1, 4)
    blend = f32(1.0 - math.exp(-max(dt, 0.0) / 0.02))
    c = np.asarray(centre, np.float32).reshape(4)
    w = [(f32(goal[k]) - c[k]) * f32(rate) for k in range(3)]
    sp = _len3(*[np.float32(x) for x in w])
    if sp > f32(limit):
        s = f32(limit) / sp
        w = [x * s for x in w]
    m = np.asarray(slot_of_members, np.int64)
    for k in range(3):
        v[m, k] = v[m, k] + (f32(w[k]) - v[m, k]) * blend
    return v


def resize(pos, old_box, new_box, spacing):
    """resizeWGSL (wgsl.js:125-157) with Sim.resizeBox's uniform (sim.js:
    742-756): on an axis that shrinks, p = newMin + clamp((p - oldMin) /
    oldSize, 0, 1) * newSize; then every axis is clamped to [d/2, box -
    d/2] (d/2 and box - d/2 formed in double, stored as float32).  Returns
    the new live positions (float32, n x 4)."""
    f32 = np.float32
    p = np.array(pos, np.float32, copy=True).reshape(-1, 4)
    half = 0.5 * spacing
    for a in range(3):
        os_, ns_ = f32(old_box[a]), f32(new_box[a])
        x = p[:, a]
        if ns_ < os_:
            t = np.minimum(np.maximum((x - f32(0.0)) / os_, f32(0.0)), f32(1.0))
            x = f32(0.0) + t * ns_
        p[:, a] = np.minimum(np.maximum(x, f32(half)), f32(new_box[a] - half))
    return p


def ssfr_kernel_g(g):
    """The SSFR splat's reading of the packed G (ssfr_wgsl.js:21-28): the
    symmetric matrix of each particle's two float4."""
    g = np.asarray(g, np.float64).reshape(-1, 2, 4)
    a, b = g[:, 0], g[:, 1]
    return np.stack([np.stack([a[:, 0], a[:, 1], a[:, 2]], -1),
                     np.stack([a[:, 1], a[:, 3], b[:, 0]], -1),
                     np.stack([a[:, 2], b[:, 0], b[:, 1]], -1)], 1)


def mesh_decode(verts, lattice):
    """The mesh vertex shader's unpacking (mesh_wgsl.js:271-283), in FP64:
    world positions (m, 3) and unit normals (m, 3) of packed vertices."""
    vd, voxel, origin = lattice
    v = np.asarray(verts, np.uint32).reshape(-1, 2).astype(np.int64)
    q = np.stack([v[:, 0] & 0xFFFF, v[:, 0] >> 16, v[:, 1] & 0xFFFF], axis=1) / 65535.0
    pos = np.array(origin) + q * (np.array(vd) - 1) * voxel
    oct_ = v[:, 1] >> 16
    e = np.stack([oct_ & 0xFF, oct_ >> 8], axis=1) / 255.0 * 2.0 - 1.0
    n = np.stack([e[:, 0], e[:, 1], 1.0 - np.abs(e[:, 0]) - np.abs(e[:, 1])], axis=1)
    neg = n[:, 2] < 0
    sx = np.where(n[:, 0] >= 0, 1.0, -1.0)
    sy = np.where(n[:, 1] >= 0, 1.0, -1.0)
    tx = (1.0 - np.abs(n[:, 1])) * sx
    ty = (1.0 - np.abs(n[:, 0])) * sy
    n[neg, 0] = tx[neg]
    n[neg, 1] = ty[neg]
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return pos, n


def xsph(mode, P, pred, pos, density, cs, normals):
    pred = np.ascontiguousarray(pred, np.float64)
    pos = np.ascontiguousarray(pos, np.float64)
    density = np.ascontiguousarray(density, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    n = len(pred)
    vel = np.zeros((n, 4))
    corr = np.zeros((n, 4))
    normal = np.zeros((n, 4))
    lib().orc_xsph(mode, C.byref(P), n, _p(pred), _p(pos), _p(density), _p(cs),
                   1 if normals else 0, _p(vel), _p(corr), _p(normal))
    return vel, corr, normal


def tension(mode, P, pred, density, normal, cs, corr):
    pred = np.ascontiguousarray(pred, np.float64)
    density = np.ascontiguousarray(density, np.float64)
    normal = np.ascontiguousarray(normal, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    corr = np.array(corr, np.float64, copy=True)
    lib().orc_tension(mode, C.byref(P), len(pred), _p(pred), _p(density), _p(normal), _p(cs),
                      _p(corr))
    return corr


def finalize(mode, pred, corr, vel):
    pred = np.ascontiguousarray(pred, np.float64)
    corr = np.ascontiguousarray(corr, np.float64)
    vel = np.array(vel, np.float64, copy=True)
    pos = np.zeros_like(pred)
    lib().orc_finalize(mode, len(pred), _p(pred), _p(corr), _p(vel), _p(pos))
    return vel, pos


def census(P, pred, cs):
    pred = np.ascontiguousarray(pred, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    n = len(pred)
    a = np.zeros(n, np.uint32)
    b = np.zeros(n, np.uint32)
    c = np.zeros(n, np.uint32)
    lib().orc_census(C.byref(P), n, _p(pred), _p(cs), _p(a), _p(b), _p(c))
    return a, b, c


def body_pre(pred, body_of, centre):
    pred = np.ascontiguousarray(pred, np.float64)
    body_of = np.ascontiguousarray(body_of, np.uint32)
    c = np.array(centre, np.float64, copy=True).reshape(-1, 3)
    lib().orc_body_pre(len(pred), _p(pred), _p(body_of), len(c), _p(c))
    return c


def body_solve(pred, body_of, rest, inv_ratio, centre, rot):
    pred = np.array(pred, np.float64, copy=True)
    body_of = np.ascontiguousarray(body_of, np.uint32)
    rest = np.ascontiguousarray(rest, np.float64)
    inv = np.ascontiguousarray(inv_ratio, np.float64)
    c = np.array(centre, np.float64, copy=True).reshape(-1, 3)
    r = np.array(rot, np.float64, copy=True).reshape(-1, 9)
    lib().orc_body_solve(len(pred), _p(pred), _p(body_of), _p(rest), len(c), _p(inv), _p(c), _p(r))
    return pred, c, r


def stats(density, vel, inv_rho0, mass):
    density = np.ascontiguousarray(density, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    out = np.zeros(4)
    lib().orc_stats(len(density), _p(density), _p(vel), C.c_double(inv_rho0), C.c_double(mass),
                    _p(out))
    return out


def aniso(mode, P, pos, cs, radius_scale=2.0, lam=0.9, kr=4.0, kn=1.0, ks=1.0, stretch=2.0,
          min_extent=0.0, neps=25, sel=None):
    """Anisotropy pass (aniso_wgsl.js:3-159) with the parameters of
    SurfaceMesh.buildAnisotropy (mesh.js:221-242); with `sel`, for those
    targets only (outputs in the order of sel)."""
    pos = np.ascontiguousarray(pos, np.float64)
    cs = np.ascontiguousarray(cs, np.uint32)
    h = P.h
    radius = radius_scale * h
    ap = np.array([lam, h, radius, math.ceil(radius_scale), kr, kn, ks / (0.15 * radius * radius),
                   stretch, min_extent, neps], np.float64)
    n = len(pos)
    if sel is not None:
        sel = np.ascontiguousarray(sel, np.int32)
        assert sel.size == 0 or (sel.min() >= 0 and sel.max() < n)
        k = len(sel)
        smooth = np.zeros((k, 4))
        g = np.zeros((k, 8))
        lib().orc_aniso_sel(mode, C.byref(P), n, _p(pos), _p(cs), _p(ap), k, _p(sel),
                            _p(smooth), _p(g))
        return smooth, g.reshape(k * 2, 4)
    smooth = np.zeros((n, 4))
    g = np.zeros((n, 8))
    lib().orc_aniso(mode, C.byref(P), n, _p(pos), _p(cs), _p(ap), _p(smooth), _p(g))
    return smooth, g.reshape(n * 2, 4)


def jacobi(mode, a):
    a = np.ascontiguousarray(a, np.float64).reshape(9)
    w = np.zeros(3)
    v = np.zeros(9)
    lib().orc_jacobi(mode, _p(a), _p(w), _p(v))
    return w, v.reshape(3, 3)


def _setup(L):
    vp = C.c_void_p
    L.orc_aniso.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp]
    L.orc_aniso_sel.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp, vp]
    L.orc_jacobi.argtypes = [C.c_int, vp, vp, vp]
    L.orc_cell_ids.argtypes = [C.c_int, C.c_void_p, C.c_int, vp, vp]
    L.orc_grid.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    L.orc_predict.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp]
    L.orc_lambda.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    L.orc_delta.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    L.orc_delta_at.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_field_iso.argtypes = [vp, vp, C.c_double, vp, C.c_double, C.c_double, C.c_int, vp,
                                vp, vp, C.c_int, vp, vp]
    L.orc_field_aniso.argtypes = [vp, C.c_double, vp, C.c_double, C.c_double, C.c_int, vp, vp,
                                  vp, vp, C.c_int, vp]
    L.orc_mesh_smooth.argtypes = [vp, C.c_int, vp, vp]
    L.orc_mesh_march.argtypes = [vp, C.c_float, C.c_uint32, vp, vp, vp]
    L.orc_mesh_march.restype = C.c_uint32
    L.orc_xsph.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp]
    L.orc_tension.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp]
    L.orc_finalize.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    L.orc_census.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    L.orc_body_pre.argtypes = [C.c_int, vp, vp, C.c_int, vp]
    L.orc_body_solve.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp]
    L.orc_stats.argtypes = [C.c_int, vp, vp, C.c_double, C.c_double, vp]


class State:
    """Particle state in the reference's live order (float4 records)."""

    def __init__(self, pos, vel, body_of, rest):
        self.pos = np.array(pos, np.float64)
        self.vel = np.array(vel, np.float64)
        self.body_of = np.array(body_of, np.uint32)
        self.rest = np.array(rest, np.float64)


def substep(mode, P, st, bnd, centre, rot, inv_ratio, iterations, tension_on, fp16=False):
    """One substep of Sim.step (sim.js:458-596) on State `st` (modified in
    place).  bnd = (bpos, bpsi, bcs) from sort_boundary.  Returns a dict of
    intermediate arrays (post-sort order) for comparisons.

    fp16=True restates the FP16-offset form of the solver iterations (not a
    reference feature): lambda and delta are evaluated at the quantized
    positions of fp16_quantize (binned cell = the grid build's cell id), and
    delta moves the unquantized positions."""
    pred, _ = predict(mode, P, st.pos, st.vel)
    key = cell_ids(mode, P, pred)
    cs, slot, order = grid(P, key)
    pos = st.pos[order]
    pred = pred[order]
    body_of = st.body_of[order]
    rest = st.rest[order]
    bpos, bpsi, bcs = bnd
    if len(centre):
        centre = body_pre(pred, body_of, centre)
    lam_last = den_last = None
    cells = key[order]
    for _ in range(max(1, iterations)):
        if fp16:
            q = fp16_quantize(P, pred, cells)
            den_last, lam_last = lambda_(mode, P, q, cs, bpos, bpsi, bcs)
            pred = delta_at(mode, P, q, pred, lam_last, cs, bpos, bpsi, bcs)
        else:
            den_last, lam_last = lambda_(mode, P, pred, cs, bpos, bpsi, bcs)
            pred = delta(mode, P, pred, lam_last, cs, bpos, bpsi, bcs)
        if len(centre):
            pred, centre, rot = body_solve(pred, body_of, rest, inv_ratio, centre, rot)
    vel, corr, normal = xsph(mode, P, pred, pos, den_last, cs, tension_on)
    if tension_on:
        corr = tension(mode, P, pred, den_last, normal, cs, corr)
    vel, pos = finalize(mode, pred, corr, vel)
    st.pos, st.vel, st.body_of, st.rest = pos, vel, body_of, rest
    return dict(key=key, cell_start=cs, slot=slot, order=order, density=den_last,
                lam=lam_last, corr=corr, normal=normal, centre=centre, rot=rot)

