"""ORACLE - test infrastructure only.  Pure-Python restatement of the
reference's CPU scene builder (src/scene.js) for small scenes, used to pin
the library's native scene builder:
  fluidBlock          scene.js:43-64
  boundaryParticles   scene.js:76-135
  sampleBody          scene.js:137-157
  buildScene          scene.js:159-315
Python floats are IEEE doubles, the arithmetic JavaScript uses; the LCG
reproduces JavaScript's double product and ToInt32 masking.
"""
import math

import numpy as np


def js_lcg(seed):
    """(seed * 1103515245 + 12345) & 0x7fffffff evaluated as JavaScript does."""
    v = float(seed) * 1103515245.0
    v = v + 12345.0
    return int(v) % (1 << 32) & 0x7FFFFFFF


def fluid_block(box, d, target, seed=12345, amp=0.002):
    bx, by, bz = box
    margin = d
    nx = max(1, math.floor(bx * 0.35 / d))
    nz = max(1, math.floor((bz - 2 * margin) / d))
    max_layers = max(1, math.floor((by - 2 * margin) / d))
    layers = min(math.ceil(target / (nx * nz)), max_layers)
    out = []
    s = seed

    def rnd():
        nonlocal s
        s = js_lcg(s)
        return (s / 0x7FFFFFFF - 0.5) * amp * d

    for iy in range(layers):
        if len(out) >= target:
            break
        for ix in range(nx):
            if len(out) >= target:
                break
            for iz in range(nz):
                if len(out) >= target:
                    break
                x = margin + (ix + 0.5) * d + rnd()
                y = margin + (iy + 0.5) * d + rnd()
                z = margin + (iz + 0.5) * d + rnd()
                out.append((x, y, z))
    return out


def boundary_particles(size, d, h, rho0):
    nb = [max(2, math.ceil(s / d) + 1) for s in size]
    step = [s / (nb[k] - 1) for k, s in enumerate(size)]
    pts = []
    for ix in range(nb[0]):
        for iy in range(nb[1]):
            for iz in range(nb[2]):
                shell = (ix == 0 or ix == nb[0] - 1 or iy == 0 or iy == nb[1] - 1 or
                         iz == 0 or iz == nb[2] - 1)
                if shell:
                    pts.append((ix * step[0], iy * step[1], iz * step[2]))
    dim = [max(1, math.ceil(s / h)) for s in size]

    def cell_of(p):
        return tuple(min(dim[k] - 1, max(0, math.floor(p[k] / h))) for k in range(3))

    buckets = {}
    for i, p in enumerate(pts):
        buckets.setdefault(cell_of(p), []).append(i)
    coef = 315 / (64 * math.pi * math.pow(h, 9))
    h2 = h * h
    psi = []
    for i, p in enumerate(pts):
        a, b, c = cell_of(p)
        s = 0.0
        for dc in (-1, 0, 1):
            if c + dc < 0 or c + dc >= dim[2]:
                continue
            for db in (-1, 0, 1):
                if b + db < 0 or b + db >= dim[1]:
                    continue
                for da in (-1, 0, 1):
                    if a + da < 0 or a + da >= dim[0]:
                        continue
                    for j in buckets.get((a + da, b + db, c + dc), ()):
                        q = pts[j]
                        dx, dy, dz = p[0] - q[0], p[1] - q[1], p[2] - q[2]
                        r2 = dx * dx + dy * dy + dz * dz
                        if r2 >= h2:
                            continue
                        t = h2 - r2
                        s += t * t * t
        s *= coef
        psi.append(rho0 / s if s > 0 else 0.0)
    return pts, psi


def sample_body(shape, size, d):
    if shape == "sphere":
        ext = [size, size, size]
    elif shape == "torus":
        ext = [size + 0.4 * size, 0.4 * size, size + 0.4 * size]
    else:
        ext = [size * 0.9, size * 0.7, size * 0.9]
    steps = [max(1, math.ceil(e / d)) for e in ext]
    local = []
    for ix in range(-steps[0], steps[0] + 1):
        for iy in range(-steps[1], steps[1] + 1):
            for iz in range(-steps[2], steps[2] + 1):
                x, y, z = ix * d, iy * d, iz * d
                if shape == "sphere":
                    inside = x * x + y * y + z * z <= size * size
                elif shape == "torus":
                    planar = math.sqrt(x * x + z * z) - size
                    inside = planar * planar + y * y <= (0.4 * size) * (0.4 * size)
                else:
                    inside = (abs(x) <= size * 0.9 and abs(y) <= size * 0.7 and
                              abs(z) <= size * 0.9)
                if inside:
                    local.append((x, y, z))
    return local


def build_scene(box, d, rho0, target, bodies, body_size):
    """buildScene for `bodies` = list of (shape, density, startY)."""
    h = 2 * d
    bx, by, bz = box
    fluid = fluid_block(box, d, target)
    n_fluid = len(fluid)
    sz = body_size if body_size > 0 else 0.12 * min(bx, by, bz)
    step = min(2.5 * sz, bx * 0.25)
    span = step * (len(bodies) - 1)
    body_pts, body_rest, body_phase, meta = [], [], [], []
    for r, (shape, dens, start_y) in enumerate(bodies):
        local = sample_body(shape, sz, d)
        count = len(local)
        if not count:
            continue
        centre = [0.5 * bx - 0.5 * span + step * r, by * start_y, bz * 0.5]
        cx = sum(p[0] for p in local) / count
        cy = sum(p[1] for p in local) / count
        cz = sum(p[2] for p in local) / count
        bid = len(meta) + 1
        for (lx, ly, lz) in local:
            body_pts.append((centre[0] + lx, centre[1] + ly, centre[2] + lz))
            body_rest.append((lx - cx, ly - cy, lz - cz))
            body_phase.append(bid)
        meta.append(dict(centre=centre, density=dens, count=count))
    n = n_fluid + len(body_pts)
    pos = np.zeros((n, 4), np.float32)
    rest = np.zeros((n, 4), np.float32)
    phase = np.zeros((n, 4), np.uint32)
    pos[:n_fluid, :3] = np.array(fluid, np.float64).reshape(-1, 3)
    if body_pts:
        pos[n_fluid:, :3] = np.array(body_pts)
        rest[n_fluid:, :3] = np.array(body_rest)
        phase[n_fluid:, 0] = body_phase
    bpts, psi = boundary_particles(box, d, h, rho0)
    # mass calibration (scene.js:232-288)
    coef = 315 / (64 * math.pi * math.pow(h, 9))
    h2 = h * h
    spiky = -45 / (math.pi * math.pow(h, 6))
    dim = [max(1, math.floor(b / h)) for b in box]
    P = pos[:, :3].astype(np.float64)

    def cell_of(i):
        return tuple(min(dim[k] - 1, max(0, math.floor(P[i, k] / h))) for k in range(3))

    buckets = {}
    for i in range(n_fluid):
        buckets.setdefault(cell_of(i), []).append(i)

    def neighbours(i):
        a, b, c = cell_of(i)
        for dc in (-1, 0, 1):
            if c + dc < 0 or c + dc >= dim[2]:
                continue
            for db in (-1, 0, 1):
                if b + db < 0 or b + db >= dim[1]:
                    continue
                for da in (-1, 0, 1):
                    if a + da < 0 or a + da >= dim[0]:
                        continue
                    for j in buckets.get((a + da, b + db, c + dc), ()):
                        if j != i:
                            yield j

    def density_at(i, m):
        rho = m * coef * h2 * h2 * h2
        for j in neighbours(i):
            r2 = ((P[i, 0] - P[j, 0]) ** 2 + (P[i, 1] - P[j, 1]) ** 2 +
                  (P[i, 2] - P[j, 2]) ** 2)
            if r2 >= h2:
                continue
            t = h2 - r2
            rho += m * coef * t * t * t
        return rho

    mass = rho0 * d * d * d
    max_rho = max((density_at(i, mass) for i in range(n_fluid)), default=0.0)
    if max_rho > 0.5 * rho0:
        mass *= rho0 / max_rho
    denom = 0.0
    volume = mass / rho0
    for i in range(n_fluid):
        if density_at(i, mass) < 0.99 * rho0:
            continue
        gx = gy = gz = sg2 = 0.0
        for j in neighbours(i):
            rx, ry, rz = P[i, 0] - P[j, 0], P[i, 1] - P[j, 1], P[i, 2] - P[j, 2]
            r2 = rx * rx + ry * ry + rz * rz
            if r2 < 1e-12 or r2 >= h2:
                continue
            r = math.sqrt(r2)
            hr = h - r
            s = volume * spiky * hr * hr / r
            gx += s * rx
            gy += s * ry
            gz += s * rz
            sg2 += s * s * r2
        denom = max(denom, gx * gx + gy * gy + gz * gz + sg2)
        if i > 2000 and denom > 0:
            break
    return dict(n=n, n_fluid=n_fluid, pos=pos, rest=rest, phase=phase,
                boundary_pts=np.array(bpts, np.float32), boundary_psi=np.array(psi, np.float32),
                bodies=meta, mass=mass, denom_rest=denom)
