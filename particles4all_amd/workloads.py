"""Synthetic workloads in lattice units (spacing d = 1, h = 2d, rho0 = 1000).

  dam_break_32     32^3 jittered block in a 64^3 box, two floating 8^3 cubes
                   of density 0.5; frame 1/60 s, 2 substeps, 4 iterations.
  block_1m         1,048,576 particles: a 128x64x128 block (x, y, z) at the
                   corner of a 256x128x128 box, half its width and half its
                   height; substep 1/120 s, 4 iterations, XSPH 0.01, surface
                   tension 0.5.
  block_4m_bodies  256x128x128 block in a 512x192x130 box plus 64 bodies
                   (32 lattice cubes, edges 6-12; 32 lattice spheres, radii
                   4-8; densities log-uniform in [0.3, 2.0]) in the space
                   above the fluid; settings as block_1m.

Jitter is uniform +-0.05d per coordinate from the reference's LCG idiom
(scene.js:52-56); mass and denomRest are calibrated on the unjittered
lattice.  `scale` shrinks the lattice for tests (box shrinks with it).

The lattice sites start one spacing inside the box (sites margin + (i +
0.5) d), as the reference's fluidBlock places them (margin = d,
scene.js:45, 61-63), so the fluid's outer layer sits 1.5 d from the wall
samples and the block starts near rest density; block_4m_bodies' box is
two spacings deeper than its lattice for that.  block_1m keeps its block
against the walls (its 128-deep block spans the 128-deep box).
"""
import numpy as np

from ._lib import (SHAPE_LATTICE_CUBE, SHAPE_LATTICE_SPHERE, default_params, lattice_scene)


def _params(box, substeps, iterations, xsph, tension, gravity=9.8, pour_budget=0, **kw):
    p = default_params(spacing=1.0, rest_density=1000.0, gravity=gravity, substeps=substeps,
                       iterations=iterations, xsph_c=xsph, surface_tension_k=tension,
                       box=box, pour_budget=pour_budget, **kw)
    return p


def dam_break_32(scale=1.0, seed=1, bodies=True, margin=1.0, **kw):
    L = max(4, int(round(32 * scale)))
    B = 2 * L
    p = _params([B, B, B], substeps=2, iterations=4, xsph=0.05, tension=0.0, **kw)
    bl = []
    if bodies:
        e = max(2, int(round(8 * scale)))
        y = 1.5 * L
        bl = [(SHAPE_LATTICE_CUBE, e, 0.5, (0.5 * L, y, 0.5 * L)),
              (SHAPE_LATTICE_CUBE, e, 0.5, (1.5 * L, y, 0.5 * L))]
    sc = lattice_scene(p, (L, L, L), origin=(margin,) * 3, jitter=0.05, seed=seed, bodies=bl)
    return p, sc, 1.0 / 60.0


def dam_break_2m(scale=1.0, seed=5, margin=1.0, **kw):
    """The dam break recipe scaled up for the anisotropy and field pass:
    a jittered 128^3 lattice (2,097,152 particles) at the corner of a 256^3
    box, substeps 2, I = 4, XSPH 0.05, no tension, no bodies (the scaled
    recipe is fluid only so the particle count is exactly 128^3)."""
    L = max(4, int(round(128 * scale)))
    B = 2 * L
    p = _params([B, B, B], substeps=2, iterations=4, xsph=0.05, tension=0.0, **kw)
    sc = lattice_scene(p, (L, L, L), origin=(margin,) * 3, jitter=0.05, seed=seed)
    return p, sc, 1.0 / 60.0


def block_1m(scale=1.0, seed=2, **kw):
    L = max(4, int(round(128 * scale)))
    p = _params([2 * L, L, L], substeps=2, iterations=4, xsph=0.01, tension=0.5, **kw)
    sc = lattice_scene(p, (L, max(2, L // 2), L), jitter=0.05, seed=seed)
    return p, sc, 1.0 / 120.0


def block_4m_bodies(scale=1.0, seed=3, n_bodies=64, margin=1.0, **kw):
    nx, ny, nz = (max(4, int(round(v * scale))) for v in (256, 128, 128))
    box = [2 * nx, ny * 1.5, nz + 2 * margin]
    p = _params(box, substeps=2, iterations=4, xsph=0.01, tension=0.5, **kw)
    rng = np.random.default_rng(seed)
    bl = []
    cols = max(1, n_bodies // 4)
    pitch_x, pitch_z = box[0] / cols, box[2] / 4
    y = ny + 0.25 * ny
    s = max(scale, 0.05)
    for k in range(n_bodies):
        cx = (k % cols + 0.5) * pitch_x
        cz = (k // cols % 4 + 0.5) * pitch_z
        dens = float(np.exp(rng.uniform(np.log(0.3), np.log(2.0))))
        if k % 2 == 0:
            edge = int(rng.integers(6, 13))
            bl.append((SHAPE_LATTICE_CUBE, max(2, round(edge * s)), dens, (cx, y, cz)))
        else:
            rad = int(rng.integers(4, 9))
            bl.append((SHAPE_LATTICE_SPHERE, max(1, round(rad * s)), dens, (cx, y, cz)))
    sc = lattice_scene(p, (nx, ny, nz), origin=(margin,) * 3, jitter=0.05, seed=seed,
                       bodies=bl)
    return p, sc, 1.0 / 120.0


def batch_scenes(n_scenes=256, seed=4, n_min=2000, n_max=200000, math=0, tension=0.0, **kw):
    """Many independent scenes of uneven size: particle counts log-uniform
    in [n_min, n_max], box extents U[32, 128] d per axis (each axis at least
    the block plus 8 d of headroom), 0-4 lattice cubes of 4-8 particle
    edges above the block, densities log-uniform in [0.3, 2]; I = 4, FP32,
    XSPH 0.05, surface tension `tension` (none by default), substep 1/120 s.
    Returns [(params, scene)]."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        n = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        side = n ** (1.0 / 3.0)
        bx = int(np.clip(round(side * rng.uniform(0.7, 1.4)), 4, 110))
        bz = int(np.clip(round(side * rng.uniform(0.7, 1.4)), 4, 110))
        by = int(np.clip(round(n / (bx * bz)), 4, 110))
        box = [float(max(rng.integers(32, 129), b + 8)) for b in (bx, by, bz)]
        box = [min(128.0, v) for v in box]
        bx, by, bz = (int(min(b, v - 8)) for b, v in zip((bx, by, bz), box))
        p = _params(box, substeps=2, iterations=4, xsph=0.05, tension=tension, math=math, **kw)
        bl = []
        for _ in range(int(rng.integers(0, 5))):
            e = int(rng.integers(4, 9))
            y = min(box[1] - e, by + 0.5 * (box[1] - by))
            cx = rng.uniform(e, box[0] - e)
            cz = rng.uniform(e, box[2] - e)
            dens = float(np.exp(rng.uniform(np.log(0.3), np.log(2.0))))
            bl.append((SHAPE_LATTICE_CUBE, e, dens, (cx, y, cz)))
        sc = lattice_scene(p, (bx, by, bz), jitter=0.05, seed=int(rng.integers(1, 1 << 30)),
                           bodies=bl)
        out.append((p, sc))
    return out


WORKLOADS = {"dam_break_32": dam_break_32, "block_1m": block_1m,
             "block_4m_bodies": block_4m_bodies, "dam_break_2m": dam_break_2m}
