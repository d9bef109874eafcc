"""Pins of the oracle's surface extraction (CPU): the field blur of
SurfaceMesh.smoothField / buildNormalField and the marching tetrahedra of
meshMarchWGSL, with the vertex unpacking of the mesh vertex shader.

* blur: constants are kept; an impulse spreads into the separable product
  of (1/4, 1/2, 1/4) (exact in float), with the clamped lattice face
  folding the outer weight back (3/4 at a corner vertex);
* one voxel: a single inside corner gives 6 triangles at the diagonal
  corners 0 and 7 and 2 elsewhere (the six tetrahedra share the 0-7
  diagonal; every other corner belongs to two of them), the complement the
  same, corners 0 and 7 inside give 2 per tetrahedron;
* a plane x = c cuts every voxel of one column into 8 triangles whose
  area sums to the cross-section and whose vertices all sit at x = c;
* a sphere: every vertex within the chord error of the sphere, decoded
  normals radial, and the mesh a closed 2-manifold of Euler
  characteristic 2 (each edge shared by exactly two triangles);
* the triangle budget keeps a prefix and leaves the total unchanged.
"""
import math

import numpy as np

from oracle import oracle as O


def lattice_of(field):
    vz, vy, vx = field.shape
    return [vx, vy, vz], 1.0, [0.0, 0.0, 0.0]


def march(field, iso, max_tris=1 << 30, nfield=None):
    vd = lattice_of(field)[0]
    nf = None if nfield is None else nfield.reshape(-1)
    return O.mesh_march(vd, iso, max_tris, field.reshape(-1), nf)


def tri_area(p):
    a, b, c = p[0::3], p[1::3], p[2::3]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)


def test_blur_constant_and_impulse():
    vd = [7, 6, 5]
    const = np.full(vd[0] * vd[1] * vd[2], 0.37, np.float32)
    assert np.array_equal(O.mesh_smooth(vd, 2, const), const)
    f = np.zeros((vd[2], vd[1], vd[0]), np.float32)
    f[2, 3, 3] = 1.0
    out = O.mesh_smooth(vd, 1, f.reshape(-1)).reshape(f.shape)
    w = {-1: 0.25, 0: 0.5, 1: 0.25}
    want = np.zeros_like(f)
    for dz in w:
        for dy in w:
            for dx in w:
                want[2 + dz, 3 + dy, 3 + dx] = w[dx] * w[dy] * w[dz]
    assert np.array_equal(out, want)
    # the clamped face folds the outer weight back onto the corner vertex
    g = np.zeros_like(f)
    g[0, 0, 0] = 1.0
    out = O.mesh_smooth(vd, 1, g.reshape(-1)).reshape(f.shape)
    assert out[0, 0, 0] == np.float32(0.75 ** 3)
    assert out[0, 0, 1] == np.float32(0.25 * 0.75 * 0.75)
    assert out[1, 1, 1] == np.float32(0.25 ** 3)
    assert out[0, 0, 2] == 0.0
    # passes compose
    two = O.mesh_smooth(vd, 2, f.reshape(-1))
    assert np.array_equal(two, O.mesh_smooth(vd, 1, O.mesh_smooth(vd, 1, f.reshape(-1))))


CORNERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


def one_voxel(inside):
    f = np.zeros((2, 2, 2), np.float32)
    for k in inside:
        x, y, z = CORNERS[k]
        f[z, y, x] = 1.0
    return f


def test_single_voxel_counts():
    for k in range(8):
        want = 6 if k in (0, 7) else 2
        assert march(one_voxel([k]), 0.5)[1] == want, k
        assert march(one_voxel([j for j in range(8) if j != k]), 0.5)[1] == want, k
    assert march(one_voxel([0, 7]), 0.5)[1] == 12
    assert march(one_voxel([]), 0.5)[1] == 0
    assert march(one_voxel(range(8)), 0.5)[1] == 0
    # d > iso is inside: a corner equal to iso is outside
    assert march(one_voxel([3]), 1.0)[1] == 0


def test_plane_cut_area_and_position():
    vx, vy, vz = 12, 9, 7
    x = np.arange(vx, dtype=np.float32)
    c = 5.3
    f = np.broadcast_to((x - c)[None, None, :], (vz, vy, vx)).astype(np.float32)
    verts, total = march(f, 0.0)
    assert total == 8 * (vy - 1) * (vz - 1)
    pos, nrm = O.mesh_decode(verts, lattice_of(f))
    q = (vx - 1) / 65535.0
    assert np.all(np.abs(pos[:, 0] - c) <= q)
    assert abs(tri_area(pos).sum() - (vy - 1) * (vz - 1)) < 1e-3
    # normal = -gradient of a field growing along +x: -x
    assert np.all(nrm[:, 0] < -0.99)


def sphere_field(n, centre, radius):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    r = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2)
    return (radius - r).astype(np.float32)


def test_sphere_closed_manifold():
    # centre and radius chosen so that no lattice vertex lies within 0.02 of
    # the sphere: every mesh vertex then sits >= 2% of an edge away from a
    # lattice vertex and is identified by its lattice edge without ambiguity
    n, centre, radius = 32, (15.49, 15.52, 15.48), 9.04
    f = sphere_field(n, centre, radius)
    assert np.abs(f).min() > 0.02
    verts, total = march(f, 0.0)
    assert len(verts) == total > 1000
    pos, nrm = O.mesh_decode(verts, lattice_of(f))
    d = pos - np.array(centre)
    rad = np.linalg.norm(d, axis=1)
    # a vertex is the linear root of the distance along a tetrahedron edge
    # (length <= sqrt 3); the distance is convex along a line with second
    # derivative <= 1/(R - 2) near the sphere, so the vertex lies inside it
    # by at most 3 / (8 (R - 2)), plus the 16-bit quantization
    q = (n - 1) / 65535.0
    assert np.all(rad - radius < 2 * q)
    assert np.all(radius - rad < 3.0 / (8 * (radius - 2)) + 2 * q)
    assert np.all(np.sum(nrm * d / rad[:, None], axis=1) > 0.99)
    area = tri_area(pos).sum()
    assert abs(area / (4 * math.pi * radius ** 2) - 1) < 0.01
    # vertex -> tetrahedron edge: the edges are lattice edges, face diagonals
    # and cube diagonals, all running from a lower to a higher corner, so an
    # edge is (lower corner, set of fractional axes)
    fr = np.abs(pos - np.round(pos))
    frac = fr > 1e-3
    assert np.all(fr[~frac] < 2 * q) and np.all(fr[frac] > 5e-3)
    corner = np.where(frac, np.floor(pos), np.round(pos)).astype(np.int64)
    mask = frac[:, 0] * 1 + frac[:, 1] * 2 + frac[:, 2] * 4
    assert np.all(mask > 0)
    key = ((corner[:, 0] * n + corner[:, 1]) * n + corner[:, 2]) * 8 + mask
    tris = key.reshape(-1, 3)
    assert np.all(tris[:, 0] != tris[:, 1]) and np.all(tris[:, 1] != tris[:, 2])
    edges = np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    V, E, F = len(np.unique(key)), len(counts), len(tris)
    assert V - E + F == 2


def test_normal_field_and_budget():
    n = 20
    f = sphere_field(n, (9.6, 10.2, 9.9), 6.3)
    nf = O.mesh_smooth([n, n, n], 2, f.reshape(-1)).reshape(f.shape)
    full, total = march(f, 0.0, nfield=nf)
    plain, total2 = march(f, 0.0)
    assert total == total2 == len(full)
    # same positions, different normals when the gradient comes from nfield
    assert np.array_equal(full.reshape(-1, 2)[:, 0], plain.reshape(-1, 2)[:, 0])
    assert np.any(full.reshape(-1, 2)[:, 1] != plain.reshape(-1, 2)[:, 1])
    cut, total3 = march(f, 0.0, max_tris=total // 3, nfield=nf)
    assert total3 == total and np.array_equal(cut, full[:total // 3])
    none, total4 = march(f, 0.0, max_tris=0, nfield=nf)
    assert total4 == total and len(none) == 0
