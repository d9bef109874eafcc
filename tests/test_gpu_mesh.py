"""GPU surface extraction against the oracle, through the C ABI.

Bars (all bit-exact, on fields generated by the GPU from a moved dam break):
the in-place field blur (smoothField) and the normal field
(buildNormalField) equal the oracle's float restatement; the marching
tetrahedra vertex buffer (packed words, voxel-ordered triangles), the total
count and the draw-indirect arguments equal the oracle's serial march, with
and without a normal field, on the isotropic and the anisotropic field; a
triangle budget keeps the leading triangles and the total; repeated marches
give the same buffer.
"""
import numpy as np
import pytest

from particles4all_amd import Sim, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def moved_dam_break(substeps=6):
    p, sc, _ = workloads.dam_break_32(scale=1.0, math=0)
    sim = Sim(0)
    sim.reset(p, sc)
    sim.substeps(1 / 120, substeps)
    return sim, p, sc


def vert_dim(sim):
    return list(sim.field_info().vert_dim)


@pytest.mark.parametrize("aniso,smooth,normal", [(False, 0, 0), (False, 1, 2), (True, 0, 2),
                                                 (True, 2, 1)])
def test_mesh_pipeline_bit_exact(aniso, smooth, normal):
    sim, p, sc = moved_dam_break()
    if aniso:
        sim.aniso()
    sim.field_configure([0, 0, 0], list(p.box), 56)
    vd = vert_dim(sim)
    sim.field_generate(anisotropic=aniso)
    f0 = sim.read("FIELD")
    sim.field_smooth(smooth)
    f1 = sim.read("FIELD")
    want1 = O.mesh_smooth(vd, smooth, f0) if smooth else f0
    assert np.array_equal(f1.view(np.uint32), want1.view(np.uint32))
    sim.mesh_normal_field(normal)
    nf = None
    if normal:
        nf = sim.read("FIELD_NORMAL")
        want_n = O.mesh_smooth(vd, normal, f1)
        assert np.array_equal(nf.view(np.uint32), want_n.view(np.uint32))
    iso = 0.4
    sim.mesh_march(iso)
    verts, total = sim.mesh()
    want, want_total = O.mesh_march(vd, iso, sim.field_info().max_triangles, f1, nf)
    print(f"aniso={aniso} smooth={smooth} normal={normal}: {total} triangles over "
          f"{np.prod(np.array(vd) - 1)} voxels")
    assert total == want_total > 1000
    assert verts.shape == want.shape
    bad = np.flatnonzero(np.any(verts != want, axis=1))
    if bad.size:
        print(f"{bad.size} triangles differ; first {bad[:6]}: gpu {verts[bad[0]]} "
              f"oracle {want[bad[0]]}")
    assert bad.size == 0
    ind = sim.read("MESH_INDIRECT")
    assert list(ind) == [3 * total, 1, 0, 0]
    # a second march of the same field gives the same buffer
    sim.mesh_march(iso)
    again, total2 = sim.mesh()
    assert total2 == total and np.array_equal(again, verts)
    sim.close()


def test_mesh_budget_and_reconfigure():
    sim, p, sc = moved_dam_break(substeps=3)
    box = list(p.box)
    sim.field_configure([0, 0, 0], box, 40)
    sim.field_generate()
    sim.mesh_normal_field(2)
    sim.mesh_march(0.4)
    full, total = sim.mesh()
    assert total == len(full) > 100
    cap = total // 3
    sim.field_configure([0, 0, 0], box, 40, max_triangles=cap)
    assert sim.field_info().max_triangles == cap
    # configure does not touch the field; the normal field is rebuilt
    sim.mesh_normal_field(2)
    sim.mesh_march(0.4)
    cut, total2 = sim.mesh()
    assert total2 == total and len(cut) == cap
    assert np.array_equal(cut, full[:cap])
    assert list(sim.read("MESH_INDIRECT")) == [3 * cap, 1, 0, 0]
    # an iso above the field's maximum gives an empty surface
    sim.mesh_march(float(sim.read("FIELD").max()) + 1.0)
    none, total3 = sim.mesh()
    assert total3 == 0 and len(none) == 0
    assert list(sim.read("MESH_INDIRECT")) == [0, 1, 0, 0]
    sim.close()


def test_mesh_vertices_on_the_surface():
    """Decoded vertices sit inside the lattice box, and the field linearly
    interpolated at each vertex along its lattice or diagonal edge is the iso
    value (checked through the oracle's decoder at the quantisation bound)."""
    sim, p, sc = moved_dam_break(substeps=3)
    sim.field_configure([0, 0, 0], list(p.box), 32)
    fi = sim.field_info()
    vd = list(fi.vert_dim)
    lat = (vd, fi.voxel, list(fi.origin))
    sim.field_generate()
    sim.mesh_march(0.4)
    verts, total = sim.mesh()
    pos, nrm = O.mesh_decode(verts, lat)
    lo = np.array(fi.origin)
    hi = lo + (np.array(vd) - 1) * fi.voxel
    assert np.all(pos >= lo - 1e-9) and np.all(pos <= hi + 1e-9)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)
    # lattice coordinates: integral axes, and the fractional ones of the
    # crossed edge (lattice edges and diagonals all run from the lower corner
    # up, so the far corner is the lower one plus 1 on each fractional axis)
    g = (pos - lo) / fi.voxel
    rg = np.round(g)
    frac = np.abs(g - rg) > 1e-3
    base = np.where(frac, np.floor(g), rg).astype(np.int64)
    t = np.where(frac, g - np.floor(g), 0).max(axis=1)
    far = base + frac
    on_edge = frac.any(axis=1)  # vertices at a lattice vertex (t ~ 0) are skipped
    assert on_edge.mean() > 0.99
    f = sim.read("FIELD").reshape(vd[2], vd[1], vd[0]).astype(np.float64)
    a = f[base[:, 2], base[:, 1], base[:, 0]]
    b = f[far[:, 2], far[:, 1], far[:, 0]]
    val = a + (b - a) * t
    q = (np.array(vd) - 1).max() / 65535.0
    tol = np.abs(b - a) * 2 * q + 1e-6
    assert np.all(np.abs(val - 0.4)[on_edge] <= tol[on_edge])
    sim.close()
