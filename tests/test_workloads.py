"""Synthetic workload builders: sizes, grid lines and boundary sets."""
import numpy as np

from emerge_amd import bc as pbc
from emerge_amd import workloads as w
from emerge_amd.mesh import Mesh3D, kuhn_box, kuhn_grid


def test_kuhn_grid_matches_kuhn_box():
    n1, t1 = kuhn_box(3, 2, 4, size=(0.3, 0.2, 0.4))
    n2, t2 = kuhn_grid(np.linspace(0, 0.3, 4), np.linspace(0, 0.2, 3), np.linspace(0, 0.4, 5))
    assert np.allclose(n1, n2) and np.array_equal(t1, t2)


def test_ndof_estimate_is_exact():
    for dims in [(11, 5, 18), (3, 2, 4), (1, 1, 1)]:
        m = Mesh3D(*kuhn_box(*dims))
        assert 2 * m.n_edges + 2 * m.n_tris == w._ndof_estimate(*dims)


def test_cavity_small_version():
    p = w.cavity(cells=(12, 8, 6))
    bcs = w.cavity_bcs(p)
    pec, port = bcs[0], bcs[1]
    assert isinstance(port, pbc.LumpedPort) and port.tri_ids.size == 4
    # probe strip is internal PEC next to the port, outer walls are PEC
    assert np.intersect1d(pec.tri_ids, port.tri_ids).size == 0
    assert pec.tri_ids.size > p.mesh.boundary_tris().size
    assert np.isclose(np.abs(p.eps_r).max(), abs(9.8 * (1 - 1e-4j)))


def test_patch_small_version():
    p = w.patch_antenna(h_slab=5e-3, n_sub=2, h_air=10e-3, inset=5e-3)
    pec, abc, port = w.patch_bcs(p)
    m = p.mesh
    # ground + patch sheets, ABC on the whole outer box, port from ground to patch
    z = m.nodes[2][m.tris[:, pec.tri_ids]]
    assert set(np.unique(np.round(z, 6))) == {0.0, round(1.52e-3, 6)}
    assert abc.tri_ids.size == m.boundary_tris().size and abc.order == 1
    pz = m.nodes[2][m.tris[:, port.tri_ids]]
    assert pz.min() == 0.0 and np.isclose(pz.max(), 1.52e-3)
    assert np.isclose(port.height, 1.52e-3)


def test_waveguide_family_sizes():
    fam = w.waveguide_family(count=4, n_range=(20_000, 40_000), seed=2)
    for p in fam:
        assert p.basis.n_field >= p.meta["target_dofs"]
        assert p.basis.n_field < p.meta["target_dofs"] + 3000


def test_family_dof_segments():
    """The mesh of every field DOF of a union of independent meshes: each
    tet's 20 DOFs lie in the tet's own mesh, and each mesh keeps exactly its
    own number of DOFs (FamilySweep's row segments)."""
    from emerge_amd import workloads as w
    from emerge_amd.basis import Nedelec2
    from emerge_amd.family import dof_segments
    from emerge_amd.mesh import Mesh3D
    parts = [w.waveguide(cells=(3, 2, k), length=5e-3 * k) for k in (2, 4, 3)]
    nodes, tets, voff, tet_mesh = [], [], [], []
    v = 0
    x0 = 0.0
    for m, p in enumerate(parts):
        xyz = p.mesh.nodes.copy()
        xyz[0] += x0 - xyz[0].min()
        x0 = xyz[0].max() + 1e-3
        nodes.append(xyz)
        tets.append(p.mesh.tets + v)
        voff.append(v)
        tet_mesh.append(np.full(p.mesh.tets.shape[1], m))
        v += p.mesh.n_nodes
    union = Mesh3D(np.concatenate(nodes, axis=1), np.concatenate(tets, axis=1))
    basis = Nedelec2(union)
    seg = dof_segments(union, basis, voff)
    tm = np.concatenate(tet_mesh)
    assert np.all(seg[basis.tet_to_field] == tm[None, :])
    assert np.array_equal(np.bincount(seg, minlength=len(parts)), [p.basis.n_field for p in parts])
