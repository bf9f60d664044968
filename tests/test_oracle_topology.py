"""Topology and DOF numbering: vectorised product code vs the set/dict oracle."""
import numpy as np
import pytest

from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import topology as otopo


@pytest.mark.parametrize("cells,jitter,seed", [((2, 2, 2), 0.0, 0), ((4, 3, 2), 0.15, 7), ((3, 1, 5), 0.1, 3)])
def test_topology_matches_oracle(cells, jitter, seed):
    nodes, tets = kuhn_box(*cells, size=(0.3, 0.2, 0.5), jitter=jitter, seed=seed)
    m = Mesh3D(nodes, tets)
    ref = otopo.pre_update(nodes, tets)
    for key in ("edges", "tris", "tet_to_edge", "tet_to_tri", "tri_to_tet", "tri_to_edge"):
        np.testing.assert_array_equal(getattr(m, key), ref[key], err_msg=key)
    np.testing.assert_array_equal(m.edge_lengths, ref["edge_lengths"])
    f = Nedelec2(m)
    maps = otopo.nedelec2_maps(ref, tets.shape[1])
    np.testing.assert_array_equal(f.tet_to_field, maps["tet_to_field"])
    np.testing.assert_array_equal(f.tri_to_field, maps["tri_to_field"])
    np.testing.assert_array_equal(f.edge_to_field, maps["edge_to_field"])
    assert f.n_field == maps["n_field"]


def test_shuffled_vertex_order_still_matches():
    """Tets given in arbitrary local vertex order (as gmsh may deliver)."""
    nodes, tets = kuhn_box(3, 2, 2, jitter=0.1, seed=2)
    rng = np.random.default_rng(5)
    tets = np.stack([rng.permutation(tets[:, t]) for t in range(tets.shape[1])], axis=1)
    m = Mesh3D(nodes, tets)
    ref = otopo.pre_update(nodes, tets)
    for key in ("tet_to_edge", "tet_to_tri", "tri_to_tet", "tri_to_edge"):
        np.testing.assert_array_equal(getattr(m, key), ref[key], err_msg=key)


@pytest.mark.parametrize("n", [1, 2, 5, 10])
def test_kuhn_closed_form_counts(n):
    m = Mesh3D(*kuhn_box(n, n, n))
    V = (n + 1) ** 3
    # Kuhn cube: 6 n^3 tets; edges: axis 3n(n+1)^2, face diagonals 3n^2(n+1), body n^3
    nE = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    # Euler: V - E + F - T = 1 for a ball
    nF = 1 - V + nE + 6 * n ** 3
    assert m.n_tets == 6 * n ** 3 and m.n_nodes == V
    assert m.n_edges == nE and m.n_tris == nF
    assert len(m.boundary_tris()) == 12 * n * n


def test_c1_counts():
    n = 60
    nE = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    nF = 1 - (n + 1) ** 3 + nE + 6 * n ** 3
    assert (nE, nF, 2 * nE + 2 * nF) == (1544580, 2613600, 8316360)


def test_tri_rowcol_matches_oracle():
    nodes, tets = kuhn_box(2, 2, 1, jitter=0.1, seed=1)
    m = Mesh3D(nodes, tets)
    f = Nedelec2(m)
    rows, cols = otopo.empty_tri_rowcol(f.tri_to_field)
    r2, c2 = f.tri_rowcol(np.arange(m.n_tris))
    np.testing.assert_array_equal(rows, r2)
    np.testing.assert_array_equal(cols, c2)
