"""Material tensors (emerge_amd.materials) pinned by what the construction
fixes: the PML stretch is 1 at the layer's inner face (no jump against the
box), 1 - j deltamax at its outer face, has a growing loss in between, gives
a uniaxial tensor diag(s, s, 1/s) for a single-axis layer and a tensor with
determinant sx sy sz for edge and corner layers; conductivity enters as
-j cond / (w eps0), so cond = w eps0 adds exactly -1j."""
import numpy as np
import pytest

from emerge_amd.materials import lossy_permittivity, pml_layer_tensors, pml_stretch, tet_centers
from emerge_amd.mesh import Mesh3D, kuhn_box
from emerge_amd.sweep import EPS0

CENTER = (0.5, 0.5, 0.5)
DIMS = (1.0, 1.0, 1.0)


def test_stretch_faces_and_growth():
    th = 0.2
    z = 1.0 + th * np.linspace(0, 1, 11)
    xyz = np.vstack([np.full_like(z, 0.5), np.full_like(z, 0.5), z])
    s = pml_stretch(xyz, CENTER, DIMS, (0, 0, 1), th, exponent=2.0, deltamax=6.0)
    assert np.all(s[:2] == 1)
    assert s[2, 0] == 1
    assert abs(s[2, -1] - (1 - 6j)) < 1e-14
    assert np.all(np.diff(-s[2].imag) > 0) and np.all(s[2].real == 1)
    # the -z side measures depth downwards from z = 0
    s2 = pml_stretch(np.vstack([[0.5], [0.5], [-th]]), CENTER, DIMS, (0, 0, -1), th, exponent=2.0, deltamax=6.0)
    assert abs(s2[2, 0] - (1 - 6j)) < 1e-14
    with pytest.raises(ValueError):
        pml_stretch(np.vstack([[0.5], [0.5], [0.9]]), CENTER, DIMS, (0, 0, 1), th)


def _layer_mesh(direction, th=0.25):
    """Kuhn grid filling the layer on the given side of the unit box."""
    lo = [0.0, 0.0, 0.0]
    size = [1.0, 1.0, 1.0]
    for a, d in enumerate(direction):
        if d > 0:
            lo[a], size[a] = 1.0, th
        elif d < 0:
            lo[a], size[a] = -th, th
    nodes, tets = kuhn_box(3, 3, 3, size=tuple(size))
    nodes = nodes + np.asarray(lo)[:, None]
    return Mesh3D(nodes, tets)


def test_single_axis_layer_is_uniaxial():
    m = _layer_mesh((0, 0, 1))
    eps, mu = pml_layer_tensors(m, np.arange(m.n_tets), CENTER, DIMS, (0, 0, 1), 0.25, er=2.0, ur=1.5)
    sz = pml_stretch(tet_centers(m), CENTER, DIMS, (0, 0, 1), 0.25)[2]
    np.testing.assert_allclose(eps[0, 0], 2.0 * sz, rtol=1e-15)
    np.testing.assert_allclose(eps[1, 1], 2.0 * sz, rtol=1e-15)
    np.testing.assert_allclose(eps[2, 2], 2.0 / sz, rtol=1e-15)
    np.testing.assert_allclose(mu / 1.5, eps / 2.0, rtol=1e-15)
    off = ~np.eye(3, dtype=bool)
    assert np.all(eps[off] == 0)


@pytest.mark.parametrize("direction", [(1, 0, -1), (-1, 1, 1)])
def test_edge_and_corner_layers_have_stretch_determinant(direction):
    m = _layer_mesh(direction)
    eps, _ = pml_layer_tensors(m, np.arange(m.n_tets), CENTER, DIMS, direction, 0.25)
    s = pml_stretch(tet_centers(m), CENTER, DIMS, direction, 0.25)
    det = eps[0, 0] * eps[1, 1] * eps[2, 2]
    np.testing.assert_allclose(det, s[0] * s[1] * s[2], rtol=1e-13)
    for a in range(3):
        assert np.all((s[a] == 1) == (direction[a] == 0))


def test_tet_centers_are_vertex_means():
    nodes, tets = kuhn_box(1, 1, 1, size=(1.0, 2.0, 3.0))
    m = Mesh3D(nodes, tets)
    c = tet_centers(m)
    for t in range(m.n_tets):
        np.testing.assert_allclose(c[:, t], m.nodes[:, m.tets[:, t]].mean(axis=1), rtol=1e-15)


def test_lossy_permittivity():
    f = 3e9
    w = 2 * np.pi * f
    e = lossy_permittivity(np.array([4.0, 4.0]), np.array([0.0, 0.01]), np.array([w * EPS0, 0.0]), f)
    assert abs(e[0] - (4 - 1j)) < 1e-14
    assert e[1] == 4 * (1 - 0.01j)


def test_conductor_split_follows_the_reference_limit():
    """cond[0, 0] > 1e7 S/m: perfect conductor; 0 < cond <= 1e7: lossy
    (ASM/assembler.py:240-288, E/settings.py:24)."""
    from emerge_amd.sweep import PECLIM, conductor_split
    assert PECLIM == 1e7
    c = np.array([0.0, 0.2, 1e7, 1.0000001e7, 5.8e7])
    pec, lossy = conductor_split(c)
    np.testing.assert_array_equal(pec, [False, False, False, True, True])
    np.testing.assert_array_equal(lossy, [0.0, 0.2, 1e7, 0.0, 0.0])
    t = np.zeros((3, 3, 5))
    for k in range(3):
        t[k, k] = c
    pec_t, lossy_t = conductor_split(t)
    np.testing.assert_array_equal(pec_t, pec)
    assert np.all(lossy_t[:, :, 3:] == 0) and lossy_t[1, 1, 1] == 0.2
