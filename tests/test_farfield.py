"""Pins of the far-field oracle (oracle/farfield.py) on the CPU.

* weighted normals: on a closed surface sum_j wn_j = 0 and, the edge-midpoint
  rule being exact for linear functions, sum_j wn_x,j x_j = the enclosed
  volume (outward orientation and the A/3 weights);
* a z-directed Hertzian dipole (exact near fields, e^{jwt}) sampled on a
  cube of one wavelength: the far field equals j Z0 k Il sin(t) / (4 pi)
  along theta, nothing radially, converging with the mesh;
* extinction: a plane wave crossing the closed cube radiates nothing outside
  (the E and H terms cancel to the discretisation error); a sign error in the
  H term (planted here) breaks both pins.
"""
import numpy as np
import pytest

from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import farfield as ff

K = 2 * np.pi   # wavelength 1


def cube_surface(n, side=1.0):
    nodes, tets = kuhn_box(n, n, n, size=(side, side, side))
    m = Mesh3D(nodes - side / 2, tets)
    bt = m.boundary_tris()
    tc = m.nodes[:, m.tets].mean(axis=1)
    tt = np.max(m.tri_to_tet[:, bt], axis=0)
    pos, wn, _ = ff.surface_geometry(m.nodes, m.tris[:, bt], None, tt, tc)
    return m, bt, pos, wn


def dipole_fields(pos):
    x, y, z = pos
    r = np.sqrt(x * x + y * y + z * z)
    th = np.arccos(z / r)
    ph = np.arctan2(y, x)
    e = np.exp(-1j * K * r)
    f1 = 1 + 1 / (1j * K * r)
    Hph = 1j * K * np.sin(th) / (4 * np.pi * r) * f1 * e
    Er = ff.Z0 * np.cos(th) / (2 * np.pi * r ** 2) * f1 * e
    Eth = 1j * ff.Z0 * K * np.sin(th) / (4 * np.pi * r) * (f1 - 1 / (K * r) ** 2) * e
    rh = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    thh = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
    phh = np.stack([-np.sin(ph), np.cos(ph), 0 * ph])
    return Er * rh + Eth * thh, Hph * phh


def dipole_error(pos, wn, flip_h=False):
    E, H = dipole_fields(pos)
    T = np.linspace(0.2, np.pi - 0.2, 9)
    P = np.full(9, 0.7)
    Eo, _ = ff.stratton_chu_ff(E, -H if flip_h else H, pos, wn, T, P, K)
    tho = np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), -np.sin(T)])
    rho = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)])
    ref = 1j * ff.Z0 * K * np.sin(T) / (4 * np.pi)
    err = np.abs((Eo * tho).sum(axis=0) - ref).max() / np.abs(ref).max()
    radial = np.abs((Eo * rho).sum(axis=0)).max() / np.abs(ref).max()
    return err, radial


def test_weighted_normals_divergence_theorem():
    _, _, pos, wn = cube_surface(4, side=0.8)
    assert np.abs(wn.sum(axis=1)).max() < 1e-14
    for a in range(3):
        assert abs((wn[a] * pos[a]).sum() - 0.8 ** 3) < 1e-14
        b = (a + 1) % 3
        assert abs((wn[a] * pos[b]).sum()) < 1e-14


def test_dipole_far_field():
    _, _, pos6, wn6 = cube_surface(6)
    e6, rad6 = dipole_error(pos6, wn6)
    _, _, pos10, wn10 = cube_surface(10)
    e10, _ = dipole_error(pos10, wn10)
    assert e6 < 1e-3 and rad6 < 1e-12
    assert e10 < e6 / 4          # converges with the mesh
    bad, _ = dipole_error(pos6, wn6, flip_h=True)
    assert bad > 0.5


def test_plane_wave_extinction():
    _, _, pos, wn = cube_surface(6)
    z = pos[2]
    E = np.stack([np.exp(-1j * K * z), 0 * z, 0 * z])
    H = np.stack([0 * z, np.exp(-1j * K * z) / ff.Z0, 0 * z])
    T = np.linspace(0.1, np.pi - 0.1, 12)
    P = np.linspace(0, 2 * np.pi, 12)
    Eo, Ho = ff.stratton_chu_ff(E, H, pos, wn, T, P, K)
    Ee, _ = ff.stratton_chu_ff(E, 0 * H, pos, wn, T, P, K)
    assert np.abs(Eo).max() < 3e-3 * np.abs(Ee).max()
    Eb, _ = ff.stratton_chu_ff(E, -H, pos, wn, T, P, K)
    assert np.abs(Eb).max() > 0.5 * np.abs(Ee).max()
    # Hout = r x Eout / Z0
    r = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)])
    np.testing.assert_allclose(Ho, np.cross(r, Eo, axis=0) / ff.Z0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("level", [0.5e-3, 2e-3])
def test_small_fields_are_left_out(level):
    """Samples at or below 1e-3 max|E| do not contribute."""
    _, _, pos, wn = cube_surface(4)
    E, H = dipole_fields(pos)
    j = 7
    E2, H2 = E.copy(), H.copy()
    mag = np.sqrt((np.abs(E) ** 2).sum(axis=0))
    E2[:, j] *= level * mag.max() / mag[j]
    H2[:, j] *= 1e6
    T, P = np.array([0.9]), np.array([0.3])
    a, _ = ff.stratton_chu_ff(E2, H2, pos, wn, T, P, K)
    keep = np.ones(pos.shape[1], dtype=bool)
    keep[j] = False
    b, _ = ff.stratton_chu_ff(E[:, keep], H[:, keep], pos[:, keep], wn[:, keep], T, P, K)
    if level < 1e-3:
        np.testing.assert_allclose(a, b, rtol=1e-12)
    else:
        assert np.abs(a - b).max() > 1e-3 * np.abs(b).max()


@pytest.mark.parametrize("origin", [None, (0.0, 0.0, 0.0)])
def test_product_surface_matches_oracle(origin):
    """emerge_amd.farfield.FarFieldSurface: same samples, same weights."""
    from emerge_amd.farfield import FarFieldSurface
    m, bt, pos, wn = cube_surface(5)
    if origin is not None:
        tc = m.nodes[:, m.tets].mean(axis=1)
        pos, wn, _ = ff.surface_geometry(m.nodes, m.tris[:, bt], origin, None, tc)
    s = FarFieldSurface(m, bt, origin=origin)
    np.testing.assert_array_equal(s.edge_centers, pos)
    np.testing.assert_allclose(s.weighted_normals, wn, rtol=0, atol=1e-15 * np.abs(wn).max())


def test_arc_on_plane_directions():
    """Directions of the cut lie in the plane, start along the projected
    reference direction and reach normal x reference at 90 degrees."""
    from emerge_amd.farfield import arc_on_plane
    n = np.array([0.0, 1.0, 1.0]) / np.sqrt(2)
    t, p = arc_on_plane((1.0, 0.3, 0.0), (0.0, 1.0, 1.0), (-180, 180), 9)
    u = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
    assert np.abs(n @ u).max() < 1e-12
    d = np.array([1.0, 0.3, 0.0]) - np.dot([1.0, 0.3, 0.0], n) * n
    e1 = d / np.linalg.norm(d)
    np.testing.assert_allclose(u[:, 4], e1, atol=1e-12)           # angle 0
    np.testing.assert_allclose(u[:, 6], np.cross(n, e1), atol=1e-12)   # +90 degrees
    with pytest.raises(ValueError):
        arc_on_plane((0.0, 1.0, 1.0), (0.0, 1.0, 1.0), (0, 90), 3)


@pytest.mark.parametrize("syms,se,sh", [
    (["Ex"], (1, -1, -1), (-1, 1, 1)),       # E wall x = 0: tangential E odd
    (["Hx"], (-1, 1, 1), (1, -1, -1)),       # H wall x = 0
    (["Ex", "Hy"], (1, 1, -1), (1, 1, -1)),
    (["Hz", "Ez"], (-1, -1, -1), (-1, -1, -1)),   # z flipped twice, then E and H once each
])
def test_symmetry_signs(syms, se, sh):
    """Component signs of the mirrored fields, as _EHSign applies them
    (MW/microwave_data.py:253-303)."""
    from emerge_amd.farfield import _signs
    a, b = _signs(syms)
    np.testing.assert_array_equal(a, se)
    np.testing.assert_array_equal(b, sh)
