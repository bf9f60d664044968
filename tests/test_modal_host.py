"""Host side of the modal port (emerge_amd.modal, no device calls) against
the boundary-mode oracle: the face's DOF map and frame, the basis values,
the E / H interpolation at arbitrary points and the power flux."""
import numpy as np

from emerge_amd import bc as pbc
from emerge_amd.mesh import Mesh3D, kuhn_box
from emerge_amd.modal import PortFace
from oracle import modal as OM
from oracle import ports as oports


def _face():
    nodes, tets = kuhn_box(5, 3, 1, size=(7e-3, 3e-3, 1e-3), jitter=0.0)
    m = Mesh3D(nodes, tets)
    tv = m.nodes[:, m.tris]
    port = np.nonzero(np.all(np.abs(tv[2]) < 1e-12, axis=0))[0]
    cs = ([1, 0, 0], [0, 1, 0], [0, 0, 1], [3.5e-3, 1.5e-3, 0.0])
    return m, port, cs


def test_face_dof_map_matches_oracle():
    m, port, cs = _face()
    face = PortFace(m, port, pbc.CoordinateSystem(*cs))
    uniq, lt, edges, t2e = OM.surface_mesh(m.nodes, m.tris, port)
    t2f, e2f, n = OM.dof_map(uniq.size, edges, lt, t2e)
    assert np.array_equal(face.node_ids, uniq) and np.array_equal(face.tris, lt)
    assert np.array_equal(face.edges, edges) and np.array_equal(face.tri_to_field, t2f)
    assert np.array_equal(face.edge_to_field, e2f) and face.n_field == n


def test_basis_fields_and_power_match_oracle():
    m, port, cs = _face()
    face = PortFace(m, port, pbc.CoordinateSystem(*cs))
    g = np.random.default_rng(5)
    field = g.normal(size=face.n_field) + 1j * g.normal(size=face.n_field)
    x = g.uniform(0.05, 6.95, 300) * 1e-3 - 3.5e-3
    y = g.uniform(0.05, 2.95, 300) * 1e-3 - 1.5e-3
    k0, beta = 150.0, 120.0
    urinv = np.ones((3, face.tris.shape[1]))
    E, H = face.fields(field, x, y, k0, beta, urinv)
    Eo = OM.interp_E(field, face.xy, face.tris, face.tri_to_field, x, y)
    Ho = OM.interp_H(field, face.xy, face.tris, face.tri_to_field, x, y, k0, beta, urinv)
    assert np.abs(E - Eo).max() <= 1e-12 * np.abs(Eo).max()
    assert np.abs(H - Ho).max() <= 1e-12 * np.abs(Ho).max()
    P = face.power_flux(field, k0, beta, urinv)
    Po = OM.power_flux(field, face.xy, face.tris, face.tri_to_field, k0, beta, urinv)
    assert abs(P - Po) <= 1e-12 * abs(Po)


def test_shared_edges():
    """Face edges shared with 3-D triangles: all of them for the face's own
    triangles; the face's boundary edges on the walls y = 0 and y = h."""
    m, port, cs = _face()
    face = PortFace(m, port, pbc.CoordinateSystem(*cs))
    assert np.array_equal(face.shared_edges(m.tris[:, port]), np.arange(face.edges.shape[1]))
    tv = m.nodes[:, m.tris]
    walls = np.nonzero(np.all(np.abs(tv[1]) < 1e-12, axis=0) | np.all(np.abs(tv[1] - 3e-3) < 1e-12, axis=0))[0]
    e = face.shared_edges(m.tris[:, walls])
    y = face.xy[1][face.edges[:, e]]
    assert e.size == 2 * 5 and np.all(np.abs(np.abs(y) - 1.5e-3) < 1e-12)


def test_face_frame_without_cs():
    """A port without a frame gets one from its face: orthonormal, z along
    the normal out of the volume (-z for the face z = 0 of the box, although
    its triangles are wound both ways), origin at the face's centroid."""
    from emerge_amd.modal import _face_frame
    m, port, _ = _face()
    cs = _face_frame(m, port)
    B = np.array([cs.xax, cs.yax, cs.zax])
    assert np.allclose(B @ B.T, np.eye(3), atol=1e-12)
    assert np.allclose(cs.zax, [0, 0, -1], atol=1e-12)
    assert np.allclose(np.cross(cs.xax, cs.yax), cs.zax, atol=1e-12)
    assert np.allclose(cs.origin, [3.5e-3, 1.5e-3, 0.0], atol=1e-12)
