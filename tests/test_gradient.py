"""Pins of the discrete gradient and nodal diagonals of the auxiliary-space
preconditioner (emerge_amd.gradient), on the CPU:

* curl-free: the oracle's curl-curl matrix annihilates G (a planted wrong
  coefficient does not);
* pointwise: sum_i (G c)_i N_i, with N_i the reference's interpolation basis
  (oracle.quadrature.basis_values), equals grad(sum_j c_j phi_j) evaluated
  from barycentric coordinates;
* diagonals: the closed-form int eps |grad phi|^2 and int_S |grad_t phi|^2
  equal g^T M g with M the exactly integrated mass matrix (quadrature of the
  basis) and the pinned Robin matrix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from emerge_amd import gradient as gr
from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import curlcurl as cc
from oracle import quadrature as qd
from oracle import robin


@pytest.fixture(scope="module")
def mesh():
    nodes, tets = kuhn_box(3, 2, 2, size=(0.02, 0.015, 0.01), jitter=0.15, seed=5)
    m = Mesh3D(nodes, tets)
    return m, Nedelec2(m)


def test_gradient_is_curl_free(mesh):
    m, f = mesh
    G = gr.discrete_gradient(m, f)
    assert G.shape == (f.n_field, m.n_nodes + m.n_edges)
    T = m.n_tets
    E, _ = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(np.ones(T) + 0j),
                                          cc.isotropic(np.ones(T) + 0j))
    rng = np.random.default_rng(0)
    x = rng.standard_normal(G.shape[1])
    Gx = G @ x
    scale = abs(E).max() * np.linalg.norm(Gx)
    assert np.linalg.norm(E @ Gx) <= 1e-13 * scale
    # planted mistake: one face coefficient off by 10% leaves a curl
    Gb = G.copy().tolil()
    face_row = m.n_edges + 3
    j = Gb.rows[face_row][0]
    Gb[face_row, j] = 1.1 * Gb[face_row, j]
    Gb = Gb.tocsr()
    assert np.linalg.norm(E @ (Gb @ x)) > 1e-6 * scale


def _bary(verts, pts):
    """Barycentric coordinates (4, P) and their gradients (4, 3) of a tet."""
    J = np.stack([verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0], verts[:, 3] - verts[:, 0]], axis=1)
    Ji = np.linalg.inv(J)
    l123 = Ji @ (pts - verts[:, :1])
    lam = np.vstack([1 - l123.sum(axis=0), l123])
    g = np.vstack([-Ji.sum(axis=0), Ji])
    return lam, g


def test_gradient_matches_nodal_function_pointwise(mesh):
    m, f = mesh
    G = gr.discrete_gradient(m, f)
    rng = np.random.default_rng(1)
    c = rng.standard_normal(G.shape[1])
    u = G @ c
    V = m.n_nodes
    for t in (0, 7, m.n_tets - 1):
        vid = m.tets[:, t]
        verts = m.nodes[:, vid]
        w = rng.dirichlet(np.ones(4), size=5).T  # (4, 5) interior points
        pts = verts @ w
        edges_g = m.edges[:, m.tet_to_edge[:, t]]
        tris_g = m.tris[:, m.tet_to_tri[:, t]]
        lem = cc.local_mapping(vid[:, None], edges_g[:, :, None])[:, :, 0]
        ltm = cc.local_mapping(vid[:, None], tris_g[:, :, None])[:, :, 0]
        N = qd.basis_values(pts, verts[0], verts[1], verts[2], lem, ltm)  # (20, 3, P)
        field = np.einsum("i,icp->cp", u[f.tet_to_field[:, t]], N)
        lam, g = _bary(verts, pts)
        ref = np.zeros_like(field)
        for a in range(4):
            ref += c[vid[a]] * g[a][:, None]
        for a in range(4):
            for b in range(a + 1, 4):
                e = m.get_edge(vid[a], vid[b])
                ref += c[V + e] * (lam[a] * g[b][:, None] + lam[b] * g[a][:, None])
        assert np.abs(field - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("kind", ["scalar", "tensor"])
def test_nodal_volume_diagonal_is_gMg(mesh, kind):
    """Scalar eps and a general complex (non-symmetric) 3x3 tensor per tet."""
    m, f = mesh
    T = m.n_tets
    rng = np.random.default_rng(2)
    if kind == "scalar":
        eps = rng.uniform(1, 4, T) * (1 - 0.01j)
        eps_t = lambda t: eps[t] * np.eye(3)  # noqa: E731
    else:
        eps = (np.eye(3)[:, :, None] * rng.uniform(1, 4, (3, 1, T))
               + 0.3 * (rng.standard_normal((3, 3, T)) + 1j * rng.standard_normal((3, 3, T))))
        eps_t = lambda t: eps[:, :, t]  # noqa: E731
    rows, cols, vals = [], [], []
    for t in range(T):
        vid = m.tets[:, t]
        _, Mt = qd.element_matrices_quadrature(m.nodes[:, vid], vid, m.edges[:, m.tet_to_edge[:, t]],
                                               m.tris[:, m.tet_to_tri[:, t]], np.eye(3), eps_t(t))
        d = f.tet_to_field[:, t]
        rows.append(np.repeat(d, 20)); cols.append(np.tile(d, 20)); vals.append(Mt.ravel())
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(f.n_field, f.n_field))
    G = gr.discrete_gradient(m, f)
    ref = np.asarray(G.multiply(M @ G).sum(axis=0)).ravel()
    got = gr.nodal_volume_diagonal(m, eps)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_nodal_surface_diagonal_is_gSg(mesh):
    m, f = mesh
    tri_ids = m.boundary_tris()[::3]
    rows, cols, vals = [], [], []
    for it in tri_ids:
        S = robin.ned2_tri_stiff(m.nodes[:, m.tris[:, it]], 1.0)
        d = f.tri_to_field[:, it]
        rows.append(np.repeat(d, 8)); cols.append(np.tile(d, 8)); vals.append(np.real(S).ravel())
    S = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(f.n_field, f.n_field))
    G = gr.discrete_gradient(m, f)
    ref = np.asarray(G.multiply(S @ G).sum(axis=0)).ravel()
    got = gr.nodal_surface_diagonal(m, tri_ids)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def _pointwise_vector_error(m, f, P, seed):
    """max |sum_i (P c)_i N_i - sum_a c_{a,d} lam_a e_d| over interior points
    of three tets, relative to the field's size."""
    rng = np.random.default_rng(seed)
    V = m.n_nodes
    c = rng.standard_normal(3 * V)
    u = P @ c
    worst = 0.0
    for t in (0, 5, m.n_tets - 1):
        vid = m.tets[:, t]
        verts = m.nodes[:, vid]
        pts = verts @ rng.dirichlet(np.ones(4), size=5).T
        edges_g = m.edges[:, m.tet_to_edge[:, t]]
        tris_g = m.tris[:, m.tet_to_tri[:, t]]
        lem = cc.local_mapping(vid[:, None], edges_g[:, :, None])[:, :, 0]
        ltm = cc.local_mapping(vid[:, None], tris_g[:, :, None])[:, :, 0]
        N = qd.basis_values(pts, verts[0], verts[1], verts[2], lem, ltm)
        field = np.einsum("i,icp->cp", u[f.tet_to_field[:, t]], N)
        lam, _ = _bary(verts, pts)
        ref = np.zeros_like(field)
        for d in range(3):
            for a in range(4):
                ref[d] += c[d * V + vid[a]] * lam[a]
        worst = max(worst, np.abs(field - ref).max() / np.abs(ref).max())
    return worst


def test_vector_interpolation_pointwise(mesh):
    """sum_i (P c)_i N_i equals the piecewise-linear field sum_a c_a lam_a
    (the second-order basis holds every linear field exactly); one
    coefficient of P off by 10% does not."""
    m, f = mesh
    P = gr.vector_interpolation(m, f)
    assert P.shape == (f.n_field, 3 * m.n_nodes)
    assert _pointwise_vector_error(m, f, P, 3) <= 1e-10
    Pb = P.tolil()
    row = f.tet_to_field[12, 0]  # a face DOF of the first sampled tet
    j = Pb.rows[row][int(np.argmax(np.abs(Pb.data[row])))]
    Pb[row, j] = 1.1 * Pb[row, j]
    assert _pointwise_vector_error(m, f, Pb.tocsr(), 3) > 1e-3


def test_vector_interpolation_constants_are_coordinate_gradients(mesh):
    """The constant field e_d = sum_a lam_a e_d is grad x_d: P 1_d = G x_d."""
    m, f = mesh
    P = gr.vector_interpolation(m, f)
    G = gr.discrete_gradient(m, f)
    V = m.n_nodes
    for d in range(3):
        ones = np.zeros(3 * V)
        ones[d * V:(d + 1) * V] = 1.0
        x = np.zeros(G.shape[1])
        x[:V] = m.nodes[d]
        a, b = P @ ones, G @ x
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("kind", ["scalar", "tensor"])
def test_vector_diagonals_are_PtMP(mesh, kind):
    """The vector-nodal diagonals equal diag(P^T K P) with K the oracle's
    curl-curl matrix (exact integrals) and diag(P^T M P) with the exactly
    integrated mass matrix, for scalar and general complex tensors."""
    m, f = mesh
    T = m.n_tets
    rng = np.random.default_rng(4)
    if kind == "scalar":
        eps = rng.uniform(1, 4, T) * (1 - 0.01j)
        imu = rng.uniform(0.5, 2, T) + 0j
        eps_t = lambda t: eps[t] * np.eye(3)  # noqa: E731
        imu_t = lambda t: np.eye(3) * imu[t]  # noqa: E731
    else:
        eps = (np.eye(3)[:, :, None] * rng.uniform(1, 4, (3, 1, T))
               + 0.3 * (rng.standard_normal((3, 3, T)) + 1j * rng.standard_normal((3, 3, T))))
        mu = (np.eye(3)[:, :, None] * rng.uniform(1, 3, (3, 1, T))
              + 0.2 * (rng.standard_normal((3, 3, T)) + 1j * rng.standard_normal((3, 3, T))))
        imu = np.moveaxis(np.linalg.inv(np.moveaxis(mu, 2, 0)), 0, 2)
        eps_t = lambda t: eps[:, :, t]  # noqa: E731
        imu_t = lambda t: imu[:, :, t]  # noqa: E731
    rows, cols, kv, mv = [], [], [], []
    for t in range(T):
        vid = m.tets[:, t]
        Kt, Mt = qd.element_matrices_quadrature(m.nodes[:, vid], vid, m.edges[:, m.tet_to_edge[:, t]],
                                                m.tris[:, m.tet_to_tri[:, t]], imu_t(t), eps_t(t))
        d = f.tet_to_field[:, t]
        rows.append(np.repeat(d, 20)); cols.append(np.tile(d, 20)); kv.append(Kt.ravel()); mv.append(Mt.ravel())
    shape = (f.n_field, f.n_field)
    r, c = np.concatenate(rows), np.concatenate(cols)
    K = sp.csr_matrix((np.concatenate(kv), (r, c)), shape=shape)
    M = sp.csr_matrix((np.concatenate(mv), (r, c)), shape=shape)
    P = gr.vector_interpolation(m, f)
    curl, mass = gr.vector_volume_diagonals(m, eps, imu)
    refk = np.asarray(P.multiply(K @ P).sum(axis=0)).ravel()
    refm = np.asarray(P.multiply(M @ P).sum(axis=0)).ravel()
    assert np.abs(curl - refk).max() <= 1e-9 * np.abs(refk).max()
    assert np.abs(mass - refm).max() <= 1e-9 * np.abs(refm).max()


def test_vector_surface_diagonal_is_PtSP(mesh):
    m, f = mesh
    tri_ids = m.boundary_tris()[::3]
    rows, cols, vals = [], [], []
    for it in tri_ids:
        S = robin.ned2_tri_stiff(m.nodes[:, m.tris[:, it]], 1.0)
        d = f.tri_to_field[:, it]
        rows.append(np.repeat(d, 8)); cols.append(np.tile(d, 8)); vals.append(np.real(S).ravel())
    S = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(f.n_field, f.n_field))
    P = gr.vector_interpolation(m, f)
    ref = np.asarray(P.multiply(S @ P).sum(axis=0)).ravel()
    got = gr.vector_surface_diagonal(m, tri_ids)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
