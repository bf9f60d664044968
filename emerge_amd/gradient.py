"""Discrete gradient from second-order nodal (Lagrange) functions into the
second-order Nedelec space, and the diagonal of the operator on that
gradient space.  Host-side set-up of the auxiliary-space preconditioner of
the Krylov solve (the solve applies ``D^-1 + G D_g^-1 G^T``).

Basis (E/elements/ned2_interp.py:243-296, in the global-id sorted frame of
each edge x<y and face p<q<r; W_ab = lam_a grad lam_b - lam_b grad lam_a):

* edge DOFs (1st, 2nd kind): ``-|xy| lam_x W_xy``, ``-|xy| lam_y W_xy``;
* face DOFs (1st, 2nd kind): ``|pr| lam_q W_pr``, ``-|pq| lam_r W_pq``;
  the third face function follows from
  ``lam_p W_qr = lam_q W_pr - lam_r W_pq``.

Nodal space: vertex hats ``lam_a`` (column a) and edge bubbles
``lam_x lam_y`` (column V + e).  With ``sum_v lam_v = 1`` on every tet,

* ``grad lam_a = sum_{edges ab} W_ba`` and
  ``W_xy = lam_x W_xy + lam_y W_xy + sum_{faces xyc} lam_c W_xy``;
* ``grad(lam_x lam_y) = (lam_x - lam_y) W_xy
  + sum_{faces xyc} (lam_x W_cy + lam_y W_cx)``,

every term of which is one edge or face DOF function (or the face
identity above), so ``G`` is exact: the curl-curl matrix annihilates it.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _whitney_and_bubbles(mesh, basis):
    """(Wexp, Ge), both (N, nE) real CSR in the global DOF numbering:
    column e of Wexp is the lowest-order Whitney function W_xy of edge
    e = (x, y), x < y, and column e of Ge is grad(lam_x lam_y)."""
    nE, nF = mesh.n_edges, mesh.n_tris
    N = basis.n_field
    L = mesh.edge_lengths
    e1 = basis.edge_to_field[0]
    e2 = basis.edge_to_field[1]
    f1 = np.arange(nF) + nE
    f2 = np.arange(nF) + 2 * nE + nF
    t2e = mesh.tri_to_edge              # (pq, qr, pr) of the sorted triangle
    s1 = L[t2e[2]]                      # |pr|
    s2 = L[t2e[0]]                      # |pq|

    # W_xy = lam_x W_xy + lam_y W_xy + sum over faces xyc of lam_c W_xy
    rows = [e1, e2]
    cols = [np.arange(nE), np.arange(nE)]
    vals = [-1.0 / L, -1.0 / L]
    # lam_c W_xy for the faces through the edge; edge k of the face:
    # k=0 (p,q), c=r: -F2/s2;  k=1 (q,r), c=p: F1/s1 + F2/s2;  k=2 (p,r), c=q: F1/s1
    wf = {0: (0.0 * s1, -1.0 / s2), 1: (1.0 / s1, 1.0 / s2), 2: (1.0 / s1, 0.0 * s2)}
    # grad(lam_x lam_y) face part: k=0: (-2/s1, -1/s2); k=1: (1/s1, -1/s2); k=2: (1/s1, 2/s2)
    bf = {0: (-2.0 / s1, -1.0 / s2), 1: (1.0 / s1, -1.0 / s2), 2: (1.0 / s1, 2.0 / s2)}
    brows = [e1, e2]
    bcols = [np.arange(nE), np.arange(nE)]
    bvals = [-1.0 / L, 1.0 / L]
    for k in range(3):
        e = t2e[k]
        a1, a2 = wf[k]
        rows += [f1, f2]
        cols += [e, e]
        vals += [a1, a2]
        b1, b2 = bf[k]
        brows += [f1, f2]
        bcols += [e, e]
        bvals += [b1, b2]
    Wexp = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, nE))
    Ge = sp.csr_matrix((np.concatenate(bvals), (np.concatenate(brows), np.concatenate(bcols))), shape=(N, nE))
    return Wexp, Ge


def discrete_gradient(mesh, basis) -> sp.csr_matrix:
    """(N, V + nE) real CSR: column a is grad lam_a, column V + e is
    grad(lam_x lam_y) of edge e = (x, y), in the global DOF numbering."""
    V, nE = mesh.n_nodes, mesh.n_edges
    Wexp, Ge = _whitney_and_bubbles(mesh, basis)
    # grad lam_a = sum over edges (x, y): +W_xy where a = y, -W_xy where a = x
    inc = sp.csr_matrix((np.concatenate([np.ones(nE), -np.ones(nE)]),
                         (np.concatenate([np.arange(nE), np.arange(nE)]),
                          np.concatenate([mesh.edges[1], mesh.edges[0]]))), shape=(nE, V))
    Gv = (Wexp @ inc).tocsr()
    G = sp.hstack([Gv, Ge], format="csr")
    G.sum_duplicates()
    G.eliminate_zeros()
    return G


def vector_interpolation(mesh, basis) -> sp.csr_matrix:
    """(N, 3 V) real CSR: column d V + a holds the Nedelec-2 coefficients of
    the continuous piecewise-linear vector field lam_a e_d (the second-order
    space contains every linear field, so the interpolation is exact).

    With e_d = sum_k x_k,d grad lam_k and sum_k grad lam_k = 0 on every tet,
    lam_a e_d = sum_{edges ak} (x_k - x_a)_d lam_a grad lam_k, and
    lam_a grad lam_k = (W_ak + grad(lam_a lam_k)) / 2, so for the edge
    e = (x, y), x < y, with t = x_y - x_x: column (a = x) gets
    t_d (W_xy + grad(lam_x lam_y)) / 2 and column (a = y) gets
    t_d (W_xy - grad(lam_x lam_y)) / 2."""
    V, nE = mesh.n_nodes, mesh.n_edges
    Wexp, Ge = _whitney_and_bubbles(mesh, basis)
    x, y = mesh.edges[0], mesh.edges[1]
    t = mesh.nodes[:, y] - mesh.nodes[:, x]
    ar = np.concatenate([np.arange(nE), np.arange(nE)])
    cols = np.concatenate([x, y])
    blocks = []
    for d in range(3):
        both = sp.csr_matrix((np.concatenate([t[d], t[d]]), (ar, cols)), shape=(nE, V))
        signed = sp.csr_matrix((np.concatenate([t[d], -t[d]]), (ar, cols)), shape=(nE, V))
        blocks.append(0.5 * (Wexp @ both) + 0.5 * (Ge @ signed))
    P = sp.hstack(blocks, format="csr")
    P.sum_duplicates()
    P.eliminate_zeros()
    return P


def _tet_gradients(nodes: np.ndarray, tets: np.ndarray):
    """Barycentric gradients (T, 4, 3) and volumes (T,) of every tet."""
    X = nodes[:, tets]                                   # (3, 4, T)
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=1)  # (3, 3, T)
    J = np.moveaxis(J, 2, 0)                             # (T, 3 coord, 3 edge)
    det = np.linalg.det(J)
    inv = np.linalg.inv(J)                               # (T, 3 edge, 3 coord): rows grad lam_1..3
    g = np.empty((tets.shape[1], 4, 3))
    g[:, 1:] = inv
    g[:, 0] = -inv.sum(axis=1)
    return g, np.abs(det) / 6.0


def nodal_volume_diagonal(mesh, eps_r: np.ndarray) -> np.ndarray:
    """int grad phi_j . (eps grad phi_j) dV for every nodal function
    (V + nE,), with eps the tet's permittivity, a scalar (T,) or a 3x3
    tensor (3, 3, T): vertex hats sum V g_a.eps g_a, edge bubbles
    V/10 (g_x.eps g_x + g_y.eps g_y) + V/20 (g_x.eps g_y + g_y.eps g_x)
    (int lam_a^2 = V/10, int lam_x lam_y = V/20)."""
    V, nE = mesh.n_nodes, mesh.n_edges
    g, vol = _tet_gradients(mesh.nodes, mesh.tets)
    eps = np.asarray(eps_r)
    out = np.zeros(V + nE, dtype=np.complex128)
    if eps.ndim == 3:
        gg = np.einsum("tia,abt,tjb->tij", g, eps, g) * vol[:, None, None]   # (T, 4, 4) g_i.eps g_j V
    else:
        gg = np.einsum("tic,tjc->tij", g, g) * (eps * vol)[:, None, None]
    for i in range(4):
        np.add.at(out, mesh.tets[i], gg[:, i, i])
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for i, j in pairs:
        eid = mesh.edge_ids(mesh.tets[i], mesh.tets[j])
        val = (gg[:, i, i] + gg[:, j, j]) / 10.0 + (gg[:, i, j] + gg[:, j, i]) / 20.0
        np.add.at(out, V + eid, val)
    return out


def nodal_surface_diagonal(mesh, tri_ids: np.ndarray) -> np.ndarray:
    """int_S |grad_t phi_j|^2 dS over the given triangles (V + nE,): vertex
    hats A |g|^2, edge bubbles A/6 (|g_x|^2 + |g_y|^2 + g_x . g_y), with g the
    in-plane barycentric gradients of each triangle."""
    V, nE = mesh.n_nodes, mesh.n_edges
    out = np.zeros(V + nE)
    if len(tri_ids) == 0:
        return out
    t = mesh.tris[:, tri_ids]                            # (3, n)
    X = mesh.nodes[:, t]                                 # (3 coord, 3 vert, n)
    a = X[:, 1] - X[:, 0]
    b = X[:, 2] - X[:, 0]
    n = np.cross(a, b, axis=0)
    area2 = np.sqrt((n ** 2).sum(axis=0))
    A = area2 / 2
    # in-plane gradients: grad lam_i = (n x e_i) / |n|^2 with e_i the edge opposite i
    nn = n / area2 ** 2
    opp = [X[:, 2] - X[:, 1], X[:, 0] - X[:, 2], X[:, 1] - X[:, 0]]
    g = [np.cross(nn, e, axis=0) for e in opp]
    for i in range(3):
        np.add.at(out, t[i], A * (g[i] ** 2).sum(axis=0))
    for i, j in [(0, 1), (1, 2), (0, 2)]:
        eid = mesh.edge_ids(t[i], t[j])
        val = A / 6.0 * ((g[i] ** 2).sum(axis=0) + (g[j] ** 2).sum(axis=0) + (g[i] * g[j]).sum(axis=0))
        np.add.at(out, V + eid, val)
    return out


def vector_volume_diagonals(mesh, eps_r, inv_mu_r=None):
    """Diagonals of the vector-nodal auxiliary space (columns d V + a, the
    order of `vector_interpolation`): (curl, mass) with
    curl = int (grad lam_a x e_d) . mu^-1 (grad lam_a x e_d) and
    mass = int lam_a^2 e_d . eps e_d = V/10 eps_dd summed over the tets of
    vertex a (int lam_a^2 = V/10).  eps_r and inv_mu_r are scalars (T,) or
    3x3 tensors (3, 3, T); inv_mu_r defaults to 1."""
    V, T = mesh.n_nodes, mesh.n_tets
    g, vol = _tet_gradients(mesh.nodes, mesh.tets)          # (T, 4, 3)
    eps = np.asarray(eps_r)
    imu = np.ones(T) if inv_mu_r is None else np.asarray(inv_mu_r)
    curl = np.zeros(3 * V, dtype=np.complex128)
    mass = np.zeros(3 * V, dtype=np.complex128)
    eye = np.eye(3)
    for d in range(3):
        ed = eye[d]
        epsd = eps[d, d] if eps.ndim == 3 else eps                      # (T,)
        for i in range(4):
            c = np.cross(g[:, i], ed)                                  # (T, 3): curl(lam_i e_d)
            if imu.ndim == 3:
                cc = np.einsum("ta,abt,tb->t", c, imu, c)
            else:
                cc = (c * c).sum(axis=1) * imu
            np.add.at(curl, d * V + mesh.tets[i], cc * vol)
            np.add.at(mass, d * V + mesh.tets[i], epsd * vol / 10.0)
    return curl, mass


def vector_surface_diagonal(mesh, tri_ids: np.ndarray) -> np.ndarray:
    """int_S |(lam_a e_d)_t|^2 dS over the given triangles (3 V,), columns
    d V + a: A/6 (1 - n_d^2) per triangle of vertex a (int lam_a^2 = A/6,
    n the unit normal)."""
    V = mesh.n_nodes
    out = np.zeros(3 * V)
    if len(tri_ids) == 0:
        return out
    t = mesh.tris[:, tri_ids]
    X = mesh.nodes[:, t]
    n = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], axis=0)
    area2 = np.sqrt((n ** 2).sum(axis=0))
    nu = n / area2
    A = area2 / 2
    for d in range(3):
        w = A / 6.0 * (1.0 - nu[d] ** 2)
        for i in range(3):
            np.add.at(out, d * V + t[i], w)
    return out
