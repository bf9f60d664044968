"""Oracle pin: brute-force quadrature of the Nedelec-2 tet basis.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The basis is evaluated exactly as the reference's field interpolation defines it
(``ned2_tet_interp``, E/elements/ned2_interp.py:141-301, with
``tet_coefficients`` at :47-78): for an edge with sorted vertices (1,2),
``N = L * F_k * (grad-part (b1 F2 - b2 F1, ...)) / (216 V^3)`` with ``F = a + bx
+ cy + dz``; for a face with sorted vertices (1,2,3), ``N_f1 = -L13 F2 (b1 F3 -
b3 F1, ...) / V1`` and ``N_f2 = L12 F3 (b1 F2 - b2 F1, ...) / V1``.  This is
independent of the closed-form element integrals in ``oracle.curlcurl``.

Integration uses a collapsed (Duffy) Gauss-Legendre product rule that is exact
for the polynomial degrees involved; curls use central differences, which are
exact for quadratic fields up to rounding.
"""
from __future__ import annotations

import numpy as np

from .curlcurl import local_mapping


def tet_coefficients(xs, ys, zs):
    """E/elements/ned2_interp.py:47-78 for one tet: a, b, c, d (4,) and V."""
    x1, x2, x3, x4 = xs
    y1, y2, y3, y4 = ys
    z1, z2, z3, z4 = zs
    V = abs(-x1*y2*z3/6 + x1*y2*z4/6 + x1*y3*z2/6 - x1*y3*z4/6 - x1*y4*z2/6 + x1*y4*z3/6
            + x2*y1*z3/6 - x2*y1*z4/6 - x2*y3*z1/6 + x2*y3*z4/6 + x2*y4*z1/6 - x2*y4*z3/6
            - x3*y1*z2/6 + x3*y1*z4/6 + x3*y2*z1/6 - x3*y2*z4/6 - x3*y4*z1/6 + x3*y4*z2/6
            + x4*y1*z2/6 - x4*y1*z3/6 - x4*y2*z1/6 + x4*y2*z3/6 + x4*y3*z1/6 - x4*y3*z2/6)
    a = np.array([x2*y3*z4 - x2*y4*z3 - x3*y2*z4 + x3*y4*z2 + x4*y2*z3 - x4*y3*z2,
                  -x1*y3*z4 + x1*y4*z3 + x3*y1*z4 - x3*y4*z1 - x4*y1*z3 + x4*y3*z1,
                  x1*y2*z4 - x1*y4*z2 - x2*y1*z4 + x2*y4*z1 + x4*y1*z2 - x4*y2*z1,
                  -x1*y2*z3 + x1*y3*z2 + x2*y1*z3 - x2*y3*z1 - x3*y1*z2 + x3*y2*z1])
    b = np.array([-y2*z3 + y2*z4 + y3*z2 - y3*z4 - y4*z2 + y4*z3,
                  y1*z3 - y1*z4 - y3*z1 + y3*z4 + y4*z1 - y4*z3,
                  -y1*z2 + y1*z4 + y2*z1 - y2*z4 - y4*z1 + y4*z2,
                  y1*z2 - y1*z3 - y2*z1 + y2*z3 + y3*z1 - y3*z2])
    c = np.array([x2*z3 - x2*z4 - x3*z2 + x3*z4 + x4*z2 - x4*z3,
                  -x1*z3 + x1*z4 + x3*z1 - x3*z4 - x4*z1 + x4*z3,
                  x1*z2 - x1*z4 - x2*z1 + x2*z4 + x4*z1 - x4*z2,
                  -x1*z2 + x1*z3 + x2*z1 - x2*z3 - x3*z1 + x3*z2])
    d = np.array([-x2*y3 + x2*y4 + x3*y2 - x3*y4 - x4*y2 + x4*y3,
                  x1*y3 - x1*y4 - x3*y1 + x3*y4 + x4*y1 - x4*y3,
                  -x1*y2 + x1*y4 + x2*y1 - x2*y4 - x4*y1 + x4*y2,
                  x1*y2 - x1*y3 - x2*y1 + x2*y3 + x3*y1 - x3*y2])
    return a, b, c, d, V


def basis_values(pts, xv, yv, zv, lem, ltm):
    """Values (20, 3, P) of the 20 basis functions at points pts (3, P), with
    the DOF order of the element matrices ([edge1 x6, face1 x4, edge2 x6,
    face2 x4]); lem (2,6) / ltm (3,4) are the local edge / face maps."""
    a, b, c, d, V = tet_coefficients(xv, yv, zv)
    x, y, z = pts
    V1 = 216 * V**3
    F = a[:, None] + b[:, None]*x + c[:, None]*y + d[:, None]*z  # (4,P)
    grad = np.stack([b, c, d], axis=1)  # (4,3)
    out = np.zeros((20, 3, pts.shape[1]))
    for ie in range(6):
        i1, i2 = lem[:, ie]
        L = np.sqrt((xv[i1]-xv[i2])**2 + (yv[i1]-yv[i2])**2 + (zv[i1]-zv[i2])**2)
        w = grad[i1][:, None]*F[i2] - grad[i2][:, None]*F[i1]  # (3,P)
        out[ie] = L * F[i1] * w / V1
        out[ie + 10] = L * F[i2] * w / V1
    for it in range(4):
        i1, i2, i3 = ltm[:, it]
        L1 = np.sqrt((xv[i1]-xv[i3])**2 + (yv[i1]-yv[i3])**2 + (zv[i1]-zv[i3])**2)
        L2 = np.sqrt((xv[i1]-xv[i2])**2 + (yv[i1]-yv[i2])**2 + (zv[i1]-zv[i2])**2)
        w13 = grad[i1][:, None]*F[i3] - grad[i3][:, None]*F[i1]
        w12 = grad[i1][:, None]*F[i2] - grad[i2][:, None]*F[i1]
        out[it + 6] = -L1 * F[i2] * w13 / V1
        out[it + 16] = L2 * F[i3] * w12 / V1
    return out


def _gauss_tet(n: int):
    """Collapsed Gauss-Legendre rule on the unit tet: points (3,P), weights (P,)."""
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1.0)
    w = 0.5 * w
    U, Vv, W = np.meshgrid(g, g, g, indexing="ij")
    WU, WV, WW = np.meshgrid(w, w, w, indexing="ij")
    xi = U
    eta = Vv * (1 - U)
    zeta = W * (1 - U) * (1 - Vv)
    jac = (1 - U)**2 * (1 - Vv)
    pts = np.stack([xi.ravel(), eta.ravel(), zeta.ravel()])
    wts = (WU * WV * WW * jac).ravel()
    return pts, wts


def element_matrices_quadrature(verts, tet_vertex_ids, edges_g, tris_g, Ms, Mm, n: int = 6,
                                step: float = 1e-4):
    """Curl-curl and mass 20x20 matrices of one tet by quadrature.

    verts (3,4); tet_vertex_ids (4,); edges_g (2,6) / tris_g (3,4) hold the
    global vertex ids of the tet's edges / faces (sorted); Ms = inv(mu_r),
    Mm = eps_r (3,3) complex.
    """
    lem = local_mapping(tet_vertex_ids[:, None], edges_g[:, :, None])[:, :, 0]
    ltm = local_mapping(tet_vertex_ids[:, None], tris_g[:, :, None])[:, :, 0]
    ref_pts, wts = _gauss_tet(n)
    v0 = verts[:, 0]
    J = np.stack([verts[:, 1] - v0, verts[:, 2] - v0, verts[:, 3] - v0], axis=1)
    detJ = abs(np.linalg.det(J))
    pts = v0[:, None] + J @ ref_pts
    xv, yv, zv = verts
    N = basis_values(pts, xv, yv, zv, lem, ltm)  # (20,3,P)
    h = step * np.max(np.abs(J))
    curl = np.zeros_like(N)
    dN = []
    for ax in range(3):
        e = np.zeros((3, 1))
        e[ax] = h
        dN.append((basis_values(pts + e, xv, yv, zv, lem, ltm)
                   - basis_values(pts - e, xv, yv, zv, lem, ltm)) / (2 * h))
    # curl = (dNz/dy - dNy/dz, dNx/dz - dNz/dx, dNy/dx - dNx/dy)
    curl[:, 0] = dN[1][:, 2] - dN[2][:, 1]
    curl[:, 1] = dN[2][:, 0] - dN[0][:, 2]
    curl[:, 2] = dN[0][:, 1] - dN[1][:, 0]
    w = wts * detJ
    MN = np.einsum("ij,ajp->aip", Mm, N)
    SC = np.einsum("ij,ajp->aip", Ms, curl)
    M = np.einsum("aip,bip,p->ab", N, MN, w)
    K = np.einsum("aip,bip,p->ab", curl, SC, w)
    return K, M
