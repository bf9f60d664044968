"""Oracle: second-order absorbing boundary correction, restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ASM/robin_abc_order2.py (``ASM/`` = src/emerge/_emerge/physics/
microwave/assembly/): per boundary triangle the 8x8 matrix
``cf * L_i L_j * (Curl_ij - Div_ij) * |Area|`` with the surface curl and
divergence of the tangential Nedelec-2 basis integrated by the module's own
6-point table (rounded to 8 digits, ``DPTS`` at :181-184), and the call of
ASM/assembler.py:342-348 (``cf = 1j * c2 / k0``).

Pinned (tests/test_oracle_surface.py) against brute-force quadrature of the
curl and divergence, taken by finite differences, of the tangential basis as
the reference's 2-D interpolation defines it (E/elements/ned2_interp.py:
509-596): agreement to the rounding of the 8-digit table.
"""
from __future__ import annotations

import numpy as np

# ASM/robin_abc_order2.py:181-184: weights, then the three barycentric rows
DPTS = np.array([[0.22338159, 0.22338159, 0.22338159, 0.10995174, 0.10995174, 0.10995174],
                 [0.10810302, 0.44594849, 0.44594849, 0.81684757, 0.09157621, 0.09157621],
                 [0.44594849, 0.44594849, 0.10810302, 0.09157621, 0.09157621, 0.81684757],
                 [0.44594849, 0.10810302, 0.44594849, 0.09157621, 0.81684757, 0.09157621]], dtype=np.float64)


def _curl_edge_1(coeff, xs, ys):  # :70-77
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    return -3*a1*b1*c2 + 3*a1*b2*c1 - 3*b1**2*c2*xs + 3*b1*b2*c1*xs - 3*b1*c1*c2*ys + 3*b2*c1**2*ys


def _curl_edge_2(coeff, xs, ys):  # :79-86
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    return -3*a2*b1*c2 + 3*a2*b2*c1 - 3*b1*b2*c2*xs - 3*b1*c2**2*ys + 3*b2**2*c1*xs + 3*b2*c1*c2*ys


def _curl_face_1(coeff, xs, ys):  # :88-96
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    a3, b3, c3 = coeff[:, 2]
    return (-b2*(c1*(a3 + b3*xs + c3*ys) - c3*(a1 + b1*xs + c1*ys))
            + c2*(b1*(a3 + b3*xs + c3*ys) - b3*(a1 + b1*xs + c1*ys))
            + 2*(b1*c3 - b3*c1)*(a2 + b2*xs + c2*ys))


def _curl_face_2(coeff, xs, ys):  # :98-106
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    a3, b3, c3 = coeff[:, 2]
    return (b3*(c1*(a2 + b2*xs + c2*ys) - c2*(a1 + b1*xs + c1*ys))
            - c3*(b1*(a2 + b2*xs + c2*ys) - b2*(a1 + b1*xs + c1*ys))
            - 2*(b1*c2 - b2*c1)*(a3 + b3*xs + c3*ys))


def _div_edge_1(coeff, xs, ys):  # :108-115
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    return (b1*(b1*(a2 + b2*xs + c2*ys) - b2*(a1 + b1*xs + c1*ys))
            + c1*(c1*(a2 + b2*xs + c2*ys) - c2*(a1 + b1*xs + c1*ys)))


def _div_edge_2(coeff, xs, ys):  # :117-124
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    return (b2*(b1*(a2 + b2*xs + c2*ys) - b2*(a1 + b1*xs + c1*ys))
            + c2*(c1*(a2 + b2*xs + c2*ys) - c2*(a1 + b1*xs + c1*ys)))


def _div_face_1(coeff, xs, ys):  # :126-134
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    a3, b3, c3 = coeff[:, 2]
    return (-b2*(b1*(a3 + b3*xs + c3*ys) - b3*(a1 + b1*xs + c1*ys))
            - c2*(c1*(a3 + b3*xs + c3*ys) - c3*(a1 + b1*xs + c1*ys)))


def _div_face_2(coeff, xs, ys):  # :136-144
    a1, b1, c1 = coeff[:, 0]
    a2, b2, c2 = coeff[:, 1]
    a3, b3, c3 = coeff[:, 2]
    return (b3*(b1*(a2 + b2*xs + c2*ys) - b2*(a1 + b1*xs + c1*ys))
            + c3*(c1*(a2 + b2*xs + c2*ys) - c2*(a1 + b1*xs + c1*ys)))


def tri_coefficients(vxs, vys):
    """:151-176: barycentric linear coefficients (sign-normalised) and area."""
    x1, x2, x3 = vxs
    y1, y2, y3 = vys
    a = np.array([x2*y3 - y2*x3, x3*y1 - y3*x1, x1*y2 - y1*x2])
    b = np.array([y2 - y3, y3 - y1, y1 - y2])
    c = np.array([x3 - x2, x1 - x3, x2 - x1])
    sA = 0.5*((x1 - x3)*(y2 - y1) - (x1 - x2)*(y3 - y1))
    s = np.sign(sA)
    return a*s, b*s, c*s, abs(sA)


def local_frame(tri_vertices):
    """:194-210: origin at vertex 0, x along edge 0->1, z the normal."""
    o = tri_vertices[:, 0]
    e1 = tri_vertices[:, 1] - o
    e2 = tri_vertices[:, 2] - o
    zh = np.cross(e1, e2)
    zh = zh / np.linalg.norm(zh)
    xh = e1 / np.linalg.norm(e1)
    yh = np.cross(zh, xh)
    yh = yh / np.linalg.norm(yh)
    basis = np.stack([xh, yh, zh])
    return basis @ (tri_vertices - o[:, None])


def abc_order_2_terms(tri_vertices, cf, local_edge_map=((0, 1, 0), (1, 2, 2))):
    """_abc_order_2_terms (:188-311) for one triangle (vertices (3, 3),
    columns in the triangle's stored order)."""
    lem = np.asarray(local_edge_map)
    lv = local_frame(tri_vertices)
    xpts, ypts = lv[0], lv[1]
    D = np.sqrt((xpts[:, None] - xpts[None, :])**2 + (ypts[:, None] - ypts[None, :])**2)
    W = DPTS[0]
    xs = xpts[0]*DPTS[1] + xpts[1]*DPTS[2] + xpts[2]*DPTS[3]
    ys = ypts[0]*DPTS[1] + ypts[1]*DPTS[2] + ypts[2]*DPTS[3]
    a, b, c, A = tri_coefficients(xpts, ypts)
    bc = np.stack([a/(2*A), b/(2*A), c/(2*A)])
    L = np.ones((8, 8))
    L[3, :] *= D[0, 2]
    L[7, :] *= D[0, 1]
    L[:, 3] *= D[0, 2]
    L[:, 7] *= D[0, 1]
    curl = np.zeros((8, len(W)))
    div = np.zeros((8, len(W)))
    curl[3] = _curl_face_1(bc, xs, ys)
    curl[7] = _curl_face_2(bc, xs, ys)
    div[3] = _div_face_1(bc, xs, ys)
    div[7] = _div_face_2(bc, xs, ys)
    for iv in range(3):
        ie = lem[:, iv]
        Le = D[ie[0], ie[1]]
        L[iv, :] *= Le
        L[iv + 4, :] *= Le
        L[:, iv] *= Le
        L[:, iv + 4] *= Le
        sub = bc[:, ie]
        curl[iv] = _curl_edge_1(sub, xs, ys)
        curl[iv + 4] = _curl_edge_2(sub, xs, ys)
        div[iv] = _div_edge_1(sub, xs, ys)
        div[iv + 4] = _div_edge_2(sub, xs, ys)
    C = (curl[:, None, :] * curl[None, :, :] * W).sum(-1)
    Dv = (div[:, None, :] * div[None, :, :] * W).sum(-1)
    return cf * L * (C - Dv) * abs(A)


def abc_order_2_matrix(nodes, tris, n_tris, tri_ids, coeff):
    """abc_order_2_matrix / _matrix_builder (:318-374): the (64 n_tris,)
    triangle-slot vector with each listed triangle's 8x8 block at 64 itri."""
    Mat = np.zeros(n_tris * 64, dtype=np.complex128)
    for itri in tri_ids:
        Mat[itri*64:itri*64 + 64] += abc_order_2_terms(nodes[:, tris[:, itri]], coeff).ravel()
    return Mat
