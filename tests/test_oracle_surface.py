"""Pins of the oracle's boundary-triangle terms against brute-force
quadrature of the tangential Nedelec-2 basis as the reference's 2-D
interpolation defines it (E/elements/ned2_interp.py:509-596):

* Robin / port matrix (oracle.robin.ned2_tri_stiff): int N_i . N_j dS;
* second-order ABC correction (oracle.abc2.abc_order_2_terms):
  int (curl N_i curl N_j - div N_i div N_j) dS, curl and div taken by
  central differences.
"""
import numpy as np
import pytest

from oracle import abc2, robin


def _tri_coefficients(xv, yv):
    x1, x2, x3 = xv
    y1, y2, y3 = yv
    a = np.array([x2*y3 - y2*x3, x3*y1 - y3*x1, x1*y2 - y1*x2])
    b = np.array([y2 - y3, y3 - y1, y1 - y2])
    c = np.array([x3 - x2, x1 - x3, x2 - x1])
    sA = 0.5*((x1 - x3)*(y2 - y1) - (x1 - x2)*(y3 - y1))
    s = np.sign(sA)
    return a*s, b*s, c*s, abs(sA)


def basis(xv, yv, x, y):
    """(8, 2, n) values of the tangential basis on the triangle with local
    vertices (xv, yv), DOF order [edge1 x3 | face1 | edge2 x3 | face2], edges
    (0,1), (1,2), (0,2), length factors included (ned2_tri_interp)."""
    a, b, c, A = _tri_coefficients(xv, yv)
    lam = [a[i] + b[i]*x + c[i]*y for i in range(3)]   # 2A * barycentric
    D = np.sqrt((xv[:, None] - xv[None, :])**2 + (yv[:, None] - yv[None, :])**2)
    L1, L2, L3 = D[0, 1], D[1, 2], D[0, 2]
    mult = [L1, L2, L3, L3, L1, L2, L3, L1]
    out = np.zeros((8, 2) + np.shape(x))
    for ie, (e1, e2) in enumerate(((0, 1), (1, 2), (0, 2))):
        wx = b[e1]*lam[e2] - b[e2]*lam[e1]
        wy = c[e1]*lam[e2] - c[e2]*lam[e1]
        out[ie] = np.array([lam[e1]*wx, lam[e1]*wy]) / (8*A**3)
        out[ie + 4] = np.array([lam[e2]*wx, lam[e2]*wy]) / (8*A**3)
    out[3] = -np.array([(b[0]*lam[2] - b[2]*lam[0])*lam[1], (c[0]*lam[2] - c[2]*lam[0])*lam[1]]) / (8*A**3)
    out[7] = np.array([(b[0]*lam[1] - b[1]*lam[0])*lam[2], (c[0]*lam[1] - c[1]*lam[0])*lam[2]]) / (8*A**3)
    for i in range(8):
        out[i] *= mult[i]
    return out, A


def tri_rule(xv, yv, n=10):
    """Collapsed Gauss-Legendre rule on the triangle (exact to degree ~2n-2)."""
    g, w = np.polynomial.legendre.leggauss(n)
    g, w = (g + 1) / 2, w / 2
    U, Vv = np.meshgrid(g, g, indexing="ij")
    WU, WV = np.meshgrid(w, w, indexing="ij")
    s, t = U, Vv * (1 - U)
    wt = (WU * WV * (1 - U)).ravel()
    s, t = s.ravel(), t.ravel()
    x = xv[0] + s*(xv[1] - xv[0]) + t*(xv[2] - xv[0])
    y = yv[0] + s*(yv[1] - yv[0]) + t*(yv[2] - yv[0])
    area2 = abs((xv[1] - xv[0])*(yv[2] - yv[0]) - (xv[2] - xv[0])*(yv[1] - yv[0]))
    return x, y, wt * area2


def _triangles():
    rng = np.random.default_rng(7)
    out = [np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]]) * 1e-3]
    for _ in range(4):
        v = rng.uniform(-1, 1, (3, 3)) * 2e-3 + rng.uniform(-1, 1, (3, 1)) * 1e-2
        out.append(v)
    return out


@pytest.mark.parametrize("k", range(5))
def test_robin_matrix_is_the_basis_mass_matrix(k):
    tv = _triangles()[k]
    lv = abc2.local_frame(tv)
    xv, yv = lv[0], lv[1]
    x, y, w = tri_rule(xv, yv)
    N, _ = basis(xv, yv, x, y)
    ref = np.einsum("iaq,jaq,q->ij", N, N, w)
    M = robin.ned2_tri_stiff(tv, 1.0)
    assert np.abs(M - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("k", range(5))
def test_abc2_matrix_is_curl_minus_div(k):
    tv = _triangles()[k]
    lv = abc2.local_frame(tv)
    xv, yv = lv[0], lv[1]
    x, y, w = tri_rule(xv, yv)
    h = 1e-6 * np.sqrt(np.abs(xv).max()**2 + np.abs(yv).max()**2)
    Nxp, _ = basis(xv, yv, x + h, y)
    Nxm, _ = basis(xv, yv, x - h, y)
    Nyp, _ = basis(xv, yv, x, y + h)
    Nym, _ = basis(xv, yv, x, y - h)
    dx = (Nxp - Nxm) / (2*h)
    dy = (Nyp - Nym) / (2*h)
    curl = dx[:, 1] - dy[:, 0]
    div = dx[:, 0] + dy[:, 1]
    ref = np.einsum("iq,jq,q->ij", curl, curl, w) - np.einsum("iq,jq,q->ij", div, div, w)
    M = abc2.abc_order_2_terms(tv, 1.0)
    # the reference integrates with its 8-digit 6-point table
    assert np.abs(M - ref).max() <= 2e-7 * np.abs(ref).max()


def test_abc2_detects_a_dropped_term():
    """A planted error (divergence of the first face function zeroed) is
    far outside the tolerance above."""
    tv = _triangles()[1]
    lv = abc2.local_frame(tv)
    M = abc2.abc_order_2_terms(tv, 1.0)
    saved = abc2._div_face_1
    try:
        abc2._div_face_1 = lambda coeff, xs, ys: 0 * xs
        Mb = abc2.abc_order_2_terms(tv, 1.0)
    finally:
        abc2._div_face_1 = saved
    assert np.abs(M - Mb).max() > 1e-3 * np.abs(M).max()
