"""Oracle: Robin / port surface terms, restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ASM/robinbc.py (``ASM/`` = src/emerge/_emerge/physics/microwave/
assembly/) and the triangle quadrature of E/mth/optimized.py:26-118.
Pinned by 2-D quadrature of the tangential basis (tests) and closed-form area
coefficients.
"""
from __future__ import annotations

import numpy as np

_FACT = [1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880]

# Dunavant table rows (multiplicity, weight, l1, l2, l3): E/mth/optimized.py:26-60
_GAUSQUADTRI = {
    1: [(1, 1, 1/3, 1/3, 1/3)],
    2: [(3, 1/3, 2/3, 1/6, 1/6)],
    3: [(1, -0.562500000000000, 1/3, 1/3, 1/3),
        (3, 0.520833333333333, 0.6, 0.2, 0.2)],
    4: [(3, 0.223381589678011, 0.108103018168070, 0.445948490915965, 0.445948490915965),
        (3, 0.109951743655322, 0.816847572980459, 0.091576213509771, 0.091576213509771)],
    5: [(1, 0.225000000000000, 0.333333333333333, 0.333333333333333, 0.333333333333333),
        (3, 0.132394152788506, 0.059715871789770, 0.470142064105115, 0.470142064105115),
        (3, 0.125939180544827, 0.797426985353087, 0.101286507323456, 0.101286507323456)],
}


def gaus_quad_tri(p: int) -> np.ndarray:
    """E/mth/optimized.py:87-118: points (4, n) rows W, L1, L2, L3; cyclic
    permutations of each orbit, the 4th point of a 6-orbit swaps L2/L3."""
    pts = []
    for n, w, l1, l2, l3 in _GAUSQUADTRI[p]:
        a, b, c = l1, l2, l3
        for k in range(n):
            if k == 3:
                a, b, c = l1, l3, l2
            pts.append([w, a, b, c])
            a, b, c = b, c, a
    return np.array(pts).T


def area_coeff(a, b, c, d) -> float:
    """ASM/robinbc.py:36-45: 2 prod(k_i!)/(sum k + 2)!."""
    k = [0] * 7
    for x in (a, b, c, d):
        k[x] += 1
    num = 1
    for i in range(1, 7):
        num *= _FACT[k[i]]
    return 2 * num / _FACT[sum(k[1:]) + 2]


AREA_COEFF_CACHE_BASE = np.zeros((5, 5, 5, 5))
for _i in range(5):
    for _j in range(5):
        for _k in range(5):
            for _l in range(5):
                AREA_COEFF_CACHE_BASE[_i, _j, _k, _l] = area_coeff(_i, _j, _k, _l)


def _normalize(a):
    return a / ((a[0]**2 + a[1]**2 + a[2]**2)**0.5)


def _cross(a, b):
    return np.array([a[1]*b[2] - a[2]*b[1], a[2]*b[0] - a[0]*b[2], a[0]*b[1] - a[1]*b[0]])


def _local_frame(glob_vertices):
    """ASM/robinbc.py:135-148: origin at vertex 1, x along edge 12, z normal."""
    orig = glob_vertices[:, 0]
    e1 = glob_vertices[:, 1] - orig
    e2 = glob_vertices[:, 2] - orig
    zhat = _normalize(_cross(e1, e2))
    xhat = _normalize(e1)
    yhat = _normalize(_cross(zhat, xhat))
    basis = np.zeros((3, 3))
    basis[0, :] = xhat
    basis[1, :] = yhat
    basis[2, :] = zhat
    return orig, basis


def _matmul3(B, data):
    out = np.zeros_like(data)
    out[0] = B[0, 0]*data[0] + B[0, 1]*data[1] + B[0, 2]*data[2]
    out[1] = B[1, 0]*data[0] + B[1, 1]*data[1] + B[1, 2]*data[2]
    out[2] = B[2, 0]*data[0] + B[2, 1]*data[1] + B[2, 2]*data[2]
    return out


def _dot2(a, b):
    return a[0]*b[0] + a[1]*b[1]


def _dist(xs, ys):
    D = np.empty((3, 3))
    for i in range(3):
        for j in range(i, 3):
            D[i, j] = np.sqrt((xs[i]-xs[j])**2 + (ys[i]-ys[j])**2 + 0.0)
            D[j, i] = D[i, j]
    return D


def _stiff_block(xs, ys, Area, AC):
    """The 8x8 matrix of ned2_tri_stiff / ned2_tri_stiff_force before the
    gamma/(2A)^2 factor (ASM/robinbc.py:199-244, 265-272)."""
    x1, x2, x3 = xs
    y1, y2, y3 = ys
    b1, b2, b3 = y2 - y3, y3 - y1, y1 - y2
    c1, c2, c3 = x3 - x2, x1 - x3, x2 - x1
    GLs = (np.array([b1, c1]), np.array([b2, c2]), np.array([b3, c3]))
    Ds = _dist(xs, ys)
    lem = np.array([[0, 1, 0], [1, 2, 2]])
    tA, tB, tC = 1, 2, 3
    GtA, GtB, GtC = GLs
    Lt1, Lt2 = Ds[2, 0], Ds[1, 0]
    M = np.zeros((8, 8), dtype=np.complex128)
    for ei in range(3):
        ei1, ei2 = lem[:, ei]
        Li = Ds[ei1, ei2]
        A, B = ei1 + 1, ei2 + 1
        GA, GB = GLs[ei1], GLs[ei2]
        for ej in range(3):
            ej1, ej2 = lem[:, ej]
            Lj = Ds[ej1, ej2]
            C, D = ej1 + 1, ej2 + 1
            GC, GD = GLs[ej1], GLs[ej2]
            DAC, DAD, DBC, DBD = _dot2(GA, GC), _dot2(GA, GD), _dot2(GB, GC), _dot2(GB, GD)
            LL = Li * Lj
            M[ei, ej] += LL*(AC[A, B, C, D]*DAC - AC[A, B, C, C]*DAD - AC[A, A, C, D]*DBC + AC[A, A, C, C]*DBD)
            M[ei, ej+4] += LL*(AC[A, B, D, D]*DAC - AC[A, B, C, D]*DAD - AC[A, A, D, D]*DBC + AC[A, A, C, D]*DBD)
            M[ei+4, ej] += LL*(AC[B, B, C, D]*DAC - AC[B, B, C, C]*DAD - AC[A, B, C, D]*DBC + AC[A, B, C, C]*DBD)
            M[ei+4, ej+4] += LL*(AC[B, B, D, D]*DAC - AC[B, B, C, D]*DAD - AC[A, B, D, D]*DBC + AC[A, B, C, D]*DBD)
        FA, FB, FC = _dot2(GA, GtC), _dot2(GA, GtA), _dot2(GB, GtC)
        FD, FE, FF = _dot2(GB, GtA), _dot2(GA, GtB), _dot2(GB, GtB)
        M[ei, 3] += Li*Lt1*(AC[A, B, tA, tB]*FA - AC[A, B, tB, tC]*FB - AC[A, A, tA, tB]*FC + AC[A, A, tB, tC]*FD)
        M[ei, 7] += Li*Lt2*(AC[A, B, tB, tC]*FB - AC[A, B, tC, tA]*FE - AC[A, A, tB, tC]*FD + AC[A, A, tC, tA]*FF)
        M[3, ei] += Lt1*Li*(AC[tA, tB, A, B]*FA - AC[tA, tB, A, A]*FC - AC[tB, tC, A, B]*FB + AC[tB, tC, A, A]*FD)
        M[7, ei] += Lt2*Li*(AC[tB, tC, A, B]*FB - AC[tB, tC, A, A]*FD - AC[tC, tA, A, B]*FE + AC[tC, tA, A, A]*FF)
        M[ei+4, 3] += Li*Lt1*(AC[B, B, tA, tB]*FA - AC[B, B, tB, tC]*FB - AC[A, B, tA, tB]*FC + AC[A, B, tB, tC]*FD)
        M[ei+4, 7] += Li*Lt2*(AC[B, B, tB, tC]*FB - AC[B, B, tC, tA]*FE - AC[A, B, tB, tC]*FD + AC[A, B, tC, tA]*FF)
        M[3, ei+4] += Lt1*Li*(AC[tA, tB, B, B]*FA - AC[tA, tB, A, B]*FC - AC[tB, tC, B, B]*FB + AC[tB, tC, A, B]*FD)
        M[7, ei+4] += Lt2*Li*(AC[tB, tC, B, B]*FB - AC[tB, tC, A, B]*FD - AC[tC, tA, B, B]*FE + AC[tC, tA, A, B]*FF)
    H1, H2, H3 = _dot2(GtA, GtC), _dot2(GtA, GtA), _dot2(GtA, GtB)
    M[3, 3] += Lt1*Lt1*(AC[tA, tB, tA, tB]*_dot2(GtC, GtC) - AC[tA, tB, tB, tC]*H1 - AC[tB, tC, tA, tB]*H1
                        + AC[tB, tC, tB, tC]*H2)
    M[3, 7] += Lt1*Lt2*(AC[tA, tB, tB, tC]*H1 - AC[tA, tB, tC, tA]*_dot2(GtB, GtC) - AC[tB, tC, tB, tC]*H2
                        + AC[tB, tC, tC, tA]*H3)
    M[7, 3] += Lt2*Lt1*(AC[tB, tC, tA, tB]*H1 - AC[tB, tC, tB, tC]*H2 - AC[tC, tA, tA, tB]*_dot2(GtB, GtC)
                        + AC[tC, tA, tB, tC]*H3)
    M[7, 7] += Lt2*Lt2*(AC[tB, tC, tB, tC]*H2 - AC[tB, tC, tC, tA]*H3 - AC[tC, tA, tB, tC]*H3
                        + AC[tC, tA, tC, tA]*_dot2(GtB, GtB))
    return M


def ned2_tri_stiff(glob_vertices, gamma):
    """ASM/robinbc.py:314-427: gamma * int N_i^t . N_j^t dS on one tri."""
    orig, basis = _local_frame(glob_vertices)
    lcs = _matmul3(basis, glob_vertices - orig[:, None])
    xs, ys = lcs[0], lcs[1]
    x1, x2, x3 = xs
    y1, y2, y3 = ys
    Area = 0.5 * np.abs((x1 - x3)*(y2 - y1) - (x1 - x2)*(y3 - y1))
    M = _stiff_block(xs, ys, Area, AREA_COEFF_CACHE_BASE * Area)
    return M * (gamma / (2 * Area)**2)


def ned2_tri_stiff_force(glob_vertices, gamma, glob_Uinc, DPTs):
    """ASM/robinbc.py:126-292: the same 8x8 matrix plus the forcing vector
    b_i = int N_i^t . U_inc dS (6-point Dunavant rule, signA orientation)."""
    orig, basis = _local_frame(glob_vertices)
    lcs = _matmul3(basis, glob_vertices - orig[:, None])
    lU = _matmul3(basis, glob_Uinc)
    xs, ys = lcs[0], lcs[1]
    x1, x2, x3 = xs
    y1, y2, y3 = ys
    a1, a2, a3 = x2*y3 - y2*x3, x3*y1 - y3*x1, x1*y2 - y1*x2
    b1, b2, b3 = y2 - y3, y3 - y1, y1 - y2
    c1, c2, c3 = x3 - x2, x1 - x3, x2 - x1
    As, Bs, Cs = np.array([a1, a2, a3]), np.array([b1, b2, b3]), np.array([c1, c2, c3])
    Ds = _dist(xs, ys)
    Area = 0.5 * np.abs((x1 - x3)*(y2 - y1) - (x1 - x2)*(y3 - y1))
    signA = -np.sign((x1 - x3)*(y2 - y1) - (x1 - x2)*(y3 - y1))
    Lt1, Lt2 = Ds[2, 0], Ds[1, 0]
    Ux, Uy = lU[0, :], lU[1, :]
    x = x1*DPTs[1, :] + x2*DPTs[2, :] + x3*DPTs[3, :]
    y = y1*DPTs[1, :] + y2*DPTs[2, :] + y3*DPTs[3, :]
    Ws = DPTs[0, :]
    M = _stiff_block(xs, ys, Area, AREA_COEFF_CACHE_BASE * Area)
    bvec = np.zeros(8, dtype=np.complex128)
    lem = np.array([[0, 1, 0], [1, 2, 2]])
    for ei in range(3):
        ei1, ei2 = lem[:, ei]
        Li = Ds[ei1, ei2]
        A1, A2 = As[ei1], As[ei2]
        B1, B2 = Bs[ei1], Bs[ei2]
        C1, C2 = Cs[ei1], Cs[ei2]
        Q = A2 + B2*x + C2*y
        Z = A1 + B1*x + C1*y
        A4 = 4 * Area**2
        Q2, Z2 = Q / A4, Z / A4
        Ar2 = 1 / (2 * Area)
        Ee1x = (B1*Q2 - B2*Z2)*Z*Ar2
        Ee1y = (C1*Q2 - C2*Z2)*Z*Ar2
        Ee2x = (B1*Q2 - B2*Z2)*Q*Ar2
        Ee2y = (C1*Q2 - C2*Z2)*Q*Ar2
        bvec[ei] += signA*Area*Li*np.sum(Ws*(Ee1x*Ux + Ee1y*Uy))
        bvec[ei + 4] += signA*Area*Li*np.sum(Ws*(Ee2x*Ux + Ee2y*Uy))
    A1, A2, A3 = As
    B1, B2, B3 = Bs
    C1, C2, C3 = Cs
    Q = A2 + B2*x + C2*y
    Z = A1 + B1*x + C1*y
    FA = 8 * Area**3
    W = (A3 + B3*x + C3*y) / FA
    W2 = Q * W
    Ef1x = Lt1*(-B1*W2 + B3*Z*Q/FA)
    Ef1y = Lt1*(-C1*W2 + C3*Z*Q/FA)
    Ef2x = Lt2*(B1*W2 - B2*Z*W)
    Ef2y = Lt2*(C1*W2 - C2*Z*W)
    bvec[3] += signA*Area*np.sum(Ws*(Ef1x*Ux + Ef1y*Uy))
    bvec[7] += signA*Area*np.sum(Ws*(Ef2x*Ux + Ef2y*Uy))
    return M * (gamma / (2 * Area)**2), bvec


def generate_points_3d(nodes, tris, DPTs, tri_ids):
    """ASM/robinbc.py:90-110: quadrature points, flattened point-major."""
    NS = tri_ids.shape[0]
    xall = np.zeros((DPTs.shape[1], NS))
    yall = np.zeros_like(xall)
    zall = np.zeros_like(xall)
    for i in range(NS):
        vid = tris[:, tri_ids[i]]
        x1, x2, x3 = nodes[0, vid]
        y1, y2, y3 = nodes[1, vid]
        z1, z2, z3 = nodes[2, vid]
        xall[:, i] = x1*DPTs[1] + x2*DPTs[2] + x3*DPTs[3]
        yall[:, i] = y1*DPTs[1] + y2*DPTs[2] + y3*DPTs[3]
        zall[:, i] = z1*DPTs[1] + z2*DPTs[2] + z3*DPTs[3]
    return xall.flatten(), yall.flatten(), zall.flatten()


def assemble_robin_bc_excited(nodes, tris, tri_to_field, n_field, Bmat, tri_ids, Ufunc, gamma, DPTs):
    """ASM/robinbc.py:445-463 + compute_bc_entries_excited (:294-311)."""
    Bvec = np.zeros(n_field, dtype=np.complex128)
    xf, yf, zf = generate_points_3d(nodes, tris, DPTs, tri_ids)
    U = np.asarray(Ufunc(xf, yf, zf)).reshape((3, DPTs.shape[1], tri_ids.shape[0]))
    for i, itri in enumerate(tri_ids):
        Bsub, bvec = ned2_tri_stiff_force(nodes[:, tris[:, itri]], gamma, U[:, :, i], DPTs)
        Bmat[itri*64:(itri + 1)*64] += Bsub.ravel()
        Bvec[tri_to_field[:, itri]] += bvec
    return Bmat, Bvec


def assemble_robin_bc(nodes, tris, Bmat, tri_ids, gamma):
    """ASM/robinbc.py:465-473 + compute_bc_entries (:429-443)."""
    for itri in tri_ids:
        Bsub = ned2_tri_stiff(nodes[:, tris[:, itri]], gamma)
        Bmat[itri*64:(itri + 1)*64] = Bmat[itri*64:(itri + 1)*64] + Bsub.ravel()
    return Bmat
