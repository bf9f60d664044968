"""Oracle: Nedelec-2 tet curl-curl (E) and mass (B) assembly, restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ASM/curlcurl.py (``ASM/`` = src/emerge/_emerge/physics/microwave/
assembly/) vectorised over tets with NumPy: the same closed form per entry,
the same local edge/face maps, the same COO layout and scipy's COO->CSR.

Pinning: ``ned2_tet_stiff_mass`` is checked against brute-force quadrature of
the basis as defined by the reference's interpolation routine
(``oracle.quadrature``), against symmetry, the gradient null space of the
curl-curl matrix and the h / h^3 scale laws; ``volume_coeff`` against exact
rational Dirichlet integrals; COO->CSR is scipy itself.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix

_FACTORIALS = np.array([1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880], dtype=np.int64)


def volume_coeff(a: int, b: int, c: int, d: int) -> float:
    """ASM/curlcurl.py:64-77: prod(k_i!)/(sum k + 3)! over letters 1..6; letter
    0 means "no factor"."""
    klmn = [0] * 7
    for x in (a, b, c, d):
        klmn[x] += 1
    num = 1
    for i in range(1, 7):
        num *= int(_FACTORIALS[klmn[i]])
    return num / int(_FACTORIALS[sum(klmn[1:]) + 3])


NFILL = 5
VOLUME_COEFF_CACHE_BASE = np.zeros((NFILL,) * 4, dtype=np.float64)  # ASM/curlcurl.py:84-92
for _i in range(NFILL):
    for _j in range(NFILL):
        for _k in range(NFILL):
            for _l in range(NFILL):
                VOLUME_COEFF_CACHE_BASE[_i, _j, _k, _l] = volume_coeff(_i, _j, _k, _l)


def tet_coefficients_bcd(xs, ys, zs):
    """ASM/curlcurl.py:99-138 vectorised: cofactors b,c,d (each (4,T)) and
    V = |det|/6 (T,).  xs/ys/zs are (4, T)."""
    x1, x2, x3, x4 = xs
    y1, y2, y3, y4 = ys
    z1, z2, z3, z4 = zs
    V = np.abs(-x1*y2*z3/6 + x1*y2*z4/6 + x1*y3*z2/6 - x1*y3*z4/6 - x1*y4*z2/6
               + x1*y4*z3/6 + x2*y1*z3/6 - x2*y1*z4/6 - x2*y3*z1/6 + x2*y3*z4/6
               + x2*y4*z1/6 - x2*y4*z3/6 - x3*y1*z2/6 + x3*y1*z4/6 + x3*y2*z1/6
               - x3*y2*z4/6 - x3*y4*z1/6 + x3*y4*z2/6 + x4*y1*z2/6 - x4*y1*z3/6
               - x4*y2*z1/6 + x4*y2*z3/6 + x4*y3*z1/6 - x4*y3*z2/6)
    b = np.stack([
        -y2*z3 + y2*z4 + y3*z2 - y3*z4 - y4*z2 + y4*z3,
        y1*z3 - y1*z4 - y3*z1 + y3*z4 + y4*z1 - y4*z3,
        -y1*z2 + y1*z4 + y2*z1 - y2*z4 - y4*z1 + y4*z2,
        y1*z2 - y1*z3 - y2*z1 + y2*z3 + y3*z1 - y3*z2])
    c = np.stack([
        x2*z3 - x2*z4 - x3*z2 + x3*z4 + x4*z2 - x4*z3,
        -x1*z3 + x1*z4 + x3*z1 - x3*z4 - x4*z1 + x4*z3,
        x1*z2 - x1*z4 - x2*z1 + x2*z4 + x4*z1 - x4*z2,
        -x1*z2 + x1*z3 + x2*z1 - x2*z3 - x3*z1 + x3*z2])
    d = np.stack([
        -x2*y3 + x2*y4 + x3*y2 - x3*y4 - x4*y2 + x4*y3,
        x1*y3 - x1*y4 - x3*y1 + x3*y4 + x4*y1 - x4*y3,
        -x1*y2 + x1*y4 + x2*y1 - x2*y4 - x4*y1 + x4*y2,
        x1*y2 - x1*y3 - x2*y1 + x2*y3 + x3*y1 - x3*y2])
    return b, c, d, V


def matinv(M: np.ndarray) -> np.ndarray:
    """E/mth/optimized.py:342-370 vectorised over the last axis of (3,3,T):
    diagonal fast path, else adjugate / det."""
    M = np.asarray(M, dtype=np.complex128)
    out = np.zeros_like(M)
    offdiag = (M[0, 1] == 0) & (M[0, 2] == 0) & (M[1, 0] == 0) & (M[1, 2] == 0) \
        & (M[2, 0] == 0) & (M[2, 1] == 0)
    dg = offdiag
    out[0, 0, dg] = 1 / M[0, 0, dg]
    out[1, 1, dg] = 1 / M[1, 1, dg]
    out[2, 2, dg] = 1 / M[2, 2, dg]
    g = ~dg
    if np.any(g):
        m = M[:, :, g]
        det = (m[0, 0]*m[1, 1]*m[2, 2] - m[0, 0]*m[1, 2]*m[2, 1] - m[0, 1]*m[1, 0]*m[2, 2]
               + m[0, 1]*m[1, 2]*m[2, 0] + m[0, 2]*m[1, 0]*m[2, 1] - m[0, 2]*m[1, 1]*m[2, 0])
        o = np.empty_like(m)
        o[0, 0] = m[1, 1]*m[2, 2] - m[1, 2]*m[2, 1]
        o[0, 1] = -m[0, 1]*m[2, 2] + m[0, 2]*m[2, 1]
        o[0, 2] = m[0, 1]*m[1, 2] - m[0, 2]*m[1, 1]
        o[1, 0] = -m[1, 0]*m[2, 2] + m[1, 2]*m[2, 0]
        o[1, 1] = m[0, 0]*m[2, 2] - m[0, 2]*m[2, 0]
        o[1, 2] = -m[0, 0]*m[1, 2] + m[0, 2]*m[1, 0]
        o[2, 0] = m[1, 0]*m[2, 1] - m[1, 1]*m[2, 0]
        o[2, 1] = -m[0, 0]*m[2, 1] + m[0, 1]*m[2, 0]
        o[2, 2] = m[0, 0]*m[1, 1] - m[0, 1]*m[1, 0]
        out[:, :, g] = o / det
    return out


def local_mapping(vertex_ids: np.ndarray, global_ids: np.ndarray) -> np.ndarray:
    """E/mth/optimized.py:293-327 vectorised: global vertex ids (ndim, n, T) ->
    tet-local index 0..3 given the tet's vertices (4, T)."""
    out = np.zeros(global_ids.shape, dtype=np.int64)
    for k in range(3, -1, -1):  # first match wins, as in the linear search
        out = np.where(global_ids == vertex_ids[k][None, None, :], k, out)
    return out


def _mv(M, v):
    """(3,3,T) @ (3,T) -> (3,T)"""
    return np.einsum("ijt,jt->it", M, v)


def _dot(a, b):
    return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]


def _cross(a, b):
    return np.stack([a[1]*b[2] - a[2]*b[1], a[2]*b[0] - a[0]*b[2], a[0]*b[1] - a[1]*b[0]])


def ned2_tet_stiff_mass(verts, edge_lengths, lem, ltm, Ms, Mm, exact_face_mass=False):
    """ASM/curlcurl.py:180-412 vectorised over T tets.

    verts (3,4,T) f64; edge_lengths (6,T); lem (2,6,T) / ltm (3,4,T) local
    edge/face maps; Ms = inv(mu_r) and Mm = eps_r as (3,3,T) complex.
    Returns Dmat (curl-curl) and Fmat (mass), each (20,20,T) complex128.

    ``exact_face_mass=True`` is for pins only: it replaces the reference's
    face-face monomial lam_B lam_C lam_D lam_F (ASM/curlcurl.py:361) by the
    integral's lam_B lam_C lam_D lam_E, which makes the mass matrix the
    exact Gram matrix of the basis (symmetric); the default reproduces the
    reference.
    """
    T = verts.shape[2]
    Dm = np.zeros((20, 20, T), dtype=np.complex128)
    Fm = np.zeros((20, 20, T), dtype=np.complex128)
    xs, ys, zs = verts[0], verts[1], verts[2]
    b, c, d, V = tet_coefficients_bcd(xs, ys, zs)
    Ds = np.sqrt((xs[:, None] - xs[None, :])**2 + (ys[:, None] - ys[None, :])**2
                 + (zs[:, None] - zs[None, :])**2)  # (4,4,T), optimized.py:330-338
    GL = np.stack([b, c, d], axis=1).astype(np.complex128)  # (4,3,T)
    KA = 1 / (6 * V)**4
    KB = 1 / (6 * V)**2
    V6 = 6 * V
    tix = np.arange(T)

    def G(idx):  # (T,) local vertex index -> (3,T) gradient vector
        return GL[idx, :, tix].T

    def VC(a, b_, c_, d_):  # letters are local index + 1; 0 = no factor
        return VOLUME_COEFF_CACHE_BASE[a, b_, c_, d_] * V6

    def pick(D, idx):  # D (4,4,T), per-tet indices
        return D[idx[0], idx[1], tix]

    for ei in range(6):
        ei1, ei2 = lem[0, ei], lem[1, ei]
        GA, GB = G(ei1), G(ei2)
        A, B = ei1 + 1, ei2 + 1
        L1 = edge_lengths[ei]
        for ej in range(6):
            ej1, ej2 = lem[0, ej], lem[1, ej]
            C, D = ej1 + 1, ej2 + 1
            GC, GD = G(ej1), G(ej2)
            L2 = edge_lengths[ej]
            mGC = _mv(Mm, GC)
            mGD = _mv(Mm, GD)
            g_ac = _dot(GA, mGC)
            g_ad = _dot(GA, mGD)
            g_bc = _dot(GB, mGC)
            g_bd = _dot(GB, mGD)
            L12 = L1 * L2
            factor = L12 * 9 * _dot(_cross(GA, GB), _mv(Ms, _cross(GC, GD)))
            Dm[ei, ej] = factor * VC(A, C, 0, 0)
            Dm[ei, ej + 10] = factor * VC(A, D, 0, 0)
            Dm[ei + 10, ej] = factor * VC(B, C, 0, 0)
            Dm[ei + 10, ej + 10] = factor * VC(B, D, 0, 0)
            Fm[ei, ej] = L12 * (VC(A, B, C, D)*g_ac - VC(A, B, C, C)*g_ad
                                - VC(A, A, C, D)*g_bc + VC(A, A, C, C)*g_bd)
            Fm[ei, ej + 10] = L12 * (VC(A, B, D, D)*g_ac - VC(A, B, C, D)*g_ad
                                     - VC(A, A, D, D)*g_bc + VC(A, A, C, D)*g_bd)
            Fm[ei + 10, ej] = L12 * (VC(B, B, C, D)*g_ac - VC(B, B, C, C)*g_ad
                                     - VC(A, B, C, D)*g_bc + VC(A, B, C, C)*g_bd)
            Fm[ei + 10, ej + 10] = L12 * (VC(B, B, D, D)*g_ac - VC(B, B, C, D)*g_ad
                                          - VC(A, B, D, D)*g_bc + VC(A, B, C, D)*g_bd)
        for ej in range(4):
            ej1, ej2, fj = ltm[0, ej], ltm[1, ej], ltm[2, ej]
            C, D, F = ej1 + 1, ej2 + 1, fj + 1
            GC, GD, GF = G(ej1), G(ej2), G(fj)
            Lab2 = pick(Ds, (ej1, ej2))
            Lac2 = pick(Ds, (ej1, fj))
            cab = _cross(GA, GB)
            h_df = _dot(cab, _mv(Ms, _cross(GC, GF)))
            h_cd = _dot(cab, _mv(Ms, _cross(GD, GF)))
            h_cf = _dot(cab, _mv(Ms, _cross(GC, GD)))
            mGF, mGC, mGD = _mv(Mm, GF), _mv(Mm, GC), _mv(Mm, GD)
            g_af, g_ac, g_ad = _dot(GA, mGF), _dot(GA, mGC), _dot(GA, mGD)
            g_bf, g_bc, g_bd = _dot(GB, mGF), _dot(GB, mGC), _dot(GB, mGD)
            Dm[ei, ej + 6] = L1*Lac2*(-6*VC(A, D, 0, 0)*h_df - 3*VC(A, C, 0, 0)*h_cd
                                      - 3*VC(A, F, 0, 0)*h_cf)
            Dm[ei, ej + 16] = L1*Lab2*(6*VC(A, F, 0, 0)*h_cf + 3*VC(A, D, 0, 0)*h_df
                                       - 3*VC(A, C, 0, 0)*h_cd)
            Dm[ei + 10, ej + 6] = L1*Lac2*(-6*VC(B, D, 0, 0)*h_df - 3*VC(B, C, 0, 0)*h_cd
                                           - 3*VC(B, F, 0, 0)*h_cf)
            Dm[ei + 10, ej + 16] = L1*Lab2*(6*VC(B, F, 0, 0)*h_cf + 3*VC(B, D, 0, 0)*h_df
                                            - 3*VC(B, C, 0, 0)*h_cd)
            Fm[ei, ej + 6] = L1*Lac2*(VC(A, B, C, D)*g_af - VC(A, B, D, F)*g_ac
                                      - VC(A, A, C, D)*g_bf + VC(A, A, D, F)*g_bc)
            Fm[ei, ej + 16] = L1*Lab2*(VC(A, B, D, F)*g_ac - VC(A, B, F, C)*g_ad
                                       - VC(A, A, D, F)*g_bc + VC(A, A, C, F)*g_bd)
            Fm[ei + 10, ej + 6] = L1*Lac2*(VC(B, B, C, D)*g_af - VC(B, B, D, F)*g_ac
                                           - VC(A, B, C, D)*g_bf + VC(A, B, D, F)*g_bc)
            Fm[ei + 10, ej + 16] = L1*Lab2*(VC(B, B, D, F)*g_ac - VC(B, B, F, C)*g_ad
                                            - VC(A, B, D, F)*g_bc + VC(A, B, F, C)*g_bd)

    # mirrored blocks (ASM/curlcurl.py:321-328)
    for X in (Dm, Fm):
        X[6:10, :6] = X[:6, 6:10].transpose(1, 0, 2)
        X[16:20, :6] = X[:6, 16:20].transpose(1, 0, 2)
        X[6:10, 10:16] = X[10:16, 6:10].transpose(1, 0, 2)
        X[16:20, 10:16] = X[10:16, 16:20].transpose(1, 0, 2)

    for ei in range(4):
        ei1, ei2, fi = ltm[0, ei], ltm[1, ei], ltm[2, ei]
        A, B, E = ei1 + 1, ei2 + 1, fi + 1
        GA, GB, GE = G(ei1), G(ei2), G(fi)
        Lac1 = pick(Ds, (ei1, fi))
        Lab1 = pick(Ds, (ei1, ei2))
        for ej in range(4):
            ej1, ej2, fj = ltm[0, ej], ltm[1, ej], ltm[2, ej]
            C, D, F = ej1 + 1, ej2 + 1, fj + 1
            GC, GD, GF = G(ej1), G(ej2), G(fj)
            Lac2 = pick(Ds, (ej1, fj))
            Lab2 = pick(Ds, (ej1, ej2))
            c_ae, c_be, c_ab = _cross(GA, GE), _cross(GB, GE), _cross(GA, GB)
            m_cf = _mv(Ms, _cross(GC, GF))
            m_df = _mv(Ms, _cross(GD, GF))
            m_cd = _mv(Ms, _cross(GC, GD))
            ae_cf, ae_df, ae_cd = _dot(c_ae, m_cf), _dot(c_ae, m_df), _dot(c_ae, m_cd)
            be_cf, be_df, be_cd = _dot(c_be, m_cf), _dot(c_be, m_df), _dot(c_be, m_cd)
            ab_cf, ab_df, ab_cd = _dot(c_ab, m_cf), _dot(c_ab, m_df), _dot(c_ab, m_cd)
            mGF, mGC, mGD = _mv(Mm, GF), _mv(Mm, GC), _mv(Mm, GD)
            g_ef, g_ec, g_ed = _dot(GE, mGF), _dot(GE, mGC), _dot(GE, mGD)
            g_af, g_ac, g_ad = _dot(GA, mGF), _dot(GA, mGC), _dot(GA, mGD)
            g_bf, g_bc, g_bd = _dot(GB, mGF), _dot(GB, mGC), _dot(GB, mGD)
            vad, vac, vaf = VC(A, D, 0, 0), VC(A, C, 0, 0), VC(A, F, 0, 0)
            vbf, vbc, vbd = VC(B, F, 0, 0), VC(B, C, 0, 0), VC(B, D, 0, 0)
            vde, vef, vce = VC(E, D, 0, 0), VC(E, F, 0, 0), VC(E, C, 0, 0)
            q1 = 2*vad*be_cf + vac*be_df + vaf*be_cd
            l12 = -2*vaf*be_cd - vad*be_cf + vac*be_df
            Dm[ei + 6, ej + 6] = Lac1*Lac2*(4*vbd*ae_cf + 2*vbc*ae_df + 2*vbf*ae_cd + q1
                                            + 2*vde*ab_cf + vce*ab_df + vef*ab_cd)
            Dm[ei + 6, ej + 16] = Lac1*Lab2*(-4*vbf*ae_cd - 2*vbd*ae_cf + 2*vbc*ae_df + l12
                                             - 2*vef*ab_cd - vde*ab_cf + vce*ab_df)
            Dm[ei + 16, ej + 6] = Lab1*Lac2*(-4*vde*ab_cf - 2*vce*ab_df - 2*vef*ab_cd
                                             - 2*vbd*ae_cf - vbc*ae_df - vbf*ae_cd + q1)
            Dm[ei + 16, ej + 16] = Lab1*Lab2*(4*vef*ab_cd + 2*vde*ab_cf - 2*vce*ab_df
                                              + 2*vbf*ae_cd + vbd*ae_cf - vbc*ae_df + l12)
            v_abcd, v_abdf = VC(A, B, C, D), VC(A, B, D, F)
            v_bcde = VC(B, C, D, E) if exact_face_mass else VC(B, C, D, F)
            v_bdef = VC(B, E, D, F)
            v_abcf, v_bcef = VC(A, B, F, C), VC(B, E, F, C)
            v_acde, v_adef = VC(E, A, C, D), VC(E, A, D, F)
            v_acef = VC(A, C, E, F)
            Fm[ei + 6, ej + 6] = Lac1*Lac2*(v_abcd*g_ef - v_abdf*g_ec - v_bcde*g_af + v_bdef*g_ac)
            Fm[ei + 6, ej + 16] = Lac1*Lab2*(v_abdf*g_ec - v_abcf*g_ed - v_bdef*g_ac + v_bcef*g_ad)
            Fm[ei + 16, ej + 6] = Lab1*Lac2*(v_bcde*g_af - v_bdef*g_ac - v_acde*g_bf + v_adef*g_bc)
            Fm[ei + 16, ej + 16] = Lab1*Lab2*(v_bdef*g_ac - v_bcef*g_ad - v_adef*g_bc + v_acef*g_bd)

    Dm *= KA[None, None, :]
    Fm *= KB[None, None, :]
    return Dm, Fm


def local_maps(tets, edges, tris, tet_to_field, nedges):
    """Local edge / face maps of every tet (ASM/curlcurl.py:36-47)."""
    gedge = edges[:, tet_to_field[:6, :]]            # (2,6,T)
    gtri = tris[:, tet_to_field[6:10, :] - nedges]   # (3,4,T)
    return local_mapping(tets, gedge), local_mapping(tets, gtri)


def matrix_builder(nodes, tets, tris, edges, all_edge_lengths, tet_to_field, tet_to_edge,
                   ur, er, chunk: int = 20000, exact_face_mass: bool = False):
    """ASM/curlcurl.py:419-466: COO data for E (curl-curl) and B (mass).

    Entry (i, j) of tet t sits at t*400 + 20*i + j with row tet_to_field[i, t]
    and column tet_to_field[j, t].  ur / er are (3,3,T) complex.
    """
    T = tets.shape[1]
    nedges = edges.shape[1]
    nnz = T * 400
    rows = np.empty(nnz, dtype=np.int64)
    cols = np.empty(nnz, dtype=np.int64)
    dataE = np.empty(nnz, dtype=np.complex128)
    dataB = np.empty(nnz, dtype=np.complex128)
    for s in range(0, T, chunk):
        e = min(T, s + chunk)
        sl = slice(s, e)
        lem, ltm = local_maps(tets[:, sl], edges, tris, tet_to_field[:, sl], nedges)
        el = all_edge_lengths[tet_to_edge[:, sl]]
        verts = nodes[:, tets[:, sl]]  # (3,4,t)
        Dm, Fm = ned2_tet_stiff_mass(verts, el, lem, ltm, matinv(ur[:, :, sl]), er[:, :, sl],
                                     exact_face_mass=exact_face_mass)
        idx = tet_to_field[:, sl]  # (20,t)
        n = e - s
        r = np.broadcast_to(idx.T[:, :, None], (n, 20, 20)).reshape(-1)
        c = np.broadcast_to(idx.T[:, None, :], (n, 20, 20)).reshape(-1)
        rows[s * 400:e * 400] = r
        cols[s * 400:e * 400] = c
        dataE[s * 400:e * 400] = Dm.transpose(2, 0, 1).reshape(-1)
        dataB[s * 400:e * 400] = Fm.transpose(2, 0, 1).reshape(-1)
    return dataE, dataB, rows, cols


def tet_mass_stiffness_matrices(mesh, tet_to_field, er, ur, exact_face_mass: bool = False):
    """ASM/curlcurl.py:145-173: global E, B as scipy CSR (COO->CSR sums
    duplicates, keeps explicit zeros, sorts columns, int32 indices).
    ``exact_face_mass`` as in ``ned2_tet_stiff_mass`` (pins only)."""
    N = 2 * mesh.n_edges + 2 * mesh.n_tris
    dataE, dataB, rows, cols = matrix_builder(mesh.nodes, mesh.tets, mesh.tris, mesh.edges,
                                              mesh.edge_lengths, tet_to_field,
                                              mesh.tet_to_edge, ur, er, exact_face_mass=exact_face_mass)
    E = coo_matrix((dataE, (rows, cols)), shape=(N, N)).tocsr()
    B = coo_matrix((dataB, (rows, cols)), shape=(N, N)).tocsr()
    return E, B


def isotropic(values: np.ndarray) -> np.ndarray:
    """(T,) complex scalars -> (3,3,T) diagonal tensors."""
    T = values.shape[0]
    out = np.zeros((3, 3, T), dtype=np.complex128)
    for k in range(3):
        out[k, k] = values
    return out
