// Second-order Nedelec tetrahedron in the canonical (global-id sorted) frame.
//
// The reference computes each 20x20 element matrix with local edge/face maps
// built from globally sorted edges and faces (ASM/curlcurl.py:36-47, 180-412,
// E/mth/optimized.py:293-327).  Relabelling the tet's vertices by the rank of
// their global id makes every one of those maps the identity, so the element
// matrices become fixed linear forms of two small Gram vectors:
//
//   M[u][v] = eps_r  /(6V)   * s_u s_v * sum_k TM[u][v][k] g_k / 5040
//   K[u][v] = mu_r^-1/(6V)^3 * s_u s_v * sum_k TK[u][v][k] h_k / 120
//
// with g_k = GL_a.GL_b (a<=b, 10 entries) and h_k = C_E.C_F (E<=F over the 6
// edge cross products C_E = GL_a x GL_b, 21 entries), GL = (b,c,d) the
// unnormalised barycentric gradients (6V grad lambda) of
// tet_coefficients_bcd (ASM/curlcurl.py:99-138), and s_u the DOF length
// factor (edge length; |ac| for the first and |ab| for the second face DOF of
// a face abc, ASM/curlcurl.py:218-245, 292-293, 336-337, 367-368).
//
// TM / TK are small integers derived here at compile time from the basis
// N_u = s_u lam_p (lam_q grad lam_r - lam_r grad lam_q) of
// E/elements/ned2_interp.py:243-296 and the monomial integral
// int lam^k dV = 6V prod(k_i!)/(sum k + 3)! (ASM/curlcurl.py:64-92).
// One deliberate deviation from the exact integral reproduces the reference:
// its face-face mass entries use the volume coefficient [B,C,D,F] where the
// integral has [B,C,D,E] (ASM/curlcurl.py:361, used at :404 and :406).
//
// DOF order inside a 20x20 block (the reference's): 0-5 edge (1st kind),
// 6-9 face (1st), 10-15 edge (2nd), 16-19 face (2nd).
#pragma once

#include <utility>

namespace emg {
namespace ned2 {

// canonical edges (a<b) and faces (a<b<c) over vertex ranks 0..3
constexpr int kEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
constexpr int kFace[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
// tet-local vertex pairs / triples in the reference's order (E/mesh3d.py:427, :431)
constexpr int kLocalEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
constexpr int kLocalFaceMissing[4] = {3, 1, 2, 0};  // vertex not on local face k

__host__ __device__ constexpr int edge_index(int a, int b) {
    // (0,1)=0 (0,2)=1 (0,3)=2 (1,2)=3 (1,3)=4 (2,3)=5
    return a > b ? edge_index(b, a) : (a == 0 ? b - 1 : (a == 1 ? b + 1 : 5));
}
__host__ __device__ constexpr int gram_index(int a, int b) {
    return a > b ? gram_index(b, a) : a * 4 - a * (a - 1) / 2 + (b - a);
}
__host__ __device__ constexpr int cross_gram_index(int e, int f) {
    return e > f ? cross_gram_index(f, e) : e * 6 - e * (e - 1) / 2 + (f - e);
}

struct Dof {
    int p, q, r;   // N = lam_p (lam_q grad lam_r - lam_r grad lam_q)
    int la, lb;    // length factor |x_la - x_lb|
};

__host__ __device__ constexpr Dof dof(int u) {
    if (u < 6) return Dof{kEdge[u][0], kEdge[u][1], kEdge[u][0], kEdge[u][0], kEdge[u][1]};
    if (u < 10) {
        const int* f = kFace[u - 6];
        return Dof{f[1], f[0], f[2], f[0], f[2]};
    }
    if (u < 16) {
        const int* e = kEdge[u - 10];
        return Dof{e[1], e[1], e[0], e[0], e[1]};
    }
    const int* f = kFace[u - 16];
    return Dof{f[2], f[1], f[0], f[0], f[1]};
}

__host__ __device__ constexpr int fact(int k) { return k <= 1 ? 1 : k * fact(k - 1); }

// prod over ranks of (multiplicity)! for a monomial of 4 barycentric letters
__host__ __device__ constexpr int mono4(int a, int b, int c, int d) {
    int m[4] = {0, 0, 0, 0};
    m[a]++; m[b]++; m[c]++; m[d]++;
    return fact(m[0]) * fact(m[1]) * fact(m[2]) * fact(m[3]);
}
__host__ __device__ constexpr int mono2(int a, int b) { return a == b ? 2 : 1; }

// mass coefficient TM[u][v][k] (integer numerator over 5040)
__host__ __device__ constexpr int tm(int u, int v, int k) {
    const Dof A = dof(u), B = dof(v);
    int acc = 0;
    if (gram_index(A.r, B.r) == k) acc += mono4(A.p, A.q, B.p, B.q);
    if (gram_index(A.r, B.q) == k) acc -= mono4(A.p, A.q, B.p, B.r);
    if (gram_index(A.q, B.r) == k) acc -= mono4(A.p, A.r, B.p, B.q);
    if (gram_index(A.q, B.q) == k) acc += mono4(A.p, A.r, B.p, B.r);
    // reference quirk (ASM/curlcurl.py:361): face-face mass rows 6..9 and
    // 16..19 against face columns 6..9 take [B,C,D,F] for [B,C,D,E]
    if (v >= 6 && v < 10 && ((u >= 6 && u < 10) || u >= 16)) {
        const int fi = (u < 10) ? u - 6 : u - 16;
        const int fj = v - 6;
        const int a = kFace[fi][0], b = kFace[fi][1], e = kFace[fi][2];
        const int c = kFace[fj][0], d = kFace[fj][1], f = kFace[fj][2];
        const int dv = mono4(b, c, d, f) - mono4(b, c, d, e);
        if (gram_index(a, f) == k) acc += (u < 10) ? -dv : dv;
    }
    return acc;
}

// curl N_u = s_u [2 lam_p C(q,r) + lam_q C(p,r) - lam_r C(p,q)],
// C(x,y) = grad lam_x x grad lam_y = sign * C_E
struct CurlTerm {
    int coef, lam, e;  // coef 0 means absent
};
__host__ __device__ constexpr CurlTerm curl_term(const Dof& D, int m) {
    int c = m == 0 ? 2 : (m == 1 ? 1 : -1);
    int lam = m == 0 ? D.p : (m == 1 ? D.q : D.r);
    int x = m == 0 ? D.q : D.p;
    int y = m == 0 ? D.r : (m == 1 ? D.r : D.q);
    if (x == y) return CurlTerm{0, 0, 0};
    if (x > y) c = -c;
    return CurlTerm{c, lam, edge_index(x, y)};
}

// curl-curl coefficient TK[u][v][k] (integer numerator over 120)
__host__ __device__ constexpr int tk(int u, int v, int k) {
    const Dof A = dof(u), B = dof(v);
    int acc = 0;
    for (int m = 0; m < 3; ++m) {
        const CurlTerm a = curl_term(A, m);
        if (a.coef == 0) continue;
        for (int n = 0; n < 3; ++n) {
            const CurlTerm b = curl_term(B, n);
            if (b.coef == 0) continue;
            if (cross_gram_index(a.e, b.e) == k) acc += a.coef * b.coef * mono2(a.lam, b.lam);
        }
    }
    return acc;
}

// Ordered tables for general (not necessarily symmetric) material tensors:
// the mass term of (row letter x, column letter y) multiplies
// grad lam_x . (eps grad lam_y), index 4x + y; the curl-curl term of (row
// cross e, column cross f) multiplies C_e . (mu^-1 C_f), index 6e + f
// (ASM/curlcurl.py:213-407: dot(G_row, Mm G_col), dot(C_row, Ms C_col)).
// The reference fills the face-row x edge-column blocks by transposing the
// edge-row x face-column ones (ASM/curlcurl.py:321-328), so there entry
// (u, v) is entry (v, u) with the products taken in the (v, u) order.
__host__ __device__ constexpr bool is_face(int u) { return (u >= 6 && u < 10) || u >= 16; }
__host__ __device__ constexpr int tmo_direct(int u, int v, int k) {
    const Dof A = dof(u), B = dof(v);
    int acc = 0;
    if (4 * A.r + B.r == k) acc += mono4(A.p, A.q, B.p, B.q);
    if (4 * A.r + B.q == k) acc -= mono4(A.p, A.q, B.p, B.r);
    if (4 * A.q + B.r == k) acc -= mono4(A.p, A.r, B.p, B.q);
    if (4 * A.q + B.q == k) acc += mono4(A.p, A.r, B.p, B.r);
    if (v >= 6 && v < 10 && ((u >= 6 && u < 10) || u >= 16)) {
        const int fi = (u < 10) ? u - 6 : u - 16;
        const int fj = v - 6;
        const int a = kFace[fi][0], b = kFace[fi][1], e = kFace[fi][2];
        const int c = kFace[fj][0], d = kFace[fj][1], f = kFace[fj][2];
        const int dv = mono4(b, c, d, f) - mono4(b, c, d, e);
        if (4 * a + f == k) acc += (u < 10) ? -dv : dv;
    }
    return acc;
}
__host__ __device__ constexpr int tmo(int u, int v, int k) {
    return (is_face(u) && !is_face(v)) ? tmo_direct(v, u, k) : tmo_direct(u, v, k);
}
__host__ __device__ constexpr int tko_direct(int u, int v, int k) {
    const Dof A = dof(u), B = dof(v);
    int acc = 0;
    for (int m = 0; m < 3; ++m) {
        const CurlTerm a = curl_term(A, m);
        if (a.coef == 0) continue;
        for (int n = 0; n < 3; ++n) {
            const CurlTerm b = curl_term(B, n);
            if (b.coef == 0) continue;
            if (6 * a.e + b.e == k) acc += a.coef * b.coef * mono2(a.lam, b.lam);
        }
    }
    return acc;
}
__host__ __device__ constexpr int tko(int u, int v, int k) {
    return (is_face(u) && !is_face(v)) ? tko_direct(v, u, k) : tko_direct(u, v, k);
}

template <int C>
__device__ __forceinline__ double madd(double acc, double x) {
    if constexpr (C == 0) {
        return acc;
    } else if constexpr (C == 1) {
        return acc + x;
    } else if constexpr (C == -1) {
        return acc - x;
    } else {
        return __builtin_fma(static_cast<double>(C), x, acc);
    }
}

template <int U, int V, int... K>
__device__ __forceinline__ double contract_k(const double* h, std::integer_sequence<int, K...>) {
    double acc = 0.0;
    ((acc = madd<tk(U, V, K)>(acc, h[K])), ...);
    return acc;
}
template <int U, int V, int... K>
__device__ __forceinline__ double contract_m(const double* g, std::integer_sequence<int, K...>) {
    double acc = 0.0;
    ((acc = madd<tm(U, V, K)>(acc, g[K])), ...);
    return acc;
}

// Length-factor index of DOF u: s_u = d[len_index(u)] with d the 6 edge
// lengths of the canonical tet.
__host__ __device__ constexpr int len_index(int u) { return edge_index(dof(u).la, dof(u).lb); }

// Row U of the unscaled element matrices: kr[v] = s_v * (TK h)[U][v] and
// mr[v] = s_v * (TM g)[U][v] (computed in separate passes to bound the
// number of live registers); s_v = d[len_index(v)].
template <int U, int... V>
__device__ __forceinline__ void stiff_row_impl(const double* h, const double* d, double* kr, int stride,
                                               std::integer_sequence<int, V...>) {
    ((kr[V * stride] = d[len_index(V)] * contract_k<U, V>(h, std::make_integer_sequence<int, 21>{})), ...);
}
template <int U, int... V>
__device__ __forceinline__ void mass_row_impl(const double* g, const double* d, double* mr, int stride,
                                              std::integer_sequence<int, V...>) {
    ((mr[V * stride] = d[len_index(V)] * contract_m<U, V>(g, std::make_integer_sequence<int, 10>{})), ...);
}

template <int U, int V, int... K>
__device__ __forceinline__ double contract_ko(const double* h, std::integer_sequence<int, K...>) {
    double acc = 0.0;
    ((acc = madd<tko(U, V, K)>(acc, h[K])), ...);
    return acc;
}
template <int U, int V, int... K>
__device__ __forceinline__ double contract_mo(const double* g, std::integer_sequence<int, K...>) {
    double acc = 0.0;
    ((acc = madd<tmo(U, V, K)>(acc, g[K])), ...);
    return acc;
}
template <int U, int... V>
__device__ __forceinline__ void stiff_row_o_impl(const double* h, const double* d, double* kr, int stride,
                                                 std::integer_sequence<int, V...>) {
    ((kr[V * stride] = d[len_index(V)] * contract_ko<U, V>(h, std::make_integer_sequence<int, 36>{})), ...);
}
template <int U, int... V>
__device__ __forceinline__ void mass_row_o_impl(const double* g, const double* d, double* mr, int stride,
                                                std::integer_sequence<int, V...>) {
    ((mr[V * stride] = d[len_index(V)] * contract_mo<U, V>(g, std::make_integer_sequence<int, 16>{})), ...);
}

// Dispatch on a wave-uniform canonical row index.
#define EMG_ROW_SWITCH(CALL)                                                            \
    switch (u) {                                                                        \
        case 0: CALL(0); break;   case 1: CALL(1); break;   case 2: CALL(2); break;     \
        case 3: CALL(3); break;   case 4: CALL(4); break;   case 5: CALL(5); break;     \
        case 6: CALL(6); break;   case 7: CALL(7); break;   case 8: CALL(8); break;     \
        case 9: CALL(9); break;   case 10: CALL(10); break; case 11: CALL(11); break;   \
        case 12: CALL(12); break; case 13: CALL(13); break; case 14: CALL(14); break;   \
        case 15: CALL(15); break; case 16: CALL(16); break; case 17: CALL(17); break;   \
        case 18: CALL(18); break; case 19: CALL(19); break; default: break;             \
    }

// out[v * stride] receives column v of row u.
__device__ __forceinline__ void stiff_row(int u, const double* h, const double* d, double* kr, int stride) {
#define EMG_K_CALL(n) stiff_row_impl<n>(h, d, kr, stride, std::make_integer_sequence<int, 20>{})
    EMG_ROW_SWITCH(EMG_K_CALL)
#undef EMG_K_CALL
}
__device__ __forceinline__ void mass_row(int u, const double* g, const double* d, double* mr, int stride) {
#define EMG_M_CALL(n) mass_row_impl<n>(g, d, mr, stride, std::make_integer_sequence<int, 20>{})
    EMG_ROW_SWITCH(EMG_M_CALL)
#undef EMG_M_CALL
}

__device__ __forceinline__ void stiff_row_o(int u, const double* h, const double* d, double* kr, int stride) {
#define EMG_K_CALL(n) stiff_row_o_impl<n>(h, d, kr, stride, std::make_integer_sequence<int, 20>{})
    EMG_ROW_SWITCH(EMG_K_CALL)
#undef EMG_K_CALL
}
__device__ __forceinline__ void mass_row_o(int u, const double* g, const double* d, double* mr, int stride) {
#define EMG_M_CALL(n) mass_row_o_impl<n>(g, d, mr, stride, std::make_integer_sequence<int, 20>{})
    EMG_ROW_SWITCH(EMG_M_CALL)
#undef EMG_M_CALL
}

// Per-tet canonical geometry: Gram vectors and DOF lengths.
struct TetGeom {
    double g[10];
    double h[21];
    double d[6];   // edge lengths; DOF u scales with d[len_index(u)]
    double inv6v;
};

// rank of each of 4 distinct global ids
__device__ __forceinline__ void ranks4(const int* w, int* rk) {
#pragma unroll
    for (int k = 0; k < 4; ++k) rk[k] = (w[0] < w[k]) + (w[1] < w[k]) + (w[2] < w[k]) + (w[3] < w[k]);
}

// x / 6 correctly rounded, as IEEE division gives it, in three operations
// instead of a full division sequence.  q = RN(x r) with r = RN(1/6) is
// within one unit in the last place of x/6; e = x - 6q is exact (one FMA);
// q + e r then differs from x/6 by at most 2^-54 of a unit in the last place
// before its single rounding.  x/6 = (x/2)/3 never lies closer than 1/6 of
// a unit to a rounding midpoint (a midpoint times 3 is an odd multiple of
// 3 * 2^-55 in the binade's units, which no double x/2 equals), so the final
// rounding returns RN(x/6) for every x whose quotient is a normal number.
__device__ __forceinline__ double div6(double x) {
    constexpr double r = 0x1.5555555555555p-3;  // RN(1/6)
    const double q = x * r;
    const double e = __builtin_fma(-q, 6.0, x);
    return __builtin_fma(e, r, q);
}

// Canonical gradient vectors gl (cofactors, 6V grad lam) and their pairwise
// crosses c, for material tensors (the Gram products then carry the tensor).
struct TetVecs {
    double gl[4][3];
    double c[6][3];
    double d[6];
    double inv6v;
};

// XL[k] = coordinates of tet-local vertex k (the mesh's own vertex order),
// rk[k] = rank of its global id.  Cofactors and volume use the reference's
// polynomials in local order (ASM/curlcurl.py:111-136) so that they round
// exactly as the reference does, then move to the canonical frame.
__device__ __forceinline__ void tet_vectors(const double (&XL)[4][3], const int (&rk)[4], TetVecs& T) {
    // The 24-term volume and the cofactors cancel heavily for small tets far
    // from the origin; evaluate them with plain IEEE multiply/add (no FMA
    // contraction) so they round exactly like the reference.
#pragma clang fp contract(off)
    const double x1 = XL[0][0], x2 = XL[1][0], x3 = XL[2][0], x4 = XL[3][0];
    const double y1 = XL[0][1], y2 = XL[1][1], y3 = XL[2][1], y4 = XL[3][1];
    const double z1 = XL[0][2], z2 = XL[1][2], z3 = XL[2][2], z4 = XL[3][2];
    double gll[4][3];
    gll[0][0] = -y2*z3 + y2*z4 + y3*z2 - y3*z4 - y4*z2 + y4*z3;
    gll[1][0] = y1*z3 - y1*z4 - y3*z1 + y3*z4 + y4*z1 - y4*z3;
    gll[2][0] = -y1*z2 + y1*z4 + y2*z1 - y2*z4 - y4*z1 + y4*z2;
    gll[3][0] = y1*z2 - y1*z3 - y2*z1 + y2*z3 + y3*z1 - y3*z2;
    gll[0][1] = x2*z3 - x2*z4 - x3*z2 + x3*z4 + x4*z2 - x4*z3;
    gll[1][1] = -x1*z3 + x1*z4 + x3*z1 - x3*z4 - x4*z1 + x4*z3;
    gll[2][1] = x1*z2 - x1*z4 - x2*z1 + x2*z4 + x4*z1 - x4*z2;
    gll[3][1] = -x1*z2 + x1*z3 + x2*z1 - x2*z3 - x3*z1 + x3*z2;
    gll[0][2] = -x2*y3 + x2*y4 + x3*y2 - x3*y4 - x4*y2 + x4*y3;
    gll[1][2] = x1*y3 - x1*y4 - x3*y1 + x3*y4 + x4*y1 - x4*y3;
    gll[2][2] = -x1*y2 + x1*y4 + x2*y1 - x2*y4 - x4*y1 + x4*y2;
    gll[3][2] = x1*y2 - x1*y3 - x2*y1 + x2*y3 + x3*y1 - x3*y2;
    // the reference's 24 terms (x*y*z)/6 in its order; div6 rounds each
    // quotient exactly as the division does
    const double vol = fabs(-div6(x1*y2*z3) + div6(x1*y2*z4) + div6(x1*y3*z2) - div6(x1*y3*z4)
                            - div6(x1*y4*z2) + div6(x1*y4*z3) + div6(x2*y1*z3) - div6(x2*y1*z4)
                            - div6(x2*y3*z1) + div6(x2*y3*z4) + div6(x2*y4*z1) - div6(x2*y4*z3)
                            - div6(x3*y1*z2) + div6(x3*y1*z4) + div6(x3*y2*z1) - div6(x3*y2*z4)
                            - div6(x3*y4*z1) + div6(x3*y4*z2) + div6(x4*y1*z2) - div6(x4*y1*z3)
                            - div6(x4*y2*z1) + div6(x4*y2*z3) + div6(x4*y3*z1) - div6(x4*y3*z2));
    T.inv6v = 1.0 / (6.0 * vol);
    // canonical frame: gl[c] / X[c] belong to the local vertex of rank c
    double gl[4][3], X[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double gv = gll[0][d], xv = XL[0][d];
#pragma unroll
            for (int k = 1; k < 4; ++k) {
                gv = rk[k] == c ? gll[k][d] : gv;
                xv = rk[k] == c ? XL[k][d] : xv;
            }
            gl[c][d] = gv;
            X[c][d] = xv;
        }
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int a = kEdge[e][0], b = kEdge[e][1];
        // (no rounding order to follow here: fused products)
        T.c[e][0] = __builtin_fma(gl[a][1], gl[b][2], -(gl[a][2] * gl[b][1]));
        T.c[e][1] = __builtin_fma(gl[a][2], gl[b][0], -(gl[a][0] * gl[b][2]));
        T.c[e][2] = __builtin_fma(gl[a][0], gl[b][1], -(gl[a][1] * gl[b][0]));
        const double dx = X[a][0] - X[b][0], dy = X[a][1] - X[b][1], dz = X[a][2] - X[b][2];
        T.d[e] = sqrt(__builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz)));
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) T.gl[k][dd] = gl[k][dd];
}

__device__ __forceinline__ void tet_geometry(const double (&XL)[4][3], const int (&rk)[4], TetGeom& G) {
#pragma clang fp contract(off)
    TetVecs T;
    tet_vectors(XL, rk, T);
    G.inv6v = T.inv6v;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a; b < 4; ++b)
            G.g[gram_index(a, b)] =
                __builtin_fma(T.gl[a][0], T.gl[b][0], __builtin_fma(T.gl[a][1], T.gl[b][1], T.gl[a][2] * T.gl[b][2]));
#pragma unroll
    for (int e = 0; e < 6; ++e)
#pragma unroll
        for (int f = e; f < 6; ++f)
            G.h[cross_gram_index(e, f)] =
                __builtin_fma(T.c[e][0], T.c[f][0], __builtin_fma(T.c[e][1], T.c[f][1], T.c[e][2] * T.c[f][2]));
#pragma unroll
    for (int e = 0; e < 6; ++e) G.d[e] = T.d[e];
}

}  // namespace ned2
}  // namespace emg
