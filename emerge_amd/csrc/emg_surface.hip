// Robin / port surface terms and the PEC-reduced system.
//
// On a boundary triangle the second-order Nedelec tet leaves 8 tangential
// trace functions (vertices 0, 1, 2 ordered by global id; DOF order of
// tri_to_field: edges (0,1), (1,2), (0,2) of the first kind, face of the
// first kind, the same edges of the second kind, face of the second kind).
// With the Whitney form W_pq = lam_p grad lam_q - lam_q grad lam_p they are
//   edge (p, q), 1st kind:  T = -L_pq lam_p W_pq,   2nd kind: T = -L_pq lam_q W_pq,
//   face, 1st kind:  T =  L_02 lam_1 W_02,   2nd kind:  T = -L_01 lam_2 W_01,
// the traces of the volume basis that the reference integrates in
// ned2_tri_stiff / ned2_tri_stiff_force (ASM/robinbc.py:126-427).  Each
// T_i is L_i times two monomials lam_a lam_b grad lam_g, so the 8x8 matrix
// int T_i . T_j dS (without its gamma factor) is a fixed linear form of the
// six products grad lam_a . grad lam_b:
//   S[i][j] = L_i L_j (Area / 360) sum_k TS[i][j][k] d_k,
// with integer TS built at compile time below from the monomial integral
// int lam^k dS = 2 Area prod(k_i!) / (sum k + 2)! (the reference's
// AREA_COEFF, ASM/robinbc.py:36-54).  The forcing of excited ports is
// b_i = -int T_i . U_inc dS on the 6-point Dunavant rule (gaus_quad_tri(4),
// E/mth/optimized.py:31-32, 87-118); the sign is the reference's signA,
// which is -1 on every triangle because its local frame is built from the
// triangle's own edges (ASM/robinbc.py:135-170).  The gradients come from
// the triangle's edge vectors E1 = P1 - P0, E2 = P2 - P0 and their Gram
// matrix (no local frame): grad lam_1 = (g22 E1 - g12 E2) / det,
// grad lam_2 = (g11 E2 - g12 E1) / det, grad lam_0 = -grad lam_1 - grad lam_2,
// det = g11 g22 - g12^2 = (2 Area)^2.
// The per-triangle values are then summed into the reduced CSR pattern /
// reduced DOF vector over host-ordered contribution lists (deterministic
// segmented sums).
#include <algorithm>

#include "emg_common.h"

namespace {

using namespace emg;

namespace trace {

// one term c * lam_a lam_b grad lam_g
struct Term {
    int c, a, b, g;
};
// trace function: length |P_la - P_lb| times two terms
struct Fn {
    int la, lb;
    Term t0, t1;
};

constexpr int kTriEdge[3][2] = {{0, 1}, {1, 2}, {0, 2}};

__host__ __device__ constexpr Fn fn(int i) {
    if (i < 3 || (i >= 4 && i < 7)) {
        const int e = i < 3 ? i : i - 4;
        const int p = kTriEdge[e][0], q = kTriEdge[e][1];
        // -lam_s W_pq = -lam_s lam_p grad lam_q + lam_s lam_q grad lam_p, s = p or q
        const int s = i < 3 ? p : q;
        return Fn{p, q, Term{-1, s, p, q}, Term{+1, s, q, p}};
    }
    if (i == 3) return Fn{0, 2, Term{+1, 1, 0, 2}, Term{-1, 1, 2, 0}};   //  lam_1 W_02
    return Fn{0, 1, Term{-1, 2, 0, 1}, Term{+1, 2, 1, 0}};                // -lam_2 W_01
}

__host__ __device__ constexpr int fact(int k) { return k <= 1 ? 1 : k * fact(k - 1); }

// prod of (multiplicity)! for a degree-4 monomial over the 3 vertices
__host__ __device__ constexpr int mono(int a, int b, int c, int d) {
    int m[3] = {0, 0, 0};
    m[a]++; m[b]++; m[c]++; m[d]++;
    return fact(m[0]) * fact(m[1]) * fact(m[2]);
}

// index of grad lam_a . grad lam_b among (00, 01, 02, 11, 12, 22)
__host__ __device__ constexpr int dot_index(int a, int b) {
    return a > b ? dot_index(b, a) : (a == 0 ? b : (a == 1 ? 2 + b : 5));
}

__host__ __device__ constexpr int pair(const Term& x, const Term& y, int k) {
    return dot_index(x.g, y.g) == k ? x.c * y.c * mono(x.a, x.b, y.a, y.b) : 0;
}

// TS[i][j][k] (numerator over 360)
__host__ __device__ constexpr int ts(int i, int j, int k) {
    const Fn A = fn(i), B = fn(j);
    return pair(A.t0, B.t0, k) + pair(A.t0, B.t1, k) + pair(A.t1, B.t0, k) + pair(A.t1, B.t1, k);
}

struct Table {
    int v[8][8][6];
    __host__ __device__ constexpr Table() : v{} {
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 6; ++k) v[i][j][k] = ts(i, j, k);
    }
};
constexpr Table kTS{};

// the Gram matrix is symmetric term by term
__host__ __device__ constexpr bool symmetric() {
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 6; ++k)
                if (kTS.v[i][j][k] != kTS.v[j][i][k]) return false;
    return true;
}
static_assert(symmetric(), "trace Gram table must be symmetric");
// edge (0,1), first kind: the |grad lam_1|^2 coefficient of
// int |-lam_0^2 grad lam_1 + lam_0 lam_1 grad lam_0|^2 is 4! (lam_0^4)
static_assert(ts(0, 0, dot_index(1, 1)) == mono(0, 0, 0, 0), "edge self term");

}  // namespace trace

// Dunavant order-4 rule as generated by gaus_quad_tri(4): two 3-orbits
// (the forcing and the port overlaps are defined on these points)
__device__ __forceinline__ void dunavant4(int k, double& w, double& l1, double& l2, double& l3) {
    const double W[2] = {0.223381589678011, 0.109951743655322};
    const double A[2] = {0.108103018168070, 0.816847572980459};
    const double B[2] = {0.445948490915965, 0.091576213509771};
    const int o = k / 3, r = k % 3;
    w = W[o];
    // cyclic shifts (a,b,b) -> (b,b,a) -> (b,a,b)
    l1 = r == 0 ? A[o] : B[o];
    l2 = r == 2 ? A[o] : B[o];
    l3 = r == 1 ? A[o] : B[o];
}

__global__ void k_tri_terms(int64_t ntri, const int32_t* __restrict__ tv, const double* __restrict__ xyz,
                            int64_t V, const double2* __restrict__ uinc, double* __restrict__ mat,
                            double2* __restrict__ force) {
    const int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (it >= ntri) return;
    double P[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int v = tv[k * ntri + it];
        P[k][0] = xyz[v];
        P[k][1] = xyz[V + v];
        P[k][2] = xyz[2 * V + v];
    }
    double E1[3], E2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        E1[d] = P[1][d] - P[0][d];
        E2[d] = P[2][d] - P[0][d];
    }
    const double g11 = E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2];
    const double g22 = E2[0] * E2[0] + E2[1] * E2[1] + E2[2] * E2[2];
    const double g12 = E1[0] * E2[0] + E1[1] * E2[1] + E1[2] * E2[2];
    const double det = g11 * g22 - g12 * g12;
    const double inv = 1.0 / det;
    const double area = 0.5 * sqrt(det);
    // grad lam_a . grad lam_b in dot_index order (00, 01, 02, 11, 12, 22)
    const double dd[6] = {(g11 + g22 - 2.0 * g12) * inv, (g12 - g22) * inv, (g12 - g11) * inv,
                          g22 * inv, -g12 * inv, g11 * inv};
    // DOF length factors: |P1-P0|, |P2-P1|, |P2-P0|
    const double e01 = sqrt(g11), e02 = sqrt(g22);
    const double e12 = sqrt((P[2][0] - P[1][0]) * (P[2][0] - P[1][0]) + (P[2][1] - P[1][1]) * (P[2][1] - P[1][1]) +
                            (P[2][2] - P[1][2]) * (P[2][2] - P[1][2]));
    auto len = [&](int a, int b) { return (a + b == 1) ? e01 : ((a + b == 3) ? e12 : e02); };
    double Ls[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const trace::Fn f = trace::fn(i);
        Ls[i] = len(f.la, f.lb);
    }
    const double scale = area / 360.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (trace::kTS.v[i][j][k] != 0) acc += double(trace::kTS.v[i][j][k]) * dd[k];
            mat[it * 64 + i * 8 + j] = Ls[i] * Ls[j] * scale * acc;
        }
    }

    if (uinc != nullptr && force != nullptr) {
        double2 fv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) fv[k] = make_double2(0.0, 0.0);
        for (int q = 0; q < 6; ++q) {
            double w, l[3];
            dunavant4(q, w, l[0], l[1], l[2]);
            const double2 ux = uinc[(it * 6 + q) * 3 + 0], uy = uinc[(it * 6 + q) * 3 + 1],
                          uz = uinc[(it * 6 + q) * 3 + 2];
            // E1 . U and E2 . U, then grad lam_g . U
            const double2 a1 = make_double2(E1[0] * ux.x + E1[1] * uy.x + E1[2] * uz.x,
                                            E1[0] * ux.y + E1[1] * uy.y + E1[2] * uz.y);
            const double2 a2 = make_double2(E2[0] * ux.x + E2[1] * uy.x + E2[2] * uz.x,
                                            E2[0] * ux.y + E2[1] * uy.y + E2[2] * uz.y);
            double2 gu[3];
            gu[1] = make_double2((g22 * a1.x - g12 * a2.x) * inv, (g22 * a1.y - g12 * a2.y) * inv);
            gu[2] = make_double2((g11 * a2.x - g12 * a1.x) * inv, (g11 * a2.y - g12 * a1.y) * inv);
            gu[0] = make_double2(-gu[1].x - gu[2].x, -gu[1].y - gu[2].y);
            const double s = -area * w;   // signA = -1
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const trace::Fn f = trace::fn(i);
                const double c0 = f.t0.c * l[f.t0.a] * l[f.t0.b];
                const double c1 = f.t1.c * l[f.t1.a] * l[f.t1.b];
                const double sl = s * Ls[i];
                fv[i].x += sl * (c0 * gu[f.t0.g].x + c1 * gu[f.t1.g].x);
                fv[i].y += sl * (c0 * gu[f.t0.g].y + c1 * gu[f.t1.g].y);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) force[it * 8 + k] = fv[k];
    }
}

// Second-order ABC correction per boundary triangle, without its
// frequency factor cf = 1j c2 / k0 (ASM/robin_abc_order2.py:188-311,
// ASM/assembler.py:342-348): L_i L_j |A| sum_q W_q (curl_i curl_j - div_i div_j)
// over the module's own 6-point table (8-digit values, :181-184), with the
// tangential surface curl and divergence of the 8 triangle basis functions
// in the triangle frame.  With lambda_i the barycentric coordinates:
//   edge (p, q):  curl = 3 (g_q x g_p) lambda_p and 3 (g_q x g_p) lambda_q,
//                 div  = |g_p|^2 lambda_q - (g_p.g_q) lambda_p and
//                        (g_p.g_q) lambda_q - |g_q|^2 lambda_p;
//   faces: the expressions below (vertices 0, 1, 2 of the sorted triangle).
__global__ void k_tri_abc2(int64_t ntri, const int32_t* __restrict__ tv, const double* __restrict__ xyz, int64_t V,
                           double* __restrict__ mat) {
    const int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (it >= ntri) return;
    double P[3][3];
    for (int k = 0; k < 3; ++k) {
        const int v = tv[k * ntri + it];
        P[k][0] = xyz[v];
        P[k][1] = xyz[V + v];
        P[k][2] = xyz[2 * V + v];
    }
    double e1[3], e2[3], zh[3], xh[3], yh[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = P[1][d] - P[0][d];
        e2[d] = P[2][d] - P[0][d];
    }
    zh[0] = e1[1] * e2[2] - e1[2] * e2[1];
    zh[1] = e1[2] * e2[0] - e1[0] * e2[2];
    zh[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double nz = sqrt(zh[0] * zh[0] + zh[1] * zh[1] + zh[2] * zh[2]);
    const double nx = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int d = 0; d < 3; ++d) {
        zh[d] /= nz;
        xh[d] = e1[d] / nx;
    }
    yh[0] = zh[1] * xh[2] - zh[2] * xh[1];
    yh[1] = zh[2] * xh[0] - zh[0] * xh[2];
    yh[2] = zh[0] * xh[1] - zh[1] * xh[0];
    const double ny = sqrt(yh[0] * yh[0] + yh[1] * yh[1] + yh[2] * yh[2]);
    double X[3], Y[3];
    for (int k = 0; k < 3; ++k) {
        const double dx = P[k][0] - P[0][0], dy = P[k][1] - P[0][1], dz = P[k][2] - P[0][2];
        X[k] = xh[0] * dx + xh[1] * dy + xh[2] * dz;
        Y[k] = (yh[0] * dx + yh[1] * dy + yh[2] * dz) / ny;
    }
    double D[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[i][j] = sqrt((X[i] - X[j]) * (X[i] - X[j]) + (Y[i] - Y[j]) * (Y[i] - Y[j]));
    const double sA = 0.5 * ((X[0] - X[2]) * (Y[1] - Y[0]) - (X[0] - X[1]) * (Y[2] - Y[0]));
    const double sg = sA > 0 ? 1.0 : (sA < 0 ? -1.0 : 0.0);
    const double A = fabs(sA);
    // barycentric lambda_i = a_i + b_i x + c_i y (coefficients over 2A)
    double a[3], b[3], c[3];
    a[0] = (X[1] * Y[2] - Y[1] * X[2]) * sg / (2 * A);
    a[1] = (X[2] * Y[0] - Y[2] * X[0]) * sg / (2 * A);
    a[2] = (X[0] * Y[1] - Y[0] * X[1]) * sg / (2 * A);
    b[0] = (Y[1] - Y[2]) * sg / (2 * A);
    b[1] = (Y[2] - Y[0]) * sg / (2 * A);
    b[2] = (Y[0] - Y[1]) * sg / (2 * A);
    c[0] = (X[2] - X[1]) * sg / (2 * A);
    c[1] = (X[0] - X[2]) * sg / (2 * A);
    c[2] = (X[1] - X[0]) * sg / (2 * A);
    const double W[6] = {0.22338159, 0.22338159, 0.22338159, 0.10995174, 0.10995174, 0.10995174};
    const double Q1[6] = {0.10810302, 0.44594849, 0.44594849, 0.81684757, 0.09157621, 0.09157621};
    const double Q2[6] = {0.44594849, 0.44594849, 0.10810302, 0.09157621, 0.09157621, 0.81684757};
    const double Q3[6] = {0.44594849, 0.10810302, 0.44594849, 0.09157621, 0.81684757, 0.09157621};
    const int ep[3] = {0, 1, 0}, eq[3] = {1, 2, 2};  // local edges (0,1), (1,2), (0,2)
    double L[8];
    for (int e = 0; e < 3; ++e) L[e] = L[e + 4] = D[ep[e]][eq[e]];
    L[3] = D[0][2];
    L[7] = D[0][1];
    double M[8][8];
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) M[i][j] = 0.0;
    for (int q = 0; q < 6; ++q) {
        const double x = X[0] * Q1[q] + X[1] * Q2[q] + X[2] * Q3[q];
        const double y = Y[0] * Q1[q] + Y[1] * Q2[q] + Y[2] * Q3[q];
        double lam[3];
        for (int k = 0; k < 3; ++k) lam[k] = a[k] + b[k] * x + c[k] * y;
        double cu[8], dv[8];
        for (int e = 0; e < 3; ++e) {
            const int p0 = ep[e], p1 = eq[e];
            const double cr = 3.0 * (b[p1] * c[p0] - b[p0] * c[p1]);
            const double gpp = b[p0] * b[p0] + c[p0] * c[p0], gqq = b[p1] * b[p1] + c[p1] * c[p1];
            const double gpq = b[p0] * b[p1] + c[p0] * c[p1];
            cu[e] = cr * lam[p0];
            cu[e + 4] = cr * lam[p1];
            dv[e] = gpp * lam[p1] - gpq * lam[p0];
            dv[e + 4] = gpq * lam[p1] - gqq * lam[p0];
        }
        cu[3] = -b[1] * (c[0] * lam[2] - c[2] * lam[0]) + c[1] * (b[0] * lam[2] - b[2] * lam[0]) +
                2.0 * (b[0] * c[2] - b[2] * c[0]) * lam[1];
        cu[7] = b[2] * (c[0] * lam[1] - c[1] * lam[0]) - c[2] * (b[0] * lam[1] - b[1] * lam[0]) -
                2.0 * (b[0] * c[1] - b[1] * c[0]) * lam[2];
        dv[3] = -b[1] * (b[0] * lam[2] - b[2] * lam[0]) - c[1] * (c[0] * lam[2] - c[2] * lam[0]);
        dv[7] = b[2] * (b[0] * lam[1] - b[1] * lam[0]) + c[2] * (c[0] * lam[1] - c[1] * lam[0]);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) M[i][j] += W[q] * (cu[i] * cu[j] - dv[i] * dv[j]);
    }
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) mat[it * 64 + i * 8 + j] = L[i] * L[j] * M[i][j] * A;
}

// out[k] = sum_{j in [seg[k], seg[k+1])} src[idx[j]]  (fixed order)
__global__ void k_seg_sum_real(int64_t n, const int32_t* __restrict__ seg, const int32_t* __restrict__ idx,
                               const double* __restrict__ src, double* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double s = 0.0;
    for (int j = seg[k]; j < seg[k + 1]; ++j) s += src[idx[j]];
    out[k] = s;
}

__global__ void k_seg_sum_cplx(int64_t n, const int32_t* __restrict__ seg, const int32_t* __restrict__ idx,
                               const int32_t* __restrict__ dst, const double2* __restrict__ src,
                               double2* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double2 s = make_double2(0.0, 0.0);
    for (int j = seg[k]; j < seg[k + 1]; ++j) {
        const double2 v = src[idx[j]];
        s.x += v.x;
        s.y += v.y;
    }
    out[dst[k]] = s;
}

// reduced pattern: per kept row, count kept columns
__global__ void k_red_count(int64_t N, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                            const int32_t* __restrict__ keep, int32_t* __restrict__ cnt) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int32_t rr = keep[r];
    if (rr < 0) return;
    int c = 0;
    for (int j = ptr[r]; j < ptr[r + 1]; ++j) c += keep[idx[j]] >= 0;
    cnt[rr] = c;
}

__global__ void k_red_fill(int64_t N, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                           const int32_t* __restrict__ keep, const int32_t* __restrict__ rptr,
                           int32_t* __restrict__ ridx, int32_t* __restrict__ rmap, int32_t* __restrict__ rdiag) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int32_t rr = keep[r];
    if (rr < 0) return;
    const int o0 = rptr[rr];
    int o = o0;
    for (int j = ptr[r]; j < ptr[r + 1]; ++j) {
        const int32_t cc = keep[idx[j]];
        if (cc < 0) continue;
        ridx[o] = cc;
        rmap[o] = j;
        ++o;
    }
    // the kept ids may come in any order: sort the row by reduced column
    // (insertion sort; rows hold a few dozen entries)
    for (int a = o0 + 1; a < o; ++a) {
        const int32_t cc = ridx[a], jj = rmap[a];
        int z = a - 1;
        while (z >= o0 && ridx[z] > cc) {
            ridx[z + 1] = ridx[z];
            rmap[z + 1] = rmap[z];
            --z;
        }
        ridx[z + 1] = cc;
        rmap[z + 1] = jj;
    }
    rdiag[rr] = -1;
    for (int a = o0; a < o; ++a)
        if (ridx[a] == rr) rdiag[rr] = a;
}

__global__ void k_gather_vals(int64_t n, const int32_t* __restrict__ map, const double2* __restrict__ src,
                              double2* __restrict__ dst64, float2* __restrict__ dst32) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double2 v = src[map[k]];
    dst64[k] = v;
    dst32[k] = make_float2(float(v.x), float(v.y));
}

inline unsigned grid_for(int64_t n, int block) { return unsigned((n + block - 1) / block); }

}  // namespace

extern "C" int emg_system_reduce(emg_ctx* c, int64_t n_keep, const int32_t* keep_ids) {
    EMG_CHECK_CTX(c);
    if (!c->have_values) return fail(c, EMG_ERR_STATE, "emg_system_reduce: assemble first");
    if (!keep_ids || n_keep < 1 || n_keep > c->n_field)
        return fail(c, EMG_ERR_ARG, "emg_system_reduce: bad keep list");
    std::vector<int32_t> keep(c->n_field, -1);
    for (int64_t k = 0; k < n_keep; ++k) {
        const int32_t r = keep_ids[k];
        if (r < 0 || r >= c->n_field || keep[r] >= 0)
            return fail(c, EMG_ERR_ARG, "emg_system_reduce: keep ids must be distinct and in [0, N)");
        keep[r] = int32_t(k);
    }
    // row segments belong to the previous row order
    c->seg_nseg = 0;
    c->seg_n = 0;
    c->seg_row_h.clear();
    hipStream_t st = c->stream;
    const int64_t N = c->n_field;
    EMG_HIP(c, c->keep.alloc(N));
    EMG_HIP(c, hipMemcpyAsync(c->keep.p, keep.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
    DBuf<int32_t> cnt;
    EMG_HIP(c, cnt.alloc(n_keep));
    k_red_count<<<grid_for(N, 256), 256, 0, st>>>(N, c->indptr.p, c->indices.p, c->keep.p, cnt.p);
    EMG_HIP(c, hipGetLastError());
    std::vector<int32_t> h(n_keep);
    EMG_HIP(c, hipMemcpyAsync(h.data(), cnt.p, sizeof(int32_t) * n_keep, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    std::vector<int32_t> rptr(n_keep + 1, 0);
    int64_t tot = 0;
    for (int64_t k = 0; k < n_keep; ++k) {
        rptr[k] = int32_t(tot);
        tot += h[k];
    }
    rptr[n_keep] = int32_t(tot);
    EMG_HIP(c, c->red_ptr.alloc(n_keep + 1));
    EMG_HIP(c, c->red_idx.alloc(tot));
    EMG_HIP(c, c->red_map.alloc(tot));
    EMG_HIP(c, c->red_diag.alloc(n_keep));
    EMG_HIP(c, hipMemcpyAsync(c->red_ptr.p, rptr.data(), sizeof(int32_t) * (n_keep + 1), hipMemcpyHostToDevice, st));
    k_red_fill<<<grid_for(N, 256), 256, 0, st>>>(N, c->indptr.p, c->indices.p, c->keep.p, c->red_ptr.p,
                                                 c->red_idx.p, c->red_map.p, c->red_diag.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, c->E64.alloc(2 * tot));
    EMG_HIP(c, c->B64.alloc(2 * tot));
    EMG_HIP(c, c->E32.alloc(2 * tot));
    EMG_HIP(c, c->B32.alloc(2 * tot));
    k_gather_vals<<<grid_for(tot, 256), 256, 0, st>>>(tot, c->red_map.p, reinterpret_cast<const double2*>(c->Ev.p),
                                                      reinterpret_cast<double2*>(c->E64.p),
                                                      reinterpret_cast<float2*>(c->E32.p));
    k_gather_vals<<<grid_for(tot, 256), 256, 0, st>>>(tot, c->red_map.p, reinterpret_cast<const double2*>(c->Bv.p),
                                                      reinterpret_cast<double2*>(c->B64.p),
                                                      reinterpret_cast<float2*>(c->B32.p));
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(st));
    c->n_red = n_keep;
    c->nnz_red = tot;
    c->have_system = true;
    c->have_bsym = false;
    c->aux_n = 0;  // the auxiliary space refers to the old reduced rows
    c->er64_state = 0;
    c->have_surface = false;
    return EMG_OK;
}

extern "C" int emg_system_refresh(emg_ctx* c) {
    EMG_CHECK_CTX(c);
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_system_refresh: no reduced system (emg_system_reduce)");
    if (!c->have_values) return fail(c, EMG_ERR_STATE, "emg_system_refresh: assemble first");
    hipStream_t st = c->stream;
    const int64_t tot = c->nnz_red;
    if (tot > 0) {
        k_gather_vals<<<grid_for(tot, 256), 256, 0, st>>>(tot, c->red_map.p, reinterpret_cast<const double2*>(c->Ev.p),
                                                          reinterpret_cast<double2*>(c->E64.p),
                                                          reinterpret_cast<float2*>(c->E32.p));
        k_gather_vals<<<grid_for(tot, 256), 256, 0, st>>>(tot, c->red_map.p, reinterpret_cast<const double2*>(c->Bv.p),
                                                          reinterpret_cast<double2*>(c->B64.p),
                                                          reinterpret_cast<float2*>(c->B32.p));
        EMG_HIP(c, hipGetLastError());
    }
    EMG_HIP(c, hipStreamSynchronize(st));
    c->have_bsym = false;  // derived from the old mass values
    c->er64_state = 0;     // E may have turned complex (or real)
    return EMG_OK;
}

extern "C" int emg_system_pattern_get(emg_ctx* c, int32_t* indptr, int32_t* indices, int64_t* nnz) {
    EMG_CHECK_CTX(c);
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_system_pattern_get: no reduced system");
    if (nnz) *nnz = c->nnz_red;
    if (indptr)
        EMG_HIP(c, hipMemcpyAsync(indptr, c->red_ptr.p, sizeof(int32_t) * (c->n_red + 1), hipMemcpyDeviceToHost, c->stream));
    if (indices)
        EMG_HIP(c, hipMemcpyAsync(indices, c->red_idx.p, sizeof(int32_t) * c->nnz_red, hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

namespace {
// append the triangles of one surface term: kind 0 = Robin / port matrix and
// forcing (k_tri_terms), kind 1 = second-order ABC correction (k_tri_abc2)
int add_surface_tris(emg_ctx* c, int kind, int64_t n_tri, const int32_t* tri_vid, const double* uinc,
                     int32_t* bc_id, int64_t* tri_offset) {
    if (c->n_nodes == 0) return fail(c, EMG_ERR_STATE, "emg_bc_add: no mesh");
    if (n_tri < 1 || !tri_vid) return fail(c, EMG_ERR_ARG, "emg_bc_add: empty triangle list");
    if (kind == 1 && uinc) return fail(c, EMG_ERR_ARG, "emg_bc_add_abc2: no forcing for an ABC term");
    for (int64_t k = 0; k < 3 * n_tri; ++k)
        if (tri_vid[k] < 0 || tri_vid[k] >= c->n_nodes) return fail(c, EMG_ERR_ARG, "emg_bc_add: vertex id out of range");
    hipStream_t st = c->stream;
    const int64_t off = c->n_tri_total;
    // grow the concatenated per-triangle buffers (preserving content)
    DBuf<double> mat;
    DBuf<double2> force;
    EMG_HIP(c, mat.alloc(64 * (off + n_tri)));
    EMG_HIP(c, force.alloc(8 * (off + n_tri)));
    EMG_HIP(c, hipMemsetAsync(force.p, 0, sizeof(double2) * 8 * (off + n_tri), st));
    if (off > 0) {
        EMG_HIP(c, hipMemcpyAsync(mat.p, c->tri_mat.p, sizeof(double) * 64 * off, hipMemcpyDeviceToDevice, st));
        EMG_HIP(c, hipMemcpyAsync(force.p, c->tri_force.p, sizeof(double2) * 8 * off, hipMemcpyDeviceToDevice, st));
    }
    DBuf<int32_t> dtv;
    DBuf<double2> du;
    EMG_HIP(c, dtv.alloc(3 * n_tri));
    EMG_HIP(c, hipMemcpyAsync(dtv.p, tri_vid, sizeof(int32_t) * 3 * n_tri, hipMemcpyHostToDevice, st));
    if (uinc) {
        EMG_HIP(c, du.alloc(18 * n_tri));
        EMG_HIP(c, hipMemcpyAsync(du.p, uinc, sizeof(double2) * 18 * n_tri, hipMemcpyHostToDevice, st));
    }
    if (kind == 0)
        k_tri_terms<<<grid_for(n_tri, 64), 64, 0, st>>>(n_tri, dtv.p, c->xyz.p, c->n_nodes, uinc ? du.p : nullptr,
                                                        mat.p + 64 * off, uinc ? force.p + 8 * off : nullptr);
    else
        k_tri_abc2<<<grid_for(n_tri, 64), 64, 0, st>>>(n_tri, dtv.p, c->xyz.p, c->n_nodes, mat.p + 64 * off);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(st));
    std::swap(c->tri_mat.p, mat.p);
    std::swap(c->tri_mat.n, mat.n);
    std::swap(c->tri_force.p, force.p);
    std::swap(c->tri_force.n, force.n);
    c->n_tri_total = off + n_tri;
    if (bc_id) *bc_id = c->n_bc;
    c->n_bc++;
    if (tri_offset) *tri_offset = off;
    c->have_surface = false;
    return EMG_OK;
}
}  // namespace

extern "C" int emg_bc_add(emg_ctx* c, int64_t n_tri, const int32_t* tri_vid, const double* uinc, int32_t* bc_id,
                          int64_t* tri_offset) {
    EMG_CHECK_CTX(c);
    return add_surface_tris(c, 0, n_tri, tri_vid, uinc, bc_id, tri_offset);
}

extern "C" int emg_bc_add_abc2(emg_ctx* c, int64_t n_tri, const int32_t* tri_vid, int32_t* bc_id,
                               int64_t* tri_offset) {
    EMG_CHECK_CTX(c);
    return add_surface_tris(c, 1, n_tri, tri_vid, nullptr, bc_id, tri_offset);
}

extern "C" int emg_tri_terms_get(emg_ctx* c, double* mat, double* force) {
    EMG_CHECK_CTX(c);
    if (c->n_tri_total == 0) return fail(c, EMG_ERR_STATE, "emg_tri_terms_get: no boundary triangles");
    if (mat)
        EMG_HIP(c, hipMemcpyAsync(mat, c->tri_mat.p, sizeof(double) * 64 * c->n_tri_total, hipMemcpyDeviceToHost, c->stream));
    if (force)
        EMG_HIP(c, hipMemcpyAsync(force, c->tri_force.p, sizeof(double2) * 8 * c->n_tri_total, hipMemcpyDeviceToHost,
                                  c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

extern "C" int emg_surface_set(emg_ctx* c, int64_t n_entries, const int32_t* s_ptr, const int32_t* s_col,
                               const int32_t* s_bc, const int32_t* seg, int64_t n_contrib, const int32_t* contrib) {
    EMG_CHECK_CTX(c);
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_surface_set: reduce the system first");
    if (n_entries < 0 || n_contrib < 0) return fail(c, EMG_ERR_ARG, "emg_surface_set: negative sizes");
    hipStream_t st = c->stream;
    for (int64_t k = 0; k < n_contrib; ++k)
        if (contrib[k] < 0 || contrib[k] >= 64 * c->n_tri_total)
            return fail(c, EMG_ERR_ARG, "emg_surface_set: contribution index out of range");
    for (int64_t k = 0; k < n_entries; ++k)
        if (s_col[k] < 0 || s_col[k] >= c->n_red || s_bc[k] < 0 || s_bc[k] >= c->n_bc)
            return fail(c, EMG_ERR_ARG, "emg_surface_set: entry column / bc out of range");
    if (s_ptr[0] != 0 || s_ptr[c->n_red] != n_entries || seg[0] != 0 || seg[n_entries] != n_contrib)
        return fail(c, EMG_ERR_ARG, "emg_surface_set: inconsistent offsets");
    EMG_HIP(c, c->s_ptr.alloc(c->n_red + 1));
    EMG_HIP(c, c->s_col.alloc(std::max<int64_t>(n_entries, 1)));
    EMG_HIP(c, c->s_bc.alloc(std::max<int64_t>(n_entries, 1)));
    EMG_HIP(c, c->s_val64.alloc(std::max<int64_t>(n_entries, 1)));
    EMG_HIP(c, c->s_val32.alloc(std::max<int64_t>(n_entries, 1)));
    EMG_HIP(c, hipMemcpyAsync(c->s_ptr.p, s_ptr, sizeof(int32_t) * (c->n_red + 1), hipMemcpyHostToDevice, st));
    if (n_entries > 0) {
        EMG_HIP(c, hipMemcpyAsync(c->s_col.p, s_col, sizeof(int32_t) * n_entries, hipMemcpyHostToDevice, st));
        EMG_HIP(c, hipMemcpyAsync(c->s_bc.p, s_bc, sizeof(int32_t) * n_entries, hipMemcpyHostToDevice, st));
        DBuf<int32_t> dseg, dcon;
        EMG_HIP(c, dseg.alloc(n_entries + 1));
        EMG_HIP(c, dcon.alloc(std::max<int64_t>(n_contrib, 1)));
        EMG_HIP(c, hipMemcpyAsync(dseg.p, seg, sizeof(int32_t) * (n_entries + 1), hipMemcpyHostToDevice, st));
        if (n_contrib > 0)
            EMG_HIP(c, hipMemcpyAsync(dcon.p, contrib, sizeof(int32_t) * n_contrib, hipMemcpyHostToDevice, st));
        k_seg_sum_real<<<grid_for(n_entries, 256), 256, 0, st>>>(n_entries, dseg.p, dcon.p, c->tri_mat.p,
                                                                 c->s_val64.p);
        EMG_HIP(c, hipGetLastError());
        std::vector<double> h(n_entries);
        EMG_HIP(c, hipMemcpyAsync(h.data(), c->s_val64.p, sizeof(double) * n_entries, hipMemcpyDeviceToHost, st));
        EMG_HIP(c, hipStreamSynchronize(st));
        std::vector<float> hf(n_entries);
        for (int64_t k = 0; k < n_entries; ++k) hf[k] = float(h[k]);
        EMG_HIP(c, hipMemcpyAsync(c->s_val32.p, hf.data(), sizeof(float) * n_entries, hipMemcpyHostToDevice, st));
    }
    EMG_HIP(c, hipStreamSynchronize(st));
    c->n_sentries = n_entries;
    c->have_surface = true;
    return EMG_OK;
}

extern "C" int emg_surface_get(emg_ctx* c, double* s_val) {
    EMG_CHECK_CTX(c);
    if (!c->have_surface) return fail(c, EMG_ERR_STATE, "emg_surface_get: no surface terms");
    if (c->n_sentries > 0 && s_val)
        EMG_HIP(c, hipMemcpyAsync(s_val, c->s_val64.p, sizeof(double) * c->n_sentries, hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

extern "C" int emg_forcing_set(emg_ctx* c, int32_t n_vectors, int64_t n_rows, const int32_t* rows,
                               const int32_t* seg, int64_t n_contrib, const int32_t* contrib) {
    EMG_CHECK_CTX(c);
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_forcing_set: reduce the system first");
    if (n_vectors < 1) return fail(c, EMG_ERR_ARG, "emg_forcing_set: need at least one vector");
    for (int64_t k = 0; k < n_rows; ++k)
        if (rows[k] < 0 || rows[k] >= int64_t(n_vectors) * c->n_red)
            return fail(c, EMG_ERR_ARG, "emg_forcing_set: row out of range");
    for (int64_t k = 0; k < n_contrib; ++k)
        if (contrib[k] < 0 || contrib[k] >= 8 * c->n_tri_total)
            return fail(c, EMG_ERR_ARG, "emg_forcing_set: contribution out of range");
    hipStream_t st = c->stream;
    EMG_HIP(c, c->forcing.alloc(2 * int64_t(n_vectors) * c->n_red));
    EMG_HIP(c, hipMemsetAsync(c->forcing.p, 0, sizeof(double) * 2 * n_vectors * c->n_red, st));
    if (n_rows > 0) {
        DBuf<int32_t> drow, dseg, dcon;
        EMG_HIP(c, drow.alloc(n_rows));
        EMG_HIP(c, dseg.alloc(n_rows + 1));
        EMG_HIP(c, dcon.alloc(std::max<int64_t>(n_contrib, 1)));
        EMG_HIP(c, hipMemcpyAsync(drow.p, rows, sizeof(int32_t) * n_rows, hipMemcpyHostToDevice, st));
        EMG_HIP(c, hipMemcpyAsync(dseg.p, seg, sizeof(int32_t) * (n_rows + 1), hipMemcpyHostToDevice, st));
        if (n_contrib > 0)
            EMG_HIP(c, hipMemcpyAsync(dcon.p, contrib, sizeof(int32_t) * n_contrib, hipMemcpyHostToDevice, st));
        k_seg_sum_cplx<<<grid_for(n_rows, 256), 256, 0, st>>>(n_rows, dseg.p, dcon.p, drow.p, c->tri_force.p,
                                                               reinterpret_cast<double2*>(c->forcing.p));
        EMG_HIP(c, hipGetLastError());
    }
    EMG_HIP(c, hipStreamSynchronize(st));
    c->n_forcing = n_vectors;
    return EMG_OK;
}

extern "C" int emg_forcing_get(emg_ctx* c, double* out) {
    EMG_CHECK_CTX(c);
    if (c->n_forcing == 0) return fail(c, EMG_ERR_STATE, "emg_forcing_get: no forcing vectors");
    EMG_HIP(c, hipMemcpyAsync(out, c->forcing.p, sizeof(double) * 2 * c->n_forcing * c->n_red, hipMemcpyDeviceToHost,
                              c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}
