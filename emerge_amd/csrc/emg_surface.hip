// Robin / port surface terms and the PEC-reduced system.
//
// Per boundary triangle the kernel evaluates the 8x8 matrix of
// ned2_tri_stiff / ned2_tri_stiff_force without its gamma factor
// (ASM/robinbc.py:314-427, 126-292) and, for excited ports, the forcing
// vector b_i = int N_i^t . U_inc dS with the 6-point Dunavant rule
// (gaus_quad_tri(4), E/mth/optimized.py:31-32, 87-118).  The per-triangle
// values are then summed into the reduced CSR pattern / reduced DOF vector
// over host-ordered contribution lists (deterministic segmented sums).
#include <algorithm>

#include "emg_common.h"

namespace {

using namespace emg;

// area coefficient 2 prod(k!)/(sum k + 2)! for 4 letters over 3 triangle
// vertices (ASM/robinbc.py:36-45); letters 1..3, 0 = none (never used here)
__device__ __forceinline__ double area_coeff4(int a, int b, int c, int d) {
    int m[4] = {0, 0, 0, 0};
    m[a]++; m[b]++; m[c]++; m[d]++;
    const int f[5] = {1, 1, 2, 6, 24};
    const double num = double(f[m[1]] * f[m[2]] * f[m[3]]);
    return 2.0 * num / 720.0;  // (4 + 2)! = 720
}

// Dunavant order-4 rule as generated by gaus_quad_tri(4): two 3-orbits
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
    double G[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int v = tv[k * ntri + it];
        G[k][0] = xyz[v];
        G[k][1] = xyz[V + v];
        G[k][2] = xyz[2 * V + v];
    }
    // local frame (ASM/robinbc.py:135-148)
    double e1[3], e2[3], zh[3], xh[3], yh[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = G[1][d] - G[0][d];
        e2[d] = G[2][d] - G[0][d];
    }
    zh[0] = e1[1] * e2[2] - e1[2] * e2[1];
    zh[1] = e1[2] * e2[0] - e1[0] * e2[2];
    zh[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double nz = sqrt(zh[0] * zh[0] + zh[1] * zh[1] + zh[2] * zh[2]);
    double nx = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int d = 0; d < 3; ++d) {
        zh[d] /= nz;
        xh[d] = e1[d] / nx;
    }
    yh[0] = zh[1] * xh[2] - zh[2] * xh[1];
    yh[1] = zh[2] * xh[0] - zh[0] * xh[2];
    yh[2] = zh[0] * xh[1] - zh[1] * xh[0];
    double ny = sqrt(yh[0] * yh[0] + yh[1] * yh[1] + yh[2] * yh[2]);
    for (int d = 0; d < 3; ++d) yh[d] /= ny;
    double xs[3], ys[3];
    for (int k = 0; k < 3; ++k) {
        const double dx = G[k][0] - G[0][0], dy = G[k][1] - G[0][1], dz = G[k][2] - G[0][2];
        xs[k] = xh[0] * dx + xh[1] * dy + xh[2] * dz;
        ys[k] = yh[0] * dx + yh[1] * dy + yh[2] * dz;
    }
    const double x1 = xs[0], x2 = xs[1], x3 = xs[2], y1 = ys[0], y2 = ys[1], y3 = ys[2];
    const double bb[3] = {y2 - y3, y3 - y1, y1 - y2};
    const double cc[3] = {x3 - x2, x1 - x3, x2 - x1};
    const double aa[3] = {x2 * y3 - y2 * x3, x3 * y1 - y3 * x1, x1 * y2 - y1 * x2};
    double Ds[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ds[i][j] = sqrt((xs[i] - xs[j]) * (xs[i] - xs[j]) + (ys[i] - ys[j]) * (ys[i] - ys[j]));
    const double sgn = (x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1);
    const double area = 0.5 * fabs(sgn);
    auto dot2 = [&](int p, int q) { return bb[p] * bb[q] + cc[p] * cc[q]; };
    auto AC = [&](int a, int b, int c, int d) { return area_coeff4(a, b, c, d) * area; };
    const int lem[2][3] = {{0, 1, 0}, {1, 2, 2}};
    double M[8][8];
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) M[i][j] = 0.0;
    const int tA = 1, tB = 2, tC = 3;
    const double Lt1 = Ds[2][0], Lt2 = Ds[1][0];
    for (int ei = 0; ei < 3; ++ei) {
        const int ei1 = lem[0][ei], ei2 = lem[1][ei];
        const double Li = Ds[ei1][ei2];
        const int A = ei1 + 1, B = ei2 + 1;
        for (int ej = 0; ej < 3; ++ej) {
            const int ej1 = lem[0][ej], ej2 = lem[1][ej];
            const double Lj = Ds[ej1][ej2];
            const int C = ej1 + 1, D = ej2 + 1;
            const double DAC = dot2(ei1, ej1), DAD = dot2(ei1, ej2), DBC = dot2(ei2, ej1), DBD = dot2(ei2, ej2);
            const double LL = Li * Lj;
            M[ei][ej] += LL * (AC(A, B, C, D) * DAC - AC(A, B, C, C) * DAD - AC(A, A, C, D) * DBC + AC(A, A, C, C) * DBD);
            M[ei][ej + 4] += LL * (AC(A, B, D, D) * DAC - AC(A, B, C, D) * DAD - AC(A, A, D, D) * DBC + AC(A, A, C, D) * DBD);
            M[ei + 4][ej] += LL * (AC(B, B, C, D) * DAC - AC(B, B, C, C) * DAD - AC(A, B, C, D) * DBC + AC(A, B, C, C) * DBD);
            M[ei + 4][ej + 4] += LL * (AC(B, B, D, D) * DAC - AC(B, B, C, D) * DAD - AC(A, B, D, D) * DBC + AC(A, B, C, D) * DBD);
        }
        const double FA = dot2(ei1, 2), FB = dot2(ei1, 0), FC = dot2(ei2, 2);
        const double FD = dot2(ei2, 0), FE = dot2(ei1, 1), FF = dot2(ei2, 1);
        M[ei][3] += Li * Lt1 * (AC(A, B, tA, tB) * FA - AC(A, B, tB, tC) * FB - AC(A, A, tA, tB) * FC + AC(A, A, tB, tC) * FD);
        M[ei][7] += Li * Lt2 * (AC(A, B, tB, tC) * FB - AC(A, B, tC, tA) * FE - AC(A, A, tB, tC) * FD + AC(A, A, tC, tA) * FF);
        M[3][ei] += Lt1 * Li * (AC(tA, tB, A, B) * FA - AC(tA, tB, A, A) * FC - AC(tB, tC, A, B) * FB + AC(tB, tC, A, A) * FD);
        M[7][ei] += Lt2 * Li * (AC(tB, tC, A, B) * FB - AC(tB, tC, A, A) * FD - AC(tC, tA, A, B) * FE + AC(tC, tA, A, A) * FF);
        M[ei + 4][3] += Li * Lt1 * (AC(B, B, tA, tB) * FA - AC(B, B, tB, tC) * FB - AC(A, B, tA, tB) * FC + AC(A, B, tB, tC) * FD);
        M[ei + 4][7] += Li * Lt2 * (AC(B, B, tB, tC) * FB - AC(B, B, tC, tA) * FE - AC(A, B, tB, tC) * FD + AC(A, B, tC, tA) * FF);
        M[3][ei + 4] += Lt1 * Li * (AC(tA, tB, B, B) * FA - AC(tA, tB, A, B) * FC - AC(tB, tC, B, B) * FB + AC(tB, tC, A, B) * FD);
        M[7][ei + 4] += Lt2 * Li * (AC(tB, tC, B, B) * FB - AC(tB, tC, A, B) * FD - AC(tC, tA, B, B) * FE + AC(tC, tA, A, B) * FF);
    }
    const double H1 = dot2(0, 2), H2 = dot2(0, 0), H3 = dot2(0, 1);
    M[3][3] += Lt1 * Lt1 * (AC(tA, tB, tA, tB) * dot2(2, 2) - AC(tA, tB, tB, tC) * H1 - AC(tB, tC, tA, tB) * H1 + AC(tB, tC, tB, tC) * H2);
    M[3][7] += Lt1 * Lt2 * (AC(tA, tB, tB, tC) * H1 - AC(tA, tB, tC, tA) * dot2(1, 2) - AC(tB, tC, tB, tC) * H2 + AC(tB, tC, tC, tA) * H3);
    M[7][3] += Lt2 * Lt1 * (AC(tB, tC, tA, tB) * H1 - AC(tB, tC, tB, tC) * H2 - AC(tC, tA, tA, tB) * dot2(1, 2) + AC(tC, tA, tB, tC) * H3);
    M[7][7] += Lt2 * Lt2 * (AC(tB, tC, tB, tC) * H2 - AC(tB, tC, tC, tA) * H3 - AC(tC, tA, tB, tC) * H3 + AC(tC, tA, tC, tA) * dot2(1, 1));
    const double coef = 1.0 / ((2.0 * area) * (2.0 * area));
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) mat[it * 64 + i * 8 + j] = M[i][j] * coef;

    if (uinc != nullptr && force != nullptr) {
        double2 fv[8];
        for (int k = 0; k < 8; ++k) fv[k] = make_double2(0.0, 0.0);
        const double signA = sgn > 0 ? -1.0 : (sgn < 0 ? 1.0 : 0.0);
        for (int q = 0; q < 6; ++q) {
            double w, l1, l2, l3;
            dunavant4(q, w, l1, l2, l3);
            const double x = x1 * l1 + x2 * l2 + x3 * l3;
            const double y = y1 * l1 + y2 * l2 + y3 * l3;
            // U_inc in the triangle frame: (xhat . U, yhat . U)
            const double2 ux = uinc[(it * 6 + q) * 3 + 0], uy = uinc[(it * 6 + q) * 3 + 1],
                          uz = uinc[(it * 6 + q) * 3 + 2];
            const double2 Ux = make_double2(xh[0] * ux.x + xh[1] * uy.x + xh[2] * uz.x,
                                            xh[0] * ux.y + xh[1] * uy.y + xh[2] * uz.y);
            const double2 Uy = make_double2(yh[0] * ux.x + yh[1] * uy.x + yh[2] * uz.x,
                                            yh[0] * ux.y + yh[1] * uy.y + yh[2] * uz.y);
            for (int ei = 0; ei < 3; ++ei) {
                const int ei1 = lem[0][ei], ei2 = lem[1][ei];
                const double Li = Ds[ei1][ei2];
                const double Q = aa[ei2] + bb[ei2] * x + cc[ei2] * y;
                const double Z = aa[ei1] + bb[ei1] * x + cc[ei1] * y;
                const double A4 = 4 * area * area;
                const double Q2 = Q / A4, Z2 = Z / A4, Ar2 = 1 / (2 * area);
                const double e1x = (bb[ei1] * Q2 - bb[ei2] * Z2) * Z * Ar2;
                const double e1y = (cc[ei1] * Q2 - cc[ei2] * Z2) * Z * Ar2;
                const double e2x = (bb[ei1] * Q2 - bb[ei2] * Z2) * Q * Ar2;
                const double e2y = (cc[ei1] * Q2 - cc[ei2] * Z2) * Q * Ar2;
                const double s = signA * area * Li * w;
                fv[ei].x += s * (e1x * Ux.x + e1y * Uy.x);
                fv[ei].y += s * (e1x * Ux.y + e1y * Uy.y);
                fv[ei + 4].x += s * (e2x * Ux.x + e2y * Uy.x);
                fv[ei + 4].y += s * (e2x * Ux.y + e2y * Uy.y);
            }
            const double Q = aa[1] + bb[1] * x + cc[1] * y;
            const double Z = aa[0] + bb[0] * x + cc[0] * y;
            const double FAr = 8 * area * area * area;
            const double W = (aa[2] + bb[2] * x + cc[2] * y) / FAr;
            const double W2 = Q * W;
            const double f1x = Lt1 * (-bb[0] * W2 + bb[2] * Z * Q / FAr);
            const double f1y = Lt1 * (-cc[0] * W2 + cc[2] * Z * Q / FAr);
            const double f2x = Lt2 * (bb[0] * W2 - bb[1] * Z * W);
            const double f2y = Lt2 * (cc[0] * W2 - cc[1] * Z * W);
            const double s = signA * area * w;
            fv[3].x += s * (f1x * Ux.x + f1y * Uy.x);
            fv[3].y += s * (f1x * Ux.y + f1y * Uy.y);
            fv[7].x += s * (f2x * Ux.x + f2y * Uy.x);
            fv[7].y += s * (f2x * Ux.y + f2y * Uy.y);
        }
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
