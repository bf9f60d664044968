// Element matrices of the boundary-mode (ModalPort) eigenproblem of a port
// face: the mixed second-order Nedelec (transverse, 8 functions) / Lagrange
// (longitudinal, 6 functions) triangle of ASM/generalized_eigen_hb.py and
// E/elements/nedleg2.py.
//
// Replaces generalized_matrix_GQ + _matrix_builder
// (ASM/generalized_eigen_hb.py:247-458).  One workgroup per triangle, one
// lane per entry (i, j) of the 14 x 14 pair: the lane evaluates its two
// basis functions at the six points of the rule and accumulates
//   A = [Att - k0^2 Btt, 0; 0, 0],  B = [Dtt, Dzt^T; Dzt, Dzz1 - k0^2 Dzz2]
// (times |area|), Att = (curl N_i, mu_zz^-1 curl N_j), Btt = (N_i, eps_t N_j),
// Dtt = (N_i, mu_t^-1 N_j), Dzt = (grad L_k, mu_t^-1 N_j),
// Dzz1 = (grad L_k, mu_t^-1 grad L_l), Dzz2 = (L_k, eps_zz L_l).  As in the
// reference, the face-face entries of Dtt are left at zero and the face rows
// of the transverse blocks copy the face columns.
#include "emg_common.h"

namespace {

using namespace emg;

// 6-point rule as tabulated (8 digits) in ASM/generalized_eigen_hb.py:236-241
__constant__ double kW6[6] = {0.22338159, 0.22338159, 0.22338159, 0.10995174, 0.10995174, 0.10995174};
__constant__ double kL6[3][6] = {{0.10810302, 0.44594849, 0.44594849, 0.81684757, 0.09157621, 0.09157621},
                                 {0.44594849, 0.44594849, 0.10810302, 0.09157621, 0.09157621, 0.81684757},
                                 {0.44594849, 0.10810302, 0.44594849, 0.09157621, 0.81684757, 0.09157621}};

// triangle edges over sorted local vertices: (0,1), (1,2), (0,2)
__device__ __forceinline__ void edge_of(int e, int& p, int& q) {
    p = (e == 1) ? 1 : 0;
    q = (e == 0) ? 1 : 2;
}

struct Tri {
    double b[3], c[3];  // grad lambda_i = (b_i, c_i)
    double a[3];
};

// lambda_i p_i x grad lambda_j - lambda_j grad lambda_i (vector) and its curl
__device__ __forceinline__ void whit(const Tri& T, const double* L, int i, int j, double& vx, double& vy,
                                     double& cz) {
    vx = L[i] * T.b[j] - L[j] * T.b[i];
    vy = L[i] * T.c[j] - L[j] * T.c[i];
    cz = 2.0 * (T.b[i] * T.c[j] - T.c[i] * T.b[j]);
}

// f w: value and curl (curl(f w) = f curl w + grad f x w)
__device__ __forceinline__ void scaled(double f, double fx, double fy, double wx, double wy, double wc,
                                       double& vx, double& vy, double& cz) {
    vx = f * wx;
    vy = f * wy;
    cz = f * wc + (fx * wy - fy * wx);
}

// transverse function k (0-2 edge first kind, 3 face 1, 4-6 edge second
// kind, 7 face 2) at barycentrics L: value (vx, vy) and curl cz
__device__ void eval_t(const Tri& T, const double* L, int k, double& vx, double& vy, double& cz) {
    if (k < 3 || (k >= 4 && k < 7)) {
        const int e = k < 3 ? k : k - 4;
        int p, q;
        edge_of(e, p, q);
        double wx, wy, wc;
        whit(T, L, p, q, wx, wy, wc);
        if (k < 3) {
            vx = wx;
            vy = wy;
            cz = wc;
        } else {
            scaled(L[p] - L[q], T.b[p] - T.b[q], T.c[p] - T.c[q], wx, wy, wc, vx, vy, cz);
        }
        return;
    }
    // face 1: lam1 w(0,2) + lam0 w(1,2); face 2: lam2 w(1,0) + lam1 w(2,0)
    const int f1 = k == 3 ? 1 : 2, i1 = k == 3 ? 0 : 1, j1 = k == 3 ? 2 : 0;
    const int f2 = k == 3 ? 0 : 1, i2 = k == 3 ? 1 : 2, j2 = k == 3 ? 2 : 0;
    double wx, wy, wc, ax, ay, ac, bx, by, bc;
    whit(T, L, i1, j1, wx, wy, wc);
    scaled(L[f1], T.b[f1], T.c[f1], wx, wy, wc, ax, ay, ac);
    whit(T, L, i2, j2, wx, wy, wc);
    scaled(L[f2], T.b[f2], T.c[f2], wx, wy, wc, bx, by, bc);
    vx = ax + bx;
    vy = ay + by;
    cz = ac + bc;
}

// longitudinal function k (0-2 vertex, 3-5 edge): value and gradient
__device__ void eval_z(const Tri& T, const double* L, int k, double& v, double& gx, double& gy) {
    if (k < 3) {
        v = L[k] * (2.0 * L[k] - 1.0);
        const double s = 4.0 * L[k] - 1.0;
        gx = T.b[k] * s;
        gy = T.c[k] * s;
        return;
    }
    int p, q;
    edge_of(k - 3, p, q);
    v = 4.0 * L[p] * L[q];
    gx = 4.0 * (T.b[p] * L[q] + T.b[q] * L[p]);
    gy = 4.0 * (T.c[p] * L[q] + T.c[q] * L[p]);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// u . (M v) for a 2x2 complex block M (row-major m00 m01 m10 m11), real u, v
__device__ __forceinline__ double2 quad2(const double2* M, double ux, double uy, double vx, double vy) {
    return make_double2(ux * (M[0].x * vx + M[1].x * vy) + uy * (M[2].x * vx + M[3].x * vy),
                        ux * (M[0].y * vx + M[1].y * vy) + uy * (M[2].y * vx + M[3].y * vy));
}

__global__ __launch_bounds__(256) void k_port_mode_elements(int64_t n, const double* __restrict__ xy,
                                                            const double2* __restrict__ eps,
                                                            const double2* __restrict__ imu, double k0,
                                                            double2* __restrict__ A, double2* __restrict__ B) {
    const int64_t t = blockIdx.x;
    if (t >= n) return;
    const int lane = threadIdx.x;
    if (lane >= 196) return;
    // the face rows of the transverse blocks copy the face columns
    int i = lane / 14, j = lane % 14;
    const bool face_row = (i == 3 || i == 7) && j < 8 && j != 3 && j != 7;
    const int ri = face_row ? j : i, rj = face_row ? i : j;

    const double x1 = xy[6 * t + 0], y1 = xy[6 * t + 1];
    const double x2 = xy[6 * t + 2], y2 = xy[6 * t + 3];
    const double x3 = xy[6 * t + 4], y3 = xy[6 * t + 5];
    // barycentric coefficients (ASM/generalized_eigen_hb.py:207-230)
    const double sA = 0.5 * ((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1));
    const double sg = sA > 0 ? 1.0 : (sA < 0 ? -1.0 : 0.0);
    const double area = fabs(sA);
    const double inv = 1.0 / (2.0 * area);
    Tri T;
    T.a[0] = sg * (x2 * y3 - y2 * x3) * inv;
    T.a[1] = sg * (x3 * y1 - y3 * x1) * inv;
    T.a[2] = sg * (x1 * y2 - y1 * x2) * inv;
    T.b[0] = sg * (y2 - y3) * inv;
    T.b[1] = sg * (y3 - y1) * inv;
    T.b[2] = sg * (y1 - y2) * inv;
    T.c[0] = sg * (x3 - x2) * inv;
    T.c[1] = sg * (x1 - x3) * inv;
    T.c[2] = sg * (x2 - x1) * inv;

    const double2* E3 = eps + 9 * t;  // eps_r, row-major 3x3
    const double2* U3 = imu + 9 * t;  // mu_r^-1
    const double2 Et[4] = {E3[0], E3[1], E3[3], E3[4]};
    const double2 Ut[4] = {U3[0], U3[1], U3[3], U3[4]};
    const double2 ezz = E3[8], uzz = U3[8];

    double2 a = make_double2(0, 0), b = make_double2(0, 0);
    for (int qd = 0; qd < 6; ++qd) {
        const double px = kL6[0][qd] * x1 + kL6[1][qd] * x2 + kL6[2][qd] * x3;
        const double py = kL6[0][qd] * y1 + kL6[1][qd] * y2 + kL6[2][qd] * y3;
        double L[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) L[v] = T.a[v] + T.b[v] * px + T.c[v] * py;
        const double w = kW6[qd];
        if (ri < 8 && rj < 8) {
            double ux, uy, uc, vx, vy, vc;
            eval_t(T, L, ri, ux, uy, uc);
            eval_t(T, L, rj, vx, vy, vc);
            const double cc = uc * vc;
            const double2 bt = quad2(Et, ux, uy, vx, vy);
            a.x += w * (uzz.x * cc - k0 * k0 * bt.x);
            a.y += w * (uzz.y * cc - k0 * k0 * bt.y);
            if (!((ri == 3 || ri == 7) && (rj == 3 || rj == 7))) {
                const double2 dt = quad2(Ut, ux, uy, vx, vy);
                b.x += w * dt.x;
                b.y += w * dt.y;
            }
        } else if (ri >= 8 && rj < 8) {
            double v, gx, gy, ux, uy, uc;
            eval_z(T, L, ri - 8, v, gx, gy);
            eval_t(T, L, rj, ux, uy, uc);
            const double2 d = quad2(Ut, gx, gy, ux, uy);
            b.x += w * d.x;
            b.y += w * d.y;
        } else if (ri < 8 && rj >= 8) {
            double v, gx, gy, ux, uy, uc;
            eval_z(T, L, rj - 8, v, gx, gy);
            eval_t(T, L, ri, ux, uy, uc);
            const double2 d = quad2(Ut, gx, gy, ux, uy);  // Dzt^T
            b.x += w * d.x;
            b.y += w * d.y;
        } else {
            double u, ugx, ugy, v, vgx, vgy;
            eval_z(T, L, ri - 8, u, ugx, ugy);
            eval_z(T, L, rj - 8, v, vgx, vgy);
            const double2 d1 = quad2(Ut, ugx, ugy, vgx, vgy);
            const double uv = u * v;
            b.x += w * (d1.x - k0 * k0 * ezz.x * uv);
            b.y += w * (d1.y - k0 * k0 * ezz.y * uv);
        }
    }
    A[196 * t + lane] = make_double2(a.x * area, a.y * area);
    B[196 * t + lane] = make_double2(b.x * area, b.y * area);
}

}  // namespace

extern "C" int emg_port_mode_elements(emg_ctx* c, int64_t n, const double* xy, const double* eps,
                                      const double* inv_mu, double k0, double* A, double* B) {
    EMG_CHECK_CTX(c);
    if (n < 1) return fail(c, EMG_ERR_ARG, "emg_port_mode_elements: need at least one triangle");
    if (!xy || !eps || !inv_mu || !A || !B) return fail(c, EMG_ERR_ARG, "emg_port_mode_elements: null array");
    if (n > (int64_t(1) << 31)) return fail(c, EMG_ERR_ARG, "emg_port_mode_elements: more than 2^31 triangles");
    hipStream_t st = c->stream;
    DBuf<double> dxy;
    DBuf<double2> de, du, dA, dB;
    EMG_HIP(c, dxy.alloc(6 * n));
    EMG_HIP(c, de.alloc(9 * n));
    EMG_HIP(c, du.alloc(9 * n));
    EMG_HIP(c, dA.alloc(196 * n));
    EMG_HIP(c, dB.alloc(196 * n));
    EMG_HIP(c, hipMemcpyAsync(dxy.p, xy, sizeof(double) * 6 * n, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(de.p, eps, sizeof(double2) * 9 * n, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(du.p, inv_mu, sizeof(double2) * 9 * n, hipMemcpyHostToDevice, st));
    k_port_mode_elements<<<unsigned(n), 256, 0, st>>>(n, dxy.p, de.p, du.p, k0, dA.p, dB.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipMemcpyAsync(A, dA.p, sizeof(double2) * 196 * n, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipMemcpyAsync(B, dB.p, sizeof(double2) * 196 * n, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    return EMG_OK;
}
