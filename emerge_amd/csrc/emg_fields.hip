// Field of a solution at arbitrary points: E and curl E (the basis of
// MWField.interpolate, MW/microwave_data.py:402-430, through
// Nedelec2.interpolate / interpolate_curl, E/elements/nedelec2.py:68-92,
// and ned2_tet_interp / ned2_tet_interp_curl, E/elements/ned2_interp.py:
// 141-507).
//
// Owner rule as the reference's: a point belongs to the LAST tet (highest
// index) whose barycentric coordinates pass sum <= 1 + 1e-8 and each >= -1e-6.
// A uniform grid (built once per mesh) lists in every cell, in ascending
// order, the tets whose bounding box -- widened by 1e-5 of the tet's size,
// beyond the tolerances -- meets the cell; every tet that can contain a point
// is in the point's cell, so a lane scans that list from its end and stops
// at the first hit.
//
// In the owner tet, with F_i = 6V lam_i = a_i + b_i x + c_i y + d_i z and
// g_i = (b_i, c_i, d_i) (tet_coefficients, ned2_interp.py:47-78), for edge
// (e1, e2) sorted by global id with DOFs m1, m2 and length L:
//   E += L (m1 F1 + m2 F2) (F2 g1 - F1 g2) / (216 V^3),
//   curl E -= L (m1 F1 + m2 F2) (g1 x g2) / (72 V^3);
// for face (i1, i2, i3) sorted by global id with DOFs m1, m2, L1 = |i1 i3|,
// L2 = |i1 i2|:
//   E += [-m1 L1 F2 (F3 g1 - F1 g3) + m2 L2 F3 (F2 g1 - F1 g2)] / (216 V^3),
//   curl E += [-m1 L1 (g2 x (F3 g1 - F1 g3) + 2 F2 g3 x g1)
//              + m2 L2 (g3 x (F2 g1 - F1 g2) + 2 F3 g2 x g1)] / (216 V^3)
// (the reference writes the curl expanded; this is the same polynomial).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "emg_common.h"

namespace {

using namespace emg;

__device__ inline void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline bool in_tet(int64_t t, double px, double py, double pz, const double* __restrict__ xyz,
                              int64_t V, const int32_t* __restrict__ tets, int64_t T) {
    const int32_t w0 = tets[t], w1 = tets[T + t], w2 = tets[2 * T + t], w3 = tets[3 * T + t];
    const double ox = xyz[w0], oy = xyz[V + w0], oz = xyz[2 * V + w0];
    const double e1[3] = {xyz[w1] - ox, xyz[V + w1] - oy, xyz[2 * V + w1] - oz};
    const double e2[3] = {xyz[w2] - ox, xyz[V + w2] - oy, xyz[2 * V + w2] - oz};
    const double e3[3] = {xyz[w3] - ox, xyz[V + w3] - oy, xyz[2 * V + w3] - oz};
    const double r[3] = {px - ox, py - oy, pz - oz};
    double n23[3], n31[3], n12[3];
    cross3(e2, e3, n23);
    cross3(e3, e1, n31);
    cross3(e1, e2, n12);
    const double det = e1[0] * n23[0] + e1[1] * n23[1] + e1[2] * n23[2];
    const double l0 = (r[0] * n23[0] + r[1] * n23[1] + r[2] * n23[2]) / det;
    const double l1 = (r[0] * n31[0] + r[1] * n31[1] + r[2] * n31[2]) / det;
    const double l2 = (r[0] * n12[0] + r[1] * n12[1] + r[2] * n12[2]) / det;
    return l0 + l1 + l2 <= 1.00000001 && l0 >= -1e-6 && l1 >= -1e-6 && l2 >= -1e-6;
}

struct Grid {
    double lo[3];
    double h;
    int32_t n[3];
};

__device__ inline int cell_axis(const Grid& g, int a, double v) {
    const double q = floor((v - g.lo[a]) / g.h);
    return q < 0.0 ? 0 : (q >= double(g.n[a]) ? g.n[a] - 1 : int(q));
}

// cells covered by tet t's widened bounding box: per-axis ranges
__device__ inline void tet_cells(const Grid& g, int64_t t, const double* __restrict__ xyz, int64_t V,
                                 const int32_t* __restrict__ tets, int64_t T, int* c0, int* c1) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = 1e300;
        mx[a] = -1e300;
    }
    for (int j = 0; j < 4; ++j) {
        const int32_t w = tets[j * T + t];
        for (int a = 0; a < 3; ++a) {
            const double v = xyz[a * V + w];
            mn[a] = fmin(mn[a], v);
            mx[a] = fmax(mx[a], v);
        }
    }
    const double size = fmax(mx[0] - mn[0], fmax(mx[1] - mn[1], mx[2] - mn[2]));
    const double pad = 1e-5 * size;
    for (int a = 0; a < 3; ++a) {
        c0[a] = cell_axis(g, a, mn[a] - pad);
        c1[a] = cell_axis(g, a, mx[a] + pad);
    }
}

__global__ void k_grid_count(int64_t T, Grid g, const double* __restrict__ xyz, int64_t V,
                             const int32_t* __restrict__ tets, int64_t* __restrict__ count) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int c0[3], c1[3];
    tet_cells(g, t, xyz, V, tets, T, c0, c1);
    count[t] = int64_t(c1[0] - c0[0] + 1) * (c1[1] - c0[1] + 1) * (c1[2] - c0[2] + 1);
}

__global__ void k_grid_fill(int64_t T, Grid g, const double* __restrict__ xyz, int64_t V,
                            const int32_t* __restrict__ tets, const int64_t* __restrict__ off,
                            uint32_t* __restrict__ key, int32_t* __restrict__ val) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int c0[3], c1[3];
    tet_cells(g, t, xyz, V, tets, T, c0, c1);
    int64_t o = off[t];
    for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
            for (int x = c0[0]; x <= c1[0]; ++x) {
                key[o] = uint32_t((int64_t(z) * g.n[1] + y) * g.n[0] + x);
                val[o] = int32_t(t);
                ++o;
            }
}

// ptr[c] = first sorted entry of cell c (lower bound), ptr[ncells] = total
__global__ void k_grid_ptr(int64_t ncells, const uint32_t* __restrict__ key, int64_t total,
                           int64_t* __restrict__ ptr) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    if (c == ncells) {
        ptr[c] = total;
        return;
    }
    int64_t lo = 0, hi = total;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (int64_t(key[mid]) < c) lo = mid + 1;
        else hi = mid;
    }
    ptr[c] = lo;
}

__global__ void k_field_points(int64_t P, const double* __restrict__ pts, const double* __restrict__ xyz, int64_t V,
                               const int32_t* __restrict__ tets, int64_t T, const int32_t* __restrict__ t2f,
                               Grid grid, const int64_t* __restrict__ gptr, const int32_t* __restrict__ gtets,
                               const double2* __restrict__ x, double2* __restrict__ Eo, double2* __restrict__ Co,
                               int32_t* __restrict__ owner) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const double px = pts[k], py = pts[P + k], pz = pts[2 * P + k];
    const int64_t cell = (int64_t(cell_axis(grid, 2, pz)) * grid.n[1] + cell_axis(grid, 1, py)) * grid.n[0] +
                         cell_axis(grid, 0, px);
    int64_t t = -1;
    for (int64_t q = gptr[cell + 1] - 1; q >= gptr[cell]; --q) {
        if (in_tet(gtets[q], px, py, pz, xyz, V, tets, T)) {
            t = gtets[q];
            break;
        }
    }
    owner[k] = int32_t(t);
    if (t < 0) {
        const double nan = __builtin_nan("");
        for (int d = 0; d < 3; ++d) {
            Eo[d * P + k] = make_double2(nan, nan);
            Co[d * P + k] = make_double2(nan, nan);
        }
        return;
    }
    int w[4];
    double xv[4], yv[4], zv[4];
    for (int j = 0; j < 4; ++j) {
        w[j] = tets[j * T + t];
        xv[j] = xyz[w[j]];
        yv[j] = xyz[V + w[j]];
        zv[j] = xyz[2 * V + w[j]];
    }
    const double x1 = xv[0], x2 = xv[1], x3 = xv[2], x4 = xv[3];
    const double y1 = yv[0], y2 = yv[1], y3 = yv[2], y4 = yv[3];
    const double z1 = zv[0], z2 = zv[1], z3 = zv[2], z4 = zv[3];
    // cofactors (tet_coefficients, E/elements/ned2_interp.py:47-78): F_j = 6V lam_j
    const double aa[4] = {x2*y3*z4 - x2*y4*z3 - x3*y2*z4 + x3*y4*z2 + x4*y2*z3 - x4*y3*z2,
                          -x1*y3*z4 + x1*y4*z3 + x3*y1*z4 - x3*y4*z1 - x4*y1*z3 + x4*y3*z1,
                          x1*y2*z4 - x1*y4*z2 - x2*y1*z4 + x2*y4*z1 + x4*y1*z2 - x4*y2*z1,
                          -x1*y2*z3 + x1*y3*z2 + x2*y1*z3 - x2*y3*z1 - x3*y1*z2 + x3*y2*z1};
    const double bb[4] = {-y2*z3 + y2*z4 + y3*z2 - y3*z4 - y4*z2 + y4*z3,
                          y1*z3 - y1*z4 - y3*z1 + y3*z4 + y4*z1 - y4*z3,
                          -y1*z2 + y1*z4 + y2*z1 - y2*z4 - y4*z1 + y4*z2,
                          y1*z2 - y1*z3 - y2*z1 + y2*z3 + y3*z1 - y3*z2};
    const double cc[4] = {x2*z3 - x2*z4 - x3*z2 + x3*z4 + x4*z2 - x4*z3,
                          -x1*z3 + x1*z4 + x3*z1 - x3*z4 - x4*z1 + x4*z3,
                          x1*z2 - x1*z4 - x2*z1 + x2*z4 + x4*z1 - x4*z2,
                          -x1*z2 + x1*z3 + x2*z1 - x2*z3 - x3*z1 + x3*z2};
    const double dd[4] = {-x2*y3 + x2*y4 + x3*y2 - x3*y4 - x4*y2 + x4*y3,
                          x1*y3 - x1*y4 - x3*y1 + x3*y4 + x4*y1 - x4*y3,
                          -x1*y2 + x1*y4 + x2*y1 - x2*y4 - x4*y1 + x4*y2,
                          x1*y2 - x1*y3 - x2*y1 + x2*y3 + x3*y1 - x3*y2};
    const double vol = fabs(xv[0] * bb[0] + xv[1] * bb[1] + xv[2] * bb[2] + xv[3] * bb[3]) / 6.0;
    const double V1 = 216.0 * vol * vol * vol;
    const double V2 = 72.0 * vol * vol * vol;
    double Fv[4], g[4][3];
    for (int j = 0; j < 4; ++j) {
        Fv[j] = aa[j] + bb[j] * px + cc[j] * py + dd[j] * pz;
        g[j][0] = bb[j];
        g[j][1] = cc[j];
        g[j][2] = dd[j];
    }
    auto dist = [&](int p, int q) {
        return sqrt((xv[p] - xv[q]) * (xv[p] - xv[q]) + (yv[p] - yv[q]) * (yv[p] - yv[q]) +
                    (zv[p] - zv[q]) * (zv[p] - zv[q]));
    };
    double2 dof[20];
    for (int i = 0; i < 20; ++i) dof[i] = x[t2f[i * T + t]];
    double E[3][2] = {{0, 0}, {0, 0}, {0, 0}}, C[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    const int le[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
    const int lf[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};
    for (int ie = 0; ie < 6; ++ie) {
        int e1 = le[ie][0], e2 = le[ie][1];
        if (w[e1] > w[e2]) { const int tmp = e1; e1 = e2; e2 = tmp; }
        const double L = dist(e1, e2);
        const double2 m1 = dof[ie], m2 = dof[ie + 10];
        const double phr = L * (m1.x * Fv[e1] + m2.x * Fv[e2]), phi = L * (m1.y * Fv[e1] + m2.y * Fv[e2]);
        double gx[3];
        cross3(g[e1], g[e2], gx);
        for (int d = 0; d < 3; ++d) {
            const double u = (g[e1][d] * Fv[e2] - g[e2][d] * Fv[e1]) / V1;
            E[d][0] += phr * u;
            E[d][1] += phi * u;
            C[d][0] -= phr * gx[d] / V2;
            C[d][1] -= phi * gx[d] / V2;
        }
    }
    for (int it = 0; it < 4; ++it) {
        int v[3] = {lf[it][0], lf[it][1], lf[it][2]};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2 - a; ++b)
                if (w[v[b]] > w[v[b + 1]]) { const int tmp = v[b]; v[b] = v[b + 1]; v[b + 1] = tmp; }
        const int i1 = v[0], i2 = v[1], i3 = v[2];
        const double L1 = dist(i1, i3), L2 = dist(i1, i2);
        const double2 m1 = dof[it + 6], m2 = dof[it + 16];
        const double F1 = Fv[i1], F2 = Fv[i2], F3 = Fv[i3];
        double a13[3], a12[3];
        for (int d = 0; d < 3; ++d) {
            a13[d] = F3 * g[i1][d] - F1 * g[i3][d];
            a12[d] = F2 * g[i1][d] - F1 * g[i2][d];
        }
        double c1[3], c2[3], x31[3], x21[3];
        cross3(g[i2], a13, c1);
        cross3(g[i3], a12, c2);
        cross3(g[i3], g[i1], x31);
        cross3(g[i2], g[i1], x21);
        for (int d = 0; d < 3; ++d) {
            const double t1 = c1[d] + 2.0 * F2 * x31[d];
            const double t2 = c2[d] + 2.0 * F3 * x21[d];
            // -m1 L1 (F2 a13, t1) + m2 L2 (F3 a12, t2), over V1
            E[d][0] += (-m1.x * L1 * F2 * a13[d] + m2.x * L2 * F3 * a12[d]) / V1;
            E[d][1] += (-m1.y * L1 * F2 * a13[d] + m2.y * L2 * F3 * a12[d]) / V1;
            C[d][0] += (-m1.x * L1 * t1 + m2.x * L2 * t2) / V1;
            C[d][1] += (-m1.y * L1 * t1 + m2.y * L2 * t2) / V1;
        }
    }
    for (int d = 0; d < 3; ++d) {
        Eo[d * P + k] = make_double2(E[d][0], E[d][1]);
        Co[d * P + k] = make_double2(C[d][0], C[d][1]);
    }
}

// Grid of about one tet per cell (edge h = cbrt(6 vol / T)), at most 1024
// cells per axis and 2^26 cells; points and tets use the same clamped cell
// index, so points outside the grid land in a border cell and find no tet.
static int build_grid(emg_ctx* c) {
    hipStream_t st = c->stream;
    const int64_t V = c->n_nodes, T = c->n_tets;
    std::vector<double> h(3 * V);
    EMG_HIP(c, hipMemcpyAsync(h.data(), c->xyz.p, sizeof(double) * 3 * V, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    Grid g;
    double ext[3];
    for (int a = 0; a < 3; ++a) {
        const auto mm = std::minmax_element(h.begin() + a * V, h.begin() + (a + 1) * V);
        g.lo[a] = *mm.first;
        ext[a] = *mm.second - *mm.first;
    }
    const double emax = std::max(ext[0], std::max(ext[1], ext[2]));
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) vol *= std::max(ext[a], 1e-9 * emax);
    double cs = std::cbrt(6.0 * vol / double(T));
    if (!(cs > 0.0)) cs = emax > 0.0 ? emax : 1.0;
    for (int pass = 0; pass < 2; ++pass) {
        int64_t cells = 1;
        for (int a = 0; a < 3; ++a) {
            g.n[a] = int32_t(std::min<double>(1024.0, std::max(1.0, std::floor(ext[a] / cs) + 1.0)));
            cells *= g.n[a];
        }
        if (cells <= (int64_t(1) << 26)) break;
        cs *= std::cbrt(double(cells) / double(int64_t(1) << 26)) * 1.01;
    }
    g.h = cs;
    const int64_t ncells = int64_t(g.n[0]) * g.n[1] * g.n[2];
    DBuf<int64_t> cnt, off;
    EMG_HIP(c, cnt.alloc(T));
    EMG_HIP(c, off.alloc(T));
    const unsigned bt = unsigned((T + 255) / 256);
    k_grid_count<<<bt, 256, 0, st>>>(T, g, c->xyz.p, V, c->tets.p, cnt.p);
    EMG_HIP(c, hipGetLastError());
    size_t sb = 0;
    EMG_HIP(c, rocprim::exclusive_scan(nullptr, sb, cnt.p, off.p, int64_t(0), size_t(T), rocprim::plus<int64_t>(), st));
    DBuf<unsigned char> tmp;
    EMG_HIP(c, tmp.alloc(std::max<size_t>(sb, 1)));
    EMG_HIP(c, rocprim::exclusive_scan(tmp.p, sb, cnt.p, off.p, int64_t(0), size_t(T), rocprim::plus<int64_t>(), st));
    int64_t last_off = 0, last_cnt = 0;
    EMG_HIP(c, hipMemcpyAsync(&last_off, off.p + (T - 1), sizeof(int64_t), hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipMemcpyAsync(&last_cnt, cnt.p + (T - 1), sizeof(int64_t), hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    const int64_t total = last_off + last_cnt;
    if (total >= (int64_t(1) << 31))
        return fail(c, EMG_ERR_ARG, "emg_field_points: point grid too large (%lld entries)", (long long)total);
    DBuf<uint32_t> k_in, k_out;
    DBuf<int32_t> v_in;
    EMG_HIP(c, k_in.alloc(total));
    EMG_HIP(c, k_out.alloc(total));
    EMG_HIP(c, v_in.alloc(total));
    EMG_HIP(c, c->grid_tets.alloc(total));
    k_grid_fill<<<bt, 256, 0, st>>>(T, g, c->xyz.p, V, c->tets.p, off.p, k_in.p, v_in.p);
    EMG_HIP(c, hipGetLastError());
    int bits = 1;
    while ((int64_t(1) << bits) < ncells) ++bits;
    size_t rb = 0;
    EMG_HIP(c, rocprim::radix_sort_pairs(nullptr, rb, k_in.p, k_out.p, v_in.p, c->grid_tets.p, size_t(total), 0u,
                                              unsigned(bits), st));
    EMG_HIP(c, tmp.alloc(std::max<size_t>(rb, 1)));
    EMG_HIP(c, rocprim::radix_sort_pairs(tmp.p, rb, k_in.p, k_out.p, v_in.p, c->grid_tets.p, size_t(total), 0u,
                                              unsigned(bits), st));
    EMG_HIP(c, c->grid_ptr.alloc(ncells + 1));
    k_grid_ptr<<<unsigned((ncells + 1 + 255) / 256), 256, 0, st>>>(ncells, k_out.p, total, c->grid_ptr.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(st));
    for (int a = 0; a < 3; ++a) {
        c->grid_lo[a] = g.lo[a];
        c->grid_n[a] = g.n[a];
    }
    c->grid_h = g.h;
    c->grid_ready = true;
    return EMG_OK;
}

}  // namespace

extern "C" int emg_field_points(emg_ctx* c, const double* field, int64_t P, const double* pts, double* E,
                                double* curlE, int32_t* owner) {
    EMG_CHECK_CTX(c);
    if (c->n_tets == 0) return fail(c, EMG_ERR_STATE, "emg_field_points: no mesh");
    if (P < 1) return fail(c, EMG_ERR_ARG, "emg_field_points: no points");
    if (!field || !pts || !E || !curlE || !owner) return fail(c, EMG_ERR_ARG, "emg_field_points: null array");
    if (!c->grid_ready) {
        const int rc = build_grid(c);
        if (rc != EMG_OK) return rc;
    }
    Grid g;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = c->grid_lo[a];
        g.n[a] = c->grid_n[a];
    }
    g.h = c->grid_h;
    hipStream_t st = c->stream;
    const int64_t N = c->n_field;
    DBuf<double>& dx = c->fp_x;
    DBuf<double>& dp = c->fp_pts;
    DBuf<double>& dE = c->fp_E;
    DBuf<double>& dC = c->fp_C;
    DBuf<int32_t>& dow = c->fp_own;
    EMG_HIP(c, dx.alloc(2 * N));
    EMG_HIP(c, dp.alloc(3 * P));
    EMG_HIP(c, dE.alloc(6 * P));
    EMG_HIP(c, dC.alloc(6 * P));
    EMG_HIP(c, dow.alloc(P));
    EMG_HIP(c, hipMemcpyAsync(dx.p, field, sizeof(double) * 2 * N, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(dp.p, pts, sizeof(double) * 3 * P, hipMemcpyHostToDevice, st));
    k_field_points<<<unsigned((P + 127) / 128), 128, 0, st>>>(
        P, dp.p, c->xyz.p, c->n_nodes, c->tets.p, c->n_tets, c->t2f.p, g, c->grid_ptr.p, c->grid_tets.p,
        reinterpret_cast<const double2*>(dx.p), reinterpret_cast<double2*>(dE.p), reinterpret_cast<double2*>(dC.p),
        dow.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipMemcpyAsync(E, dE.p, sizeof(double) * 6 * P, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipMemcpyAsync(curlE, dC.p, sizeof(double) * 6 * P, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipMemcpyAsync(owner, dow.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    return EMG_OK;
}
