// Port overlap integrals for S-parameter extraction.
//
// A probe is a list of points per port, each with the tet that owns it, a
// complex vector c and a weight w; the device returns, per solved system and
// port, O = sum_i w_i (E(p_i) . conj(c_i)) with E interpolated from the
// solution in the owner tet (ned2_tet_interp, E/elements/ned2_interp.py:
// 141-301), and the k0-independent N = sum_i w_i |c_i|^2.
//   * Mode-overlap ports (sparam_field_power / sparam_mode_power,
//     MW/sparam.py:71-137, through surface_integral, E/mth/integrals.py:
//     60-92): the 6 Dunavant points of every port triangle, c = the mode
//     shape (e_mode = q(k0) shape), w = weight x area.
//   * Voltage-line ports (MW/microwave_3d.py:1271-1292 with
//     Line.line_integral, E/coord.py:149-153): the points of each line,
//     c = the line step vector, w = the composite-rule weight / number of
//     lines, so O is the mean line voltage.
// The host combines the scalars.
#include <algorithm>

#include "emg_common.h"
#include "emg_cplx.h"

namespace {

using namespace emg;

// partial[k][f] = w[k] * (E_f(p_k) . conj(c_k)) for probe point k, E_f from
// the reduced solution x[row][f] (eliminated DOFs = 0).
__global__ void k_port_overlap(int64_t npts, int F, const int32_t* __restrict__ owner,
                               const double* __restrict__ pts, const double2* __restrict__ shape,
                               const double* __restrict__ wa, const double* __restrict__ xyz, int64_t V,
                               const int32_t* __restrict__ tets, int64_t T, const int32_t* __restrict__ t2f,
                               const int32_t* __restrict__ keep, const double2* __restrict__ x,
                               double2* __restrict__ partial) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    const int t = owner[k];
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
    // tet_coefficients (E/elements/ned2_interp.py:47-78)
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
    const double vol = fabs(x1 * bb[0] + x2 * bb[1] + x3 * bb[2] + x4 * bb[3]) / 6.0;
    const double V1 = 216.0 * vol * vol * vol;
    // local edge / face maps: sorted global ids -> local vertex index
    const int le[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
    const int lf[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};
    auto dist = [&](int p, int q) {
        return sqrt((xv[p] - xv[q]) * (xv[p] - xv[q]) + (yv[p] - yv[q]) * (yv[p] - yv[q]) +
                    (zv[p] - zv[q]) * (zv[p] - zv[q]));
    };
    for (int f = 0; f < F; ++f) {
        double2 acc = make_double2(0.0, 0.0);
        double2 dof[20];
        for (int i = 0; i < 20; ++i) {
            const int r = keep[t2f[i * T + t]];
            dof[i] = r >= 0 ? x[int64_t(r) * F + f] : make_double2(0.0, 0.0);
        }
        {
            const double px = pts[k * 3 + 0], py = pts[k * 3 + 1], pz = pts[k * 3 + 2];
            double Fv[4];
            for (int j = 0; j < 4; ++j) Fv[j] = aa[j] + bb[j] * px + cc[j] * py + dd[j] * pz;
            double2 E[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
            for (int ie = 0; ie < 6; ++ie) {
                int e1 = le[ie][0], e2 = le[ie][1];
                if (w[e1] > w[e2]) { const int tmp = e1; e1 = e2; e2 = tmp; }
                const double L = dist(e1, e2);
                const double2 m1 = dof[ie], m2 = dof[ie + 10];
                const double2 amp = make_double2(L * (m1.x * Fv[e1] + m2.x * Fv[e2]), L * (m1.y * Fv[e1] + m2.y * Fv[e2]));
                const double g[3] = {(bb[e1] * Fv[e2] - bb[e2] * Fv[e1]) / V1, (cc[e1] * Fv[e2] - cc[e2] * Fv[e1]) / V1,
                                     (dd[e1] * Fv[e2] - dd[e2] * Fv[e1]) / V1};
                for (int d = 0; d < 3; ++d) {
                    E[d].x += amp.x * g[d];
                    E[d].y += amp.y * g[d];
                }
            }
            for (int it = 0; it < 4; ++it) {
                int v[3] = {lf[it][0], lf[it][1], lf[it][2]};
                // sort the face's local vertices by global id
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2 - a; ++b)
                        if (w[v[b]] > w[v[b + 1]]) { const int tmp = v[b]; v[b] = v[b + 1]; v[b + 1] = tmp; }
                const int i1 = v[0], i2 = v[1], i3 = v[2];
                const double L1 = dist(i1, i3), L2 = dist(i1, i2);
                const double2 q1 = make_double2(dof[it + 6].x * L1 * Fv[i2], dof[it + 6].y * L1 * Fv[i2]);
                const double2 q2 = make_double2(dof[it + 16].x * L2 * Fv[i3], dof[it + 16].y * L2 * Fv[i3]);
                const double g13[3] = {bb[i1] * Fv[i3] - bb[i3] * Fv[i1], cc[i1] * Fv[i3] - cc[i3] * Fv[i1],
                                       dd[i1] * Fv[i3] - dd[i3] * Fv[i1]};
                const double g12[3] = {bb[i1] * Fv[i2] - bb[i2] * Fv[i1], cc[i1] * Fv[i2] - cc[i2] * Fv[i1],
                                       dd[i1] * Fv[i2] - dd[i2] * Fv[i1]};
                for (int d = 0; d < 3; ++d) {
                    E[d].x += (-q1.x * g13[d] + q2.x * g12[d]) / V1;
                    E[d].y += (-q1.y * g13[d] + q2.y * g12[d]) / V1;
                }
            }
            double2 dotv = make_double2(0.0, 0.0);
            for (int d = 0; d < 3; ++d) {
                const double2 sh = shape[k * 3 + d];
                dotv.x += E[d].x * sh.x + E[d].y * sh.y;  // E . conj(c)
                dotv.y += E[d].y * sh.x - E[d].x * sh.y;
            }
            const double wq = wa[k];
            acc.x += wq * dotv.x;
            acc.y += wq * dotv.y;
        }
        partial[k * F + f] = acc;
    }
}

// out[f][p] = sum over the points k of port p (in order) of partial[k][f]
__global__ void k_port_sum(int n_ports, int F, const int64_t* __restrict__ start, const double2* __restrict__ partial,
                           double2* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_ports * F) return;
    const int p = t / F, f = t % F;
    double2 s = make_double2(0.0, 0.0);
    for (int64_t k = start[p]; k < start[p + 1]; ++k) {
        s.x += partial[k * F + f].x;
        s.y += partial[k * F + f].y;
    }
    out[int64_t(f) * n_ports + p] = s;
}

// per-point k0-independent norm w |c|^2 (as .x)
__global__ void k_port_norm_pt(int64_t npts, const double2* __restrict__ shape, const double* __restrict__ wa,
                               double2* __restrict__ part) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    double v = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double2 z = shape[k * 3 + d];
        v += z.x * z.x + z.y * z.y;
    }
    part[k] = make_double2(wa[k] * v, 0.0);
}

}  // namespace

namespace emg {

// overlap[f][p] = sum over port p's points (in order) of partial[k][f]
int port_overlaps(emg_ctx* c, int F, const double2* x, std::vector<double2>& out) {
    out.assign(size_t(F) * c->n_ports, make_double2(0.0, 0.0));
    if (c->n_ports == 0) return EMG_OK;
    DBuf<double2> part;
    EMG_HIP(c, part.alloc(c->n_probe_pts * F));
    k_port_overlap<<<unsigned((c->n_probe_pts + 63) / 64), 64, 0, c->stream>>>(
        c->n_probe_pts, F, c->probe_tet.p, c->probe_pts.p, c->probe_shape.p, c->probe_wa.p, c->xyz.p, c->n_nodes,
        c->tets.p, c->n_tets, c->t2f.p, c->keep.p, x, part.p);
    EMG_HIP(c, hipGetLastError());
    DBuf<double2> dout;
    DBuf<int64_t> dstart;
    EMG_HIP(c, dout.alloc(size_t(F) * c->n_ports));
    EMG_HIP(c, dstart.alloc(c->n_ports + 1));
    EMG_HIP(c, hipMemcpyAsync(dstart.p, c->port_pt_start.data(), sizeof(int64_t) * (c->n_ports + 1),
                              hipMemcpyHostToDevice, c->stream));
    k_port_sum<<<unsigned((c->n_ports * F + 63) / 64), 64, 0, c->stream>>>(c->n_ports, F, dstart.p, part.p, dout.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipMemcpyAsync(out.data(), dout.p, sizeof(double2) * out.size(), hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

}  // namespace emg

extern "C" int emg_port_points_set(emg_ctx* c, int32_t n_ports, const int64_t* point_counts,
                                   const int32_t* owner_tet, const double* pts, const double* vec, const double* w) {
    EMG_CHECK_CTX(c);
    if (c->n_tets == 0) return emg::fail(c, EMG_ERR_STATE, "emg_port_points_set: no mesh");
    if (n_ports < 1 || !point_counts || !owner_tet || !pts || !vec || !w)
        return emg::fail(c, EMG_ERR_ARG, "emg_port_points_set: bad arguments");
    int64_t n = 0;
    std::vector<int64_t> start(n_ports + 1, 0);
    for (int p = 0; p < n_ports; ++p) {
        if (point_counts[p] < 1) return emg::fail(c, EMG_ERR_ARG, "emg_port_points_set: empty port");
        n += point_counts[p];
        start[p + 1] = n;
    }
    for (int64_t k = 0; k < n; ++k)
        if (owner_tet[k] < 0 || owner_tet[k] >= c->n_tets)
            return emg::fail(c, EMG_ERR_ARG, "emg_port_points_set: owner tet out of range");
    hipStream_t s = c->stream;
    EMG_HIP(c, c->probe_tet.alloc(n));
    EMG_HIP(c, c->probe_pts.alloc(3 * n));
    EMG_HIP(c, c->probe_shape.alloc(3 * n));
    EMG_HIP(c, c->probe_wa.alloc(n));
    EMG_HIP(c, hipMemcpyAsync(c->probe_tet.p, owner_tet, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->probe_pts.p, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->probe_shape.p, vec, sizeof(double2) * 3 * n, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->probe_wa.p, w, sizeof(double) * n, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    c->n_ports = n_ports;
    c->n_probe_pts = n;
    c->port_pt_start = start;
    // k0-independent norms: per point on the device, then per port in order
    DBuf<double2> part, dout;
    DBuf<int64_t> dstart;
    EMG_HIP(c, part.alloc(n));
    EMG_HIP(c, dout.alloc(n_ports));
    EMG_HIP(c, dstart.alloc(n_ports + 1));
    EMG_HIP(c, hipMemcpyAsync(dstart.p, start.data(), sizeof(int64_t) * (n_ports + 1), hipMemcpyHostToDevice, s));
    k_port_norm_pt<<<unsigned((n + 63) / 64), 64, 0, s>>>(n, c->probe_shape.p, c->probe_wa.p, part.p);
    k_port_sum<<<unsigned((n_ports + 63) / 64), 64, 0, s>>>(n_ports, 1, dstart.p, part.p, dout.p);
    EMG_HIP(c, hipGetLastError());
    std::vector<double2> h(n_ports);
    EMG_HIP(c, hipMemcpyAsync(h.data(), dout.p, sizeof(double2) * n_ports, hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    c->port_norm.assign(n_ports, 0.0);
    for (int p = 0; p < n_ports; ++p) c->port_norm[p] = h[p].x;
    return EMG_OK;
}

extern "C" int emg_port_probe_set(emg_ctx* c, int32_t n_ports, const int64_t* tri_counts, const int32_t* owner_tet,
                                  const double* pts, const double* shape, const double* wa) {
    EMG_CHECK_CTX(c);
    if (n_ports < 1 || !tri_counts || !owner_tet || !pts || !shape || !wa)
        return emg::fail(c, EMG_ERR_ARG, "emg_port_probe_set: bad arguments");
    int64_t n = 0;
    std::vector<int64_t> cnt(n_ports);
    for (int p = 0; p < n_ports; ++p) {
        if (tri_counts[p] < 1) return emg::fail(c, EMG_ERR_ARG, "emg_port_probe_set: empty port");
        n += tri_counts[p];
        cnt[p] = 6 * tri_counts[p];
    }
    // the 6 Dunavant points of a triangle share its owner tet
    std::vector<int32_t> own(6 * n);
    for (int64_t k = 0; k < n; ++k)
        for (int q = 0; q < 6; ++q) own[6 * k + q] = owner_tet[k];
    return emg_port_points_set(c, n_ports, cnt.data(), own.data(), pts, shape, wa);
}

extern "C" int emg_port_norms_get(emg_ctx* c, double* norms) {
    EMG_CHECK_CTX(c);
    if (c->n_ports == 0) return emg::fail(c, EMG_ERR_STATE, "emg_port_norms_get: no port probes");
    for (int p = 0; p < c->n_ports; ++p) norms[p] = c->port_norm[p];
    return EMG_OK;
}
