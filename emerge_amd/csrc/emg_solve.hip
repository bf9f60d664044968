// Batched restarted GMRES for the complex Helmholtz systems
//     A_f x_f = b_f,   A_f = E - k2_f B + sum_bc gamma_{bc,f} S_bc
// on the PEC-reduced pattern (frequency sweep), or for one CSR matrix with
// several right-hand sides (Solver plug-in path).
//
// Replaces the per-frequency direct solve of SolveRoutine.solve
// (E/solver.py:1241-1326; SuperLU/UMFPACK/PARDISO/cuDSS, :535-773) and the
// per-frequency K = E - k0^2 B + B_p combination of
// Assembler.assemble_freq_matrix (ASM/assembler.py:255, 349-350).
//
// Algorithm (published methods): GMRES(m) (Saad & Schultz 1986) with right
// Jacobi scaling, classical Gram-Schmidt with one re-orthogonalisation,
// complex Givens rotations; every restart recomputes the true residual
// b - A x in FP64, so with an FP32 inner solver the outer loop is mixed
// precision GMRES-IR (Carson & Higham 2018).  The F systems of a batch are
// stored interleaved ([row][f]) so one pass over the matrix serves all of
// them.  All reductions run in a fixed order: results are reproducible.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "emg_common.h"
#include "emg_cplx.h"

namespace {

using namespace emg;

template <typename R>
struct Op {
    int64_t n = 0;
    const int32_t* ptr = nullptr;
    const int32_t* idx = nullptr;
    const cplx<R>* E = nullptr;    // matrix values (or the stiffness part)
    const cplx<R>* B = nullptr;    // optional mass part, scaled by -k2[f]
    const int32_t* sptr = nullptr; // optional surface entries, CSR over rows
    const int32_t* scol = nullptr;
    const int32_t* sbc = nullptr;
    const R* sval = nullptr;
    const int32_t* diag = nullptr; // slot of the diagonal per row (-1 if none)
};

inline unsigned grid_for(int64_t n, int block) { return unsigned((n + block - 1) / block); }

// y = A_f (dinv o x) for every f of the batch; one wavefront per row, lanes
// split as (entry group, f).
template <typename R, int F>
__global__ __launch_bounds__(256) void k_spmv(Op<R> op, const R* __restrict__ k2, const cplx<R>* __restrict__ gam,
                                              const cplx<R>* __restrict__ x, const cplx<R>* __restrict__ dinv,
                                              cplx<R>* __restrict__ y) {
    constexpr int G = 64 / F;
    const int lane = threadIdx.x & 63;
    const int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (row >= op.n) return;
    const int f = lane % F, g = lane / F;
    const R kk = op.B ? k2[f] : R(0);
    cplx<R> acc = cmk<R>(R(0), R(0));
    const int j1 = op.ptr[row + 1];
    for (int j = op.ptr[row] + g; j < j1; j += G) {
        const int64_t c = op.idx[j];
        cplx<R> a = op.E[j];
        if (op.B) {
            const cplx<R> b = op.B[j];
            a.x -= kk * b.x;
            a.y -= kk * b.y;
        }
        cplx<R> xv = x[c * F + f];
        if (dinv) xv = cmul(xv, dinv[c * F + f]);
        acc = cfma(a, xv, acc);
    }
#pragma unroll
    for (int o = 32; o >= F; o >>= 1) {
        acc.x += __shfl_xor(acc.x, o);
        acc.y += __shfl_xor(acc.y, o);
    }
    if (g == 0) {
        if (op.sptr) {
            for (int j = op.sptr[row]; j < op.sptr[row + 1]; ++j) {
                const int64_t c = op.scol[j];
                cplx<R> xv = x[c * F + f];
                if (dinv) xv = cmul(xv, dinv[c * F + f]);
                const cplx<R> gv = gam[op.sbc[j] * F + f];
                const cplx<R> coef = cmk<R>(gv.x * op.sval[j], gv.y * op.sval[j]);
                acc = cfma(coef, xv, acc);
            }
        }
        y[row * F + f] = acc;
    }
}

// dinv[r][f] = 1 / A_f[r][r]
template <typename R, int F>
__global__ void k_jacobi(Op<R> op, const R* __restrict__ k2, const cplx<R>* __restrict__ gam,
                         cplx<R>* __restrict__ dinv) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= op.n * F) return;
    const int64_t row = q / F;
    const int f = int(q % F);
    double2 d = make_double2(0.0, 0.0);
    const int s = op.diag[row];
    if (s >= 0) {
        d = to_d(op.E[s]);
        if (op.B) {
            const double2 b = to_d(op.B[s]);
            d.x -= double(k2[f]) * b.x;
            d.y -= double(k2[f]) * b.y;
        }
    }
    if (op.sptr) {
        for (int j = op.sptr[row]; j < op.sptr[row + 1]; ++j) {
            if (op.scol[j] != row) continue;
            const double2 gv = to_d(gam[op.sbc[j] * F + f]);
            d.x += gv.x * double(op.sval[j]);
            d.y += gv.y * double(op.sval[j]);
        }
    }
    const double m2 = d.x * d.x + d.y * d.y;
    double2 inv = m2 > 0 ? make_double2(d.x / m2, -d.y / m2) : make_double2(1.0, 0.0);
    dinv[q] = from_d<R>(inv);
}

// Stage 1 of batched inner products: each wavefront covers 512 consecutive
// [row][f] elements and writes partial[wave][i][f] = sum conj(V_i) w.
template <typename R, int F>
__global__ __launch_bounds__(256) void k_dots(int64_t nq, int nv, const cplx<R>* __restrict__ V, int64_t ldv,
                                              const cplx<R>* __restrict__ w, double2* __restrict__ partial) {
    constexpr int P = 8;
    const int64_t wave = (int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t base = wave * 64 * P;
    if (base >= nq) return;
    double2 wv[P];
    bool ok[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int64_t q = base + p * 64 + lane;
        ok[p] = q < nq;
        wv[p] = ok[p] ? to_d(w[q]) : make_double2(0.0, 0.0);
    }
    for (int i = 0; i < nv; ++i) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (!ok[p]) continue;
            const double2 v = to_d(V[i * ldv + base + p * 64 + lane]);
            acc.x += v.x * wv[p].x + v.y * wv[p].y;
            acc.y += v.x * wv[p].y - v.y * wv[p].x;
        }
#pragma unroll
        for (int o = 32; o >= F; o >>= 1) {
            acc.x += __shfl_xor(acc.x, o);
            acc.y += __shfl_xor(acc.y, o);
        }
        if (lane < F) partial[(wave * nv + i) * F + lane] = acc;
    }
}

// Stage 2: out[i][f] (+)= sum over waves in order.
__global__ void k_dots_final(int64_t nwaves, int nv, int F, const double2* __restrict__ partial,
                             double2* __restrict__ out, int accumulate) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * F) return;
    const int i = t / F, f = t % F;
    double2 s = make_double2(0.0, 0.0);
    for (int64_t wv = 0; wv < nwaves; ++wv) {
        const double2 v = partial[(wv * nv + i) * F + f];
        s.x += v.x;
        s.y += v.y;
    }
    if (accumulate) {
        out[t].x += s.x;
        out[t].y += s.y;
    } else {
        out[t] = s;
    }
}

// w[q] -= sum_i V_i[q] h[i][f]
template <typename R, int F>
__global__ void k_axpy_multi(int64_t nq, int nv, const cplx<R>* __restrict__ V, int64_t ldv,
                             const double2* __restrict__ h, cplx<R>* __restrict__ w) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    double2 acc = to_d(w[q]);
    for (int i = 0; i < nv; ++i) {
        const double2 v = to_d(V[i * ldv + q]);
        const double2 c = h[i * F + f];
        acc.x -= v.x * c.x - v.y * c.y;
        acc.y -= v.x * c.y + v.y * c.x;
    }
    w[q] = from_d<R>(acc);
}

// dst[q] = src[q] * inv_norm[f]  (zero when the norm vanishes)
template <typename R, int F>
__global__ void k_scale(int64_t nq, const cplx<R>* __restrict__ src, const double2* __restrict__ nrm2,
                        cplx<R>* __restrict__ dst) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    const double nn = sqrt(fmax(nrm2[f].x, 0.0));
    const double s = nn > 1e-300 ? 1.0 / nn : 0.0;
    const double2 v = to_d(src[q]);
    dst[q] = from_d<R>(make_double2(v.x * s, v.y * s));
}

// One Arnoldi column per system: combine the two CGS passes and the new
// norm, apply previous Givens rotations, create the new one, update g.
__global__ void k_givens(int j, int m, int F, const double2* __restrict__ h, const double2* __restrict__ nrm2,
                         double2* __restrict__ H, double* __restrict__ cs, double2* __restrict__ sn,
                         double2* __restrict__ g, double* __restrict__ res) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double2 col[2];
    // rotate column j in place in H (entries 0..j), keeping the sub-diagonal aside
    for (int i = 0; i <= j; ++i) H[(int64_t(j) * (m + 1) + i) * F + f] = h[i * F + f];
    double2 sub = make_double2(sqrt(fmax(nrm2[f].x, 0.0)), 0.0);
    for (int i = 0; i < j; ++i) {
        col[0] = H[(int64_t(j) * (m + 1) + i) * F + f];
        col[1] = H[(int64_t(j) * (m + 1) + i + 1) * F + f];
        const double c = cs[i * F + f];
        const double2 s = sn[i * F + f];
        const double2 a = make_double2(c * col[0].x + (s.x * col[1].x - s.y * col[1].y),
                                       c * col[0].y + (s.x * col[1].y + s.y * col[1].x));
        const double2 b = make_double2(-(s.x * col[0].x + s.y * col[0].y) + c * col[1].x,
                                       -(s.x * col[0].y - s.y * col[0].x) + c * col[1].y);
        H[(int64_t(j) * (m + 1) + i) * F + f] = a;
        H[(int64_t(j) * (m + 1) + i + 1) * F + f] = b;
    }
    const double2 a = H[(int64_t(j) * (m + 1) + j) * F + f];
    const double2 b = sub;
    const double na = sqrt(a.x * a.x + a.y * a.y), nb = sqrt(b.x * b.x + b.y * b.y);
    double c;
    double2 s, r;
    if (nb == 0.0) {
        c = 1.0;
        s = make_double2(0.0, 0.0);
        r = a;
    } else if (na == 0.0) {
        c = 0.0;
        s = make_double2(1.0, 0.0);  // conj(b)/|b| scaled below
        s = make_double2(b.x / nb, -b.y / nb);
        r = make_double2(nb, 0.0);
    } else {
        const double t = sqrt(na * na + nb * nb);
        const double2 ph = make_double2(a.x / na, a.y / na);
        c = na / t;
        // s = ph * conj(b) / t
        s = make_double2((ph.x * b.x + ph.y * b.y) / t, (ph.y * b.x - ph.x * b.y) / t);
        r = make_double2(ph.x * t, ph.y * t);
    }
    cs[j * F + f] = c;
    sn[j * F + f] = s;
    H[(int64_t(j) * (m + 1) + j) * F + f] = r;
    const double2 gj = g[j * F + f];
    // g_{j+1} = -conj(s) g_j ; g_j = c g_j
    g[(j + 1) * F + f] = make_double2(-(s.x * gj.x + s.y * gj.y), -(s.x * gj.y - s.y * gj.x));
    g[j * F + f] = make_double2(c * gj.x, c * gj.y);
    const double2 gn = g[(j + 1) * F + f];
    res[f] = sqrt(gn.x * gn.x + gn.y * gn.y);
}

// y = H(0:k,0:k)^-1 g(0:k) per system
__global__ void k_backsolve(int k, int m, int F, const double2* __restrict__ H, const double2* __restrict__ g,
                            double2* __restrict__ y) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int i = k - 1; i >= 0; --i) {
        double2 s = g[i * F + f];
        for (int l = i + 1; l < k; ++l) {
            const double2 hv = H[(int64_t(l) * (m + 1) + i) * F + f];
            const double2 yv = y[l * F + f];
            s.x -= hv.x * yv.x - hv.y * yv.y;
            s.y -= hv.x * yv.y + hv.y * yv.x;
        }
        const double2 d = H[(int64_t(i) * (m + 1) + i) * F + f];
        const double dd = d.x * d.x + d.y * d.y;
        y[i * F + f] = dd > 0 ? make_double2((s.x * d.x + s.y * d.y) / dd, (s.y * d.x - s.x * d.y) / dd)
                              : make_double2(0.0, 0.0);
    }
}

// x64[q] += dinv[q] * sum_i V_i[q] y[i][f]
template <typename R, int F>
__global__ void k_update(int64_t nq, int k, const cplx<R>* __restrict__ V, int64_t ldv, const double2* __restrict__ y,
                         const cplx<R>* __restrict__ dinv, double2* __restrict__ x) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < k; ++i) {
        const double2 v = to_d(V[i * ldv + q]);
        const double2 c = y[i * F + f];
        acc.x += v.x * c.x - v.y * c.y;
        acc.y += v.x * c.y + v.y * c.x;
    }
    const double2 d = to_d(dinv[q]);
    x[q].x += d.x * acc.x - d.y * acc.y;
    x[q].y += d.x * acc.y + d.y * acc.x;
}

// r = b - t  (FP64); also writes the FP32/64 copy scaled by 1/||r|| later
__global__ void k_sub(int64_t nq, const double2* __restrict__ b, const double2* __restrict__ t, double2* __restrict__ r) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    r[q] = make_double2(b[q].x - t[q].x, b[q].y - t[q].y);
}

template <typename R>
__global__ void k_cast(int64_t nq, const double2* __restrict__ src, cplx<R>* __restrict__ dst) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    dst[q] = from_d<R>(src[q]);
}

__global__ void k_zero2(int64_t n, double2* __restrict__ p) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) p[q] = make_double2(0.0, 0.0);
}

// b[n][f] = sum_p coef[f][p] * forcing[p][n] for the batch [f0, f0+F)
__global__ void k_build_rhs(int64_t n, int F, int np, const double2* __restrict__ forcing,
                            const double2* __restrict__ coef, double2* __restrict__ b) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n * F) return;
    const int64_t row = q / F;
    const int f = int(q % F);
    double2 s = make_double2(0.0, 0.0);
    for (int p = 0; p < np; ++p) {
        const double2 cv = coef[f * np + p];
        const double2 v = forcing[int64_t(p) * n + row];
        s.x += cv.x * v.x - cv.y * v.y;
        s.y += cv.x * v.y + cv.y * v.x;
    }
    b[q] = s;
}

// out[f][n] = in[n][f]  or the reverse
__global__ void k_transpose(int64_t n, int F, const double2* __restrict__ in, double2* __restrict__ out, int to_fn) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n * F) return;
    if (to_fn) {
        const int64_t row = q / F;
        const int f = int(q % F);
        out[int64_t(f) * n + row] = in[q];
    } else {
        const int f = int(q / n);
        const int64_t row = q % n;
        out[row * F + f] = in[q];
    }
}

struct SolveParams {
    double tol = 1e-8;
    int restart = 100;
    int max_cycles = 100;
    double inner_rel = 1e-5;  // inner (per-cycle) reduction target for FP32
    int check_every = 10;
};

struct BatchStats {
    std::vector<double> relres;
    std::vector<int64_t> iters;
    int cycles = 0;
};

// Scratch of one batch solve.
template <typename R>
struct Work {
    DBuf<cplx<R>> V, w, dinv;
    DBuf<double2> r64, t64, H, sn, g, hbuf, y, partial, nrm;
    DBuf<double> cs, res;
    DBuf<R> k2;
    DBuf<cplx<R>> gam;
    DBuf<double2> gam64;
    DBuf<double> k264;
};

template <typename R, int F>
int run_batch(emg_ctx* c, const Op<R>& op, const Op<double>& op64, const double* k2_h /*F*/,
              const double2* gam_h /*nbc x F*/, int nbc, const double2* b64 /*device [n][F]*/,
              double2* x64 /*device [n][F]*/, const SolveParams& P, BatchStats& st) {
    hipStream_t s = c->stream;
    const int64_t n = op.n, nq = n * F;
    const int m = P.restart;
    Work<R> W;
    EMG_HIP(c, W.V.alloc(int64_t(m + 1) * nq));
    EMG_HIP(c, W.w.alloc(nq));
    EMG_HIP(c, W.dinv.alloc(nq));
    EMG_HIP(c, W.r64.alloc(nq));
    EMG_HIP(c, W.t64.alloc(nq));
    EMG_HIP(c, W.H.alloc(int64_t(m) * (m + 1) * F));
    EMG_HIP(c, W.sn.alloc(int64_t(m) * F));
    EMG_HIP(c, W.cs.alloc(int64_t(m) * F));
    EMG_HIP(c, W.g.alloc(int64_t(m + 1) * F));
    EMG_HIP(c, W.hbuf.alloc(int64_t(m + 1) * F));
    EMG_HIP(c, W.y.alloc(int64_t(m) * F));
    EMG_HIP(c, W.nrm.alloc(F));
    EMG_HIP(c, W.res.alloc(F));
    const int64_t nwaves = (nq + 511) / 512;
    EMG_HIP(c, W.partial.alloc(nwaves * (m + 1) * F));
    EMG_HIP(c, W.k2.alloc(F));
    EMG_HIP(c, W.k264.alloc(F));
    EMG_HIP(c, W.gam.alloc(std::max(1, nbc) * F));
    EMG_HIP(c, W.gam64.alloc(std::max(1, nbc) * F));
    {
        std::vector<R> k2r(F);
        std::vector<cplx<R>> gr(std::max(1, nbc) * F);
        std::vector<double2> g64(std::max(1, nbc) * F, make_double2(0, 0));
        for (int f = 0; f < F; ++f) k2r[f] = R(k2_h[f]);
        for (int b = 0; b < nbc; ++b)
            for (int f = 0; f < F; ++f) {
                gr[b * F + f] = from_d<R>(gam_h[b * F + f]);
                g64[b * F + f] = gam_h[b * F + f];
            }
        if (nbc == 0) gr[0] = cmk<R>(R(0), R(0));
        EMG_HIP(c, hipMemcpyAsync(W.k2.p, k2r.data(), sizeof(R) * F, hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(W.k264.p, k2_h, sizeof(double) * F, hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(W.gam.p, gr.data(), sizeof(cplx<R>) * gr.size(), hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(W.gam64.p, g64.data(), sizeof(double2) * g64.size(), hipMemcpyHostToDevice, s));
    }
    const unsigned gq = grid_for(nq, 256);
    const unsigned grow = grid_for(n, 4);
    const unsigned gdots = grid_for(nwaves * 64, 256);
    k_jacobi<R, F><<<gq, 256, 0, s>>>(op, W.k2.p, W.gam.p, W.dinv.p);
    EMG_HIP(c, hipGetLastError());

    auto dots = [&](const cplx<R>* V, int nv, const cplx<R>* w, double2* out, int acc) {
        k_dots<R, F><<<gdots, 256, 0, s>>>(nq, nv, V, nq, w, W.partial.p);
        k_dots_final<<<grid_for(int64_t(nv) * F, 64), 64, 0, s>>>(nwaves, nv, F, W.partial.p, out, acc);
    };
    auto dots64 = [&](const double2* a, double2* out) {
        k_dots<double, F><<<gdots, 256, 0, s>>>(nq, 1, a, nq, a, W.partial.p);
        k_dots_final<<<grid_for(F, 64), 64, 0, s>>>(nwaves, 1, F, W.partial.p, out, 0);
    };

    // ||b||
    std::vector<double2> hb(F), hr(F);
    DBuf<double2> bn;
    EMG_HIP(c, bn.alloc(F));
    dots64(b64, bn.p);
    EMG_HIP(c, hipMemcpyAsync(hb.data(), bn.p, sizeof(double2) * F, hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    std::vector<double> bnorm(F);
    for (int f = 0; f < F; ++f) bnorm[f] = std::sqrt(std::max(hb[f].x, 0.0));

    st.relres.assign(F, 1.0);
    st.iters.assign(F, 0);
    std::vector<double> hres(F);
    for (int cyc = 0; cyc < P.max_cycles; ++cyc) {
        // true residual in FP64
        k_spmv<double, F><<<grow, 256, 0, s>>>(op64, W.k264.p, W.gam64.p, x64, nullptr, W.t64.p);
        k_sub<<<gq, 256, 0, s>>>(nq, b64, W.t64.p, W.r64.p);
        dots64(W.r64.p, W.nrm.p);
        EMG_HIP(c, hipMemcpyAsync(hr.data(), W.nrm.p, sizeof(double2) * F, hipMemcpyDeviceToHost, s));
        EMG_HIP(c, hipStreamSynchronize(s));
        EMG_HIP(c, hipGetLastError());
        bool all = true;
        std::vector<double> rn(F);
        for (int f = 0; f < F; ++f) {
            rn[f] = std::sqrt(std::max(hr[f].x, 0.0));
            st.relres[f] = bnorm[f] > 0 ? rn[f] / bnorm[f] : 0.0;
            if (!(st.relres[f] <= P.tol)) all = false;
        }
        st.cycles = cyc;
        if (all) break;
        // V0 = r / ||r||, g = (||r||, 0, ...)
        k_cast<R><<<gq, 256, 0, s>>>(nq, W.r64.p, W.w.p);
        k_scale<R, F><<<gq, 256, 0, s>>>(nq, W.w.p, W.nrm.p, W.V.p);
        {
            std::vector<double2> g0(int64_t(m + 1) * F, make_double2(0, 0));
            for (int f = 0; f < F; ++f) g0[f] = make_double2(rn[f], 0.0);
            EMG_HIP(c, hipMemcpyAsync(W.g.p, g0.data(), sizeof(double2) * g0.size(), hipMemcpyHostToDevice, s));
        }
        std::vector<double> target(F);
        for (int f = 0; f < F; ++f) target[f] = std::max(0.5 * P.tol * bnorm[f], P.inner_rel * rn[f]);
        int k = 0;
        for (int j = 0; j < m; ++j) {
            cplx<R>* Vj = W.V.p + int64_t(j) * nq;
            k_spmv<R, F><<<grow, 256, 0, s>>>(op, W.k2.p, W.gam.p, Vj, W.dinv.p, W.w.p);
            dots(W.V.p, j + 1, W.w.p, W.hbuf.p, 0);
            k_axpy_multi<R, F><<<gq, 256, 0, s>>>(nq, j + 1, W.V.p, nq, W.hbuf.p, W.w.p);
            // second pass into y (scratch), accumulated onto hbuf
            dots(W.V.p, j + 1, W.w.p, W.y.p, 0);
            k_axpy_multi<R, F><<<gq, 256, 0, s>>>(nq, j + 1, W.V.p, nq, W.y.p, W.w.p);
            k_dots_final<<<grid_for(int64_t(j + 1) * F, 64), 64, 0, s>>>(nwaves, j + 1, F, W.partial.p, W.hbuf.p, 1);
            dots(W.w.p, 1, W.w.p, W.nrm.p, 0);
            k_scale<R, F><<<gq, 256, 0, s>>>(nq, W.w.p, W.nrm.p, W.V.p + int64_t(j + 1) * nq);
            k_givens<<<1, 64, 0, s>>>(j, m, F, W.hbuf.p, W.nrm.p, W.H.p, W.cs.p, W.sn.p, W.g.p, W.res.p);
            k = j + 1;
            if (k % P.check_every == 0 || k == m) {
                EMG_HIP(c, hipMemcpyAsync(hres.data(), W.res.p, sizeof(double) * F, hipMemcpyDeviceToHost, s));
                EMG_HIP(c, hipStreamSynchronize(s));
                EMG_HIP(c, hipGetLastError());
                bool done = true;
                for (int f = 0; f < F; ++f)
                    if (!(hres[f] <= target[f])) done = false;
                if (done) break;
            }
        }
        for (int f = 0; f < F; ++f) st.iters[f] += k;
        k_backsolve<<<1, 64, 0, s>>>(k, m, F, W.H.p, W.g.p, W.y.p);
        k_update<R, F><<<gq, 256, 0, s>>>(nq, k, W.V.p, nq, W.y.p, W.dinv.p, x64);
        EMG_HIP(c, hipGetLastError());
        st.cycles = cyc + 1;
    }
    return EMG_OK;
}

template <typename R>
int dispatch_batch(emg_ctx* c, int F, const Op<R>& op, const Op<double>& op64, const double* k2, const double2* gam,
                   int nbc, const double2* b64, double2* x64, const SolveParams& P, BatchStats& st) {
    switch (F) {
        case 1: return run_batch<R, 1>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 2: return run_batch<R, 2>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 4: return run_batch<R, 4>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 8: return run_batch<R, 8>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 16: return run_batch<R, 16>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 32: return run_batch<R, 32>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        default: return fail(c, EMG_ERR_ARG, "batch width %d not in {1,2,4,8,16,32}", F);
    }
}

int pow2_block(int64_t remaining, int fblock) {
    int F = 1;
    while (F * 2 <= fblock && F * 2 <= remaining) F *= 2;
    return F;
}

}  // namespace

extern "C" int emg_sweep_solve(emg_ctx* c, int64_t nbatch, const double* k2, const double* gamma,
                               const double* rhs_coef, int precision, int fblock, double tol, int restart,
                               int max_cycles, double* x_out, double* relres_out, int64_t* iters_out) {
    EMG_CHECK_CTX(c);
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: no reduced system");
    if (!c->have_surface) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: surface terms not set");
    if (c->n_forcing < 1) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: no forcing vectors");
    if (nbatch < 1 || !k2 || !rhs_coef || !x_out || (c->n_bc > 0 && !gamma))
        return fail(c, EMG_ERR_ARG, "emg_sweep_solve: bad arguments");
    if (restart < 1 || restart > 1000 || max_cycles < 1 || !(tol > 0))
        return fail(c, EMG_ERR_ARG, "emg_sweep_solve: bad solver parameters");
    if (precision != 0 && precision != 1) return fail(c, EMG_ERR_ARG, "emg_sweep_solve: precision must be 0 or 1");
    fblock = std::max(1, std::min(fblock, 32));
    hipStream_t s = c->stream;
    const int64_t n = c->n_red;
    const int nbc = c->n_bc, np = c->n_forcing;
    Op<double> op64;
    op64.n = n;
    op64.ptr = c->red_ptr.p;
    op64.idx = c->red_idx.p;
    op64.E = reinterpret_cast<const double2*>(c->E64.p);
    op64.B = reinterpret_cast<const double2*>(c->B64.p);
    op64.diag = c->red_diag.p;
    if (c->n_sentries > 0) {
        op64.sptr = c->s_ptr.p;
        op64.scol = c->s_col.p;
        op64.sbc = c->s_bc.p;
        op64.sval = c->s_val64.p;
    }
    Op<float> op32;
    op32.n = n;
    op32.ptr = op64.ptr;
    op32.idx = op64.idx;
    op32.E = reinterpret_cast<const float2*>(c->E32.p);
    op32.B = reinterpret_cast<const float2*>(c->B32.p);
    op32.diag = op64.diag;
    if (c->n_sentries > 0) {
        op32.sptr = op64.sptr;
        op32.scol = op64.scol;
        op32.sbc = op64.sbc;
        op32.sval = c->s_val32.p;
    }
    SolveParams P;
    P.tol = tol;
    P.restart = restart;
    P.max_cycles = max_cycles;
    P.inner_rel = precision == 1 ? 1e-5 : 1e-13;
    DBuf<double2> b64, x64, xt, coef;
    int64_t done = 0;
    while (done < nbatch) {
        const int F = pow2_block(nbatch - done, fblock);
        EMG_HIP(c, b64.alloc(n * F));
        EMG_HIP(c, x64.alloc(n * F));
        EMG_HIP(c, xt.alloc(n * F));
        EMG_HIP(c, coef.alloc(int64_t(F) * np));
        EMG_HIP(c, hipMemcpyAsync(coef.p, rhs_coef + 2 * done * np, sizeof(double2) * F * np, hipMemcpyHostToDevice, s));
        k_build_rhs<<<grid_for(n * F, 256), 256, 0, s>>>(n, F, np, reinterpret_cast<const double2*>(c->forcing.p),
                                                         coef.p, b64.p);
        k_zero2<<<grid_for(n * F, 256), 256, 0, s>>>(n * F, x64.p);
        std::vector<double2> gam(std::max(1, nbc) * F);
        for (int b = 0; b < nbc; ++b)
            for (int f = 0; f < F; ++f)
                gam[b * F + f] = make_double2(gamma[2 * ((done + f) * nbc + b)], gamma[2 * ((done + f) * nbc + b) + 1]);
        BatchStats st;
        const int rc = precision == 1
                           ? dispatch_batch<float>(c, F, op32, op64, k2 + done, gam.data(), nbc, b64.p, x64.p, P, st)
                           : dispatch_batch<double>(c, F, op64, op64, k2 + done, gam.data(), nbc, b64.p, x64.p, P, st);
        if (rc != EMG_OK) return rc;
        k_transpose<<<grid_for(n * F, 256), 256, 0, s>>>(n, F, x64.p, xt.p, 1);
        EMG_HIP(c, hipMemcpyAsync(x_out + 2 * done * n, xt.p, sizeof(double2) * n * F, hipMemcpyDeviceToHost, s));
        EMG_HIP(c, hipStreamSynchronize(s));
        for (int f = 0; f < F; ++f) {
            if (relres_out) relres_out[done + f] = st.relres[f];
            if (iters_out) iters_out[done + f] = st.iters[f];
        }
        done += F;
    }
    return EMG_OK;
}

__global__ void k_find_diag(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                            int32_t* __restrict__ diag) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int d = -1;
    for (int j = ptr[r]; j < ptr[r + 1]; ++j)
        if (idx[j] == r) d = j;
    diag[r] = d;
}

extern "C" int emg_csr_set(emg_ctx* c, int64_t n, const int32_t* indptr, const int32_t* indices, const double* vals) {
    EMG_CHECK_CTX(c);
    if (n < 1 || !indptr || !indices || !vals) return fail(c, EMG_ERR_ARG, "emg_csr_set: null or empty matrix");
    const int64_t nnz = indptr[n];
    if (indptr[0] != 0 || nnz < 1) return fail(c, EMG_ERR_ARG, "emg_csr_set: bad indptr");
    for (int64_t r = 0; r < n; ++r)
        if (indptr[r + 1] < indptr[r]) return fail(c, EMG_ERR_ARG, "emg_csr_set: indptr not monotone");
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || indices[k] >= n) return fail(c, EMG_ERR_ARG, "emg_csr_set: column out of range");
    hipStream_t s = c->stream;
    EMG_HIP(c, c->g_ptr.alloc(n + 1));
    EMG_HIP(c, c->g_idx.alloc(nnz));
    EMG_HIP(c, c->g_diag.alloc(n));
    EMG_HIP(c, c->g_v64.alloc(2 * nnz));
    EMG_HIP(c, c->g_v32.alloc(2 * nnz));
    EMG_HIP(c, hipMemcpyAsync(c->g_ptr.p, indptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_idx.p, indices, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_v64.p, vals, sizeof(double) * 2 * nnz, hipMemcpyHostToDevice, s));
    std::vector<float> v32(2 * nnz);
    for (int64_t k = 0; k < 2 * nnz; ++k) v32[k] = float(vals[k]);
    EMG_HIP(c, hipMemcpyAsync(c->g_v32.p, v32.data(), sizeof(float) * 2 * nnz, hipMemcpyHostToDevice, s));
    k_find_diag<<<grid_for(n, 256), 256, 0, s>>>(n, c->g_ptr.p, c->g_idx.p, c->g_diag.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(s));
    c->g_n = n;
    c->g_nnz = nnz;
    return EMG_OK;
}

extern "C" int emg_csr_solve(emg_ctx* c, int32_t nrhs, const double* b, int precision, double tol, int restart,
                             int max_cycles, double* x, double* relres_out, int64_t* iters_out) {
    EMG_CHECK_CTX(c);
    if (c->g_n == 0) return fail(c, EMG_ERR_STATE, "emg_csr_solve: no matrix (call emg_csr_set)");
    if (nrhs < 1 || !b || !x) return fail(c, EMG_ERR_ARG, "emg_csr_solve: bad arguments");
    if (restart < 1 || restart > 1000 || max_cycles < 1 || !(tol > 0))
        return fail(c, EMG_ERR_ARG, "emg_csr_solve: bad solver parameters");
    if (precision != 0 && precision != 1) return fail(c, EMG_ERR_ARG, "emg_csr_solve: precision must be 0 or 1");
    hipStream_t s = c->stream;
    const int64_t n = c->g_n;
    Op<double> op64;
    op64.n = n;
    op64.ptr = c->g_ptr.p;
    op64.idx = c->g_idx.p;
    op64.E = reinterpret_cast<const double2*>(c->g_v64.p);
    op64.diag = c->g_diag.p;
    Op<float> op32;
    op32.n = n;
    op32.ptr = op64.ptr;
    op32.idx = op64.idx;
    op32.E = reinterpret_cast<const float2*>(c->g_v32.p);
    op32.diag = op64.diag;
    SolveParams P;
    P.tol = tol;
    P.restart = restart;
    P.max_cycles = max_cycles;
    P.inner_rel = precision == 1 ? 1e-5 : 1e-13;
    DBuf<double2> bin, b64, x64, xt;
    int64_t done = 0;
    int rc = EMG_OK;
    bool converged = true;
    while (done < nrhs) {
        const int F = pow2_block(nrhs - done, 32);
        EMG_HIP(c, bin.alloc(n * F));
        EMG_HIP(c, b64.alloc(n * F));
        EMG_HIP(c, x64.alloc(n * F));
        EMG_HIP(c, xt.alloc(n * F));
        EMG_HIP(c, hipMemcpyAsync(bin.p, b + 2 * done * n, sizeof(double2) * n * F, hipMemcpyHostToDevice, s));
        k_transpose<<<grid_for(n * F, 256), 256, 0, s>>>(n, F, bin.p, b64.p, 0);
        k_zero2<<<grid_for(n * F, 256), 256, 0, s>>>(n * F, x64.p);
        std::vector<double> k2(F, 0.0);
        double2 dummy = make_double2(0, 0);
        BatchStats st;
        rc = precision == 1 ? dispatch_batch<float>(c, F, op32, op64, k2.data(), &dummy, 0, b64.p, x64.p, P, st)
                            : dispatch_batch<double>(c, F, op64, op64, k2.data(), &dummy, 0, b64.p, x64.p, P, st);
        if (rc != EMG_OK) return rc;
        k_transpose<<<grid_for(n * F, 256), 256, 0, s>>>(n, F, x64.p, xt.p, 1);
        EMG_HIP(c, hipMemcpyAsync(x + 2 * done * n, xt.p, sizeof(double2) * n * F, hipMemcpyDeviceToHost, s));
        EMG_HIP(c, hipStreamSynchronize(s));
        for (int f = 0; f < F; ++f) {
            if (relres_out) relres_out[done + f] = st.relres[f];
            if (iters_out) iters_out[done + f] = st.iters[f];
            if (!(st.relres[f] <= tol)) converged = false;
        }
        done += F;
    }
    if (!converged) return fail(c, EMG_ERR_CONVERGE, "emg_csr_solve: tolerance %.1e not reached", tol);
    return EMG_OK;
}
