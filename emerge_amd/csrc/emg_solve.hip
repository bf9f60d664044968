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
#include <cstdlib>
#include <cstring>

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
    // complex128 COCG only: the mass part read as complex64 by the Krylov
    // SpMV (half its bytes; the refinement's residual keeps B in FP64)
    const float2* B32 = nullptr;
    const int32_t* sptr = nullptr; // optional surface entries, CSR over rows
    const int32_t* scol = nullptr;
    const int32_t* sbc = nullptr;
    const R* sval = nullptr;
    const int32_t* diag = nullptr; // slot of the diagonal per row (-1 if none)
    // auxiliary space of the COCG preconditioner D^-1 + G W D_g^-1 G^T
    // (aux_n functions; 0 = Jacobi only): G (n x aux_n) and G^T as CSR,
    // D_g = dconst - k2 dvol + sum_s gam[s] dsurf[s] per function, W the
    // per-function weights (dconst 0 and W 1 when not set)
    int64_t aux_n = 0;
    int aux_nsurf = 0;
    const int32_t* gptr = nullptr;
    const int32_t* gidx = nullptr;
    const R* gval = nullptr;
    const int32_t* tptr = nullptr;
    const int32_t* tidx = nullptr;
    const R* tval = nullptr;
    const double2* dvol = nullptr;
    const double* dsurf = nullptr;  // [aux_nsurf][aux_n]
    const double2* dconst = nullptr; // [aux_n] complex, or null
    const double* wt = nullptr;     // [aux_n] or null
    // complex64 COCG only: the stiffness values in FP64 when they are real
    // (real mu_r); the SpMV then accumulates in FP64
    const double* Er64 = nullptr;
    // with Er64: the gradient-space correction also runs in FP64 (G values)
    const double* gval_d = nullptr;
    const double* tval_d = nullptr;
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

// Stage 1 of batched inner products: wavefront (chunk, i) with i =
// blockIdx.y covers 512 consecutive [row][f] elements of vector i and writes
// partial[i][chunk][f] = sum conj(V_i) w.  All (vector, chunk) pairs run in
// parallel.
template <typename R, int F>
__global__ __launch_bounds__(256) void k_dots(int64_t nq, int64_t nchunk, const cplx<R>* __restrict__ V,
                                              int64_t ldv, const cplx<R>* __restrict__ w,
                                              double2* __restrict__ partial) {
    constexpr int P = 8;
    const int64_t chunk = (int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int i = blockIdx.y;
    const int lane = threadIdx.x & 63;
    if (chunk >= nchunk) return;
    const int64_t base = chunk * 64 * P;
    const cplx<R>* Vi = V + i * ldv;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int64_t q = base + p * 64 + lane;
        if (q < nq) {
            const double2 v = to_d(Vi[q]);
            const double2 x = to_d(w[q]);
            acc.x += v.x * x.x + v.y * x.y;
            acc.y += v.x * x.y - v.y * x.x;
        }
    }
#pragma unroll
    for (int o = 32; o >= F; o >>= 1) {
        acc.x += __shfl_xor(acc.x, o);
        acc.y += __shfl_xor(acc.y, o);
    }
    if (lane < F) partial[(int64_t(i) * nchunk + chunk) * F + lane] = acc;
}

// Stage 2: out[i][f] (+)= sum over chunks in order; one wavefront per (i, f)
// group of 64 chunks strided, then a fixed-order tree.
__global__ void k_dots_final(int64_t nchunk, int nv, int F, const double2* __restrict__ partial,
                             double2* __restrict__ out, int accumulate) {
    const int t = blockIdx.x;  // (i, f)
    if (t >= nv * F) return;
    const int i = t / F, f = t % F;
    const int lane = threadIdx.x;  // 64 lanes
    double2 s = make_double2(0.0, 0.0);
    for (int64_t c = lane; c < nchunk; c += 64) {
        const double2 v = partial[(int64_t(i) * nchunk + c) * F + f];
        s.x += v.x;
        s.y += v.y;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s.x += __shfl_xor(s.x, o);
        s.y += __shfl_xor(s.y, o);
    }
    if (lane == 0) {
        if (accumulate) {
            out[t].x += s.x;
            out[t].y += s.y;
        } else {
            out[t] = s;
        }
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
#pragma unroll 8
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
#pragma unroll 8
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
    int restart = 100;        // GMRES: Krylov dimension; IDR: matvecs per refinement cycle
    int max_cycles = 100;
    double inner_rel = 1e-5;  // inner (per-cycle) reduction target for FP32
    int check_every = 10;
    int method = 0;           // 0 = GMRES(m), 1 = IDR(s)
    int idr_s = 4;
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
    const int64_t nchunk = (nq + 511) / 512;
    EMG_HIP(c, W.partial.alloc(nchunk * (m + 1) * F));
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
    const unsigned gdots = grid_for(nchunk * 64, 256);
    k_jacobi<R, F><<<gq, 256, 0, s>>>(op, W.k2.p, W.gam.p, W.dinv.p);
    EMG_HIP(c, hipGetLastError());

    auto dots = [&](const cplx<R>* V, int nv, const cplx<R>* w, double2* out, int acc) {
        k_dots<R, F><<<dim3(gdots, nv), 256, 0, s>>>(nq, nchunk, V, nq, w, W.partial.p);
        k_dots_final<<<nv * F, 64, 0, s>>>(nchunk, nv, F, W.partial.p, out, acc);
    };
    auto dots64 = [&](const double2* a, double2* out) {
        k_dots<double, F><<<dim3(gdots, 1), 256, 0, s>>>(nq, nchunk, a, nq, a, W.partial.p);
        k_dots_final<<<F, 64, 0, s>>>(nchunk, 1, F, W.partial.p, out, 0);
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
            k_dots_final<<<(j + 1) * F, 64, 0, s>>>(nchunk, j + 1, F, W.partial.p, W.hbuf.p, 1);
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

// ------------------------------------------------------------------ IDR(s)
// Induced dimension reduction IDR(s) with bi-orthogonalisation (van Gijzen &
// Sonneveld 2011, "Algorithm 913"), right Jacobi scaling and the
// "maintaining the convergence" choice of omega (threshold 0.7).  Short
// recurrences: per matrix-vector product O(s) vector updates, independent of
// the iteration count.

// w = r - sum_{j<cnt} c[j][f] * G_{k+j}; w = dinv o w  (v of Alg. 913)
template <typename R, int F>
__global__ void k_idr_v(int64_t nq, int k, int cnt, const cplx<R>* __restrict__ G, int64_t ldv,
                        const double2* __restrict__ c, const cplx<R>* __restrict__ r, const cplx<R>* __restrict__ dinv,
                        cplx<R>* __restrict__ v) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    double2 a = to_d(r[q]);
    for (int j = 0; j < cnt; ++j) {
        const double2 g = to_d(G[(k + j) * ldv + q]);
        const double2 cj = c[j * F + f];
        a.x -= cj.x * g.x - cj.y * g.y;
        a.y -= cj.x * g.y + cj.y * g.x;
    }
    const double2 d = to_d(dinv[q]);
    v[q] = from_d<R>(make_double2(d.x * a.x - d.y * a.y, d.x * a.y + d.y * a.x));
}

// U_k = sum_{j<cnt} c[j][f] U_{k+j} + om[f] v   (in place for slot k)
template <typename R, int F>
__global__ void k_idr_u(int64_t nq, int k, int cnt, cplx<R>* __restrict__ U, int64_t ldv,
                        const double2* __restrict__ c, const double2* __restrict__ om, const cplx<R>* __restrict__ v) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    const double2 o = om[f];
    const double2 vv = to_d(v[q]);
    double2 a = make_double2(o.x * vv.x - o.y * vv.y, o.x * vv.y + o.y * vv.x);
    for (int j = 0; j < cnt; ++j) {
        const double2 u = to_d(U[(k + j) * ldv + q]);
        const double2 cj = c[j * F + f];
        a.x += cj.x * u.x - cj.y * u.y;
        a.y += cj.x * u.y + cj.y * u.x;
    }
    U[k * ldv + q] = from_d<R>(a);
}

// G_k -= sum_{i<k} a[i][f] G_i ; U_k -= sum_{i<k} a[i][f] U_i
template <typename R, int F>
__global__ void k_idr_biorth(int64_t nq, int k, cplx<R>* __restrict__ G, cplx<R>* __restrict__ U, int64_t ldv,
                             const double2* __restrict__ a) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    double2 g = to_d(G[k * ldv + q]), u = to_d(U[k * ldv + q]);
    for (int i = 0; i < k; ++i) {
        const double2 ai = a[i * F + f];
        const double2 gi = to_d(G[i * ldv + q]), ui = to_d(U[i * ldv + q]);
        g.x -= ai.x * gi.x - ai.y * gi.y;
        g.y -= ai.x * gi.y + ai.y * gi.x;
        u.x -= ai.x * ui.x - ai.y * ui.y;
        u.y -= ai.x * ui.y + ai.y * ui.x;
    }
    G[k * ldv + q] = from_d<R>(g);
    U[k * ldv + q] = from_d<R>(u);
}

// r -= beta G_k ; x64 += beta U_k   (beta per system)
template <typename R, int F>
__global__ void k_idr_rx(int64_t nq, const cplx<R>* __restrict__ Gk, const cplx<R>* __restrict__ Uk,
                         const double2* __restrict__ beta, cplx<R>* __restrict__ r, double2* __restrict__ x) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    const double2 b = beta[f];
    const double2 g = to_d(Gk[q]), u = to_d(Uk[q]);
    double2 rr = to_d(r[q]);
    rr.x -= b.x * g.x - b.y * g.y;
    rr.y -= b.x * g.y + b.y * g.x;
    r[q] = from_d<R>(rr);
    x[q].x += b.x * u.x - b.y * u.y;
    x[q].y += b.x * u.y + b.y * u.x;
}

// x64 += om (dinv o r) ; r -= om t
template <typename R, int F>
__global__ void k_idr_omega_update(int64_t nq, const double2* __restrict__ om, const cplx<R>* __restrict__ dinv,
                                   const cplx<R>* __restrict__ t, cplx<R>* __restrict__ r, double2* __restrict__ x) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = int(q % F);
    const double2 o = om[f];
    const double2 rv = to_d(r[q]), d = to_d(dinv[q]), tv = to_d(t[q]);
    const double2 v = make_double2(d.x * rv.x - d.y * rv.y, d.x * rv.y + d.y * rv.x);
    x[q].x += o.x * v.x - o.y * v.y;
    x[q].y += o.x * v.y + o.y * v.x;
    r[q] = from_d<R>(make_double2(rv.x - (o.x * tv.x - o.y * tv.y), rv.y - (o.x * tv.y + o.y * tv.x)));
}

__device__ __forceinline__ double2 cdivd(double2 a, double2 b) {
    const double m = b.x * b.x + b.y * b.y;
    if (!(m > 0)) return make_double2(0.0, 0.0);
    return make_double2((a.x * b.x + a.y * b.y) / m, (a.y * b.x - a.x * b.y) / m);
}

// forward substitution: c = M[k:,k:]^-1 fvec[k:]  (M lower triangular,
// M[i][j] at (i*s + j)*F + f)
__global__ void k_idr_solve_c(int k, int s, int F, const double2* __restrict__ M, const double2* __restrict__ fv,
                              double2* __restrict__ c) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int j = 0; j < s - k; ++j) {
        double2 a = fv[(k + j) * F + f];
        for (int i = 0; i < j; ++i) {
            const double2 m = M[((k + j) * s + (k + i)) * F + f];
            const double2 ci = c[i * F + f];
            a.x -= m.x * ci.x - m.y * ci.y;
            a.y -= m.x * ci.y + m.y * ci.x;
        }
        c[j * F + f] = cdivd(a, M[((k + j) * s + (k + j)) * F + f]);
    }
}

// biorthogonalisation coefficients from d_l = p_l^H g_k (l < k):
// a_l = (d_l - sum_{i<l} M[l][i] a_i) / M[l][l]
__global__ void k_idr_alpha(int k, int s, int F, const double2* __restrict__ M, const double2* __restrict__ d,
                            double2* __restrict__ a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int l = 0; l < k; ++l) {
        double2 v = d[l * F + f];
        for (int i = 0; i < l; ++i) {
            const double2 m = M[(l * s + i) * F + f];
            const double2 ai = a[i * F + f];
            v.x -= m.x * ai.x - m.y * ai.y;
            v.y -= m.x * ai.y + m.y * ai.x;
        }
        a[l * F + f] = cdivd(v, M[(l * s + l) * F + f]);
    }
}

// column k of M from d_j = p_j^H g_k (all j); beta = f_k / M_kk;
// f[j] -= beta M[j][k] for j > k
__global__ void k_idr_column(int k, int s, int F, const double2* __restrict__ d, double2* __restrict__ M,
                             double2* __restrict__ fv, double2* __restrict__ beta) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int j = k; j < s; ++j) M[(j * s + k) * F + f] = d[j * F + f];
    const double2 b = cdivd(fv[k * F + f], M[(k * s + k) * F + f]);
    beta[f] = b;
    for (int j = k + 1; j < s; ++j) {
        const double2 m = M[(j * s + k) * F + f];
        fv[j * F + f].x -= b.x * m.x - b.y * m.y;
        fv[j * F + f].y -= b.x * m.y + b.y * m.x;
    }
}

// om = (t^H r) / (t^H t), enlarged to keep |cos| >= 0.7; tt / tr / rr given
// as dot results ("maintaining the convergence", Sleijpen & van der Vorst)
__global__ void k_idr_omega(int F, const double2* __restrict__ tt, const double2* __restrict__ rt,
                            const double2* __restrict__ rr, double2* __restrict__ om) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const double ntt = tt[f].x, nrr = rr[f].x;
    const double2 tr = make_double2(rt[f].x, -rt[f].y);  // t^H r = conj(r^H t)
    double2 o = ntt > 0 ? make_double2(tr.x / ntt, tr.y / ntt) : make_double2(0.0, 0.0);
    const double denom = sqrt(fmax(ntt, 0.0)) * sqrt(fmax(nrr, 0.0));
    const double rho = denom > 0 ? sqrt(tr.x * tr.x + tr.y * tr.y) / denom : 1.0;
    if (rho < 0.7 && rho > 0) {
        o.x *= 0.7 / rho;
        o.y *= 0.7 / rho;
    }
    om[f] = o;
}

__global__ void k_set_c(int64_t n, double2* __restrict__ p, double2 v) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

template <typename R, int F>
int run_idr(emg_ctx* c, const Op<R>& op, const Op<double>& op64, const double* k2_h, const double2* gam_h, int nbc,
            const double2* b64, double2* x64, const SolveParams& P, BatchStats& st) {
    hipStream_t s = c->stream;
    const int64_t n = op.n, nq = n * F;
    const int S = P.idr_s;
    DBuf<cplx<R>> Pm, G, U, r, v, t, dinv;
    DBuf<double2> r64, t64, Mm, fv, cvec, avec, beta, om, dbuf, tt, rt, rr, partial, bn;
    EMG_HIP(c, Pm.alloc(int64_t(S) * nq));
    EMG_HIP(c, G.alloc(int64_t(S) * nq));
    EMG_HIP(c, U.alloc(int64_t(S) * nq));
    EMG_HIP(c, r.alloc(nq));
    EMG_HIP(c, v.alloc(nq));
    EMG_HIP(c, t.alloc(2 * nq));  // t followed by a copy slot for r (dots of t against [t, r])
    EMG_HIP(c, dinv.alloc(nq));
    EMG_HIP(c, r64.alloc(nq));
    EMG_HIP(c, t64.alloc(nq));
    EMG_HIP(c, Mm.alloc(int64_t(S) * S * F));
    EMG_HIP(c, fv.alloc(int64_t(S) * F));
    EMG_HIP(c, cvec.alloc(int64_t(S) * F));
    EMG_HIP(c, avec.alloc(int64_t(S) * F));
    EMG_HIP(c, beta.alloc(F));
    EMG_HIP(c, om.alloc(F));
    EMG_HIP(c, dbuf.alloc(int64_t(S + 2) * F));
    EMG_HIP(c, tt.alloc(2 * F));
    EMG_HIP(c, rt.alloc(F));
    EMG_HIP(c, rr.alloc(F));
    EMG_HIP(c, bn.alloc(F));
    const int64_t nchunk = (nq + 511) / 512;
    EMG_HIP(c, partial.alloc(nchunk * (S + 2) * F));
    DBuf<R> k2d;
    DBuf<double> k264;
    DBuf<cplx<R>> gam;
    DBuf<double2> gam64;
    EMG_HIP(c, k2d.alloc(F));
    EMG_HIP(c, k264.alloc(F));
    EMG_HIP(c, gam.alloc(std::max(1, nbc) * F));
    EMG_HIP(c, gam64.alloc(std::max(1, nbc) * F));
    {
        std::vector<R> k2r(F);
        std::vector<cplx<R>> gr(std::max(1, nbc) * F, cmk<R>(R(0), R(0)));
        std::vector<double2> g64(std::max(1, nbc) * F, make_double2(0, 0));
        for (int f = 0; f < F; ++f) k2r[f] = R(k2_h[f]);
        for (int b = 0; b < nbc; ++b)
            for (int f = 0; f < F; ++f) {
                gr[b * F + f] = from_d<R>(gam_h[b * F + f]);
                g64[b * F + f] = gam_h[b * F + f];
            }
        EMG_HIP(c, hipMemcpyAsync(k2d.p, k2r.data(), sizeof(R) * F, hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(k264.p, k2_h, sizeof(double) * F, hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(gam.p, gr.data(), sizeof(cplx<R>) * gr.size(), hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(gam64.p, g64.data(), sizeof(double2) * g64.size(), hipMemcpyHostToDevice, s));
        // shadow space: fixed pseudo-random complex vectors, same for every system
        std::vector<cplx<R>> ph(size_t(S) * nq);
        uint64_t z = 0x9E3779B97F4A7C15ull;
        auto next = [&]() {
            z ^= z << 13;
            z ^= z >> 7;
            z ^= z << 17;
            return double(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        };
        for (int i = 0; i < S; ++i)
            for (int64_t row = 0; row < n; ++row) {
                const double a = next(), b = next();
                for (int f = 0; f < F; ++f) ph[size_t(i) * nq + row * F + f] = cmk<R>(R(a), R(b));
            }
        EMG_HIP(c, hipMemcpyAsync(Pm.p, ph.data(), sizeof(cplx<R>) * ph.size(), hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipStreamSynchronize(s));
    }
    const unsigned gq = grid_for(nq, 256);
    const unsigned grow = grid_for(n, 4);
    const unsigned gdots = grid_for(nchunk * 64, 256);
    k_jacobi<R, F><<<gq, 256, 0, s>>>(op, k2d.p, gam.p, dinv.p);
    EMG_HIP(c, hipGetLastError());
    auto dots = [&](const cplx<R>* V, int nv, const cplx<R>* w, double2* out) {
        k_dots<R, F><<<dim3(gdots, nv), 256, 0, s>>>(nq, nchunk, V, nq, w, partial.p);
        k_dots_final<<<nv * F, 64, 0, s>>>(nchunk, nv, F, partial.p, out, 0);
    };
    auto dots64 = [&](const double2* a, double2* out) {
        k_dots<double, F><<<dim3(gdots, 1), 256, 0, s>>>(nq, nchunk, a, nq, a, partial.p);
        k_dots_final<<<F, 64, 0, s>>>(nchunk, 1, F, partial.p, out, 0);
    };
    std::vector<double2> hb(F), hr(F);
    dots64(b64, bn.p);
    EMG_HIP(c, hipMemcpyAsync(hb.data(), bn.p, sizeof(double2) * F, hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    std::vector<double> bnorm(F);
    for (int f = 0; f < F; ++f) bnorm[f] = std::sqrt(std::max(hb[f].x, 0.0));
    st.relres.assign(F, 1.0);
    st.iters.assign(F, 0);
    const int64_t max_mv = int64_t(P.restart) * P.max_cycles;  // total matvec budget
    int64_t mv = 0;
    for (int cyc = 0; cyc < P.max_cycles && mv < max_mv; ++cyc) {
        k_spmv<double, F><<<grow, 256, 0, s>>>(op64, k264.p, gam64.p, x64, nullptr, t64.p);
        k_sub<<<gq, 256, 0, s>>>(nq, b64, t64.p, r64.p);
        dots64(r64.p, rr.p);
        EMG_HIP(c, hipMemcpyAsync(hr.data(), rr.p, sizeof(double2) * F, hipMemcpyDeviceToHost, s));
        EMG_HIP(c, hipStreamSynchronize(s));
        EMG_HIP(c, hipGetLastError());
        bool all = true;
        std::vector<double> target(F);
        for (int f = 0; f < F; ++f) {
            const double rn = std::sqrt(std::max(hr[f].x, 0.0));
            st.relres[f] = bnorm[f] > 0 ? rn / bnorm[f] : 0.0;
            if (!(st.relres[f] <= P.tol)) all = false;
            target[f] = std::max(0.5 * P.tol * bnorm[f], P.inner_rel * rn);
        }
        st.cycles = cyc;
        if (all) break;
        // inner IDR(s) on A d = r64, accumulating d into x64
        k_cast<R><<<gq, 256, 0, s>>>(nq, r64.p, r.p);
        EMG_HIP(c, hipMemsetAsync(G.p, 0, sizeof(cplx<R>) * S * nq, s));
        EMG_HIP(c, hipMemsetAsync(U.p, 0, sizeof(cplx<R>) * S * nq, s));
        {
            std::vector<double2> eye(size_t(S) * S * F, make_double2(0, 0));
            for (int i = 0; i < S; ++i)
                for (int f = 0; f < F; ++f) eye[(size_t(i) * S + i) * F + f] = make_double2(1, 0);
            EMG_HIP(c, hipMemcpyAsync(Mm.p, eye.data(), sizeof(double2) * eye.size(), hipMemcpyHostToDevice, s));
        }
        k_set_c<<<1, 64, 0, s>>>(F, om.p, make_double2(1.0, 0.0));
        int64_t inner = 0;
        bool done = false;
        while (!done && inner < P.restart && mv < max_mv) {
            dots(Pm.p, S, r.p, fv.p);  // f = P^H r
            for (int k = 0; k < S; ++k) {
                k_idr_solve_c<<<1, 64, 0, s>>>(k, S, F, Mm.p, fv.p, cvec.p);
                k_idr_v<R, F><<<gq, 256, 0, s>>>(nq, k, S - k, G.p, nq, cvec.p, r.p, dinv.p, v.p);
                k_idr_u<R, F><<<gq, 256, 0, s>>>(nq, k, S - k, U.p, nq, cvec.p, om.p, v.p);
                k_spmv<R, F><<<grow, 256, 0, s>>>(op, k2d.p, gam.p, U.p + int64_t(k) * nq, nullptr,
                                                   G.p + int64_t(k) * nq);
                if (k > 0) {
                    dots(Pm.p, k, G.p + int64_t(k) * nq, dbuf.p);
                    k_idr_alpha<<<1, 64, 0, s>>>(k, S, F, Mm.p, dbuf.p, avec.p);
                    k_idr_biorth<R, F><<<gq, 256, 0, s>>>(nq, k, G.p, U.p, nq, avec.p);
                }
                dots(Pm.p, S, G.p + int64_t(k) * nq, dbuf.p);
                k_idr_column<<<1, 64, 0, s>>>(k, S, F, dbuf.p, Mm.p, fv.p, beta.p);
                k_idr_rx<R, F><<<gq, 256, 0, s>>>(nq, G.p + int64_t(k) * nq, U.p + int64_t(k) * nq, beta.p, r.p, x64);
                ++inner;
                ++mv;
            }
            // omega step: t = A (dinv o r)
            k_spmv<R, F><<<grow, 256, 0, s>>>(op, k2d.p, gam.p, r.p, dinv.p, t.p);
            EMG_HIP(c, hipMemcpyAsync(t.p + nq, r.p, sizeof(cplx<R>) * nq, hipMemcpyDeviceToDevice, s));
            dots(t.p, 2, t.p, tt.p);      // [t^H t, r^H t]
            dots(r.p, 1, r.p, rr.p);      // r^H r
            k_idr_omega<<<1, 64, 0, s>>>(F, tt.p, tt.p + F, rr.p, om.p);
            k_idr_omega_update<R, F><<<gq, 256, 0, s>>>(nq, om.p, dinv.p, t.p, r.p, x64);
            ++inner;
            ++mv;
            // recursive residual check
            dots(r.p, 1, r.p, rr.p);
            EMG_HIP(c, hipMemcpyAsync(hr.data(), rr.p, sizeof(double2) * F, hipMemcpyDeviceToHost, s));
            EMG_HIP(c, hipStreamSynchronize(s));
            EMG_HIP(c, hipGetLastError());
            done = true;
            for (int f = 0; f < F; ++f)
                if (!(std::sqrt(std::max(hr[f].x, 0.0)) <= target[f])) done = false;
        }
        for (int f = 0; f < F; ++f) st.iters[f] += inner;
        st.cycles = cyc + 1;
    }
    return EMG_OK;
}


// -------------------------------------------------------------------- COCG
// Conjugate orthogonal conjugate gradients (van der Vorst & Melissen, IEEE
// Trans. Magn. 26(2), 1990) with Jacobi preconditioning, on the complex-
// symmetric part A_s = (A + A^T)/2 of the system, restarted inside FP64
// iterative refinement on the true A.  The reference's mass matrix is not
// symmetric (its face-face block uses the monomial of ASM/curlcurl.py:361),
// so COCG alone would stall; as an inner solver for the refinement
// r = b - A x, x += A_s^{-1} r it converges, each cycle gaining the factor
// ||A_s^{-1}(A - A_s)||.  Per iteration: one SpMV (the direction
// p = z + beta p formed on the fly) writing partial sums of p^T q, one
// single-workgroup kernel that adds them in a fixed order and sets alpha,
// one update pass writing partial sums of r^T z and ||r||^2, and one that
// sets beta and the stop flags.  Scalars stay on the device, blocks of
// iterations are replayed as one HIP graph, and the host polls the
// per-system stop flags between blocks.

constexpr int kCgThreads = 1024;    // threads per workgroup of the COCG kernels
constexpr int kCgMaxBlocks = 1024;  // grid cap of the SpMV (partial sums per kernel)
constexpr int kCgUpdBlocks = 256;   // grid cap of the update pass

struct CgState {
    double2* rho;    // [S F] r^T z
    double2* beta;   // [S F]
    double2* alpha;  // [S F]
    double* rr;      // [S F] ||r||^2 of the (scaled) inner residual
    double* tgt2;    // [S F] inner stop: rr <= tgt2
    int* done;       // [S F] inner solve finished (converged or broke down)
    double2* part;   // [kCgMaxBlocks][2F] partial sums
    // row segments with Krylov scalars of their own (S = nseg; 1: one set
    // per system).  Segment s holds rows [srow[s], srow[s+1]); the streaming
    // kernels give it workgroups [sblk[s], sblk[s+1]), so that every
    // workgroup's partial sums belong to one segment.
    int nseg = 1;
    const int* sblk = nullptr;      // [S + 1]
    const int64_t* srow = nullptr;  // [S + 1]
    const int* rseg = nullptr;      // [n] segment of each row (nullptr when S = 1)
    // [S] every system of the segment has stopped (set by k_cg_beta, cleared
    // by the first pass of each inner solve): the SpMV, the update and the
    // correction skip its rows until the next inner solve
    int* sdone = nullptr;
};

// The segment's inner solves have all stopped: its rows are skipped.  A
// stopped system has alpha = 0, so skipping leaves x, r and z as they are.
__device__ __forceinline__ bool seg_skip(const CgState& st, int s) {
    return st.nseg > 1 && st.sdone[s] != 0;
}

// Scalar slot (segment, system) of element k = row F + f.
__device__ __forceinline__ int seg_slot(const int* rseg, int64_t k, int F) {
    return (rseg ? rseg[k / F] * F : 0) + int(k % F);
}

// The [row][f] elements a workgroup of a streaming COCG kernel walks: with
// one segment the whole vector, grid-strided; otherwise the elements of its
// own segment, strided over that segment's workgroups.  k0 and the stride
// are multiples of kCgThreads, so thread t always holds system t % F.
struct SegRange {
    int64_t k0, k1, stride;
    int s;
};
__device__ __forceinline__ SegRange seg_range(const CgState& st, int64_t nq, int F) {
    SegRange r;
    if (st.nseg <= 1) {
        r.k0 = int64_t(blockIdx.x) * kCgThreads;
        r.k1 = nq;
        r.stride = int64_t(gridDim.x) * kCgThreads;
        r.s = 0;
        return r;
    }
    int s = 0;
    while (s + 1 < st.nseg && int(blockIdx.x) >= st.sblk[s + 1]) ++s;
    const int lb = int(blockIdx.x) - st.sblk[s], nb = st.sblk[s + 1] - st.sblk[s];
    r.k0 = st.srow[s] * F + int64_t(lb) * kCgThreads;
    r.k1 = st.srow[s + 1] * F;
    r.stride = int64_t(nb) * kCgThreads;
    r.s = s;
    return r;
}

// slot of (col, row) for every slot of a structurally symmetric CSR pattern
__global__ void k_tmap(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                       int32_t* __restrict__ tmap, int32_t* __restrict__ err) {
    const int64_t row = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= n) return;
    for (int j = ptr[row]; j < ptr[row + 1]; ++j) {
        const int c = idx[j];
        int lo = ptr[c], hi = ptr[c + 1] - 1, found = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int v = idx[mid];
            if (v < row) lo = mid + 1;
            else if (v > row) hi = mid - 1;
            else { found = mid; break; }
        }
        if (found < 0) {
            atomicExch(err, 1);
            found = j;
        }
        tmap[j] = found;
    }
}

// s = (v + v^T) / 2, FP64 and FP32 copies (exactly symmetric: the sum is
// formed from the same two operands for both slots)
__global__ void k_symmetrize(int64_t nnz, const int32_t* __restrict__ tmap, const double2* __restrict__ v,
                             double2* __restrict__ s64, float2* __restrict__ s32) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const double2 a = v[k], b = v[tmap[k]];
    const double2 m = make_double2((a.x + b.x) * 0.5, (a.y + b.y) * 0.5);
    s64[k] = m;
    s32[k] = make_float2(float(m.x), float(m.y));
}

int build_symmetric(emg_ctx* c, int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* idx, const double* v,
                    DBuf<double>& s64, DBuf<float>& s32) {
    hipStream_t s = c->stream;
    DBuf<int32_t> tmap, err;
    EMG_HIP(c, tmap.alloc(std::max<int64_t>(nnz, 1)));
    EMG_HIP(c, err.alloc(1));
    EMG_HIP(c, s64.alloc(2 * std::max<int64_t>(nnz, 1)));
    EMG_HIP(c, s32.alloc(2 * std::max<int64_t>(nnz, 1)));
    EMG_HIP(c, hipMemsetAsync(err.p, 0, sizeof(int32_t), s));
    k_tmap<<<grid_for(n, 256), 256, 0, s>>>(n, ptr, idx, tmap.p, err.p);
    k_symmetrize<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, tmap.p, reinterpret_cast<const double2*>(v),
                                                    reinterpret_cast<double2*>(s64.p),
                                                    reinterpret_cast<float2*>(s32.p));
    EMG_HIP(c, hipGetLastError());
    int32_t herr = 0;
    EMG_HIP(c, hipMemcpyAsync(&herr, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    if (herr) return fail(c, EMG_ERR_ARG, "COCG needs a structurally symmetric pattern");
    return EMG_OK;
}

// Fixed-order sum of the nblk x nv partials (layout [block][v], nv a power
// of two <= 64) by one workgroup; the sum of v is left in red[v].
__device__ void cg_sum_partials(const double2* part, int nblk, int nv, double2* red) {
    const int tid = threadIdx.x;
    const int parts = kCgThreads / nv;
    const int v = tid % nv, pi = tid / nv;
    double2 s = make_double2(0.0, 0.0);
#pragma unroll 8
    for (int b = pi; b < nblk; b += parts) {
        const double2 t = part[int64_t(b) * nv + v];
        s.x += t.x;
        s.y += t.y;
    }
    red[tid] = s;
    __syncthreads();
    for (int h = parts / 2; h >= 1; h >>= 1) {
        if (pi < h) {
            red[tid].x += red[tid + h * nv].x;
            red[tid].y += red[tid + h * nv].y;
        }
        __syncthreads();
    }
}

// Two consecutive complex values as one 16-byte load.
__device__ __forceinline__ void load_pair(const float2* p, float2& a, float2& b) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    a = make_float2(v.x, v.y);
    b = make_float2(v.z, v.w);
}
__device__ __forceinline__ void load_pair(const double2* p, double2& a, double2& b) {
    a = p[0];
    b = p[1];
}

// (ar + i ai)[v] += (E[j] - k2[v] B[j]) z[idx[j]][f0 + v] for the V systems
// of a lane (real and imaginary parts in separate arrays: they stay in
// registers for complex128 too)
template <typename R, int F, int V, typename A>
__device__ __forceinline__ void cg_entry(const Op<R>& op, const cplx<R>* __restrict__ z, const R (&kk)[V], int f0,
                                         int j, A (&ar)[V], A (&ai)[V]) {
    const int64_t c = op.idx[j];
    // matrix values are loaded before the gathers so that all are in flight
    cplx<R> e = cmk<R>(R(0), R(0));
    double er = 0.0;
    if constexpr (sizeof(A) > sizeof(R)) er = op.Er64[j];
    else if (sizeof(R) == sizeof(double) && op.Er64) e = cmk<R>(R(op.Er64[j]), R(0));
    else e = op.E[j];
    cplx<R> b = cmk<R>(R(0), R(0));
    if (sizeof(R) == sizeof(double) && op.B32) {
        const float2 t = op.B32[j];
        b = cmk<R>(R(t.x), R(t.y));
    } else if (op.B) {
        b = op.B[j];
    }
    const cplx<R>* zc = z + c * F + f0;
    cplx<R> zv[V];
    if constexpr (V % 2 == 0) {
#pragma unroll
        for (int v = 0; v < V; v += 2) load_pair(zc + v, zv[v], zv[v + 1]);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) zv[v] = zc[v];
    }
    if constexpr (sizeof(A) > sizeof(R)) {
        // real FP64 stiffness, FP64 accumulation: E annihilates gradients, and
        // rounding E to FP32 would leave eps32 |E| |z| in rows that cancel
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double mr = er - double(kk[v]) * double(b.x), mi = -double(kk[v]) * double(b.y);
            ar[v] += mr * double(zv[v].x) - mi * double(zv[v].y);
            ai[v] += mr * double(zv[v].y) + mi * double(zv[v].x);
        }
    } else {
        (void)er;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const R mr = e.x - kk[v] * b.x, mi = e.y - kk[v] * b.y;
            ar[v] += mr * zv[v].x - mi * zv[v].y;
            ai[v] += mr * zv[v].y + mi * zv[v].x;
        }
    }
}

// q = A_s p and per-workgroup partial sums of p^T q.  q is formed from p
// every iteration (not by the recurrence q = A z + beta q, whose drift from
// A p breaks complex64 COCG on strongly indefinite systems).
// A lane covers V consecutive systems of the batch (V/2 16-byte gathers per
// matrix entry) and every GR-th entry of its row, two entries per step so
// that their gathers overlap.  A row gets LPR = FG * GR lanes with GR = 16
// entries in flight (8 at F >= 32; rows hold ~40 entries), so small batches
// put several rows in one wavefront instead of leaving lanes idle.  The butterfly
// reduction stays inside the row's lanes and leaves the row sum in all GR
// lanes of its systems, so the row end (surface terms, q store, p^T q) is
// spread one system per lane, and those lanes load their p before the
// gathers.  Control flow is uniform: every wavefront walks the same number
// of row groups and lanes past the last row only add zeros.
// Lane layout of k_cg_spmv: V systems per lane (8 complex64 or 4 complex128
// per lane, 64 bytes; 4 with FP64 accumulators), FG = F / V lanes per entry,
// GR entries of a row in flight (GRX, or by default 8 at F >= 32 and 16
// below), LPR = FG GR lanes per row and 64 / LPR rows per wavefront.
template <typename R, int F, typename A, int GRX>
struct SpmvShape {
    static constexpr int VM = sizeof(A) == 4 ? 8 : 4;
    static constexpr int V = F >= VM ? VM : F;
    static constexpr int FG = F / V;
    static constexpr int GW = 64 / FG;
    static constexpr int GR = GRX > 0 ? GRX : (GW >= 16 ? (F >= 32 ? 8 : 16) : GW);
    // a row's system owners need GR >= V lanes per entry group
    static constexpr bool valid = GR >= V && GR <= GW && (GR & (GR - 1)) == 0;
};

template <typename R, int F, typename A = R, int GRX = 0>
__global__ __launch_bounds__(kCgThreads) void k_cg_spmv(Op<R> op, const R* __restrict__ k2,
                                                        const cplx<R>* __restrict__ gam,
                                                        const cplx<R>* __restrict__ p, cplx<R>* __restrict__ q,
                                                        CgState st) {
    using Sh = SpmvShape<R, F, A, GRX>;
    static_assert(Sh::valid, "SpMV lane layout");
    constexpr int V = Sh::V;                    // systems per lane
    constexpr int FG = Sh::FG;                  // lanes per entry
    constexpr int GR = Sh::GR;                  // entries in flight per row (GR >= V)
    constexpr int LPR = FG * GR;                // lanes per row
    constexpr int RPW = 64 / LPR;               // rows per wavefront
    constexpr int W = kCgThreads / 64;
    __shared__ double2 wsum[W][F];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rs = lane / LPR, l = lane % LPR;
    const int fg = l % FG, g = l / FG, f0 = fg * V;
    const bool own = g < V;                     // row end of system f0 + g
    const int fo = f0 + (own ? g : 0);
    R kk[V];
#pragma unroll
    for (int v = 0; v < V; ++v) kk[v] = op.B ? k2[f0 + v] : R(0);
    double2 dot = make_double2(0.0, 0.0);
    // XCD-aware rows: workgroups are dealt round-robin over the 8 XCDs;
    // XCD x walks rows [x n/8, (x+1) n/8) so that its L2 holds the part of
    // z those rows gather (the grid is a multiple of 8)
    const int64_t per = (op.n + 7) / 8;
    const int64_t xcd = blockIdx.x % 8;
    const int64_t r0 = xcd * per, r1 = min(op.n, r0 + per);
    const int64_t nw = int64_t(gridDim.x / 8) * W;
    for (int64_t base = r0 + (int64_t(blockIdx.x / 8) * W + wv) * RPW; base < r1; base += nw * RPW) {
        const int64_t row = base + rs;
        const bool valid = row < r1 && !(st.nseg > 1 && st.sdone[st.rseg[row]]);
        const int64_t o = row * F + fo;
        R px = R(0), py = R(0);
        if (own && valid) {
            const cplx<R> pv = p[o];
            px = pv.x;
            py = pv.y;
        }
        A ar[V], ai[V], br[V], bi[V];
#pragma unroll
        for (int v = 0; v < V; ++v) ar[v] = ai[v] = br[v] = bi[v] = A(0);
        const int j1 = valid ? op.ptr[row + 1] : 0;
        int j = valid ? op.ptr[row] + g : 0;
        for (; j + GR < j1; j += 2 * GR) {
            cg_entry<R, F, V, A>(op, p, kk, f0, j, ar, ai);
            cg_entry<R, F, V, A>(op, p, kk, f0, j + GR, br, bi);
        }
        if (j < j1) cg_entry<R, F, V, A>(op, p, kk, f0, j, ar, ai);
        // the pick is done on scalars: a select of vector-type structs is
        // lowered to an indexed stack load for complex128
        A ax = A(0), ay = A(0);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            A tx = ar[v] + br[v], ty = ai[v] + bi[v];
#pragma unroll
            for (int off = LPR / 2; off >= FG; off >>= 1) {
                tx += __shfl_xor(tx, off);
                ty += __shfl_xor(ty, off);
            }
            ax = (v == g) ? tx : ax;
            ay = (v == g) ? ty : ay;
        }
        cplx<R> a = cmk<R>(R(ax), R(ay));
        if (own && valid) {
            if (op.sptr) {
                for (int k = op.sptr[row]; k < op.sptr[row + 1]; ++k) {
                    const cplx<R> gv = gam[op.sbc[k] * F + fo];
                    const R sv = op.sval[k];
                    a = cfma(cmk<R>(gv.x * sv, gv.y * sv), p[int64_t(op.scol[k]) * F + fo], a);
                }
            }
            q[o] = a;
            const double2 t = to_d(cmul(cmk<R>(px, py), a));
            dot.x += t.x;
            dot.y += t.y;
        }
    }
    // the row slots of a wavefront hold partial sums of the same systems
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1) {
        dot.x += __shfl_xor(dot.x, off);
        dot.y += __shfl_xor(dot.y, off);
    }
    if (own && rs == 0) wsum[wv][fo] = dot;
    __syncthreads();
    if (threadIdx.x < F) {
        double2 sum = make_double2(0.0, 0.0);
        for (int w = 0; w < W; ++w) {
            sum.x += wsum[w][threadIdx.x].x;
            sum.y += wsum[w][threadIdx.x].y;
        }
        st.part[int64_t(blockIdx.x) * F + threadIdx.x] = sum;
    }
}

// k_cg_spmv with GR entries in flight per row (emg_solver_set_spmv_lanes;
// 0 or a layout the batch width does not allow: the default)
template <typename R, int F, typename A>
void launch_cg_spmv(int gr, unsigned grid, hipStream_t s, const Op<R>& op, const R* k2, const cplx<R>* gam,
                    const cplx<R>* p, cplx<R>* q, const CgState& cs) {
    switch (gr) {
        case 4:
            if constexpr (SpmvShape<R, F, A, 4>::valid) {
                k_cg_spmv<R, F, A, 4><<<grid, kCgThreads, 0, s>>>(op, k2, gam, p, q, cs);
                return;
            }
            break;
        case 8:
            if constexpr (SpmvShape<R, F, A, 8>::valid) {
                k_cg_spmv<R, F, A, 8><<<grid, kCgThreads, 0, s>>>(op, k2, gam, p, q, cs);
                return;
            }
            break;
        case 16:
            if constexpr (SpmvShape<R, F, A, 16>::valid) {
                k_cg_spmv<R, F, A, 16><<<grid, kCgThreads, 0, s>>>(op, k2, gam, p, q, cs);
                return;
            }
            break;
        case 32:
            if constexpr (SpmvShape<R, F, A, 32>::valid) {
                k_cg_spmv<R, F, A, 32><<<grid, kCgThreads, 0, s>>>(op, k2, gam, p, q, cs);
                return;
            }
            break;
        default:
            break;
    }
    k_cg_spmv<R, F, A, 0><<<grid, kCgThreads, 0, s>>>(op, k2, gam, p, q, cs);
}

// alpha = rho / p^T q (one workgroup per segment; partials of k_cg_spmv,
// or of k_seg_dot with several segments)
template <int F>
__global__ __launch_bounds__(kCgThreads) void k_cg_alpha(int nblk, CgState st) {
    __shared__ double2 red[kCgThreads];
    const int sg = blockIdx.x;
    const int b0 = st.nseg > 1 ? st.sblk[sg] : 0;
    const int nb = st.nseg > 1 ? st.sblk[sg + 1] - b0 : nblk;
    cg_sum_partials(st.part + int64_t(b0) * F, nb, F, red);
    if (threadIdx.x < F) {
        const int ff = sg * F + threadIdx.x;
        const double2 pq = red[threadIdx.x];
        double2 al = make_double2(0.0, 0.0);
        if (!st.done[ff]) {
            const double m2 = pq.x * pq.x + pq.y * pq.y;
            if (m2 > 0 && isfinite(m2)) al = cdivd(st.rho[ff], pq);
            else st.done[ff] = 1;  // breakdown: the refinement restarts this system
        }
        st.alpha[ff] = al;
    }
}

// Workgroup sums of the [row][f] products a_k b_k (conj(a_k) b_k with
// CONJ) over the elements seg_range gives it: part[blockIdx][F].
template <typename R, int F, bool CONJ>
__global__ __launch_bounds__(kCgThreads) void k_seg_dot(int64_t nq, const cplx<R>* __restrict__ a,
                                                        const cplx<R>* __restrict__ b, CgState st,
                                                        double2* __restrict__ part, int skip) {
    __shared__ double2 red[kCgThreads];
    const int tid = threadIdx.x;
    const SegRange sr = seg_range(st, nq, F);
    double2 acc = make_double2(0.0, 0.0);
    const int64_t k1 = skip && seg_skip(st, sr.s) ? sr.k0 : sr.k1;  // stopped: zero sums
    for (int64_t k = sr.k0 + tid; k < k1; k += sr.stride) {
        const double2 u = to_d(a[k]), v = to_d(b[k]);
        if constexpr (CONJ) {
            acc.x += u.x * v.x + u.y * v.y;
            acc.y += u.x * v.y - u.y * v.x;
        } else {
            acc.x += u.x * v.x - u.y * v.y;
            acc.y += u.x * v.y + u.y * v.x;
        }
    }
    const int parts = kCgThreads / F, pi = tid / F;
    red[tid] = acc;
    __syncthreads();
    for (int h = parts / 2; h >= 1; h >>= 1) {
        if (pi < h) {
            red[tid].x += red[tid + h * F].x;
            red[tid].y += red[tid + h * F].y;
        }
        __syncthreads();
    }
    if (tid < F) part[int64_t(blockIdx.x) * F + tid] = red[tid];
}

// out[s F + f] = sum of the k_seg_dot partials of segment s (one workgroup
// per segment)
template <int F>
__global__ __launch_bounds__(kCgThreads) void k_seg_dot_final(const double2* __restrict__ part, CgState st,
                                                              double2* __restrict__ out) {
    __shared__ double2 red[kCgThreads];
    const int sg = blockIdx.x;
    const int b0 = st.sblk[sg];
    cg_sum_partials(part + int64_t(b0) * F, st.sblk[sg + 1] - b0, F, red);
    if (threadIdx.x < F) out[sg * F + threadIdx.x] = red[threadIdx.x];
}

// p = z + beta p (beta = 0 on the first iteration of an inner solve)
template <typename R, int F>
__global__ void k_cg_pupd(int64_t nq, const cplx<R>* __restrict__ z, cplx<R>* __restrict__ p, CgState st) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    if (st.nseg > 1 && st.sdone[st.rseg[k / F]]) return;  // stopped segment: p is not read again
    const cplx<R> b = from_d<R>(st.beta[seg_slot(st.rseg, k, F)]);
    p[k] = cfma(b, p[k], z[k]);
}

// dx += alpha p; r -= alpha q; z = dinv o r; per-workgroup partial sums of
// r^T z and ||r||^2.  init: only z and the sums of the starting residual
// (dx, p and q zeroed).
template <typename R, int F>
__global__ __launch_bounds__(kCgThreads) void k_cg_update(int64_t nq, cplx<R>* __restrict__ p, cplx<R>* __restrict__ q,
                                                          const cplx<R>* __restrict__ dinv, cplx<R>* __restrict__ r,
                                                          cplx<R>* __restrict__ z, cplx<R>* __restrict__ dx, CgState st,
                                                          int init, int aux) {
    __shared__ double2 red[kCgThreads];
    const int tid = threadIdx.x;
    const int f = tid % F;  // blockDim and the grid stride are multiples of F
    const SegRange sr = seg_range(st, nq, F);
    const cplx<R> al = init ? cmk<R>(R(0), R(0)) : from_d<R>(st.alpha[sr.s * F + f]);
    double2 rz = make_double2(0.0, 0.0);
    double rr = 0.0;
    // a new inner solve runs every segment; later passes skip stopped ones
    // (their sums stay zero)
    if (init && st.nseg > 1 && blockIdx.x == 0 && tid < st.nseg) st.sdone[tid] = 0;
    const int64_t k1 = !init && seg_skip(st, sr.s) ? sr.k0 : sr.k1;
    for (int64_t k = sr.k0 + tid; k < k1; k += sr.stride) {
        cplx<R> rv = r[k];
        if (init) {
            dx[k] = cmk<R>(R(0), R(0));
            p[k] = cmk<R>(R(0), R(0));
            q[k] = cmk<R>(R(0), R(0));
        } else {
            dx[k] = cfma(al, p[k], dx[k]);
            const cplx<R> aq = cmul(al, q[k]);
            rv.x -= aq.x;
            rv.y -= aq.y;
            r[k] = rv;
        }
        if (!aux) {
            // with the auxiliary space, z and r^T z come from k_aux_z
            const cplx<R> zv = cmul(dinv[k], rv);
            z[k] = zv;
            const double2 t = to_d(cmul(rv, zv));
            rz.x += t.x;
            rz.y += t.y;
        }
        rr += cabs2(rv);
    }
    // fixed-order workgroup sums per f: threads f, f+F, f+2F, ...
    const int parts = kCgThreads / F, pi = tid / F;
    red[tid] = rz;
    __syncthreads();
    for (int h = parts / 2; h >= 1; h >>= 1) {
        if (pi < h) {
            red[tid].x += red[tid + h * F].x;
            red[tid].y += red[tid + h * F].y;
        }
        __syncthreads();
    }
    if (tid < F && !aux) st.part[int64_t(blockIdx.x) * 2 * F + tid] = red[tid];
    __syncthreads();
    red[tid] = make_double2(rr, 0.0);
    __syncthreads();
    for (int h = parts / 2; h >= 1; h >>= 1) {
        if (pi < h) red[tid].x += red[tid + h * F].x;
        __syncthreads();
    }
    if (tid < F) st.part[int64_t(blockIdx.x) * 2 * F + F + tid] = red[tid];
}

// rho' and ||r||^2 from the partials of k_cg_update; beta = rho'/rho and
// the stop flags (one workgroup)
template <int F>
__global__ __launch_bounds__(kCgThreads) void k_cg_beta(int nblk, CgState st, int init) {
    __shared__ double2 red[kCgThreads];
    const int tid = threadIdx.x;
    const int sg = blockIdx.x;  // one workgroup per segment
    const int b0 = st.nseg > 1 ? st.sblk[sg] : 0;
    const int nb = st.nseg > 1 ? st.sblk[sg + 1] - b0 : nblk;
    cg_sum_partials(st.part + int64_t(b0) * 2 * F, nb, 2 * F, red);
    if (tid < F) {
        const int sf = sg * F + tid;
        const double2 rho1 = red[tid];
        const double rrv = red[F + tid].x;
        st.rr[sf] = rrv;
        int dn = init ? 0 : st.done[sf];
        double2 be = make_double2(0.0, 0.0);
        if (!(rrv <= st.tgt2[sf])) {
            const double2 rho0 = st.rho[sf];
            const double m2 = rho0.x * rho0.x + rho0.y * rho0.y;
            if (!init && !dn) {
                if (m2 > 0 && isfinite(m2)) be = cdivd(rho1, rho0);
                else dn = 1;
            }
        } else {
            dn = 1;
        }
        if (!isfinite(rrv)) dn = 1;
        st.done[sf] = dn;
        st.rho[sf] = rho1;
        st.beta[sf] = be;
    }
    if (st.nseg > 1) {
        __shared__ int live;
        if (tid == 0) live = 0;
        __syncthreads();
        if (tid < F && !st.done[sg * F + tid]) atomicOr(&live, 1);
        __syncthreads();
        if (tid == 0) st.sdone[sg] = live ? 0 : 1;
    }
}

// r = cast(r64 / ||r64||) per system
template <typename R, int F>
__global__ void k_cast_scaled(int64_t nq, const double2* __restrict__ src, const double2* __restrict__ nrm,
                              cplx<R>* __restrict__ dst, const int* __restrict__ rseg) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const double n2 = nrm[seg_slot(rseg, k, F)].x;
    const double sc = n2 > 0 ? 1.0 / sqrt(n2) : 0.0;
    dst[k] = from_d<R>(make_double2(src[k].x * sc, src[k].y * sc));
}

// x64 += ||r64|| dx per system
template <typename R, int F>
__global__ void k_cg_accum(int64_t nq, const cplx<R>* __restrict__ dx, const double2* __restrict__ nrm,
                           double2* __restrict__ x, const int* __restrict__ rseg) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const double sc = sqrt(fmax(nrm[seg_slot(rseg, k, F)].x, 0.0));
    const double2 d = to_d(dx[k]);
    x[k].x += sc * d.x;
    x[k].y += sc * d.y;
}

// D_g^-1 per nodal function and system: 1 / (-k2[f] dvol[j] + sum_s gam[s][f] dsurf[s][j])
template <typename R, int F, typename A = R>
__global__ void k_aux_dg(Op<R> op, const double* __restrict__ k2, const double2* __restrict__ gam,
                         cplx<A>* __restrict__ dginv) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= op.aux_n * F) return;
    const int64_t j = k / F;
    const int f = int(k % F);
    const double2 dv = op.dvol[j];
    double dr = -k2[f] * dv.x, di = -k2[f] * dv.y;
    if (op.dconst) {
        dr += op.dconst[j].x;
        di += op.dconst[j].y;
    }
    for (int sb = 0; sb < op.aux_nsurf; ++sb) {
        const double w = op.dsurf[int64_t(sb) * op.aux_n + j];
        dr += gam[sb * F + f].x * w;
        di += gam[sb * F + f].y * w;
    }
    const double m2 = dr * dr + di * di;
    const double wj = op.wt ? op.wt[j] : 1.0;
    dginv[k] = m2 > 0 ? from_d<A>(make_double2(wj * dr / m2, -wj * di / m2)) : cmk<A>(A(0), A(0));
}

// t = D_g^-1 o (G^T r), layout [nodal][f].  A nodal row of G^T (a vertex
// gathers from every DOF of its edge star, 60-100 entries) is split over
// LPR lanes: F lanes per entry, G = LPR / F entries in flight, reduced by a
// fixed-order butterfly; 64 / LPR rows per wavefront.
template <typename R, int F, typename A = R>
__global__ __launch_bounds__(256) void k_aux_gt(Op<R> op, const cplx<R>* __restrict__ r,
                                                const cplx<A>* __restrict__ dginv, cplx<A>* __restrict__ t) {
    constexpr int LPR = F >= 16 ? 64 : 16;
    constexpr int G = LPR / F;
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63;
    const int l = lane % LPR;
    const int f = l % F, g = l / F;
    const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const int64_t j = wave * RPW + lane / LPR;
    A ax = A(0), ay = A(0);
    if (j < op.aux_n) {
        const int e1 = op.tptr[j + 1];
        for (int e = op.tptr[j] + g; e < e1; e += G) {
            A w;
            if constexpr (sizeof(A) > sizeof(R)) w = op.tval_d[e];
            else w = op.tval[e];
            const cplx<R> v = r[int64_t(op.tidx[e]) * F + f];
            ax += w * A(v.x);
            ay += w * A(v.y);
        }
    }
#pragma unroll
    for (int off = LPR / 2; off >= F; off >>= 1) {
        ax += __shfl_xor(ax, off);
        ay += __shfl_xor(ay, off);
    }
    if (j < op.aux_n && g == 0) t[j * F + f] = cmul(dginv[j * F + f], cmk<A>(ax, ay));
}

// The same for batches of F >= 8 systems: a lane gathers V consecutive
// systems of an entry (64 bytes, as the SpMV), F / V lanes per entry and
// 64 V / F entries in flight per nodal row, one row per wavefront; the
// butterfly leaves each row sum in the lanes of its systems and lane g of a
// system group writes system f0 + g.
template <typename R, int F, typename A = R>
__global__ __launch_bounds__(256) void k_aux_gt_v(Op<R> op, const cplx<R>* __restrict__ r,
                                                  const cplx<A>* __restrict__ dginv, cplx<A>* __restrict__ t) {
    constexpr int VM = sizeof(A) == 4 ? 8 : 4;  // 4 systems per lane with FP64 accumulators
    constexpr int V = F >= VM ? VM : F;
    constexpr int FG = F / V;
    constexpr int G = 64 / FG;
    const int lane = threadIdx.x & 63;
    const int fg = lane % FG, g = lane / FG, f0 = fg * V;
    const int64_t j = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    A ax[V], ay[V];
#pragma unroll
    for (int v = 0; v < V; ++v) ax[v] = ay[v] = A(0);
    if (j < op.aux_n) {
        const int e1 = op.tptr[j + 1];
        for (int e = op.tptr[j] + g; e < e1; e += G) {
            A w;
            if constexpr (sizeof(A) > sizeof(R)) w = op.tval_d[e];
            else w = op.tval[e];
            const cplx<R>* rc = r + int64_t(op.tidx[e]) * F + f0;
            cplx<R> rv[V];
#pragma unroll
            for (int v = 0; v < V; v += 2) load_pair(rc + v, rv[v], rv[v + 1]);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                ax[v] += w * A(rv[v].x);
                ay[v] += w * A(rv[v].y);
            }
        }
    }
    A sx = A(0), sy = A(0);
#pragma unroll
    for (int v = 0; v < V; ++v) {
        A tx = ax[v], ty = ay[v];
#pragma unroll
        for (int off = 32; off >= FG; off >>= 1) {
            tx += __shfl_xor(tx, off);
            ty += __shfl_xor(ty, off);
        }
        sx = (v == g) ? tx : sx;
        sy = (v == g) ? ty : sy;
    }
    if (j < op.aux_n && g < V) {
        const int64_t o = j * F + f0 + g;
        t[o] = cmul(dginv[o], cmk<A>(sx, sy));
    }
}

// z = dinv o r + G t and per-workgroup partial sums of r^T z (same grid,
// stride and partial layout as k_cg_update, which leaves the ||r||^2 half)
template <typename R, int F, typename A = R>
__global__ __launch_bounds__(kCgThreads) void k_aux_z(int64_t nq, Op<R> op, const cplx<R>* __restrict__ dinv,
                                                      const cplx<R>* __restrict__ r, const cplx<A>* __restrict__ t,
                                                      cplx<R>* __restrict__ z, CgState st) {
    __shared__ double2 red[kCgThreads];
    const int tid = threadIdx.x;
    double2 rz = make_double2(0.0, 0.0);
    const SegRange sr = seg_range(st, nq, F);
    const int64_t k1 = seg_skip(st, sr.s) ? sr.k0 : sr.k1;  // cleared by the first pass of an inner solve
    for (int64_t k = sr.k0 + tid; k < k1; k += sr.stride) {
        const int64_t row = k / F;
        const int f = int(k % F);
        const cplx<R> rv = r[k];
        const cplx<R> dz = cmul(dinv[k], rv);
        A zx = A(dz.x), zy = A(dz.y);
        for (int e = op.gptr[row]; e < op.gptr[row + 1]; ++e) {
            A w;
            if constexpr (sizeof(A) > sizeof(R)) w = op.gval_d[e];
            else w = op.gval[e];
            const cplx<A> v = t[int64_t(op.gidx[e]) * F + f];
            zx += w * v.x;
            zy += w * v.y;
        }
        const cplx<R> zv = cmk<R>(R(zx), R(zy));
        z[k] = zv;
        const double2 d = to_d(cmul(rv, zv));
        rz.x += d.x;
        rz.y += d.y;
    }
    const int parts = kCgThreads / F, pi = tid / F;
    red[tid] = rz;
    __syncthreads();
    for (int h = parts / 2; h >= 1; h >>= 1) {
        if (pi < h) {
            red[tid].x += red[tid + h * F].x;
            red[tid].y += red[tid + h * F].y;
        }
        __syncthreads();
    }
    if (tid < F) st.part[int64_t(blockIdx.x) * 2 * F + tid] = red[tid];
}

// minimal-residual refinement step: with d = r_old - r_new = A (s dx) of the
// last correction s dx (s = ||r_old||), take x += (w - 1) s dx and
// r = r_old - w d, w = d^H r_old / d^H d per system, so that no step
// increases ||r|| (w = 1 is the plain refinement step)
template <typename R, int F>
__global__ void k_mr_apply(int64_t nq, const cplx<R>* __restrict__ dx, const double2* __restrict__ nrm_prev,
                           const double2* __restrict__ om, double2* __restrict__ x,
                           const double2* __restrict__ r_old, const double2* __restrict__ d,
                           double2* __restrict__ r, const int* __restrict__ rseg) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const int f = seg_slot(rseg, k, F);
    const double2 w = om[f];
    const double sc = sqrt(fmax(nrm_prev[f].x, 0.0));
    const double2 s = to_d(dx[k]);
    const double ar = (w.x - 1.0) * sc, ai = w.y * sc;
    x[k].x += ar * s.x - ai * s.y;
    x[k].y += ar * s.y + ai * s.x;
    const double2 dk = d[k], ro = r_old[k];
    r[k] = make_double2(ro.x - (w.x * dk.x - w.y * dk.y), ro.y - (w.x * dk.y + w.y * dk.x));
}

// undo the last refinement step of the flagged systems: x64 -= w ||r64_prev|| dx
template <typename R, int F>
__global__ void k_cg_rollback(int64_t nq, const cplx<R>* __restrict__ dx, const double2* __restrict__ nrm,
                              const double2* __restrict__ om, const int* __restrict__ flag, double2* __restrict__ x,
                              const int* __restrict__ rseg) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const int sf = seg_slot(rseg, k, F);
    if (!flag[sf]) return;
    const double sc = sqrt(fmax(nrm[sf].x, 0.0));
    const double2 w = om[sf];
    const double2 d = to_d(dx[k]);
    x[k].x -= sc * (w.x * d.x - w.y * d.y);
    x[k].y -= sc * (w.x * d.y + w.y * d.x);
}

template <typename R, int F>
int run_cocg(emg_ctx* c, const Op<R>& op, const Op<double>& op64, const double* k2_h, const double2* gam_h, int nbc,
             const double2* b64, double2* x64, const SolveParams& P, BatchStats& st) {
    hipStream_t s = c->stream;
    const int64_t n = op.n, nq = n * F;
    // row segments (emg_solver_set_segments) of this system: S sets of
    // Krylov scalars per batch, NS = S F inner systems
    const bool segd = c->seg_nseg > 1 && c->seg_n == n && op.ptr == c->red_ptr.p;  // the sweep's system only
    const int S = segd ? c->seg_nseg : 1;
    const int NS = S * F;
    DBuf<cplx<R>> pv, q, r, z, dx, dinv;
    DBuf<double2> r64, t64, nrm, bn, partial, rho, beta, alpha;
    DBuf<double> rr, tgt2, k264;
    DBuf<int> done;
    DBuf<R> k2d;
    DBuf<cplx<R>> gam;
    DBuf<double2> gam64;
    for (DBuf<cplx<R>>* b : {&pv, &q, &r, &z, &dx, &dinv}) EMG_HIP(c, b->alloc(nq));
    const bool aux = op.aux_n > 0;
    // complex64 with an FP64 real stiffness: SpMV and the gradient-space
    // correction accumulate in FP64 (the correction's t kept in FP64)
    const bool safe = sizeof(R) < sizeof(double) && op.Er64 != nullptr;
    DBuf<cplx<R>> tg, dginv;
    DBuf<double2> tg64, dginv64;
    if (aux && !safe) {
        EMG_HIP(c, tg.alloc(op.aux_n * F));
        EMG_HIP(c, dginv.alloc(op.aux_n * F));
    }
    if (aux && safe) {
        EMG_HIP(c, tg64.alloc(op.aux_n * F));
        EMG_HIP(c, dginv64.alloc(op.aux_n * F));
    }
    EMG_HIP(c, r64.alloc(nq));
    EMG_HIP(c, t64.alloc(nq));
    for (DBuf<double2>* b : {&nrm, &bn, &rho, &beta, &alpha}) EMG_HIP(c, b->alloc(NS));
    EMG_HIP(c, rr.alloc(NS));
    EMG_HIP(c, tgt2.alloc(NS));
    EMG_HIP(c, done.alloc(NS));
    const int64_t nchunk = (nq + 511) / 512;
    EMG_HIP(c, partial.alloc(std::max<int64_t>(nchunk * F, int64_t(kCgMaxBlocks) * 2 * F)));
    EMG_HIP(c, k2d.alloc(F));
    EMG_HIP(c, k264.alloc(F));
    EMG_HIP(c, gam.alloc(std::max(1, nbc) * F));
    EMG_HIP(c, gam64.alloc(std::max(1, nbc) * F));
    {
        std::vector<R> k2r(F);
        std::vector<cplx<R>> gr(std::max(1, nbc) * F, cmk<R>(R(0), R(0)));
        std::vector<double2> g64(std::max(1, nbc) * F, make_double2(0, 0));
        for (int f = 0; f < F; ++f) k2r[f] = R(k2_h[f]);
        for (int b = 0; b < nbc; ++b)
            for (int f = 0; f < F; ++f) {
                gr[b * F + f] = from_d<R>(gam_h[b * F + f]);
                g64[b * F + f] = gam_h[b * F + f];
            }
        EMG_HIP(c, hipMemcpyAsync(k2d.p, k2r.data(), sizeof(R) * F, hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(k264.p, k2_h, sizeof(double) * F, hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(gam.p, gr.data(), sizeof(cplx<R>) * gr.size(), hipMemcpyHostToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(gam64.p, g64.data(), sizeof(double2) * g64.size(), hipMemcpyHostToDevice, s));
    }
    CgState cs;
    cs.rho = rho.p;
    cs.beta = beta.p;
    cs.alpha = alpha.p;
    cs.rr = rr.p;
    cs.tgt2 = tgt2.p;
    cs.done = done.p;
    cs.part = partial.p;
    const unsigned gq = grid_for(nq, 256);
    // streaming grid: gb workgroups over the whole vector, or per segment a
    // share of them proportional to its rows (at least one)
    // (one workgroup per CU, or for large systems up to kCgMaxBlocks so that
    // each thread keeps about 32 elements: more gathers in flight in k_aux_z)
    const int64_t gwant = std::max<int64_t>(kCgUpdBlocks, nq / (int64_t(kCgThreads) * 32));
    const int64_t gcap = kCgMaxBlocks - (segd ? S : 0);
    const unsigned gb0 = unsigned(std::min<int64_t>(std::min(gwant, gcap), std::max<int64_t>(1, (nq + kCgThreads - 1) / kCgThreads)));
    DBuf<int> sblk, sdone;
    unsigned gs = gb0;
    const int* rseg = nullptr;
    if (segd) {
        std::vector<int> sb(S + 1, 0);
        for (int g = 0; g < S; ++g) {
            const int64_t rows = c->seg_row_h[g + 1] - c->seg_row_h[g];
            const int64_t cap = std::max<int64_t>(1, (rows * F + kCgThreads - 1) / kCgThreads);
            const int64_t want = int64_t(gb0) * rows / n;  // sum <= gb0, so sb[S] <= gb0 + S
            sb[g + 1] = sb[g] + int(std::max<int64_t>(1, std::min(cap, want)));
        }
        if (sb[S] > kCgMaxBlocks) return fail(c, EMG_ERR_ARG, "COCG: %d segments need too many workgroups", S);
        EMG_HIP(c, sblk.alloc(S + 1));
        EMG_HIP(c, sdone.alloc(S));
        EMG_HIP(c, hipMemcpyAsync(sblk.p, sb.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, s));
        gs = unsigned(sb[S]);
        cs.nseg = S;
        cs.sblk = sblk.p;
        cs.srow = c->seg_row.p;
        cs.rseg = c->seg_rseg.p;
        cs.sdone = sdone.p;
        rseg = c->seg_rseg.p;
    }
    const unsigned grow = grid_for(n, 4);
    const unsigned gdots = grid_for(nchunk * 64, 256);
    // SpMV grid: at most 1,024 workgroups and at most 16,384 partial sums
    // (ga F), which the one-workgroup alpha reduction reads every iteration
    // (WR-90 sweep at F = 64: 11.6 us per reduction with 1,024 workgroups;
    // 349 -> 376 freq-points/s with 256)
    const int64_t gacap = std::min<int64_t>(kCgMaxBlocks, std::max(8, 16384 / F));
    const unsigned ga = unsigned(8 * std::min<int64_t>(gacap / 8, std::max<int64_t>(1, (n + 127) / 128)));
    const unsigned gb = gs;
    k_jacobi<R, F><<<gq, 256, 0, s>>>(op, k2d.p, gam.p, dinv.p);
    const unsigned gaux = aux ? grid_for(op.aux_n * F, 256) : 1;
    // k_aux_gt: 64 / LPR nodal rows per wavefront, 4 wavefronts per workgroup
    const int rpw_gt = F >= 16 ? 1 : 4;
    const unsigned ggt = aux ? grid_for((op.aux_n + rpw_gt - 1) / rpw_gt, 4) : 1;
    // k_aux_gt_v: one nodal row per wavefront, 4 wavefronts per workgroup
    const unsigned ggv = aux ? grid_for(op.aux_n, 4) : 1;
    if (aux && !safe) k_aux_dg<R, F><<<gaux, 256, 0, s>>>(op, k264.p, gam64.p, dginv.p);
    if (aux && safe) k_aux_dg<R, F, double><<<gaux, 256, 0, s>>>(op, k264.p, gam64.p, dginv64.p);
    EMG_HIP(c, hipGetLastError());
    // z = M r with M = D^-1 (+ G D_g^-1 G^T)
    auto precondition = [&]() {
        if (!aux) return;
        if (safe) {
            if constexpr (F >= 8)
                k_aux_gt_v<R, F, double><<<ggv, 256, 0, s>>>(op, r.p, dginv64.p, tg64.p);
            else
                k_aux_gt<R, F, double><<<ggt, 256, 0, s>>>(op, r.p, dginv64.p, tg64.p);
            k_aux_z<R, F, double><<<gb, kCgThreads, 0, s>>>(nq, op, dinv.p, r.p, tg64.p, z.p, cs);
        } else {
            if constexpr (F >= 8)
                k_aux_gt_v<R, F><<<ggv, 256, 0, s>>>(op, r.p, dginv.p, tg.p);
            else
                k_aux_gt<R, F><<<ggt, 256, 0, s>>>(op, r.p, dginv.p, tg.p);
            k_aux_z<R, F><<<gb, kCgThreads, 0, s>>>(nq, op, dinv.p, r.p, tg.p, z.p, cs);
        }
    };
    // a^H b per inner system (per segment and system with segments)
    auto dot64 = [&](const double2* a, const double2* b, double2* out) {
        if (segd) {
            k_seg_dot<double, F, true><<<gs, kCgThreads, 0, s>>>(nq, a, b, cs, partial.p, 0);
            k_seg_dot_final<F><<<S, kCgThreads, 0, s>>>(partial.p, cs, out);
        } else {
            k_dots<double, F><<<dim3(gdots, 1), 256, 0, s>>>(nq, nchunk, a, nq, b, partial.p);
            k_dots_final<<<F, 64, 0, s>>>(nchunk, 1, F, partial.p, out, 0);
        }
    };
    auto dots64 = [&](const double2* a, double2* out) { dot64(a, a, out); };
    std::vector<double2> hb(NS), hr(NS);
    dots64(b64, bn.p);
    EMG_HIP(c, hipMemcpyAsync(hb.data(), bn.p, sizeof(double2) * NS, hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    std::vector<double> bnorm(NS);
    for (int f = 0; f < NS; ++f) bnorm[f] = std::sqrt(std::max(hb[f].x, 0.0));
    st.relres.assign(NS, 1.0);
    st.iters.assign(NS, 0);
    // stop flags land in pinned host memory; the graph of `blk` iterations
    // is built on first use and freed on every exit path
    struct Res {
        int* h = nullptr;
        double* hr = nullptr;
        hipGraphExec_t g = nullptr;
        ~Res() {
            if (g) (void)hipGraphExecDestroy(g);
            if (h) (void)hipHostFree(h);
            if (hr) (void)hipHostFree(hr);
        }
    } res;
    EMG_HIP(c, hipHostMalloc(reinterpret_cast<void**>(&res.h), sizeof(int) * NS, hipHostMallocDefault));
    EMG_HIP(c, hipHostMalloc(reinterpret_cast<void**>(&res.hr), sizeof(double) * NS, hipHostMallocDefault));
    int* hdone = res.h;
    double* hrr = res.hr;
    // an inner solve whose residual has not halved for this many iterations
    // has stagnated (FP32 loss of bi-orthogonality): it is stopped and the
    // refinement restarts it from the FP64 residual
    const int stall = 2000;
    // EMG_COCG_GRAPH=0: plain launches instead of graph replay (same kernels;
    // for tools that do not follow graph launches)
    const char* ge = getenv("EMG_COCG_GRAPH");
    const bool use_graph = !(ge && ge[0] == '0');
    hipGraphExec_t& graph = res.g;
    const int blk = std::max(2, (P.check_every + 1) & ~1);
    std::vector<double> ht(NS);
    // refinement guard: a system whose FP64 residual grows past 4x its best,
    // or has not improved for 3 cycles, is given up (reported unconverged)
    // (a step that made the residual worse is undone first)
    std::vector<double> best_out(NS, HUGE_VAL), prev_rel(NS, 1.0);
    std::vector<int> cyc_best(NS, 0), gave_up(NS, 0), rb(NS, 0);
    DBuf<double2> nrm_prev, r_old, dn, dd, omd;
    DBuf<int> rbd;
    EMG_HIP(c, nrm_prev.alloc(NS));
    EMG_HIP(c, rbd.alloc(NS));
    EMG_HIP(c, r_old.alloc(nq));
    EMG_HIP(c, dn.alloc(NS));
    EMG_HIP(c, dd.alloc(NS));
    EMG_HIP(c, omd.alloc(NS));
    std::vector<double2> hdn(NS), hdd(NS), hom(NS, make_double2(1.0, 0.0));
    std::vector<int> stepped(NS, 0);  // the system took an inner correction last cycle
    EMG_HIP(c, hipMemcpyAsync(omd.p, hom.data(), sizeof(double2) * NS, hipMemcpyHostToDevice, s));
    for (int cyc = 0; cyc < P.max_cycles; ++cyc) {
        if (cyc > 0) EMG_HIP(c, hipMemcpyAsync(nrm_prev.p, nrm.p, sizeof(double2) * NS, hipMemcpyDeviceToDevice, s));
        k_spmv<double, F><<<grow, 256, 0, s>>>(op64, k264.p, gam64.p, x64, nullptr, t64.p);
        k_sub<<<gq, 256, 0, s>>>(nq, b64, t64.p, r64.p);
        if (cyc > 0) {
            // minimal-residual scaling of the last correction (t64 <- d)
            k_sub<<<gq, 256, 0, s>>>(nq, r_old.p, r64.p, t64.p);
            dot64(t64.p, r_old.p, dn.p);
            dot64(t64.p, t64.p, dd.p);
            EMG_HIP(c, hipMemcpyAsync(hdn.data(), dn.p, sizeof(double2) * NS, hipMemcpyDeviceToHost, s));
            EMG_HIP(c, hipMemcpyAsync(hdd.data(), dd.p, sizeof(double2) * NS, hipMemcpyDeviceToHost, s));
            EMG_HIP(c, hipStreamSynchronize(s));
            for (int f = 0; f < NS; ++f) {
                const double den = hdd[f].x;
                hom[f] = (stepped[f] && den > 0 && std::isfinite(den) && std::isfinite(hdn[f].x) &&
                          std::isfinite(hdn[f].y))
                             ? make_double2(hdn[f].x / den, hdn[f].y / den)
                             : make_double2(1.0, 0.0);
            }
            EMG_HIP(c, hipMemcpyAsync(omd.p, hom.data(), sizeof(double2) * NS, hipMemcpyHostToDevice, s));
            k_mr_apply<R, F><<<gq, 256, 0, s>>>(nq, dx.p, nrm_prev.p, omd.p, x64, r_old.p, t64.p, r64.p, rseg);
        }
        dots64(r64.p, nrm.p);
        EMG_HIP(c, hipMemcpyAsync(r_old.p, r64.p, sizeof(double2) * nq, hipMemcpyDeviceToDevice, s));
        EMG_HIP(c, hipMemcpyAsync(hr.data(), nrm.p, sizeof(double2) * NS, hipMemcpyDeviceToHost, s));
        EMG_HIP(c, hipStreamSynchronize(s));
        EMG_HIP(c, hipGetLastError());
        bool all = true;
        for (int f = 0; f < NS; ++f) {
            const double rn = std::sqrt(std::max(hr[f].x, 0.0));
            st.relres[f] = bnorm[f] > 0 ? rn / bnorm[f] : 0.0;
            rb[f] = 0;
            if (!gave_up[f]) {
                if (st.relres[f] < best_out[f]) {
                    best_out[f] = st.relres[f];
                    cyc_best[f] = cyc;
                } else if (!(st.relres[f] <= 4.0 * best_out[f])) {
                    gave_up[f] = 1;
                    rb[f] = 1;
                    st.relres[f] = prev_rel[f];
                } else if (cyc - cyc_best[f] >= 3) {
                    gave_up[f] = 1;
                }
            }
            prev_rel[f] = st.relres[f];
            const bool ok = st.relres[f] <= P.tol || gave_up[f];
            if (!ok) all = false;
            // inner target relative to the scaled residual (norm 1); a
            // finished system gets a target it already meets
            const double t = ok || rn == 0 ? 2.0 : std::max(P.inner_rel, 0.5 * P.tol * bnorm[f] / rn);
            ht[f] = t * t;
        }
        st.cycles = cyc;
        if (std::find(rb.begin(), rb.end(), 1) != rb.end()) {
            EMG_HIP(c, hipMemcpyAsync(rbd.p, rb.data(), sizeof(int) * NS, hipMemcpyHostToDevice, s));
            k_cg_rollback<R, F><<<gq, 256, 0, s>>>(nq, dx.p, nrm_prev.p, omd.p, rbd.p, x64, rseg);
            EMG_HIP(c, hipGetLastError());
        }
        if (all) break;
        for (int f = 0; f < NS; ++f) stepped[f] = ht[f] < 1.0;
        EMG_HIP(c, hipMemcpyAsync(tgt2.p, ht.data(), sizeof(double) * NS, hipMemcpyHostToDevice, s));
        k_cast_scaled<R, F><<<gq, 256, 0, s>>>(nq, r64.p, nrm.p, r.p, rseg);
        k_cg_update<R, F><<<gb, kCgThreads, 0, s>>>(nq, pv.p, q.p, dinv.p, r.p, z.p, dx.p, cs, 1, int(aux));
        precondition();
        k_cg_beta<F><<<S, kCgThreads, 0, s>>>(int(gb), cs, 1);
        EMG_HIP(c, hipGetLastError());
        auto iterations = [&]() {
            for (int i = 0; i < blk; ++i) {
                k_cg_pupd<R, F><<<gq, 256, 0, s>>>(nq, z.p, pv.p, cs);
                if (safe)
                    launch_cg_spmv<R, F, double>(c->spmv_gr, ga, s, op, k2d.p, gam.p, pv.p, q.p, cs);
                else
                    launch_cg_spmv<R, F, R>(c->spmv_gr, ga, s, op, k2d.p, gam.p, pv.p, q.p, cs);
                // with segments p^T q is summed per segment by its own pass
                // (the SpMV's workgroups span segments)
                if (segd) k_seg_dot<R, F, false><<<gs, kCgThreads, 0, s>>>(nq, pv.p, q.p, cs, partial.p, 1);
                k_cg_alpha<F><<<S, kCgThreads, 0, s>>>(int(ga), cs);
                k_cg_update<R, F><<<gb, kCgThreads, 0, s>>>(nq, pv.p, q.p, dinv.p, r.p, z.p, dx.p, cs, 0, int(aux));
                precondition();
                k_cg_beta<F><<<S, kCgThreads, 0, s>>>(int(gb), cs, 0);
            }
        };
        if (!graph && use_graph) {
            // one graph = `blk` iterations
            EMG_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            iterations();
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &g);
            if (ec != hipSuccess) return fail(c, EMG_ERR_HIP, "COCG graph capture failed: %s", hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return fail(c, EMG_ERR_HIP, "COCG graph instantiate failed: %s", hipGetErrorString(ei));
        }
        int it = 0;
        std::vector<int> active(NS), it_best(NS, 0);
        std::vector<double> best(NS, 1.0);
        // systems with an inner target this cycle (finished and given-up
        // ones got a target they already meet)
        for (int f = 0; f < NS; ++f) active[f] = ht[f] < 1.0;
        // a block runs until its slowest system stops: once the first system
        // has converged at iteration k1, complex64 systems still running past
        // max(4 k1, k1 + 400) are stopped and restarted by the refinement
        // (complex64 inner solves can wander; complex128 ones near the end of
        // the refinement have loose targets, and cutting them would stall
        // the slower systems)
        int first_done = -1;
        while (it < P.restart) {
            if (use_graph) EMG_HIP(c, hipGraphLaunch(graph, s));
            else iterations();
            EMG_HIP(c, hipGetLastError());
            it += blk;
            for (int f = 0; f < NS; ++f) st.iters[f] += active[f] ? blk : 0;
            EMG_HIP(c, hipMemcpyAsync(hdone, done.p, sizeof(int) * NS, hipMemcpyDeviceToHost, s));
            EMG_HIP(c, hipMemcpyAsync(hrr, rr.p, sizeof(double) * NS, hipMemcpyDeviceToHost, s));
            EMG_HIP(c, hipStreamSynchronize(s));
            bool fin = true, stop = false;
            if (first_done < 0)
                for (int f = 0; f < NS; ++f)
                    if (active[f] && hdone[f]) first_done = it;
            const bool cut = sizeof(R) < sizeof(double) && first_done >= 0 &&
                             it > std::max(4 * first_done, first_done + 400);
            for (int f = 0; f < NS; ++f) {
                if (cut && !hdone[f] && active[f]) {
                    hdone[f] = 1;
                    stop = true;
                }
                if (!hdone[f] && active[f]) {
                    if (hrr[f] < 0.25 * best[f]) {
                        best[f] = hrr[f];
                        it_best[f] = it;
                    } else if (it - it_best[f] > stall) {
                        hdone[f] = 1;
                        stop = true;
                    }
                }
                if (hdone[f]) active[f] = 0;
                else fin = false;
            }
            if (fin) break;
            if (stop) EMG_HIP(c, hipMemcpyAsync(done.p, hdone, sizeof(int) * NS, hipMemcpyHostToDevice, s));
        }
        if (getenv("EMG_COCG_TRACE")) {
            std::fprintf(stderr, "[cocg] F=%d cycle %d: %d inner iterations, relres", F, cyc, it);
            for (int f = 0; f < NS; ++f) std::fprintf(stderr, " %.1e", st.relres[f]);
            std::fprintf(stderr, "\n");
        }
        k_cg_accum<R, F><<<gq, 256, 0, s>>>(nq, dx.p, nrm.p, x64, rseg);
        EMG_HIP(c, hipGetLastError());
        st.cycles = cyc + 1;
    }
    if (segd) {
        // a system's figures are those of its slowest segment (each segment's
        // residual relative to its own right-hand side)
        for (int f = 0; f < F; ++f)
            for (int g = 1; g < S; ++g) {
                st.relres[f] = std::max(st.relres[f], st.relres[g * F + f]);
                st.iters[f] = std::max(st.iters[f], st.iters[g * F + f]);
            }
        st.relres.resize(F);
        st.iters.resize(F);
    }
    return EMG_OK;
}

template <typename R, int F>
int run_method(emg_ctx* c, const Op<R>& op, const Op<double>& op64, const double* k2, const double2* gam, int nbc,
               const double2* b64, double2* x64, const SolveParams& P, BatchStats& st) {
    if (P.method == 1) {
        // FP32 IDR(s) recurrences lose bi-orthogonality over the thousands of
        // steps these systems need and break down; IDR runs in FP64 only.
        if (sizeof(R) != sizeof(double)) return fail(c, EMG_ERR_ARG, "IDR(s) requires FP64 (precision 0)");
        return run_idr<R, F>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
    }
    if (P.method == 2) return run_cocg<R, F>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
    return run_batch<R, F>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
}

template <typename R>
int dispatch_batch(emg_ctx* c, int F, const Op<R>& op, const Op<double>& op64, const double* k2, const double2* gam,
                   int nbc, const double2* b64, double2* x64, const SolveParams& P, BatchStats& st) {
#define run_batch run_method
    switch (F) {
        case 1: return run_batch<R, 1>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 2: return run_batch<R, 2>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 4: return run_batch<R, 4>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 8: return run_batch<R, 8>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 16: return run_batch<R, 16>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 32: return run_batch<R, 32>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        case 64:
            // the widest block: COCG only (GMRES and IDR keep per-system bases)
            if (P.method != 2) return fail(c, EMG_ERR_ARG, "batch width 64 needs COCG");
            return run_cocg<R, 64>(c, op, op64, k2, gam, nbc, b64, x64, P, st);
        default: return fail(c, EMG_ERR_ARG, "batch width %d not in {1,2,4,8,16,32,64}", F);
    }
#undef run_batch
}

int pow2_block(int64_t remaining, int fblock) {
    int F = 1;
    while (F * 2 <= fblock && F * 2 <= remaining) F *= 2;
    return F;
}

}  // namespace

namespace {
// real parts of complex values; flag set when any imaginary part is nonzero
__global__ void k_real_part(int64_t n, const double2* __restrict__ v, double* __restrict__ re,
                            int* __restrict__ nonreal) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double2 x = v[k];
    re[k] = x.x;
    if (x.y != 0.0) *nonreal = 1;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) b[k] = float(a[k]);
}
}  // namespace

extern "C" int emg_solver_set_spmv_lanes(emg_ctx* c, int entries_in_flight) {
    EMG_CHECK_CTX(c);
    const int g = entries_in_flight;
    if (g != 0 && g != 4 && g != 8 && g != 16 && g != 32)
        return fail(c, EMG_ERR_ARG, "emg_solver_set_spmv_lanes: %d not in {0, 4, 8, 16, 32}", g);
    c->spmv_gr = g;
    return EMG_OK;
}

namespace {
__global__ void k_row_segments(int nseg, const int64_t* __restrict__ row0, int32_t* __restrict__ rseg) {
    const int g = blockIdx.y;
    if (g >= nseg) return;
    for (int64_t r = row0[g] + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < row0[g + 1];
         r += int64_t(gridDim.x) * blockDim.x)
        rseg[r] = g;
}
}  // namespace

extern "C" int emg_solver_set_segments(emg_ctx* c, int32_t nseg, const int64_t* row_start) {
    EMG_CHECK_CTX(c);
    if (nseg <= 1) {
        c->seg_nseg = 0;
        c->seg_n = 0;
        c->seg_row_h.clear();
        return EMG_OK;
    }
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_solver_set_segments: no reduced system");
    if (nseg > 256 || !row_start) return fail(c, EMG_ERR_ARG, "emg_solver_set_segments: 2..256 segments");
    const int64_t n = c->n_red;
    if (row_start[0] != 0 || row_start[nseg] != n)
        return fail(c, EMG_ERR_ARG, "emg_solver_set_segments: segments must cover rows 0..%lld", (long long)n);
    for (int g = 0; g < nseg; ++g)
        if (row_start[g + 1] <= row_start[g])
            return fail(c, EMG_ERR_ARG, "emg_solver_set_segments: empty or decreasing segment %d", g);
    hipStream_t s = c->stream;
    c->seg_row_h.assign(row_start, row_start + nseg + 1);
    EMG_HIP(c, c->seg_row.alloc(nseg + 1));
    EMG_HIP(c, c->seg_rseg.alloc(n));
    EMG_HIP(c, hipMemcpyAsync(c->seg_row.p, c->seg_row_h.data(), sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, s));
    k_row_segments<<<dim3(64, nseg), 256, 0, s>>>(nseg, c->seg_row.p, c->seg_rseg.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(s));
    c->seg_nseg = nseg;
    c->seg_n = n;
    return EMG_OK;
}

extern "C" int emg_solver_set_safe32(emg_ctx* c, int on) {
    EMG_CHECK_CTX(c);
    c->safe32 = on != 0;
    return EMG_OK;
}

extern "C" int emg_solver_set_escalation(emg_ctx* c, int on) {
    EMG_CHECK_CTX(c);
    c->escalate = on != 0;
    return EMG_OK;
}

extern "C" int emg_sweep_escalated(emg_ctx* c, int64_t* n) {
    EMG_CHECK_CTX(c);
    if (!n) return fail(c, EMG_ERR_ARG, "emg_sweep_escalated: null output");
    *n = c->n_escalated;
    return EMG_OK;
}

// D_g[j] = g_j^T A g_j for the columns g_j of G (rows of G^T, sorted):
// the Galerkin projection of a given matrix onto the gradient space (for a
// Maxwell operator E - k0^2 B + sum gamma S, E g = 0 and this equals
// -k0^2 g^T B g + sum gamma g^T S g).  One lane per nodal function; the
// entries of A's rows are matched against g_j's support by binary search.
__global__ void k_galerkin_diag(int64_t n_nodal, const int32_t* __restrict__ tptr, const int32_t* __restrict__ tidx,
                                const double* __restrict__ tval, const int32_t* __restrict__ ptr,
                                const int32_t* __restrict__ idx, const double2* __restrict__ a,
                                double2* __restrict__ dg) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_nodal) return;
    const int b0 = tptr[j], b1 = tptr[j + 1];
    double sr = 0.0, si = 0.0;
    for (int q = b0; q < b1; ++q) {
        const int k = tidx[q];
        const double gk = tval[q];
        double ar = 0.0, ai = 0.0;
        for (int e = ptr[k]; e < ptr[k + 1]; ++e) {
            const int l = idx[e];
            int lo = b0, hi = b1 - 1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int v = tidx[mid];
                if (v < l) lo = mid + 1;
                else if (v > l) hi = mid - 1;
                else {
                    const double2 av = a[e];
                    ar += av.x * tval[mid];
                    ai += av.y * tval[mid];
                    break;
                }
            }
        }
        sr += gk * ar;
        si += gk * ai;
    }
    dg[j] = make_double2(sr, si);
}

extern "C" int emg_aux_set(emg_ctx* c, int64_t n_nodal, const int32_t* g_ptr, const int32_t* g_idx,
                           const double* g_val, const int32_t* gt_ptr, const int32_t* gt_idx, const double* gt_val,
                           const double* dvol, int n_surf, const double* dsurf) {
    EMG_CHECK_CTX(c);
    if (n_nodal == 0) {
        c->aux_n = 0;
        c->aux_nsurf = 0;
        return EMG_OK;
    }
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_aux_set: no reduced system");
    if (n_nodal < 0 || n_nodal >= (int64_t(1) << 31) || !g_ptr || !g_idx || !g_val || !gt_ptr || !gt_idx || !gt_val ||
        !dvol || n_surf < 0 || (n_surf > 0 && !dsurf))
        return fail(c, EMG_ERR_ARG, "emg_aux_set: bad arguments");
    const int64_t n = c->n_red;
    const int64_t nnz = g_ptr[n];
    if (g_ptr[0] != 0 || nnz < 0 || gt_ptr[0] != 0 || gt_ptr[n_nodal] != nnz)
        return fail(c, EMG_ERR_ARG, "emg_aux_set: G and G^T row pointers disagree");
    for (int64_t i = 0; i < n; ++i)
        if (g_ptr[i + 1] < g_ptr[i]) return fail(c, EMG_ERR_ARG, "emg_aux_set: G row pointers not monotone");
    for (int64_t j = 0; j < n_nodal; ++j)
        if (gt_ptr[j + 1] < gt_ptr[j]) return fail(c, EMG_ERR_ARG, "emg_aux_set: G^T row pointers not monotone");
    for (int64_t e = 0; e < nnz; ++e)
        if (g_idx[e] < 0 || g_idx[e] >= n_nodal || gt_idx[e] < 0 || gt_idx[e] >= n)
            return fail(c, EMG_ERR_ARG, "emg_aux_set: column index out of range");
    hipStream_t s = c->stream;
    const int64_t nz = std::max<int64_t>(nnz, 1);
    EMG_HIP(c, c->aux_gptr.alloc(n + 1));
    EMG_HIP(c, c->aux_tptr.alloc(n_nodal + 1));
    EMG_HIP(c, c->aux_gidx.alloc(nz));
    EMG_HIP(c, c->aux_tidx.alloc(nz));
    EMG_HIP(c, c->aux_gval64.alloc(nz));
    EMG_HIP(c, c->aux_tval64.alloc(nz));
    EMG_HIP(c, c->aux_gval32.alloc(nz));
    EMG_HIP(c, c->aux_tval32.alloc(nz));
    EMG_HIP(c, c->aux_dvol.alloc(2 * n_nodal));
    EMG_HIP(c, c->aux_dsurf.alloc(std::max<int64_t>(1, int64_t(n_surf) * n_nodal)));
    EMG_HIP(c, hipMemcpyAsync(c->aux_gptr.p, g_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_tptr.p, gt_ptr, sizeof(int32_t) * (n_nodal + 1), hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_gidx.p, g_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_tidx.p, gt_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_gval64.p, g_val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_tval64.p, gt_val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_dvol.p, dvol, sizeof(double) * 2 * n_nodal, hipMemcpyHostToDevice, s));
    if (n_surf > 0)
        EMG_HIP(c, hipMemcpyAsync(c->aux_dsurf.p, dsurf, sizeof(double) * n_surf * n_nodal, hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        k_to_f32<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, c->aux_gval64.p, c->aux_gval32.p);
        k_to_f32<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, c->aux_tval64.p, c->aux_tval32.p);
    }
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(s));
    c->aux_n = n_nodal;
    c->aux_nsurf = n_surf;
    c->aux_have_terms = false;
    return EMG_OK;
}

extern "C" int emg_aux_set_terms(emg_ctx* c, const double* dconst, const double* weight) {
    EMG_CHECK_CTX(c);
    if (c->aux_n == 0) return fail(c, EMG_ERR_STATE, "emg_aux_set_terms: no auxiliary space (emg_aux_set)");
    if (!dconst && !weight) {
        c->aux_have_terms = false;
        return EMG_OK;
    }
    const int64_t na = c->aux_n;
    std::vector<double> dc(2 * na, 0.0), wt(na, 1.0);
    for (int64_t j = 0; j < na; ++j) {
        if (dconst) {
            dc[2 * j] = dconst[2 * j];
            dc[2 * j + 1] = dconst[2 * j + 1];
        }
        if (weight) wt[j] = weight[j];
        if (!std::isfinite(dc[2 * j]) || !std::isfinite(dc[2 * j + 1]) || !std::isfinite(wt[j]) || wt[j] < 0)
            return fail(c, EMG_ERR_ARG, "emg_aux_set_terms: non-finite term or negative weight at %lld",
                        (long long)j);
    }
    hipStream_t s = c->stream;
    EMG_HIP(c, c->aux_dconst.alloc(2 * na));
    EMG_HIP(c, c->aux_wt.alloc(na));
    EMG_HIP(c, hipMemcpyAsync(c->aux_dconst.p, dc.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->aux_wt.p, wt.data(), sizeof(double) * na, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    c->aux_have_terms = true;
    return EMG_OK;
}

extern "C" int emg_sweep_solve(emg_ctx* c, int64_t nbatch, const double* k2, const double* gamma,
                               const double* rhs_coef, int precision, int fblock, double tol, int restart,
                               int max_cycles, double* x_out, double* relres_out, int64_t* iters_out,
                               double* overlap_out) {
    EMG_CHECK_CTX(c);
    if (!c->have_system) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: no reduced system");
    if (overlap_out && c->n_ports == 0) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: no port probes set");
    if (!c->have_surface) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: surface terms not set");
    if (c->n_forcing < 1) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: no forcing vectors");
    if (nbatch < 1 || !k2 || !rhs_coef || (c->n_bc > 0 && !gamma))
        return fail(c, EMG_ERR_ARG, "emg_sweep_solve: bad arguments");
    if (restart < 1 || (c->krylov_method == 0 && restart > 1000) || restart > 1000000 || max_cycles < 1 || !(tol > 0))
        return fail(c, EMG_ERR_ARG, "emg_sweep_solve: bad solver parameters");
    if (precision < 0 || precision > 2) return fail(c, EMG_ERR_ARG, "emg_sweep_solve: precision must be 0, 1 or 2");
    if (precision == 2 && c->krylov_method != 2)
        return fail(c, EMG_ERR_ARG, "emg_sweep_solve: precision 2 (complex64 mass values) needs COCG");
    if (c->seg_nseg > 1 && c->seg_n == c->n_red && c->krylov_method != 2)
        return fail(c, EMG_ERR_ARG, "emg_sweep_solve: row segments need COCG");
    fblock = std::max(1, std::min(fblock, c->krylov_method == 2 ? 64 : 32));
    hipStream_t s = c->stream;
    const int64_t n = c->n_red;
    const int nbc = c->n_bc, np = c->n_forcing;
    c->n_escalated = 0;
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
    if (c->krylov_method == 2) {
        // COCG iterates on the symmetric part: E - k2 (B + B^T)/2 + surface
        // (E and the surface terms are symmetric already)
        if (!c->have_bsym) {
            const int rc = build_symmetric(c, n, c->nnz_red, c->red_ptr.p, c->red_idx.p, c->B64.p, c->Bs64, c->Bs32);
            if (rc != EMG_OK) return rc;
            c->have_bsym = true;
        }
        op32.B = reinterpret_cast<const float2*>(c->Bs32.p);
    }
    if (c->krylov_method == 2) {
        // a real reduced stiffness (real mu_r) is kept in FP64 for the
        // complex64 COCG SpMV (same 8 bytes per entry as complex64)
        if (c->er64_state == 0) {
            DBuf<int> flag;
            EMG_HIP(c, flag.alloc(1));
            EMG_HIP(c, hipMemsetAsync(flag.p, 0, sizeof(int), s));
            EMG_HIP(c, c->E64r.alloc(std::max<int64_t>(c->nnz_red, 1)));
            if (c->nnz_red > 0)
                k_real_part<<<grid_for(c->nnz_red, 256), 256, 0, s>>>(
                    c->nnz_red, reinterpret_cast<const double2*>(c->E64.p), c->E64r.p, flag.p);
            EMG_HIP(c, hipGetLastError());
            int hf = 0;
            EMG_HIP(c, hipMemcpyAsync(&hf, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
            EMG_HIP(c, hipStreamSynchronize(s));
            c->er64_state = hf ? 2 : 1;
            if (hf) c->E64r.release();
        }
        if (c->er64_state == 1 && c->safe32) op32.Er64 = c->E64r.p;
    }
    Op<double> op64s = op64;
    if (c->krylov_method == 2) op64s.B = reinterpret_cast<const double2*>(c->Bs64.p);
    // complex128 COCG reads the real stiffness as 8-byte reals (of 36 bytes
    // per entry, 28 remain)
    if (c->krylov_method == 2 && c->er64_state == 1) op64s.Er64 = c->E64r.p;
    if (c->krylov_method == 2 && c->aux_n > 0) {
        if (c->aux_nsurf != nbc)
            return fail(c, EMG_ERR_STATE, "emg_sweep_solve: auxiliary space has %d surface diagonals, system has %d",
                        c->aux_nsurf, nbc);
        auto wire = [&](auto& o, auto gv, auto tv) {
            o.aux_n = c->aux_n;
            o.aux_nsurf = c->aux_nsurf;
            o.gptr = c->aux_gptr.p;
            o.gidx = c->aux_gidx.p;
            o.gval = gv;
            o.tptr = c->aux_tptr.p;
            o.tidx = c->aux_tidx.p;
            o.tval = tv;
            o.dvol = reinterpret_cast<const double2*>(c->aux_dvol.p);
            o.dsurf = c->aux_dsurf.p;
            o.dconst = c->aux_have_terms ? reinterpret_cast<const double2*>(c->aux_dconst.p) : nullptr;
            o.wt = c->aux_have_terms ? c->aux_wt.p : nullptr;
        };
        wire(op32, c->aux_gval32.p, c->aux_tval32.p);
        wire(op64s, c->aux_gval64.p, c->aux_tval64.p);
        op32.gval_d = c->aux_gval64.p;
        op32.tval_d = c->aux_tval64.p;
    }
    // precision 2: complex128 Krylov vectors, the SpMV's symmetric mass read
    // as complex64 (the FP64 refinement corrects the perturbation)
    Op<double> op64m = op64s;
    if (precision == 2) {
        if (!c->have_bsym) return fail(c, EMG_ERR_STATE, "emg_sweep_solve: symmetric mass part not built");
        op64m.B32 = reinterpret_cast<const float2*>(c->Bs32.p);
    }
    SolveParams P;
    P.tol = tol;
    P.restart = restart;
    P.max_cycles = max_cycles;
    P.inner_rel = (precision != 0 || c->krylov_method == 2) ? c->inner_rel32 : 1e-13;
    P.method = c->krylov_method;
    P.idr_s = c->idr_s;
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
                           : dispatch_batch<double>(c, F, precision == 2 ? op64m : op64s, op64, k2 + done, gam.data(),
                                                    nbc, b64.p, x64.p, P, st);
        if (rc != EMG_OK) return rc;
        if (precision == 1 && c->krylov_method == 2 && c->escalate) {
            // systems the complex64 inner solves could not finish continue with
            // complex128 inner solves from their current iterate (the
            // refinement recomputes b - A x in FP64; finished systems skip)
            int bad = 0;
            for (int f = 0; f < F; ++f) bad += !(st.relres[f] <= tol);
            if (bad > 0) {
                BatchStats st2;
                const int rc2 = dispatch_batch<double>(c, F, op64s, op64, k2 + done, gam.data(), nbc, b64.p, x64.p,
                                                       P, st2);
                if (rc2 != EMG_OK) return rc2;
                for (int f = 0; f < F; ++f) {
                    st.iters[f] += st2.iters[f];
                    st.relres[f] = st2.relres[f];
                }
                c->n_escalated += bad;
            }
        }
        if (overlap_out) {
            std::vector<double2> ov;
            const int orc = port_overlaps(c, F, x64.p, ov);
            if (orc != EMG_OK) return orc;
            std::memcpy(overlap_out + 2 * done * c->n_ports, ov.data(), sizeof(double2) * ov.size());
        }
        if (x_out) {
            k_transpose<<<grid_for(n * F, 256), 256, 0, s>>>(n, F, x64.p, xt.p, 1);
            EMG_HIP(c, hipMemcpyAsync(x_out + 2 * done * n, xt.p, sizeof(double2) * n * F, hipMemcpyDeviceToHost, s));
        }
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
    if (c->g_n != n) c->g_aux_n = 0;   // G belongs to a matrix of another size
    c->g_n = n;
    c->g_have_sym = false;
    c->g_aux_fresh = false;
    c->g_nnz = nnz;
    return EMG_OK;
}

extern "C" int emg_csr_aux_set(emg_ctx* c, int64_t n_nodal, const int32_t* g_ptr, const int32_t* g_idx,
                               const double* g_val, const int32_t* gt_ptr, const int32_t* gt_idx,
                               const double* gt_val) {
    EMG_CHECK_CTX(c);
    if (n_nodal == 0) {
        c->g_aux_n = 0;
        return EMG_OK;
    }
    if (c->g_n == 0) return fail(c, EMG_ERR_STATE, "emg_csr_aux_set: no matrix (call emg_csr_set)");
    if (n_nodal < 0 || n_nodal >= (int64_t(1) << 31) || !g_ptr || !g_idx || !g_val || !gt_ptr || !gt_idx || !gt_val)
        return fail(c, EMG_ERR_ARG, "emg_csr_aux_set: bad arguments");
    const int64_t n = c->g_n;
    const int64_t nnz = g_ptr[n];
    if (g_ptr[0] != 0 || nnz < 0 || gt_ptr[0] != 0 || gt_ptr[n_nodal] != nnz)
        return fail(c, EMG_ERR_ARG, "emg_csr_aux_set: G and G^T row pointers disagree");
    for (int64_t i = 0; i < n; ++i)
        if (g_ptr[i + 1] < g_ptr[i]) return fail(c, EMG_ERR_ARG, "emg_csr_aux_set: G row pointers not monotone");
    for (int64_t j = 0; j < n_nodal; ++j) {
        if (gt_ptr[j + 1] < gt_ptr[j]) return fail(c, EMG_ERR_ARG, "emg_csr_aux_set: G^T row pointers not monotone");
        for (int64_t e = gt_ptr[j] + 1; e < gt_ptr[j + 1]; ++e)
            if (gt_idx[e] <= gt_idx[e - 1]) return fail(c, EMG_ERR_ARG, "emg_csr_aux_set: G^T rows not sorted");
    }
    for (int64_t e = 0; e < nnz; ++e)
        if (g_idx[e] < 0 || g_idx[e] >= n_nodal || gt_idx[e] < 0 || gt_idx[e] >= n)
            return fail(c, EMG_ERR_ARG, "emg_csr_aux_set: column index out of range");
    hipStream_t s = c->stream;
    const int64_t nz = std::max<int64_t>(nnz, 1);
    EMG_HIP(c, c->g_aux_gptr.alloc(n + 1));
    EMG_HIP(c, c->g_aux_tptr.alloc(n_nodal + 1));
    EMG_HIP(c, c->g_aux_gidx.alloc(nz));
    EMG_HIP(c, c->g_aux_tidx.alloc(nz));
    EMG_HIP(c, c->g_aux_gval64.alloc(nz));
    EMG_HIP(c, c->g_aux_tval64.alloc(nz));
    EMG_HIP(c, c->g_aux_gval32.alloc(nz));
    EMG_HIP(c, c->g_aux_tval32.alloc(nz));
    EMG_HIP(c, c->g_aux_dg.alloc(2 * n_nodal));
    EMG_HIP(c, hipMemcpyAsync(c->g_aux_gptr.p, g_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_aux_tptr.p, gt_ptr, sizeof(int32_t) * (n_nodal + 1), hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_aux_gidx.p, g_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_aux_tidx.p, gt_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_aux_gval64.p, g_val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(c->g_aux_tval64.p, gt_val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        k_to_f32<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, c->g_aux_gval64.p, c->g_aux_gval32.p);
        k_to_f32<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, c->g_aux_tval64.p, c->g_aux_tval32.p);
    }
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(s));
    c->g_aux_n = n_nodal;
    c->g_aux_fresh = false;
    return EMG_OK;
}

extern "C" int emg_csr_solve(emg_ctx* c, int32_t nrhs, const double* b, int precision, double tol, int restart,
                             int max_cycles, double* x, double* relres_out, int64_t* iters_out) {
    EMG_CHECK_CTX(c);
    if (c->g_n == 0) return fail(c, EMG_ERR_STATE, "emg_csr_solve: no matrix (call emg_csr_set)");
    if (nrhs < 1 || !b || !x) return fail(c, EMG_ERR_ARG, "emg_csr_solve: bad arguments");
    if (restart < 1 || (c->krylov_method == 0 && restart > 1000) || restart > 1000000 || max_cycles < 1 || !(tol > 0))
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
    Op<double> op64s = op64;
    if (c->krylov_method == 2) {
        if (!c->g_have_sym) {
            const int rc = build_symmetric(c, n, c->g_nnz, c->g_ptr.p, c->g_idx.p, c->g_v64.p, c->g_s64, c->g_s32);
            if (rc != EMG_OK) return rc;
            c->g_have_sym = true;
        }
        op64s.E = reinterpret_cast<const double2*>(c->g_s64.p);
        op32.E = reinterpret_cast<const float2*>(c->g_s32.p);
    }
    // gradient-space correction (COCG): D_g = diag(G^T A G) of this matrix,
    // applied through D_g = -k2 dvol with k2 = -1 (the generic matrix has no
    // separate mass part, so k2 enters nothing else)
    const bool gaux = c->krylov_method == 2 && c->g_aux_n > 0;
    if (gaux) {
        if (!c->g_aux_fresh) {
            k_galerkin_diag<<<grid_for(c->g_aux_n, 256), 256, 0, s>>>(
                c->g_aux_n, c->g_aux_tptr.p, c->g_aux_tidx.p, c->g_aux_tval64.p, c->g_ptr.p, c->g_idx.p,
                reinterpret_cast<const double2*>(c->g_v64.p), reinterpret_cast<double2*>(c->g_aux_dg.p));
            EMG_HIP(c, hipGetLastError());
            c->g_aux_fresh = true;
        }
        auto wire = [&](auto& o, auto gv, auto tv) {
            o.aux_n = c->g_aux_n;
            o.aux_nsurf = 0;
            o.gptr = c->g_aux_gptr.p;
            o.gidx = c->g_aux_gidx.p;
            o.gval = gv;
            o.tptr = c->g_aux_tptr.p;
            o.tidx = c->g_aux_tidx.p;
            o.tval = tv;
            o.dvol = reinterpret_cast<const double2*>(c->g_aux_dg.p);
            o.dsurf = nullptr;
        };
        wire(op32, c->g_aux_gval32.p, c->g_aux_tval32.p);
        wire(op64s, c->g_aux_gval64.p, c->g_aux_tval64.p);
    }
    SolveParams P;
    P.tol = tol;
    P.restart = restart;
    P.max_cycles = max_cycles;
    P.inner_rel = (precision == 1 || c->krylov_method == 2) ? c->inner_rel32 : 1e-13;
    P.method = c->krylov_method;
    P.idr_s = c->idr_s;
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
        std::vector<double> k2(F, gaux ? -1.0 : 0.0);
        double2 dummy = make_double2(0, 0);
        BatchStats st;
        rc = precision == 1 ? dispatch_batch<float>(c, F, op32, op64, k2.data(), &dummy, 0, b64.p, x64.p, P, st)
                            : dispatch_batch<double>(c, F, op64s, op64, k2.data(), &dummy, 0, b64.p, x64.p, P, st);
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

extern "C" int emg_solver_config(emg_ctx* c, int method, int idr_s, double inner_rel32) {
    EMG_CHECK_CTX(c);
    if (method < 0 || method > 2)
        return fail(c, EMG_ERR_ARG, "emg_solver_config: method must be 0 (GMRES), 1 (IDR) or 2 (COCG)");
    if (idr_s < 1 || idr_s > 16) return fail(c, EMG_ERR_ARG, "emg_solver_config: idr_s must be in [1, 16]");
    if (!(inner_rel32 > 0 && inner_rel32 < 1)) return fail(c, EMG_ERR_ARG, "emg_solver_config: inner_rel32 in (0,1)");
    c->krylov_method = method;
    c->idr_s = idr_s;
    c->inner_rel32 = inner_rel32;
    return EMG_OK;
}
