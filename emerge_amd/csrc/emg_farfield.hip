// Stratton-Chu far field from tangential surface fields (MW/sc.py:41-132).
//
// Samples j (edge centres of a closed surface) carry E_j, H_j, position x_j
// and the area-weighted normal n_j.  Samples with |E_j| <= 1e-3 max|E| are
// left out (sc.py:58-61).  For a direction r (theta, phi):
//   Eout = Q r x [sum_j (n x E)_j G_j - Z0 r x sum_j (n x H)_j G_j],
//   G_j = exp(j k0 r . x_j),  Q = -j k0 / (4 pi),  Hout = r x Eout / Z0
// (the reference applies r x and Z0 r x per sample; both are linear, so the
// sums are taken first).
//
// Layout: one 128-byte record per sample {x, y, z, keep, (n x E) re/im x3,
// (n x H) re/im x3}.  k_ff_sum gives every lane one direction and walks its
// block's chunk of samples through LDS, 64 records (8 KB) per tile, so each
// record is read from HBM once per block and from LDS by every lane; every
// lane reads the same record at a time, so the keep test is uniform.  The
// blocks of one direction tile cover disjoint sample chunks and write 12
// partial sums per direction; k_ff_finish adds the chunks in fixed order
// (results do not depend on the launch shape's timing) and applies Q r x.
// Cost: one FP64 sincos and 24 FMAs per (direction, sample): bound by the
// FP64 vector pipe, not by memory.
#include <algorithm>
#include <mutex>

#include "emg_common.h"

namespace {

using namespace emg;

constexpr int kFFThreads = 256;      // directions per block
constexpr int kFFTile = 64;          // sample records per LDS tile
constexpr int kRec = 16;             // doubles per record
constexpr double kZ0 = 376.73031366857;   // E/const.py:3
constexpr double kLR = 0.001;             // MW/sc.py:25

__global__ void k_ff_emag(int64_t M, const double2* __restrict__ E, double* __restrict__ emag) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const double2 a = E[j], b = E[M + j], c = E[2 * M + j];
    emag[j] = sqrt((a.x * a.x + a.y * a.y) + (b.x * b.x + b.y * b.y) + (c.x * c.x + c.y * c.y));
}

// One block: max over all samples (order-independent, exact).
__global__ void k_ff_max(int64_t M, const double* __restrict__ emag, double* __restrict__ out) {
    __shared__ double red[1024];
    double m = 0.0;
    for (int64_t j = threadIdx.x; j < M; j += blockDim.x) m = fmax(m, emag[j]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_ff_prep(int64_t M, const double2* __restrict__ E, const double2* __restrict__ H,
                          const double* __restrict__ pos, const double* __restrict__ wn,
                          const double* __restrict__ emag, const double* __restrict__ emax,
                          double* __restrict__ rec) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const bool keep = emag[j] > emax[0] * kLR;
    const double nx = wn[j], ny = wn[M + j], nz = wn[2 * M + j];
    const double2 ex = E[j], ey = E[M + j], ez = E[2 * M + j];
    const double2 hx = H[j], hy = H[M + j], hz = H[2 * M + j];
    double* r = rec + j * kRec;
    r[0] = pos[j];
    r[1] = pos[M + j];
    r[2] = pos[2 * M + j];
    r[3] = keep ? 1.0 : 0.0;
    // n x E
    r[4] = ny * ez.x - nz * ey.x;
    r[5] = ny * ez.y - nz * ey.y;
    r[6] = nz * ex.x - nx * ez.x;
    r[7] = nz * ex.y - nx * ez.y;
    r[8] = nx * ey.x - ny * ex.x;
    r[9] = nx * ey.y - ny * ex.y;
    // n x H
    r[10] = ny * hz.x - nz * hy.x;
    r[11] = ny * hz.y - nz * hy.y;
    r[12] = nz * hx.x - nx * hz.x;
    r[13] = nz * hx.y - nx * hz.y;
    r[14] = nx * hy.x - ny * hx.x;
    r[15] = nx * hy.y - ny * hx.y;
}

__device__ inline void ff_direction(double th, double ph, double& rx, double& ry, double& rz) {
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    rx = st * cp;
    ry = st * sp;
    rz = ct;
}

// part[(chunk * 12 + q) * N + d]: q = 0..5 sum (n x E) G re/im per axis,
// q = 6..11 the same for n x H.
__global__ void __launch_bounds__(kFFThreads)
k_ff_sum(int64_t M, int64_t N, int64_t per_chunk, const double* __restrict__ rec,
         const double* __restrict__ theta, const double* __restrict__ phi, double k0,
         double* __restrict__ part) {
    __shared__ double tile[kFFTile * kRec];
    const int64_t d = int64_t(blockIdx.x) * kFFThreads + threadIdx.x;
    const int64_t dd = d < N ? d : N - 1;   // lanes past N compute a copy and store nothing
    double rx, ry, rz;
    ff_direction(theta[dd], phi[dd], rx, ry, rz);
    const double kx = k0 * rx, ky = k0 * ry, kz = k0 * rz;
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) acc[q] = 0.0;
    const int64_t c0 = int64_t(blockIdx.y) * per_chunk;
    const int64_t c1 = std::min<int64_t>(M, c0 + per_chunk);
    for (int64_t base = c0; base < c1; base += kFFTile) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kFFTile * kRec / kFFThreads; ++i) {
            const int idx = i * kFFThreads + threadIdx.x;
            const int64_t s = base + idx / kRec;
            tile[idx] = s < c1 ? rec[base * kRec + idx] : 0.0;   // flag 0 past the chunk
        }
        __syncthreads();
        const int cnt = int(std::min<int64_t>(kFFTile, c1 - base));
        for (int k = 0; k < cnt; ++k) {
            const double* r = tile + k * kRec;
            if (r[3] == 0.0) continue;      // same record in every lane: uniform
            double s, c;
            sincos(kx * r[0] + ky * r[1] + kz * r[2], &s, &c);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double are = r[4 + 2 * q], aim = r[5 + 2 * q];
                acc[2 * q] += are * c - aim * s;
                acc[2 * q + 1] += are * s + aim * c;
            }
        }
    }
    if (d < N) {
#pragma unroll
        for (int q = 0; q < 12; ++q) part[(int64_t(blockIdx.y) * 12 + q) * N + d] = acc[q];
    }
}

__global__ void k_ff_finish(int64_t N, int chunks, const double* __restrict__ part,
                            const double* __restrict__ theta, const double* __restrict__ phi, double k0,
                            double2* __restrict__ Eout, double2* __restrict__ Hout) {
    const int64_t d = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (d >= N) return;
    double a[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) a[q] = 0.0;
    for (int c = 0; c < chunks; ++c) {
#pragma unroll
        for (int q = 0; q < 12; ++q) a[q] += part[(int64_t(c) * 12 + q) * N + d];
    }
    double rx, ry, rz;
    ff_direction(theta[d], phi[d], rx, ry, rz);
    // v = sE - Z0 r x sH   (re, im per axis)
    const double hxr = a[6], hxi = a[7], hyr = a[8], hyi = a[9], hzr = a[10], hzi = a[11];
    const double vxr = a[0] - kZ0 * (ry * hzr - rz * hyr), vxi = a[1] - kZ0 * (ry * hzi - rz * hyi);
    const double vyr = a[2] - kZ0 * (rz * hxr - rx * hzr), vyi = a[3] - kZ0 * (rz * hxi - rx * hzi);
    const double vzr = a[4] - kZ0 * (rx * hyr - ry * hxr), vzi = a[5] - kZ0 * (rx * hyi - ry * hxi);
    // w = r x v, Eout = Q w with Q = -j k0 / (4 pi): (re, im) -> (q im, -q re)
    const double wxr = ry * vzr - rz * vyr, wxi = ry * vzi - rz * vyi;
    const double wyr = rz * vxr - rx * vzr, wyi = rz * vxi - rx * vzi;
    const double wzr = rx * vyr - ry * vxr, wzi = rx * vyi - ry * vxi;
    const double q = k0 / (4.0 * 3.14159265358979323846);
    const double2 ex = make_double2(q * wxi, -q * wxr);
    const double2 ey = make_double2(q * wyi, -q * wyr);
    const double2 ez = make_double2(q * wzi, -q * wzr);
    Eout[d] = ex;
    Eout[N + d] = ey;
    Eout[2 * N + d] = ez;
    Hout[d] = make_double2((ry * ez.x - rz * ey.x) / kZ0, (ry * ez.y - rz * ey.y) / kZ0);
    Hout[N + d] = make_double2((rz * ex.x - rx * ez.x) / kZ0, (rz * ex.y - rx * ez.y) / kZ0);
    Hout[2 * N + d] = make_double2((rx * ey.x - ry * ex.x) / kZ0, (rx * ey.y - ry * ex.y) / kZ0);
}

inline unsigned blocks_for(int64_t n, int b) { return unsigned((n + b - 1) / b); }

// Launch plan of k_ff_sum: about 2048 blocks over (direction tiles x sample
// chunks), whole LDS tiles per chunk, partial sums capped at 256 MB.
struct FFPlan {
    int64_t dtiles, chunks, per;
};
FFPlan ff_plan(int64_t M, int64_t N) {
    FFPlan p{};
    p.dtiles = (N + kFFThreads - 1) / kFFThreads;
    int64_t chunks = std::max<int64_t>(1, (2048 + p.dtiles - 1) / p.dtiles);
    chunks = std::min<int64_t>(chunks, (M + kFFTile - 1) / kFFTile);
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, (int64_t(256) << 20) / (96 * N)));
    chunks = std::min<int64_t>(chunks, 65535);
    int64_t per = (M + chunks - 1) / chunks;
    per = (per + kFFTile - 1) / kFFTile * kFFTile;
    p.per = per;
    p.chunks = (M + per - 1) / per;
    return p;
}

// The kernels of one far-field evaluation on device buffers: emag[M],
// emax[1], rec[kRec * M] and part[chunks * 12 * N] are scratch.
void ff_launch(hipStream_t st, int64_t M, const double* E, const double* H, const double* pos, const double* wn,
               int64_t N, const double* th, const double* ph, double k0, const FFPlan& p, double* emag, double* emax,
               double* rec, double* part, double* Eout, double* Hout) {
    const double2* E2 = reinterpret_cast<const double2*>(E);
    const double2* H2 = reinterpret_cast<const double2*>(H);
    k_ff_emag<<<blocks_for(M, 256), 256, 0, st>>>(M, E2, emag);
    k_ff_max<<<1, 1024, 0, st>>>(M, emag, emax);
    k_ff_prep<<<blocks_for(M, 256), 256, 0, st>>>(M, E2, H2, pos, wn, emag, emax, rec);
    const dim3 grid{unsigned(p.dtiles), unsigned(p.chunks), 1u};
    k_ff_sum<<<grid, kFFThreads, 0, st>>>(M, N, p.per, rec, th, ph, k0, part);
    k_ff_finish<<<blocks_for(N, 256), 256, 0, st>>>(N, int(p.chunks), part, th, ph, k0,
                                                    reinterpret_cast<double2*>(Eout),
                                                    reinterpret_cast<double2*>(Hout));
}

// Scratch of emg_farfield_noalloc: a file-scope device array that every GPU
// the library runs on holds until the process ends.  In host code its name
// is only a stand-in; the device address comes from hipGetSymbolAddress,
// taken once per device by emg::ff_scratch_setup (called by emg_create with
// the device current).
constexpr int64_t kScratchDoubles = int64_t(8) << 20;   // 64 MiB
constexpr int64_t kScratchAlign = 32;                    // doubles (256 bytes)
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

inline int64_t round_up(int64_t n, int64_t a) { return (n + a - 1) / a * a; }

}  // namespace

namespace emg {
void ff_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch[device] = static_cast<double*>(p);
    else (void)hipGetLastError();   // emg_farfield_noalloc reports the missing scratch
}
}  // namespace emg

extern "C" int emg_farfield(emg_ctx* c, int64_t M, const double* E, const double* H, const double* pos,
                            const double* wn, int64_t N, const double* theta, const double* phi, double k0,
                            double* Eout, double* Hout) {
    EMG_CHECK_CTX(c);
    if (M < 1 || N < 1) return fail(c, EMG_ERR_ARG, "emg_farfield: need at least one sample and one direction");
    if (!E || !H || !pos || !wn || !theta || !phi || !Eout || !Hout)
        return fail(c, EMG_ERR_ARG, "emg_farfield: null array");
    if (M > (int64_t(1) << 31) || N > (int64_t(1) << 31))
        return fail(c, EMG_ERR_ARG, "emg_farfield: more than 2^31 samples or directions");
    hipStream_t st = c->stream;
    DBuf<double>& dE = c->ff_E;
    DBuf<double>& dH = c->ff_H;
    DBuf<double>& dpos = c->ff_pos;
    DBuf<double>& dwn = c->ff_wn;
    DBuf<double>& dth = c->ff_th;
    DBuf<double>& dph = c->ff_ph;
    DBuf<double>& emag = c->ff_emag;
    DBuf<double>& emax = c->ff_emax;
    DBuf<double>& rec = c->ff_rec;
    DBuf<double>& part = c->ff_part;
    DBuf<double>& dEo = c->ff_Eo;
    DBuf<double>& dHo = c->ff_Ho;
    EMG_HIP(c, dE.alloc(6 * M));
    EMG_HIP(c, dH.alloc(6 * M));
    EMG_HIP(c, dpos.alloc(3 * M));
    EMG_HIP(c, dwn.alloc(3 * M));
    EMG_HIP(c, dth.alloc(N));
    EMG_HIP(c, dph.alloc(N));
    EMG_HIP(c, emag.alloc(M));
    EMG_HIP(c, emax.alloc(1));
    EMG_HIP(c, rec.alloc(kRec * M));
    EMG_HIP(c, hipMemcpyAsync(dE.p, E, sizeof(double) * 6 * M, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(dH.p, H, sizeof(double) * 6 * M, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(dpos.p, pos, sizeof(double) * 3 * M, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(dwn.p, wn, sizeof(double) * 3 * M, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(dth.p, theta, sizeof(double) * N, hipMemcpyHostToDevice, st));
    EMG_HIP(c, hipMemcpyAsync(dph.p, phi, sizeof(double) * N, hipMemcpyHostToDevice, st));
    const FFPlan plan = ff_plan(M, N);
    EMG_HIP(c, part.alloc(size_t(plan.chunks) * 12 * N));
    EMG_HIP(c, dEo.alloc(6 * N));
    EMG_HIP(c, dHo.alloc(6 * N));
    ff_launch(st, M, dE.p, dH.p, dpos.p, dwn.p, N, dth.p, dph.p, k0, plan, emag.p, emax.p, rec.p, part.p, dEo.p,
              dHo.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipMemcpyAsync(Eout, dEo.p, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipMemcpyAsync(Hout, dHo.p, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    return EMG_OK;
}

extern "C" int emg_farfield_noalloc(emg_ctx* c, int64_t M, const double* E, const double* H, const double* pos,
                                    const double* wn, int64_t N, const double* theta, const double* phi, double k0,
                                    double* Eout, double* Hout) {
    EMG_CHECK_CTX(c);
    if (M < 1 || N < 1)
        return fail(c, EMG_ERR_ARG, "emg_farfield_noalloc: need at least one sample and one direction");
    if (!E || !H || !pos || !wn || !theta || !phi || !Eout || !Hout)
        return fail(c, EMG_ERR_ARG, "emg_farfield_noalloc: null array");
    if (M > (int64_t(1) << 31) || N > (int64_t(1) << 31))
        return fail(c, EMG_ERR_ARG, "emg_farfield_noalloc: more than 2^31 samples or directions");
    double* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    if (!base) return fail(c, EMG_ERR_HIP, "emg_farfield_noalloc: no scratch address for device %d", c->device);
    const FFPlan plan = ff_plan(M, N);
    // scratch layout: emax | emag[M] | rec[kRec M] | part[chunks 12 N], each 256-byte aligned
    const int64_t o_emag = kScratchAlign;
    const int64_t o_rec = o_emag + round_up(M, kScratchAlign);
    const int64_t o_part = o_rec + round_up(kRec * M, kScratchAlign);
    const int64_t need = o_part + plan.chunks * 12 * N;
    if (need > kScratchDoubles)
        return fail(c, EMG_ERR_ARG,
                    "emg_farfield_noalloc: %lld samples x %lld directions need %.1f MiB of scratch, the limit is %lld MiB",
                    (long long)M, (long long)N, double(need) * 8.0 / (1 << 20),
                    (long long)(kScratchDoubles * 8 >> 20));
    ff_launch(c->stream, M, E, H, pos, wn, N, theta, phi, k0, plan, base + o_emag, base, base + o_rec,
              base + o_part, Eout, Hout);
    EMG_HIP(c, hipGetLastError());
    return EMG_OK;
}
