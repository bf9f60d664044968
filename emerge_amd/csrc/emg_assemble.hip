// Row-owner assembly of the Nedelec-2 curl-curl (E) and mass (B) matrices
// straight into CSR values.
//
// Replaces ned2_tet_stiff_mass + _matrix_builder + scipy's COO->CSR
// duplicate sum (ASM/curlcurl.py:145-173, 180-412, 419-466).  Each workgroup
// owns a chunk of consecutive CSR rows.  Every lane takes one incidence (tet,
// canonical row u) of those rows, computes the tet's canonical geometry and
// row u of the element matrices into LDS; then every CSR slot of the chunk is
// summed over its precomputed contributor list (LDS offsets of (lane,
// column) values), which is in
// increasing tet order -- the order in which scipy's COO->CSR duplicate sum
// adds the reference's COO entries.  Results are bitwise reproducible (no
// atomics) and no element matrix or COO list ever goes to HBM.  HBM traffic
// per tet: ~1.5 KB of schedule, the vertex/material gathers (mostly L2 hits)
// and the written CSR values.
#include "emg_common.h"
#include "emg_ned2.h"

#include <type_traits>

namespace {

using namespace emg;

// One workgroup per chunk; two passes (curl-curl, then mass).  Slot sums
// are segmented over the contributor list: lane l takes codes
// [20 l, 20 l + 20), finishes the slots that end in its range and passes the
// partial sum of a slot continuing from lane l-1 to that slot's first lane
// (LDS carry), which adds the carries in lane order.  Every lane does the
// same work whatever the slot sizes.  The finished slot values are held in
// registers until every lane has read the element rows, then written into
// the same LDS one component at a time (the chunk's slots are one
// contiguous CSR range of at most kAsmMaxSlots entries) and copied out with
// coalesced stores: the wavefront writes one run of consecutive values per
// store instead of 64 scattered pieces.
//
// kRealK: every tet's 1/mu_r is real (the common case), so the curl-curl
// values are real: that pass stages its rows already multiplied by the
// lane's real factor and sums one 8-byte LDS read per code (the imaginary
// parts are exact zeros).
template <bool kRealK>
__global__ __launch_bounds__(kAsmThreads) __attribute__((amdgpu_waves_per_eu(kRealK ? 3 : 2, kRealK ? 3 : 2))) void k_assemble_rows(
    const Chunk* __restrict__ chunks, int64_t n_chunks, const int32_t* __restrict__ inc_tet,
    const uint8_t* __restrict__ inc_u, const int4* __restrict__ inc_vid, const uint16_t* __restrict__ grp_slot,
    const uint16_t* __restrict__ contrib, const double* __restrict__ xyz, int64_t V,
    const double2* __restrict__ eps, const double2* __restrict__ inu, double2* __restrict__ Ev,
    double2* __restrict__ Bv) {
    // element rows of every lane (20 x kAsmThreads doubles, 40 KB) while the
    // sums run; one component of the chunk's finished slot values
    // (at most kAsmMaxSlots doubles) afterwards.  49 KB of LDS in all: three
    // workgroups per CU.
    __shared__ double sbuf[20 * kAsmThreads];
    __shared__ double2 fac[kAsmThreads];       // complex row factor of every lane
    __shared__ double2 carry[kAsmThreads];     // partial sum of a slot continued from the previous lane
    __shared__ int closed[kAsmThreads];        // whether that slot ends in this lane's range
    __shared__ int wslot[kAsmThreads / 64 + 1];
    static_assert(kAsmThreads % 64 == 0, "whole wavefronts");
    double* stage = sbuf;  // stage[c * kAsmThreads + lane]

    // XCD-aware order: workgroups are dealt round-robin over the 8 XCDs, so
    // give each XCD a contiguous range of chunks (neighbouring rows share
    // tets and vertices in that XCD's L2).
    const int64_t per = (n_chunks + 7) / 8;
    const int64_t b = blockIdx.x;
    const int64_t ci = (b % 8) * per + b / 8;
    if (ci >= n_chunks) return;
    const Chunk ch = chunks[ci];
    const int tid = threadIdx.x;
    const bool active = tid < ch.ninc;

    int u = 0;
    ned2::TetGeom G;
    double2 kf = make_double2(0, 0), mf = make_double2(0, 0);
    uint32_t cw[10];  // this lane's 20 contributor codes, two per word
    int slot_first = 0;
#pragma unroll
    for (int r = 0; r < 10; ++r) cw[r] = 0;
    if (active) {
        const int64_t inc = int64_t(ch.inc0) + tid;
        // 20 codes = 40 bytes at 40*inc: 8-byte aligned
        const uint2* cp = reinterpret_cast<const uint2*>(contrib + inc * 20);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const uint2 v = cp[r];
            cw[2 * r] = v.x;
            cw[2 * r + 1] = v.y;
        }
        slot_first = grp_slot[inc];
        const int t = inc_tet[inc];
        u = inc_u[inc];
        const int4 wv = inc_vid[inc];
        const int w[4] = {wv.x, wv.y, wv.z, wv.w};
        int rk[4];
        ned2::ranks4(w, rk);
        double X[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            X[k][0] = xyz[w[k]];
            X[k][1] = xyz[V + w[k]];
            X[k][2] = xyz[2 * V + w[k]];
        }
        const double2 mu = inu[t];
        const double2 ep = eps[t];
        ned2::tet_geometry(X, rk, G);
        double su = 0.0;
#pragma unroll
        for (int k = 0; k < 20; ++k) su = (k == u) ? G.d[ned2::len_index(k)] : su;
        const double i3 = G.inv6v * G.inv6v * G.inv6v;
        const double ck = su * i3 * (1.0 / 120.0);
        const double cm = su * G.inv6v * (1.0 / 5040.0);
        kf = make_double2(mu.x * ck, mu.y * ck);
        mf = make_double2(ep.x * cm, ep.y * cm);
    }
    const bool cont = (cw[0] & kCodeCont) != 0;
    const bool tail_open = active && (cw[9] & (uint32_t(kCodeLast) << 16)) == 0;
    // first slot that starts in each wavefront's code range (wslot[4] =
    // nslot); read after the first barrier
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) wslot[wave] = active ? slot_first + (cont ? 1 : 0) : ch.nslot;
    if (tid == 0) wslot[kAsmThreads / 64] = ch.nslot;

    // pass 0: curl-curl -> Ev; pass 1: mass -> Bv
    auto run_pass = [&](auto pass_tag) __attribute__((always_inline)) {
        constexpr int pass = decltype(pass_tag)::value;
        constexpr bool real = pass == 0 && kRealK;
        if (active) {
            // wave-uniform loop over the distinct canonical rows of this wavefront
            double d[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) d[k] = real ? G.d[k] * kf.x : G.d[k];
            bool pending = true;
            while (true) {
                const uint64_t m = __ballot(pending);
                if (m == 0) break;
                const int leader = __ffsll((unsigned long long)m) - 1;
                const int u0 = __builtin_amdgcn_readfirstlane(__shfl(u, leader));
                if (pending && u == u0) {
                    // The row arithmetic is loop invariant and speculatable;
                    // marking its inputs as rewritten on every trip (an empty
                    // asm with in/out operands: no instruction) stops the
                    // compiler from hoisting all 20 switch cases out of this
                    // loop.
#pragma unroll
                    for (int k = 0; k < 6; ++k) asm volatile("" : "+v"(d[k]));
                    if constexpr (pass == 0) {
#pragma unroll
                        for (int k = 0; k < 21; ++k) asm volatile("" : "+v"(G.h[k]));
                        ned2::stiff_row(u0, G.h, d, stage + tid, kAsmThreads);
                    } else {
#pragma unroll
                        for (int k = 0; k < 10; ++k) asm volatile("" : "+v"(G.g[k]));
                        ned2::mass_row(u0, G.g, d, stage + tid, kAsmThreads);
                    }
                    pending = false;
                }
            }
            if constexpr (!real) fac[tid] = pass == 0 ? kf : mf;
        }
        __syncthreads();
        // Segmented sums, branch-free.  res[r] is the running sum of code r's
        // slot within this lane's range (the slot's value where code r is
        // its last).  A lane whose first codes continue the previous lane's
        // slot (cont) sums that prefix into (cre, cim) and hands it over
        // through carry[].  The codes are marked as rewritten here so that
        // the LDS addresses derived from them are formed in this phase, not
        // hoisted out of the pass loop and held through the element rows.
        // The reads go in batches of 5 codes, so that each batch's LDS
        // round trip is paid once.
        uint32_t cv[10];
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            cv[r] = cw[r];
            asm volatile("" : "+v"(cv[r]));
        }
        double2 res[20];
        int slot = slot_first;
        bool pre = cont;  // still inside the slot continued from the previous lane
        double re = 0.0, im = 0.0;
        if (active) {
            double cre = 0.0, cim = 0.0;
#pragma unroll
            for (int b0 = 0; b0 < 20; b0 += 5) {
                double xs[5];
                double2 fs[5];
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    const int r = b0 + j;
                    const uint32_t off = (cv[r >> 1] >> (16 * (r & 1))) & kCodeOffset;
                    xs[j] = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(stage) + off);
                    if constexpr (!real) fs[j] = fac[(off >> 3) & (kAsmThreads - 1)];
                }
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    const int r = b0 + j;
                    const uint32_t code = (cv[r >> 1] >> (16 * (r & 1))) & 0xffffu;
                    if constexpr (real) {
                        re += xs[j];
                    } else {
                        re = __builtin_fma(fs[j].x, xs[j], re);
                        im = __builtin_fma(fs[j].y, xs[j], im);
                    }
                    res[r] = make_double2(re, im);
                    const bool last = (code & kCodeLast) != 0;
                    cre = pre ? re : cre;
                    if constexpr (!real) cim = pre ? im : cim;
                    pre = pre && !last;
                    slot += last ? 1 : 0;
                    re = last ? 0.0 : re;
                    if constexpr (!real) im = last ? 0.0 : im;
                }
            }
            if (cont) {
                carry[tid] = make_double2(cre, cim);
                closed[tid] = pre ? 0 : 1;  // 0: the whole range lies inside that slot
            }
        }
        __syncthreads();  // element rows read, carries written: sbuf may now take the slot values
        // a slot that starts in this lane's range and runs past it: add the
        // following lanes' partial sums in order (the chunk's last code
        // always ends a slot, so the walk stops inside the chunk)
        if (tail_open && !pre) {
            for (int q = tid + 1; q < ch.ninc; ++q) {
                const double2 cq = carry[q];
                re += cq.x;
                im += cq.y;
                if (closed[q]) break;
            }
        }
        // Each wavefront writes out the slots that start in its own code
        // range, [wslot[w], wslot[w + 1]): one component at a time (8 bytes
        // per slot) it stages them in its own 20 x 64 stage entries (room
        // for 1280 slots; a wavefront's range holds at most 1280 codes), the
        // last code of every slot that starts in a lane's range writing it (a
        // masked store), reads them back in slot order and stores whole
        // 16-byte values, (v, 0) for the real curl-curl pass: every store
        // instruction writes one run of consecutive values.  No other
        // wavefront touches these entries until the next pass's element rows
        // (written by this wavefront after its reads), so this phase needs
        // no workgroup barrier.
        double2* out = (pass == 0 ? Ev : Bv) + ch.slot0;
        const int wbase = wslot[wave];
        const int wn = wslot[wave + 1] - wbase;
        constexpr int kPer = 20;  // 1280 / 64
        auto at = [&](int j) __attribute__((always_inline)) { return (j >> 6) * kAsmThreads + (wave << 6) + (j & 63); };
        auto stage_component = [&](auto comp_tag, double (&v)[kPer]) __attribute__((always_inline)) {
            constexpr int comp = decltype(comp_tag)::value;
            if (active) {
                int s = slot_first;
                bool p = cont;
#pragma unroll
                for (int r = 0; r < 20; ++r) {
                    const uint32_t code = (cv[r >> 1] >> (16 * (r & 1))) & 0xffffu;
                    const bool last = (code & kCodeLast) != 0;
                    if (last && !p) sbuf[at(s - wbase)] = comp == 0 ? res[r].x : res[r].y;
                    p = p && !last;
                    s += last ? 1 : 0;
                }
            }
            if (tail_open && !pre) sbuf[at(slot - wbase)] = comp == 0 ? re : im;
            // the lanes of one wavefront read each other's writes: LDS
            // operations of a wavefront complete in order
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < kPer; ++t) v[t] = (t * 64 + lane < wn) ? sbuf[t * kAsmThreads + (wave << 6) + lane] : 0.0;
            __builtin_amdgcn_wave_barrier();
        };
        double vr[kPer], vi[kPer];
        stage_component(std::integral_constant<int, 0>{}, vr);
        if constexpr (real) {
#pragma unroll
            for (int t = 0; t < kPer; ++t) vi[t] = 0.0;
        } else {
            stage_component(std::integral_constant<int, 1>{}, vi);
        }
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int j = t * 64 + lane;
            if (j < wn) out[wbase + j] = make_double2(vr[t], vi[t]);
        }
    };
    run_pass(std::integral_constant<int, 0>{});
    run_pass(std::integral_constant<int, 1>{});
}


// General 3x3 complex material tensors (anisotropic media, PML): the same
// schedule and slot sums, with the material inside the Gram products,
// h_ef = C_e . (mu^-1 C_f) and g_xy = grad_x . (eps grad_y), contracted with
// the ordered tables.  Four passes: (curl-curl, mass) x (real, imaginary
// part); each pass writes one component of the CSR values.  Tensors are
// stored per component, t fastest ((3,3,T) as the reference's arrays).
__global__ __launch_bounds__(kAsmThreads) void k_assemble_rows_tensor(
    const Chunk* __restrict__ chunks, int64_t n_chunks, const int32_t* __restrict__ inc_tet,
    const uint8_t* __restrict__ inc_u, const int4* __restrict__ inc_vid, const uint16_t* __restrict__ grp_slot,
    const uint16_t* __restrict__ contrib, const double* __restrict__ xyz, int64_t V, int64_t T,
    const double2* __restrict__ eps9, const double2* __restrict__ imu9, double* __restrict__ Ev,
    double* __restrict__ Bv) {
    __shared__ double stage[20][kAsmThreads];  // element row (one component) of every lane
    __shared__ double fac[kAsmThreads];        // real row factor of every lane
    __shared__ double carry[kAsmThreads];      // partial sum of a slot continued from the previous lane
    __shared__ int closed[kAsmThreads];        // whether that slot ends in this lane's range

    const int64_t per = (n_chunks + 7) / 8;
    const int64_t b = blockIdx.x;
    const int64_t ci = (b % 8) * per + b / 8;
    if (ci >= n_chunks) return;
    const Chunk ch = chunks[ci];
    const int tid = threadIdx.x;
    const bool active = tid < ch.ninc;

    int u = 0, t = 0;
    ned2::TetVecs G;
    double kf = 0.0, mf = 0.0;
    uint32_t cw[10];
    int slot_first = 0;
#pragma unroll
    for (int r = 0; r < 10; ++r) cw[r] = 0;
    if (active) {
        const int64_t inc = int64_t(ch.inc0) + tid;
        const uint2* cp = reinterpret_cast<const uint2*>(contrib + inc * 20);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const uint2 v = cp[r];
            cw[2 * r] = v.x;
            cw[2 * r + 1] = v.y;
        }
        slot_first = grp_slot[inc];
        t = inc_tet[inc];
        u = inc_u[inc];
        const int4 wv = inc_vid[inc];
        const int w[4] = {wv.x, wv.y, wv.z, wv.w};
        int rk[4];
        ned2::ranks4(w, rk);
        double X[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            X[k][0] = xyz[w[k]];
            X[k][1] = xyz[V + w[k]];
            X[k][2] = xyz[2 * V + w[k]];
        }
        ned2::tet_vectors(X, rk, G);
        double su = 0.0;
#pragma unroll
        for (int k = 0; k < 20; ++k) su = (k == u) ? G.d[ned2::len_index(k)] : su;
        kf = su * G.inv6v * G.inv6v * G.inv6v * (1.0 / 120.0);
        mf = su * G.inv6v * (1.0 / 5040.0);
    }
    const bool cont = (cw[0] & kCodeCont) != 0;
    const bool tail_open = active && (cw[9] & (uint32_t(kCodeLast) << 16)) == 0;

#pragma unroll 1
    for (int pass = 0; pass < 4; ++pass) {
        const int mat = pass >> 1, part = pass & 1;
        if (active) {
            // this pass's component of the material tensor, then its Gram products
            double M[3][3];
            const double2* src = mat == 0 ? imu9 : eps9;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const double2 v = src[int64_t(k) * T + t];
                M[k / 3][k % 3] = part == 0 ? v.x : v.y;
            }
            double hg[36];
            if (mat == 0) {
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    const double w0 = M[0][0] * G.c[f][0] + M[0][1] * G.c[f][1] + M[0][2] * G.c[f][2];
                    const double w1 = M[1][0] * G.c[f][0] + M[1][1] * G.c[f][1] + M[1][2] * G.c[f][2];
                    const double w2 = M[2][0] * G.c[f][0] + M[2][1] * G.c[f][1] + M[2][2] * G.c[f][2];
#pragma unroll
                    for (int e = 0; e < 6; ++e) hg[6 * e + f] = G.c[e][0] * w0 + G.c[e][1] * w1 + G.c[e][2] * w2;
                }
            } else {
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const double w0 = M[0][0] * G.gl[y][0] + M[0][1] * G.gl[y][1] + M[0][2] * G.gl[y][2];
                    const double w1 = M[1][0] * G.gl[y][0] + M[1][1] * G.gl[y][1] + M[1][2] * G.gl[y][2];
                    const double w2 = M[2][0] * G.gl[y][0] + M[2][1] * G.gl[y][1] + M[2][2] * G.gl[y][2];
#pragma unroll
                    for (int x = 0; x < 4; ++x) hg[4 * x + y] = G.gl[x][0] * w0 + G.gl[x][1] * w1 + G.gl[x][2] * w2;
                }
#pragma unroll
                for (int k = 16; k < 36; ++k) hg[k] = 0.0;
            }
            bool pending = true;
            while (true) {
                const uint64_t m = __ballot(pending);
                if (m == 0) break;
                const int leader = __ffsll((unsigned long long)m) - 1;
                const int u0 = __builtin_amdgcn_readfirstlane(__shfl(u, leader));
                if (pending && u == u0) {
                    // inputs marked as rewritten on every trip: the switch
                    // cases stay inside this loop (see k_assemble_rows)
#pragma unroll
                    for (int k = 0; k < 6; ++k) asm volatile("" : "+v"(G.d[k]));
                    if (mat == 0) {
#pragma unroll
                        for (int k = 0; k < 36; ++k) asm volatile("" : "+v"(hg[k]));
                        ned2::stiff_row_o(u0, hg, G.d, &stage[0][tid], kAsmThreads);
                    } else {
#pragma unroll
                        for (int k = 0; k < 16; ++k) asm volatile("" : "+v"(hg[k]));
                        ned2::mass_row_o(u0, hg, G.d, &stage[0][tid], kAsmThreads);
                    }
                    pending = false;
                }
            }
            fac[tid] = mat == 0 ? kf : mf;
        }
        __syncthreads();
        double* out = (mat == 0 ? Ev : Bv) + 2 * int64_t(ch.slot0) + part;  // component `part` of slot 0
        int slot = slot_first;
        bool first = true;
        double acc = 0.0;
        if (active) {
#pragma unroll
            for (int r = 0; r < 20; ++r) {
                const uint32_t code = (cw[r >> 1] >> (16 * (r & 1))) & 0xffffu;
                const uint32_t off = code & kCodeOffset;
                const double x = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(&stage[0][0]) + off);
                acc = __builtin_fma(fac[(off >> 3) & (kAsmThreads - 1)], x, acc);
                if (code & kCodeLast) {
                    if (first && cont) {
                        carry[tid] = acc;
                        closed[tid] = 1;
                    } else {
                        out[2 * slot] = acc;
                    }
                    ++slot;
                    acc = 0.0;
                    first = false;
                }
            }
            if (tail_open && first && cont) {
                carry[tid] = acc;
                closed[tid] = 0;
            }
        }
        __syncthreads();
        if (tail_open && !(first && cont)) {
            for (int q = tid + 1; q < ch.ninc; ++q) {
                acc += carry[q];
                if (closed[q]) break;
            }
            out[2 * slot] = acc;
        }
    }
}

// 3x3 complex inverse per tet (matinv, E/mth/optimized.py:342-370: exact
// reciprocals on a diagonal tensor, adjugate / determinant otherwise);
// (3,3,T) component layout
__global__ void k_matinv(int64_t T, const double2* __restrict__ in, double2* __restrict__ out) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double2 M[3][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k / 3][k % 3] = in[int64_t(k) * T + t];
    auto mul = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    auto sub = [](double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); };
    auto add = [](double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); };
    auto neg = [](double2 a) { return make_double2(-a.x, -a.y); };
    auto inv = [](double2 a) {
        const double m = a.x * a.x + a.y * a.y;
        return make_double2(a.x / m, -a.y / m);
    };
    auto zero = [](double2 a) { return a.x == 0.0 && a.y == 0.0; };
    double2 R[3][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k / 3][k % 3] = make_double2(0.0, 0.0);
    if (zero(M[0][1]) && zero(M[0][2]) && zero(M[1][0]) && zero(M[1][2]) && zero(M[2][0]) && zero(M[2][1])) {
        R[0][0] = inv(M[0][0]);
        R[1][1] = inv(M[1][1]);
        R[2][2] = inv(M[2][2]);
    } else {
        // the reference's term order, left to right
        const double2 det = sub(add(add(sub(sub(mul(mul(M[0][0], M[1][1]), M[2][2]), mul(mul(M[0][0], M[1][2]), M[2][1])),
                                            mul(mul(M[0][1], M[1][0]), M[2][2])),
                                        mul(mul(M[0][1], M[1][2]), M[2][0])),
                                    mul(mul(M[0][2], M[1][0]), M[2][1])),
                                mul(mul(M[0][2], M[1][1]), M[2][0]));
        R[0][0] = sub(mul(M[1][1], M[2][2]), mul(M[1][2], M[2][1]));
        R[0][1] = add(neg(mul(M[0][1], M[2][2])), mul(M[0][2], M[2][1]));
        R[0][2] = sub(mul(M[0][1], M[1][2]), mul(M[0][2], M[1][1]));
        R[1][0] = add(neg(mul(M[1][0], M[2][2])), mul(M[1][2], M[2][0]));
        R[1][1] = sub(mul(M[0][0], M[2][2]), mul(M[0][2], M[2][0]));
        R[1][2] = add(neg(mul(M[0][0], M[1][2])), mul(M[0][2], M[1][0]));
        R[2][0] = sub(mul(M[1][0], M[2][1]), mul(M[1][1], M[2][0]));
        R[2][1] = add(neg(mul(M[0][0], M[2][1])), mul(M[0][1], M[2][0]));
        R[2][2] = sub(mul(M[0][0], M[1][1]), mul(M[0][1], M[1][0]));
        // out / det as numpy's complex division
        const double dm = det.x * det.x + det.y * det.y;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const double2 a = R[k / 3][k % 3];
            R[k / 3][k % 3] = make_double2((a.x * det.x + a.y * det.y) / dm, (a.y * det.x - a.x * det.y) / dm);
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) out[int64_t(k) * T + t] = R[k / 3][k % 3];
}

}  // namespace

extern "C" int emg_assemble(emg_ctx* c) {
    EMG_CHECK_CTX(c);
    if (!c->have_pattern) return fail(c, EMG_ERR_STATE, "emg_assemble: pattern not built");
    if (!c->have_material) return fail(c, EMG_ERR_STATE, "emg_assemble: materials not set");
    EMG_HIP(c, c->Ev.alloc(2 * c->nnz));
    EMG_HIP(c, c->Bv.alloc(2 * c->nnz));
    const int64_t per = (c->n_chunks + 7) / 8;
    if (c->have_tensor) {
        k_assemble_rows_tensor<<<unsigned(per * 8), kAsmThreads, 0, c->stream>>>(
            c->chunks.p, c->n_chunks, c->inc_tet.p, c->inc_u.p, reinterpret_cast<const int4*>(c->inc_vid.p),
            c->grp_slot.p, c->contrib.p, c->xyz.p, c->n_nodes, c->n_tets,
            reinterpret_cast<const double2*>(c->eps9.p), reinterpret_cast<const double2*>(c->imu9.p), c->Ev.p,
            c->Bv.p);
        EMG_HIP(c, hipGetLastError());
        c->have_values = true;
        return EMG_OK;
    }
    auto kern = c->inu_real ? k_assemble_rows<true> : k_assemble_rows<false>;
    kern<<<unsigned(per * 8), kAsmThreads, 0, c->stream>>>(
        c->chunks.p, c->n_chunks, c->inc_tet.p, c->inc_u.p, reinterpret_cast<const int4*>(c->inc_vid.p),
        c->grp_slot.p, c->contrib.p, c->xyz.p, c->n_nodes, reinterpret_cast<const double2*>(c->eps.p),
        reinterpret_cast<const double2*>(c->inu.p),
        reinterpret_cast<double2*>(c->Ev.p), reinterpret_cast<double2*>(c->Bv.p));
    EMG_HIP(c, hipGetLastError());
    c->have_values = true;
    return EMG_OK;
}

extern "C" int emg_material_set_tensor(emg_ctx* c, const double* eps_r, const double* mu_r) {
    EMG_CHECK_CTX(c);
    if (c->n_tets == 0) return fail(c, EMG_ERR_STATE, "emg_material_set_tensor: no mesh");
    if (!eps_r || !mu_r) return fail(c, EMG_ERR_ARG, "emg_material_set_tensor: null tensor");
    const int64_t T = c->n_tets;
    hipStream_t s = c->stream;
    DBuf<double> mu;
    EMG_HIP(c, c->eps9.alloc(18 * T));
    EMG_HIP(c, c->imu9.alloc(18 * T));
    EMG_HIP(c, mu.alloc(18 * T));
    EMG_HIP(c, hipMemcpyAsync(c->eps9.p, eps_r, sizeof(double) * 18 * T, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(mu.p, mu_r, sizeof(double) * 18 * T, hipMemcpyHostToDevice, s));
    k_matinv<<<unsigned((T + 255) / 256), 256, 0, s>>>(T, reinterpret_cast<const double2*>(mu.p),
                                                        reinterpret_cast<double2*>(c->imu9.p));
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(s));
    c->have_material = true;
    c->have_tensor = true;
    c->have_values = false;
    return EMG_OK;
}
