// Row-owner assembly of the Nedelec-2 curl-curl (E) and mass (B) matrices
// straight into CSR values.
//
// Replaces ned2_tet_stiff_mass + _matrix_builder + scipy's COO->CSR
// duplicate sum (ASM/curlcurl.py:145-173, 180-412, 419-466).  Each workgroup
// owns a chunk of consecutive CSR rows.  Every lane takes one incidence (tet,
// canonical row u) of those rows, computes the tet's canonical geometry and
// row u of the element matrices into LDS; then every CSR slot of the chunk is
// summed over its precomputed contributor list (lane, column), which is in
// increasing tet order -- the order in which scipy's COO->CSR duplicate sum
// adds the reference's COO entries.  Results are bitwise reproducible (no
// atomics) and no element matrix or COO list ever goes to HBM.  HBM traffic
// per tet: ~1.5 KB of schedule, the vertex/material gathers (mostly L2 hits)
// and the written CSR values.
#include "emg_common.h"
#include "emg_ned2.h"

namespace {

using namespace emg;

// One workgroup per chunk; two passes (curl-curl, then mass).  Slot sums
// are segmented over the contributor list: lane l takes codes
// [20 l, 20 l + 20), finishes the slots that end in its range and passes the
// partial sum of a slot continuing from lane l-1 to that slot's first lane
// (LDS carry), which adds the carries in lane order.  Every lane does the
// same work whatever the slot sizes.
__global__ __launch_bounds__(kAsmThreads) void k_assemble_rows(
    const Chunk* __restrict__ chunks, int64_t n_chunks, const int32_t* __restrict__ inc_tet,
    const uint8_t* __restrict__ inc_u, const int4* __restrict__ inc_vid, const uint16_t* __restrict__ grp_slot,
    const uint16_t* __restrict__ contrib, const double* __restrict__ xyz, int64_t V,
    const double2* __restrict__ eps, const double2* __restrict__ inu, double2* __restrict__ Ev,
    double2* __restrict__ Bv) {
    __shared__ double stage[20][kAsmThreads];  // element row of every lane
    __shared__ double2 fac[kAsmThreads];       // complex row factor of every lane
    __shared__ double2 carry[kAsmThreads];     // partial sum of a slot continued from the previous lane
    __shared__ int closed[kAsmThreads];        // whether that slot ends in this lane's range

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

    // pass 0: curl-curl -> Ev; pass 1: mass -> Bv
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        if (active) {
            // wave-uniform loop over the distinct canonical rows of this wavefront
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
                    for (int k = 0; k < 6; ++k) asm volatile("" : "+v"(G.d[k]));
                    if (pass == 0) {
#pragma unroll
                        for (int k = 0; k < 21; ++k) asm volatile("" : "+v"(G.h[k]));
                        ned2::stiff_row(u0, G.h, G.d, &stage[0][tid], kAsmThreads);
                    } else {
#pragma unroll
                        for (int k = 0; k < 10; ++k) asm volatile("" : "+v"(G.g[k]));
                        ned2::mass_row(u0, G.g, G.d, &stage[0][tid], kAsmThreads);
                    }
                    pending = false;
                }
            }
            fac[tid] = make_double2(pass == 0 ? kf.x : mf.x, pass == 0 ? kf.y : mf.y);
        }
        __syncthreads();
        double2* out = (pass == 0 ? Ev : Bv) + ch.slot0;
        int slot = slot_first;
        bool first = true;
        double re = 0.0, im = 0.0;
        if (active) {
#pragma unroll
            for (int r = 0; r < 20; ++r) {
                const uint32_t code = (cw[r >> 1] >> (16 * (r & 1))) & 0xffffu;
                const int ln = (code >> 5) & 0xff;
                const double x = stage[code & 31][ln];
                const double2 f = fac[ln];
                re = __builtin_fma(f.x, x, re);
                im = __builtin_fma(f.y, x, im);
                if (code & kCodeLast) {
                    if (first && cont) {
                        carry[tid] = make_double2(re, im);
                        closed[tid] = 1;
                    } else {
                        out[slot] = make_double2(re, im);
                    }
                    ++slot;
                    re = 0.0;
                    im = 0.0;
                    first = false;
                }
            }
            if (tail_open && first && cont) {  // the whole range lies inside one slot
                carry[tid] = make_double2(re, im);
                closed[tid] = 0;
            }
        }
        __syncthreads();
        // a slot that starts in this lane's range and runs past it: add the
        // following lanes' partial sums in order (the chunk's last code
        // always ends a slot, so the walk stops inside the chunk)
        if (tail_open && !(first && cont)) {
            for (int q = tid + 1; q < ch.ninc; ++q) {
                const double2 cq = carry[q];
                re += cq.x;
                im += cq.y;
                if (closed[q]) break;
            }
            out[slot] = make_double2(re, im);
        }
    }
}

}  // namespace

extern "C" int emg_assemble(emg_ctx* c) {
    EMG_CHECK_CTX(c);
    if (!c->have_pattern) return fail(c, EMG_ERR_STATE, "emg_assemble: pattern not built");
    if (!c->have_material) return fail(c, EMG_ERR_STATE, "emg_assemble: materials not set");
    EMG_HIP(c, c->Ev.alloc(2 * c->nnz));
    EMG_HIP(c, c->Bv.alloc(2 * c->nnz));
    const int64_t per = (c->n_chunks + 7) / 8;
    k_assemble_rows<<<unsigned(per * 8), kAsmThreads, 0, c->stream>>>(
        c->chunks.p, c->n_chunks, c->inc_tet.p, c->inc_u.p, reinterpret_cast<const int4*>(c->inc_vid.p),
        c->grp_slot.p, c->contrib.p, c->xyz.p, c->n_nodes, reinterpret_cast<const double2*>(c->eps.p),
        reinterpret_cast<const double2*>(c->inu.p),
        reinterpret_cast<double2*>(c->Ev.p), reinterpret_cast<double2*>(c->Bv.p));
    EMG_HIP(c, hipGetLastError());
    c->have_values = true;
    return EMG_OK;
}
