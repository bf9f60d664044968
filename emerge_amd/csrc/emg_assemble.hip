// Row-owner assembly of the Nedelec-2 curl-curl (E) and mass (B) matrices
// straight into CSR values.
//
// Replaces ned2_tet_stiff_mass + _matrix_builder + scipy's COO->CSR
// duplicate sum (ASM/curlcurl.py:145-173, 180-412, 419-466).  Each workgroup
// owns a chunk of consecutive CSR rows.  Every lane takes one incidence (tet,
// canonical row u) of those rows, computes the tet's canonical geometry and
// row u of both element matrices, and adds it into an LDS copy of the chunk's
// CSR slots through the precomputed scatter offsets.  Lanes add in phases of
// their rank within the row (= tet order), so every CSR value is summed in a
// fixed order: results are bitwise reproducible and no element matrix or COO
// list ever goes to HBM.  HBM traffic per tet: ~1 KB of schedule + the
// written CSR values + cached vertex/material gathers.
#include "emg_common.h"
#include "emg_ned2.h"

namespace {

using namespace emg;

__global__ __launch_bounds__(kAsmThreads) void k_assemble_rows(
    const Chunk* __restrict__ chunks, int64_t n_chunks, const int32_t* __restrict__ inc_tet,
    const uint8_t* __restrict__ inc_u, const uint16_t* __restrict__ slot_beg, const uint16_t* __restrict__ contrib,
    const double* __restrict__ xyz, int64_t V, const int32_t* __restrict__ tets, int64_t T,
    const double2* __restrict__ eps, const double2* __restrict__ inu, double2* __restrict__ Ev,
    double2* __restrict__ Bv) {
    __shared__ double stage[20][kAsmThreads];  // element row of every lane
    __shared__ double fac[2][kAsmThreads];     // complex row factor of every lane
    __shared__ uint2 s_code4[5 * kAsmThreads];  // contributor codes of the chunk (u16, 4 per uint2)
    __shared__ uint16_t s_beg[kAsmMaxSlots + 1];  // first contributor of every slot, + sentinel

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
    const int ncontrib = 20 * ch.ninc;

    // Stage the chunk's schedule in LDS with coalesced loads, issued before
    // the geometry gathers so both are in flight together; the two slot
    // reductions below then touch only LDS.  The contributor list starts at
    // 40*inc0 bytes (8-byte aligned) and holds a multiple of 4 codes.
    {
        const uint2* cb4 = reinterpret_cast<const uint2*>(contrib + int64_t(ch.inc0) * 20);
        for (int k = tid; k < ncontrib / 4; k += kAsmThreads) s_code4[k] = cb4[k];
        for (int k = tid; k < ch.nslot; k += kAsmThreads) s_beg[k] = slot_beg[int64_t(ch.slot0) + k];
        if (tid == 0) s_beg[ch.nslot] = uint16_t(ncontrib);
    }
    const uint16_t* s_code = reinterpret_cast<const uint16_t*>(s_code4);

    int u = 0;
    ned2::TetGeom G;
    double e_re = 0, e_im = 0, m_re = 0, m_im = 0;
    if (active) {
        const int64_t inc = int64_t(ch.inc0) + tid;
        const int t = inc_tet[inc];
        u = inc_u[inc];
        int w[4], rk[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = tets[k * T + t];
        ned2::ranks4(w, rk);
        double X[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            X[k][0] = xyz[w[k]];
            X[k][1] = xyz[V + w[k]];
            X[k][2] = xyz[2 * V + w[k]];
        }
        ned2::tet_geometry(X, rk, G);
        double su = 0.0;
#pragma unroll
        for (int k = 0; k < 20; ++k) su = (k == u) ? G.s[k] : su;
        const double i3 = G.inv6v * G.inv6v * G.inv6v;
        const double2 mu = inu[t];
        const double2 ep = eps[t];
        const double ck = su * i3 * (1.0 / 120.0);
        const double cm = su * G.inv6v * (1.0 / 5040.0);
        e_re = mu.x * ck;
        e_im = mu.y * ck;
        m_re = ep.x * cm;
        m_im = ep.y * cm;
    }

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
                    // opaque copies of the inputs stop the compiler from
                    // hoisting all 20 switch cases out of this loop.
                    double sv[20];
#pragma unroll
                    for (int k = 0; k < 20; ++k) {
                        sv[k] = G.s[k];
                        asm volatile("" : "+v"(sv[k]));
                    }
                    if (pass == 0) {
                        double hv[21];
#pragma unroll
                        for (int k = 0; k < 21; ++k) {
                            hv[k] = G.h[k];
                            asm volatile("" : "+v"(hv[k]));
                        }
                        ned2::stiff_row(u0, hv, sv, &stage[0][tid], kAsmThreads);
                    } else {
                        double gv[10];
#pragma unroll
                        for (int k = 0; k < 10; ++k) {
                            gv[k] = G.g[k];
                            asm volatile("" : "+v"(gv[k]));
                        }
                        ned2::mass_row(u0, gv, sv, &stage[0][tid], kAsmThreads);
                    }
                    pending = false;
                }
            }
            fac[0][tid] = pass == 0 ? e_re : m_re;
            fac[1][tid] = pass == 0 ? e_im : m_im;
        }
        __syncthreads();
        double2* out = pass == 0 ? Ev : Bv;
        for (int k = tid; k < ch.nslot; k += kAsmThreads) {
            const int j0 = s_beg[k];
            const int j1 = s_beg[k + 1];
            double re = 0.0, im = 0.0;
            for (int j = j0; j < j1; ++j) {
                const int code = s_code[j];
                const int ln = code >> 5;
                const double x = stage[code & 31][ln];
                re = __builtin_fma(fac[0][ln], x, re);
                im = __builtin_fma(fac[1][ln], x, im);
            }
            out[int64_t(ch.slot0) + k] = make_double2(re, im);
        }
        __syncthreads();
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
        c->chunks.p, c->n_chunks, c->inc_tet.p, c->inc_u.p, c->slot_beg.p, c->contrib.p, c->xyz.p, c->n_nodes,
        c->tets.p,
        c->n_tets, reinterpret_cast<const double2*>(c->eps.p), reinterpret_cast<const double2*>(c->inu.p),
        reinterpret_cast<double2*>(c->Ev.p), reinterpret_cast<double2*>(c->Bv.p));
    EMG_HIP(c, hipGetLastError());
    c->have_values = true;
    return EMG_OK;
}
