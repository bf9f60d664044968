// One-time pattern build: global CSR pattern of E/B, COO->CSR scatter map and
// the row-owner assembly schedule.
//
// The pattern reproduces scipy's coo_matrix((data,(rows,cols))).tocsr() for
// the COO layout of ASM/curlcurl.py:459-465 (entry (i,j) of tet t at
// t*400+20i+j, row tet_to_field[i,t], col tet_to_field[j,t]): the unique
// (row, col) pairs in row-major, column-sorted order, explicit zeros kept.
// That is exactly a stable radix sort of packed (row<<nb | col) keys followed
// by run-length unique.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "emg_common.h"
#include "emg_ned2.h"

namespace {

using namespace emg;

__global__ void k_coo_keys(const int32_t* __restrict__ t2f, int64_t T, int nb, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ vals) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= T * 20) return;
    const int64_t t = id / 20;
    const int i = int(id - t * 20);
    const uint64_t r = uint32_t(t2f[i * T + t]);
    const int64_t base = t * 400 + int64_t(i) * 20;
    for (int j = 0; j < 20; ++j) {
        const uint64_t col = uint32_t(t2f[j * T + t]);
        keys[base + j] = (r << nb) | col;
        vals[base + j] = uint32_t(base + j);
    }
}

__global__ void k_inc_keys(const int32_t* __restrict__ t2f, int64_t T, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ vals) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= T * 20) return;
    const int64_t t = id / 20;
    const int i = int(id - t * 20);
    keys[id] = uint32_t(t2f[i * T + t]);
    vals[id] = uint32_t(id);
}

__global__ void k_heads(const uint64_t* __restrict__ keys, int64_t M, int32_t* __restrict__ head) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= M) return;
    head[id] = (id == 0 || keys[id] != keys[id - 1]) ? 1 : 0;
}

__global__ void k_fill(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                       const int32_t* __restrict__ head, const int32_t* __restrict__ slotp1, int64_t M, int nb,
                       int32_t* __restrict__ indices, int32_t* __restrict__ rowof, int32_t* __restrict__ cmap) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= M) return;
    const int32_t s = slotp1[id] - 1;
    cmap[vals[id]] = s;
    if (head[id]) {
        const uint64_t k = keys[id];
        indices[s] = int32_t(k & ((uint64_t(1) << nb) - 1));
        rowof[s] = int32_t(k >> nb);
    }
}

// rowof[0..n) sorted ascending; writes ptr[0..N] so that ptr[r] = first s with
// rowof[s] >= r.  Thread s in [0, n] fills the rows (rowof[s-1], rowof[s]].
template <typename RowT>
__global__ void k_rowptr(const RowT* __restrict__ rowof, int64_t n, int64_t N, int32_t* __restrict__ ptr) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s > n) return;
    const int64_t cur = s < n ? int64_t(rowof[s]) : N;
    const int64_t prev = s == 0 ? -1 : int64_t(rowof[s - 1]);
    for (int64_t r = prev + 1; r <= cur; ++r) ptr[r] = int32_t(s);
}

// local DOF -> canonical DOF of a tet, from the global-id ranks of its vertices
__device__ __forceinline__ void canonical_map(const int* w, int* umap) {
    int rk[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) rk[k] = (w[0] < w[k]) + (w[1] < w[k]) + (w[2] < w[k]) + (w[3] < w[k]);
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int ce = ned2::edge_index(rk[ned2::kLocalEdge[e][0]], rk[ned2::kLocalEdge[e][1]]);
        umap[e] = ce;
        umap[e + 10] = ce + 10;
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int cf = 3 - rk[ned2::kLocalFaceMissing[f]];
        umap[f + 6] = cf + 6;
        umap[f + 16] = cf + 16;
    }
}

// One workgroup per chunk.  Per incidence: its tet, vertex ids and canonical
// row u (incidences stably reordered by u inside the chunk, so that a
// wavefront evaluates few distinct element rows).  The chunk's 20*ninc
// contributor codes (stage byte offset 8 * (canonical column * kAsmThreads +
// incidence lane), flags in the low bits) are grouped by
// CSR slot, each slot's list in rank order within the row (= increasing tet
// id), which fixes the summation order.  Flags: kCodeLast on the last code of
// every slot, kCodeCont on a code at a group start (index % 20 == 0) that
// continues the previous group's slot.  Per group of 20 codes: the
// chunk-relative slot of its first code.
__global__ __launch_bounds__(kAsmThreads) void k_schedule(
    const Chunk* __restrict__ chunks, const uint32_t* __restrict__ inc_code, const uint32_t* __restrict__ inc_row,
    const int32_t* __restrict__ incptr, const int32_t* __restrict__ tets, const int32_t* __restrict__ cmap, int64_t T,
    int32_t* __restrict__ out_tet, uint8_t* __restrict__ out_u, int4* __restrict__ out_vid,
    uint16_t* __restrict__ grp_slot, uint16_t* __restrict__ contrib, int32_t* __restrict__ err) {
    __shared__ int su[kAsmThreads];
    __shared__ int cnt[kAsmMaxSlots + 1];
    __shared__ int fillc[kAsmMaxSlots];
    __shared__ uint16_t tcode[20 * kAsmThreads];
    __shared__ uint8_t trank[20 * kAsmThreads];
    __shared__ uint16_t gslot[kAsmThreads];
    const Chunk ch = chunks[blockIdx.x];
    const int l = threadIdx.x;
    for (int k = l; k <= ch.nslot; k += kAsmThreads) cnt[k] = 0;
    for (int k = l; k < ch.nslot; k += kAsmThreads) fillc[k] = 0;
    int u = 1 << 20;
    int t = 0, rank = 0;
    int off[20];
    int w[4] = {0, 0, 0, 0};
    if (l < ch.ninc) {
        const int64_t inc = int64_t(ch.inc0) + l;
        const uint32_t code = inc_code[inc];
        t = int(code / 20);
        const int i = int(code % 20);
        const int row = int(inc_row[inc]);
        rank = int(inc - incptr[row]);
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = tets[k * T + t];
        int umap[20], vmap[20];
        canonical_map(w, umap);
#pragma unroll
        for (int k = 0; k < 20; ++k) vmap[umap[k]] = k;
        u = umap[i];
        const int64_t base = int64_t(t) * 400 + int64_t(i) * 20;
        for (int v = 0; v < 20; ++v) {
            const int32_t o = cmap[base + vmap[v]] - ch.slot0;
            if (o < 0 || o >= ch.nslot) atomicExch(err, 1);
            off[v] = (o < 0 || o >= ch.nslot) ? 0 : o;
        }
        if (rank > 255) atomicExch(err, 2);
    }
    su[l] = u;
    __syncthreads();
    int pos = 0;
    if (l < ch.ninc) {
        for (int m = 0; m < ch.ninc; ++m) pos += (su[m] < u) || (su[m] == u && m < l);
        for (int v = 0; v < 20; ++v) atomicAdd(&cnt[off[v]], 1);
    }
    __syncthreads();
    if (l == 0) {
        int run = 0;
        for (int k = 0; k < ch.nslot; ++k) {
            const int c = cnt[k];
            cnt[k] = run;
            run += c;
        }
        cnt[ch.nslot] = run;
        if (run != 20 * ch.ninc) atomicExch(err, 3);
    }
    __syncthreads();
    if (l < ch.ninc) {
        for (int v = 0; v < 20; ++v) {
            const int idx = cnt[off[v]] + atomicAdd(&fillc[off[v]], 1);
            tcode[idx] = uint16_t((v * kAsmThreads + pos) << 3);
            trank[idx] = uint8_t(rank);
        }
    }
    __syncthreads();
    for (int k = l; k < ch.nslot; k += kAsmThreads) {
        const int b = cnt[k], e = cnt[k + 1];
        for (int a = b + 1; a < e; ++a) {  // insertion sort by rank (lists are short)
            const uint16_t cc = tcode[a];
            const uint8_t rr = trank[a];
            int z = a - 1;
            while (z >= b && trank[z] > rr) {
                tcode[z + 1] = tcode[z];
                trank[z + 1] = trank[z];
                --z;
            }
            tcode[z + 1] = cc;
            trank[z + 1] = rr;
        }
        tcode[e - 1] |= kCodeLast;  // every slot has at least one contributor
        for (int g = (b + 19) / 20; 20 * g < e; ++g) {
            gslot[g] = uint16_t(k);
            if (20 * g > b) tcode[20 * g] |= kCodeCont;
        }
    }
    __syncthreads();
    if (l < ch.ninc) {
        out_tet[int64_t(ch.inc0) + pos] = t;
        out_u[int64_t(ch.inc0) + pos] = uint8_t(u);
        out_vid[int64_t(ch.inc0) + pos] = make_int4(w[0], w[1], w[2], w[3]);
    }
    for (int j = l; j < 20 * ch.ninc; j += kAsmThreads) contrib[int64_t(ch.inc0) * 20 + j] = tcode[j];
    if (l < ch.ninc) grp_slot[int64_t(ch.inc0) + l] = gslot[l];
}

inline unsigned grid_for(int64_t n, int block) { return unsigned((n + block - 1) / block); }

}  // namespace

extern "C" int emg_pattern_build(emg_ctx* c, int64_t* nnz_out) {
    EMG_CHECK_CTX(c);
    if (c->n_tets == 0) return fail(c, EMG_ERR_STATE, "emg_pattern_build: no mesh");
    const int64_t T = c->n_tets, N = c->n_field, M = 400 * T, MI = 20 * T;
    int nb = 1;
    while ((int64_t(1) << nb) < N) ++nb;
    hipStream_t st = c->stream;

    // ---- COO keys, sort, unique -> indices / indptr / scatter map
    DBuf<uint64_t> kin, kout;
    DBuf<uint32_t> vin, vout;
    EMG_HIP(c, kin.alloc(M));
    EMG_HIP(c, kout.alloc(M));
    EMG_HIP(c, vin.alloc(M));
    EMG_HIP(c, vout.alloc(M));
    k_coo_keys<<<grid_for(MI, 256), 256, 0, st>>>(c->t2f.p, T, nb, kin.p, vin.p);
    EMG_HIP(c, hipGetLastError());
    size_t tmp_bytes = 0;
    EMG_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin.p, kout.p, vin.p, vout.p, size_t(M), 0u, unsigned(2 * nb), st));
    DBuf<unsigned char> tmp;
    EMG_HIP(c, tmp.alloc(tmp_bytes));
    EMG_HIP(c, rocprim::radix_sort_pairs(tmp.p, tmp_bytes, kin.p, kout.p, vin.p, vout.p, size_t(M), 0u, unsigned(2 * nb), st));
    kin.release();
    vin.release();
    DBuf<int32_t> head, slotp1;
    EMG_HIP(c, head.alloc(M));
    EMG_HIP(c, slotp1.alloc(M));
    k_heads<<<grid_for(M, 256), 256, 0, st>>>(kout.p, M, head.p);
    EMG_HIP(c, hipGetLastError());
    size_t scan_bytes = 0;
    EMG_HIP(c, rocprim::inclusive_scan(nullptr, scan_bytes, head.p, slotp1.p, size_t(M), rocprim::plus<int32_t>(), st));
    if (scan_bytes > tmp.n) {
        tmp.release();
        EMG_HIP(c, tmp.alloc(scan_bytes));
    }
    EMG_HIP(c, rocprim::inclusive_scan(tmp.p, scan_bytes, head.p, slotp1.p, size_t(M), rocprim::plus<int32_t>(), st));
    int32_t nnz32 = 0;
    EMG_HIP(c, hipMemcpyAsync(&nnz32, slotp1.p + (M - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    const int64_t nnz = nnz32;
    DBuf<int32_t> rowof;
    EMG_HIP(c, c->indices.alloc(nnz));
    EMG_HIP(c, c->indptr.alloc(N + 1));
    EMG_HIP(c, c->cmap.alloc(M));
    EMG_HIP(c, rowof.alloc(nnz));
    k_fill<<<grid_for(M, 256), 256, 0, st>>>(kout.p, vout.p, head.p, slotp1.p, M, nb, c->indices.p, rowof.p,
                                             c->cmap.p);
    EMG_HIP(c, hipGetLastError());
    k_rowptr<int32_t><<<grid_for(nnz + 1, 256), 256, 0, st>>>(rowof.p, nnz, N, c->indptr.p);
    EMG_HIP(c, hipGetLastError());
    EMG_HIP(c, hipStreamSynchronize(st));
    kout.release();
    vout.release();
    head.release();
    slotp1.release();
    rowof.release();

    // ---- incidences (tet, local row) grouped by global row, tet ascending
    DBuf<uint32_t> ikin, ikout, ivin, ivout;
    EMG_HIP(c, ikin.alloc(MI));
    EMG_HIP(c, ikout.alloc(MI));
    EMG_HIP(c, ivin.alloc(MI));
    EMG_HIP(c, ivout.alloc(MI));
    k_inc_keys<<<grid_for(MI, 256), 256, 0, st>>>(c->t2f.p, T, ikin.p, ivin.p);
    EMG_HIP(c, hipGetLastError());
    size_t itmp = 0;
    EMG_HIP(c, rocprim::radix_sort_pairs(nullptr, itmp, ikin.p, ikout.p, ivin.p, ivout.p, size_t(MI), 0u, unsigned(nb), st));
    if (itmp > tmp.n) {
        tmp.release();
        EMG_HIP(c, tmp.alloc(itmp));
    }
    EMG_HIP(c, rocprim::radix_sort_pairs(tmp.p, itmp, ikin.p, ikout.p, ivin.p, ivout.p, size_t(MI), 0u, unsigned(nb), st));
    DBuf<int32_t> incptr;
    EMG_HIP(c, incptr.alloc(N + 1));
    k_rowptr<uint32_t><<<grid_for(MI + 1, 256), 256, 0, st>>>(ikout.p, MI, N, incptr.p);
    EMG_HIP(c, hipGetLastError());

    // ---- greedy row chunks on the host
    std::vector<int32_t> h_indptr(N + 1), h_incptr(N + 1);
    EMG_HIP(c, hipMemcpyAsync(h_indptr.data(), c->indptr.p, sizeof(int32_t) * (N + 1), hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipMemcpyAsync(h_incptr.data(), incptr.p, sizeof(int32_t) * (N + 1), hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    std::vector<Chunk> hch;
    int64_t max_rank = 0;
    for (int64_t r = 0; r < N;) {
        const int64_t r0 = r;
        int64_t ninc = 0, nslot = 0, nrank = 0;
        while (r < N) {
            const int64_t di = h_incptr[r + 1] - h_incptr[r];
            const int64_t ds = h_indptr[r + 1] - h_indptr[r];
            if (di > kAsmThreads || ds > kAsmMaxSlots)
                return fail(c, EMG_ERR_ARG, "emg_pattern_build: row %lld has %lld incidences / %lld entries "
                            "(limits %d / %d)", (long long)r, (long long)di, (long long)ds, kAsmThreads, kAsmMaxSlots);
            if (r > r0 && (ninc + di > kAsmThreads || nslot + ds > kAsmMaxSlots)) break;
            ninc += di;
            nslot += ds;
            nrank = std::max(nrank, di);
            ++r;
        }
        Chunk ch{};
        ch.inc0 = h_incptr[r0];
        ch.ninc = int32_t(ninc);
        ch.slot0 = h_indptr[r0];
        ch.nslot = int32_t(nslot);
        ch.nrank = int32_t(nrank);
        hch.push_back(ch);
        max_rank = std::max(max_rank, nrank);
    }
    EMG_HIP(c, c->chunks.alloc(hch.size()));
    EMG_HIP(c, hipMemcpyAsync(c->chunks.p, hch.data(), sizeof(Chunk) * hch.size(), hipMemcpyHostToDevice, st));
    EMG_HIP(c, c->inc_tet.alloc(MI));
    EMG_HIP(c, c->inc_u.alloc(MI));
    EMG_HIP(c, c->inc_vid.alloc(4 * MI));
    EMG_HIP(c, c->grp_slot.alloc(MI));
    EMG_HIP(c, c->contrib.alloc(MI * 20));
    DBuf<int32_t> derr;
    EMG_HIP(c, derr.alloc(1));
    EMG_HIP(c, hipMemsetAsync(derr.p, 0, sizeof(int32_t), st));
    k_schedule<<<unsigned(hch.size()), kAsmThreads, 0, st>>>(c->chunks.p, ivout.p, ikout.p, incptr.p, c->tets.p,
                                                             c->cmap.p, T, c->inc_tet.p, c->inc_u.p,
                                                             reinterpret_cast<int4*>(c->inc_vid.p),
                                                             c->grp_slot.p, c->contrib.p, derr.p);
    EMG_HIP(c, hipGetLastError());
    int32_t herr = 0;
    EMG_HIP(c, hipMemcpyAsync(&herr, derr.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EMG_HIP(c, hipStreamSynchronize(st));
    if (herr) return fail(c, EMG_ERR_STATE, "emg_pattern_build: inconsistent schedule (code %d)", herr);
    c->nnz = nnz;
    c->n_chunks = int64_t(hch.size());
    c->n_inc = MI;
    c->max_rank = max_rank;
    c->have_pattern = true;
    c->have_values = false;
    c->have_system = false;  // the reduced map indexes the old pattern
    if (nnz_out) *nnz_out = nnz;
    return EMG_OK;
}
