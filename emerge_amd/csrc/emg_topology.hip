// Mesh topology on the device: unique edges and triangles of a tet mesh and
// the maps between them, bit-exact with the reference's construction.
//
// Replaces Mesh3D._pre_update (E/mesh3d.py:386-498): edges and triangles are
// the unique sorted vertex tuples of the tets' 6 edges / 4 faces, in
// lexicographic order (:388-399); tet_to_edge / tet_to_tri follow the tet's
// local vertex pairs (0,1),(0,2),(0,3),(1,2),(3,1),(2,3) and triples
// (0,1,2),(0,2,3),(0,3,1),(1,2,3) (:426-442); tri_to_edge holds the sorted
// triangle's (i1,i2),(i2,i3),(i1,i3) (:475-484); tri_to_tet the adjacent tets
// in increasing tet order, -1234 for a missing second tet (:449-453, _MISSING_ID
// at :26); edge lengths sqrt(sum((n[:,e0]-n[:,e1])**2)) (:497).
//
// Keys: an edge (a<b) is a*V + b, a triangle (a<b<c) is (a*V + b)*V + c
// (64-bit, V < 2^21); rocPRIM radix sorts, a head-flag scan compacts the
// unique keys (so they come out in lexicographic order), and every map is a
// binary search in the sorted keys.  tri_to_tet takes the min / max tet id
// per triangle with atomics (the reference's first / second in tet order).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <climits>
#include <cstdint>

#include "emg_common.h"

namespace {

using namespace emg;

constexpr int kMissing = -1234;

__device__ __forceinline__ void sort4(int64_t v[4]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3 - i; ++j)
            if (v[j] > v[j + 1]) {
                const int64_t t = v[j];
                v[j] = v[j + 1];
                v[j + 1] = t;
            }
}

__device__ __forceinline__ uint64_t edge_key(int64_t a, int64_t b, int64_t nv) {
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    return uint64_t(lo) * uint64_t(nv) + uint64_t(hi);
}

__device__ __forceinline__ uint64_t tri_key(int64_t a, int64_t b, int64_t c, int64_t nv) {
    int64_t v[3] = {a, b, c};
    if (v[0] > v[1]) { const int64_t t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1] > v[2]) { const int64_t t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0] > v[1]) { const int64_t t = v[0]; v[0] = v[1]; v[1] = t; }
    return (uint64_t(v[0]) * uint64_t(nv) + uint64_t(v[1])) * uint64_t(nv) + uint64_t(v[2]);
}

// first index i with a[i] >= key (a sorted ascending)
__device__ __forceinline__ int64_t lower_bound(const uint64_t* __restrict__ a, int64_t n, uint64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_topo_keys(int64_t T, int64_t nv, const int32_t* __restrict__ tets, uint64_t* __restrict__ ekeys,
                            uint64_t* __restrict__ tkeys, int32_t* __restrict__ err) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t s[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        s[k] = tets[int64_t(k) * T + t];
        if (s[k] < 0 || s[k] >= nv) *err = 1;
    }
    sort4(s);
    const int pe[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
#pragma unroll
    for (int k = 0; k < 6; ++k) ekeys[int64_t(k) * T + t] = edge_key(s[pe[k][0]], s[pe[k][1]], nv);
    const int pt[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
#pragma unroll
    for (int k = 0; k < 4; ++k) tkeys[int64_t(k) * T + t] = tri_key(s[pt[k][0]], s[pt[k][1]], s[pt[k][2]], nv);
}

__global__ void k_heads(int64_t n, const uint64_t* __restrict__ keys, int32_t* __restrict__ head) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_compact(int64_t n, const uint64_t* __restrict__ keys, const int32_t* __restrict__ head,
                          const int32_t* __restrict__ pos, uint64_t* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    out[pos[i] - 1] = keys[i];
}

__global__ void k_tet_maps(int64_t T, int64_t nv, const int32_t* __restrict__ tets, const uint64_t* __restrict__ ek,
                           int64_t nE, const uint64_t* __restrict__ tk, int64_t nF, int32_t* __restrict__ t2e,
                           int32_t* __restrict__ t2t) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = tets[int64_t(k) * T + t];
    const int le[2][6] = {{0, 0, 0, 1, 3, 2}, {1, 2, 3, 2, 1, 3}};
#pragma unroll
    for (int k = 0; k < 6; ++k) t2e[int64_t(k) * T + t] = int32_t(lower_bound(ek, nE, edge_key(v[le[0][k]], v[le[1][k]], nv)));
    const int lt[3][4] = {{0, 0, 0, 1}, {1, 2, 3, 2}, {2, 3, 1, 3}};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        t2t[int64_t(k) * T + t] = int32_t(lower_bound(tk, nF, tri_key(v[lt[0][k]], v[lt[1][k]], v[lt[2][k]], nv)));
}

__global__ void k_tri_out(int64_t nF, int64_t nv, const uint64_t* __restrict__ tk, const uint64_t* __restrict__ ek,
                          int64_t nE, int32_t* __restrict__ tris, int32_t* __restrict__ tri_to_edge) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nF) return;
    const uint64_t key = tk[i];
    const uint64_t V = uint64_t(nv);
    const int64_t a = int64_t(key / (V * V)), b = int64_t((key / V) % V), c = int64_t(key % V);
    tris[i] = int32_t(a);
    tris[nF + i] = int32_t(b);
    tris[2 * nF + i] = int32_t(c);
    tri_to_edge[i] = int32_t(lower_bound(ek, nE, edge_key(a, b, nv)));
    tri_to_edge[nF + i] = int32_t(lower_bound(ek, nE, edge_key(b, c, nv)));
    tri_to_edge[2 * nF + i] = int32_t(lower_bound(ek, nE, edge_key(a, c, nv)));
}

__global__ void k_edge_out(int64_t nE, int64_t nv, const uint64_t* __restrict__ ek, const double* __restrict__ xyz,
                           int32_t* __restrict__ edges, double* __restrict__ len) {
#pragma clang fp contract(off)
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nE) return;
    const int64_t a = int64_t(ek[i] / uint64_t(nv)), b = int64_t(ek[i] % uint64_t(nv));
    edges[i] = int32_t(a);
    edges[nE + i] = int32_t(b);
    // ((dx^2 + dy^2) + dz^2) with every operation rounded (no fused
    // multiply-add): the reference's sum over axis 0, then a correctly
    // rounded square root
    const double dx = xyz[a] - xyz[b];
    const double dy = xyz[nv + a] - xyz[nv + b];
    const double dz = xyz[2 * nv + a] - xyz[2 * nv + b];
    const double sx = dx * dx, sy = dy * dy, sz = dz * dz;
    len[i] = sqrt((sx + sy) + sz);
}

__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ p, int32_t v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_tri_tets(int64_t T, const int32_t* __restrict__ t2t, int32_t* __restrict__ first,
                           int32_t* __restrict__ second, int32_t* __restrict__ count) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t tri = t2t[int64_t(k) * T + t];
        atomicMin(&first[tri], int32_t(t));
        atomicMax(&second[tri], int32_t(t));
        atomicAdd(&count[tri], 1);
    }
}

__global__ void k_tri_tets_final(int64_t nF, const int32_t* __restrict__ count, int32_t* __restrict__ second,
                                 int32_t* __restrict__ err) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nF) return;
    const int32_t n = count[i];
    if (n == 1) second[i] = kMissing;
    if (n > 2) *err = 2;
}

// unique sorted keys of `in` (n entries) into `out`; returns the count
int unique_keys(emg_ctx* c, int64_t n, DBuf<uint64_t>& in, DBuf<uint64_t>& out, int64_t& count) {
    hipStream_t s = c->stream;
    DBuf<uint64_t> sorted;
    DBuf<int32_t> head, pos;
    DBuf<unsigned char> tmp;
    EMG_HIP(c, sorted.alloc(n));
    EMG_HIP(c, head.alloc(n));
    EMG_HIP(c, pos.alloc(n));
    size_t sb = 0, cb = 0;
    EMG_HIP(c, rocprim::radix_sort_keys(nullptr, sb, in.p, sorted.p, size_t(n), 0u, 64u, s));
    EMG_HIP(c, rocprim::inclusive_scan(nullptr, cb, head.p, pos.p, size_t(n), rocprim::plus<int32_t>(), s));
    EMG_HIP(c, tmp.alloc(std::max(sb, cb)));
    EMG_HIP(c, rocprim::radix_sort_keys(tmp.p, sb, in.p, sorted.p, size_t(n), 0u, 64u, s));
    k_heads<<<unsigned((n + 255) / 256), 256, 0, s>>>(n, sorted.p, head.p);
    EMG_HIP(c, rocprim::inclusive_scan(tmp.p, cb, head.p, pos.p, size_t(n), rocprim::plus<int32_t>(), s));
    int32_t hc = 0;
    EMG_HIP(c, hipMemcpyAsync(&hc, pos.p + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    count = hc;
    EMG_HIP(c, out.alloc(std::max<int64_t>(count, 1)));
    k_compact<<<unsigned((n + 255) / 256), 256, 0, s>>>(n, sorted.p, head.p, pos.p, out.p);
    EMG_HIP(c, hipGetLastError());
    return EMG_OK;
}

// DOF maps of the second-order Nedelec space from the device topology
// (Nedelec2.__init__, E/elements/nedelec2.py:42-58): blocks
// [edge 1st | face 1st | edge 2nd | face 2nd] of sizes nE, nF, nE, nF.
__global__ void k_dof_maps(int64_t T, int64_t nE, int64_t nF, const int32_t* __restrict__ t2e,
                           const int32_t* __restrict__ t2t, const int32_t* __restrict__ r2e,
                           int32_t* __restrict__ t2f, int32_t* __restrict__ r2f, int32_t* __restrict__ e2f) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int32_t oF = int32_t(nE), oE2 = int32_t(nE + nF), oF2 = int32_t(2 * nE + nF);
    if (i < T) {
        for (int k = 0; k < 6; ++k) {
            const int32_t e = t2e[k * T + i];
            t2f[k * T + i] = e;
            t2f[(10 + k) * T + i] = e + oE2;
        }
        for (int k = 0; k < 4; ++k) {
            const int32_t f = t2t[k * T + i];
            t2f[(6 + k) * T + i] = f + oF;
            t2f[(16 + k) * T + i] = f + oF2;
        }
    }
    if (i < nF) {
        for (int k = 0; k < 3; ++k) {
            const int32_t e = r2e[k * nF + i];
            r2f[k * nF + i] = e;
            r2f[(4 + k) * nF + i] = e + oE2;
        }
        r2f[3 * nF + i] = int32_t(i) + oF;
        r2f[7 * nF + i] = int32_t(i) + oF2;
    }
    if (i < nE) {
        e2f[i] = int32_t(i);
        e2f[nE + i] = int32_t(i) + oE2;
    }
}

// removed[d] = 1 for both DOFs of every edge of a PEC triangle and all 8 DOFs
// of the triangle (ASM/assembler.py:290-307); every writer stores the same 1
__global__ void k_pec_mark(int64_t n, const int32_t* __restrict__ tri, int64_t nE, int64_t nF,
                           const int32_t* __restrict__ r2e, int32_t* __restrict__ keep) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t t = tri[k];
    for (int j = 0; j < 3; ++j) {
        const int64_t e = r2e[j * nF + t];
        keep[e] = 0;
        keep[e + nE + nF] = 0;
    }
    keep[t + nE] = 0;
    keep[t + 2 * nE + nF] = 0;
}

__global__ void k_keep_scatter(int64_t N, const int32_t* __restrict__ keep, const int32_t* __restrict__ pos,
                               int32_t* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < N && keep[i]) out[pos[i]] = int32_t(i);
}

}  // namespace

extern "C" int emg_topology_build(emg_ctx* c, int64_t V, const double* xyz_soa, int64_t T, const int32_t* tets,
                                  int64_t* n_edges, int64_t* n_tris) {
    EMG_CHECK_CTX(c);
    if (V < 4 || T < 1 || !xyz_soa || !tets || !n_edges || !n_tris)
        return fail(c, EMG_ERR_ARG, "emg_topology_build: bad arguments");
    if (V >= (int64_t(1) << 21)) return fail(c, EMG_ERR_ARG, "emg_topology_build: V >= 2^21 (64-bit triangle keys)");
    if (6 * T >= (int64_t(1) << 31)) return fail(c, EMG_ERR_ARG, "emg_topology_build: too many tets for int32 scans");
    for (int64_t i = 0; i < 4 * T; ++i)
        if (tets[i] < 0 || tets[i] >= V) return fail(c, EMG_ERR_ARG, "emg_topology_build: vertex id out of range");
    hipStream_t s = c->stream;
    DBuf<int32_t> dt, err;
    DBuf<uint64_t> ek_all, tk_all;
    EMG_HIP(c, c->topo_xyz.alloc(3 * V));
    EMG_HIP(c, dt.alloc(4 * T));
    EMG_HIP(c, err.alloc(1));
    EMG_HIP(c, ek_all.alloc(6 * T));
    EMG_HIP(c, tk_all.alloc(4 * T));
    EMG_HIP(c, hipMemcpyAsync(c->topo_xyz.p, xyz_soa, sizeof(double) * 3 * V, hipMemcpyHostToDevice, s));
    EMG_HIP(c, hipMemcpyAsync(dt.p, tets, sizeof(int32_t) * 4 * T, hipMemcpyHostToDevice, s));
    k_fill_i32<<<1, 64, 0, s>>>(1, err.p, 0);
    const unsigned gt = unsigned((T + 255) / 256);
    k_topo_keys<<<gt, 256, 0, s>>>(T, V, dt.p, ek_all.p, tk_all.p, err.p);
    EMG_HIP(c, hipGetLastError());
    int64_t nE = 0, nF = 0;
    int rc = unique_keys(c, 6 * T, ek_all, c->topo_ekeys, nE);
    if (rc != EMG_OK) return rc;
    rc = unique_keys(c, 4 * T, tk_all, c->topo_tkeys, nF);
    if (rc != EMG_OK) return rc;
    EMG_HIP(c, c->topo_edges.alloc(2 * nE));
    EMG_HIP(c, c->topo_len.alloc(nE));
    EMG_HIP(c, c->topo_tris.alloc(3 * nF));
    EMG_HIP(c, c->topo_t2e.alloc(6 * T));
    EMG_HIP(c, c->topo_t2t.alloc(4 * T));
    EMG_HIP(c, c->topo_r2e.alloc(3 * nF));
    EMG_HIP(c, c->topo_r2t.alloc(2 * nF));
    DBuf<int32_t> cnt;
    EMG_HIP(c, cnt.alloc(nF));
    k_tet_maps<<<gt, 256, 0, s>>>(T, V, dt.p, c->topo_ekeys.p, nE, c->topo_tkeys.p, nF, c->topo_t2e.p, c->topo_t2t.p);
    k_tri_out<<<unsigned((nF + 255) / 256), 256, 0, s>>>(nF, V, c->topo_tkeys.p, c->topo_ekeys.p, nE, c->topo_tris.p,
                                                          c->topo_r2e.p);
    k_edge_out<<<unsigned((nE + 255) / 256), 256, 0, s>>>(nE, V, c->topo_ekeys.p, c->topo_xyz.p, c->topo_edges.p,
                                                           c->topo_len.p);
    int32_t* first = c->topo_r2t.p;
    int32_t* second = c->topo_r2t.p + nF;
    const unsigned gf = unsigned((nF + 255) / 256);
    k_fill_i32<<<gf, 256, 0, s>>>(nF, first, INT32_MAX);
    k_fill_i32<<<gf, 256, 0, s>>>(nF, second, -1);
    k_fill_i32<<<gf, 256, 0, s>>>(nF, cnt.p, 0);
    k_tri_tets<<<gt, 256, 0, s>>>(T, c->topo_t2t.p, first, second, cnt.p);
    k_tri_tets_final<<<gf, 256, 0, s>>>(nF, cnt.p, second, err.p);
    EMG_HIP(c, hipGetLastError());
    int32_t herr = 0;
    EMG_HIP(c, hipMemcpyAsync(&herr, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    if (herr == 1) return fail(c, EMG_ERR_ARG, "emg_topology_build: vertex id out of range");
    if (herr == 2) return fail(c, EMG_ERR_ARG, "emg_topology_build: non-manifold mesh (a triangle in > 2 tets)");
    c->topo_T = T;
    c->topo_nE = nE;
    c->topo_nF = nF;
    *n_edges = nE;
    *n_tris = nF;
    return EMG_OK;
}

extern "C" int emg_topology_get(emg_ctx* c, int32_t* edges, int32_t* tris, int32_t* tet_to_edge, int32_t* tet_to_tri,
                                int32_t* tri_to_edge, int32_t* tri_to_tet, double* edge_lengths) {
    EMG_CHECK_CTX(c);
    if (c->topo_T == 0) return fail(c, EMG_ERR_STATE, "emg_topology_get: no topology built");
    hipStream_t s = c->stream;
    const int64_t T = c->topo_T, nE = c->topo_nE, nF = c->topo_nF;
    auto get = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        return dst ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
    };
    EMG_HIP(c, get(edges, c->topo_edges.p, sizeof(int32_t) * 2 * nE));
    EMG_HIP(c, get(tris, c->topo_tris.p, sizeof(int32_t) * 3 * nF));
    EMG_HIP(c, get(tet_to_edge, c->topo_t2e.p, sizeof(int32_t) * 6 * T));
    EMG_HIP(c, get(tet_to_tri, c->topo_t2t.p, sizeof(int32_t) * 4 * T));
    EMG_HIP(c, get(tri_to_edge, c->topo_r2e.p, sizeof(int32_t) * 3 * nF));
    EMG_HIP(c, get(tri_to_tet, c->topo_r2t.p, sizeof(int32_t) * 2 * nF));
    EMG_HIP(c, get(edge_lengths, c->topo_len.p, sizeof(double) * nE));
    EMG_HIP(c, hipStreamSynchronize(s));
    return EMG_OK;
}

extern "C" int emg_dof_maps(emg_ctx* c, int32_t* tet_to_field, int32_t* tri_to_field, int32_t* edge_to_field) {
    EMG_CHECK_CTX(c);
    if (c->topo_T == 0) return fail(c, EMG_ERR_STATE, "emg_dof_maps: no topology built");
    const int64_t T = c->topo_T, nE = c->topo_nE, nF = c->topo_nF;
    if (2 * (nE + nF) > INT_MAX) return fail(c, EMG_ERR_ARG, "emg_dof_maps: more than 2^31 DOFs");
    hipStream_t s = c->stream;
    DBuf<int32_t> t2f, r2f, e2f;
    EMG_HIP(c, t2f.alloc(20 * T));
    EMG_HIP(c, r2f.alloc(8 * nF));
    EMG_HIP(c, e2f.alloc(2 * nE));
    const int64_t n = std::max(T, std::max(nE, nF));
    k_dof_maps<<<unsigned((n + 255) / 256), 256, 0, s>>>(T, nE, nF, c->topo_t2e.p, c->topo_t2t.p, c->topo_r2e.p,
                                                         t2f.p, r2f.p, e2f.p);
    EMG_HIP(c, hipGetLastError());
    if (tet_to_field)
        EMG_HIP(c, hipMemcpyAsync(tet_to_field, t2f.p, sizeof(int32_t) * 20 * T, hipMemcpyDeviceToHost, s));
    if (tri_to_field)
        EMG_HIP(c, hipMemcpyAsync(tri_to_field, r2f.p, sizeof(int32_t) * 8 * nF, hipMemcpyDeviceToHost, s));
    if (edge_to_field)
        EMG_HIP(c, hipMemcpyAsync(edge_to_field, e2f.p, sizeof(int32_t) * 2 * nE, hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    return EMG_OK;
}

extern "C" int emg_pec_keep(emg_ctx* c, int64_t n_pec, const int32_t* pec_tris, int64_t* n_keep, int32_t* keep_ids) {
    EMG_CHECK_CTX(c);
    if (c->topo_T == 0) return fail(c, EMG_ERR_STATE, "emg_pec_keep: no topology built");
    if (n_pec < 0 || (n_pec > 0 && !pec_tris) || !n_keep) return fail(c, EMG_ERR_ARG, "emg_pec_keep: bad arguments");
    const int64_t nE = c->topo_nE, nF = c->topo_nF, N = 2 * (nE + nF);
    if (N > INT_MAX) return fail(c, EMG_ERR_ARG, "emg_pec_keep: more than 2^31 DOFs");
    for (int64_t k = 0; k < n_pec; ++k)
        if (pec_tris[k] < 0 || pec_tris[k] >= nF)
            return fail(c, EMG_ERR_ARG, "emg_pec_keep: triangle id %lld out of range", (long long)pec_tris[k]);
    hipStream_t s = c->stream;
    DBuf<int32_t> keep, pos, tri, out;
    EMG_HIP(c, keep.alloc(N));
    EMG_HIP(c, pos.alloc(N));
    const unsigned bn = unsigned((N + 255) / 256);
    k_fill_i32<<<bn, 256, 0, s>>>(N, keep.p, 1);
    if (n_pec > 0) {
        EMG_HIP(c, tri.alloc(n_pec));
        EMG_HIP(c, hipMemcpyAsync(tri.p, pec_tris, sizeof(int32_t) * n_pec, hipMemcpyHostToDevice, s));
        k_pec_mark<<<unsigned((n_pec + 255) / 256), 256, 0, s>>>(n_pec, tri.p, nE, nF, c->topo_r2e.p, keep.p);
    }
    EMG_HIP(c, hipGetLastError());
    size_t sb = 0;
    EMG_HIP(c, rocprim::exclusive_scan(nullptr, sb, keep.p, pos.p, int32_t(0), size_t(N), rocprim::plus<int32_t>(), s));
    DBuf<unsigned char> tmp;
    EMG_HIP(c, tmp.alloc(std::max<size_t>(sb, 1)));
    EMG_HIP(c, rocprim::exclusive_scan(tmp.p, sb, keep.p, pos.p, int32_t(0), size_t(N), rocprim::plus<int32_t>(), s));
    int32_t last_pos = 0, last_keep = 0;
    EMG_HIP(c, hipMemcpyAsync(&last_pos, pos.p + (N - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipMemcpyAsync(&last_keep, keep.p + (N - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EMG_HIP(c, hipStreamSynchronize(s));
    const int64_t nk = int64_t(last_pos) + last_keep;
    *n_keep = nk;
    if (keep_ids && nk > 0) {
        EMG_HIP(c, out.alloc(nk));
        k_keep_scatter<<<bn, 256, 0, s>>>(N, keep.p, pos.p, out.p);
        EMG_HIP(c, hipGetLastError());
        EMG_HIP(c, hipMemcpyAsync(keep_ids, out.p, sizeof(int32_t) * nk, hipMemcpyDeviceToHost, s));
        EMG_HIP(c, hipStreamSynchronize(s));
    }
    return EMG_OK;
}
