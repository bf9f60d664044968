// Batched dense element matrices: the 20x20 curl-curl (K) and mass (M)
// matrices of every tet, as ned2_tet_stiff_mass computes them one tet at a
// time (ASM/curlcurl.py:180-412), for isotropic complex eps_r and 1/mu_r.
//
// In the canonical frame (vertices by global-id rank) both are fixed linear
// forms of the tet's Gram vectors (emg_ned2.h): for the 400 entries
// e = 20 u + v,
//   K_raw[e][t] = sum_k TK[e][k] h_k(t)   (21 cross Gram products, padded to 24)
//   M_raw[e][t] = sum_k TM[e][k] g_k(t)   (10 Gram products, padded to 12)
// then K = s_u s_v mu^-1 K_raw / (120 (6V)^3), M = s_u s_v eps M_raw / (5040 (6V)).
// Over a batch of tets these are two dense products (400 x 24)(24 x T) and
// (400 x 12)(12 x T).  Two kernels compute them:
//
// * k_element_scalar: one tet per lane, the products unrolled at compile time
//   over the nonzero table entries only (about 5 of 24 per K entry);
// * k_element_mfma: FP64 matrix cores (v_mfma_f64_16x16x4f64), 64 tets per
//   wavefront in four 16-tet tiles, the dense tables as the A operand (16
//   entries x 4 products per instruction), the tets' Gram vectors from LDS
//   as the B operand: 150 + 75 instructions per tile.
//
// Output layout: entry-major, K[e][t] and M[e][t] complex128 (consecutive
// tets of one entry are contiguous, so both kernels store with coalesced
// 16-byte lanes), canonical DOF order; the host maps canonical DOF u to the
// reference's local DOF order per tet.  12.8 KB are written per tet, so both
// kernels are bound by HBM writes, not by the FP64 pipes.
#include "emg_common.h"
#include "emg_ned2.h"

namespace {

using namespace emg;

constexpr int kElTets = 64;   // tets per workgroup (one wavefront)
constexpr int kHK = 24;       // cross Gram products padded to a multiple of 4
constexpr int kGK = 12;       // Gram products padded to a multiple of 4

struct DenseTables {
    double tk[400][kHK];
    double tm[400][kGK];
    constexpr DenseTables() : tk{}, tm{} {
        for (int u = 0; u < 20; ++u)
            for (int v = 0; v < 20; ++v) {
                for (int k = 0; k < 21; ++k) tk[20 * u + v][k] = double(ned2::tk(u, v, k));
                for (int k = 0; k < 10; ++k) tm[20 * u + v][k] = double(ned2::tm(u, v, k));
            }
    }
};
constexpr DenseTables kTablesInit{};           // evaluated at compile time
__device__ const DenseTables kTables = kTablesInit;

typedef double v4d __attribute__((ext_vector_type(4)));

// Per-tet canonical geometry of tet t (the assembly's own routine).
__device__ __forceinline__ void tet_canonical(const double* __restrict__ xyz, int64_t V,
                                              const int32_t* __restrict__ tets, int64_t T, int64_t t,
                                              ned2::TetGeom& G) {
    int w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = tets[k * T + t];
    int rk[4];
    ned2::ranks4(w, rk);
    double X[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        X[k][0] = xyz[w[k]];
        X[k][1] = xyz[V + w[k]];
        X[k][2] = xyz[2 * V + w[k]];
    }
    ned2::tet_geometry(X, rk, G);
}

template <int E>
__device__ __forceinline__ double entry_k(const ned2::TetGeom& G) {
    constexpr int U = E / 20, V = E % 20;
    return G.d[ned2::len_index(U)] * G.d[ned2::len_index(V)] *
           ned2::contract_k<U, V>(G.h, std::make_integer_sequence<int, 21>{});
}
template <int E>
__device__ __forceinline__ double entry_m(const ned2::TetGeom& G) {
    constexpr int U = E / 20, V = E % 20;
    return G.d[ned2::len_index(U)] * G.d[ned2::len_index(V)] *
           ned2::contract_m<U, V>(G.g, std::make_integer_sequence<int, 10>{});
}
// entries OFF .. OFF + sizeof...(E) - 1 of K (or M) of tet t
template <int OFF, bool MASS, int... E>
__device__ __forceinline__ void store_entries(const ned2::TetGeom& G, double2 f, double2* __restrict__ out, int64_t T,
                                              int64_t t, std::integer_sequence<int, E...>) {
    ((out[int64_t(E + OFF) * T + t] = [&] {
         const double r = MASS ? entry_m<E + OFF>(G) : entry_k<E + OFF>(G);
         return make_double2(f.x * r, f.y * r);
     }()),
     ...);
}

__global__ __launch_bounds__(kElTets) void k_element_scalar(int64_t T, const double* __restrict__ xyz, int64_t V,
                                                            const int32_t* __restrict__ tets,
                                                            const double2* __restrict__ eps,
                                                            const double2* __restrict__ inu, double2* __restrict__ K,
                                                            double2* __restrict__ M) {
    const int64_t t = int64_t(blockIdx.x) * kElTets + threadIdx.x;
    if (t >= T) return;
    ned2::TetGeom G;
    tet_canonical(xyz, V, tets, T, t, G);
    const double i3 = G.inv6v * G.inv6v * G.inv6v / 120.0;
    const double i1 = G.inv6v / 5040.0;
    const double2 mu = inu[t], ep = eps[t];
    const double2 fk = make_double2(mu.x * i3, mu.y * i3), fm = make_double2(ep.x * i1, ep.y * i1);
    store_entries<0, false>(G, fk, K, T, t, std::make_integer_sequence<int, 200>{});
    store_entries<200, false>(G, fk, K, T, t, std::make_integer_sequence<int, 200>{});
    store_entries<0, true>(G, fm, M, T, t, std::make_integer_sequence<int, 200>{});
    store_entries<200, true>(G, fm, M, T, t, std::make_integer_sequence<int, 200>{});
}

__global__ __launch_bounds__(kElTets) void k_element_mfma(int64_t T, const double* __restrict__ xyz, int64_t V,
                                                          const int32_t* __restrict__ tets,
                                                          const double2* __restrict__ eps,
                                                          const double2* __restrict__ inu, double2* __restrict__ K,
                                                          double2* __restrict__ M) {
    __shared__ double H[kHK][kElTets];
    __shared__ double Gs[kGK][kElTets];
    __shared__ double Ds[6][kElTets];
    __shared__ double2 FK[kElTets], FM[kElTets];
    const int l = threadIdx.x;
    const int64_t tb = int64_t(blockIdx.x) * kElTets;
    const int64_t t = tb + l;
    {
        ned2::TetGeom G;
        double2 fk = make_double2(0, 0), fm = make_double2(0, 0);
        if (t < T) {
            tet_canonical(xyz, V, tets, T, t, G);
            const double i3 = G.inv6v * G.inv6v * G.inv6v / 120.0;
            const double i1 = G.inv6v / 5040.0;
            const double2 mu = inu[t], ep = eps[t];
            fk = make_double2(mu.x * i3, mu.y * i3);
            fm = make_double2(ep.x * i1, ep.y * i1);
        } else {
#pragma unroll
            for (int k = 0; k < 21; ++k) G.h[k] = 0.0;
#pragma unroll
            for (int k = 0; k < 10; ++k) G.g[k] = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) G.d[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < kHK; ++k) H[k][l] = k < 21 ? G.h[k] : 0.0;
#pragma unroll
        for (int k = 0; k < kGK; ++k) Gs[k][l] = k < 10 ? G.g[k] : 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) Ds[k][l] = G.d[k];
        FK[l] = fk;
        FM[l] = fm;
    }
    __syncthreads();
    // lane layout of v_mfma_f64_16x16x4f64 (checked on the device): A[i][k]
    // with i = l % 16, k = l / 16; B[k][j] with k = l / 16, j = l % 16;
    // accumulator r of lane l is D[l / 16 + 4 r][l % 16], r = 0..3
    const int li = l % 16, lk = l / 16;
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        const int nks = pass == 0 ? kHK / 4 : kGK / 4;
        double2* out = pass == 0 ? K : M;
#pragma unroll 1
        for (int rt = 0; rt < 25; ++rt) {
            v4d acc[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < nks; ++ks) {
                const int e = rt * 16 + li, k = ks * 4 + lk;
                const double a = pass == 0 ? kTables.tk[e][k] : kTables.tm[e][k];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double bv = pass == 0 ? H[k][q * 16 + li] : Gs[k][q * 16 + li];
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[q], 0, 0, 0);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = q * 16 + li;
                const int64_t tt = tb + j;
                if (tt >= T) continue;
                const double2 f = pass == 0 ? FK[j] : FM[j];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = rt * 16 + lk + 4 * r;
                    const int u = e / 20, v = e % 20;
                    const double s = Ds[ned2::len_index(u)][j] * Ds[ned2::len_index(v)][j] * acc[q][r];
                    out[int64_t(e) * T + tt] = make_double2(f.x * s, f.y * s);
                }
            }
        }
    }
}

}  // namespace

extern "C" int emg_element_matrices(emg_ctx* c, int use_mfma, double* K, double* M) {
    EMG_CHECK_CTX(c);
    if (c->n_tets == 0) return fail(c, EMG_ERR_STATE, "emg_element_matrices: no mesh");
    if (!c->have_material || c->have_tensor)
        return fail(c, EMG_ERR_STATE, "emg_element_matrices: needs isotropic materials (emg_material_set)");
    if (use_mfma != 0 && use_mfma != 1) return fail(c, EMG_ERR_ARG, "emg_element_matrices: use_mfma must be 0 or 1");
    const int64_t T = c->n_tets;
    EMG_HIP(c, c->el_K.alloc(2 * 400 * T));
    EMG_HIP(c, c->el_M.alloc(2 * 400 * T));
    const unsigned grid = unsigned((T + kElTets - 1) / kElTets);
    auto* Kd = reinterpret_cast<double2*>(c->el_K.p);
    auto* Md = reinterpret_cast<double2*>(c->el_M.p);
    const auto* ep = reinterpret_cast<const double2*>(c->eps.p);
    const auto* mu = reinterpret_cast<const double2*>(c->inu.p);
    if (use_mfma)
        k_element_mfma<<<grid, kElTets, 0, c->stream>>>(T, c->xyz.p, c->n_nodes, c->tets.p, ep, mu, Kd, Md);
    else
        k_element_scalar<<<grid, kElTets, 0, c->stream>>>(T, c->xyz.p, c->n_nodes, c->tets.p, ep, mu, Kd, Md);
    EMG_HIP(c, hipGetLastError());
    if (K) EMG_HIP(c, hipMemcpyAsync(K, c->el_K.p, sizeof(double) * 2 * 400 * T, hipMemcpyDeviceToHost, c->stream));
    if (M) EMG_HIP(c, hipMemcpyAsync(M, c->el_M.p, sizeof(double) * 2 * 400 * T, hipMemcpyDeviceToHost, c->stream));
    if (K || M) EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}
