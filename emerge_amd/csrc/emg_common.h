// Internal definitions shared by the HIP translation units of libemerge_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/emerge_hip.h"

namespace emg {

// Owning device buffer (hipMalloc / hipFree), sized in elements of T.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        if (count == n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = count;
        return hipSuccess;
    }
};

// Work unit of the row-owner assembly: a run of consecutive CSR rows whose
// incidences fit one workgroup and whose CSR slots fit the LDS accumulator.
struct Chunk {
    int32_t inc0;   // first incidence (index into inc arrays)
    int32_t ninc;   // number of incidences (<= kAsmThreads)
    int32_t slot0;  // first CSR slot of the chunk (= indptr[row0])
    int32_t nslot;  // number of CSR slots (<= kAsmMaxSlots)
    int32_t nrank;  // max incidences per row in the chunk
    int32_t pad[3];
};

constexpr int kAsmThreads = 256;
// Contributor code (16 bits): bits 3-15 hold the LDS byte offset
// 8 * (v * kAsmThreads + lane) of element-row value v of incidence lane in
// the assembly kernels' stage (at most 8 * (20 * 256) - 8 = 40952), bits 0-1
// the flags below.
constexpr uint16_t kCodeLast = 0x0001;  // contributor code flag: last of its CSR slot
constexpr uint16_t kCodeCont = 0x0002;  // flag: group start that continues the previous group's slot
constexpr uint32_t kCodeOffset = 0xfff8u;  // mask of the stage byte offset
constexpr int kAsmMaxSlots = 4096;

}  // namespace emg

struct emg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;

    // mesh
    int64_t n_nodes = 0, n_tets = 0, n_field = 0;
    emg::DBuf<double> xyz;        // 3 x V SoA
    emg::DBuf<int32_t> tets;      // 4 x T
    emg::DBuf<int32_t> t2f;       // 20 x T

    // pattern
    bool have_pattern = false;
    int64_t nnz = 0;
    emg::DBuf<int32_t> indptr;    // N + 1
    emg::DBuf<int32_t> indices;   // nnz
    emg::DBuf<int32_t> cmap;      // 400 x T COO -> CSR slot
    // row-owner assembly schedule
    emg::DBuf<emg::Chunk> chunks;
    int64_t n_chunks = 0, n_inc = 0, max_rank = 0;
    emg::DBuf<int32_t> inc_tet;   // tet id per incidence (chunk-local order: by canonical row)
    emg::DBuf<uint8_t> inc_u;     // canonical element row of the incidence
    emg::DBuf<int32_t> inc_vid;   // the incidence's tet vertex ids, 4 per incidence (16-byte records)
    emg::DBuf<uint16_t> grp_slot; // per group of 20 contributor codes: chunk-relative slot of its first
    emg::DBuf<uint16_t> contrib;  // 20 per incidence: (lane << 5) | canonical column | flags, grouped by slot

    // materials and values
    bool have_material = false;
    emg::DBuf<double> eps;        // 2 x T (re, im)
    emg::DBuf<double> inu;        // 2 x T
    bool inu_real = false;        // every 1/mu_r has a zero imaginary part
    bool have_tensor = false;     // general tensors set (emg_material_set_tensor)
    emg::DBuf<double> eps9;       // complex (3,3,T): eps_r per component, t fastest
    emg::DBuf<double> imu9;       // complex (3,3,T): mu_r^-1
    emg::DBuf<double> Ev;         // 2 x nnz
    emg::DBuf<double> Bv;         // 2 x nnz
    bool have_values = false;

    // PEC-reduced system (rows/cols of eliminated DOFs removed)
    bool have_system = false;
    int64_t n_red = 0, nnz_red = 0;
    emg::DBuf<int32_t> keep;      // N: reduced index or -1
    emg::DBuf<int32_t> red_ptr, red_idx, red_map, red_diag;
    emg::DBuf<double> E64, B64;   // 2 x nnz_red
    emg::DBuf<float> E32, B32;    // 2 x nnz_red
    bool have_bsym = false;       // symmetric part (B + B^T)/2 of the reduced mass values built
    emg::DBuf<double> Bs64;       // 2 x nnz_red
    emg::DBuf<float> Bs32;        // 2 x nnz_red

    // boundary triangles (all BCs concatenated) and reduced surface terms
    int64_t n_tri_total = 0;
    int32_t n_bc = 0;
    emg::DBuf<double> tri_mat;    // 64 per triangle (no gamma)
    emg::DBuf<double2> tri_force; // 8 per triangle (zero when not excited)
    bool have_surface = false;
    int64_t n_sentries = 0;
    emg::DBuf<int32_t> s_ptr, s_col, s_bc;
    emg::DBuf<double> s_val64;
    emg::DBuf<float> s_val32;
    int32_t n_forcing = 0;
    emg::DBuf<double> forcing;    // 2 x n_forcing x n_red

    // port probes for S-parameter overlaps (mode-overlap and voltage-line ports)
    int32_t n_ports = 0;
    int64_t n_probe_pts = 0;
    std::vector<int64_t> port_pt_start;   // n_ports + 1
    emg::DBuf<int32_t> probe_tet;         // owning tet per probe point
    emg::DBuf<double> probe_pts;          // [n][3]
    emg::DBuf<double2> probe_shape;       // [n][3] complex vector c
    emg::DBuf<double> probe_wa;           // [n] weight
    std::vector<double> port_norm;        // sum w |c|^2 per port

    // mesh topology built on the device (emg_topology_build)
    int64_t topo_T = 0, topo_nE = 0, topo_nF = 0;
    emg::DBuf<double> topo_xyz, topo_len;
    emg::DBuf<uint64_t> topo_ekeys, topo_tkeys;
    emg::DBuf<int32_t> topo_edges, topo_tris, topo_t2e, topo_t2t, topo_r2e, topo_r2t;

    // uniform grid over the tets for point location (emg_field_points),
    // built on first use per mesh: cell c lists its tets in ascending order
    // at grid_tets[grid_ptr[c] .. grid_ptr[c + 1])
    bool grid_ready = false;
    double grid_lo[3] = {0, 0, 0};
    double grid_h = 1.0;
    int32_t grid_n[3] = {1, 1, 1};
    emg::DBuf<int64_t> grid_ptr;
    emg::DBuf<int32_t> grid_tets;

    // buffers of emg_field_points and emg_farfield, kept between calls (a
    // free would synchronise the device)
    emg::DBuf<double> fp_x, fp_pts, fp_E, fp_C;
    emg::DBuf<int32_t> fp_own;
    emg::DBuf<double> ff_E, ff_H, ff_pos, ff_wn, ff_th, ff_ph, ff_emag, ff_emax, ff_rec, ff_part, ff_Eo, ff_Ho;

    // Krylov method of the solves (emg_solver_config)
    int krylov_method = 0;        // 0 = GMRES(m), 1 = IDR(s), 2 = COCG on the symmetric part
    int spmv_gr = 0;              // COCG SpMV: entries of a row in flight (0 = default for the batch width)
    // row segments of the reduced system with Krylov scalars of their own
    // (emg_solver_set_segments; independent meshes side by side)
    int seg_nseg = 0;
    int64_t seg_n = 0;                  // reduced rows the segments were set for
    std::vector<int64_t> seg_row_h;     // nseg + 1 first rows
    emg::DBuf<int64_t> seg_row;
    emg::DBuf<int32_t> seg_rseg;        // segment of each row
    int idr_s = 4;
    double inner_rel32 = 1e-5;    // FP32 inner reduction per refinement step

    // auxiliary gradient space of the COCG preconditioner (emg_aux_set)
    int64_t n_escalated = 0;      // systems of the last emg_sweep_solve finished with FP64 inner solves
    bool escalate = true;         // continue stalled complex64 systems with complex128 inner solves
    bool safe32 = false;          // complex64 COCG with FP64 real stiffness and FP64 accumulation
    int er64_state = 0;           // 0 unknown, 1 reduced E is real (E64r holds it), 2 complex
    emg::DBuf<double> E64r;       // nnz_red: real part of the reduced E values
    int64_t aux_n = 0;            // nodal functions; 0 = Jacobi only
    int aux_nsurf = 0;
    emg::DBuf<int32_t> aux_gptr, aux_gidx, aux_tptr, aux_tidx;
    emg::DBuf<double> aux_gval64, aux_tval64;
    emg::DBuf<float> aux_gval32, aux_tval32;
    emg::DBuf<double> aux_dvol;   // 2 x aux_n
    emg::DBuf<double> aux_dsurf;  // aux_nsurf x aux_n
    bool aux_have_terms = false;  // emg_aux_set_terms: frequency-independent diagonal part and weights
    emg::DBuf<double> aux_dconst; // 2 x aux_n (complex)
    emg::DBuf<double> aux_wt;     // aux_n

    // generic CSR system handed in by the Solver plug-in
    int64_t g_n = 0, g_nnz = 0;
    emg::DBuf<int32_t> g_ptr, g_idx, g_diag;
    emg::DBuf<double> g_v64;
    emg::DBuf<float> g_v32;
    bool g_have_sym = false;      // symmetric part (A + A^T)/2 of the generic matrix built
    emg::DBuf<double> g_s64;
    emg::DBuf<float> g_s32;
    // auxiliary gradient space of the generic matrix (emg_csr_aux_set): G and
    // G^T, and D_g = diag(G^T A G) computed on the device from the matrix
    // batched element matrices (emg_element_matrices): [400][T] complex
    emg::DBuf<double> el_K, el_M;
    int64_t g_aux_n = 0;
    bool g_aux_fresh = false;     // D_g belongs to the current generic matrix
    emg::DBuf<int32_t> g_aux_gptr, g_aux_gidx, g_aux_tptr, g_aux_tidx;
    emg::DBuf<double> g_aux_gval64, g_aux_tval64, g_aux_dg;
    emg::DBuf<float> g_aux_gval32, g_aux_tval32;
};

namespace emg {

// Takes the device address of the no-allocation far field's scratch array
// for the current device, once per device (emg_farfield.hip).
void ff_scratch_setup(int device);

inline int fail(emg_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

}  // namespace emg

namespace emg {
// S-parameter port overlaps of the F solutions x ([row][f], reduced rows):
// out[f * n_ports + p] (defined in emg_sparam.hip)
int port_overlaps(emg_ctx* c, int F, const double2* x, std::vector<double2>& out);
}  // namespace emg

#define EMG_HIP(ctx, call)                                                      \
    do {                                                                        \
        hipError_t _e = (call);                                                 \
        if (_e != hipSuccess)                                                   \
            return emg::fail((ctx), _e == hipErrorOutOfMemory ? EMG_ERR_NOMEM   \
                                                               : EMG_ERR_HIP,  \
                             "%s failed: %s (%s:%d)", #call,                    \
                             hipGetErrorString(_e), __FILE__, __LINE__);        \
    } while (0)

#define EMG_CHECK_CTX(ctx)                                                      \
    do {                                                                        \
        if (!(ctx)) return EMG_ERR_ARG;                                         \
        (void)hipSetDevice((ctx)->device);                                      \
    } while (0)
