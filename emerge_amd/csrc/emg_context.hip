// Context lifecycle, error reporting, timing and host<->device transfers.
#include "emg_common.h"

#include <cstring>
#include <new>

extern "C" {

const char* emg_version(void) { return "0.1.0"; }

int emg_create(int device, emg_ctx** out) {
    if (!out) return EMG_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EMG_ERR_HIP;
    if (device < 0 || device >= ndev) return EMG_ERR_ARG;
    emg_ctx* c = new (std::nothrow) emg_ctx();
    if (!c) return EMG_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return EMG_ERR_HIP;
    }
    emg::ff_scratch_setup(device);
    *out = c;
    return EMG_OK;
}

int emg_destroy(emg_ctx* c) {
    if (!c) return EMG_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;  // DBuf members free their device memory
    return EMG_OK;
}

int emg_last_error(const emg_ctx* c, char* buf, size_t len) {
    if (!c || !buf || len == 0) return EMG_ERR_ARG;
    std::strncpy(buf, c->err.c_str(), len - 1);
    buf[len - 1] = '\0';
    return EMG_OK;
}

int emg_timer_start(emg_ctx* c) {
    EMG_CHECK_CTX(c);
    EMG_HIP(c, hipEventRecord(c->ev0, c->stream));
    return EMG_OK;
}

int emg_timer_stop(emg_ctx* c, float* ms) {
    EMG_CHECK_CTX(c);
    if (!ms) return emg::fail(c, EMG_ERR_ARG, "emg_timer_stop: null output");
    EMG_HIP(c, hipEventRecord(c->ev1, c->stream));
    EMG_HIP(c, hipEventSynchronize(c->ev1));
    EMG_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    return EMG_OK;
}

int emg_synchronize(emg_ctx* c) {
    EMG_CHECK_CTX(c);
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

int emg_mesh_set(emg_ctx* c, int64_t n_nodes, const double* xyz, int64_t n_tets,
                 const int32_t* tets, const int32_t* t2f, int64_t n_field) {
    EMG_CHECK_CTX(c);
    if (!xyz || !tets || !t2f || n_nodes < 4 || n_tets < 1 || n_field < 1)
        return emg::fail(c, EMG_ERR_ARG, "emg_mesh_set: empty or null mesh arrays");
    if (n_tets * 400 >= (int64_t(1) << 31) || n_field >= (int64_t(1) << 31) || n_nodes >= (int64_t(1) << 31))
        return emg::fail(c, EMG_ERR_ARG, "emg_mesh_set: mesh too large for int32 indexing (T=%lld, N=%lld)",
                         (long long)n_tets, (long long)n_field);
    for (int64_t k = 0; k < 4 * n_tets; ++k)
        if (tets[k] < 0 || tets[k] >= n_nodes)
            return emg::fail(c, EMG_ERR_ARG, "emg_mesh_set: tets[%lld]=%d out of range [0,%lld)", (long long)k,
                             tets[k], (long long)n_nodes);
    for (int64_t k = 0; k < 20 * n_tets; ++k)
        if (t2f[k] < 0 || t2f[k] >= n_field)
            return emg::fail(c, EMG_ERR_ARG, "emg_mesh_set: tet_to_field[%lld]=%d out of range [0,%lld)",
                             (long long)k, t2f[k], (long long)n_field);
    EMG_HIP(c, c->xyz.alloc(3 * n_nodes));
    EMG_HIP(c, c->tets.alloc(4 * n_tets));
    EMG_HIP(c, c->t2f.alloc(20 * n_tets));
    EMG_HIP(c, hipMemcpyAsync(c->xyz.p, xyz, sizeof(double) * 3 * n_nodes, hipMemcpyHostToDevice, c->stream));
    EMG_HIP(c, hipMemcpyAsync(c->tets.p, tets, sizeof(int32_t) * 4 * n_tets, hipMemcpyHostToDevice, c->stream));
    EMG_HIP(c, hipMemcpyAsync(c->t2f.p, t2f, sizeof(int32_t) * 20 * n_tets, hipMemcpyHostToDevice, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    c->n_nodes = n_nodes;
    c->n_tets = n_tets;
    c->n_field = n_field;
    c->have_pattern = false;
    c->have_material = false;
    c->have_values = false;
    c->have_system = false;
    c->grid_ready = false;
    return EMG_OK;
}

int emg_material_set(emg_ctx* c, const double* eps, const double* inu) {
    EMG_CHECK_CTX(c);
    if (c->n_tets == 0) return emg::fail(c, EMG_ERR_STATE, "emg_material_set: no mesh");
    if (!eps || !inu) return emg::fail(c, EMG_ERR_ARG, "emg_material_set: null material arrays");
    EMG_HIP(c, c->eps.alloc(2 * c->n_tets));
    EMG_HIP(c, c->inu.alloc(2 * c->n_tets));
    EMG_HIP(c, hipMemcpyAsync(c->eps.p, eps, sizeof(double) * 2 * c->n_tets, hipMemcpyHostToDevice, c->stream));
    EMG_HIP(c, hipMemcpyAsync(c->inu.p, inu, sizeof(double) * 2 * c->n_tets, hipMemcpyHostToDevice, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    // all 1/mu_r real: the curl-curl pass of the assembly sums real values
    bool real = true;
    for (int64_t t = 0; t < c->n_tets && real; ++t) real = inu[2 * t + 1] == 0.0;
    c->inu_real = real;
    c->have_material = true;
    c->have_tensor = false;
    return EMG_OK;
}

int emg_pattern_get(emg_ctx* c, int32_t* indptr, int32_t* indices) {
    EMG_CHECK_CTX(c);
    if (!c->have_pattern) return emg::fail(c, EMG_ERR_STATE, "emg_pattern_get: pattern not built");
    if (!indptr || !indices) return emg::fail(c, EMG_ERR_ARG, "emg_pattern_get: null output");
    EMG_HIP(c, hipMemcpyAsync(indptr, c->indptr.p, sizeof(int32_t) * (c->n_field + 1), hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipMemcpyAsync(indices, c->indices.p, sizeof(int32_t) * c->nnz, hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

int emg_scatter_map_get(emg_ctx* c, int32_t* map) {
    EMG_CHECK_CTX(c);
    if (!c->have_pattern) return emg::fail(c, EMG_ERR_STATE, "emg_scatter_map_get: pattern not built");
    if (!map) return emg::fail(c, EMG_ERR_ARG, "emg_scatter_map_get: null output");
    EMG_HIP(c, hipMemcpyAsync(map, c->cmap.p, sizeof(int32_t) * 400 * c->n_tets, hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

int emg_pattern_stats(emg_ctx* c, int64_t* n_chunks, int64_t* n_inc, int64_t* max_rank) {
    EMG_CHECK_CTX(c);
    if (!c->have_pattern) return emg::fail(c, EMG_ERR_STATE, "emg_pattern_stats: pattern not built");
    if (n_chunks) *n_chunks = c->n_chunks;
    if (n_inc) *n_inc = c->n_inc;
    if (max_rank) *max_rank = c->max_rank;
    return EMG_OK;
}

int emg_values_get(emg_ctx* c, double* E, double* B) {
    EMG_CHECK_CTX(c);
    if (!c->have_values) return emg::fail(c, EMG_ERR_STATE, "emg_values_get: nothing assembled");
    if (!E || !B) return emg::fail(c, EMG_ERR_ARG, "emg_values_get: null output");
    EMG_HIP(c, hipMemcpyAsync(E, c->Ev.p, sizeof(double) * 2 * c->nnz, hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipMemcpyAsync(B, c->Bv.p, sizeof(double) * 2 * c->nnz, hipMemcpyDeviceToHost, c->stream));
    EMG_HIP(c, hipStreamSynchronize(c->stream));
    return EMG_OK;
}

}  // extern "C"
