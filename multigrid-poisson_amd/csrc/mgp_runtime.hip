// Device, memory, stream, graph and event entry points of the mgp library.
// They replace the reference's OpenMP device-memory calls
// (include/devicearray.h:60, 70-76, 106-113, 115-130: omp_target_alloc/free/memcpy)
// and add what its host-blocking launches never needed: streams, graph
// capture of whole cycles, and events for timing.
#include <hip/hip_runtime.h>

#include <cstring>

#include "mgp.h"
#include "mgp_sweep.h"

static inline int st(hipError_t e) { return e == hipSuccess ? MGP_OK : (int)e; }

extern "C" {

int mgp_device_count(int* count) {
    if (!count) return MGP_ERR_ARG;
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return (int)e;
    }
    *count = n;
    return n > 0 ? MGP_OK : MGP_ERR_NODEV;
}

int mgp_set_device(int device) { return st(hipSetDevice(device)); }

int mgp_malloc(void** ptr, size_t bytes) {
    if (!ptr) return MGP_ERR_ARG;
    *ptr = nullptr;
    return st(hipMalloc(ptr, bytes ? bytes : 1));
}

int mgp_free(void* ptr) { return ptr ? st(hipFree(ptr)) : MGP_OK; }

int mgp_host_alloc(void** ptr, size_t bytes) {
    if (!ptr) return MGP_ERR_ARG;
    *ptr = nullptr;
    return st(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
}

int mgp_host_free(void* ptr) { return ptr ? st(hipHostFree(ptr)) : MGP_OK; }

int mgp_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return st(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
}

int mgp_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return st(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
}

int mgp_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    return st(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
}

int mgp_stream_create(void** stream) {
    if (!stream) return MGP_ERR_ARG;
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *stream = s;
    return st(e);
}

int mgp_stream_destroy(void* stream) { return stream ? st(hipStreamDestroy((hipStream_t)stream)) : MGP_OK; }

int mgp_stream_sync(void* stream) { return st(hipStreamSynchronize((hipStream_t)stream)); }

int mgp_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return MGP_ERR_ARG;
    return st(hipMemGetInfo(free_bytes, total_bytes));
}

int mgp_graph_begin(void* stream) {
    if (!stream) return MGP_ERR_ARG;  // the legacy default stream cannot be captured
    return st(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
}

int mgp_graph_end(void* stream, void** exec) {
    if (!stream || !exec) return MGP_ERR_ARG;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
    if (e != hipSuccess) return (int)e;
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    *exec = x;
    return st(e);
}

int mgp_graph_launch(void* exec, void* stream) {
    if (!exec) return MGP_ERR_ARG;
    return st(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream));
}

int mgp_graph_destroy(void* exec) { return exec ? st(hipGraphExecDestroy((hipGraphExec_t)exec)) : MGP_OK; }

int mgp_event_create(void** event) {
    if (!event) return MGP_ERR_ARG;
    hipEvent_t ev = nullptr;
    const hipError_t e = hipEventCreate(&ev);
    *event = ev;
    return st(e);
}

int mgp_event_destroy(void* event) { return event ? st(hipEventDestroy((hipEvent_t)event)) : MGP_OK; }

int mgp_event_record(void* event, void* stream) { return st(hipEventRecord((hipEvent_t)event, (hipStream_t)stream)); }

int mgp_event_sync(void* event) { return st(hipEventSynchronize((hipEvent_t)event)); }

int mgp_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!ms) return MGP_ERR_ARG;
    return st(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
}

int mgp_tuning_set(const char* key, int64_t value) {
    if (!key || value < 0) return MGP_ERR_ARG;
    if (!strcmp(key, "sweep.target_wgs") && value >= 1) mgp::g_sweep_target_wgs = value;
    else if (!strcmp(key, "sweep.min_chunk") && value >= 1) mgp::g_sweep_min_chunk = value;
    else if (!strcmp(key, "sweep2.chunks_per_tile") && value >= 1 && value <= 1024) mgp::g_sweep2_cpt = value;
    else if (!strcmp(key, "residual.target_wgs") && value <= (1 << 20)) mgp::g_resid_target = value;
    else if (!strcmp(key, "residual_fw.target_wgs") && value <= (1 << 20)) mgp::g_fw_target = value;
    else if (!strcmp(key, "prolong.target_wgs") && value <= (1 << 20)) mgp::g_prolong_target = value;
    else if (!strcmp(key, "update_residual.target_wgs") && value <= (1 << 20)) mgp::g_upd_target = value;
    else if (!strcmp(key, "update_residual.rows") && (value == 8 || value == 16)) mgp::g_hilo_rows = value;
    else if (!strcmp(key, "sweep.small_points")) mgp::g_small_points = value;
    else if (!strcmp(key, "sweep.small_min_chunk") && value >= 1) mgp::g_small_min_chunk = value;
    else if (!strcmp(key, "sweep.pair_min_points")) mgp::g_pair_min_points = value;
    else if (!strcmp(key, "sweep.waves") && (value == 4 || value == 8 || value == 16)) mgp::g_sweep_waves = value;
    else if (!strcmp(key, "sweep.kernel") && (value == 1 || value == 2)) mgp::g_sweep_kernel = value;
    else if (!strcmp(key, "sweep.rows") && (value == 0 || value == 8 || value == 12 || value == 16))
        mgp::g_sweep_rows = value;
    else if (!strcmp(key, "sweep.depth") && (value == 1 || value == 2)) mgp::g_sweep_depth = value;
    else if (!strcmp(key, "sweep.pair") && (value == 0 || value == 1)) mgp::g_sweep_pair = value;
    else if (!strcmp(key, "sweep2.rows") &&
             (value == 0 || value == 12 || value == 16 || value == 20 || value == 24))
        mgp::g_sweep2_rows = value;
    else if (!strcmp(key, "sweep2.prolong") && value <= 2) mgp::g_sweep2_prolong = value;
    else if (!strcmp(key, "sweep2.pack") && value <= 3) mgp::g_sweep2_pack = value;
    else if (!strcmp(key, "sweep2.target_waves") && value >= 1 && value <= (1 << 20)) mgp::g_sweep2_waves = value;
    else if (!strcmp(key, "residual_fw.rows") && (value == 0 || value == 4 || value == 6)) mgp::g_fw_rows = value;
    else if (!strcmp(key, "residual_fw.depth") && (value == 0 || value == 1 || value == 2)) mgp::g_fw_depth = value;
    else if (!strcmp(key, "batch.blocks") && value >= 1 && value <= (1 << 20)) mgp::g_batch_blocks = value;
    else if (!strcmp(key, "residual.kernel") && (value == 1 || value == 2)) mgp::g_resid_kernel = value;
    else if (!strcmp(key, "residual.rows") && (value == 8 || value == 16)) mgp::g_resid_rows = value;
    else if (!strcmp(key, "prolong.kernel") && (value == 1 || value == 2)) mgp::g_prolong_kernel = value;
    else return MGP_ERR_ARG;
    return MGP_OK;
}

int mgp_tuning_get(const char* key, int64_t* value) {
    if (!key || !value) return MGP_ERR_ARG;
    if (!strcmp(key, "sweep.target_wgs")) *value = mgp::g_sweep_target_wgs;
    else if (!strcmp(key, "sweep.min_chunk")) *value = mgp::g_sweep_min_chunk;
    else if (!strcmp(key, "sweep2.chunks_per_tile")) *value = mgp::g_sweep2_cpt;
    else if (!strcmp(key, "residual.target_wgs")) *value = mgp::g_resid_target;
    else if (!strcmp(key, "residual_fw.target_wgs")) *value = mgp::g_fw_target;
    else if (!strcmp(key, "prolong.target_wgs")) *value = mgp::g_prolong_target;
    else if (!strcmp(key, "update_residual.target_wgs")) *value = mgp::g_upd_target;
    else if (!strcmp(key, "update_residual.rows")) *value = mgp::g_hilo_rows;
    else if (!strcmp(key, "sweep.small_points")) *value = mgp::g_small_points;
    else if (!strcmp(key, "sweep.small_min_chunk")) *value = mgp::g_small_min_chunk;
    else if (!strcmp(key, "sweep.pair_min_points")) *value = mgp::g_pair_min_points;
    else if (!strcmp(key, "sweep.waves")) *value = mgp::g_sweep_waves;
    else if (!strcmp(key, "sweep.kernel")) *value = mgp::g_sweep_kernel;
    else if (!strcmp(key, "sweep.rows")) *value = mgp::g_sweep_rows;
    else if (!strcmp(key, "sweep.depth")) *value = mgp::g_sweep_depth;
    else if (!strcmp(key, "sweep.pair")) *value = mgp::g_sweep_pair;
    else if (!strcmp(key, "sweep2.rows")) *value = mgp::g_sweep2_rows;
    else if (!strcmp(key, "sweep2.prolong")) *value = mgp::g_sweep2_prolong;
    else if (!strcmp(key, "sweep2.pack")) *value = mgp::g_sweep2_pack;
    else if (!strcmp(key, "sweep2.target_waves")) *value = mgp::g_sweep2_waves;
    else if (!strcmp(key, "residual_fw.rows")) *value = mgp::g_fw_rows;
    else if (!strcmp(key, "residual_fw.depth")) *value = mgp::g_fw_depth;
    else if (!strcmp(key, "batch.blocks")) *value = mgp::g_batch_blocks;
    else if (!strcmp(key, "residual.kernel")) *value = mgp::g_resid_kernel;
    else if (!strcmp(key, "residual.rows")) *value = mgp::g_resid_rows;
    else if (!strcmp(key, "prolong.kernel")) *value = mgp::g_prolong_kernel;
    else return MGP_ERR_ARG;
    return MGP_OK;
}

const char* mgp_error_string(int code) {
    switch (code) {
    case MGP_OK: return "success";
    case MGP_ERR_ARG: return "invalid argument";
    case MGP_ERR_NODEV: return "no GPU device";
    case MGP_ERR_STATE: return "call out of order";
    default: return hipGetErrorString((hipError_t)code);
    }
}

const char* mgp_version(void) { return "mgp 0.1 (gfx950)"; }

}  // extern "C"
