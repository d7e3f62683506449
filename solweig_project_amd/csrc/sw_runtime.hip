// libsolweig_hip runtime entry points: device memory with allocation
// counters, pinned host memory, copies, streams, events, stream capture into
// executable graphs, and the composite SVF-sweep / day-loop entry points that
// enqueue a whole stage with one call.
//
// The day loop issues, per step, the sun march (daytime steps only) and the
// flux kernel on the caller's stream; it contains no host synchronisation,
// allocation or memset, so it can be captured with sw_graph_begin/_end and
// replayed.
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "sw_internal.h"

namespace {

std::atomic<int64_t> g_live_bytes{0};
std::atomic<int64_t> g_n_alloc{0};
std::atomic<int64_t> g_n_free{0};

// sizes of live blocks, so sw_free keeps the live-byte counter exact
std::mutex g_mu;
std::unordered_map<void *, int64_t> g_sizes;

__global__ void k_fill_f32(float *__restrict__ p, float v, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

extern "C" {

int sw_device_name(int device, char *buf, int len) {
    if (!buf || len <= 0) return swi_set_err(SW_E_ARG, "sw_device_name: empty buffer");
    hipDeviceProp_t p;
    int rc = swi_hip_check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    if (rc) return rc;
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return SW_OK;
}

int sw_mem_info(int64_t *free_bytes, int64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return swi_set_err(SW_E_ARG, "sw_mem_info: NULL pointer");
    size_t f = 0, t = 0;
    int rc = swi_hip_check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return rc;
}

int sw_alloc(void **ptr, int64_t bytes) {
    if (!ptr || bytes < 0) return swi_set_err(SW_E_ARG, "sw_alloc: invalid arguments");
    *ptr = nullptr;
    void *p = nullptr;
    int rc = swi_hip_check(hipMalloc(&p, (size_t)(bytes > 0 ? bytes : 1)), "hipMalloc");
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_sizes[p] = bytes;
    }
    g_live_bytes += bytes;
    g_n_alloc += 1;
    *ptr = p;
    return SW_OK;
}

int sw_free(void *ptr) {
    if (!ptr) return SW_OK;
    int64_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_sizes.find(ptr);
        if (it == g_sizes.end()) return swi_set_err(SW_E_ARG, "sw_free: pointer was not allocated by sw_alloc");
        bytes = it->second;
        g_sizes.erase(it);
    }
    int rc = swi_hip_check(hipFree(ptr), "hipFree");
    if (rc) return rc;
    g_live_bytes -= bytes;
    g_n_free += 1;
    return SW_OK;
}

int sw_alloc_stats(int64_t *live_bytes, int64_t *n_alloc, int64_t *n_free) {
    if (!live_bytes || !n_alloc || !n_free) return swi_set_err(SW_E_ARG, "sw_alloc_stats: NULL pointer");
    *live_bytes = g_live_bytes.load();
    *n_alloc = g_n_alloc.load();
    *n_free = g_n_free.load();
    return SW_OK;
}

int sw_host_alloc(void **ptr, int64_t bytes) {
    if (!ptr || bytes < 0) return swi_set_err(SW_E_ARG, "sw_host_alloc: invalid arguments");
    *ptr = nullptr;
    return swi_hip_check(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault), "hipHostMalloc");
}

int sw_host_free(void *ptr) {
    if (!ptr) return SW_OK;
    return swi_hip_check(hipHostFree(ptr), "hipHostFree");
}

int sw_copy(void *dst, const void *src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return swi_set_err(SW_E_ARG, "sw_copy: invalid arguments");
    if (bytes == 0) return SW_OK;
    return swi_hip_check(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault), "hipMemcpy");
}

int sw_copy_async(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src)))
        return swi_set_err(SW_E_ARG, "sw_copy_async: invalid arguments");
    if (bytes == 0) return SW_OK;
    return swi_hip_check(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, swi_stream(stream)),
                         "hipMemcpyAsync");
}

int sw_copy2d_async(void *dst, int64_t dpitch, const void *src, int64_t spitch, int64_t width_bytes,
                    int64_t height, void *stream) {
    if (!dst || !src || width_bytes < 0 || height < 0 || dpitch < width_bytes || spitch < width_bytes)
        return swi_set_err(SW_E_ARG, "sw_copy2d_async: invalid arguments");
    if (width_bytes == 0 || height == 0) return SW_OK;
    return swi_hip_check(hipMemcpy2DAsync(dst, (size_t)dpitch, src, (size_t)spitch, (size_t)width_bytes,
                                          (size_t)height, hipMemcpyDefault, swi_stream(stream)),
                         "hipMemcpy2DAsync");
}

int sw_copy_to_host(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return swi_set_err(SW_E_ARG, "sw_copy_to_host: invalid arguments");
    if (bytes == 0) return SW_OK;
    // "device to device without compute units": pinned host memory is device-addressable, so this
    // runs on the DMA engines and leaves every CU to the solves on other streams
    return swi_hip_check(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDeviceNoCU, swi_stream(stream)),
                         "sw_copy_to_host: hipMemcpyAsync");
}

// Upload from pinned host memory by the compute units: each thread reads 16-byte words of the host
// buffer over PCIe (device-addressable pinned memory) and writes them to HBM; the last bytes of a
// size that is not a multiple of 16 are copied by the first threads one byte each.
__global__ void k_copy_from_host(uint4 *__restrict__ dst, const uint4 *__restrict__ src, long n16,
                                 unsigned char *__restrict__ dtail, const unsigned char *__restrict__ stail,
                                 int ntail) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntail) dtail[t] = stail[t];
}

int sw_copy_to_device_cu(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src)))
        return swi_set_err(SW_E_ARG, "sw_copy_to_device_cu: invalid arguments");
    if (bytes == 0) return SW_OK;
    if ((((uintptr_t)dst) & 15u) || (((uintptr_t)src) & 15u))
        return swi_set_err(SW_E_ARG, "sw_copy_to_device_cu: source and destination must be 16-byte aligned");
    // the kernel dereferences src on the device: only pinned (device-addressable) host memory qualifies
    hipPointerAttribute_t pa;
    if (hipPointerGetAttributes(&pa, src) != hipSuccess || pa.type != hipMemoryTypeHost) {
        (void)hipGetLastError();
        return swi_set_err(SW_E_ARG, "sw_copy_to_device_cu: the source is not pinned host memory");
    }
    const long n16 = (long)(bytes / 16);
    const int ntail = (int)(bytes - n16 * 16);
    long blocks = (n16 + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy_from_host, dim3((unsigned)blocks), dim3(256), 0, swi_stream(stream), (uint4 *)dst,
                       (const uint4 *)src, n16, (unsigned char *)dst + n16 * 16,
                       (const unsigned char *)src + n16 * 16, ntail);
    return swi_hip_check(hipGetLastError(), "k_copy_from_host");
}

int sw_fill_f32(float *dst, float value, int64_t n, void *stream) {
    if (n < 0 || (n > 0 && !dst)) return swi_set_err(SW_E_ARG, "sw_fill_f32: invalid arguments");
    if (n == 0) return SW_OK;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, swi_stream(stream), dst, value,
                       (long)n);
    return swi_hip_check(hipGetLastError(), "k_fill_f32");
}

int sw_stream_create(void **stream) {
    if (!stream) return swi_set_err(SW_E_ARG, "sw_stream_create: NULL pointer");
    hipStream_t s = nullptr;
    int rc = swi_hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    *stream = s;
    return rc;
}

int sw_stream_destroy(void *stream) {
    if (!stream) return SW_OK;
    return swi_hip_check(hipStreamDestroy(swi_stream(stream)), "hipStreamDestroy");
}

int sw_event_create(void **event) {
    if (!event) return swi_set_err(SW_E_ARG, "sw_event_create: NULL pointer");
    hipEvent_t e = nullptr;
    int rc = swi_hip_check(hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return rc;
}

int sw_event_destroy(void *event) {
    if (!event) return SW_OK;
    return swi_hip_check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}

int sw_event_record(void *event, void *stream) {
    if (!event) return swi_set_err(SW_E_ARG, "sw_event_record: NULL event");
    return swi_hip_check(hipEventRecord((hipEvent_t)event, swi_stream(stream)), "hipEventRecord");
}

int sw_event_sync(void *event) {
    if (!event) return swi_set_err(SW_E_ARG, "sw_event_sync: NULL event");
    return swi_hip_check(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

int sw_event_elapsed(void *start, void *end, float *ms) {
    if (!start || !end || !ms) return swi_set_err(SW_E_ARG, "sw_event_elapsed: NULL pointer");
    return swi_hip_check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end), "hipEventElapsedTime");
}

int sw_stream_wait_event(void *stream, void *event) {
    if (!event) return swi_set_err(SW_E_ARG, "sw_stream_wait_event: NULL event");
    return swi_hip_check(hipStreamWaitEvent(swi_stream(stream), (hipEvent_t)event, 0), "hipStreamWaitEvent");
}

int sw_graph_begin(void *stream) {
    if (!stream) return swi_set_err(SW_E_ARG, "sw_graph_begin: capture needs a created stream");
    return swi_hip_check(hipStreamBeginCapture(swi_stream(stream), hipStreamCaptureModeThreadLocal),
                         "hipStreamBeginCapture");
}

int sw_graph_end(void *stream, void **graph_exec) {
    if (!stream || !graph_exec) return swi_set_err(SW_E_ARG, "sw_graph_end: NULL pointer");
    *graph_exec = nullptr;
    hipGraph_t g = nullptr;
    int rc = swi_hip_check(hipStreamEndCapture(swi_stream(stream), &g), "hipStreamEndCapture");
    if (rc) return rc;
    hipGraphExec_t ge = nullptr;
    rc = swi_hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    *graph_exec = ge;
    return SW_OK;
}

int sw_graph_launch(void *graph_exec, void *stream) {
    if (!graph_exec) return swi_set_err(SW_E_ARG, "sw_graph_launch: NULL graph");
    return swi_hip_check(hipGraphLaunch((hipGraphExec_t)graph_exec, swi_stream(stream)), "hipGraphLaunch");
}

int sw_graph_destroy(void *graph_exec) {
    if (!graph_exec) return SW_OK;
    return swi_hip_check(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

int sw_svf_sweep(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                 const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                 const int32_t *nann, const int32_t *dirm, int32_t *counts, void *hpad, uint32_t *masks,
                 float *fields, int rows, int cols, void *stream) {
    return sw_svf_sweep_cfg(hpack, stats, offs, tab, vegdsm2, w_iso, w_aniso, ring, nann, dirm, counts, hpad, masks,
                            fields, rows, cols, nullptr, stream);
}

int sw_svf_sweep_cfg(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                     const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                     const int32_t *nann, const int32_t *dirm, int32_t *counts, void *hpad, uint32_t *masks,
                     float *fields, int rows, int cols, const sw_march_config_t *cfg, void *stream) {
    int rc = sw_svf_march_cfg(hpack, offs, tab, counts, hpad, masks, stats, rows, cols, cfg, stream);
    if (rc) return rc;
    return sw_svf_fold(masks, vegdsm2, w_iso, w_aniso, ring, nann, dirm, fields, rows, cols, stream);
}

static int solve_day(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model,
                     const int32_t *sun_tab, const int32_t *is_day, int t0, int t1, float *out, int out_mask,
                     int walk_halo, int reset_state, void **events, void *stream) {
    if (!tile || !steps || !model || !sun_tab || !is_day || !out)
        return swi_set_err(SW_E_ARG, "sw_solve_day: NULL pointer");
    int rc = swi_check_shape(tile->rows, tile->cols, "sw_solve_day");
    if (rc) return rc;
    if (t0 < 0 || t1 < t0) return swi_set_err(SW_E_ARG, "sw_solve_day: invalid step range");
    if (out_mask <= 0 || out_mask > 255) return swi_set_err(SW_E_ARG, "sw_solve_day: out_mask selects nothing");
    const int rows = tile->rows, cols = tile->cols;
    const long n = (long)rows * cols;
    const long stride = (long)__builtin_popcount((unsigned)out_mask) * n;
    if (reset_state) {
        rc = sw_fill_f32(tile->state, 0.f, 6 * (int64_t)n, stream);
        if (rc) return rc;
    }
    for (int i = t0; i < t1; ++i) {
        void **ev = events ? events + 3 * (i - t0) : nullptr;
        if (ev) {
            rc = swi_hip_check(hipEventRecord((hipEvent_t)ev[0], swi_stream(stream)), "hipEventRecord");
            if (rc) return rc;
        }
        if (is_day[i - t0]) {
            rc = sw_sun_step(tile->hpack, tile->walls, tile->aspect_rad, tile->cls, tile->buildings, steps, model,
                             sun_tab, tile->stats, tile->src, i, rows, cols, stream);
            if (rc) return rc;
        }
        if (ev) {
            rc = swi_hip_check(hipEventRecord((hipEvent_t)ev[1], swi_stream(stream)), "hipEventRecord");
            if (rc) return rc;
        }
        rc = sw_flux_step(tile->masks, tile->src, tile->svfbuveg, tile->tan_asvf, tile->tan_svfalfa,
                          tile->walkface, tile->buildings, tile->cls, steps, model, tile->state,
                          out + (long)(i - t0) * stride, out_mask, i, is_day[i - t0] ? walk_halo : -1, rows,
                          cols, stream);
        if (rc) return rc;
        if (ev) {
            rc = swi_hip_check(hipEventRecord((hipEvent_t)ev[2], swi_stream(stream)), "hipEventRecord");
            if (rc) return rc;
        }
    }
    return SW_OK;
}

int sw_solve_day(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model, const int32_t *sun_tab,
                 const int32_t *is_day, int t0, int t1, float *out, int out_mask, int walk_halo, int reset_state,
                 void *stream) {
    return solve_day(tile, steps, model, sun_tab, is_day, t0, t1, out, out_mask, walk_halo, reset_state, nullptr,
                     stream);
}

int sw_solve_day_timed(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model,
                       const int32_t *sun_tab, const int32_t *is_day, void **events, int t0, int t1, float *out,
                       int out_mask, int walk_halo, int reset_state, void *stream) {
    if (!events) return swi_set_err(SW_E_ARG, "sw_solve_day_timed: NULL events");
    return solve_day(tile, steps, model, sun_tab, is_day, t0, t1, out, out_mask, walk_halo, reset_state, events,
                     stream);
}

}  // extern "C"
