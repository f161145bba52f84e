// Context lifetime: device binding, the context's own stream, error text.
#include <algorithm>
#include <new>

#include "common.hpp"

namespace {

// Plain streaming copy: the calibration kernel for HBM byte counters and the
// achievable-bandwidth figure the roofline fractions are read against.
// 16-byte lanes and four blocks per CU, about 5.6 TB/s on MI355X.  Ordinary
// (temporal) loads and stores, so its byte counters read like the solvers'.
typedef float lane16 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy(const lane16 *__restrict__ src,
                                              lane16 *__restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" {

fl_status fl_probe_copy(fl_ctx *ctx, int64_t bytes, int32_t iterations, double *ms_per_copy) {
    if (!ctx || bytes < 16 || iterations < 1 || !ms_per_copy)
        return fl::set_error(ctx, FL_ERR_INVALID_ARGUMENT, "invalid copy probe arguments");
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)bytes / sizeof(lane16);
    lane16 *a = nullptr, *b = nullptr;
    FL_HIP_TRY(ctx, hipMalloc(&a, n * sizeof(lane16)));
    if (hipMalloc(&b, n * sizeof(lane16)) != hipSuccess) {
        (void)hipFree(a);
        return fl::set_error(ctx, FL_ERR_OUT_OF_MEMORY, "copy probe allocation failed");
    }
    hipEvent_t e0, e1;
    (void)hipMemsetAsync(a, 0, n * sizeof(lane16), ctx->stream);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cu * 4);
    k_copy<<<blocks, 256, 0, ctx->stream>>>(a, b, n);  // warm-up
    (void)hipEventRecord(e0, ctx->stream);
    for (int it = 0; it < iterations; ++it) k_copy<<<blocks, 256, 0, ctx->stream>>>(a, b, n);
    (void)hipEventRecord(e1, ctx->stream);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return fl::set_error(ctx, FL_ERR_HIP, "copy probe failed");
    *ms_per_copy = ms / iterations;
    return FL_OK;
}

fl_status fl_create(int device, fl_ctx **out) {
    if (!out) return FL_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return FL_ERR_HIP;
    if (device < 0 || device >= count) return FL_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return FL_ERR_HIP;
    fl_ctx *ctx = new (std::nothrow) fl_ctx();
    if (!ctx) return FL_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cu = prop.multiProcessorCount;
    // Default-priority, non-blocking stream (one per context).
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return FL_ERR_HIP;
    }
    // With the device current, look up its copy of the allocation-free
    // solve buffers (fl_pcg_jacobi_noalloc retries if this lookup failed).
    fl::noalloc_init(device);
    *out = ctx;
    return FL_OK;
}

static void ctx_free(fl_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

void fl_destroy(fl_ctx *ctx) {
    if (!ctx || ctx->closed) return;
    ctx->closed = true;
    if (ctx->refs == 0) ctx_free(ctx);
}

const char *fl_last_error(const fl_ctx *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

const char *fl_version(void) { return "flursys-hip 0.1.0 (gfx950)"; }

}  // extern "C"

namespace fl {
void ctx_release(fl_ctx *ctx) {
    if (!ctx) return;
    ctx->refs -= 1;
    if (ctx->refs == 0 && ctx->closed) ctx_free(ctx);
}
}  // namespace fl
