// Context, memory, level-1 and graph entry points of the sov_hip C-ABI.
// Reference counterparts: src/l0/backend_hip.cpp:56-175 (hipMalloc/hipMemcpy,
// rocBLAS level 1, the clamp kernel) and the CUDA backend's abs-max kernel
// (src/l0/backend_cuda.cu:61-80).  Here every reduction is a fixed-order
// two-stage wave64 reduction in FP64, and nothing synchronises the device.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "sov_hip_internal.hpp"

namespace sovhip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? SOV_OUT_OF_MEMORY : SOV_BACKEND_ERROR;
}

int launch_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b > 65535LL * 64) b = 65535LL * 64;   // grid-stride beyond this
    return static_cast<int>(b < 1 ? 1 : b);
}

hipError_t check_launch() { return hipGetLastError(); }

// ---------------------------------------------------------------------------
// elementwise / reduction kernels
// ---------------------------------------------------------------------------

__global__ void k_fill_zero_u32(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_fill_zero_u8(uint8_t* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

__global__ void k_scal(int64_t n, double a, double* __restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= a;
}

// Clamp with the reference's comparison order (backend_host.cpp:174-182): a
// NaN stays NaN, a null bound is skipped.
__global__ void k_project_box(int64_t n, const double* __restrict__ lo,
                              const double* __restrict__ hi, double* __restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = x[i];
        if (lo != nullptr && v < lo[i]) v = lo[i];
        if (hi != nullptr && v > hi[i]) v = hi[i];
        x[i] = v;
    }
}

__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ s, float* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = static_cast<float>(s[i]);
}

__global__ void k_f32_to_f64(int64_t n, const float* __restrict__ s, double* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = static_cast<double>(s[i]);
}

// Stage 1 of dot / amax: each block reduces a grid-strided share to one value.
// mode 0: sum x*y, mode 1: max |x|.
__global__ void k_reduce_stage1(int64_t n, const double* __restrict__ x,
                                const double* __restrict__ y, int mode,
                                double* __restrict__ partial) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (mode == 0) acc += x[i] * y[i];
        else acc = nmax(acc, fabs(x[i]));
    }
    double s[1] = {acc}, mx[1] = {acc};
    if (mode == 0) block_reduce_write<1, 0>(s, mx, partial + blockIdx.x);
    else block_reduce_write<0, 1>(s, mx, partial + blockIdx.x);
}

// Stage 2: one block reduces `count` partials in a fixed order.
__global__ void k_reduce_stage2(int64_t count, const double* __restrict__ partial, int mode,
                                double* __restrict__ out) {
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += blockDim.x)
        acc = (mode == 0) ? acc + partial[i] : nmax(acc, partial[i]);
    double s[1] = {acc}, mx[1] = {acc};
    if (mode == 0) block_reduce_write<1, 0>(s, mx, out);
    else block_reduce_write<0, 1>(s, mx, out);
}

constexpr int kRedBlocks = 1024;

static int reduce_to_host(sov_hip_ctx* ctx, int64_t n, const double* x, const double* y, int mode,
                          double* result) {
    int rc = ctx->ensure_red(kRedBlocks + 8);
    if (rc != SOV_OK) return rc;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > kRedBlocks) blocks = kRedBlocks;
    if (blocks < 1) blocks = 1;
    k_reduce_stage1<<<(int)blocks, kBlock, 0, ctx->stream>>>(n, x, y, mode, ctx->d_red);
    k_reduce_stage2<<<1, kBlock, 0, ctx->stream>>>(blocks, ctx->d_red, mode, ctx->d_red + kRedBlocks);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, ctx->d_red + kRedBlocks, sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *result = ctx->h_pinned[0];
    return SOV_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace sovhip

using namespace sovhip;

int sov_hip_ctx::ensure_red(size_t count) {
    if (count <= d_red_count) return SOV_OK;
    if (capturing) return fail(SOV_INVALID_ARGUMENT, "scratch growth during graph capture");
    if (d_red != nullptr) (void)hipFree(d_red);
    d_red = nullptr;
    d_red_count = 0;
    SOV_HIP_TRY(hipMalloc(&d_red, count * sizeof(double)));
    d_red_count = count;
    return SOV_OK;
}

extern "C" {

const char* sov_hip_last_error(void) { return sovhip::g_last_error.c_str(); }

int sov_hip_device_count(int* count) {
    if (count == nullptr) return fail(SOV_INVALID_ARGUMENT, "count is null");
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        (void)hipGetLastError();
        return SOV_OK;   // no runtime device: zero devices, not an error
    }
    *count = c;
    return SOV_OK;
}

int sov_hip_ctx_create(int device, sov_hip_ctx** out) {
    if (out == nullptr) return fail(SOV_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    sov_hip_device_count(&count);
    if (device < 0 || device >= count)
        return fail(SOV_UNSUPPORTED, "no HIP device " + std::to_string(device) + " (found " +
                                         std::to_string(count) + ")");
    SOV_HIP_TRY(hipSetDevice(device));
    auto* ctx = new (std::nothrow) sov_hip_ctx();
    if (ctx == nullptr) return fail(SOV_OUT_OF_MEMORY, "ctx");
    ctx->device = device;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->name = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
                    std::to_string(prop.multiProcessorCount) + " CUs, " +
                    std::to_string(prop.totalGlobalMem >> 30) + " GiB)";
        ctx->cus = prop.multiProcessorCount;
        ctx->l2_bytes = prop.l2CacheSize;
        ctx->lds_max = static_cast<int64_t>(prop.sharedMemPerBlock);
    } else {
        ctx->name = "HIP device " + std::to_string(device);
    }
    // Default-priority stream (non-blocking w.r.t. the null stream).
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&ctx->h_pinned), 64 * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e != hipSuccess) {
        const int rc = hip_fail(e, "sov_hip_ctx_create");
        sov_hip_ctx_destroy(ctx);
        return rc;
    }
    const int src = static_scratch_address(device, &ctx->check_scratch);
    if (src != SOV_OK) {
        sov_hip_ctx_destroy(ctx);
        return src;
    }
    *out = ctx;
    return SOV_OK;
}

void sov_hip_ctx_destroy(sov_hip_ctx* ctx) {
    if (ctx == nullptr) return;
    DeviceGuard g(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_red) (void)hipFree(ctx->d_red);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* sov_hip_ctx_device_name(const sov_hip_ctx* ctx) {
    return ctx ? ctx->name.c_str() : "";
}

int sov_hip_ctx_props(const sov_hip_ctx* ctx, int* cus, int64_t* l2, int64_t* mall) {
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "ctx is null");
    if (cus) *cus = ctx->cus;
    if (l2) *l2 = ctx->l2_bytes;
    if (mall) *mall = ctx->mall_bytes;
    return SOV_OK;
}

int sov_hip_malloc(sov_hip_ctx* ctx, size_t bytes, void** out) {
    if (ctx == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "malloc args");
    *out = nullptr;
    if (bytes == 0) return SOV_OK;
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "malloc during graph capture");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipMalloc(out, bytes));
    return SOV_OK;
}

int sov_hip_free(sov_hip_ctx* ctx, void* ptr) {
    if (ptr == nullptr) return SOV_OK;
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "free without ctx");
    DeviceGuard g(ctx->device);
    // Work queued on this context's stream may still read ptr.
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    SOV_HIP_TRY(hipFree(ptr));
    return SOV_OK;
}

int sov_hip_memcpy_h2d(sov_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SOV_OK;
    if (ctx == nullptr || dst == nullptr || src == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "memcpy_h2d args");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_memcpy_d2h(sov_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SOV_OK;
    if (ctx == nullptr || dst == nullptr || src == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "memcpy_d2h args");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_memcpy_d2d(sov_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SOV_OK;
    if (ctx == nullptr || dst == nullptr || src == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "memcpy_d2d args");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SOV_OK;
}

int sov_hip_fill_zero(sov_hip_ctx* ctx, void* dst, size_t bytes) {
    if (bytes == 0) return SOV_OK;
    if (ctx == nullptr || dst == nullptr) return fail(SOV_INVALID_ARGUMENT, "fill_zero args");
    DeviceGuard g(ctx->device);
    // A fill kernel rather than hipMemsetAsync, so this is safe under capture.
    if (bytes % 4 == 0 && (reinterpret_cast<uintptr_t>(dst) % 4) == 0) {
        const int64_t n = static_cast<int64_t>(bytes / 4);
        k_fill_zero_u32<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(static_cast<uint32_t*>(dst), n);
    } else {
        const int64_t n = static_cast<int64_t>(bytes);
        k_fill_zero_u8<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(static_cast<uint8_t*>(dst), n);
    }
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_sync(sov_hip_ctx* ctx) {
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "sync without ctx");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_axpy_f64(sov_hip_ctx* ctx, int64_t n, double alpha, const double* x, double* y) {
    if (n < 0) return fail(SOV_INVALID_ARGUMENT, "axpy: n < 0");
    if (n == 0) return SOV_OK;
    if (ctx == nullptr || x == nullptr || y == nullptr) return fail(SOV_INVALID_ARGUMENT, "axpy args");
    DeviceGuard g(ctx->device);
    k_axpy<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, alpha, x, y);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_scal_f64(sov_hip_ctx* ctx, int64_t n, double alpha, double* x) {
    if (n < 0) return fail(SOV_INVALID_ARGUMENT, "scal: n < 0");
    if (n == 0) return SOV_OK;
    if (ctx == nullptr || x == nullptr) return fail(SOV_INVALID_ARGUMENT, "scal args");
    DeviceGuard g(ctx->device);
    k_scal<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, alpha, x);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_dot_f64(sov_hip_ctx* ctx, int64_t n, const double* x, const double* y, double* result) {
    if (n < 0 || result == nullptr) return fail(SOV_INVALID_ARGUMENT, "dot args");
    if (n == 0) { *result = 0.0; return SOV_OK; }
    if (ctx == nullptr || x == nullptr || y == nullptr) return fail(SOV_INVALID_ARGUMENT, "dot args");
    DeviceGuard g(ctx->device);
    return reduce_to_host(ctx, n, x, y, 0, result);
}

int sov_hip_nrm2_f64(sov_hip_ctx* ctx, int64_t n, const double* x, double* result) {
    double d = 0.0;
    const int rc = sov_hip_dot_f64(ctx, n, x, x, &d);
    if (rc != SOV_OK) return rc;
    *result = std::sqrt(d);
    return SOV_OK;
}

int sov_hip_amax_f64(sov_hip_ctx* ctx, int64_t n, const double* x, double* result) {
    if (n < 0 || result == nullptr) return fail(SOV_INVALID_ARGUMENT, "amax args");
    if (n == 0) { *result = 0.0; return SOV_OK; }
    if (ctx == nullptr || x == nullptr) return fail(SOV_INVALID_ARGUMENT, "amax args");
    DeviceGuard g(ctx->device);
    return reduce_to_host(ctx, n, x, nullptr, 1, result);
}

int sov_hip_project_box_f64(sov_hip_ctx* ctx, int64_t n, const double* lo, const double* hi,
                            double* x) {
    if (n < 0) return fail(SOV_INVALID_ARGUMENT, "project_box: n < 0");
    if (n == 0) return SOV_OK;
    if (ctx == nullptr || x == nullptr) return fail(SOV_INVALID_ARGUMENT, "project_box args");
    DeviceGuard g(ctx->device);
    k_project_box<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, lo, hi, x);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_convert_f64_f32(sov_hip_ctx* ctx, int64_t n, const double* src, float* dst) {
    if (n <= 0) return n < 0 ? fail(SOV_INVALID_ARGUMENT, "convert: n < 0") : SOV_OK;
    if (ctx == nullptr || src == nullptr || dst == nullptr) return fail(SOV_INVALID_ARGUMENT, "convert args");
    DeviceGuard g(ctx->device);
    k_f64_to_f32<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, src, dst);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_convert_f32_f64(sov_hip_ctx* ctx, int64_t n, const float* src, double* dst) {
    if (n <= 0) return n < 0 ? fail(SOV_INVALID_ARGUMENT, "convert: n < 0") : SOV_OK;
    if (ctx == nullptr || src == nullptr || dst == nullptr) return fail(SOV_INVALID_ARGUMENT, "convert args");
    DeviceGuard g(ctx->device);
    k_f32_to_f64<<<launch_blocks(n), kBlock, 0, ctx->stream>>>(n, src, dst);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_graph_begin(sov_hip_ctx* ctx) {
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "graph_begin without ctx");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "already capturing");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return SOV_OK;
}

int sov_hip_graph_end(sov_hip_ctx* ctx, sov_hip_graph** out) {
    if (ctx == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "graph_end args");
    if (!ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "not capturing");
    DeviceGuard g(ctx->device);
    ctx->capturing = false;
    auto* gr = new (std::nothrow) sov_hip_graph();
    if (gr == nullptr) return fail(SOV_OUT_OF_MEMORY, "graph");
    hipError_t e = hipStreamEndCapture(ctx->stream, &gr->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        sov_hip_graph_destroy(gr);
        return hip_fail(e, "sov_hip_graph_end");
    }
    *out = gr;
    return SOV_OK;
}

int sov_hip_graph_launch(sov_hip_ctx* ctx, sov_hip_graph* gr) {
    if (ctx == nullptr || gr == nullptr) return fail(SOV_INVALID_ARGUMENT, "graph_launch args");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipGraphLaunch(gr->exec, ctx->stream));
    return SOV_OK;
}

void sov_hip_graph_destroy(sov_hip_graph* gr) {
    if (gr == nullptr) return;
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
}

int sov_hip_timer_start(sov_hip_ctx* ctx) {
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "timer without ctx");
    SOV_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    return SOV_OK;
}

int sov_hip_timer_stop(sov_hip_ctx* ctx, double* ms) {
    if (ctx == nullptr || ms == nullptr) return fail(SOV_INVALID_ARGUMENT, "timer args");
    SOV_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    SOV_HIP_TRY(hipEventSynchronize(ctx->ev1));
    float f = 0.f;
    SOV_HIP_TRY(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = f;
    return SOV_OK;
}

struct sov_hip_copier {
    int device = 0;
    hipStream_t stream = nullptr;   // default priority, non-blocking
    hipEvent_t ready = nullptr;     // recorded on the context stream
    hipEvent_t done = nullptr;      // recorded on the copier stream
};

int sov_hip_copier_create(sov_hip_ctx* ctx, sov_hip_copier** out) {
    if (ctx == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "copier_create args");
    *out = nullptr;
    DeviceGuard g(ctx->device);
    auto* c = new (std::nothrow) sov_hip_copier();
    if (c == nullptr) return fail(SOV_OUT_OF_MEMORY, "copier");
    c->device = ctx->device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        sov_hip_copier_destroy(c);
        return hip_fail(e, "sov_hip_copier_create");
    }
    *out = c;
    return SOV_OK;
}

void sov_hip_copier_destroy(sov_hip_copier* c) {
    if (c == nullptr) return;
    DeviceGuard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ready) (void)hipEventDestroy(c->ready);
    if (c->done) (void)hipEventDestroy(c->done);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int sov_hip_copier_d2h(sov_hip_ctx* ctx, sov_hip_copier* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SOV_OK;
    if (ctx == nullptr || c == nullptr || dst == nullptr || src == nullptr)
        return fail(SOV_INVALID_ARGUMENT, "copier_d2h args");
    if (ctx->capturing) return fail(SOV_INVALID_ARGUMENT, "copier_d2h during capture");
    DeviceGuard g(ctx->device);
    // after everything already queued on the context stream
    SOV_HIP_TRY(hipEventRecord(c->ready, ctx->stream));
    SOV_HIP_TRY(hipStreamWaitEvent(c->stream, c->ready, 0));
    SOV_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SOV_HIP_TRY(hipEventRecord(c->done, c->stream));
    return SOV_OK;
}

int sov_hip_copier_wait(sov_hip_copier* c) {
    if (c == nullptr) return fail(SOV_INVALID_ARGUMENT, "copier_wait args");
    SOV_HIP_TRY(hipEventSynchronize(c->done));
    return SOV_OK;
}

int sov_hip_host_alloc(sov_hip_ctx* ctx, size_t bytes, void** out) {
    if (ctx == nullptr || out == nullptr) return fail(SOV_INVALID_ARGUMENT, "host_alloc args");
    *out = nullptr;
    if (bytes == 0) return SOV_OK;
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipHostMalloc(out, bytes));
    return SOV_OK;
}

int sov_hip_host_free(sov_hip_ctx* ctx, void* p) {
    if (p == nullptr) return SOV_OK;
    if (ctx == nullptr) return fail(SOV_INVALID_ARGUMENT, "host_free without ctx");
    DeviceGuard g(ctx->device);
    SOV_HIP_TRY(hipHostFree(p));
    return SOV_OK;
}

struct sov_hip_events {
    std::vector<hipEvent_t> ev;
};

int sov_hip_events_create(sov_hip_ctx* ctx, int count, sov_hip_events** out) {
    if (ctx == nullptr || out == nullptr || count <= 0) return fail(SOV_INVALID_ARGUMENT, "events_create args");
    DeviceGuard g(ctx->device);
    auto* e = new (std::nothrow) sov_hip_events();
    if (e == nullptr) return fail(SOV_OUT_OF_MEMORY, "events");
    e->ev.assign(static_cast<size_t>(count), nullptr);
    for (auto& x : e->ev) {
        const hipError_t r = hipEventCreate(&x);
        if (r != hipSuccess) {
            sov_hip_events_destroy(e);
            return hip_fail(r, "hipEventCreate");
        }
    }
    *out = e;
    return SOV_OK;
}

int sov_hip_events_record(sov_hip_ctx* ctx, sov_hip_events* e, int i) {
    if (ctx == nullptr || e == nullptr || i < 0 || i >= (int)e->ev.size())
        return fail(SOV_INVALID_ARGUMENT, "events_record args");
    SOV_HIP_TRY(hipEventRecord(e->ev[static_cast<size_t>(i)], ctx->stream));
    return SOV_OK;
}

int sov_hip_events_elapsed(sov_hip_events* e, int i, int j, double* ms) {
    if (e == nullptr || ms == nullptr || i < 0 || j < 0 || i >= (int)e->ev.size() || j >= (int)e->ev.size())
        return fail(SOV_INVALID_ARGUMENT, "events_elapsed args");
    float f = 0.f;
    SOV_HIP_TRY(hipEventElapsedTime(&f, e->ev[static_cast<size_t>(i)], e->ev[static_cast<size_t>(j)]));
    *ms = f;
    return SOV_OK;
}

void sov_hip_events_destroy(sov_hip_events* e) {
    if (e == nullptr) return;
    for (auto& x : e->ev)
        if (x) (void)hipEventDestroy(x);
    delete e;
}

}  // extern "C"
