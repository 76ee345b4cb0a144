// runtime.hip -- error state, per-device context and scratch arena.
#include <atomic>
#include <math.h>
#include <mutex>
#include <stdarg.h>
#include <stdio.h>
#include <vector>

#include "wbx_internal.h"

namespace wbx {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *last_error() { return g_err; }

static std::mutex g_ctx_mu;
static std::vector<DeviceCtx *> g_ctx;

DeviceCtx *ctx_current() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed: no HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if ((int)g_ctx.size() <= dev) g_ctx.resize(dev + 1, nullptr);
    if (!g_ctx[dev]) {
        DeviceCtx *c = new DeviceCtx();
        c->device = dev;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            set_error("hipStreamCreate failed on device %d", dev);
            return nullptr;
        }
        if (hipEventCreateWithFlags(&c->last, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            set_error("hipEventCreate failed on device %d", dev);
            return nullptr;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) c->n_cu = prop.multiProcessorCount;
        g_ctx[dev] = c;
    }
    return g_ctx[dev];
}

// Grow-only arenas.  Callers serialise on the device lock (api layer) and make
// their stream wait for the previous user of the arena before touching it.
static int arena_reserve(Workspace &a, size_t bytes, void **out) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (a.size < bytes) {
        if (a.ptr) {
            if (hipDeviceSynchronize() != hipSuccess)
                return fail(WBX_EDEVICE, "hipDeviceSynchronize failed");
            (void)hipFree(a.ptr);
            a.ptr = nullptr;
            a.size = 0;
        }
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess)
            return fail(WBX_ENOMEM, "device allocation of %zu bytes failed", bytes);
        a.ptr = p;
        a.size = bytes;
    }
    *out = a.ptr;
    return WBX_OK;
}

int ws_reserve(DeviceCtx *ctx, size_t bytes, void **out) { return arena_reserve(ctx->ws, bytes, out); }

int ws_in_reserve(DeviceCtx *ctx, size_t bytes, void **out) {
    return arena_reserve(ctx->ws_in, bytes, out);
}

int ws_dv_reserve(DeviceCtx *ctx, size_t bytes, void **out) {
    return arena_reserve(ctx->ws_dv, bytes, out);
}

// FP32 requests whose DP has more rows or columns than this run with float64
// accumulation (FP32_ACC64).  0 disables the promotion.
static std::atomic<int64_t> g_acc64_len{2048};

void set_acc64_length(int64_t len) { g_acc64_len.store(len); }

void promote_long(wbx_metric *m, int64_t len) {
    const int64_t lim = g_acc64_len.load();
    if (m->precision == WBX_FP32 && lim > 0 && len > lim) m->precision = WBX_FP32_ACC64;
}

// WDTW weights 1 / (1 + exp(-g (i - n/2))) with the host libm exp, as
// WeightedDtwMetric.reset (distance/_elastic.pyx:3334-3341) computes them,
// copied into the device vectors w64 / w32 (n elements each).
int fill_weights(int64_t n, double g, double *w64, float *w32, double *w_min) {
    std::vector<double> h((size_t)n);
    std::vector<float> hf((size_t)n);
    double lo = WBX_INF_D;
    for (int64_t i = 0; i < n; ++i) {
        h[i] = 1.0 / (1.0 + exp(-g * ((double)i - (double)n / 2.0)));
        hf[i] = (float)h[i];
        lo = h[i] < lo ? h[i] : lo;
    }
    WBX_HIP_CHECK(hipMemcpy(w64, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    WBX_HIP_CHECK(hipMemcpy(w32, hf.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    if (w_min) *w_min = lo;
    return WBX_OK;
}

int cached_weights(DeviceCtx *ctx, int64_t n, double g, const double **w64, const float **w32) {
    for (const DeviceCtx::Weights &e : ctx->weights)
        if (e.n == n && e.g == g) {
            *w64 = e.w64;
            *w32 = e.w32;
            return WBX_OK;
        }
    // A bounded cache: past 64 MiB of vectors the device is drained once and
    // every entry released (a workload cycling through that many (n, g) pairs
    // pays one synchronisation per 64 MiB, not one per change).
    const size_t bytes = (size_t)n * (sizeof(double) + sizeof(float));
    if (ctx->weight_bytes + bytes > ((size_t)64 << 20) && !ctx->weights.empty()) {
        WBX_HIP_CHECK(hipDeviceSynchronize());
        for (DeviceCtx::Weights &e : ctx->weights) {
            (void)hipFree(e.w64);
            (void)hipFree(e.w32);
        }
        ctx->weights.clear();
        ctx->weight_bytes = 0;
    }
    DeviceCtx::Weights e{n, g, nullptr, nullptr, 0.0, 0.0f};
    if (hipMalloc(&e.w64, sizeof(double) * n) != hipSuccess) return fail(WBX_ENOMEM, "weights allocation failed");
    if (hipMalloc(&e.w32, sizeof(float) * n) != hipSuccess) {
        (void)hipFree(e.w64);
        return fail(WBX_ENOMEM, "weights allocation failed");
    }
    const int rc = fill_weights(n, g, e.w64, e.w32, &e.w_min);
    e.w32_min = (float)e.w_min;  // rounding is monotone: the smallest float is the rounded smallest
    if (rc) {
        (void)hipFree(e.w64);
        (void)hipFree(e.w32);
        return rc;
    }
    ctx->weights.push_back(e);
    ctx->weight_bytes += bytes;
    *w64 = e.w64;
    *w32 = e.w32;
    return WBX_OK;
}

double cached_weights_min(const DeviceCtx *ctx, const void *w) {
    if (!ctx) return 0.0;
    for (const DeviceCtx::Weights &e : ctx->weights) {
        if ((const void *)e.w64 == w) return e.w_min;
        if ((const void *)e.w32 == w) return (double)e.w32_min;
    }
    return 0.0;
}

}  // namespace wbx
