// runtime.hip -- error state, per-device context and scratch arena.
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

// Grow-only arena.  Callers serialise on the device lock (api layer) and make
// their stream wait for the previous user of the arena before touching it.
int ws_reserve(DeviceCtx *ctx, size_t bytes, void **out) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (ctx->ws.size < bytes) {
        if (ctx->ws.ptr) {
            if (hipDeviceSynchronize() != hipSuccess)
                return fail(WBX_EDEVICE, "hipDeviceSynchronize failed");
            (void)hipFree(ctx->ws.ptr);
            ctx->ws.ptr = nullptr;
            ctx->ws.size = 0;
        }
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess)
            return fail(WBX_ENOMEM, "device allocation of %zu bytes failed", bytes);
        ctx->ws.ptr = p;
        ctx->ws.size = bytes;
    }
    *out = ctx->ws.ptr;
    return WBX_OK;
}

}  // namespace wbx
