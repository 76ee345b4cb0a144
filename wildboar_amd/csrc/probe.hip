// probe.hip -- VALU throughput probe for the roofline of the DTW kernels.
//
// Every lane runs 16 independent chains of the DTW cell mix
//     d = x - b ; m = min3(a, b, d) ; a = fma(d, d, m)
// (v_sub_f32, v_min3_f32, v_fmac_f32) with no memory traffic except one store
// per lane, on enough waves to fill every SIMD.  The result is the achievable
// rate of exactly the instructions a DTW cell needs.
#include "wbx_internal.h"

namespace wbx {

constexpr int PROBE_CHAINS = 16;

__global__ __launch_bounds__(256) void k_valu_probe(float *__restrict__ out, int iters, float seed) {
    float a[PROBE_CHAINS], b[PROBE_CHAINS];
#pragma unroll
    for (int c = 0; c < PROBE_CHAINS; ++c) {
        a[c] = seed * (float)(threadIdx.x + c);
        b[c] = seed + (float)c;
    }
    float x = seed * 0.5f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) {
            const float d = x - b[c];
            const float m = fminf(fminf(a[c], b[c]), d);
            a[c] = __builtin_fmaf(d, d, m);
        }
        x += 1e-7f;
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < PROBE_CHAINS; ++c) s += a[c];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace wbx

using namespace wbx;

extern "C" int wbx_prof_valu_peak(double *lane_instr_per_s) {
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    const int blocks = (ctx->n_cu > 0 ? ctx->n_cu : 256) * 32;
    const int iters = 20000;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)blocks * 256) != hipSuccess)
        return fail(WBX_ENOMEM, "probe buffer allocation failed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = WBX_OK;
    float ms = 0.0f;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(WBX_EDEVICE, "event creation failed");
    } else {
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, ctx->stream, out, 200, 1.0f);
        (void)hipEventRecord(e0, ctx->stream);
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, ctx->stream, out, iters, 1.0f);
        (void)hipEventRecord(e1, ctx->stream);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(WBX_EDEVICE, "probe timing failed");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc) return rc;
    const double instr = (double)blocks * 256.0 * iters * PROBE_CHAINS * 3.0;
    *lane_instr_per_s = instr / (ms * 1e-3);
    return WBX_OK;
}
