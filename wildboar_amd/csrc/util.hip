// util.hip -- strided row conversion, fills and max-|x| reduction.
#include "wbx_internal.h"

namespace wbx {

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert_rows(const S *__restrict__ src, int64_t n,
                                                      int64_t len, int64_t sstride,
                                                      D *__restrict__ dst, int64_t dstride) {
    const int64_t total = n * len;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * 256) {
        const int64_t i = q / len, c = q - i * len;
        dst[i * dstride + c] = (D)src[i * sstride + c];
    }
}

template <typename S, typename D>
int launch_convert_rows(const S *src, int64_t n, int64_t len, int64_t sstride, D *dst,
                        int64_t dstride, hipStream_t st) {
    const int64_t total = n * len;
    if (total <= 0) return WBX_OK;
    int64_t nb = (total + 255) / 256;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL((k_convert_rows<S, D>), dim3((unsigned)nb), dim3(256), 0, st, src, n, len,
                       sstride, dst, dstride);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_convert_rows<double, float>(const double *, int64_t, int64_t, int64_t, float *,
                                                int64_t, hipStream_t);
template int launch_convert_rows<float, double>(const float *, int64_t, int64_t, int64_t, double *,
                                                int64_t, hipStream_t);
template int launch_convert_rows<double, double>(const double *, int64_t, int64_t, int64_t,
                                                 double *, int64_t, hipStream_t);
template int launch_convert_rows<float, float>(const float *, int64_t, int64_t, int64_t, float *,
                                               int64_t, hipStream_t);

__global__ __launch_bounds__(256) void k_round_f32_rows(const double *__restrict__ src, int64_t n,
                                                        int64_t len, int64_t sstride,
                                                        double *__restrict__ dst, int64_t dstride) {
    const int64_t total = n * len;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * 256) {
        const int64_t i = q / len, c = q - i * len;
        dst[i * dstride + c] = (double)(float)src[i * sstride + c];
    }
}

int launch_round_f32_rows(const double *src, int64_t n, int64_t len, int64_t sstride, double *dst,
                          int64_t dstride, hipStream_t st) {
    const int64_t total = n * len;
    if (total <= 0) return WBX_OK;
    int64_t nb = (total + 255) / 256;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_round_f32_rows, dim3((unsigned)nb), dim3(256), 0, st, src, n, len, sstride,
                       dst, dstride);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

// average_slope (distance/_elastic.pyx:3220-3225) of every row, in float64:
// dst[i][c] = ((q[c+1] - q[c]) + ((q[c+2] - q[c]) / 2)) / 2, c < len - 2.
template <typename S>
__global__ __launch_bounds__(256) void k_slope_rows(const S *__restrict__ src, int64_t n,
                                                    int64_t len, int64_t sstride,
                                                    double *__restrict__ dst) {
    const int64_t dl = len - 2;
    const int64_t total = n * dl;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * 256) {
        const int64_t i = g / dl, c = g - i * dl;
        const S *q = src + i * sstride + c;
        const double q0 = (double)q[0], q1 = (double)q[1], q2 = (double)q[2];
        dst[g] = ((q1 - q0) + ((q2 - q0) / 2)) / 2;
    }
}

template <typename S>
int launch_slope_rows(const S *src, int64_t n, int64_t len, int64_t sstride, double *dst,
                      hipStream_t st) {
    const int64_t total = n * (len - 2);
    if (total <= 0) return WBX_OK;
    int64_t nb = (total + 255) / 256;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL((k_slope_rows<S>), dim3((unsigned)nb), dim3(256), 0, st, src, n, len,
                       sstride, dst);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}
template int launch_slope_rows<double>(const double *, int64_t, int64_t, int64_t, double *,
                                       hipStream_t);
template int launch_slope_rows<float>(const float *, int64_t, int64_t, int64_t, double *,
                                      hipStream_t);

__global__ __launch_bounds__(256) void k_slope_packed(const double *__restrict__ S,
                                                      const int64_t *__restrict__ off,
                                                      const int64_t *__restrict__ len, int64_t n_s,
                                                      const int64_t *__restrict__ out_off,
                                                      int64_t max_len, double *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= n_s * max_len) return;
    const int64_t j = g / max_len, c = g - j * max_len;
    if (c >= len[j] - 2) return;
    const double *q = S + off[j] + c;
    out[out_off[j] + c] = ((q[1] - q[0]) + ((q[2] - q[0]) / 2)) / 2;
}

int launch_slope_packed(const double *S, const int64_t *off, const int64_t *len, int64_t n_s,
                        const int64_t *out_off, int64_t max_len, double *out, hipStream_t st) {
    const int64_t total = n_s * max_len;
    if (total <= 0) return WBX_OK;
    if ((total + 255) / 256 > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "too many patterns");
    hipLaunchKernelGGL(k_slope_packed, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, S,
                       off, len, n_s, out_off, max_len, out);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

__global__ __launch_bounds__(256) void k_fill_f64(double *__restrict__ dst, int64_t n, double v) {
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * 256)
        dst[q] = v;
}

int launch_fill_f64(double *dst, int64_t n, double value, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    int64_t nb = (n + 255) / 256;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)nb), dim3(256), 0, st, dst, n, value);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

__global__ __launch_bounds__(256) void k_fill_i32(int *__restrict__ dst, int64_t n, int v) {
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * 256)
        dst[q] = v;
}

int launch_fill_i32(int *dst, int64_t n, int value, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    int64_t nb = (n + 255) / 256;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)nb), dim3(256), 0, st, dst, n, value);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

// Non-negative floats order like their bit patterns, so an unsigned atomicMax
// on the bits of fabsf(v) is a max reduction.
template <typename S>
__global__ __launch_bounds__(256) void k_absmax_rows(const S *__restrict__ src, int64_t n,
                                                     int64_t len, int64_t stride,
                                                     unsigned *__restrict__ out) {
    float m = 0.0f;
    const int64_t total = n * len;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * 256) {
        const int64_t i = q / len, c = q - i * len;
        // round up so the float bound is never below the double value
        const double v = fabs((double)src[i * stride + c]);
        float f = (float)v;
        if ((double)f < v) f = __uint_as_float(__float_as_uint(f) + 1u);  // next float up
        m = fmaxf(m, f);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

template <typename S>
int launch_absmax_rows(const S *src, int64_t n, int64_t len, int64_t stride, unsigned *out,
                       hipStream_t st) {
    const int64_t total = n * len;
    if (total <= 0) return WBX_OK;
    int64_t nb = (total + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_absmax_rows<S>, dim3((unsigned)nb), dim3(256), 0, st, src, n, len, stride,
                       out);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_absmax_rows<double>(const double *, int64_t, int64_t, int64_t, unsigned *,
                                        hipStream_t);
template int launch_absmax_rows<float>(const float *, int64_t, int64_t, int64_t, unsigned *,
                                       hipStream_t);

}  // namespace wbx
