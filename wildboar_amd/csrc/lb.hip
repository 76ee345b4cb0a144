// lb.hip -- DTW lower bounds: the warping envelope, LB_Keogh and LB_Kim.
//
// k_envelope: lower[k] / upper[k] = min / max of x over [k - r, k + r]
//   clipped to the series -- what find_min_max's deque pass computes
//   (distance/_elastic.pyx:98-150, used by _dtw_envelop :1076-1092).  Min and
//   max are exact, so the window scan gives the same values.
// k_lb_keogh: out[i, j] = sqrt(sum_t d_t) with d_t = (x - u)^2 above the
//   envelope, (x - l)^2 below it, 0 inside; summed in t order
//   (cumulative_bound, _elastic.pyx:229-260, with mean 0 and std 1, as
//   _dtw_lb_keogh calls it :1095-1116).
// k_lb_kim: DtwKimLowerBound.transform (distance/lb.py:206-312): squared
//   differences of the first and last three points, in the reference's order
//   of minima and additions.
#pragma clang fp contract(off)

#include "wbx_internal.h"

namespace wbx {

namespace {

__global__ __launch_bounds__(256) void k_envelope(const double *__restrict__ X, int64_t n,
                                                  int64_t stride, int t, int r,
                                                  double *__restrict__ lower,
                                                  double *__restrict__ upper) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= n * t) return;
    const int64_t i = g / t;
    const int k = (int)(g - i * t);
    const double *__restrict__ x = X + i * stride;
    const int a = k - r > 0 ? k - r : 0;
    const int b = k + r < t - 1 ? k + r : t - 1;
    double lo = x[a], hi = x[a];
    for (int q = a + 1; q <= b; ++q) {
        const double v = x[q];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    lower[g] = lo;
    upper[g] = hi;
}

__device__ __forceinline__ double sq(double a, double b) {
    const double s = a - b;
    return s * s;
}

__global__ __launch_bounds__(256) void k_lb_keogh(const double *__restrict__ X, int64_t nx,
                                                  int64_t x_stride,
                                                  const double *__restrict__ lower,
                                                  const double *__restrict__ upper,
                                                  int64_t ny, int t,
                                                  double *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nx * ny) return;
    const int64_t i = g / ny, j = g - i * ny;
    const double *__restrict__ x = X + i * x_stride;
    const double *__restrict__ lo = lower + j * t;
    const double *__restrict__ up = upper + j * t;
    double m = 0.0;
    for (int q = 0; q < t; ++q) {
        const double v = x[q];
        double d = 0.0;
        if (v > up[q])
            d = sq(v, up[q]);
        else if (v < lo[q])
            d = sq(v, lo[q]);
        m += d;
    }
    out[g] = sqrt(m);
}

__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }

// out[i, j] for query row i (y) and fitted row j (x).
__global__ __launch_bounds__(256) void k_lb_kim(const double *__restrict__ Y, int64_t ny,
                                                int64_t y_stride, const double *__restrict__ X,
                                                int64_t nx, int64_t x_stride, int t,
                                                double *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= ny * nx) return;
    const int64_t i = g / nx, j = g - i * nx;
    const double *__restrict__ y = Y + i * y_stride;
    const double *__restrict__ x = X + j * x_stride;
    const double x0 = x[0], x0_ = x[t - 1], y0 = y[0], y0_ = y[t - 1];
    double d = sq(x0, y0) + sq(x0_, y0_);
    if (t > 1) {
        const double x1 = x[1], x1_ = x[t - 2], y1 = y[1], y1_ = y[t - 2];
        d += dmin(sq(x1, y0), dmin(sq(x0, y1), sq(x1, y1)));
        d += dmin(sq(x1_, y1), dmin(sq(x0_, y1_), sq(x1_, y1_)));
        if (t > 2) {
            const double x2 = x[2], x2_ = x[t - 3], y2 = y[2], y2_ = y[t - 3];
            d += dmin(sq(x0, y2), dmin(sq(x1, y2), dmin(sq(x2, y2), dmin(sq(x2, y1), sq(x2, y0)))));
            d += dmin(sq(x0_, y2_),
                      dmin(sq(x1_, y2_), dmin(sq(x2_, y2_), dmin(sq(x2_, y1_), sq(x2_, y0_)))));
        }
    }
    out[g] = d;
}

inline dim3 grid_for(int64_t total) { return dim3((unsigned)((total + 255) / 256)); }

}  // namespace

int launch_envelope(const double *X, int64_t n, int64_t stride, int t, int r, double *lower,
                    double *upper, hipStream_t st) {
    if (n * t <= 0) return WBX_OK;
    if ((n * t + 255) / 256 > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "envelope too large");
    hipLaunchKernelGGL(k_envelope, grid_for(n * t), dim3(256), 0, st, X, n, stride, t, r, lower,
                       upper);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_lb_keogh(const double *X, int64_t nx, int64_t x_stride, const double *lower,
                    const double *upper, int64_t ny, int t, double *out, hipStream_t st) {
    if (nx * ny <= 0) return WBX_OK;
    if ((nx * ny + 255) / 256 > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "LB_Keogh too large");
    hipLaunchKernelGGL(k_lb_keogh, grid_for(nx * ny), dim3(256), 0, st, X, nx, x_stride, lower,
                       upper, ny, t, out);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_lb_kim(const double *Y, int64_t ny, int64_t y_stride, const double *X, int64_t nx,
                  int64_t x_stride, int t, double *out, hipStream_t st) {
    if (nx * ny <= 0) return WBX_OK;
    if ((nx * ny + 255) / 256 > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "LB_Kim too large");
    hipLaunchKernelGGL(k_lb_kim, grid_for(nx * ny), dim3(256), 0, st, Y, ny, y_stride, X, nx,
                       x_stride, t, out);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
