// lb.hip -- DTW lower bounds: the warping envelope, LB_Keogh and LB_Kim.
//
// k_envelope: lower[k] / upper[k] = min / max of x over [k - r, k + r]
//   clipped to the series -- what find_min_max's deque pass computes
//   (distance/_elastic.pyx:98-150, used by _dtw_envelop :1076-1092).  Min and
//   max are exact, so the window scan gives the same values.
// k_lb_keogh: out[i, j] = sqrt(sum_t d_t) with d_t = (x - u)^2 above the
//   envelope, (x - l)^2 below it, 0 inside; summed in t order, 64 x 64 pairs
//   per block over LDS tiles
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

// 64 x 64 pairs per block: the 64 query rows and the 64 envelopes' upper and
// lower rows staged in LDS in chunks of 16 timesteps, 4 x 4 pairs per thread
// (rows ty + 16 i, envelopes tx + 16 j).  Each pair's sum runs over t in order
// with the same terms as one thread per pair (a query read by every lane of a
// wave, each envelope row by one lane, was 64 cache lines per step).
constexpr int kLbB = 64, kLbK = 16;

__global__ __launch_bounds__(256) void k_lb_keogh(const double *__restrict__ X, int64_t nx,
                                                  int64_t x_stride,
                                                  const double *__restrict__ lower,
                                                  const double *__restrict__ upper,
                                                  int64_t ny, int t,
                                                  double *__restrict__ out) {
    __shared__ double sx[kLbB][kLbK + 1];
    __shared__ double su[kLbB][kLbK + 1];
    __shared__ double sl[kLbB][kLbK + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int64_t a0 = (int64_t)blockIdx.y * kLbB, b0 = (int64_t)blockIdx.x * kLbB;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    constexpr int PER = kLbB * kLbK / 256;  // values of each tile per thread
    for (int c0 = 0; c0 < t; c0 += kLbK) {
        const int cn = t - c0 < kLbK ? t - c0 : kLbK;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = threadIdx.x + 256 * u;
            const int rr = q / kLbK, cc = q % kLbK;
            double vx = 0.0, vu = 0.0, vl = 0.0;
            if (cc < cn) {
                if (a0 + rr < nx) vx = X[(a0 + rr) * x_stride + c0 + cc];
                if (b0 + rr < ny) {
                    vu = upper[(b0 + rr) * t + c0 + cc];
                    vl = lower[(b0 + rr) * t + c0 + cc];
                }
            }
            sx[rr][cc] = vx;
            su[rr][cc] = vu;
            sl[rr][cc] = vl;
        }
        __syncthreads();
        for (int cc = 0; cc < cn; ++cc) {
            double xv[4], uv[4], lv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                xv[i] = sx[ty + 16 * i][cc];
                uv[i] = su[tx + 16 * i][cc];
                lv[i] = sl[tx + 16 * i][cc];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double v = xv[i];
                    double d = 0.0;
                    if (v > uv[j])
                        d = sq(v, uv[j]);
                    else if (v < lv[j])
                        d = sq(v, lv[j]);
                    acc[i][j] += d;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t a = a0 + ty + 16 * i;
        if (a >= nx) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t b = b0 + tx + 16 * j;
            if (b < ny) out[a * ny + b] = sqrt(acc[i][j]);
        }
    }
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
    const int64_t gx = (ny + kLbB - 1) / kLbB, gy = (nx + kLbB - 1) / kLbB;
    if (gx > 0x7fffffff || gy > 65535) return fail(WBX_EUNSUPPORTED, "LB_Keogh too large");
    hipLaunchKernelGGL(k_lb_keogh, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, X, nx,
                       x_stride, lower, upper, ny, t, out);
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
