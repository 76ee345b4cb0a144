// wbx_internal.h -- shared definitions for the gfx950 DTW / distance kernels.
//
// Band geometry follows wildboar's dtw_distance (distance/_elastic.pyx:869-940):
//   w        = max(floor(min(lx, ly) * r), 1)                 (_compute_r, :1924)
//   min_len  = max(0, lx - ly),  max_len = max(0, ly - lx) + w
//   row i spans columns [max(0, i - min_len - w + 1), min(ly, i + max_len))
// In band coordinates a cell (i, j) lives in slot k = j - i + off with
// off = min_len + w - 1, k in [0, B), B = min_len + max_len + w - 1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/wbx.h"

#define WBX_INF_F __builtin_inff()
#define WBX_INF_D __builtin_inf()

namespace wbx {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define WBX_HIP_CHECK(expr)                                                  \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess)                                                \
            return ::wbx::fail(WBX_EDEVICE, "%s failed: %s (%s:%d)", #expr,  \
                               hipGetErrorString(_e), __FILE__, __LINE__);   \
    } while (0)

// ---------------------------------------------------------------- geometry
struct Band {
    int lx, ly, w, off, B;
};

__host__ __device__ inline int band_w(int lx, int ly, double r) {
    int m = lx < ly ? lx : ly;
    double v = floor((double)m * r);
    return (int)(v > 1.0 ? v : 1.0);
}

// The band of a pair for width w (w_fixed > 0), else for fraction r.
__host__ __device__ inline Band make_band(int lx, int ly, double r, int w_fixed = 0) {
    Band b;
    b.lx = lx;
    b.ly = ly;
    b.w = w_fixed > 0 ? w_fixed : band_w(lx, ly, r);
    int min_len = lx > ly ? lx - ly : 0;
    int max_len = (ly > lx ? ly - lx : 0) + b.w;
    b.off = min_len + b.w - 1;
    b.B = min_len + max_len + b.w - 1;
    return b;
}

// Exact number of band cells of the reference's loops (row 0 included).
__host__ __device__ inline int64_t band_cells(int lx, int ly, double r) {
    int w = band_w(lx, ly, r);
    int min_len = lx > ly ? lx - ly : 0;
    int max_len = (ly > lx ? ly - lx : 0) + w;
    int64_t cells = ly < max_len ? ly : max_len;  // row 0
    for (int i = 1; i < lx; i++) {
        int js = i - min_len - w + 1;
        if (js < 0) js = 0;
        int je = i + max_len;
        if (je > ly) je = ly;
        cells += je - js;
    }
    return cells;
}

// ---------------------------------------------------------------- launches
// The logical index of block b of an n-block launch such that each XCD runs a
// contiguous range: workgroups are dealt to the 8 XCDs of an MI355X round
// robin (block b to XCD b % 8), so XCD x holds blocks x, x + 8, ... and gets
// logical blocks [off(x), off(x) + count(x)).  Pairs that share a series then
// meet in one XCD's L2 instead of every XCD reading every series.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t n) {
    constexpr int64_t kXcd = 8;
    const int64_t q = n / kXcd, r = n % kXcd;
    const int64_t x = b % kXcd;
    return x * q + (x < r ? x : r) + b / kXcd;
}

// ---------------------------------------------------------------- pairs
// How a flat pair index p maps onto (a, b) = (row series, column series).
enum PairMode : int {
    PAIR_GRID = 0,      // p -> (p / ny, p % ny)          pairwise X x Y
    PAIR_TRIANGLE = 1,  // p -> (i, j), i < j, row-major  singleton X x X
    PAIR_DIAG = 2,      // p -> (p, p)                    paired
    PAIR_LIST = 3,      // p -> (pa[p], pb[p])            explicit list
};

struct PairSpec {
    int mode;
    int64_t n_pairs;
    int64_t ny;  // GRID: columns; TRIANGLE: n
    const int32_t *pa, *pb;
};

__host__ __device__ inline void decode_pair(const PairSpec &ps, int64_t p,
                                            int64_t &a, int64_t &b) {
    switch (ps.mode) {
        case PAIR_GRID:
            a = p / ps.ny;
            b = p - a * ps.ny;
            break;
        case PAIR_TRIANGLE: {
            const int64_t n = ps.ny;
            const double q = (double)(2 * n - 1);
            int64_t i = (int64_t)floor((q - sqrt(q * q - 8.0 * (double)p)) * 0.5);
            if (i < 0) i = 0;
            while (i > 0 && i * (2 * n - i - 1) / 2 > p) i--;
            while ((i + 1) * (2 * n - i - 2) / 2 <= p) i++;
            a = i;
            b = p - i * (2 * n - i - 1) / 2 + i + 1;
            break;
        }
        case PAIR_DIAG:
            a = p;
            b = p;
            break;
        default:
            a = ps.pa[p];
            b = ps.pb[p];
            break;
    }
}

// Series set: base pointer + stride between series, optional per-series length.
template <typename T>
struct Series {
    const T *base;
    int64_t stride;
    int len;              // fixed length when lens == nullptr
    const int32_t *lens;  // per-series length (EOS-trimmed) or nullptr
    __device__ inline const T *row(int64_t i) const { return base + i * stride; }
    __device__ inline int length(int64_t i) const { return lens ? lens[i] : len; }
};

// ---------------------------------------------------------------- context
struct Workspace {
    void *ptr = nullptr;
    size_t size = 0;
};

// Per-device state: device id, a default stream, a grow-only scratch arena,
// the event that marks the end of the arena's last user, and a one-entry
// cache of WDTW weights.
struct DeviceCtx {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t last = nullptr;
    Workspace ws;
    Workspace ws_in;  // inputs rounded to float32 for the FP32_ACC64 mode
    Workspace ws_dv;  // average slopes of the inputs (derivative metrics)
    int n_cu = 0;
    // WDTW weight vectors by (n, g): entries are never freed while the
    // context lives (kernels in flight may read any of them), so switching
    // between parameters does not synchronise the device.
    struct Weights {
        int64_t n;
        double g;
        double *w64;
        float *w32;
        double w_min;  // smallest weight (the lane kernels need > 0)
        float w32_min;  // smallest float32 weight
    };
    std::vector<Weights> weights;
    size_t weight_bytes = 0;
    // pairs whose distance argmin with a lower bound computed (measurement)
    int64_t prune_evaluated = 0;
};

DeviceCtx *ctx_current();              // current device's context (created lazily)
// FP32 with float64 accumulation past a DP length: past ~2048 rows the float32
// path cost drifts beyond a 1e-5 relative error (default 2048, 0 = never).
void set_acc64_length(int64_t len);
void promote_long(wbx_metric *m, int64_t len);
int ws_reserve(DeviceCtx *ctx, size_t bytes, void **out);
int ws_in_reserve(DeviceCtx *ctx, size_t bytes, void **out);
int ws_dv_reserve(DeviceCtx *ctx, size_t bytes, void **out);
// WDTW weights for width n and penalty g, computed on the host with libm exp
// and cached on the device (one entry per (n, g) per device).
int cached_weights(DeviceCtx *ctx, int64_t n, double g, const double **w64, const float **w32);
// Smallest weight of a cached float64 or float32 vector (0 if w is not one).
double cached_weights_min(const DeviceCtx *ctx, const void *w);
// Fills a caller-owned pair of device vectors with the weights for (n, g).
int fill_weights(int64_t n, double g, double *w64, float *w32, double *w_min);

// ---------------------------------------------------------------- launchers
// Plain one-pair-per-block DTW (any band, any lengths, WDTW quirk): writes the
// squared cost of each pair into sq[p].  rp (optional) gives r per x row.
// padded_iter: the measurement baseline that walks every pair over the whole
// padded length (xs.len == ys.len), same results.  w_fixed > 0: the band
// width of every pair.  penalty != NULL: the ADTW recursion with *penalty.
template <typename T>
int launch_plain_dtw(const PairSpec &ps, Series<T> xs, Series<T> ys, double r,
                     const T *weights, T *sq, hipStream_t st, const double *rp = nullptr,
                     bool padded_iter = false, int w_fixed = 0,
                     const double *penalty = nullptr);

// Euclidean: one thread per pair, sequential sum in the reference's order.
template <typename T>
int launch_euclid_pairs(const PairSpec &ps, Series<T> xs, Series<T> ys, T *sq,
                        hipStream_t st);

// Fast FP32 lane-per-pair kernel for pairs that share one geometry.
// Returns WBX_EUNSUPPORTED when the band is too wide for it.
// w_fixed > 0: band width of every pair; penalty != NULL: ADTW.
int launch_lane_dtw(const PairSpec &ps, Series<float> xs, Series<float> ys,
                    double r, const float *weights, float *sq, hipStream_t st, int w_fixed = 0,
                    const double *penalty = nullptr);
bool lane_supported(int lx, int ly, double r, int weighted, int w_fixed = 0);

// Group kernel (group.hip): one geometry per launch, n lanes per pair, 64 / n
// pairs per wave; any band width, DTW and WDTW (every weight > 0), float32
// (FMA) or float64 (the reference's arithmetic, bit-exact).
struct GroupGeo {
    int lx, ly, off, B, kf;  // band geometry shared by every pair
    int n, a, P, R;          // lanes per pair, lanes with R + 1 slots, pairs per wave, strip size
    int nw;                  // weight vector length
};
bool group_plan(int B, int lx, bool fp64, bool weighted, GroupGeo *g);

// Differentiable DTW / WDTW with band checkpoints (train_seg.hip); mode 0:
// all pairs, 1: paired.
int seg_train_sizes(const wbx_metric *m, int mode, int64_t nx, int64_t ny, int64_t tx, int64_t ty,
                    int64_t *state_bytes, int64_t *scratch_bytes);
int seg_train_forward(const wbx_metric *m, int mode, const float *X, int64_t nx, int64_t x_stride,
                      int64_t tx, const float *Y, int64_t ny, int64_t y_stride, int64_t ty,
                      double *D, void *state, hipStream_t st);
int seg_train_backward(const wbx_metric *m, int mode, const float *X, int64_t nx,
                       int64_t x_stride, int64_t tx, const float *Y, int64_t ny, int64_t y_stride,
                       int64_t ty, const double *G, const void *state, float *dX, int64_t ldx,
                       float *dY, int64_t ldy, void *scratch, hipStream_t st);

// Full-matrix kernel (full.hip): DTW whose band covers the whole matrix
// (w >= min(lx, ly), the default r = 1), one geometry per launch, lanes own
// columns; float32 (FMA) or float64 (bit-exact).
struct FullGeo {
    int lx, ly;
    int n, P, S;  // lanes per pair, pairs per wave, columns per lane
};
bool full_plan(int lx, int ly, bool fp64, FullGeo *g);
template <typename T>
int launch_full_dtw(const PairSpec &ps, Series<T> xs, Series<T> ys, T *sq, hipStream_t st);
template <typename T>
int launch_group_dtw(const PairSpec &ps, Series<T> xs, Series<T> ys, double r, const T *weights,
                     int n_weights, T *sq, hipStream_t st, int w_fixed = 0);

// sqrt + scatter of per-pair squared costs into the driver's output layout.
template <typename T>
int launch_finish(const PairSpec &ps, const T *sq, double *out, int64_t ld,
                  hipStream_t st);

// argmin with a lower bound (prune.hip): the seed pairs (first k and the
// smallest bounds of every row), then every pair that the seeds cannot prove
// the reference skips; pairs are appended to (pa, pb) at *cnt.
int prune_k_max();
int launch_prune_seed(const double *lb, int64_t nx, int64_t ny, int64_t k, int32_t *pa,
                      int32_t *pb, unsigned long long *cnt, hipStream_t st);
int launch_prune_scan(const double *D, const double *lb, int64_t nx, int64_t ny, int64_t k,
                      int32_t *pa, int32_t *pb, unsigned long long *cnt, hipStream_t st);

// Heap replay (_ArgMinBatch acceptance rule) per x row over a distance matrix.
// heaps: nx * k * 16 bytes of scratch.
int launch_argmin_replay(const double *D, int64_t nx, int64_t ny, int64_t k,
                         const double *lower_bound, void *heaps, int64_t *idx, double *dist,
                         hipStream_t st);

// End-of-series lengths: len[i] = index of the first EOS of row i (or t).
template <typename T>
int launch_eos_lengths(const T *X, int64_t n, int64_t stride, int t, int32_t *len,
                       hipStream_t st);

// Conversions and fills.
int launch_f64_to_f32(const double *src, float *dst, int64_t n, hipStream_t st);
int launch_f32_to_f64(const float *src, double *dst, int64_t n, hipStream_t st);
// dst[i * dstride + c] = (D)src[i * sstride + c] for i < n, c < len.
template <typename S, typename D>
int launch_convert_rows(const S *src, int64_t n, int64_t len, int64_t sstride, D *dst,
                        int64_t dstride, hipStream_t st);
int launch_fill_f64(double *dst, int64_t n, double value, hipStream_t st);
int launch_fill_i32(int *dst, int64_t n, int value, hipStream_t st);
// dst[i][c] = (double)(float)src[i][c]: float64 rows rounded to float32 values.
int launch_round_f32_rows(const double *src, int64_t n, int64_t len, int64_t sstride, double *dst,
                          int64_t dstride, hipStream_t st);
// dst (n rows of len - 2, packed) = average_slope of every row, in float64.
template <typename S>
int launch_slope_rows(const S *src, int64_t n, int64_t len, int64_t sstride, double *dst,
                      hipStream_t st);
// *out = max |src[i * stride + c]| as float bits (out must be zeroed first).
template <typename S>
int launch_absmax_rows(const S *src, int64_t n, int64_t len, int64_t stride, unsigned *out,
                       hipStream_t st);

// ---------------------------------------------------------------- subsequence
struct Shapelets {
    const void *S;        // packed values (element type of the launch)
    const int64_t *off;   // device: start of shapelet j
    const int64_t *len;   // device: length of shapelet j
    int64_t n;
};
// Tasks whose band class is at most lane_cap (0: none) are left to the
// lane64 kernels.
// band_add: the band of a pattern of length L is that of L + band_add
// (derivative metrics); penalty != NULL: ADTW.  Patterns of length < 1 are
// left to the caller.
template <typename T>
int launch_subseq_exact(int kind, double r, Shapelets sh, int max_len, Series<T> xs, int64_t n,
                        const T *weights, double *dist, int64_t *idx, hipStream_t st,
                        int lane_cap = 0, int band_add = 0, const double *penalty = nullptr);

// ------------------------------------------------ z-normalised scans
// pairwise_subsequence_distance with scale=True for kind WBX_EUCLIDEAN
// (ScaledEuclideanSubsequenceMetric) or WBX_DTW (ScaledDtwSubsequenceMetric,
// band fraction r): float64 patterns S (device, s_off / s_len device, hl the
// host lengths, total = packed length), pattern mean / std (device), float64
// series X (n rows of length t, stride xstride).  ws holds scaled_ws_bytes.
// wide_only: the DTW band in scratch memory for every width.
size_t scaled_ws_bytes(int kind, const int64_t *hl, int64_t n_s, int64_t total, int64_t n,
                       int64_t t);
int launch_scaled_subseq(int kind, double r, const double *S, const int64_t *s_off,
                         const int64_t *s_len, const int64_t *hl, int64_t n_s, int64_t total,
                         const double *s_mean, const double *s_std, const double *X,
                         int64_t xstride, int64_t n, int64_t t, double *dist, int64_t *idx,
                         char *ws, bool wide_only, hipStream_t st);
// pairwise_subsequence_distance with scale=True for the wrapped metrics
// (ScaledSubsequenceMetricWrap): kind WBX_WDTW (weights sized t), WBX_DDTW,
// WBX_WDDTW (weights sized t - 2) or WBX_ADTW (penalty >= 0).
size_t scaled_wrap_ws_bytes(const int64_t *hl, int64_t n_s, int64_t total, int64_t n,
                            int64_t t);
int launch_scaled_wrap(int kind, double r, const double *weights, double penalty,
                       const double *S, const int64_t *s_off, const int64_t *s_len,
                       const int64_t *hl, int64_t n_s, int64_t total, const double *s_mean,
                       const double *s_std, const double *X, int64_t xstride, int64_t n,
                       int64_t t, double *dist, int64_t *idx, char *ws, hipStream_t st);
// argmin_subsequence_distance(scale=True): the distance of every window of
// series i against pattern i, both normalised (pattern by fast_mean_std,
// windows by incremental statistics), +inf past the pattern's last window.
// The same kernels give the wrap's matching profile (match = true: the
// squared distance where the distance is below threshold, else +inf;
// patterns i or 0 by `paired`, their mean / std given).
size_t scaled_wrap_profile_ws_bytes(const int64_t *hl, int64_t n_s, int64_t n, int64_t total,
                                    int64_t t);
// norm: NORM_NONE (the raw series and patterns), NORM_GIVEN (s_mean / s_std,
// windows by incremental statistics), NORM_FAST (patterns by fast_mean_std).
// mode: 0 the distance sqrt(cost), 1 the wrap's match (distance^2 where the
// distance < threshold, else +inf), 2 the squared cost.  The band is
// _compute_r(Lw + band_add, r) with Lw the compared length.
enum { NORM_NONE = 0, NORM_GIVEN = 1, NORM_FAST = 2 };
int launch_scaled_wrap_profile(int kind, double r, const double *weights, double penalty,
                               const double *S, const int64_t *s_off, const int64_t *s_len,
                               const int64_t *hl, int64_t n_s, int64_t total, int norm,
                               const double *s_mean, const double *s_std, int paired,
                               const double *X, int64_t xstride, int64_t n, int64_t t,
                               int64_t nwp, int mode, int band_add, double threshold,
                               double *prof, char *ws, hipStream_t st);
// Window profiles of the reference's scaled matching at `threshold`:
// prof (n, nwp) holds the squared distance the matching records for window
// s of series i against pattern (paired ? i : 0), +inf where it records
// nothing or past the pattern's last window.
size_t scaled_profile_ws_bytes(int kind, const int64_t *hl, int64_t n_s, int64_t total,
                               int paired, int64_t n, int64_t t);
int launch_scaled_profile(int kind, double r, const double *S, const int64_t *s_off,
                          const int64_t *s_len, const int64_t *hl, int64_t n_s, int64_t total,
                          const double *s_mean, const double *s_std, int paired, const double *X,
                          int64_t xstride, int64_t n, int64_t t, int64_t nwp, double threshold,
                          double *prof, char *ws, bool wide_only, hipStream_t st);

// --------------------------------------------------------- warping paths
// dtw_mapping's path of each pair (X row p, Y row p), band width w, as
// per-row column ranges rows[p][2 i .. 2 i + 1]; moves: n * lx * B bytes.
// weights (nullable): w[|i - j|] for every cell, max(lx, ly) values.
int launch_path_rows(const double *X, int64_t xstride, const double *Y, int64_t ystride,
                     int64_t n, int lx, int ly, int w, const double *weights,
                     unsigned char *moves, int32_t *rows, hipStream_t st);
// mean[m] = sum of the X values aligned to m / their count (rows from
// launch_path_rows with the mean as X), in the reference's order.
int launch_mm_update(const int32_t *rows, const double *X, int64_t xstride, int64_t n, int lm,
                     const double *sample_weight, double *mean, hipStream_t st);
// One stochastic-subgradient step along the path rows of (mean, x), in place.
int launch_ssg_update(const int32_t *rows, const double *x, double w, double lr, int lm,
                      double *mean, hipStream_t st);

// ----------------------------------------------------------- lower bounds
int launch_envelope(const double *X, int64_t n, int64_t stride, int t, int r, double *lower,
                    double *upper, hipStream_t st);
int launch_lb_keogh(const double *X, int64_t nx, int64_t x_stride, const double *lower,
                    const double *upper, int64_t ny, int t, double *out, hipStream_t st);
int launch_lb_kim(const double *Y, int64_t ny, int64_t y_stride, const double *X, int64_t nx,
                  int64_t x_stride, int t, double *out, hipStream_t st);

// ------------------------------------------------------- float64 lane kernels
// Register-band FP64 DTW/WDTW, bit-identical to the reference.  lane64_class
// gives the register class (a multiple of 8) for band width B, or 0 when B is
// too wide.  Weighted launches need every weight > 0.
int lane64_class(int B, bool weighted);
// The class of a pairwise / paired band (unweighted rows up to 104 slots).
int lane64_pairs_class(int B, bool weighted);
int lane64_cap(bool weighted);
int launch_lane64_pairs(const PairSpec &ps, Series<double> xs, Series<double> ys, double r,
                        const double *weights, double *sq, hipStream_t st, int w_fixed = 0,
                        const double *penalty = nullptr);
int launch_profile_lane64(double r, Shapelets sh, int paired, int min_len, int max_len,
                          Series<double> xs, int64_t n, int64_t nwp, const double *weights,
                          double *prof, hipStream_t st);
int launch_subseq_lane64(double r, Shapelets sh, int min_len, int max_len, Series<double> xs,
                         int64_t n, const double *weights, double *dist, int64_t *idx,
                         hipStream_t st, int band_add = 0, const double *penalty = nullptr);
int launch_subseq_short(const int64_t *len, int64_t n_s, int64_t n, double *dist, int64_t *idx,
                        hipStream_t st);
// Average slopes of packed patterns: pattern j (offset off[j], length len[j])
// becomes len[j] - 2 values at out_off[j] (nothing for len[j] < 3).
int launch_slope_packed(const double *S, const int64_t *off, const int64_t *len, int64_t n_s,
                        const int64_t *out_off, int64_t max_len, double *out, hipStream_t st);
// Distance profiles (subseq_match.hip): prof[q * nwp + s] = squared cost of
// window s of series q against pattern (paired ? q : 0), +inf past its last.
// lane_ok: the lane64 kernels may take the DTW/WDTW rows whose class fits.
int launch_profile(int kind, double r, Shapelets sh, int paired, int min_len, int max_len,
                   Series<double> xs, int64_t n, int64_t nwp, const double *weights, bool lane_ok,
                   double *prof, hipStream_t st);
int launch_sqrt_inplace(double *p, int64_t n, hipStream_t st);

// ---------------------------------------------------------------- re-rank
struct ErrBound {
    double e_up, e_dn, A, e64;
};
ErrBound make_err_bound(int lx, int ly, int weighted, double absmax, int inputs_exact);
int launch_argmin_candidates(const float *sq, int64_t nx, int64_t ny, int k, const ErrBound &eb,
                             int64_t *counts, int64_t *offs, int32_t *pa, int32_t *pb,
                             int64_t *total_host, hipStream_t st);

}  // namespace wbx
