/*
 * wbx.h -- C ABI of the MI355X (gfx950) DTW / distance library (libwbx.so).
 *
 * Drop-in for the batch drivers of wildboar's Cython metric layer
 * (src/wildboar/distance/_cdistance.pyx) for metric = "euclidean", "dtw" and
 * "wdtw".  Plain C types only: pointers, sizes, status codes.  No exceptions
 * cross this boundary, nothing here calls Python, and every function may be
 * called with the GIL released.
 *
 * Conventions
 *   - Return value: WBX_OK (0) or a negative WBX_E* code; the message is in
 *     wbx_last_error() (thread-local).
 *   - Inputs are borrowed; outputs are caller-owned and fully written.
 *   - Time-series arrays follow the reference's TSArray layout
 *     (src/wildboar/utils/__init__.pxd:4): C-contiguous (n, dims, t) float64;
 *     `dim` selects the row &X[i, dim, 0].
 *   - Calls on one device are serialised by a per-device lock; distinct
 *     devices run concurrently.
 */
#ifndef WBX_H
#define WBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WBX_OK 0
#define WBX_EINVAL (-1)        /* bad argument: maps to ValueError        */
#define WBX_ENOMEM (-2)        /* device/host allocation failed           */
#define WBX_EDEVICE (-3)       /* HIP runtime error / no device           */
#define WBX_EUNSUPPORTED (-4)  /* valid request this build cannot serve   */

/* Metric kinds: the names of the reference's _METRICS registry
 * (distance/_distance.py:185-204) this library serves. */
#define WBX_EUCLIDEAN 0 /* EuclideanMetric,  distance/_metric.pyx:381-409   */
#define WBX_DTW 1       /* DtwMetric,        distance/_elastic.pyx:3126-3217 */
#define WBX_WDTW 2      /* WeightedDtwMetric, distance/_elastic.pyx:3322-3341 */
#define WBX_DDTW 3      /* DerivativeDtwMetric, distance/_elastic.pyx:3228-3320:
                           DTW on average_slope series (len - 2); pairwise and
                           paired band by min(x_len, y_len), argmin by
                           min(x_len - 2, y_len - 2); series shorter than 3 give
                           0 (argmin: never accepted, zero outputs)            */
#define WBX_WDDTW 4     /* WeightedDerivativeDtwMetric, _elastic.pyx:3404-3430:
                           as WBX_DDTW with WDTW weights sized
                           max(x_len, y_len) - 2                               */
#define WBX_ADTW 5      /* AmercingDtwMetric, distance/_elastic.pyx:3344-3401,
                           adtw_distance :943-1008: penalty p on the up and
                           left moves outside row 0                            */

/* Arithmetic modes. */
#define WBX_FP64 0        /* float64, bit-identical to the reference            */
#define WBX_FP32 1        /* float32 cells; argmin/labels certified exact      */
#define WBX_FP32_ACC64 2  /* float32 data, float64 recursion: inputs rounded to
                             float32 once, then bit-identical to the reference
                             on those rounded values                            */

/* Input element types for the device-pointer entry points. */
#define WBX_F64 0
#define WBX_F32 1

typedef struct wbx_metric {
    int32_t kind;      /* WBX_EUCLIDEAN .. WBX_ADTW                           */
    int32_t precision; /* WBX_FP64 / WBX_FP32 / WBX_FP32_ACC64                */
    double r;          /* Sakoe-Chiba window fraction in [0, 1] (elastic)    */
    double g;          /* WDTW / WDDTW penalty control, >= 0                  */
    double p;          /* ADTW amercing penalty (not NaN, not -inf)           */
} wbx_metric;

/* ----------------------------------------------------------------- runtime */

/* Select and initialise a device (creates its stream and scratch arena). */
int wbx_init(int device);
/* Thread-local message of the last failed call ("" if none). */
const char *wbx_last_error(void);
/* Library version string. */
const char *wbx_version(void);
/* Number of HIP devices visible (0 when there is none). */
int wbx_device_count(void);

/* No counterpart in the reference (tuning / measurement).  Kernel family:
 * 0 = fastest available (default), 1 = the plain one-pair-per-block kernel
 * for every DTW request (the "plain" baseline of the benchmarks), 2 = the
 * group kernel (n lanes per pair) wherever it applies. */
int wbx_set_policy(int policy);
/* No counterpart in the reference (the reference computes in float64).
 * WBX_FP32 requests of pairwise / paired / argmin / batched pairs whose
 * series are longer than `len` run as WBX_FP32_ACC64 (float32 data, float64
 * accumulation): past ~2048 steps the float32 path cost drifts beyond a
 * 1e-5 relative error.  Default 2048; 0 never promotes. */
int wbx_set_acc64_length(int64_t len);
/* No counterpart in the reference (measurement).  While enabled, HIP events
 * are recorded on the launch stream around every distance kernel;
 * wbx_prof_take waits for them, returns the summed milliseconds, the number
 * of kernels and the kind of the last one (1 euclidean, 2 lane-per-pair DTW,
 * 3 plain DTW, 4 batched DTW, 5 subsequence scan, 6 group of lanes per pair,
 * 7 full-matrix DTW), and clears the record. */
int wbx_prof_enable(int on);
int wbx_prof_take(double *total_ms, int64_t *launches, int *last_kind);
/* No counterpart in the reference (measurement).  The number of pairs whose
 * distance argmin with a lower bound has computed on the current device since
 * the last call (the reference evaluates only pairs whose bound is below its
 * running threshold; the library computes those and a few seeds per row,
 * _cdistance.pyx:1329-1340), then clears it. */
int wbx_prof_pruned(int64_t *evaluated);
/* No counterpart in the reference (measurement).  Runs a VALU probe made of
 * the DTW cell's instructions (sub, min3, fma) on every SIMD of the current
 * device and returns the achieved lane-instructions per second. */
int wbx_prof_valu_peak(double *lane_instr_per_s);

/* ------------------------------------------- host-pointer drop-ins (float64)
 * Synchronous.  Inputs and outputs are host memory.  These replace the
 * reference's batch drivers one for one.                                    */

/* Replaces _pairwise_distance (distance/_cdistance.pyx:1184-1205) and, when
 * Y == NULL, _singleton_pairwise_distance (_cdistance.pyx:1248-1267).
 * out is (nx, ny) (or (nx, nx) for Y == NULL, zero diagonal, i<j computed
 * as D(X[i], X[j]) and mirrored). */
int wbx_pairwise(const wbx_metric *metric, const double *X, int64_t nx,
                 int64_t dims, int64_t tx, const double *Y, int64_t ny,
                 int64_t ty, int64_t dim, double *out);

/* Replaces _paired_distance (distance/_cdistance.pyx:1632-1652):
 * out[i] = D(Y[i], X[i]) where X, Y are paired_distance's x and y
 * (the reference swaps the arguments; this matters for WDTW). */
int wbx_paired(const wbx_metric *metric, const double *X, const double *Y,
               int64_t n, int64_t dims, int64_t t, int64_t dim, double *out);

/* Replaces _argmin_distance (distance/_cdistance.pyx:1348-1378): the k
 * nearest Y rows of each X row, in the reference's heap order (Heap,
 * utils/_misc.pyx:18-107).  lower_bound (nx, ny) may be NULL; with one and
 * k <= 16 only the distances the reference evaluates (bound below its running
 * threshold, :1329-1340) and a few seeds per row are computed, in float64.
 * idx is (nx, k) int64, dist is (nx, k). */
int wbx_argmin(const wbx_metric *metric, const double *X, int64_t nx,
               int64_t dims, int64_t tx, const double *Y, int64_t ny,
               int64_t ty, int64_t dim, int64_t k, const double *lower_bound,
               int64_t *idx, double *dist);

/* Replaces _pairwise_subsequence_distance (distance/_cdistance.pyx:1018-1037):
 * shapelets are packed back to back in S (shapelet j starts at s_off[j] and
 * has s_len[j] values); dist and idx are (n, n_s): minimum distance over the
 * windows of X[i, dim] and the first window index attaining it. */
int wbx_subsequence_pairwise(const wbx_metric *metric, const double *S,
                             const int64_t *s_off, const int64_t *s_len,
                             int64_t n_s, const double *X, int64_t n,
                             int64_t dims, int64_t t, int64_t dim, double *dist,
                             int64_t *idx);

/* Replaces _argmin_subsequence_distance (distance/_cdistance.pyx:1572-1594,
 * non-scaled _ArgminSubsequenceDistance :1408-1434; any metric kind, adtw
 * with penalty >= 0): pattern i (s_len[i]
 * values at S + s_off[i]) against series i of X (n series), the k best
 * windows in the reference's heap order (Heap, utils/_misc.pyx:18-107).  idx
 * and dist are (n, k).  k <= t - longest pattern + 1. */
int wbx_argmin_subsequence(const wbx_metric *metric, const double *S,
                           const int64_t *s_off, const int64_t *s_len,
                           int64_t n, const double *X, int64_t dims, int64_t t,
                           int64_t dim, int64_t k, int64_t *idx, double *dist);

/* pairwise_subsequence_distance with scale=True (distance/_distance.py:
 * 543-636; "scaled_" + metric, :280-300) through
 * _pairwise_subsequence_distance (_cdistance.pyx:1018-1037), replacing
 * ScaledEuclideanSubsequenceMetric (metric->kind WBX_EUCLIDEAN,
 * distance/_metric.pyx:96-160, scaled_euclidean_distance :561-651),
 * ScaledDtwSubsequenceMetric (WBX_DTW with band fraction r,
 * distance/_elastic.pyx:1928-2203, scaled_dtw_subsequence_distance :343-482,
 * its LB_Kim / LB_Keogh cascade included) and ScaledSubsequenceMetricWrap
 * (_cdistance.pyx:470-567) around WBX_WDTW, WBX_DDTW, WBX_WDDTW or WBX_ADTW
 * (penalty >= 0): windows normalised with incremental statistics, the first
 * window of smallest wrapped distance.  Every window is z-normalised with
 * the reference's running mean and std; s_mean / s_std are each pattern's
 * mean and std as the reference's Subsequence holds them (from_array,
 * _cdistance.pyx:453-467: np.mean, np.std, std <= 1e-13 stored as 0; a 0 std
 * is used as 1.0, persistent_distance :283-299).  Float64 arithmetic for
 * every precision.  Outputs as wbx_subsequence_pairwise; a series with no
 * accepted window gets index 0 (the reference leaves it unset). */
int wbx_scaled_subsequence_pairwise(const wbx_metric *metric, const double *S,
                                    const int64_t *s_off, const int64_t *s_len,
                                    const double *s_mean, const double *s_std,
                                    int64_t n_s, const double *X, int64_t n,
                                    int64_t dims, int64_t t, int64_t dim,
                                    double *dist, int64_t *idx);

/* The window profiles behind subsequence_match / paired_subsequence_match
 * with scale=True (distance/_distance.py:732-1079): the windows the
 * reference's scaled matching records at `threshold` --
 * scaled_euclidean_distance_matches (distance/_metric.pyx:798-866) for
 * metric->kind WBX_EUCLIDEAN, scaled_dtw_matches (distance/_elastic.pyx:
 * 484-619, its LB_Kim / LB_Keogh cascade against threshold^2 included) for
 * WBX_DTW, ScaledSubsequenceMetricWrap._matches (_cdistance.pyx:547-615:
 * windows whose distance is below threshold, strict) for WBX_WDTW,
 * WBX_DDTW, WBX_WDDTW and WBX_ADTW (penalty >= 0).  prof (n, nwp), nwp = t - shortest pattern + 1 (pattern 0 alone
 * when not paired), holds the squared distance recorded for window s of
 * series i against pattern (paired ? i : 0) -- sqrt of it is the reported
 * distance -- and +inf where nothing is recorded or past the pattern's last
 * window.  threshold = +inf records every window (the reference's distance
 * profile).  s_mean / s_std as for wbx_scaled_subsequence_pairwise. */
int wbx_scaled_subsequence_profile(const wbx_metric *metric, const double *S,
                                   const int64_t *s_off, const int64_t *s_len,
                                   const double *s_mean, const double *s_std,
                                   int64_t n_s, int paired, const double *X,
                                   int64_t n, int64_t dims, int64_t t,
                                   int64_t dim, int64_t nwp, double threshold,
                                   double *prof);

/* argmin_subsequence_distance with scale=True (distance/_distance.py:
 * 1738-1756) through _ScaledArgminSubsequenceDistance
 * (distance/_cdistance.pyx:1451-1549): pattern i, normalised with its
 * fast_mean_std (utils/_stats.pyx:22-42; std 0 read as 1.0), against the
 * windows of series i normalised with incremental statistics, through the
 * metric's _eadistance (any kind; adtw with penalty >= 0) and the k-Heap; idx
 * and dist (n, k) in heap order.  k <= t - longest pattern + 1. */
int wbx_scaled_argmin_subsequence(const wbx_metric *metric, const double *S,
                                  const int64_t *s_off, const int64_t *s_len,
                                  int64_t n, const double *X, int64_t dims,
                                  int64_t t, int64_t dim, int64_t k,
                                  int64_t *idx, double *dist);

/* The per-series window costs behind _subsequence_match and
 * _paired_subsequence_match (distance/_cdistance.pyx:1065-1141, windows with
 * cost <= threshold^2 as dtw_subsequence_matches, _elastic.pyx:658-698,
 * euclidean_distance_matches, _metric.pyx:762-795, ddtw_ / adtw_subsequence_
 * matches, _elastic.pyx:737-866 -- band by the pattern's own length, nothing
 * below length 3 for the derivative kinds, adtw penalty >= 0): prof (n, nwp),
 * nwp = t - shortest pattern + 1, holds the squared cost of window s of
 * series i against pattern (paired ? i : 0), +inf past the pattern's last
 * window. */
int wbx_subsequence_profile(const wbx_metric *metric, const double *S,
                            const int64_t *s_off, const int64_t *s_len,
                            int64_t n_s, int paired, const double *X, int64_t n,
                            int64_t dims, int64_t t, int64_t dim, int64_t nwp,
                            double *prof);

/* ------------------------------------------- device-pointer entry points
 * Inputs are device pointers of element type `dtype` (WBX_F64 / WBX_F32);
 * series i of X starts at X + i * x_stride.  Outputs are device memory.
 * Work is enqueued on `stream` (a hipStream_t; NULL = HIP's default null
 * stream) and the call returns once it is enqueued on that stream.  Calls
 * that must read a count back to the host (argmin in FP32 mode, the
 * subsequence scan) synchronise `stream` before returning.                  */

/* Pairwise / singleton (Y == NULL) as wbx_pairwise; out is (nx, ny) f64. */
int wbx_dev_pairwise(const wbx_metric *metric, int dtype, const void *X,
                     int64_t nx, int64_t x_stride, int64_t tx, const void *Y,
                     int64_t ny, int64_t y_stride, int64_t ty, double *out,
                     void *stream);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): wbx_dev_pairwise with the same arguments and results, but its
 * staged inputs and per-pair costs live in a fixed per-GPU device array of
 * 64 MiB, so it calls no hipMalloc / hipFree / memset.  Sizes needing more
 * scratch return WBX_EINVAL; wdtw, ddtw, wddtw and FP32_ACC64 return
 * WBX_EUNSUPPORTED.  Calls that overlap on one GPU share its array and race. */
int wbx_dev_pairwise_noalloc(const wbx_metric *metric, int dtype, const void *X,
                             int64_t nx, int64_t x_stride, int64_t tx,
                             const void *Y, int64_t ny, int64_t y_stride,
                             int64_t ty, double *out, void *stream);

/* Paired as wbx_paired: out[i] = D(Y[i], X[i]). */
int wbx_dev_paired(const wbx_metric *metric, int dtype, const void *X,
                   int64_t x_stride, const void *Y, int64_t y_stride,
                   int64_t n, int64_t t, double *out, void *stream);

/* Argmin as wbx_argmin (heap order), device outputs idx (nx, k) int64 and
 * dist (nx, k) f64. */
int wbx_dev_argmin(const wbx_metric *metric, int dtype, const void *X,
                   int64_t nx, int64_t x_stride, int64_t tx, const void *Y,
                   int64_t ny, int64_t y_stride, int64_t ty, int64_t k,
                   const double *lower_bound, int64_t *idx, double *dist,
                   void *stream);

/* Subsequence scan as wbx_subsequence_pairwise; S is packed shapelets of
 * element type dtype, s_off / s_len are device int64 arrays. */
int wbx_dev_subsequence_pairwise(const wbx_metric *metric, int dtype,
                                 const void *S, const int64_t *s_off,
                                 const int64_t *s_len, int64_t n_s,
                                 const void *X, int64_t n, int64_t x_stride,
                                 int64_t t, double *dist, int64_t *idx,
                                 void *stream);

/* Device form of wbx_scaled_subsequence_pairwise: S packed patterns of
 * element type dtype; s_off / s_len (int64) and s_mean / s_std (float64) are
 * device arrays.  Synchronises `stream` (pattern lengths are read back). */
int wbx_dev_scaled_subsequence_pairwise(const wbx_metric *metric, int dtype,
                                        const void *S, const int64_t *s_off,
                                        const int64_t *s_len,
                                        const double *s_mean,
                                        const double *s_std, int64_t n_s,
                                        const void *X, int64_t n,
                                        int64_t x_stride, int64_t t,
                                        double *dist, int64_t *idx,
                                        void *stream);

/* Device form of wbx_scaled_subsequence_profile (prof device (n, nwp);
 * synchronises `stream`). */
int wbx_dev_scaled_subsequence_profile(const wbx_metric *metric, int dtype,
                                       const void *S, const int64_t *s_off,
                                       const int64_t *s_len,
                                       const double *s_mean,
                                       const double *s_std, int64_t n_s,
                                       int paired, const void *X, int64_t n,
                                       int64_t x_stride, int64_t t,
                                       int64_t nwp, double threshold,
                                       double *prof, void *stream);

/* Device form of wbx_scaled_argmin_subsequence (idx / dist device (n, k));
 * synchronises `stream`. */
int wbx_dev_scaled_argmin_subsequence(const wbx_metric *metric, int dtype,
                                      const void *S, const int64_t *s_off,
                                      const int64_t *s_len, const void *X,
                                      int64_t n, int64_t x_stride, int64_t t,
                                      int64_t k, int64_t *idx, double *dist,
                                      void *stream);

/* Device forms of wbx_argmin_subsequence (n patterns, one per series;
 * idx / dist device (n, k)) and wbx_subsequence_profile (prof device
 * (n, nwp)); S is packed patterns of element type dtype, s_off / s_len are
 * device int64 arrays.  Both synchronise `stream` (pattern lengths are read
 * back). */
int wbx_dev_argmin_subsequence(const wbx_metric *metric, int dtype,
                               const void *S, const int64_t *s_off,
                               const int64_t *s_len, const void *X, int64_t n,
                               int64_t x_stride, int64_t t, int64_t k,
                               int64_t *idx, double *dist, void *stream);
int wbx_dev_subsequence_profile(const wbx_metric *metric, int dtype,
                                const void *S, const int64_t *s_off,
                                const int64_t *s_len, int64_t n_s, int paired,
                                const void *X, int64_t n, int64_t x_stride,
                                int64_t t, int64_t nwp, double *prof,
                                void *stream);

/* ----------------------------------------------------------- lower bounds */

/* DTW envelope of n float64 device rows (dtw_envelop, distance/dtw.py:155-184;
 * _dtw_envelop + find_min_max, _elastic.pyx:1076-1092, 98-150): lower/upper
 * (n, t) packed = min/max over [k - r, k + r]; 0 <= r < t. */
int wbx_dev_dtw_envelope(const double *X, int64_t n, int64_t stride, int64_t t, int64_t r,
                         double *lower, double *upper, void *stream);
/* LB_Keogh of every X row against every envelope row (_dtw_lb_keogh,
 * _elastic.pyx:1095-1116; cumulative_bound :229-260): out (nx, ny). */
int wbx_dev_lb_keogh(const double *X, int64_t nx, int64_t x_stride, const double *lower,
                     const double *upper, int64_t ny, int64_t t, double *out, void *stream);
/* DtwKimLowerBound.transform (distance/lb.py:206-312): out[i, j] for query
 * Y[i] and fitted X[j], (ny, nx). */
int wbx_dev_lb_kim(const double *Y, int64_t ny, int64_t y_stride, const double *X, int64_t nx,
                   int64_t x_stride, int64_t t, double *out, void *stream);

/* ------------------------------------------------------- warping paths */

/* dtw_mapping's optimal path (distance/dtw.py:362-427, cells of
 * _dtw_alignment, _elastic.pyx:1011-1073) for the pairs (X row p, Y row p),
 * p < n, float64 on the device, warp size w (dtw.py:_compute_warp_size:
 * max(floor(max(tx, ty) * r), 1)).  A row-major path covers one contiguous
 * column range per row: rows[p * tx * 2 + 2 i] and [.. + 2 i + 1] are the
 * first and last column of row i.  x_stride 0 aligns one X row with every Y
 * row. */
int wbx_dev_dtw_path(const double *X, int64_t x_stride, int64_t tx, const double *Y,
                     int64_t y_stride, int64_t ty, int64_t n, int64_t w, const double *weights,
                     int32_t *rows, void *stream);
/* One step of the stochastic subgradient mean, _ssg_dtw_average
 * (distance/dtw.py:597-647, the body of its inner loop, :618-633): with
 * rows the path of (mean, x) from wbx_dev_dtw_path (per-row column ranges),
 * z[m] = sum over the path columns c of row m, in order, of
 * mean[m] - x[c] * weight, then mean[m] -= learning_rate * z[m], in place.
 * mean (tm values), rows (tm x 2 int32) and x (tx values) are device
 * pointers. */
int wbx_dev_dtw_ssg_update(double *mean, int64_t tm, const int32_t *rows,
                           const double *x, int64_t tx, double weight,
                           double learning_rate, void *stream);

/* One majorize-minimize step of _mm_dtw_average (distance/dtw.py:640-666):
 * aligns `mean` (tm values) with each X row (n rows of tx), then
 * mean_out[m] = sum over rows i and aligned columns x (in order) of
 * X[i, x] * w_i, divided by the summed weights w_i; sample_weight may be
 * NULL (all 1).  weights: the wdtw_alignment cell weights as in
 * wbx_dev_dtw_path, or NULL.  All pointers are device pointers. */
int wbx_dev_dtw_mm_update(const double *mean, int64_t tm, const double *X, int64_t x_stride,
                          int64_t tx, int64_t n, int64_t w, const double *weights,
                          const double *sample_weight, double *mean_out, void *stream);

/* ------------------------------------------- batched uneven pairs
 * No single reference driver: this is the batch form of DtwMetric._distance
 * (distance/_elastic.pyx:3167-3186) over pairs of different lengths, each with
 * its own r.  Series are padded to t_pad with the end-of-series marker of
 * utils/variable_len.py:17-115 (which the reference's distance API rejects,
 * utils/validation.py:574-577); out[p] = D(X[p][:len_x[p]], Y[p][:len_y[p]]). */

/* len[i] = index of the first end-of-series marker of row i (t if none). */
int wbx_dev_eos_lengths(int dtype, const void *X, int64_t n, int64_t stride,
                        int64_t t, int32_t *len, void *stream);

/* Host-side plan: lengths and per-pair r are host arrays (r_per_pair NULL =
 * metric->r for all).  policy 0 = fastest kernels, 1 = plain kernel only,
 * 2 = plain kernel walking every pair over the whole padded length (a
 * measurement baseline; same results).
 * The plan owns its device buffers; free it with wbx_batch_free. */
int wbx_batch_create(const wbx_metric *metric, int64_t n_pairs, int64_t t_pad,
                     const int32_t *len_x, const int32_t *len_y,
                     const double *r_per_pair, int policy, void **plan);
int wbx_batch_free(void *plan);
/* Launch shape of the strip kernel: at most waves_per_cu persistent waves
 * per CU (1..32; default 12 for float32, 8 for float64 strips).  n_streams
 * (1..4) is accepted and ignored: a run uses the caller's stream only.
 * Drops a captured graph.  The results do not depend on it. */
int wbx_batch_set_launch(void *plan, int waves_per_cu, int n_streams);
/* Autotuner: times `reps` strip launches alone on `stream` with the caller's
 * data for waves per CU in {4, 6, 8, 10, 12, 16} and keeps the fastest.
 * tuned_ms = the best time, or -1 when the choice came from the
 * process-wide cache of plans with the same signature, or when the plan has
 * no strip pairs.  Results do not depend on it.  Drops a captured graph. */
int wbx_batch_autotune(void *plan, int dtype, const void *X, int64_t x_stride,
                       const void *Y, int64_t y_stride, void *stream, int reps,
                       float *tuned_ms);
/* Strip size class k of the plan (the pairs of one strip size R, in queue
 * order, widest first): R, pairs, the strip launch's persistent waves, stream
 * index (always 0) and the tuned launch time (ms, -1 if not autotuned); k < 0
 * only reports the number of classes. */
int wbx_batch_bucket_info(void *plan, int k, int *n_buckets, int *R, int *pairs,
                          int *waves, int *stream, float *ms);
/* Algorithmic band cells of the plan and how many pairs each kernel takes. */
int wbx_batch_info(void *plan, int64_t *cells, int64_t *n_strip,
                   int64_t *n_plain);
/* Enqueue one run on `stream` (device pointers, dtype WBX_F32 / WBX_F64).
 * use_graph != 0: capture the run into a hipGraph the first time these
 * arguments are seen and replay the graph on later calls with the same
 * arguments.  out is n_pairs float64. */
int wbx_batch_run(void *plan, int dtype, const void *X, int64_t x_stride,
                  const void *Y, int64_t y_stride, double *out, void *stream,
                  int use_graph);

/* ------------------------------------------- differentiable DTW
 * No single reference driver: DTW distances with gradients along the optimal
 * path of dtw_mapping's backtrack (distance/dtw.py:362-427, ties diag > up >
 * left), for DTW barycenter averaging (the role of dtw_average,
 * distance/dtw.py:429-690) and the counterfactual explainers.  FP32, metric
 * dtw or wdtw (WeightedDtwMetric, distance/_elastic.pyx:3322-3341: weights
 * sized max(tx, ty), each path cell's term weighted as the forward weighs
 * it), all series of X of length tx and of Y of length ty.
 * mode bit 0 clear: all pairs, D is (nx, ny); set: D[p] = DTW(X[p], Y[p]),
 * nx == ny.  Mode bit 1 set: the forward keeps only the cost rows at
 * checkpoints 32 anti-diagonal steps apart and the backward recomputes each
 * band of steps from its checkpoint; WDTW and bands wider than 112 slots
 * always take that form, DTW bands of one slot never do.
 * The forward writes `state` (state_bytes) that the backward reads; the
 * backward needs `scratch` (scratch_bytes).  G is dL/dD in D's layout; dX
 * (nx, tx) and dY (ny, ty) are float32 with row strides ldx, ldy. */
int wbx_dtw_train_sizes(const wbx_metric *metric, int mode, int64_t nx,
                        int64_t ny, int64_t tx, int64_t ty,
                        int64_t *state_bytes, int64_t *scratch_bytes);
int wbx_dev_dtw_train_forward(const wbx_metric *metric, int mode,
                              const float *X, int64_t nx, int64_t x_stride,
                              int64_t tx, const float *Y, int64_t ny,
                              int64_t y_stride, int64_t ty, double *D,
                              void *state, void *stream);
int wbx_dev_dtw_train_backward(const wbx_metric *metric, int mode,
                               const float *X, int64_t nx, int64_t x_stride,
                               int64_t tx, const float *Y, int64_t ny,
                               int64_t y_stride, int64_t ty, const double *G,
                               const void *state, float *dX, int64_t ldx,
                               float *dY, int64_t ldy, void *scratch,
                               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* WBX_H */
