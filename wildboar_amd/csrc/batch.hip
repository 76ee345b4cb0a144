// batch.hip -- DTW over a batch of (x, y) pairs of uneven lengths.
//
// Inputs are two arrays of series padded to a common width with wildboar's
// end-of-series marker (utils/variable_len.py:17-115: a NaN whose float32 bit
// pattern ends in 0x9); pair p is (X[p][:lx_p], Y[p][:ly_p]) with its own
// window fraction r_p.  out[p] = DTW(X[p], Y[p]) with DtwMetric's band
// (distance/_elastic.pyx:3167-3186, w = _compute_r(min(lx, ly), r_p)).
//
// One run is, on the caller's stream only:
//   k_zero_i32     clears the queue counter,
//   (conversion)   float64 inputs to float32 (FP32) or float32 values
//                  (FP32_ACC64) into the plan's buffers,
//   k_stage_rows   copies the few pairs at the ends of the arrays whose reads
//                  could leave them into padded rows,
//   k_strip_batch  (strip.hip) one persistent launch over every strip pair,
//                  waves taking pairs from one queue sorted by strip size and
//                  cost, float32 or float64,
//   plain kernel   (plain.hip) the pairs no strip size covers, WDTW, and the
//                  "plain per-pair" baseline policies,
//   finish         sqrt into out.
// No event or second stream is involved, so a captured run is one chain.
#include <algorithm>
#include <map>
#include <math.h>
#include <mutex>
#include <numeric>
#include <string.h>
#include <type_traits>
#include <vector>

#include "strip.h"

namespace wbx {

// ------------------------------------------------------------------ EOS
template <typename T>
__device__ __forceinline__ bool is_eos(T v);
template <>
__device__ __forceinline__ bool is_eos<float>(float v) {
    return v != v && (__float_as_uint(v) & 0xFu) == 9u;
}
template <>
__device__ __forceinline__ bool is_eos<double>(double v) {
    // the reference tests the float32 image of the value
    return v != v && (__float_as_uint((float)v) & 0xFu) == 9u;
}

template <typename T>
__global__ __launch_bounds__(256) void k_eos_lengths(const T *__restrict__ X, int64_t n,
                                                     int64_t stride, int t,
                                                     int32_t *__restrict__ len) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const T *row = X + i * stride;
    int l = t;
    for (int c = 0; c < t; ++c)
        if (is_eos<T>(row[c])) {
            l = c;
            break;
        }
    len[i] = l;
}

template <typename T>
int launch_eos_lengths(const T *X, int64_t n, int64_t stride, int t, int32_t *len,
                       hipStream_t st) {
    if (n <= 0) return WBX_OK;
    hipLaunchKernelGGL(k_eos_lengths<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X,
                       n, stride, t, len);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}
template int launch_eos_lengths<float>(const float *, int64_t, int64_t, int, int32_t *,
                                       hipStream_t);
template int launch_eos_lengths<double>(const double *, int64_t, int64_t, int, int32_t *,
                                        hipStream_t);

// ------------------------------------------------------------------ plan
// Staged rows: x and y of pair pairs[e] to rows 2e, 2e + 1 of sb (padding
// untouched).
template <typename T>
__global__ __launch_bounds__(256) void k_stage_rows(const T *__restrict__ X, int64_t xs,
                                                    const T *__restrict__ Y, int64_t ys,
                                                    const int32_t *__restrict__ pairs, int t,
                                                    T *__restrict__ sb, int sstride, int spad) {
    const int e = blockIdx.y >> 1, side = blockIdx.y & 1;
    const int64_t q = pairs[e];
    const T *src = side ? Y + q * ys : X + q * xs;
    T *dst = sb + (int64_t)blockIdx.y * sstride + spad;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < t; j += gridDim.x * 256) dst[j] = src[j];
}

__global__ void k_zero_i32(int *p, int n) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Host-side plan: per-pair geometry, routing (strip kernel, or the plain
// kernel), the sorted queue, device buffers, and the graph captured from the
// last run.
struct BatchPlan {
    wbx_metric metric{};
    bool fp64 = false;  // strip pairs run the float64 kernel (FP64, FP32_ACC64)
    int64_t n_pairs = 0, t_pad = 0;
    int policy = 0;
    // per strip size: R, pairs (the queue is ordered by R, then cost)
    std::vector<int> class_R, class_count;
    // per register tier: queue range, waves per CU, persistent waves, tuned ms
    int tier_begin[kStripTiers] = {}, tier_count[kStripTiers] = {};
    int tier_wpc[kStripTiers] = {}, tier_waves[kStripTiers] = {};
    float tier_ms[kStripTiers] = {};
    int n_strip = 0;
    StripItem *d_items = nullptr;
    int *d_counter = nullptr;  // one per tier
    int64_t n_plain = 0;
    int32_t *d_plain_list = nullptr;
    int n_stage = 0, stage_stride = 0, stage_pad = 0;
    int32_t *d_stage_pairs = nullptr;
    void *d_stage = nullptr;
    void *d_zero = nullptr;  // zero buffer read by the lanes past a pair's band
    int zero_pad = 0;
    int64_t zero_len = 0;  // elements of each half of d_zero (zeros, then +inf)
    int32_t *d_lx = nullptr, *d_ly = nullptr;
    double *d_r = nullptr;
    void *d_sq = nullptr;        // strip pairs' squared costs (T)
    void *d_sq_plain = nullptr;  // plain pairs' squared costs (T)
    void *d_xconv = nullptr, *d_yconv = nullptr;
    // WDTW weights sized by t_pad, owned by the plan (a captured graph keeps
    // these pointers, so they live exactly as long as the plan)
    double *d_w64 = nullptr;
    float *d_w32 = nullptr;
    // completion of the last enqueued run or replay: recorded after every
    // run outside stream capture, waited on before anything a run reads or
    // writes (graph, buffers) is destroyed
    hipEvent_t ev_done = nullptr;
    bool pending = false;
    int64_t cells = 0, slots = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    const void *g_x = nullptr, *g_y = nullptr;
    double *g_out = nullptr;
    hipStream_t g_stream = nullptr;
    int g_dtype = -1;
    int64_t g_xs = 0, g_ys = 0;
    int n_cu = 256;
};

#define RC(x)                \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

template <typename T>
static int dmalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess)
        return fail(WBX_ENOMEM, "device allocation of %zu bytes failed", n * sizeof(T));
    return WBX_OK;
}

// Waits for the plan's last recorded run (it may still be executing the
// graph or reading the plan's buffers).
static void plan_wait(BatchPlan *b) {
    if (b->pending && b->ev_done) (void)hipEventSynchronize(b->ev_done);
    b->pending = false;
}

static void plan_drop_graph(BatchPlan *b) {
    plan_wait(b);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static void plan_free(BatchPlan *b) {
    if (!b) return;
    plan_drop_graph(b);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    void *ptrs[] = {b->d_items, b->d_counter, b->d_plain_list, b->d_lx,  b->d_ly,
                    b->d_r,     b->d_sq,      b->d_sq_plain,   b->d_xconv, b->d_yconv,
                    b->d_stage_pairs, b->d_stage, b->d_zero, b->d_w64, b->d_w32};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete b;
}

// Persistent waves of each tier's launch: at most its waves per CU per CU, at
// most one per pair of the tier.
static int tier_waves(const BatchPlan *b, int t, int wpc) {
    return (int)std::min<int64_t>(b->tier_count[t], (int64_t)b->n_cu * wpc);
}
static void plan_assign(BatchPlan *b) {
    for (int t = 0; t < kStripTiers; ++t) b->tier_waves[t] = tier_waves(b, t, b->tier_wpc[t]);
}
// Default launch shape: waves per CU by the tiers' register budgets.
static void plan_default_shape(BatchPlan *b) {
    for (int t = 0; t < kStripTiers; ++t) {
        b->tier_wpc[t] = b->fp64 ? 8 : 12;  // measured best on the uneven-pairs workload
        b->tier_ms[t] = -1.0f;
    }
}

static int plan_build(BatchPlan *b, const int32_t *lx, const int32_t *ly, const double *r) {
    const int64_t n = b->n_pairs;
    const bool dtw = b->metric.kind == WBX_DTW;
    std::vector<StripItem> items;
    std::vector<int64_t> cost;
    std::vector<int32_t> plain, staged;
    int64_t stage_hi = 0;
    int Rmax = 0;
    // the strip kernel addresses a dense copy of the arrays with 32-bit offsets
    const bool dense_ok = n * b->t_pad >= 4 && n * b->t_pad <= (int64_t)0x7fffffff;
    for (int64_t q = 0; q < n; ++q) {
        const int a = lx[q], c = ly[q];
        const double rq = r ? r[q] : b->metric.r;
        b->cells += band_cells(a, c, rq);
        // DTW is exactly symmetric: put the shorter series on the rows, which
        // keeps band slots inside the matrix.
        const bool swap = dtw && a > c;
        const int rx = swap ? c : a, ry = swap ? a : c;
        const Band bd = make_band(rx, ry, rq);
        int R = 0, nl = 0, al = 0;
        const bool strip = dtw && b->policy == 0 && rx >= 1 && ry >= 1 && dense_ok &&
                           strip_split_t(bd.B, b->fp64, &R, &nl, &al);
        if (!strip) {
            plain.push_back((int32_t)q);
            continue;
        }
        StripItem it{};
        it.pair = (int32_t)q;
        it.swap = swap ? 1 : 0;
        it.lx = rx;
        it.ly = ry;
        it.off = bd.off;
        it.B = bd.B;
        it.n = nl;
        it.a = al;
        it.R = R;
        it.stage = -1;
        if (!strip_reads_inside(q, n, b->t_pad, rx, bd.off, R)) {
            it.stage = (int32_t)staged.size();
            staged.push_back((int32_t)q);
            stage_hi = std::max(stage_hi, strip_read_hi(rx, bd.off, R));
        }
        items.push_back(it);
        // wave steps x VALU per lane step
        cost.push_back((int64_t)(rx + nl) * (int64_t)(3 * (R + 1) + 8));
        b->slots += (int64_t)rx * bd.B;
        Rmax = std::max(Rmax, R);
    }
    // Queue order: strip size descending (a wave's code stays the same size
    // class for long stretches, and the last pairs are the narrow, cheap
    // ones), then cost descending (longest first).
    std::vector<int> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int u, int v) {
        if (items[u].R != items[v].R) return items[u].R > items[v].R;
        if (cost[u] != cost[v]) return cost[u] > cost[v];
        return u < v;
    });
    std::vector<StripItem> sorted;
    sorted.reserve(items.size());
    for (int k : order) {
        const StripItem &it = items[k];
        if (b->class_R.empty() || b->class_R.back() != it.R) {
            b->class_R.push_back(it.R);
            b->class_count.push_back(0);
        }
        b->class_count.back()++;
        sorted.push_back(it);
    }
    b->n_strip = (int)sorted.size();
    b->n_plain = (int64_t)plain.size();
    for (int q = 0; q < b->n_strip; ++q) {
        const int t = strip_tier(sorted[q].R, b->fp64);
        if (b->tier_count[t] == 0) b->tier_begin[t] = q;
        b->tier_count[t]++;
    }
    plan_default_shape(b);
    plan_assign(b);
    const size_t esz = b->fp64 ? sizeof(double) : sizeof(float);
    RC(dmalloc(&b->d_items, sorted.size()));
    RC(dmalloc(&b->d_counter, kStripTiers));
    RC(dmalloc(&b->d_plain_list, plain.size()));
    RC(dmalloc(&b->d_lx, (size_t)n));
    RC(dmalloc(&b->d_ly, (size_t)n));
    RC(dmalloc(&b->d_r, (size_t)n));
    RC(dmalloc((char **)&b->d_sq, (size_t)n * esz));
    RC(dmalloc((char **)&b->d_sq_plain, std::max<size_t>(plain.size(), 1) * 8));
    // zero buffer: a pair's reads span [lo, hi] around its row start
    // (strip_read_lo / strip_read_hi with off <= t_pad, R <= Rmax)
    b->zero_pad = (int)(b->t_pad + 128);
    const int64_t zlen = b->zero_pad + strip_read_hi((int)b->t_pad, 0, std::max(Rmax, 2)) + 64;
    // zeros (x of the lanes past a band), then +inf (their y)
    b->zero_len = zlen;
    RC(dmalloc((char **)&b->d_zero, (size_t)(2 * zlen) * esz));
    {
        std::vector<unsigned char> zi((size_t)(2 * zlen) * esz, 0);
        for (int64_t q = zlen; q < 2 * zlen; ++q) {
            if (b->fp64) {
                const double v = WBX_INF_D;
                memcpy(&zi[(size_t)q * esz], &v, esz);
            } else {
                const float v = WBX_INF_F;
                memcpy(&zi[(size_t)q * esz], &v, esz);
            }
        }
        WBX_HIP_CHECK(hipMemcpy(b->d_zero, zi.data(), zi.size(), hipMemcpyHostToDevice));
    }
    if (!staged.empty()) {
        // a staged row has t_pad + 64 elements before it (-64 - off >= -64 - t_pad)
        // and reads reach stage_hi past its start
        const int64_t pad = (b->t_pad + 64 + 3) / 4 * 4;
        const int64_t stride = (pad + std::max<int64_t>(stage_hi + 1, b->t_pad) + 3) / 4 * 4;
        if (2 * (int64_t)staged.size() * stride + pad > (int64_t)0x7fffffff)
            return fail(WBX_EINVAL, "batch: staged rows exceed 32-bit offsets");
        b->n_stage = (int)staged.size();
        b->stage_stride = (int)stride;
        b->stage_pad = (int)pad;
        const size_t total = (size_t)(2 * staged.size() * stride + pad);
        RC(dmalloc(&b->d_stage_pairs, staged.size()));
        RC(dmalloc((char **)&b->d_stage, total * esz));
        WBX_HIP_CHECK(hipMemset(b->d_stage, 0, total * esz));
        WBX_HIP_CHECK(hipMemcpy(b->d_stage_pairs, staged.data(), sizeof(int32_t) * staged.size(),
                                hipMemcpyHostToDevice));
    }
    if (!sorted.empty())
        WBX_HIP_CHECK(hipMemcpy(b->d_items, sorted.data(), sizeof(StripItem) * sorted.size(),
                                hipMemcpyHostToDevice));
    if (!plain.empty())
        WBX_HIP_CHECK(hipMemcpy(b->d_plain_list, plain.data(), sizeof(int32_t) * plain.size(),
                                hipMemcpyHostToDevice));
    WBX_HIP_CHECK(hipMemcpy(b->d_lx, lx, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    WBX_HIP_CHECK(hipMemcpy(b->d_ly, ly, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    std::vector<double> hr((size_t)n);
    for (int64_t q = 0; q < n; ++q) hr[q] = r ? r[q] : b->metric.r;
    WBX_HIP_CHECK(hipMemcpy(b->d_r, hr.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    return WBX_OK;
}

// Inputs in the compute type C: float32 for FP32, float64 for FP64, float64
// holding float32 values for FP32_ACC64 (converted into the plan's buffers
// with dense rows when the caller's type or mode needs it).
template <typename C, typename T>
static int plan_inputs(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                       hipStream_t st, const C **Xc, int64_t *xsc, const C **Yc, int64_t *ysc) {
    const int64_t n = b->n_pairs, t = b->t_pad;
    *Xc = (const C *)X;
    *Yc = (const C *)Y;
    *xsc = xs;
    *ysc = ys;
    const bool round = std::is_same<C, double>::value && b->metric.precision == WBX_FP32_ACC64 &&
                       std::is_same<T, double>::value;
    if (round) {
        RC(launch_round_f32_rows((const double *)X, n, t, xs, (double *)b->d_xconv, t, st));
        RC(launch_round_f32_rows((const double *)Y, n, t, ys, (double *)b->d_yconv, t, st));
    } else if (!std::is_same<C, T>::value) {
        RC((launch_convert_rows<T, C>(X, n, t, xs, (C *)b->d_xconv, t, st)));
        RC((launch_convert_rows<T, C>(Y, n, t, ys, (C *)b->d_yconv, t, st)));
    } else {
        return WBX_OK;
    }
    *Xc = (const C *)b->d_xconv;
    *Yc = (const C *)b->d_yconv;
    *xsc = *ysc = t;
    return WBX_OK;
}

// Arrays the strip kernel reads: 32-bit element offsets (dense copies when
// the caller's strides need more) and the staged rows.
template <typename C>
static int plan_strip_arrays(BatchPlan *b, const C *Xc, int64_t xs, const C *Yc, int64_t ys,
                             hipStream_t st, StripArrays<C> *ar) {
    const int64_t n = b->n_pairs, t = b->t_pad;
    const C *xb = Xc, *yb = Yc;
    int64_t xst = xs, yst = ys;
    const int64_t lim = (int64_t)0x7fffffff;
    if ((n - 1) * xst + t > lim) {
        RC((launch_convert_rows<C, C>(Xc, n, t, xst, (C *)b->d_xconv, t, st)));
        xb = (const C *)b->d_xconv;
        xst = t;
    }
    if ((n - 1) * yst + t > lim) {
        RC((launch_convert_rows<C, C>(Yc, n, t, yst, (C *)b->d_yconv, t, st)));
        yb = (const C *)b->d_yconv;
        yst = t;
    }
    if (b->n_stage > 0) {
        const unsigned gx = (unsigned)std::min<int64_t>((t + 255) / 256, 64);
        hipLaunchKernelGGL(k_stage_rows<C>, dim3(gx, (unsigned)(2 * b->n_stage)), dim3(256), 0, st,
                           xb, xst, yb, yst, b->d_stage_pairs, (int)t, (C *)b->d_stage,
                           b->stage_stride, b->stage_pad);
        WBX_HIP_CHECK(hipGetLastError());
    }
    *ar = StripArrays<C>{xb,  yb, (const C *)b->d_stage, (const C *)b->d_zero,
                         (const C *)b->d_zero + b->zero_len, (int)xst, (int)yst,
                         b->stage_stride, b->stage_pad, b->zero_pad};
    return WBX_OK;
}

// Tier t's strip launch alone (its queue counter must be zero).
template <typename C>
static int plan_strip_launch(BatchPlan *b, const StripArrays<C> &ar, int t, int waves,
                             hipStream_t st) {
    return launch_strip_batch<C>(t, b->d_items + b->tier_begin[t], b->tier_count[t],
                                 b->d_counter + t, ar, (C *)b->d_sq, waves, st);
}

static void plan_zero_counters(BatchPlan *b, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, b->d_counter, kStripTiers);
}

// Everything one run enqueues on st (graph-capturable: kernels only, no
// memset, no event).
template <typename C, typename T>
static int plan_enqueue_c(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                          double *out, hipStream_t st, const C *w) {
    const int64_t n = b->n_pairs, t = b->t_pad;
    const C *Xc, *Yc;
    int64_t xsc, ysc;
    RC((plan_inputs<C, T>(b, X, xs, Y, ys, st, &Xc, &xsc, &Yc, &ysc)));
    if (b->n_strip > 0) {
        StripArrays<C> ar;
        RC(plan_strip_arrays<C>(b, Xc, xsc, Yc, ysc, st, &ar));
        plan_zero_counters(b, st);
        // widest strips first: each launch's queue is longest-first, so the
        // launches' tails are its cheapest pairs
        for (int t = 0; t < kStripTiers; ++t)
            if (b->tier_count[t] > 0) RC(plan_strip_launch<C>(b, ar, t, b->tier_waves[t], st));
        PairSpec all{PAIR_DIAG, n, n, nullptr, nullptr};
        RC(launch_finish<C>(all, (const C *)b->d_sq, out, 0, st));
    }
    if (b->n_plain > 0) {
        PairSpec pl{PAIR_LIST, b->n_plain, 0, b->d_plain_list, b->d_plain_list};
        Series<C> sx{Xc, xsc, (int)t, b->d_lx}, sy{Yc, ysc, (int)t, b->d_ly};
        C *sqp = (C *)b->d_sq_plain;
        RC(launch_plain_dtw<C>(pl, sx, sy, b->metric.r, w, sqp, st, b->d_r, b->policy == 2));
        RC(launch_finish<C>(pl, sqp, out, 0, st));
    }
    return WBX_OK;
}

template <typename T>
static int plan_enqueue(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                        double *out, hipStream_t st) {
    if (b->fp64)
        return plan_enqueue_c<double, T>(b, X, xs, Y, ys, out, st, b->d_w64);
    return plan_enqueue_c<float, T>(b, X, xs, Y, ys, out, st, b->d_w32);
}

}  // namespace wbx

using namespace wbx;

// ===================================================================== C ABI
extern "C" {

int wbx_dev_eos_lengths(int dtype, const void *X, int64_t n, int64_t stride, int64_t t,
                        int32_t *len, void *stream) {
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    if (n < 0 || t < 1) return fail(WBX_EINVAL, "bad sizes");
    hipStream_t st = (hipStream_t)stream;
    return dtype == WBX_F64 ? launch_eos_lengths<double>((const double *)X, n, stride, (int)t, len, st)
                            : launch_eos_lengths<float>((const float *)X, n, stride, (int)t, len, st);
}

int wbx_batch_create(const wbx_metric *metric, int64_t n_pairs, int64_t t_pad,
                     const int32_t *len_x, const int32_t *len_y, const double *r_per_pair,
                     int policy, void **out_plan) {
    if (!metric || !out_plan || !len_x || !len_y) return fail(WBX_EINVAL, "NULL argument");
    if (metric->kind != WBX_DTW && metric->kind != WBX_WDTW)
        return fail(WBX_EINVAL, "batched DTW needs metric dtw or wdtw");
    if (metric->precision < WBX_FP64 || metric->precision > WBX_FP32_ACC64)
        return fail(WBX_EINVAL, "unknown precision %d", metric->precision);
    if (n_pairs < 1 || t_pad < 1 || t_pad > (1 << 20)) return fail(WBX_EINVAL, "bad sizes");
    if (policy < 0 || policy > 2) return fail(WBX_EINVAL, "policy must be 0, 1 or 2");
    for (int64_t q = 0; q < n_pairs; ++q) {
        if (len_x[q] < 1 || len_x[q] > t_pad || len_y[q] < 1 || len_y[q] > t_pad)
            return fail(WBX_EINVAL, "pair %lld: lengths (%d, %d) outside [1, %lld]", (long long)q,
                        len_x[q], len_y[q], (long long)t_pad);
        const double rq = r_per_pair ? r_per_pair[q] : metric->r;
        if (!(rq >= 0.0 && rq <= 1.0))
            return fail(WBX_EINVAL, "pair %lld: r == %g, must be in [0, 1]", (long long)q, rq);
    }
    if (metric->kind == WBX_WDTW && !(metric->g >= 0.0))
        return fail(WBX_EINVAL, "g == %g, must be >= 0", metric->g);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    BatchPlan *b = new BatchPlan();
    b->metric = *metric;
    int32_t longest = 0;
    for (int64_t q = 0; q < n_pairs; ++q) longest = std::max(longest, std::max(len_x[q], len_y[q]));
    promote_long(&b->metric, longest);
    b->fp64 = b->metric.precision != WBX_FP32;
    b->n_pairs = n_pairs;
    b->t_pad = t_pad;
    b->policy = policy;
    b->n_cu = ctx->n_cu > 0 ? ctx->n_cu : 256;
    int rc = plan_build(b, len_x, len_y, r_per_pair);
    if (!rc && hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess)
        rc = fail(WBX_EDEVICE, "batch: event creation failed");
    if (!rc && metric->kind == WBX_WDTW) {
        // weights sized by the padded array width, as the reference sizes
        // them by the array's n_timestep
        rc = dmalloc(&b->d_w64, (size_t)t_pad);
        if (!rc) rc = dmalloc(&b->d_w32, (size_t)t_pad);
        if (!rc) rc = fill_weights(t_pad, metric->g, b->d_w64, b->d_w32, nullptr);
    }
    if (!rc) {
        const size_t conv = (size_t)n_pairs * (size_t)t_pad * 8;
        if (hipMalloc(&b->d_xconv, conv) != hipSuccess || hipMalloc(&b->d_yconv, conv) != hipSuccess)
            rc = fail(WBX_ENOMEM, "batch conversion buffers (%zu bytes) failed", 2 * conv);
    }
    if (rc) {
        plan_free(b);
        return rc;
    }
    *out_plan = b;
    return WBX_OK;
}

// n_streams is kept for the interface and ignored: the run uses the caller's
// stream only.
int wbx_batch_set_launch(void *plan, int waves_per_cu, int n_streams) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    if (waves_per_cu < 1 || waves_per_cu > 32)
        return fail(WBX_EINVAL, "waves_per_cu == %d, must be in [1, 32]", waves_per_cu);
    if (n_streams < 1 || n_streams > 4)
        return fail(WBX_EINVAL, "n_streams == %d, must be in [1, 4]", n_streams);
    for (int t = 0; t < kStripTiers; ++t) {
        b->tier_wpc[t] = waves_per_cu;
        b->tier_ms[t] = -1.0f;
    }
    plan_drop_graph(b);
    plan_assign(b);
    return WBX_OK;
}

}  // extern "C"

namespace wbx {

// Autotuner: the persistent waves per CU of each tier's launch, timed on the
// caller's data with that launch alone on the caller's stream; the fastest
// is kept.  Choices are cached for the process by the plan's signature
// (strip sizes and counts, precision, sizes, device), so plans of the same
// workload reuse them without timing.
struct TierTune {
    int wpc[kStripTiers];
    float ms[kStripTiers];
};
static std::mutex g_tune_mu;
static std::map<uint64_t, TierTune> g_tune_cache;

static uint64_t plan_signature(const BatchPlan *b, int device) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix((uint64_t)device);
    mix((uint64_t)b->fp64);
    mix((uint64_t)b->n_pairs);
    mix((uint64_t)b->t_pad);
    mix((uint64_t)b->slots);
    for (size_t k = 0; k < b->class_R.size(); ++k) {
        mix((uint64_t)b->class_R[k]);
        mix((uint64_t)b->class_count[k]);
    }
    return h;
}

template <typename C, typename T>
static int plan_autotune_c(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                           hipStream_t st, int reps) {
    static const int kWpc[] = {4, 6, 8, 10, 12, 16};
    const C *Xc, *Yc;
    int64_t xsc, ysc;
    RC((plan_inputs<C, T>(b, X, xs, Y, ys, st, &Xc, &xsc, &Yc, &ysc)));
    StripArrays<C> ar;
    RC(plan_strip_arrays<C>(b, Xc, xsc, Yc, ysc, st, &ar));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(WBX_EDEVICE, "autotune: event creation failed");
    }
    int rc = WBX_OK;
    for (int t = 0; t < kStripTiers && !rc; ++t) {
        if (b->tier_count[t] == 0) continue;
        int best_w = b->tier_wpc[t], last_waves = -1;
        float best = -1.0f;
        for (int w : kWpc) {
            const int waves = tier_waves(b, t, w);
            if (waves == last_waves) continue;  // fewer pairs than this fills
            last_waves = waves;
            auto launch = [&]() -> int {
                plan_zero_counters(b, st);
                return plan_strip_launch<C>(b, ar, t, waves, st);
            };
            rc = launch();  // warm-up
            if (rc) break;
            (void)hipEventRecord(e0, st);
            for (int q = 0; q < reps && !rc; ++q) rc = launch();
            (void)hipEventRecord(e1, st);
            if (rc) break;
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(WBX_EDEVICE, "autotune: run failed");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)reps;
            if (best < 0.0f || ms < best) {
                best = ms;
                best_w = w;
            }
        }
        b->tier_wpc[t] = best_w;
        b->tier_ms[t] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace wbx

extern "C" {

int wbx_batch_autotune(void *plan, int dtype, const void *X, int64_t x_stride, const void *Y,
                       int64_t y_stride, void *stream, int reps, float *tuned_ms) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    if (!X || !Y) return fail(WBX_EINVAL, "NULL array");
    if (dtype != WBX_F32 && dtype != WBX_F64) return fail(WBX_EINVAL, "unknown dtype %d", dtype);
    if (x_stride < b->t_pad || y_stride < b->t_pad)
        return fail(WBX_EINVAL, "row strides (%lld, %lld) below t_pad %lld", (long long)x_stride,
                    (long long)y_stride, (long long)b->t_pad);
    if (reps < 1 || reps > 100) return fail(WBX_EINVAL, "reps == %d, must be in [1, 100]", reps);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    float total = -1.0f;
    if (b->n_strip > 0) {
        plan_drop_graph(b);  // the timing runs reuse the plan's buffers
        const uint64_t key = plan_signature(b, ctx->device);
        bool cached = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto f = g_tune_cache.find(key);
            if (f != g_tune_cache.end()) {
                for (int t = 0; t < kStripTiers; ++t) {
                    b->tier_wpc[t] = f->second.wpc[t];
                    b->tier_ms[t] = f->second.ms[t];
                }
                cached = true;
            }
        }
        if (!cached) {
            hipStream_t st = (hipStream_t)stream;
            int rc;
            if (b->fp64)
                rc = dtype == WBX_F64
                         ? plan_autotune_c<double, double>(b, (const double *)X, x_stride,
                                                           (const double *)Y, y_stride, st, reps)
                         : plan_autotune_c<double, float>(b, (const float *)X, x_stride,
                                                          (const float *)Y, y_stride, st, reps);
            else
                rc = dtype == WBX_F64
                         ? plan_autotune_c<float, double>(b, (const double *)X, x_stride,
                                                          (const double *)Y, y_stride, st, reps)
                         : plan_autotune_c<float, float>(b, (const float *)X, x_stride,
                                                         (const float *)Y, y_stride, st, reps);
            if (rc) return rc;
            total = 0.0f;
            TierTune tt{};
            for (int t = 0; t < kStripTiers; ++t) {
                if (b->tier_ms[t] > 0.0f) total += b->tier_ms[t];
                tt.wpc[t] = b->tier_wpc[t];
                tt.ms[t] = b->tier_ms[t];
            }
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[key] = tt;
        }
        plan_assign(b);
    }
    if (tuned_ms) *tuned_ms = total;
    return WBX_OK;
}

// Strip size class k of the plan: R, pairs, the persistent waves of its
// tier's launch, stream index (always 0: the launches run on the caller's
// stream) and the tuned time of its tier's launch (ms, -1 if not autotuned).
int wbx_batch_bucket_info(void *plan, int k, int *n_buckets, int *R, int *pairs, int *waves,
                          int *stream, float *ms) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    const int nb = (int)b->class_R.size();
    if (n_buckets) *n_buckets = nb;
    if (k < 0) return WBX_OK;
    if (k >= nb) return fail(WBX_EINVAL, "bucket %d of %d", k, nb);
    if (R) *R = b->class_R[k];
    if (pairs) *pairs = b->class_count[k];
    const int t = strip_tier(b->class_R[k], b->fp64);
    if (waves) *waves = b->tier_waves[t];
    if (stream) *stream = 0;
    if (ms) *ms = b->tier_ms[t];
    return WBX_OK;
}

int wbx_batch_free(void *plan) {
    plan_free((BatchPlan *)plan);
    return WBX_OK;
}

int wbx_batch_info(void *plan, int64_t *cells, int64_t *n_strip, int64_t *n_plain) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    if (cells) *cells = b->cells;
    if (n_strip) *n_strip = b->n_strip;
    if (n_plain) *n_plain = b->n_plain;
    return WBX_OK;
}

// Enqueue one run on `stream`.  use_graph: capture the run into a hipGraph
// the first time these (X, Y, out, stream, dtype) are seen and replay it on
// later calls with the same arguments.
int wbx_batch_run(void *plan, int dtype, const void *X, int64_t x_stride, const void *Y,
                  int64_t y_stride, double *out, void *stream, int use_graph) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    if (!X || !Y || !out) return fail(WBX_EINVAL, "NULL array");
    if (dtype != WBX_F32 && dtype != WBX_F64) return fail(WBX_EINVAL, "unknown dtype %d", dtype);
    if (x_stride < b->t_pad || y_stride < b->t_pad)
        return fail(WBX_EINVAL, "row strides (%lld, %lld) below t_pad %lld", (long long)x_stride,
                    (long long)y_stride, (long long)b->t_pad);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = (hipStream_t)stream;
    // Marks the run for plan_wait; a run enqueued into a caller's capture
    // has no completion this plan can wait on.
    auto mark_done = [&]() -> int {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        WBX_HIP_CHECK(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone) return WBX_OK;
        WBX_HIP_CHECK(hipEventRecord(b->ev_done, st));
        b->pending = true;
        return WBX_OK;
    };
    auto enqueue = [&](hipStream_t s) -> int {
        return dtype == WBX_F64
                   ? plan_enqueue<double>(b, (const double *)X, x_stride, (const double *)Y,
                                          y_stride, out, s)
                   : plan_enqueue<float>(b, (const float *)X, x_stride, (const float *)Y, y_stride,
                                         out, s);
    };
    if (!use_graph) {
        RC(enqueue(st));
        return mark_done();
    }
    if (st == nullptr)
        return fail(WBX_EINVAL, "graph capture needs a stream other than the null stream");
    const bool same = b->exec && b->g_x == X && b->g_y == Y && b->g_out == out &&
                      b->g_stream == st && b->g_dtype == dtype && b->g_xs == x_stride &&
                      b->g_ys == y_stride;
    if (!same) {
        plan_drop_graph(b);
        WBX_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(WBX_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
        b->graph = g;
        WBX_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->g_x = X;
        b->g_y = Y;
        b->g_out = out;
        b->g_stream = st;
        b->g_dtype = dtype;
        b->g_xs = x_stride;
        b->g_ys = y_stride;
    }
    WBX_HIP_CHECK(hipGraphLaunch(b->exec, st));
    return mark_done();
}

}  // extern "C"
