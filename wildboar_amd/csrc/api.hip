// api.hip -- the extern "C" entry points declared in include/wbx.h.
//
// Dispatch per request:
//   FP64 (parity)  euclidean -> exact tiled sum; dtw/wdtw -> plain kernel.
//   FP32           euclidean -> FP32 tiled sum; dtw -> lane-per-pair register
//                  kernel when the band fits (B <= 112), plain FP32 kernel
//                  otherwise; wdtw -> plain FP32 kernel.
//   argmin FP32    FP32 distances -> certified candidate set -> exact FP64
//                  re-evaluation of the candidates -> heap replay, so indices
//                  and distances equal the reference's.
#include <algorithm>
#include <atomic>
#include <initializer_list>
#include <math.h>
#include <mutex>
#include <stdlib.h>
#include <string.h>
#include <type_traits>
#include <vector>

#include "subseq_fast.h"
#include "subseq_mfma.h"
#include "wbx_internal.h"

namespace wbx {
const char *last_error();
}

using namespace wbx;

namespace {

std::atomic<int> g_policy{0};  // 0 fastest kernels, 1 plain kernel only, 2 group kernel preferred
std::mutex g_dev_mu[64];

struct DevGuard {
    std::unique_lock<std::mutex> lk;
    explicit DevGuard(int dev) : lk(g_dev_mu[dev & 63]) {}
};

// ------------------------------------------------------ allocation-free scratch
// Workspace of wbx_dev_pairwise_noalloc: a fixed array in device memory, one
// copy per GPU the library runs on (held until the process ends).  Host code
// never uses the array's name as an address: each device's copy lives at its
// own address, found once per device with hipGetSymbolAddress (with that
// device current) and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};

// Address of the current device's scratch (device `dev` must be current).
int scratch_setup(int dev, void **out) {
    if (dev < 0 || dev >= 64) return fail(WBX_EUNSUPPORTED, "device %d: at most 64 devices", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(WBX_EDEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    if (out) *out = g_scratch_ptr[dev];
    return WBX_OK;
}

// Carves one workspace reservation into aligned regions.
struct Carve {
    size_t total = 0;
    size_t take(size_t bytes) {
        size_t o = total;
        total += (bytes + 255) & ~(size_t)255;
        return o;
    }
};

int check_metric(const wbx_metric *m) {
    if (!m) return fail(WBX_EINVAL, "metric is NULL");
    if (m->kind < WBX_EUCLIDEAN || m->kind > WBX_ADTW)
        return fail(WBX_EINVAL, "unknown metric kind %d", m->kind);
    if (m->precision < WBX_FP64 || m->precision > WBX_FP32_ACC64)
        return fail(WBX_EINVAL, "unknown precision %d", m->precision);
    if (m->kind != WBX_EUCLIDEAN && !(m->r >= 0.0 && m->r <= 1.0))
        return fail(WBX_EINVAL, "r == %g, must be in [0, 1]", m->r);
    // WeightedDtwMetric checks g >= 0 (_elastic.pyx:3326); the derivative
    // variant takes any float (:3406-3408)
    if (m->kind == WBX_WDTW && !(m->g >= 0.0))
        return fail(WBX_EINVAL, "g == %g, must be >= 0", m->g);
    // -inf would meet +inf (a cell outside the band) as NaN
    if (m->kind == WBX_ADTW && !(m->p > -WBX_INF_D))
        return fail(WBX_EINVAL, "p == %g, must be a number above -inf", m->p);
    return WBX_OK;
}

// Every metric kind is served for subsequences.
int check_subsequence_metric(const wbx_metric *m, bool scan = false) {
    (void)scan;
    return check_metric(m);
}

hipStream_t pick_stream(DeviceCtx *ctx, void *stream) {
    (void)ctx;
    return (hipStream_t)stream;  // NULL is the default (null) stream
}

// FP32_ACC64 (float32 data, float64 accumulation): float64 inputs are rounded
// to float32 once, into the input arena, and every call then runs the float64
// recursion (float32 inputs already hold float32 values).  The metric handed
// on has precision FP64.  Call after begin(): the arena follows its ordering.
struct InArray {
    const void **ptr;
    int64_t n, len;
    int64_t *stride;
};

int acc64_prepare(DeviceCtx *ctx, wbx_metric *m, int dtype, std::initializer_list<InArray> arrays,
                  hipStream_t st) {
    if (m->precision != WBX_FP32_ACC64) return WBX_OK;
    m->precision = WBX_FP64;
    if (dtype != WBX_F64) return WBX_OK;
    size_t total = 0;
    for (const InArray &a : arrays)
        if (a.ptr && *a.ptr) total += (size_t)a.n * (size_t)a.len;
    void *wsp = nullptr;
    int rc = ws_in_reserve(ctx, total * sizeof(double), &wsp);
    if (rc) return rc;
    double *dst = (double *)wsp;
    for (const InArray &a : arrays) {
        if (!a.ptr || !*a.ptr) continue;
        rc = launch_round_f32_rows((const double *)*a.ptr, a.n, a.len, *a.stride, dst, a.len, st);
        if (rc) return rc;
        *a.ptr = dst;
        *a.stride = a.len;
        dst += (size_t)a.n * (size_t)a.len;
    }
    return WBX_OK;
}

// DDTW / WDDTW (DerivativeDtwMetric, distance/_elastic.pyx:3228-3320): each
// input becomes its average slope (float64, from the values as given, or as
// rounded by acc64_prepare), of length len - 2, and the metric becomes DTW /
// WDTW on those series (WDDTW's weights are sized by the slope lengths,
// :3415-3430).  The caller's element type becomes float64.  Needs every
// length >= 3.  Call after acc64_prepare.
struct DvArray {
    const void **ptr;
    int64_t n;
    int64_t *len;
    int64_t *stride;
};

int deriv_prepare(DeviceCtx *ctx, wbx_metric *m, int *dtype, std::initializer_list<DvArray> arrays,
                  hipStream_t st) {
    if (m->kind != WBX_DDTW && m->kind != WBX_WDDTW) return WBX_OK;
    m->kind = m->kind == WBX_DDTW ? WBX_DTW : WBX_WDTW;
    size_t total = 0;
    for (const DvArray &a : arrays)
        if (a.ptr && *a.ptr) {
            if (*a.len < 3) return fail(WBX_EINVAL, "derivative metrics need length >= 3");
            total += (size_t)a.n * (size_t)(*a.len - 2);
        }
    void *wsp = nullptr;
    int rc = ws_dv_reserve(ctx, total * sizeof(double), &wsp);
    if (rc) return rc;
    double *dst = (double *)wsp;
    for (const DvArray &a : arrays) {
        if (!a.ptr || !*a.ptr) continue;
        rc = *dtype == WBX_F64
                 ? launch_slope_rows<double>((const double *)*a.ptr, a.n, *a.len, *a.stride, dst, st)
                 : launch_slope_rows<float>((const float *)*a.ptr, a.n, *a.len, *a.stride, dst, st);
        if (rc) return rc;
        *a.ptr = dst;
        *a.len -= 2;
        *a.stride = *a.len;
        dst += (size_t)a.n * (size_t)*a.len;
    }
    *dtype = WBX_F64;
    return WBX_OK;
}

// Orders this call after the arena's previous user (possibly on another stream).
int begin(DeviceCtx *ctx, hipStream_t st) {
    WBX_HIP_CHECK(hipStreamWaitEvent(st, ctx->last, 0));
    return WBX_OK;
}
int end(DeviceCtx *ctx, hipStream_t st) {
    WBX_HIP_CHECK(hipEventRecord(ctx->last, st));
    return WBX_OK;
}

template <typename T>
Series<T> series(const T *base, int64_t stride, int64_t len) {
    Series<T> s;
    s.base = base;
    s.stride = stride;
    s.len = (int)len;
    s.lens = nullptr;
    return s;
}

// Optional timing of the distance kernels: HIP events recorded on the launch
// stream around each one, summed on request (wbx_prof_take).
struct ProfState {
    std::mutex mu;
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<int> kind;
} g_prof;

// Set while the allocation-free entry point runs on this thread: no timing
// events are created then.
thread_local bool g_no_alloc = false;

struct ProfScope {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    int k;
    ProfScope(hipStream_t s, int kind_) : st(s), k(kind_) {
        if (g_no_alloc) return;
        std::lock_guard<std::mutex> lk(g_prof.mu);
        if (!g_prof.on) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        (void)hipEventRecord(a, st);
    }
    ~ProfScope() {
        if (!a) return;
        (void)hipEventRecord(b, st);
        std::lock_guard<std::mutex> lk(g_prof.mu);
        g_prof.ev.push_back({a, b});
        g_prof.kind.push_back(k);
    }
};

enum { KERN_EUCLID = 1, KERN_LANE = 2, KERN_PLAIN = 3, KERN_BATCH = 4, KERN_SUBSEQ = 5,
       KERN_GROUP = 6, KERN_FULL = 7 };

// The lane kernels multiply +inf (columns outside a series) by the weights,
// so they take WDTW only when every cached weight (float64 or float32) is > 0.
bool lane64_ok(const DeviceCtx *ctx, const void *weights) {
    return !weights || cached_weights_min(ctx, weights) > 0.0;
}

// Squared costs of the pairs of ps.  m->kind is EUCLIDEAN, DTW, WDTW or ADTW
// (the derivative kinds arrive here as DTW / WDTW on slopes); w_fixed > 0
// replaces every pair's band width.
template <typename T>
int run_pairs(const wbx_metric *m, const PairSpec &ps, Series<T> xs, Series<T> ys,
              const T *weights, T *sq, hipStream_t st, int w_fixed = 0) {
    if (m->kind == WBX_EUCLIDEAN) {
        ProfScope pf(st, KERN_EUCLID);
        return launch_euclid_pairs<T>(ps, xs, ys, sq, st);
    }
    const double *pen = m->kind == WBX_ADTW ? &m->p : nullptr;
    const int policy = g_policy.load();
    // Register-band kernels need one geometry, and every weight > 0 (a zero
    // weight would meet the +inf of a column outside the series as NaN).
    if (policy != 1 && !xs.lens && !ys.lens && (!weights || lane64_ok(ctx_current(), weights))) {
        const Band bd = make_band(xs.len, ys.len, m->r, w_fixed);
        const int B = bd.B;
        const bool wt = weights != nullptr;
        // the band is the whole matrix (w >= min(lx, ly), r = 1): lanes own
        // columns, no slot outside the matrix
        FullGeo fg{};
        if (!wt && !pen && bd.w >= (xs.len < ys.len ? xs.len : ys.len) && policy != 2 &&
            full_plan(xs.len, ys.len, !std::is_same<T, float>::value, &fg)) {
            ProfScope pf(st, KERN_FULL);
            return launch_full_dtw<T>(ps, xs, ys, sq, st);
        }
        // one pair per lane while that is faster than a group of lanes per
        // pair (measured on MI355X, 1000 x 1000 x 256: float32 DTW lanes up
        // to 112 slots; float32 WDTW and float64 lanes up to 40 slots -- at
        // 49 slots the group kernel is 1.1-1.4x faster, at 23 the lanes
        // 1.6-1.8x); ADTW stays on the lanes (no group form)
        bool lane;
        if constexpr (std::is_same<T, float>::value)
            lane = lane_supported(xs.len, ys.len, m->r, wt, w_fixed) && (!wt || B <= 40);
        else
            lane = lane64_pairs_class(B, wt) != 0 && (B <= 40 || pen);
        GroupGeo gg{};
        const bool group = !pen && group_plan(B, bd.lx, !std::is_same<T, float>::value, wt, &gg);
        if (policy == 2 && group) lane = false;  // group kernel wherever it applies
        if (lane) {
            ProfScope pf(st, KERN_LANE);
            if constexpr (std::is_same<T, float>::value)
                return launch_lane_dtw(ps, xs, ys, m->r, weights, sq, st, w_fixed, pen);
            else
                return launch_lane64_pairs(ps, xs, ys, m->r, weights, sq, st, w_fixed, pen);
        }
        // wider bands and the remaining weighted ones: a group of lanes per pair
        if (group) {
            ProfScope pf(st, KERN_GROUP);
            const int nw = xs.len > ys.len ? xs.len : ys.len;
            return launch_group_dtw<T>(ps, xs, ys, m->r, weights, nw, sq, st, w_fixed);
        }
    }
    ProfScope pf(st, KERN_PLAIN);
    return launch_plain_dtw<T>(ps, xs, ys, m->r, weights, sq, st, nullptr, false, w_fixed, pen);
}

// Inputs of element type `dtype` brought to the compute type T (dense copy
// when a conversion is needed, the caller's memory otherwise).
template <typename T>
struct Staged {
    const T *ptr;
    int64_t stride;
};

template <typename T>
int stage(int dtype, const void *src, int64_t n, int64_t len, int64_t stride, char *ws,
          size_t off, Staged<T> *out, hipStream_t st) {
    const bool same = (dtype == WBX_F64) == std::is_same<T, double>::value;
    if (same) {
        out->ptr = (const T *)src;
        out->stride = stride;
        return WBX_OK;
    }
    T *dst = (T *)(ws + off);
    int rc = (dtype == WBX_F64)
                 ? launch_convert_rows<double, T>((const double *)src, n, len, stride, dst, len, st)
                 : launch_convert_rows<float, T>((const float *)src, n, len, stride, dst, len, st);
    out->ptr = dst;
    out->stride = len;
    return rc;
}

size_t stage_bytes(int dtype, bool want_double, int64_t n, int64_t len) {
    const bool same = (dtype == WBX_F64) == want_double;
    return same ? 0 : (size_t)n * (size_t)len * (want_double ? 8 : 4);
}

#define RC(x)              \
    do {                   \
        int _rc = (x);     \
        if (_rc) return _rc; \
    } while (0)

// -------------------------------------------------------------- pairwise
// Workspace of one pairwise call: staged inputs (when the element type is
// not the compute type) and the squared cost of every pair.
struct PairwiseLayout {
    PairSpec ps{};
    size_t o_x = 0, o_y = 0, o_sq = 0, total = 0;
};

template <typename T>
PairwiseLayout pairwise_layout(int dtype, int64_t nx, int64_t tx, int64_t ny, int64_t ty,
                               bool singleton) {
    PairwiseLayout L;
    if (singleton) {
        L.ps.mode = PAIR_TRIANGLE;
        L.ps.ny = nx;
        L.ps.n_pairs = nx * (nx - 1) / 2;
    } else {
        L.ps.mode = PAIR_GRID;
        L.ps.ny = ny;
        L.ps.n_pairs = nx * ny;
    }
    const bool dbl = std::is_same<T, double>::value;
    Carve cv;
    L.o_x = cv.take(stage_bytes(dtype, dbl, nx, tx));
    L.o_y = singleton ? 0 : cv.take(stage_bytes(dtype, dbl, ny, ty));
    L.o_sq = cv.take((size_t)L.ps.n_pairs * sizeof(T));
    L.total = cv.total;
    return L;
}

// Everything one pairwise call enqueues, in the workspace `ws` of L.total
// bytes (kernels only: no allocation, no memset).
template <typename T>
int pairwise_run(const PairwiseLayout &L, const wbx_metric *m, int dtype, const void *X,
                 int64_t nx, int64_t xstr, int64_t tx, const void *Y, int64_t ystr, int64_t ty,
                 bool singleton, const T *w, char *ws, double *out, hipStream_t st, int w_fixed) {
    Staged<T> sx, sy;
    RC(stage<T>(dtype, X, nx, tx, xstr, ws, L.o_x, &sx, st));
    if (singleton)
        sy = sx;
    else
        RC(stage<T>(dtype, Y, L.ps.ny, ty, ystr, ws, L.o_y, &sy, st));
    T *sq = (T *)(ws + L.o_sq);
    RC(run_pairs<T>(m, L.ps, series<T>(sx.ptr, sx.stride, tx), series<T>(sy.ptr, sy.stride, ty), w,
                    sq, st, w_fixed));
    if (singleton) RC(launch_fill_f64(out, nx * nx, 0.0, st));
    return launch_finish<T>(L.ps, sq, out, singleton ? nx : L.ps.ny, st);
}

template <typename T>
int pairwise_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *X, int64_t nx,
                  int64_t xstr, int64_t tx, const void *Y, int64_t ny, int64_t ystr, int64_t ty,
                  bool singleton, double *out, hipStream_t st, int w_fixed) {
    const PairwiseLayout L = pairwise_layout<T>(dtype, nx, tx, ny, ty, singleton);
    void *wsp = nullptr;
    RC(ws_reserve(ctx, L.total, &wsp));
    const T *w = nullptr;
    if (m->kind == WBX_WDTW) {
        const double *w64;
        const float *w32;
        RC(cached_weights(ctx, tx > ty ? tx : ty, m->g, &w64, &w32));
        w = std::is_same<T, double>::value ? (const T *)w64 : (const T *)w32;
    }
    return pairwise_run<T>(L, m, dtype, X, nx, xstr, tx, Y, ystr, ty, singleton, w, (char *)wsp,
                           out, st, w_fixed);
}

// -------------------------------------------------------------- argmin
// argmin with a lower bound: the float64 distances of the pairs the
// reference evaluates (prune.hip), then the heap replay with its skip rule.
int argmin_pruned(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *X, int64_t nx,
                  int64_t xstr, int64_t tx, const void *Y, int64_t ny, int64_t ystr, int64_t ty,
                  int64_t k, const double *lb, int64_t *idx, double *dist, hipStream_t st) {
    const int64_t np = nx * ny;
    Carve cv;
    const size_t o_x64 = cv.take(stage_bytes(dtype, true, nx, tx));
    const size_t o_y64 = cv.take(stage_bytes(dtype, true, ny, ty));
    const size_t o_d64 = cv.take((size_t)np * 8);
    const size_t o_heap = cv.take((size_t)nx * (size_t)k * 16);
    const size_t o_pa = cv.take((size_t)np * 4);
    const size_t o_pb = cv.take((size_t)np * 4);
    const size_t o_sq = cv.take((size_t)np * 8);
    const size_t o_cnt = cv.take(8);
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, y64;
    RC(stage<double>(dtype, X, nx, tx, xstr, ws, o_x64, &x64, st));
    RC(stage<double>(dtype, Y, ny, ty, ystr, ws, o_y64, &y64, st));
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (m->kind == WBX_WDTW) RC(cached_weights(ctx, tx > ty ? tx : ty, m->g, &w64, &w32));
    double *D64 = (double *)(ws + o_d64);
    int32_t *pa = (int32_t *)(ws + o_pa), *pb = (int32_t *)(ws + o_pb);
    double *sq = (double *)(ws + o_sq);
    unsigned long long *cnt = (unsigned long long *)(ws + o_cnt);
    wbx_metric m64 = *m;
    m64.precision = WBX_FP64;
    RC(launch_fill_f64(D64, np, __builtin_nan(""), st));  // NaN: not computed
    // pass 0: the seeds; pass 1: the pairs the seeds cannot exclude
    for (int pass = 0; pass < 2; ++pass) {
        RC(launch_fill_i32((int *)cnt, 2, 0, st));
        RC(pass == 0 ? launch_prune_seed(lb, nx, ny, k, pa, pb, cnt, st)
                     : launch_prune_scan(D64, lb, nx, ny, k, pa, pb, cnt, st));
        unsigned long long total = 0;
        WBX_HIP_CHECK(hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, st));
        WBX_HIP_CHECK(hipStreamSynchronize(st));
        ctx->prune_evaluated += (int64_t)total;
        if (total == 0) continue;
        PairSpec pl{PAIR_LIST, (int64_t)total, ny, pa, pb};
        RC(run_pairs<double>(&m64, pl, series<double>(x64.ptr, x64.stride, tx),
                             series<double>(y64.ptr, y64.stride, ty), w64, sq, st));
        RC(launch_finish<double>(pl, sq, D64, ny, st));
    }
    return launch_argmin_replay(D64, nx, ny, k, lb, ws + o_heap, idx, dist, st);
}

int argmin_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *X, int64_t nx,
                int64_t xstr, int64_t tx, const void *Y, int64_t ny, int64_t ystr, int64_t ty,
                int64_t k, const double *lb, int64_t *idx, double *dist, hipStream_t st) {
    const int64_t np = nx * ny;
    // a lower bound: only the distances the reference evaluates (exact, float64)
    if (lb && k <= prune_k_max() && g_policy.load() != 1)
        return argmin_pruned(ctx, m, dtype, X, nx, xstr, tx, Y, ny, ystr, ty, k, lb, idx, dist, st);
    // the certified FP32 candidate bound covers the DTW recursion; ADTW runs
    // in float64 throughout.  The candidates are the pairs the sequence
    // without a lower bound pushes; with a bound the reference pushes others
    // (the first k it evaluates), so bounded calls take the float64 matrix.
    const bool certified = m->precision == WBX_FP32 && k <= 32 && m->kind != WBX_ADTW && !lb;
    Carve cv;
    const size_t o_x64 = cv.take(stage_bytes(dtype, true, nx, tx));
    const size_t o_y64 = cv.take(stage_bytes(dtype, true, ny, ty));
    const size_t o_d64 = cv.take((size_t)np * 8);
    const size_t o_heap = cv.take((size_t)nx * (size_t)k * 16);
    size_t o_x32 = 0, o_y32 = 0, o_sq32 = 0, o_cnt = 0, o_off = 0, o_pa = 0, o_pb = 0, o_sq64 = 0,
           o_mx = 0;
    if (certified) {
        o_x32 = cv.take(stage_bytes(dtype, false, nx, tx));
        o_y32 = cv.take(stage_bytes(dtype, false, ny, ty));
        o_sq32 = cv.take((size_t)np * 4);
        o_cnt = cv.take((size_t)nx * 8);
        o_off = cv.take((size_t)(nx + 1) * 8);
        o_pa = cv.take((size_t)np * 4);
        o_pb = cv.take((size_t)np * 4);
        o_mx = cv.take(8);
    } else {
        o_sq64 = cv.take((size_t)np * 8);
    }
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, y64;
    RC(stage<double>(dtype, X, nx, tx, xstr, ws, o_x64, &x64, st));
    RC(stage<double>(dtype, Y, ny, ty, ystr, ws, o_y64, &y64, st));
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (m->kind == WBX_WDTW) RC(cached_weights(ctx, tx > ty ? tx : ty, m->g, &w64, &w32));
    double *D64 = (double *)(ws + o_d64);
    if (!certified) {
        PairSpec ps{PAIR_GRID, np, ny, nullptr, nullptr};
        double *sq = (double *)(ws + o_sq64);
        wbx_metric m64 = *m;
        m64.precision = WBX_FP64;
        RC(run_pairs<double>(&m64, ps, series<double>(x64.ptr, x64.stride, tx),
                             series<double>(y64.ptr, y64.stride, ty), w64, sq, st));
        RC(launch_finish<double>(ps, sq, D64, ny, st));
    } else {
        Staged<float> x32, y32;
        RC(stage<float>(dtype, X, nx, tx, xstr, ws, o_x32, &x32, st));
        RC(stage<float>(dtype, Y, ny, ty, ystr, ws, o_y32, &y32, st));
        float *sq32 = (float *)(ws + o_sq32);
        PairSpec ps{PAIR_GRID, np, ny, nullptr, nullptr};
        RC(run_pairs<float>(m, ps, series<float>(x32.ptr, x32.stride, tx),
                            series<float>(y32.ptr, y32.stride, ty), w32, sq32, st));
        double absmax = 0.0;
        const bool exact_inputs = dtype == WBX_F32;
        if (!exact_inputs) {
            unsigned *mx = (unsigned *)(ws + o_mx);
            WBX_HIP_CHECK(hipMemsetAsync(mx, 0, 4, st));
            RC(launch_absmax_rows<double>((const double *)X, nx, tx, xstr, mx, st));
            RC(launch_absmax_rows<double>((const double *)Y, ny, ty, ystr, mx, st));
            unsigned hb = 0;
            WBX_HIP_CHECK(hipMemcpyAsync(&hb, mx, 4, hipMemcpyDeviceToHost, st));
            WBX_HIP_CHECK(hipStreamSynchronize(st));
            float f;
            memcpy(&f, &hb, 4);
            absmax = (double)f;
        }
        ErrBound eb = make_err_bound((int)tx, (int)ty, m->kind == WBX_WDTW, absmax, exact_inputs);
        int64_t total = 0;
        int32_t *pa = (int32_t *)(ws + o_pa), *pb = (int32_t *)(ws + o_pb);
        RC(launch_argmin_candidates(sq32, nx, ny, (int)k, eb, (int64_t *)(ws + o_cnt),
                                    (int64_t *)(ws + o_off), pa, pb, &total, st));
        RC(launch_fill_f64(D64, np, WBX_INF_D, st));
        // exact FP64 values for the candidates; sq64 reuses the sq32+cnt area
        // only if large enough, so give it the pa/pb-sized tail instead.
        double *sq64 = (double *)(ws + o_sq32);  // np*4 bytes >= total*8 when total <= np/2
        double *tmp = nullptr;
        if ((size_t)total * 8 > (size_t)np * 4) {
            if (hipMalloc(&tmp, (size_t)total * 8) != hipSuccess)
                return fail(WBX_ENOMEM, "candidate buffer allocation failed");
            sq64 = tmp;
        }
        PairSpec pl{PAIR_LIST, total, ny, pa, pb};
        wbx_metric m64 = *m;
        m64.precision = WBX_FP64;
        int rc = run_pairs<double>(&m64, pl, series<double>(x64.ptr, x64.stride, tx),
                                   series<double>(y64.ptr, y64.stride, ty), w64, sq64, st);
        if (!rc) rc = launch_finish<double>(pl, sq64, D64, ny, st);
        if (tmp) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(tmp);
        }
        if (rc) return rc;
    }
    return launch_argmin_replay(D64, nx, ny, k, lb, ws + o_heap, idx, dist, st);
}

// -------------------------------------------------------------- subsequence
// FP32 scan + certified float64 finish (subseq_fast.hip).  Shapelets are
// grouped into band classes (BP = band slots rounded up to 8); each class is
// one launch over every sample.
int subseq_fast_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                     const int64_t *s_off, const int64_t *s_len, int64_t n_s,
                     const std::vector<int64_t> &hl, const void *X, int64_t n, int64_t xstr,
                     int64_t t, int64_t total, int64_t max_len, double *dist, int64_t *idx,
                     hipStream_t st) {
    const bool euclid = m->kind == WBX_EUCLIDEAN;
    // classes by band width B = 2 w - 1 (odd, <= 63): one FP32 scan launch
    // each; class 0 = bands too wide for the scan (exact rescan)
    std::vector<std::vector<int32_t>> cls(64);
    // euclidean shapelets up to kMfmaMaxLen long: the matrix-core scan
    const bool use_mfma = euclid && g_policy.load() != 1;
    std::vector<int32_t> mf_ids;
    for (int64_t j = 0; j < n_s; ++j) {
        int c = 1;
        if (use_mfma && hl[j] <= kMfmaMaxLen) {
            mf_ids.push_back((int32_t)j);
            continue;
        }
        if (!euclid) {
            const int w = band_w((int)hl[j], (int)hl[j], m->r);
            c = 2 * w - 1 <= 63 ? 2 * w - 1 : 0;
        }
        cls[c].push_back((int32_t)j);
    }
    // all class lists back to back (one upload); class c's tasks are n * size
    int64_t max_tasks = 0;
    std::vector<int32_t> lists;
    std::vector<int64_t> cbeg(65, 0);
    for (int c = 0; c < 64; ++c) {
        cbeg[c] = (int64_t)lists.size();
        lists.insert(lists.end(), cls[c].begin(), cls[c].end());
        max_tasks = std::max<int64_t>(max_tasks, n * (int64_t)cls[c].size());
    }
    cbeg[64] = (int64_t)lists.size();
    Carve cv;
    const size_t o_x32 = cv.take(stage_bytes(dtype, false, n, t));
    const size_t o_s32 = cv.take(stage_bytes(dtype, false, 1, total));
    const size_t o_x64 = cv.take(stage_bytes(dtype, true, n, t));
    const size_t o_s64 = cv.take(stage_bytes(dtype, true, 1, total));
    const size_t o_cand = cv.take((size_t)max_tasks * sizeof(SubCand));
    const size_t o_cost = cv.take((size_t)max_tasks * kMaxCand * 8);
    const size_t o_ovf = cv.take((size_t)n * (size_t)n_s * 4);
    const size_t o_list = cv.take((size_t)n_s * 4);
    const size_t o_cnt = cv.take(80 * 4);  // 64 class counters, then the |x| max
    MfmaPlan mf;
    size_t o_mf = 0;
    if (!mf_ids.empty()) {
        mfma_plan(mf_ids, hl, t, &mf);
        o_mf = cv.take(mfma_ws_bytes(mf, n));
    }
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<float> x32, s32;
    Staged<double> x64, s64;
    RC(stage<float>(dtype, X, n, t, xstr, ws, o_x32, &x32, st));
    RC(stage<float>(dtype, S, 1, total, total, ws, o_s32, &s32, st));
    RC(stage<double>(dtype, X, n, t, xstr, ws, o_x64, &x64, st));
    RC(stage<double>(dtype, S, 1, total, total, ws, o_s64, &s64, st));
    int *cnt = (int *)(ws + o_cnt);
    RC(launch_fill_i32(cnt, 80, 0, st));
    double absmax = 0.0;
    const bool exact_inputs = dtype == WBX_F32;
    if (!mf_ids.empty()) {
        ProfScope pf(st, KERN_SUBSEQ);
        RC(mfma_scan(mf, ws + o_mf, exact_inputs ? 1 : 0, x32.ptr, x32.stride, s32.ptr, x64.ptr,
                     x64.stride, s64.ptr, s_off, s_len, n, n_s, dist, idx, st));
    }
    bool any_cls = false;
    for (int c = 0; c < 64; ++c) any_cls = any_cls || !cls[c].empty();
    if (!any_cls) return WBX_OK;
    if (!exact_inputs) {
        unsigned *mx = (unsigned *)(cnt + 64);
        RC(launch_absmax_rows<double>((const double *)X, n, t, xstr, mx, st));
        RC(launch_absmax_rows<double>((const double *)S, 1, total, total, mx, st));
        unsigned hb = 0;
        WBX_HIP_CHECK(hipMemcpyAsync(&hb, mx, 4, hipMemcpyDeviceToHost, st));
        WBX_HIP_CHECK(hipStreamSynchronize(st));
        float f;
        memcpy(&f, &hb, 4);
        absmax = (double)f;
    }
    int32_t *list = (int32_t *)(ws + o_list);
    SubCand *cands = (SubCand *)(ws + o_cand);
    double *cost = (double *)(ws + o_cost);
    int32_t *ovf = (int32_t *)(ws + o_ovf);
    WBX_HIP_CHECK(hipMemcpyAsync(list, lists.data(), sizeof(int32_t) * lists.size(),
                                 hipMemcpyHostToDevice, st));
    // FP32 scan, FP64 finish of the candidates and selection for every class;
    // tasks whose candidates overflowed are listed per class and rescanned
    // exactly after one read-back of the class counters.
    for (int c = 1; c < 64; ++c) {
        const std::vector<int32_t> &v = cls[c];
        if (v.empty()) continue;
        int Lmax = 0;
        for (int32_t j : v) Lmax = std::max<int>(Lmax, (int)hl[j]);
        // path cells P: 2 L - 1 for the DTW recursion, L summed terms for euclidean
        const int P = euclid ? Lmax + 1 : Lmax;
        const ErrBound eb = make_err_bound(P, P, 0, absmax, exact_inputs);
        const double e64 = eb.e64;
        SubGeo sg;
        sg.S = s32.ptr;
        sg.off = s_off;
        sg.len = s_len;
        sg.X = x32.ptr;
        sg.xstride = x32.stride;
        sg.T = (int)t;
        sg.r = m->r;
        sg.rho = (float)(eb.e_up / eb.e_dn * (1.0 + 3.0 * e64) * (1.0 + 1e-6));
        sg.A = (float)(eb.A * (1.0 + 1e-6));
        TaskMap tm{list + cbeg[c], (int32_t)v.size()};
        const int64_t nt = n * (int64_t)v.size();
        {
            ProfScope pf(st, KERN_SUBSEQ);
            RC(launch_subseq_fast(m->kind, c, tm, nt, sg, cands, st));
        }
        RC(launch_subseq_exact_cands(m->kind, c, tm, cands, nt, s64.ptr, s_off, s_len, x64.ptr,
                                     x64.stride, (int)t, m->r, cost, st));
        RC(launch_subseq_select(tm, cands, nt, cost, n_s, dist, idx, ovf + n * cbeg[c], cnt + c,
                                st));
    }
    int novf[64];
    WBX_HIP_CHECK(hipMemcpyAsync(novf, cnt, sizeof novf, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    for (int c = 1; c < 64; ++c) {
        if (cls[c].empty() || novf[c] == 0) continue;
        TaskMap tm{list + cbeg[c], (int32_t)cls[c].size()};
        RC(launch_subseq_rescan(tm, ovf + n * cbeg[c], cnt + c, novf[c], m->kind, s64.ptr, s_off,
                                s_len, x64.ptr, x64.stride, (int)t, m->r, (int)max_len, n_s, dist,
                                idx, st));
    }
    if (!cls[0].empty()) {  // bands too wide for the FP32 kernels: exact rescan
        const std::vector<int32_t> &v = cls[0];
        const int64_t nt = n * (int64_t)v.size();
        std::vector<int32_t> all((size_t)nt);
        for (int64_t q = 0; q < nt; ++q) all[q] = (int32_t)q;
        WBX_HIP_CHECK(hipMemcpyAsync(ovf, all.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice,
                                     st));
        const int ntv = (int)nt;
        WBX_HIP_CHECK(hipMemcpyAsync(cnt, &ntv, 4, hipMemcpyHostToDevice, st));
        TaskMap tm{list + cbeg[0], (int32_t)v.size()};
        RC(launch_subseq_rescan(tm, ovf, cnt, ntv, m->kind, s64.ptr, s_off, s_len, x64.ptr,
                                x64.stride, (int)t, m->r, (int)max_len, n_s, dist, idx, st));
    }
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

// ddtw / wddtw / adtw subsequence scans, float64.  DerivativeDtwSubsequenceMetric
// (distance/_elastic.pyx:2443-2540, ddtw_subsequence_distance :780-818):
// DTW / WDTW of the pattern's average slope against the slopes of each window
// -- a slice of the series' slopes -- with the band of the pattern's own
// length, 0 below length 3; WDDTW's weights are sized by the series' slope
// length (:2530-2540).  AmercingDtwSubsequenceMetric (:2293-2350,
// adtw_subsequence_distance :701-734).
int subseq_other_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                      const int64_t *s_off, const int64_t *s_len, int64_t n_s,
                      const std::vector<int64_t> &hl, const void *X, int64_t n, int64_t xstr,
                      int64_t t, int64_t total, int64_t max_len, double *dist, int64_t *idx,
                      hipStream_t st) {
    const bool deriv = m->kind != WBX_ADTW;
    std::vector<int64_t> nl((size_t)n_s), no((size_t)n_s);
    int64_t ntot = 0, nmin = t, nmax = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        nl[j] = deriv ? (hl[j] >= 3 ? hl[j] - 2 : 0) : hl[j];
        no[j] = ntot;
        ntot += nl[j];
        if (nl[j] > 0) {
            nmin = nl[j] < nmin ? nl[j] : nmin;
            nmax = nl[j] > nmax ? nl[j] : nmax;
        }
    }
    const int64_t tt = deriv ? t - 2 : t;
    Carve cv;
    const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
    const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
    const size_t o_xs = deriv ? cv.take((size_t)n * (size_t)(tt > 0 ? tt : 0) * 8) : 0;
    const size_t o_ss = deriv ? cv.take((size_t)ntot * 8) : 0;
    const size_t o_meta = cv.take((size_t)n_s * 16);
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, s64;
    RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
    RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
    int64_t *d_len = (int64_t *)(ws + o_meta), *d_off = d_len + n_s;
    WBX_HIP_CHECK(hipMemcpyAsync(d_len, nl.data(), sizeof(int64_t) * n_s, hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(d_off, no.data(), sizeof(int64_t) * n_s, hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors end with this call
    RC(launch_subseq_short(d_len, n_s, n, dist, idx, st));
    if (nmax == 0) return WBX_OK;  // every pattern shorter than 3
    const double *xp = x64.ptr, *sp = s64.ptr;
    int64_t xs_stride = x64.stride;
    const int64_t *soff = s_off;
    if (deriv) {
        double *xsl = (double *)(ws + o_xs), *ssl = (double *)(ws + o_ss);
        RC(launch_slope_rows<double>(x64.ptr, n, t, x64.stride, xsl, st));
        RC(launch_slope_packed(s64.ptr, s_off, s_len, n_s, d_off, max_len, ssl, st));
        xp = xsl;
        sp = ssl;
        xs_stride = tt;
        soff = d_off;
    }
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (m->kind == WBX_WDDTW) RC(cached_weights(ctx, tt, m->g, &w64, &w32));
    const int kind = m->kind == WBX_DDTW ? WBX_DTW : (m->kind == WBX_WDDTW ? WBX_WDTW : WBX_ADTW);
    const double *pen = m->kind == WBX_ADTW ? &m->p : nullptr;
    const int band_add = deriv ? 2 : 0;
    Shapelets sh{sp, soff, d_len, n_s};
    const Series<double> xs = series<double>(xp, xs_stride, tt);
    ProfScope pf(st, KERN_SUBSEQ);
    int lane_cap = 0;
    if (g_policy.load() != 1 && lane64_ok(ctx, w64)) {
        RC(launch_subseq_lane64(m->r, sh, (int)nmin, (int)nmax, xs, n, w64, dist, idx, st, band_add,
                                pen));
        lane_cap = lane64_cap(w64 != nullptr);
        if (lane64_class(2 * band_w((int)nmax + band_add, (int)nmax + band_add, m->r) - 1,
                         w64 != nullptr))
            return WBX_OK;
    }
    return launch_subseq_exact<double>(kind, m->r, sh, (int)nmax, xs, n, w64, dist, idx, st,
                                       lane_cap, band_add, pen);
}

int subseq_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                const int64_t *s_off, const int64_t *s_len, int64_t n_s, const void *X,
                int64_t n, int64_t xstr, int64_t t, double *dist, int64_t *idx, hipStream_t st) {
    // Shapelet lengths drive the LDS size: read them back.
    std::vector<int64_t> hl((size_t)n_s), ho((size_t)n_s);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(ho.data(), s_off, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t max_len = 0, min_len = t, total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (hl[j] < 1 || hl[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)hl[j],
                        (long long)t);
        if (hl[j] > max_len) max_len = hl[j];
        if (hl[j] < min_len) min_len = hl[j];
        if (ho[j] + hl[j] > total) total = ho[j] + hl[j];
    }
    wbx_metric mm = *m;
    int64_t s_stride = total;
    RC(acc64_prepare(ctx, &mm, dtype, {{&S, 1, total, &s_stride}, {&X, n, t, &xstr}}, st));
    m = &mm;
    if (m->kind > WBX_WDTW)
        return subseq_other_impl(ctx, m, dtype, S, s_off, s_len, n_s, hl, X, n, xstr, t, total,
                                 max_len, dist, idx, st);
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (m->kind == WBX_WDTW) RC(cached_weights(ctx, t, m->g, &w64, &w32));
    if (m->precision == WBX_FP64 || m->kind == WBX_WDTW) {
        // Exact path (FP64 arithmetic on float64-upcast values).
        Carve cv;
        const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
        const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
        void *wsp = nullptr;
        RC(ws_reserve(ctx, cv.total, &wsp));
        char *ws = (char *)wsp;
        Staged<double> x64, s64;
        RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
        RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
        Shapelets sh{s64.ptr, s_off, s_len, n_s};
        ProfScope pf(st, KERN_SUBSEQ);
        const Series<double> xs = series<double>(x64.ptr, x64.stride, t);
        int lane_cap = 0;
        if (m->kind != WBX_EUCLIDEAN && g_policy.load() != 1 && lane64_ok(ctx, w64)) {
            RC(launch_subseq_lane64(m->r, sh, (int)min_len, (int)max_len, xs, n, w64, dist, idx,
                                    st));
            lane_cap = lane64_cap(w64 != nullptr);
            if (lane64_class(2 * band_w((int)max_len, (int)max_len, m->r) - 1, w64 != nullptr))
                return WBX_OK;
        }
        return launch_subseq_exact<double>(m->kind, m->r, sh, (int)max_len, xs, n, w64, dist, idx,
                                           st, lane_cap);
    }
    return subseq_fast_impl(ctx, m, dtype, S, s_off, s_len, n_s, hl, X, n, xstr, t, total,
                            max_len, dist, idx, st);
}

// -------------------------------------------------------------- profiles
// Squared cost of every window of series q of X against pattern
// (paired ? q : 0) of S: (n, nwp) with nwp = t - shortest pattern + 1.
// Exact float64 arithmetic for every precision (FP32_ACC64 on the inputs
// rounded to float32); WDTW weights are sized by t (the series length, as
// WeightedDtwSubsequenceMetric and Metric.reset(t, s) size them).
int profile_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                 const int64_t *s_off, const int64_t *s_len, int64_t n_s, int paired,
                 const void *X, int64_t n, int64_t xstr, int64_t t, int64_t nwp, double *prof,
                 int64_t *max_len_out, hipStream_t st) {
    if (paired && n_s != n) return fail(WBX_EINVAL, "paired patterns need one per series");
    std::vector<int64_t> hl((size_t)n_s), ho((size_t)n_s);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(ho.data(), s_off, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t max_len = 0, min_len = t, total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (hl[j] < 1 || hl[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)hl[j],
                        (long long)t);
        if (ho[j] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        max_len = hl[j] > max_len ? hl[j] : max_len;
        min_len = hl[j] < min_len ? hl[j] : min_len;
        if (ho[j] + hl[j] > total) total = ho[j] + hl[j];
    }
    if (nwp != t - min_len + 1)
        return fail(WBX_EINVAL, "profile width %lld, expected %lld", (long long)nwp,
                    (long long)(t - min_len + 1));
    if (max_len_out) *max_len_out = max_len;
    wbx_metric mm = *m;
    int64_t s_stride = total;
    RC(acc64_prepare(ctx, &mm, dtype, {{&S, 1, total, &s_stride}, {&X, n, t, &xstr}}, st));
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (mm.kind == WBX_WDTW) RC(cached_weights(ctx, t, mm.g, &w64, &w32));
    if (mm.kind > WBX_WDTW) {
        // DerivativeDtw- / WeightedDerivativeDtw- / AmercingDtwSubsequenceMetric
        // ._matches (_elastic.pyx:2493-2520, 2330-2350; ddtw_ / adtw_subsequence_
        // matches :821-866, :737-777): the full squared cost of every window,
        // band by the pattern's own length, WDDTW weights sized t - 2; their
        // row-minimum abandon never drops a window whose cost is within the
        // threshold (penalty >= 0)
        if (mm.kind == WBX_ADTW && !(mm.p >= 0.0))
            return fail(WBX_EUNSUPPORTED, "adtw window profiles need a penalty >= 0");
        if (mm.kind == WBX_WDDTW && t > 2) RC(cached_weights(ctx, t - 2, mm.g, &w64, &w32));
        Carve cv;
        const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
        const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
        const size_t o_k = cv.take(scaled_wrap_profile_ws_bytes(hl.data(), n_s, n, total, t));
        void *wsp = nullptr;
        RC(ws_reserve(ctx, cv.total, &wsp));
        char *ws = (char *)wsp;
        Staged<double> x64, s64;
        RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
        RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
        ProfScope pf(st, KERN_SUBSEQ);
        const bool deriv = mm.kind != WBX_ADTW;
        return launch_scaled_wrap_profile(mm.kind, mm.r, w64, deriv ? 0.0 : mm.p, s64.ptr, s_off,
                                          s_len, hl.data(), n_s, total, NORM_NONE, nullptr,
                                          nullptr, paired, x64.ptr, x64.stride, n, t, nwp, 2,
                                          deriv ? 2 : 0, WBX_INF_D, prof, ws + o_k, st);
    }
    Carve cv;
    const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
    const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, s64;
    RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
    RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
    Shapelets sh{s64.ptr, s_off, s_len, n_s};
    ProfScope pf(st, KERN_SUBSEQ);
    return launch_profile(mm.kind, mm.r, sh, paired, (int)min_len, (int)max_len,
                          series<double>(x64.ptr, x64.stride, t), n, nwp, w64,
                          g_policy.load() != 1 && lane64_ok(ctx, w64), prof, st);
}

// Host-pointer wrapper helpers.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (bytes == 0) bytes = 8;
        if (hipMalloc(&p, bytes) != hipSuccess)
            return fail(WBX_ENOMEM, "device allocation of %zu bytes failed", bytes);
        return WBX_OK;
    }
};

int upload(DevBuf &b, const void *h, size_t bytes, hipStream_t st) {
    RC(b.alloc(bytes));
    if (h && bytes) WBX_HIP_CHECK(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
    return WBX_OK;
}

int check_sizes(int64_t nx, int64_t tx) {
    if (nx < 1) return fail(WBX_EINVAL, "need at least one series, got %lld", (long long)nx);
    if (tx < 1) return fail(WBX_EINVAL, "need at least one timestep, got %lld", (long long)tx);
    if (tx > (1 << 20)) return fail(WBX_EUNSUPPORTED, "series longer than 2^20");
    return WBX_OK;
}

// ------------------------------------------------- z-normalised scans
// pairwise_subsequence_distance with scale=True (_distance.py:543-636 picks
// "scaled_" + metric, :280-300): ScaledEuclideanSubsequenceMetric
// (kind WBX_EUCLIDEAN) and ScaledDtwSubsequenceMetric (WBX_DTW), float64 for
// every precision (FP32_ACC64 on the inputs rounded to float32).
int scaled_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                const int64_t *s_off, const int64_t *s_len, const double *s_mean,
                const double *s_std, int64_t n_s, const void *X, int64_t n, int64_t xstr,
                int64_t t, double *dist, int64_t *idx, hipStream_t st) {
    std::vector<int64_t> hl((size_t)n_s), ho((size_t)n_s);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(ho.data(), s_off, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (hl[j] < 1 || hl[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)hl[j],
                        (long long)t);
        if (ho[j] < 0) return fail(WBX_EINVAL, "negative shapelet offset");
        if (ho[j] + hl[j] > total) total = ho[j] + hl[j];
    }
    wbx_metric mm = *m;
    int64_t s_stride = total;
    RC(acc64_prepare(ctx, &mm, dtype, {{&S, 1, total, &s_stride}, {&X, n, t, &xstr}}, st));
    Carve cv;
    const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
    const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
    const size_t o_k = cv.take(scaled_ws_bytes(mm.kind, hl.data(), n_s, total, n, t));
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, s64;
    RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
    RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
    ProfScope pf(st, KERN_SUBSEQ);
    return launch_scaled_subseq(mm.kind, mm.r, s64.ptr, s_off, s_len, hl.data(), n_s, total,
                                s_mean, s_std, x64.ptr, x64.stride, n, t, dist, idx, ws + o_k,
                                g_policy.load() == 1, st);
}

// The wrapped metrics of the same scan (ScaledSubsequenceMetricWrap,
// _cdistance.pyx:470-567, "scaled_" + wdtw / ddtw / wddtw / adtw,
// _distance.py:151-160): the wrapped metric is reset with the series array
// (wrap.reset(X, X)), so WDTW weights are sized t and WDDTW weights t - 2.
int scaled_wrap_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                     const int64_t *s_off, const int64_t *s_len, const double *s_mean,
                     const double *s_std, int64_t n_s, const void *X, int64_t n, int64_t xstr,
                     int64_t t, double *dist, int64_t *idx, hipStream_t st) {
    if (m->kind == WBX_ADTW && !(m->p >= 0.0))
        return fail(WBX_EUNSUPPORTED, "scaled adtw with a negative penalty is not served");
    std::vector<int64_t> hl((size_t)n_s), ho((size_t)n_s);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(ho.data(), s_off, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (hl[j] < 1 || hl[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)hl[j],
                        (long long)t);
        if (ho[j] < 0) return fail(WBX_EINVAL, "negative shapelet offset");
        if (ho[j] + hl[j] > total) total = ho[j] + hl[j];
    }
    wbx_metric mm = *m;
    int64_t s_stride = total;
    RC(acc64_prepare(ctx, &mm, dtype, {{&S, 1, total, &s_stride}, {&X, n, t, &xstr}}, st));
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (mm.kind == WBX_WDTW) RC(cached_weights(ctx, t, mm.g, &w64, &w32));
    if (mm.kind == WBX_WDDTW && t > 2) RC(cached_weights(ctx, t - 2, mm.g, &w64, &w32));
    Carve cv;
    const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
    const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
    const size_t o_k = cv.take(scaled_wrap_ws_bytes(hl.data(), n_s, total, n, t));
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, s64;
    RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
    RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
    ProfScope pf(st, KERN_SUBSEQ);
    return launch_scaled_wrap(mm.kind, mm.r, w64, mm.kind == WBX_ADTW ? mm.p : 0.0, s64.ptr,
                              s_off, s_len, hl.data(), n_s, total, s_mean, s_std, x64.ptr,
                              x64.stride, n, t, dist, idx, ws + o_k, st);
}

// Window profiles of the reference's scaled matching (subsequence_match /
// paired_subsequence_match with scale=True, _distance.py:732-1079 through
// ScaledEuclideanSubsequenceMetric._matches, _metric.pyx:141-167, and
// ScaledDtwSubsequenceMetric._matches, _elastic.pyx:2112-2149).
int scaled_profile_impl(DeviceCtx *ctx, const wbx_metric *m, int dtype, const void *S,
                        const int64_t *s_off, const int64_t *s_len, const double *s_mean,
                        const double *s_std, int64_t n_s, int paired, const void *X, int64_t n,
                        int64_t xstr, int64_t t, int64_t nwp, double threshold, double *prof,
                        hipStream_t st) {
    if (paired && n_s != n) return fail(WBX_EINVAL, "paired patterns need one per series");
    std::vector<int64_t> hl((size_t)n_s), ho((size_t)n_s);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(ho.data(), s_off, sizeof(int64_t) * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t min_len = t, total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (hl[j] < 1 || hl[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)hl[j],
                        (long long)t);
        if (ho[j] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        if (paired || j == 0) min_len = hl[j] < min_len ? hl[j] : min_len;
        if (ho[j] + hl[j] > total) total = ho[j] + hl[j];
    }
    if (nwp != t - min_len + 1)
        return fail(WBX_EINVAL, "profile width %lld, expected %lld", (long long)nwp,
                    (long long)(t - min_len + 1));
    if (threshold != threshold) return fail(WBX_EINVAL, "threshold is NaN");
    wbx_metric mm = *m;
    int64_t s_stride = total;
    RC(acc64_prepare(ctx, &mm, dtype, {{&S, 1, total, &s_stride}, {&X, n, t, &xstr}}, st));
    const bool wrap = mm.kind != WBX_EUCLIDEAN && mm.kind != WBX_DTW;
    if (wrap && mm.kind == WBX_ADTW && !(mm.p >= 0.0))
        return fail(WBX_EUNSUPPORTED, "scaled adtw with a negative penalty is not served");
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (mm.kind == WBX_WDTW) RC(cached_weights(ctx, t, mm.g, &w64, &w32));
    if (mm.kind == WBX_WDDTW && t > 2) RC(cached_weights(ctx, t - 2, mm.g, &w64, &w32));
    Carve cv;
    const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
    const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
    const size_t o_k =
        cv.take(wrap ? scaled_wrap_profile_ws_bytes(hl.data(), n_s, n, total, t)
                     : scaled_profile_ws_bytes(mm.kind, hl.data(), n_s, total, paired ? 1 : 0,
                                               n, t));
    void *wsp = nullptr;
    RC(ws_reserve(ctx, cv.total, &wsp));
    char *ws = (char *)wsp;
    Staged<double> x64, s64;
    RC(stage<double>(dtype, X, n, t, xstr, ws, o_x, &x64, st));
    RC(stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st));
    ProfScope pf(st, KERN_SUBSEQ);
    if (wrap)  // ScaledSubsequenceMetricWrap._matches: distance < threshold
        return launch_scaled_wrap_profile(mm.kind, mm.r, w64, mm.kind == WBX_ADTW ? mm.p : 0.0,
                                          s64.ptr, s_off, s_len, hl.data(), n_s, total,
                                          NORM_GIVEN, s_mean, s_std, paired ? 1 : 0, x64.ptr,
                                          x64.stride, n, t, nwp, 1, 0, threshold, prof, ws + o_k,
                                          st);
    return launch_scaled_profile(mm.kind, mm.r, s64.ptr, s_off, s_len, hl.data(), n_s, total,
                                 s_mean, s_std, paired ? 1 : 0, x64.ptr, x64.stride, n, t, nwp,
                                 threshold, prof, ws + o_k, g_policy.load() == 1, st);
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

const char *wbx_last_error(void) { return last_error(); }

const char *wbx_version(void) { return "wbx 0.1.0 (gfx950)"; }

int wbx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int wbx_init(int device) {
    WBX_HIP_CHECK(hipSetDevice(device));
    if (!ctx_current()) return WBX_EDEVICE;
    return scratch_setup(device, nullptr);
}

// Not part of the reference interface: selects the kernel family (0 = the
// fastest available, 1 = the plain one-pair-per-block kernel everywhere,
// 2 = the group kernel wherever it applies; for tests and measurements).
int wbx_set_policy(int policy) {
    if (policy < 0 || policy > 2) return fail(WBX_EINVAL, "policy must be 0, 1 or 2");
    g_policy.store(policy);
    return WBX_OK;
}

// Not part of the reference interface: kernel timing for benchmarks.
// wbx_prof_enable(1) starts recording HIP events around every distance kernel;
// wbx_prof_take waits for them and returns the summed milliseconds, the number
// of launches and the kind of the last one (1 euclid, 2 lane, 3 plain), then
// clears the record.
// Not part of the reference interface: the series length above which FP32
// requests of pairwise / paired / argmin run with float64 accumulation
// (FP32_ACC64); default 2048, 0 = never.
int wbx_prof_pruned(int64_t *evaluated) {
    if (!evaluated) return fail(WBX_EINVAL, "NULL argument");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    *evaluated = ctx->prune_evaluated;
    ctx->prune_evaluated = 0;
    return WBX_OK;
}

int wbx_set_acc64_length(int64_t len) {
    if (len < 0) return fail(WBX_EINVAL, "length must be >= 0");
    set_acc64_length(len);
    return WBX_OK;
}

int wbx_prof_enable(int on) {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.on = on != 0;
    return WBX_OK;
}

int wbx_prof_take(double *total_ms, int64_t *launches, int *last_kind) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<int> kinds;
    {
        std::lock_guard<std::mutex> lk(g_prof.mu);
        ev.swap(g_prof.ev);
        kinds.swap(g_prof.kind);
    }
    double tot = 0.0;
    int rc = WBX_OK;
    for (auto &e : ev) {
        float ms = 0.0f;
        if (hipEventSynchronize(e.second) != hipSuccess ||
            hipEventElapsedTime(&ms, e.first, e.second) != hipSuccess)
            rc = fail(WBX_EDEVICE, "event timing failed");
        tot += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (int64_t)ev.size();
    if (last_kind) *last_kind = kinds.empty() ? 0 : kinds.back();
    return rc;
}

int wbx_dev_pairwise(const wbx_metric *metric, int dtype, const void *X, int64_t nx,
                     int64_t x_stride, int64_t tx, const void *Y, int64_t ny, int64_t y_stride,
                     int64_t ty, double *out, void *stream) {
    RC(check_metric(metric));
    RC(check_sizes(nx, tx));
    const bool singleton = (Y == nullptr);
    if (singleton) {
        ny = nx;
        ty = tx;
        y_stride = x_stride;
    } else {
        RC(check_sizes(ny, ty));
    }
    if (metric->kind == WBX_EUCLIDEAN && !singleton && tx != ty)
        return fail(WBX_EINVAL, "Illegal n_timestep (%lld != %lld) for non-elastic distance measure",
                    (long long)tx, (long long)ty);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    wbx_metric mm = *metric;
    promote_long(&mm, tx > ty ? tx : ty);
    int rc = acc64_prepare(ctx, &mm, dtype,
                           {{&X, nx, tx, &x_stride}, {singleton ? nullptr : &Y, ny, ty, &y_stride}},
                           st);
    int w_fixed = 0;
    if (!rc && (mm.kind == WBX_DDTW || mm.kind == WBX_WDDTW)) {
        if ((tx < ty ? tx : ty) < 3) {
            // DerivativeDtwMetric._distance returns 0 below length 3
            rc = launch_fill_f64(out, singleton ? nx * nx : nx * ny, 0.0, st);
            int rc2 = end(ctx, st);
            return rc ? rc : rc2;
        }
        // the pairwise band comes from the series' own lengths (_elastic.pyx:3276)
        w_fixed = band_w((int)tx, (int)ty, mm.r);
        rc = deriv_prepare(ctx, &mm, &dtype,
                           {{&X, nx, &tx, &x_stride}, {singleton ? nullptr : &Y, ny, &ty, &y_stride}},
                           st);
        if (singleton) ty = tx;
    }
    if (singleton) y_stride = x_stride;
    if (!rc)
        rc = mm.precision == WBX_FP64
                 ? pairwise_impl<double>(ctx, &mm, dtype, X, nx, x_stride, tx, Y, ny, y_stride, ty,
                                         singleton, out, st, w_fixed)
                 : pairwise_impl<float>(ctx, &mm, dtype, X, nx, x_stride, tx, Y, ny, y_stride, ty,
                                        singleton, out, st, w_fixed);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

// wbx_dev_pairwise without allocation: the staged inputs and the per-pair
// costs go to the fixed per-device scratch (kScratchBytes), so nothing is
// allocated or freed and no memset is enqueued.  Metrics that need weights or
// slopes (wdtw, ddtw, wddtw) and the FP32_ACC64 mode are refused.  Calls that
// overlap on one GPU share its scratch and race.
int wbx_dev_pairwise_noalloc(const wbx_metric *metric, int dtype, const void *X, int64_t nx,
                             int64_t x_stride, int64_t tx, const void *Y, int64_t ny,
                             int64_t y_stride, int64_t ty, double *out, void *stream) {
    RC(check_metric(metric));
    RC(check_sizes(nx, tx));
    if (dtype != WBX_F32 && dtype != WBX_F64) return fail(WBX_EINVAL, "unknown dtype %d", dtype);
    if (!X || !out) return fail(WBX_EINVAL, "NULL array");
    const bool singleton = (Y == nullptr);
    if (singleton) {
        ny = nx;
        ty = tx;
        y_stride = x_stride;
    } else {
        RC(check_sizes(ny, ty));
    }
    if (x_stride < tx || y_stride < ty) return fail(WBX_EINVAL, "row stride below the series length");
    if (metric->kind == WBX_WDTW || metric->kind == WBX_DDTW || metric->kind == WBX_WDDTW)
        return fail(WBX_EUNSUPPORTED,
                    "the allocation-free pairwise serves euclidean, dtw and adtw (this metric "
                    "needs weights or slopes: use wbx_dev_pairwise)");
    if (metric->precision == WBX_FP32_ACC64)
        return fail(WBX_EUNSUPPORTED, "the allocation-free pairwise serves FP64 and FP32");
    if (metric->kind == WBX_EUCLIDEAN && !singleton && tx != ty)
        return fail(WBX_EINVAL, "Illegal n_timestep (%lld != %lld) for non-elastic distance measure",
                    (long long)tx, (long long)ty);
    int dev = 0;
    WBX_HIP_CHECK(hipGetDevice(&dev));
    void *ws = nullptr;
    RC(scratch_setup(dev, &ws));
    const bool dbl = metric->precision == WBX_FP64;
    const PairwiseLayout L = dbl ? pairwise_layout<double>(dtype, nx, tx, ny, ty, singleton)
                                 : pairwise_layout<float>(dtype, nx, tx, ny, ty, singleton);
    if (L.total > kScratchBytes)
        return fail(WBX_EINVAL,
                    "allocation-free pairwise of %lld x %lld series needs %zu bytes of scratch; "
                    "the maximum is %zu (use wbx_dev_pairwise)",
                    (long long)nx, (long long)ny, L.total, kScratchBytes);
    hipStream_t st = (hipStream_t)stream;
    g_no_alloc = true;
    const int rc =
        dbl ? pairwise_run<double>(L, metric, dtype, X, nx, x_stride, tx, Y, y_stride, ty,
                                   singleton, nullptr, (char *)ws, out, st, 0)
            : pairwise_run<float>(L, metric, dtype, X, nx, x_stride, tx, Y, y_stride, ty, singleton,
                                  nullptr, (char *)ws, out, st, 0);
    g_no_alloc = false;
    return rc;
}

int wbx_dev_paired(const wbx_metric *metric, int dtype, const void *X, int64_t x_stride,
                   const void *Y, int64_t y_stride, int64_t n, int64_t t, double *out,
                   void *stream) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    // _paired_distance binds its rows to the caller's y: out[i] = D(Y[i], X[i]).
    PairSpec ps{PAIR_DIAG, n, n, nullptr, nullptr};
    wbx_metric mm = *metric;
    promote_long(&mm, t);
    int rc = acc64_prepare(ctx, &mm, dtype, {{&X, n, t, &x_stride}, {&Y, n, t, &y_stride}}, st);
    int w_fixed = 0;
    if (!rc && (mm.kind == WBX_DDTW || mm.kind == WBX_WDDTW)) {
        if (t < 3) {
            rc = launch_fill_f64(out, n, 0.0, st);
            int rc2 = end(ctx, st);
            return rc ? rc : rc2;
        }
        w_fixed = band_w((int)t, (int)t, mm.r);
        int64_t ty = t;
        rc = deriv_prepare(ctx, &mm, &dtype, {{&X, n, &t, &x_stride}, {&Y, n, &ty, &y_stride}}, st);
    }
    if (rc) {
        int rc2 = end(ctx, st);
        return rc ? rc : rc2;
    }
    metric = &mm;
    const bool dbl = metric->precision == WBX_FP64;
    Carve cv;
    const size_t o_x = cv.take(stage_bytes(dtype, dbl, n, t));
    const size_t o_y = cv.take(stage_bytes(dtype, dbl, n, t));
    const size_t o_sq = cv.take((size_t)n * 8);
    void *wsp = nullptr;
    rc = ws_reserve(ctx, cv.total, &wsp);
    if (!rc) {
        char *ws = (char *)wsp;
        const double *w64 = nullptr;
        const float *w32 = nullptr;
        if (metric->kind == WBX_WDTW) rc = cached_weights(ctx, t, metric->g, &w64, &w32);
        if (!rc && dbl) {
            Staged<double> sx, sy;
            rc = stage<double>(dtype, X, n, t, x_stride, ws, o_x, &sx, st);
            if (!rc) rc = stage<double>(dtype, Y, n, t, y_stride, ws, o_y, &sy, st);
            if (!rc)
                rc = run_pairs<double>(metric, ps, series<double>(sy.ptr, sy.stride, t),
                                       series<double>(sx.ptr, sx.stride, t), w64,
                                       (double *)(ws + o_sq), st, w_fixed);
            if (!rc) rc = launch_finish<double>(ps, (double *)(ws + o_sq), out, n, st);
        } else if (!rc) {
            Staged<float> sx, sy;
            rc = stage<float>(dtype, X, n, t, x_stride, ws, o_x, &sx, st);
            if (!rc) rc = stage<float>(dtype, Y, n, t, y_stride, ws, o_y, &sy, st);
            if (!rc)
                rc = run_pairs<float>(metric, ps, series<float>(sy.ptr, sy.stride, t),
                                      series<float>(sx.ptr, sx.stride, t), w32,
                                      (float *)(ws + o_sq), st, w_fixed);
            if (!rc) rc = launch_finish<float>(ps, (float *)(ws + o_sq), out, n, st);
        }
    }
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

int wbx_dev_argmin(const wbx_metric *metric, int dtype, const void *X, int64_t nx,
                   int64_t x_stride, int64_t tx, const void *Y, int64_t ny, int64_t y_stride,
                   int64_t ty, int64_t k, const double *lower_bound, int64_t *idx, double *dist,
                   void *stream) {
    RC(check_metric(metric));
    RC(check_sizes(nx, tx));
    RC(check_sizes(ny, ty));
    if (k < 1) return fail(WBX_EINVAL, "k must be >= 1, got %lld", (long long)k);
    if (k > ny) k = ny;
    if (metric->kind == WBX_EUCLIDEAN && tx != ty)
        return fail(WBX_EINVAL, "Illegal n_timestep (%lld != %lld) for non-elastic distance measure.",
                    (long long)tx, (long long)ty);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    wbx_metric mm = *metric;
    promote_long(&mm, tx > ty ? tx : ty);
    int rc = acc64_prepare(ctx, &mm, dtype, {{&X, nx, tx, &x_stride}, {&Y, ny, ty, &y_stride}}, st);
    if (!rc && (mm.kind == WBX_DDTW || mm.kind == WBX_WDDTW)) {
        if ((tx < ty ? tx : ty) < 3) {
            // DerivativeDtwMetric._eadistance never accepts a series shorter
            // than 3: the driver's zero-initialised outputs stay
            // (_cdistance.pyx:1360-1361)
            rc = launch_fill_f64(dist, nx * k, 0.0, st);
            if (!rc) rc = launch_fill_f64((double *)idx, nx * k, 0.0, st);
            int rc2 = end(ctx, st);
            return rc ? rc : rc2;
        }
        // argmin bands by the slope lengths (_elastic.pyx:3308): no override
        rc = deriv_prepare(ctx, &mm, &dtype, {{&X, nx, &tx, &x_stride}, {&Y, ny, &ty, &y_stride}},
                           st);
    }
    if (!rc)
        rc = argmin_impl(ctx, &mm, dtype, X, nx, x_stride, tx, Y, ny, y_stride, ty, k, lower_bound,
                         idx, dist, st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

int wbx_dev_scaled_subsequence_profile(const wbx_metric *metric, int dtype, const void *S,
                                       const int64_t *s_off, const int64_t *s_len,
                                       const double *s_mean, const double *s_std, int64_t n_s,
                                       int paired, const void *X, int64_t n, int64_t x_stride,
                                       int64_t t, int64_t nwp, double threshold, double *prof,
                                       void *stream) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (n_s > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "too many patterns");
    if (!S || !s_off || !s_len || !s_mean || !s_std || !X || !prof)
        return fail(WBX_EINVAL, "NULL argument");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    int rc = scaled_profile_impl(ctx, metric, dtype, S, s_off, s_len, s_mean, s_std, n_s,
                                 paired ? 1 : 0, X, n, x_stride, t, nwp, threshold, prof, st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

}  // extern "C"

namespace {
// argmin_subsequence_distance through the wrapped-metric kernels: norm
// NORM_FAST for scale=True (_ScaledArgminSubsequenceDistance), NORM_NONE for
// the unscaled ddtw / wddtw / adtw (_ArgminSubsequenceDistance,
// _cdistance.pyx:1408-1434, the metric's _eadistance: derivative band by
// min(x_len - 2, y_len - 2)).
int wrap_argmin(const wbx_metric *metric, int dtype, const void *S, const int64_t *s_off,
                const int64_t *s_len, const void *X, int64_t n, int64_t x_stride, int64_t t,
                int64_t k, int64_t *idx, double *dist, void *stream, int norm) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (!S || !s_off || !s_len || !X || !idx || !dist) return fail(WBX_EINVAL, "NULL argument");
    if (k < 1) return fail(WBX_EINVAL, "k == %lld, must be >= 1", (long long)k);
    if (metric->kind == WBX_ADTW && !(metric->p >= 0.0))
        return fail(WBX_EUNSUPPORTED, "scaled adtw with a negative penalty is not served");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    std::vector<int64_t> hl((size_t)n), ho((size_t)n);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(ho.data(), s_off, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t min_len = t, max_len = 0, total = 0;
    for (int64_t q = 0; q < n; ++q) {
        if (hl[q] < 1 || hl[q] > t)
            return fail(WBX_EINVAL, "the longest subsequence must be shorter than samples.");
        if (ho[q] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        min_len = hl[q] < min_len ? hl[q] : min_len;
        max_len = hl[q] > max_len ? hl[q] : max_len;
        if (ho[q] + hl[q] > total) total = ho[q] + hl[q];
    }
    if (k > t - max_len + 1) return fail(WBX_EINVAL, "k must be less x.shape[-1] - y.shape[-1] + 1.");
    const int64_t nwp = t - min_len + 1;
    wbx_metric mm = *metric;
    int64_t s_stride = total;
    RC(acc64_prepare(ctx, &mm, dtype, {{&S, 1, total, &s_stride}, {&X, n, t, &x_stride}}, st));
    // Metric.reset(t, s) sizes the weights by the wider array: t
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (mm.kind == WBX_WDTW) RC(cached_weights(ctx, t, mm.g, &w64, &w32));
    if (mm.kind == WBX_WDDTW && t > 2) RC(cached_weights(ctx, t - 2, mm.g, &w64, &w32));
    // the profile and the heaps live in their own allocation; the workspace
    // holds the staged inputs and the scan's buffers
    const size_t prof_bytes = ((size_t)n * nwp * 8 + 255) & ~(size_t)255;
    void *buf = nullptr;
    if (hipMalloc(&buf, prof_bytes + (size_t)n * k * 16) != hipSuccess)
        return fail(WBX_ENOMEM, "profile allocation of %zu bytes failed", prof_bytes);
    double *prof = (double *)buf;
    int rc = WBX_OK;
    {
        Carve cv;
        const size_t o_x = cv.take(stage_bytes(dtype, true, n, t));
        const size_t o_s = cv.take(stage_bytes(dtype, true, 1, total));
        const size_t o_k = cv.take(scaled_wrap_profile_ws_bytes(hl.data(), n, n, total, t));
        void *wsp = nullptr;
        rc = ws_reserve(ctx, cv.total, &wsp);
        char *ws = (char *)wsp;
        Staged<double> x64{}, s64{};
        if (!rc) rc = stage<double>(dtype, X, n, t, x_stride, ws, o_x, &x64, st);
        if (!rc) rc = stage<double>(dtype, S, 1, total, total, ws, o_s, &s64, st);
        if (!rc) {
            ProfScope pf(st, KERN_SUBSEQ);
            rc = launch_scaled_wrap_profile(mm.kind, mm.r, w64, mm.kind == WBX_ADTW ? mm.p : 0.0,
                                            s64.ptr, s_off, s_len, hl.data(), n, total, norm,
                                            nullptr, nullptr, 1, x64.ptr, x64.stride, n, t, nwp,
                                            0, 0, WBX_INF_D, prof, ws + o_k, st);
        }
    }
    if (!rc) rc = launch_argmin_replay(prof, n, nwp, k, nullptr, (char *)buf + prof_bytes, idx,
                                       dist, st);
    const int rc2 = end(ctx, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(buf);
    return rc ? rc : rc2;
}
}  // namespace

extern "C" {

int wbx_dev_scaled_argmin_subsequence(const wbx_metric *metric, int dtype, const void *S,
                                      const int64_t *s_off, const int64_t *s_len, const void *X,
                                      int64_t n, int64_t x_stride, int64_t t, int64_t k,
                                      int64_t *idx, double *dist, void *stream) {
    return wrap_argmin(metric, dtype, S, s_off, s_len, X, n, x_stride, t, k, idx, dist, stream,
                       NORM_FAST);
}

int wbx_dev_scaled_subsequence_pairwise(const wbx_metric *metric, int dtype, const void *S,
                                        const int64_t *s_off, const int64_t *s_len,
                                        const double *s_mean, const double *s_std, int64_t n_s,
                                        const void *X, int64_t n, int64_t x_stride, int64_t t,
                                        double *dist, int64_t *idx, void *stream) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (n_s > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "too many shapelets");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    const bool direct = metric->kind == WBX_EUCLIDEAN || metric->kind == WBX_DTW;
    int rc = direct ? scaled_impl(ctx, metric, dtype, S, s_off, s_len, s_mean, s_std, n_s, X, n,
                                  x_stride, t, dist, idx, st)
                    : scaled_wrap_impl(ctx, metric, dtype, S, s_off, s_len, s_mean, s_std, n_s,
                                       X, n, x_stride, t, dist, idx, st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

int wbx_dev_subsequence_pairwise(const wbx_metric *metric, int dtype, const void *S,
                                 const int64_t *s_off, const int64_t *s_len, int64_t n_s,
                                 const void *X, int64_t n, int64_t x_stride, int64_t t,
                                 double *dist, int64_t *idx, void *stream) {
    RC(check_subsequence_metric(metric, true));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    int rc = subseq_impl(ctx, metric, dtype, S, s_off, s_len, n_s, X, n, x_stride, t, dist, idx,
                         st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

// ------------------------------------------------------- warping paths
// Pairs per launch so that the move bytes stay within 512 MiB.
static int64_t path_chunk(int64_t n, int64_t lx, int64_t B) {
    const int64_t per = lx * B;
    int64_t c = ((int64_t)512 << 20) / (per > 0 ? per : 1);
    if (c < 1) c = 1;
    return c < n ? c : n;
}

int wbx_dev_dtw_path(const double *X, int64_t x_stride, int64_t tx, const double *Y,
                     int64_t y_stride, int64_t ty, int64_t n, int64_t w, const double *weights,
                     int32_t *rows, void *stream) {
    if (!X || !Y || !rows) return fail(WBX_EINVAL, "NULL argument");
    RC(check_sizes(n, tx));
    RC(check_sizes(n, ty));
    if (w < 1) return fail(WBX_EINVAL, "warp size must be >= 1, got %lld", (long long)w);
    if (w > (tx > ty ? tx : ty)) w = tx > ty ? tx : ty;
    if (x_stride < 0 || y_stride < 0) return fail(WBX_EINVAL, "negative stride");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    const Band bd = make_band((int)tx, (int)ty, 0.0, (int)w);
    const int64_t chunk = path_chunk(n, tx, bd.B);
    void *wsp = nullptr;
    int rc = ws_reserve(ctx, (size_t)chunk * (size_t)tx * (size_t)bd.B, &wsp);
    for (int64_t p0 = 0; !rc && p0 < n; p0 += chunk) {
        const int64_t c = n - p0 < chunk ? n - p0 : chunk;
        rc = launch_path_rows(X + p0 * x_stride, x_stride, Y + p0 * y_stride, y_stride, c, (int)tx,
                              (int)ty, (int)w, weights, (unsigned char *)wsp, rows + p0 * tx * 2,
                              st);
    }
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

int wbx_dev_dtw_ssg_update(double *mean, int64_t tm, const int32_t *rows, const double *x,
                           int64_t tx, double weight, double learning_rate, void *stream) {
    if (!mean || !rows || !x) return fail(WBX_EINVAL, "NULL argument");
    RC(check_sizes(1, tm));
    RC(check_sizes(1, tx));
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    int rc = launch_ssg_update(rows, x, weight, learning_rate, (int)tm, mean, st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

int wbx_dev_dtw_mm_update(const double *mean, int64_t tm, const double *X, int64_t x_stride,
                          int64_t tx, int64_t n, int64_t w, const double *weights,
                          const double *sample_weight, double *mean_out, void *stream) {
    if (!mean || !X || !mean_out) return fail(WBX_EINVAL, "NULL argument");
    RC(check_sizes(n, tx));
    RC(check_sizes(1, tm));
    if (w < 1) return fail(WBX_EINVAL, "warp size must be >= 1, got %lld", (long long)w);
    if (w > (tm > tx ? tm : tx)) w = tm > tx ? tm : tx;
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    const Band bd = make_band((int)tm, (int)tx, 0.0, (int)w);
    const int64_t chunk = path_chunk(n, tm, bd.B);
    Carve cv;
    const size_t o_rows = cv.take((size_t)n * (size_t)tm * 8);
    const size_t o_mv = cv.take((size_t)chunk * (size_t)tm * (size_t)bd.B);
    void *wsp = nullptr;
    int rc = ws_reserve(ctx, cv.total, &wsp);
    int32_t *rows = wsp ? (int32_t *)((char *)wsp + o_rows) : nullptr;
    // every pair aligns the same mean (stride 0) with its series
    for (int64_t p0 = 0; !rc && p0 < n; p0 += chunk) {
        const int64_t c = n - p0 < chunk ? n - p0 : chunk;
        rc = launch_path_rows(mean, 0, X + p0 * x_stride, x_stride, c, (int)tm, (int)tx, (int)w,
                              weights, (unsigned char *)wsp + o_mv, rows + p0 * tm * 2, st);
    }
    if (!rc) rc = launch_mm_update(rows, X, x_stride, n, (int)tm, sample_weight, mean_out, st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

// ----------------------------------------------------------- lower bounds
int wbx_dev_dtw_envelope(const double *X, int64_t n, int64_t stride, int64_t t, int64_t r,
                         double *lower, double *upper, void *stream) {
    if (!X || !lower || !upper) return fail(WBX_EINVAL, "NULL argument");
    RC(check_sizes(n, t));
    if (r < 0 || r >= t) return fail(WBX_EINVAL, "invalid r");
    if (stride < t) return fail(WBX_EINVAL, "stride %lld < length %lld", (long long)stride,
                                (long long)t);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    return launch_envelope(X, n, stride, (int)t, (int)r, lower, upper, st);
}

int wbx_dev_lb_keogh(const double *X, int64_t nx, int64_t x_stride, const double *lower,
                     const double *upper, int64_t ny, int64_t t, double *out, void *stream) {
    if (!X || !lower || !upper || !out) return fail(WBX_EINVAL, "NULL argument");
    RC(check_sizes(nx, t));
    RC(check_sizes(ny, t));
    if (x_stride < t) return fail(WBX_EINVAL, "stride %lld < length %lld", (long long)x_stride,
                                  (long long)t);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    return launch_lb_keogh(X, nx, x_stride, lower, upper, ny, (int)t, out, st);
}

int wbx_dev_lb_kim(const double *Y, int64_t ny, int64_t y_stride, const double *X, int64_t nx,
                   int64_t x_stride, int64_t t, double *out, void *stream) {
    if (!X || !Y || !out) return fail(WBX_EINVAL, "NULL argument");
    RC(check_sizes(nx, t));
    RC(check_sizes(ny, t));
    if (x_stride < t || y_stride < t) return fail(WBX_EINVAL, "stride below the length");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    return launch_lb_kim(Y, ny, y_stride, X, nx, x_stride, (int)t, out, st);
}

// ------------------------------------------------------------ host drop-ins
int wbx_pairwise(const wbx_metric *metric, const double *X, int64_t nx, int64_t dims, int64_t tx,
                 const double *Y, int64_t ny, int64_t ty, int64_t dim, double *out) {
    RC(check_metric(metric));
    RC(check_sizes(nx, tx));
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    const bool singleton = Y == nullptr;
    DevBuf dx, dy, dout;
    RC(upload(dx, X, sizeof(double) * nx * dims * tx, st));
    if (!singleton) RC(upload(dy, Y, sizeof(double) * ny * dims * ty, st));
    const int64_t ocols = singleton ? nx : ny;
    RC(dout.alloc(sizeof(double) * nx * ocols));
    RC(wbx_dev_pairwise(metric, WBX_F64, (const double *)dx.p + dim * tx, nx, dims * tx, tx,
                        singleton ? nullptr : (const double *)dy.p + dim * ty, ny, dims * ty, ty,
                        (double *)dout.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(out, dout.p, sizeof(double) * nx * ocols, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_paired(const wbx_metric *metric, const double *X, const double *Y, int64_t n,
               int64_t dims, int64_t t, int64_t dim, double *out) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, dy, dout;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(dy, Y, sizeof(double) * n * dims * t, st));
    RC(dout.alloc(sizeof(double) * n));
    RC(wbx_dev_paired(metric, WBX_F64, (const double *)dx.p + dim * t, dims * t,
                      (const double *)dy.p + dim * t, dims * t, n, t, (double *)dout.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(out, dout.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_argmin(const wbx_metric *metric, const double *X, int64_t nx, int64_t dims, int64_t tx,
               const double *Y, int64_t ny, int64_t ty, int64_t dim, int64_t k,
               const double *lower_bound, int64_t *idx, double *dist) {
    RC(check_metric(metric));
    RC(check_sizes(nx, tx));
    RC(check_sizes(ny, ty));
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    if (k < 1) return fail(WBX_EINVAL, "k must be >= 1, got %lld", (long long)k);
    if (k > ny) k = ny;
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, dy, dlb, didx, ddist;
    RC(upload(dx, X, sizeof(double) * nx * dims * tx, st));
    RC(upload(dy, Y, sizeof(double) * ny * dims * ty, st));
    if (lower_bound) RC(upload(dlb, lower_bound, sizeof(double) * nx * ny, st));
    RC(didx.alloc(sizeof(int64_t) * nx * k));
    RC(ddist.alloc(sizeof(double) * nx * k));
    RC(wbx_dev_argmin(metric, WBX_F64, (const double *)dx.p + dim * tx, nx, dims * tx, tx,
                      (const double *)dy.p + dim * ty, ny, dims * ty, ty, k,
                      lower_bound ? (const double *)dlb.p : nullptr, (int64_t *)didx.p,
                      (double *)ddist.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(idx, didx.p, sizeof(int64_t) * nx * k, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(dist, ddist.p, sizeof(double) * nx * k, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_dev_subsequence_profile(const wbx_metric *metric, int dtype, const void *S,
                                const int64_t *s_off, const int64_t *s_len, int64_t n_s,
                                int paired, const void *X, int64_t n, int64_t x_stride, int64_t t,
                                int64_t nwp, double *prof, void *stream) {
    RC(check_subsequence_metric(metric));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (!S || !s_off || !s_len || !X || !prof) return fail(WBX_EINVAL, "NULL argument");
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    int rc = profile_impl(ctx, metric, dtype, S, s_off, s_len, n_s, paired ? 1 : 0, X, n, x_stride,
                          t, nwp, prof, nullptr, st);
    int rc2 = end(ctx, st);
    return rc ? rc : rc2;
}

int wbx_dev_argmin_subsequence(const wbx_metric *metric, int dtype, const void *S,
                               const int64_t *s_off, const int64_t *s_len, const void *X,
                               int64_t n, int64_t x_stride, int64_t t, int64_t k, int64_t *idx,
                               double *dist, void *stream) {
    RC(check_subsequence_metric(metric));
    if (metric->kind > WBX_WDTW)
        return wrap_argmin(metric, dtype, S, s_off, s_len, X, n, x_stride, t, k, idx, dist,
                           stream, NORM_NONE);
    RC(check_sizes(n, t));
    if (!S || !s_off || !s_len || !X || !idx || !dist) return fail(WBX_EINVAL, "NULL argument");
    if (k < 1) return fail(WBX_EINVAL, "k == %lld, must be >= 1", (long long)k);
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    DevGuard g(ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    RC(begin(ctx, st));
    std::vector<int64_t> hl((size_t)n);
    WBX_HIP_CHECK(hipMemcpyAsync(hl.data(), s_len, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    int64_t min_len = t, max_len = 0;
    for (int64_t q = 0; q < n; ++q) {
        if (hl[q] < 1 || hl[q] > t)
            return fail(WBX_EINVAL, "the longest subsequence must be shorter than samples.");
        min_len = hl[q] < min_len ? hl[q] : min_len;
        max_len = hl[q] > max_len ? hl[q] : max_len;
    }
    if (k > t - max_len + 1) return fail(WBX_EINVAL, "k must be less x.shape[-1] - y.shape[-1] + 1.");
    const int64_t nwp = t - min_len + 1;
    // the profile and the heaps live in their own allocation: profile_impl
    // stages its inputs in the shared workspace
    const size_t prof_bytes = ((size_t)n * nwp * 8 + 255) & ~(size_t)255;
    void *buf = nullptr;
    if (hipMalloc(&buf, prof_bytes + (size_t)n * k * 16) != hipSuccess)
        return fail(WBX_ENOMEM, "profile allocation of %zu bytes failed", prof_bytes);
    double *prof = (double *)buf;
    int rc = profile_impl(ctx, metric, dtype, S, s_off, s_len, n, 1, X, n, x_stride, t, nwp, prof,
                          nullptr, st);
    // the heap keeps square roots, as Metric._eadistance returns them
    if (!rc) rc = launch_sqrt_inplace(prof, n * nwp, st);
    if (!rc) rc = launch_argmin_replay(prof, n, nwp, k, nullptr, (char *)buf + prof_bytes, idx,
                                       dist, st);
    const int rc2 = end(ctx, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(buf);
    return rc ? rc : rc2;
}

int wbx_subsequence_profile(const wbx_metric *metric, const double *S, const int64_t *s_off,
                            const int64_t *s_len, int64_t n_s, int paired, const double *X,
                            int64_t n, int64_t dims, int64_t t, int64_t dim, int64_t nwp,
                            double *prof) {
    RC(check_subsequence_metric(metric));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    int64_t total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (s_len[j] < 1 || s_len[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)s_len[j],
                        (long long)t);
        if (s_off[j] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        if (s_off[j] + s_len[j] > total) total = s_off[j] + s_len[j];
    }
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, ds, doff, dlen, dprof;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(ds, S, sizeof(double) * total, st));
    RC(upload(doff, s_off, sizeof(int64_t) * n_s, st));
    RC(upload(dlen, s_len, sizeof(int64_t) * n_s, st));
    RC(dprof.alloc(sizeof(double) * n * nwp));
    RC(wbx_dev_subsequence_profile(metric, WBX_F64, ds.p, (const int64_t *)doff.p,
                                   (const int64_t *)dlen.p, n_s, paired,
                                   (const double *)dx.p + dim * t, n, dims * t, t, nwp,
                                   (double *)dprof.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(prof, dprof.p, sizeof(double) * n * nwp, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_argmin_subsequence(const wbx_metric *metric, const double *S, const int64_t *s_off,
                           const int64_t *s_len, int64_t n, const double *X, int64_t dims,
                           int64_t t, int64_t dim, int64_t k, int64_t *idx, double *dist) {
    RC(check_subsequence_metric(metric));
    RC(check_sizes(n, t));
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    if (k < 1) return fail(WBX_EINVAL, "k == %lld, must be >= 1", (long long)k);
    int64_t total = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (s_len[j] < 1 || s_len[j] > t)
            return fail(WBX_EINVAL, "the longest subsequence must be shorter than samples.");
        if (s_off[j] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        if (s_off[j] + s_len[j] > total) total = s_off[j] + s_len[j];
    }
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, ds, doff, dlen, didx, ddist;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(ds, S, sizeof(double) * total, st));
    RC(upload(doff, s_off, sizeof(int64_t) * n, st));
    RC(upload(dlen, s_len, sizeof(int64_t) * n, st));
    RC(didx.alloc(sizeof(int64_t) * n * k));
    RC(ddist.alloc(sizeof(double) * n * k));
    RC(wbx_dev_argmin_subsequence(metric, WBX_F64, ds.p, (const int64_t *)doff.p,
                                  (const int64_t *)dlen.p, (const double *)dx.p + dim * t, n,
                                  dims * t, t, k, (int64_t *)didx.p, (double *)ddist.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(idx, didx.p, sizeof(int64_t) * n * k, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(dist, ddist.p, sizeof(double) * n * k, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_scaled_subsequence_profile(const wbx_metric *metric, const double *S,
                                   const int64_t *s_off, const int64_t *s_len,
                                   const double *s_mean, const double *s_std, int64_t n_s,
                                   int paired, const double *X, int64_t n, int64_t dims,
                                   int64_t t, int64_t dim, int64_t nwp, double threshold,
                                   double *prof) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    int64_t total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (s_len[j] < 1 || s_len[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)s_len[j],
                        (long long)t);
        if (s_off[j] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        if (s_off[j] + s_len[j] > total) total = s_off[j] + s_len[j];
    }
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, ds, doff, dlen, dmean, dstd, dprof;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(ds, S, sizeof(double) * total, st));
    RC(upload(doff, s_off, sizeof(int64_t) * n_s, st));
    RC(upload(dlen, s_len, sizeof(int64_t) * n_s, st));
    RC(upload(dmean, s_mean, sizeof(double) * n_s, st));
    RC(upload(dstd, s_std, sizeof(double) * n_s, st));
    RC(dprof.alloc(sizeof(double) * n * nwp));
    RC(wbx_dev_scaled_subsequence_profile(
        metric, WBX_F64, ds.p, (const int64_t *)doff.p, (const int64_t *)dlen.p,
        (const double *)dmean.p, (const double *)dstd.p, n_s, paired,
        (const double *)dx.p + dim * t, n, dims * t, t, nwp, threshold, (double *)dprof.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(prof, dprof.p, sizeof(double) * n * nwp, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_scaled_argmin_subsequence(const wbx_metric *metric, const double *S,
                                  const int64_t *s_off, const int64_t *s_len, int64_t n,
                                  const double *X, int64_t dims, int64_t t, int64_t dim,
                                  int64_t k, int64_t *idx, double *dist) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    int64_t total = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (s_len[j] < 1 || s_len[j] > t)
            return fail(WBX_EINVAL, "the longest subsequence must be shorter than samples.");
        if (s_off[j] < 0) return fail(WBX_EINVAL, "negative pattern offset");
        if (s_off[j] + s_len[j] > total) total = s_off[j] + s_len[j];
    }
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, ds, doff, dlen, didx, ddist;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(ds, S, sizeof(double) * total, st));
    RC(upload(doff, s_off, sizeof(int64_t) * n, st));
    RC(upload(dlen, s_len, sizeof(int64_t) * n, st));
    RC(didx.alloc(sizeof(int64_t) * n * k));
    RC(ddist.alloc(sizeof(double) * n * k));
    RC(wbx_dev_scaled_argmin_subsequence(metric, WBX_F64, ds.p, (const int64_t *)doff.p,
                                         (const int64_t *)dlen.p, (const double *)dx.p + dim * t,
                                         n, dims * t, t, k, (int64_t *)didx.p, (double *)ddist.p,
                                         st));
    WBX_HIP_CHECK(hipMemcpyAsync(idx, didx.p, sizeof(int64_t) * n * k, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(dist, ddist.p, sizeof(double) * n * k, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_scaled_subsequence_pairwise(const wbx_metric *metric, const double *S,
                                    const int64_t *s_off, const int64_t *s_len,
                                    const double *s_mean, const double *s_std, int64_t n_s,
                                    const double *X, int64_t n, int64_t dims, int64_t t,
                                    int64_t dim, double *dist, int64_t *idx) {
    RC(check_metric(metric));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    int64_t total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (s_len[j] < 1 || s_len[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)s_len[j],
                        (long long)t);
        if (s_off[j] < 0) return fail(WBX_EINVAL, "negative shapelet offset");
        if (s_off[j] + s_len[j] > total) total = s_off[j] + s_len[j];
    }
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, ds, doff, dlen, dmean, dstd, ddist, didx;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(ds, S, sizeof(double) * total, st));
    RC(upload(doff, s_off, sizeof(int64_t) * n_s, st));
    RC(upload(dlen, s_len, sizeof(int64_t) * n_s, st));
    RC(upload(dmean, s_mean, sizeof(double) * n_s, st));
    RC(upload(dstd, s_std, sizeof(double) * n_s, st));
    RC(ddist.alloc(sizeof(double) * n * n_s));
    RC(didx.alloc(sizeof(int64_t) * n * n_s));
    RC(wbx_dev_scaled_subsequence_pairwise(
        metric, WBX_F64, ds.p, (const int64_t *)doff.p, (const int64_t *)dlen.p,
        (const double *)dmean.p, (const double *)dstd.p, n_s, (const double *)dx.p + dim * t, n,
        dims * t, t, (double *)ddist.p, (int64_t *)didx.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(dist, ddist.p, sizeof(double) * n * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(idx, didx.p, sizeof(int64_t) * n * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

int wbx_subsequence_pairwise(const wbx_metric *metric, const double *S, const int64_t *s_off,
                             const int64_t *s_len, int64_t n_s, const double *X, int64_t n,
                             int64_t dims, int64_t t, int64_t dim, double *dist, int64_t *idx) {
    RC(check_subsequence_metric(metric, true));
    RC(check_sizes(n, t));
    if (n_s < 1) return fail(WBX_EINVAL, "Subsequence y cannot be empty.");
    if (dims < 1 || dim < 0 || dim >= dims)
        return fail(WBX_EINVAL, "The parameter dim must be 0 <= dim < n_dims");
    int64_t total = 0;
    for (int64_t j = 0; j < n_s; ++j) {
        if (s_len[j] < 1 || s_len[j] > t)
            return fail(WBX_EINVAL, "Invalid subsequnce shape (%lld > %lld)", (long long)s_len[j],
                        (long long)t);
        if (s_off[j] < 0) return fail(WBX_EINVAL, "negative shapelet offset");
        if (s_off[j] + s_len[j] > total) total = s_off[j] + s_len[j];
    }
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = ctx->stream;
    DevBuf dx, ds, doff, dlen, ddist, didx;
    RC(upload(dx, X, sizeof(double) * n * dims * t, st));
    RC(upload(ds, S, sizeof(double) * total, st));
    RC(upload(doff, s_off, sizeof(int64_t) * n_s, st));
    RC(upload(dlen, s_len, sizeof(int64_t) * n_s, st));
    RC(ddist.alloc(sizeof(double) * n * n_s));
    RC(didx.alloc(sizeof(int64_t) * n * n_s));
    RC(wbx_dev_subsequence_pairwise(metric, WBX_F64, ds.p, (const int64_t *)doff.p,
                                    (const int64_t *)dlen.p, n_s, (const double *)dx.p + dim * t, n,
                                    dims * t, t, (double *)ddist.p, (int64_t *)didx.p, st));
    WBX_HIP_CHECK(hipMemcpyAsync(dist, ddist.p, sizeof(double) * n * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipMemcpyAsync(idx, didx.p, sizeof(int64_t) * n * n_s, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    return WBX_OK;
}

}  // extern "C"
