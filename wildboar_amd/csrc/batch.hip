// batch.hip -- DTW over a batch of (x, y) pairs of uneven lengths.
//
// Inputs are two arrays of series padded to a common width with wildboar's
// end-of-series marker (utils/variable_len.py:17-115: a NaN whose float32 bit
// pattern ends in 0x9); pair p is (X[p][:lx_p], Y[p][:ly_p]) with its own
// window fraction r_p.  out[p] = DTW(X[p], Y[p]) with DtwMetric's band
// (distance/_elastic.pyx:3167-3186, w = _compute_r(min(lx, ly), r_p)).
//
// Kernels
//  k_eos_lengths  index of the first EOS of each series.
//  k_strip_dtw    one pair per 64-lane wave, two pairs interleaved.  Lane p
//                 owns a strip of R or R+1 consecutive band slots (band
//                 coordinates: slot k <-> column j = i + k - off) and walks the
//                 rows of its pair; the left input of its first slot comes from
//                 lane p-1 (same row) and the up input of its last slot from
//                 lane p+1 (previous row), each by one DPP shift of what the
//                 neighbour computed one step earlier.  Lane p handles row i of
//                 stream z at physical step 2 (i + lag) + (p + z) % 2 with
//                 lag = floor((p + z) / 2): the row chain crosses one lane per
//                 step and every lane is busy every step (even lanes on one
//                 pair while odd lanes are on the other).  Waves are persistent
//                 and pull pairs from a counter over a list sorted by cost.
//  plain kernel   (plain.hip) with per-pair lengths and r: FP64, WDTW, bands
//                 the strip sizes do not cover, and the "plain per-pair"
//                 baseline.
#include <algorithm>
#include <math.h>
#include <numeric>
#include <string.h>
#include <type_traits>
#include <vector>

#include "wbx_internal.h"

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

// ------------------------------------------------------------------ strip
struct StripItem {
    int32_t pair;   // output slot
    int32_t swap;   // 1: rows come from Y[pair], columns from X[pair]
    int32_t lx, ly; // lengths along rows / columns
    int32_t off, B; // band offset and width (slots)
    int32_t n, a;   // lanes used; the first a lanes own R+1 slots, the rest R
};

__device__ __forceinline__ float dpp_from_left(float v) {  // lane p <- lane p-1; lane 0 <- +inf
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_from_right(float v) {  // lane p <- lane p+1; lane 63 <- +inf
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x130, 0xF, 0xF, false));
}

// Four consecutive floats at linear offset `off` of an array (hi = array
// length - 4): one 16-byte load when it lies inside the array, otherwise each
// element clamped on its own.  Elements outside the array are never used
// (rows outside a series are not computed, columns outside it become +inf).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ f4u ld4(const float *__restrict__ base, int off, int hi) {
    if (off >= 0 && off <= hi) return *(const f4u *)(base + off);
    f4u v;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = off + u;
        v[u] = base[e < 0 ? 0 : (e > hi + 3 ? hi + 3 : e)];
    }
    return v;
}

// y[j0 .. j0 + 3] of a series of length ly at offset yo, +inf outside [0, ly).
__device__ __forceinline__ void ld4y(const float *__restrict__ base, int yo, int hi, int j0, int ly,
                                     float *out) {
    const f4u v = ld4(base, yo + j0, hi);
#pragma unroll
    for (int u = 0; u < 4; ++u) out[u] = (unsigned)(j0 + u) < (unsigned)ly ? v[u] : WBX_INF_F;
}

// Arrays of one launch: base, row stride and the last offset a 4-float load
// may start at.
struct StripArrays {
    const float *xb, *yb;
    int xstride, ystride, xhi, yhi;
};

template <int R, int U>
struct StripState {
    static_assert(U == 4, "the strip kernel loads 4 rows / columns at a time");
    static constexpr int S = R + 1;  // slots per lane (slot R is a phantom in short lanes)
    static constexpr int W = S + U - 1;
    float row[S];
    float yw[W];  // yw[q] = y[i + k0 - off + q] at the start of a group
    float f, l;   // first / last real slot of the last computed row
    int i;        // this lane's row at the start of the group
    int lx, ly, k0, kf_local, off;
    bool shortl;
    const float *Xa, *Ya;  // arrays holding the row series and the column series
    int xo, yo, xhi, yhi;  // their offsets in those arrays, last 4-load offsets
};

template <int R, int U>
__device__ __forceinline__ void strip_load_window(StripState<R, U> &st) {
    constexpr int W = StripState<R, U>::W;
    float tmp[(W + 3) / 4 * 4];
#pragma unroll
    for (int q = 0; q < W; q += 4)
        ld4y(st.Ya, st.yo, st.yhi, st.i + st.k0 - st.off + q, st.ly, tmp + q);
#pragma unroll
    for (int q = 0; q < W; ++q) st.yw[q] = tmp[q];
}

// Row -1 of item `it` at lane p (lag = this lane's row delay).
template <int R, int U>
__device__ __forceinline__ void strip_init(StripState<R, U> &st, const StripItem &it, int p,
                                           int lag, const StripArrays &ar) {
    constexpr int S = R + 1;
    const bool unused = p >= it.n;
    st.ly = it.ly;
    st.off = it.off;
    st.i = -lag;
    st.lx = unused ? 0 : it.lx;  // unused lanes never compute; their outputs stay +inf
    st.shortl = p >= it.a;
    st.k0 = p < it.a ? p * S : it.a * S + (p - it.a) * R;
    const int nslots = st.shortl ? R : S;
    const int kf = it.ly - it.lx + it.off;
    st.kf_local = (!unused && kf >= st.k0 && kf < st.k0 + nslots) ? kf - st.k0 : -1;
    st.Xa = it.swap ? ar.yb : ar.xb;
    st.Ya = it.swap ? ar.xb : ar.yb;
    st.xo = it.pair * (it.swap ? ar.ystride : ar.xstride);
    st.yo = it.pair * (it.swap ? ar.xstride : ar.ystride);
    st.xhi = it.swap ? ar.yhi : ar.xhi;
    st.yhi = it.swap ? ar.xhi : ar.yhi;
#pragma unroll
    for (int m = 0; m < S; ++m) st.row[m] = (!unused && st.k0 + m == it.off) ? 0.0f : WBX_INF_F;
    st.f = st.row[0];
    st.l = st.shortl ? st.row[R - 1] : st.row[R];
    strip_load_window<R, U>(st);
}

template <int R, int U>
__device__ __forceinline__ void strip_row(StripState<R, U> &st, int u, float xi, float Lin,
                                          float Uin) {
    constexpr int S = R + 1;
    float left = Lin;
#pragma unroll
    for (int m = 0; m < S; ++m) {
        float up;
        if (m == S - 1)
            up = Uin;
        else if (m == S - 2)
            up = st.shortl ? Uin : st.row[m + 1];
        else
            up = st.row[m + 1];
        const float dg = st.row[m];  // cell (i-1, j-1) has the same band slot
        const float mn = fminf(fminf(up, left), dg);
        const float d = xi - st.yw[m + u];
        const float c = __builtin_fmaf(d, d, mn);
        st.row[m] = c;
        left = c;
    }
    st.f = st.row[0];
    st.l = st.shortl ? st.row[R - 1] : st.row[R];
}

template <int R, int U>
__device__ __forceinline__ void strip_capture(const StripState<R, U> &st, int pair,
                                              float *__restrict__ sq) {
    if (st.kf_local < 0) return;
    float res = WBX_INF_F;
#pragma unroll
    for (int m = 0; m < R + 1; ++m)
        if (m == st.kf_local) res = st.row[m];
    sq[pair] = res;
}

template <int R, int U>
__device__ __forceinline__ void strip_park(StripState<R, U> &st, const StripArrays &ar) {
    st.lx = 0;
    st.i = 1;
    st.ly = 1;
    st.k0 = 0;
    st.off = 0;
    st.kf_local = -1;
    st.shortl = true;
    st.f = st.l = WBX_INF_F;
    st.Xa = ar.xb;
    st.Ya = ar.xb;
    st.xo = st.yo = 0;
    st.xhi = st.yhi = ar.xhi;
#pragma unroll
    for (int m = 0; m < R + 1; ++m) st.row[m] = WBX_INF_F;
#pragma unroll
    for (int q = 0; q < StripState<R, U>::W; ++q) st.yw[q] = WBX_INF_F;
}

// Persistent kernel: two pair streams z = 0, 1 per wave; lane p keeps stream z
// in register set (p + z) % 2, so at every physical step (which works on one
// register set) half the lanes advance stream 0 and the other half stream 1,
// and a lane's neighbours always hold the same stream in the other set.
template <int R, int U>
__global__ __launch_bounds__(64) void k_strip_dtw(const StripItem *__restrict__ items, int n_items,
                                                  int *__restrict__ counter, StripArrays ar,
                                                  float *__restrict__ sq) {
    using St = StripState<R, U>;
    constexpr int W = St::W;
    const int p = threadIdx.x;
    St S[2];
    strip_park<R, U>(S[0], ar);
    strip_park<R, U>(S[1], ar);
    int item[2] = {-1, -1};
    int clock[2] = {0, 0}, end[2] = {0, 0};
    bool done[2] = {false, false};
    bool fresh = true;
    f4u xv[2];
    while (true) {
#pragma unroll
        for (int z = 0; z < 2; ++z) {
            if (!done[z] && clock[z] >= end[z]) {
                const int sset = (p + z) & 1;
                if (item[z] >= 0) {
                    const int pr = items[item[z]].pair;
                    if (sset == 0)
                        strip_capture<R, U>(S[0], pr, sq);
                    else
                        strip_capture<R, U>(S[1], pr, sq);
                }
                int nxt = 0;
                if (p == 0) nxt = atomicAdd(counter, 1);
                nxt = __shfl(nxt, 0);
                if (nxt >= n_items) {
                    done[z] = true;
                    item[z] = -1;
                    if (sset == 0)
                        strip_park<R, U>(S[0], ar);
                    else
                        strip_park<R, U>(S[1], ar);
                } else {
                    const StripItem it = items[nxt];
                    item[z] = nxt;
                    const int lag = (p + z) >> 1;
                    if (sset == 0)
                        strip_init<R, U>(S[0], it, p, lag, ar);
                    else
                        strip_init<R, U>(S[1], it, p, lag, ar);
                    clock[z] = 0;
                    fresh = true;
                    // lane 63 finishes row lx-1 at clock lx + 30 + z
                    end[z] = ((it.lx + 32 + U - 1) / U) * U;
                }
            }
        }
        if (done[0] && done[1]) break;
        if (fresh) {  // a stream (re)started: its prefetched x values are stale
#pragma unroll
            for (int s = 0; s < 2; ++s) xv[s] = ld4(S[s].Xa, S[s].xo + S[s].i, S[s].xhi);
            fresh = false;
        }
        // x and the new y window entries of the next group (used one group later)
        f4u xn[2];
        float yn[2][U];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            xn[s] = ld4(S[s].Xa, S[s].xo + S[s].i + U, S[s].xhi);
            ld4y(S[s].Ya, S[s].yo, S[s].yhi, S[s].i + U + S[s].k0 - S[s].off + (W - U), S[s].ly,
                 yn[s]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const float Lin = dpp_from_left(S[1 - s].l);
                const float Uin = dpp_from_right(S[1 - s].f);
                const int i = S[s].i + u;
                if ((unsigned)i < (unsigned)S[s].lx) strip_row<R, U>(S[s], u, xv[s][u], Lin, Uin);
            }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int q = 0; q < W - U; ++q) S[s].yw[q] = S[s].yw[q + U];
#pragma unroll
            for (int u = 0; u < U; ++u) S[s].yw[W - U + u] = yn[s][u];
            xv[s] = xn[s];
            S[s].i += U;
        }
        clock[0] += U;
        clock[1] += U;
    }
}

template <int R, int U>
static void launch_strip(const StripItem *items, int n_items, int *counter, StripArrays ar,
                         float *sq, int waves, hipStream_t st) {
    hipLaunchKernelGGL((k_strip_dtw<R, U>), dim3((unsigned)waves), dim3(64), 0, st, items, n_items,
                       counter, ar, sq);
}

// Strip sizes compiled in.  A pair of band B uses the smallest R with
// 64 (R + 1) >= B that admits a split n R <= B <= n (R + 1), n <= 64.
static const int kStripR[] = {2, 3, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40};
constexpr int kNumStripR = sizeof(kStripR) / sizeof(kStripR[0]);

static int launch_strip_r(int R, const StripItem *items, int n_items, int *counter,
                          StripArrays ar, float *sq, int waves, hipStream_t st) {
    switch (R) {
        case 2: launch_strip<2, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 3: launch_strip<3, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 4: launch_strip<4, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 6: launch_strip<6, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 8: launch_strip<8, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 10: launch_strip<10, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 12: launch_strip<12, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 14: launch_strip<14, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 16: launch_strip<16, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 20: launch_strip<20, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 24: launch_strip<24, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 28: launch_strip<28, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 32: launch_strip<32, 4>(items, n_items, counter, ar, sq, waves, st); break;
        case 40: launch_strip<40, 4>(items, n_items, counter, ar, sq, waves, st); break;
        default: return fail(WBX_EUNSUPPORTED, "strip size %d not compiled", R);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

static bool strip_split(int B, int min_r, int *R_out, int *n_out, int *a_out) {
    for (int q = 0; q < kNumStripR; ++q) {
        const int R = kStripR[q];
        if (R < min_r || 64 * (R + 1) < B) continue;
        const int n = (B + R) / (R + 1);  // ceil(B / (R + 1))
        if (n * R > B || n > 64) continue;
        *R_out = R;
        *n_out = n;
        *a_out = B - n * R;
        return true;
    }
    return false;
}

__global__ void k_zero_i32(int *p, int n) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ plan
// Host-side plan: per-pair geometry, routing (strip buckets by strip size R,
// or the plain kernel), cost-sorted work lists, device buffers, and the graph
// captured from the last run.
struct BatchPlan {
    wbx_metric metric{};
    int64_t n_pairs = 0, t_pad = 0;
    int policy = 0;
    std::vector<int> bucket_R, bucket_begin, bucket_count, bucket_waves, bucket_stream;
    // Buckets run concurrently on up to kAux streams forked from the caller's
    // stream (default priority) and joined back with events.
    static constexpr int kAux = 4;
    int n_aux = 0;
    hipStream_t aux[kAux] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kAux] = {};
    int n_strip = 0;
    StripItem *d_items = nullptr;
    int *d_counters = nullptr;
    int64_t n_plain = 0;
    int32_t *d_plain_list = nullptr;
    int32_t *d_lx = nullptr, *d_ly = nullptr;
    double *d_r = nullptr;
    float *d_sq = nullptr;
    double *d_sq_plain = nullptr;
    void *d_xconv = nullptr, *d_yconv = nullptr;
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

static void plan_free(BatchPlan *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (int a = 0; a < b->n_aux; ++a) {
        if (b->aux[a]) (void)hipStreamDestroy(b->aux[a]);
        if (b->ev_join[a]) (void)hipEventDestroy(b->ev_join[a]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    void *ptrs[] = {b->d_items, b->d_counters, b->d_plain_list, b->d_lx, b->d_ly,
                    b->d_r,     b->d_sq,       b->d_sq_plain,   b->d_xconv, b->d_yconv};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete b;
}

static int plan_build(BatchPlan *b, const int32_t *lx, const int32_t *ly, const double *r) {
    const int64_t n = b->n_pairs;
    const bool fp32 = b->metric.precision == WBX_FP32;
    const bool dtw = b->metric.kind == WBX_DTW;
    std::vector<StripItem> items;
    std::vector<int> itemR;
    std::vector<int64_t> cost;
    std::vector<int32_t> plain;
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
        const bool strip = fp32 && dtw && b->policy == 0 && rx >= 1 && ry >= 1 && dense_ok &&
                           strip_split(bd.B, 2, &R, &nl, &al);
        if (!strip) {
            plain.push_back((int32_t)q);
            continue;
        }
        StripItem it;
        it.pair = (int32_t)q;
        it.swap = swap ? 1 : 0;
        it.lx = rx;
        it.ly = ry;
        it.off = bd.off;
        it.B = bd.B;
        it.n = nl;
        it.a = al;
        items.push_back(it);
        itemR.push_back(R);
        cost.push_back((int64_t)(rx + 32) * (int64_t)(R + 1));
        b->slots += (int64_t)rx * bd.B;
    }
    std::vector<int> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int u, int v) {
        if (itemR[u] != itemR[v]) return itemR[u] < itemR[v];
        return cost[u] > cost[v];
    });
    std::vector<StripItem> sorted;
    sorted.reserve(items.size());
    for (size_t q = 0; q < order.size(); ++q) {
        const int R = itemR[order[q]];
        if (b->bucket_R.empty() || b->bucket_R.back() != R) {
            b->bucket_R.push_back(R);
            b->bucket_begin.push_back((int)q);
            b->bucket_count.push_back(0);
        }
        b->bucket_count.back()++;
        sorted.push_back(items[order[q]]);
    }
    // Persistent waves per bucket (two pairs per wave in flight), and the
    // buckets spread over the auxiliary streams longest-total-cost first.
    std::vector<int64_t> bcost(b->bucket_R.size(), 0);
    for (size_t k = 0; k < b->bucket_R.size(); ++k) {
        const int waves = (b->bucket_count[k] + 1) / 2;
        const int cap = b->n_cu * 4;
        b->bucket_waves.push_back(waves < cap ? waves : cap);
        for (int q = 0; q < b->bucket_count[k]; ++q)
            bcost[k] += cost[order[b->bucket_begin[k] + q]];
    }
    b->n_aux = (int)std::min<size_t>(BatchPlan::kAux, b->bucket_R.size());
    b->bucket_stream.assign(b->bucket_R.size(), 0);
    {
        std::vector<int> by(b->bucket_R.size());
        std::iota(by.begin(), by.end(), 0);
        std::sort(by.begin(), by.end(), [&](int u, int v) { return bcost[u] > bcost[v]; });
        std::vector<int64_t> load((size_t)std::max(b->n_aux, 1), 0);
        for (int k : by) {
            const int a = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            b->bucket_stream[k] = a;
            load[a] += bcost[k];
        }
    }
    b->n_strip = (int)sorted.size();
    b->n_plain = (int64_t)plain.size();
    RC(dmalloc(&b->d_items, sorted.size()));
    RC(dmalloc(&b->d_counters, b->bucket_R.size()));
    RC(dmalloc(&b->d_plain_list, plain.size()));
    RC(dmalloc(&b->d_lx, (size_t)n));
    RC(dmalloc(&b->d_ly, (size_t)n));
    RC(dmalloc(&b->d_r, (size_t)n));
    RC(dmalloc(&b->d_sq, (size_t)n));
    RC(dmalloc(&b->d_sq_plain, plain.size()));
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

// Everything one run enqueues (graph-capturable: kernels only, no memset).
template <typename T>
static int plan_enqueue(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                        double *out, hipStream_t st, const double *w64, const float *w32) {
    const int64_t n = b->n_pairs, t = b->t_pad;
    PairSpec pl{PAIR_LIST, b->n_plain, 0, b->d_plain_list, b->d_plain_list};
    if (b->metric.precision != WBX_FP32) {
        // FP64, or FP32_ACC64: float64 recursion on inputs rounded to float32
        const double *X64 = (const double *)X, *Y64 = (const double *)Y;
        int64_t xs64 = xs, ys64 = ys;
        if (!std::is_same<T, double>::value) {
            RC((launch_convert_rows<T, double>(X, n, t, xs, (double *)b->d_xconv, t, st)));
            RC((launch_convert_rows<T, double>(Y, n, t, ys, (double *)b->d_yconv, t, st)));
            X64 = (const double *)b->d_xconv;
            Y64 = (const double *)b->d_yconv;
            xs64 = ys64 = t;
        } else if (b->metric.precision == WBX_FP32_ACC64) {
            RC(launch_round_f32_rows((const double *)X, n, t, xs, (double *)b->d_xconv, t, st));
            RC(launch_round_f32_rows((const double *)Y, n, t, ys, (double *)b->d_yconv, t, st));
            X64 = (const double *)b->d_xconv;
            Y64 = (const double *)b->d_yconv;
            xs64 = ys64 = t;
        }
        Series<double> sx{X64, xs64, (int)t, b->d_lx}, sy{Y64, ys64, (int)t, b->d_ly};
        RC(launch_plain_dtw<double>(pl, sx, sy, b->metric.r, w64, b->d_sq_plain, st, b->d_r));
        return launch_finish<double>(pl, b->d_sq_plain, out, 0, st);
    }
    const float *X32 = (const float *)X, *Y32 = (const float *)Y;
    int64_t xs32 = xs, ys32 = ys;
    if (!std::is_same<T, float>::value) {
        RC((launch_convert_rows<T, float>(X, n, t, xs, (float *)b->d_xconv, t, st)));
        RC((launch_convert_rows<T, float>(Y, n, t, ys, (float *)b->d_yconv, t, st)));
        X32 = (const float *)b->d_xconv;
        Y32 = (const float *)b->d_yconv;
        xs32 = ys32 = t;
    }
    Series<float> sx{X32, xs32, (int)t, b->d_lx}, sy{Y32, ys32, (int)t, b->d_ly};
    if (!b->bucket_R.empty()) {
        const int nb = (int)b->bucket_R.size();
        hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, st,
                           b->d_counters, nb);
        WBX_HIP_CHECK(hipEventRecord(b->ev_fork, st));
        for (int a = 0; a < b->n_aux; ++a) WBX_HIP_CHECK(hipStreamWaitEvent(b->aux[a], b->ev_fork, 0));
        // 32-bit element offsets; dense copies when the caller's strides need more
        const float *xb = X32, *yb = Y32;
        int64_t xst = xs32, yst = ys32;
        const int64_t lim = (int64_t)0x7fffffff;
        if ((n - 1) * xst + t > lim) {
            RC((launch_convert_rows<float, float>(X32, n, t, xst, (float *)b->d_xconv, t, st)));
            xb = (const float *)b->d_xconv;
            xst = t;
        }
        if ((n - 1) * yst + t > lim) {
            RC((launch_convert_rows<float, float>(Y32, n, t, yst, (float *)b->d_yconv, t, st)));
            yb = (const float *)b->d_yconv;
            yst = t;
        }
        StripArrays ar{xb, yb, (int)xst, (int)yst, (int)((n - 1) * xst + t - 4),
                       (int)((n - 1) * yst + t - 4)};
        for (int k = 0; k < nb; ++k)
            RC(launch_strip_r(b->bucket_R[k], b->d_items + b->bucket_begin[k], b->bucket_count[k],
                              b->d_counters + k, ar, b->d_sq, b->bucket_waves[k],
                              b->aux[b->bucket_stream[k]]));
    }
    // the plain kernel overlaps the strip buckets on the caller's stream
    float *sqp = (float *)b->d_sq_plain;
    if (b->n_plain > 0) RC(launch_plain_dtw<float>(pl, sx, sy, b->metric.r, w32, sqp, st, b->d_r));
    if (!b->bucket_R.empty()) {
        for (int a = 0; a < b->n_aux; ++a) {
            WBX_HIP_CHECK(hipEventRecord(b->ev_join[a], b->aux[a]));
            WBX_HIP_CHECK(hipStreamWaitEvent(st, b->ev_join[a], 0));
        }
        PairSpec all{PAIR_DIAG, n, n, nullptr, nullptr};
        RC(launch_finish<float>(all, b->d_sq, out, 0, st));
    }
    if (b->n_plain > 0) RC(launch_finish<float>(pl, sqp, out, 0, st));
    return WBX_OK;
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
    b->n_pairs = n_pairs;
    b->t_pad = t_pad;
    b->policy = policy;
    b->n_cu = ctx->n_cu > 0 ? ctx->n_cu : 256;
    int rc = plan_build(b, len_x, len_y, r_per_pair);
    if (!rc && b->n_aux > 0) {
        bool ok = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
        for (int a = 0; ok && a < b->n_aux; ++a)
            ok = hipStreamCreateWithFlags(&b->aux[a], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&b->ev_join[a], hipEventDisableTiming) == hipSuccess;
        if (!ok) rc = fail(WBX_EDEVICE, "batch: stream or event creation failed");
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
    DeviceCtx *ctx = ctx_current();
    if (!ctx) return WBX_EDEVICE;
    hipStream_t st = (hipStream_t)stream;
    const double *w64 = nullptr;
    const float *w32 = nullptr;
    if (b->metric.kind == WBX_WDTW) {
        // weights sized by the padded array width, as the reference sizes
        // them by the array's n_timestep
        const int rc = cached_weights(ctx, b->t_pad, b->metric.g, &w64, &w32);
        if (rc) return rc;
    }
    auto enqueue = [&](hipStream_t s) -> int {
        return dtype == WBX_F64
                   ? plan_enqueue<double>(b, (const double *)X, x_stride, (const double *)Y,
                                          y_stride, out, s, w64, w32)
                   : plan_enqueue<float>(b, (const float *)X, x_stride, (const float *)Y, y_stride,
                                         out, s, w64, w32);
    };
    if (!use_graph) return enqueue(st);
    if (st == nullptr)
        return fail(WBX_EINVAL, "graph capture needs a stream other than the null stream");
    const bool same = b->exec && b->g_x == X && b->g_y == Y && b->g_out == out &&
                      b->g_stream == st && b->g_dtype == dtype && b->g_xs == x_stride &&
                      b->g_ys == y_stride;
    if (!same) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
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
    return WBX_OK;
}

}  // extern "C"
