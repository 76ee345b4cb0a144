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
#include <map>
#include <math.h>
#include <mutex>
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
    int32_t stage;  // -1: read the caller's arrays; else the pair's padded staged rows
};

// Lane p <- lane p-1 (wave_shr:1) / lane p+1 (wave_shl:1); lane 0 / lane 63
// (no source) get +inf.
__device__ __forceinline__ float dpp_from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x130, 0xF, 0xF, false));
}

// Four consecutive floats at linear offset `off` of an array, one 16-byte
// load.  The plan gives the strip kernel only reads that lie inside the arrays
// (strip_reads_inside, staged rows, the zero buffer), so there is no bounds
// handling here and every load is one instruction on every path: the waits
// the compiler places before the rows then cover only loads whose results the
// rows read.
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ f4u ld4(const float *__restrict__ base, int off) {
    return *(const f4u *)(base + off);
}

// Arrays of one launch: X and Y (row strides in elements), the staged rows
// (pairs whose reads could leave X or Y: x row of stage e at
// spad + 2 e sstride, y row sstride later, with room for every read on both
// sides), and kZeroBuf zeros read by lanes outside a pair's band and by idle
// streams.
struct StripArrays {
    const float *xb, *yb, *sb, *zb;
    int xstride, ystride, sstride, spad;
};
constexpr int kZeroBuf = 64;

template <int R, int U>
struct StripState {
    static_assert(U == 4, "the strip kernel loads 4 rows / columns at a time");
    static constexpr int S = R + 1;  // slots per lane (slot R is a phantom in short lanes)
    static constexpr int W = S + U - 1;
    float row[S];
    float yw[W];  // yw[q] = y[i + k0 - off + q] at the start of a group
    float l;      // last real slot of the last computed row
    int i;        // this lane's row at the start of the group
    int adv;      // rows per group: U while a pair runs, 0 when parked
    int lx, ly, k0, kf_local, off;
    bool shortl;
    const float *Xa, *Ya;  // arrays holding the row series and the column series
    int xo, yo;            // the pair's offsets in those arrays
};

// A lane without a pair: x = 0 against columns that are all +inf (ly = 0), so
// every cell it computes is +inf; it reads the same few zeros every group.
template <int R, int U>
__device__ __forceinline__ void strip_park(StripState<R, U> &st, const StripArrays &ar) {
    st.lx = 0;
    st.ly = 0;
    st.i = 0;
    st.adv = 0;
    st.k0 = 0;
    st.off = 0;
    st.kf_local = -1;
    st.shortl = true;
    st.l = WBX_INF_F;
    st.Xa = ar.zb;
    st.Ya = ar.zb;
    st.xo = st.yo = 0;
#pragma unroll
    for (int m = 0; m < R + 1; ++m) st.row[m] = WBX_INF_F;
#pragma unroll
    for (int q = 0; q < StripState<R, U>::W; ++q) st.yw[q] = WBX_INF_F;
}

// y window of the lane's first group, +inf outside [0, ly).
template <int R, int U>
__device__ __forceinline__ void strip_load_window(StripState<R, U> &st) {
    constexpr int W = StripState<R, U>::W;
    float tmp[(W + 3) / 4 * 4];
#pragma unroll
    for (int q = 0; q < W; q += 4) {
        const int j0 = st.i + st.k0 - st.off + q;
        const f4u v = ld4(st.Ya, st.yo + j0);
#pragma unroll
        for (int u = 0; u < 4; ++u) tmp[q + u] = (unsigned)(j0 + u) < (unsigned)st.ly ? v[u] : WBX_INF_F;
    }
#pragma unroll
    for (int q = 0; q < W; ++q) st.yw[q] = tmp[q];
}

// Row -1 of item `it` at lane p (lag = this lane's row delay, at most 63);
// returns x of the lane's first group of rows.
template <int R, int U>
__device__ __forceinline__ f4u strip_init(StripState<R, U> &st, const StripItem &it, int p,
                                          int lag, const StripArrays &ar) {
    constexpr int S = R + 1;
    if (p >= it.n) {  // lanes past the band
        strip_park<R, U>(st, ar);
        return ld4(ar.zb, 0);
    }
    st.shortl = p >= it.a;
    st.k0 = p < it.a ? p * S : it.a * S + (p - it.a) * R;
    const int nslots = st.shortl ? R : S;
    const int kf = it.ly - it.lx + it.off;
    st.kf_local = (kf >= st.k0 && kf < st.k0 + nslots) ? kf - st.k0 : -1;
    st.ly = it.ly;
    st.off = it.off;
    st.i = -lag;
    st.adv = U;
    st.lx = it.lx;
    if (it.stage < 0) {
        st.Xa = it.swap ? ar.yb : ar.xb;
        st.Ya = it.swap ? ar.xb : ar.yb;
        st.xo = it.pair * (it.swap ? ar.ystride : ar.xstride);
        st.yo = it.pair * (it.swap ? ar.xstride : ar.ystride);
    } else {
        const int xr = ar.spad + 2 * it.stage * ar.sstride, yr = xr + ar.sstride;
        st.Xa = st.Ya = ar.sb;
        st.xo = it.swap ? yr : xr;
        st.yo = it.swap ? xr : yr;
    }
#pragma unroll
    for (int m = 0; m < S; ++m) st.row[m] = st.k0 + m == it.off ? 0.0f : WBX_INF_F;
    st.l = st.shortl ? st.row[R - 1] : st.row[R];
    strip_load_window<R, U>(st);
    return ld4(st.Xa, st.xo + st.i);
}

// Slot 0 of row i: computed by every lane (also lanes outside their rows,
// whose result is discarded), so that the first cell of the right neighbour's
// row can be passed left with every lane active.  SEL variants write the
// difference as y - x: (y - x)^2 == (x - y)^2 exactly, and the two forms keep
// the compiler from hoisting the differences of both variants above their
// branch.
template <int R, int U, bool SEL>
__device__ __forceinline__ float strip_first(const StripState<R, U> &st, int u, float xi,
                                             float Lin) {
    const float mn = fminf(fminf(st.row[1], Lin), st.row[0]);
    const float d = SEL ? st.yw[u] - xi : xi - st.yw[u];
    return __builtin_fmaf(d, d, mn);
}

// Slots 1 .. R of row i given slot 0 (c0) and the right neighbour's slot 0
// of row i - 1 (Uin).  SEL: keep the old row where `act` is false (lanes
// before their first row or past their last); the chain runs through the
// computed values either way.
template <int R, int U, bool SEL>
__device__ __forceinline__ void strip_rest(StripState<R, U> &st, int u, float xi, float c0,
                                           float Uin, bool act) {
    constexpr int S = R + 1;
    float left = c0;
    st.row[0] = SEL ? (act ? c0 : st.row[0]) : c0;
#pragma unroll
    for (int m = 1; m < S; ++m) {
        float up;
        if (m == S - 1)
            up = Uin;
        else if (m == S - 2)
            up = st.shortl ? Uin : st.row[m + 1];
        else
            up = st.row[m + 1];
        const float dg = st.row[m];  // cell (i-1, j-1) has the same band slot
        const float mn = fminf(fminf(up, left), dg);
        const float d = SEL ? st.yw[m + u] - xi : xi - st.yw[m + u];
        const float c = __builtin_fmaf(d, d, mn);
        st.row[m] = SEL ? (act ? c : st.row[m]) : c;  // the old slot m is dead after cell m
        left = c;
    }
    st.l = st.shortl ? st.row[R - 1] : st.row[R];
}

// One group of U rows of both streams.  STEADY: every lane of the band is
// inside its rows for all U rows of both streams (lanes past the band compute
// +inf), so no cell needs a select.
template <int R, int U, bool STEADY>
__device__ __forceinline__ void strip_group(StripState<R, U> (&S)[2], const f4u (&xv)[2]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        float c0[2];
        bool act[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const float Lin = dpp_from_left(S[s].l);
            act[s] = (unsigned)(S[s].i + u) < (unsigned)S[s].lx;
            c0[s] = strip_first<R, U, !STEADY>(S[s], u, xv[s][u], Lin);
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const float Uin = dpp_from_right(STEADY || act[s] ? c0[s] : S[s].row[0]);
            strip_rest<R, U, !STEADY>(S[s], u, xv[s][u], c0[s], Uin, act[s]);
        }
        // one row step at a time: the two streams' chains interleave, and
        // the differences of later rows are not all hoisted (register use)
        __builtin_amdgcn_sched_barrier(0);
    }
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

// Persistent kernel: each wave runs two independent pair streams, one per
// register set.  In a stream, lane p works on row i - p while lane 0 works on
// row i: its left input (the left neighbour's last slot of the same row) was
// made one step earlier, and its up input for the last slot (the right
// neighbour's slot 0 of the row above) is made at the start of the same step,
// so a step is one dependent chain of R + 1 cells per stream and the two
// streams' chains overlap.  A stream takes its next pair from the bucket's
// queue as soon as its current one is done.
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
    int first_all[2] = {0, 0}, last_all[2] = {0, 0};  // steady clocks: [first_all, last_all]
    bool done[2] = {false, false};
    f4u xv[2] = {ld4(ar.zb, 0), ld4(ar.zb, 0)};
    while (true) {
#pragma unroll
        for (int z = 0; z < 2; ++z) {
            if (!done[z] && clock[z] >= end[z]) {
                if (item[z] >= 0) strip_capture<R, U>(S[z], items[item[z]].pair, sq);
                int nxt = 0;
                if (p == 0) nxt = atomicAdd(counter, 1);
                nxt = __shfl(nxt, 0);
                if (nxt >= n_items) {
                    done[z] = true;
                    item[z] = -1;
                    strip_park<R, U>(S[z], ar);
                    xv[z] = ld4(ar.zb, 0);
                    first_all[z] = 0;
                    last_all[z] = 0x7fffffff;  // parked lanes compute +inf: always steady
                } else {
                    const StripItem it = items[nxt];
                    item[z] = nxt;
                    xv[z] = strip_init<R, U>(S[z], it, p, p, ar);
                    clock[z] = 0;
                    // lane it.n - 1 finishes row lx - 1 at clock lx + n - 2
                    end[z] = ((it.lx + it.n - 1 + U - 1) / U) * U;
                    // a group at clock c is steady when lane n - 1 is at a row
                    // >= 0 (c >= n - 1) and lane 0 ends it at a row < lx
                    first_all[z] = it.n - 1;
                    last_all[z] = it.lx - U;
                }
            }
        }
        if (done[0] && done[1]) break;
        // x and the new y window entries of the next group: issued here, used
        // after this group's rows (the +inf of columns outside the series is
        // applied there too, so no instruction waits on these loads early)
        f4u xn[2], yraw[2];
        int yj0[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            xn[s] = ld4(S[s].Xa, S[s].xo + S[s].i + S[s].adv);
            yj0[s] = S[s].i + S[s].adv + S[s].k0 - S[s].off + (W - U);
            yraw[s] = ld4(S[s].Ya, S[s].yo + yj0[s]);
        }
        // (both variants in one kernel exceed 256 registers from R = 24 on)
        const bool steady = R <= 20 && clock[0] >= first_all[0] && clock[0] <= last_all[0] &&
                            clock[1] >= first_all[1] && clock[1] <= last_all[1];
        if (steady)
            strip_group<R, U, true>(S, xv);
        else
            strip_group<R, U, false>(S, xv);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int q = 0; q < W - U; ++q) S[s].yw[q] = S[s].yw[q + U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                S[s].yw[W - U + u] =
                    (unsigned)(yj0[s] + u) < (unsigned)S[s].ly ? yraw[s][u] : WBX_INF_F;
            xv[s] = xn[s];
            S[s].i += S[s].adv;
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

// Reads of the strip kernel relative to a pair's row start.  Lane p (lag p)
// runs clocks 0 .. end - 1 with end <= lx + 66 and reads x rows
// [-p, end - p + 3] and y columns j = i + k0 - off + q' with i in
// [-p, end - 4 - p], k0 - p <= 63 R and q' <= W - 1: every element offset lies
// in [-64 - off, max(lx + 69, lx + 65 + 63 R + W - off)] (W = R + 4).
static int64_t strip_read_lo(int off) { return -64 - (int64_t)off; }
static int64_t strip_read_hi(int lx, int off, int R) {
    const int64_t W = R + 4;
    return std::max<int64_t>(lx + 69, lx + 65 + 63 * (int64_t)R + W - off);
}
// Whether they stay inside n x t arrays of any row stride >= t for pair q
// (pairs at the two ends of the arrays can fail this: they are staged).
static bool strip_reads_inside(int64_t q, int64_t n, int64_t t, int lx, int off, int R) {
    return q * t + strip_read_lo(off) >= 0 && strip_read_hi(lx, off, R) <= (n - 1 - q) * t + t - 1;
}

// Staged rows: x and y of pair pairs[e] to rows 2e, 2e + 1 of sb (padding
// untouched).
__global__ __launch_bounds__(256) void k_stage_rows(const float *__restrict__ X, int64_t xs,
                                                    const float *__restrict__ Y, int64_t ys,
                                                    const int32_t *__restrict__ pairs, int t,
                                                    float *__restrict__ sb, int sstride, int spad) {
    const int e = blockIdx.y >> 1, side = blockIdx.y & 1;
    const int64_t q = pairs[e];
    const float *src = side ? Y + q * ys : X + q * xs;
    float *dst = sb + (int64_t)blockIdx.y * sstride + spad;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < t; j += gridDim.x * 256) dst[j] = src[j];
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
    std::vector<int64_t> bucket_cost;  // sum over the bucket's pairs of (rows + lanes) (R + 1)
    // launch shape: persistent waves per CU (per bucket once autotuned) and
    // streams; bucket_ms: measured time of each bucket alone (autotuner)
    std::vector<int> bucket_wpc;
    std::vector<float> bucket_ms;
    int waves_per_cu = 8, n_streams = kAux;
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
    int n_stage = 0, stage_stride = 0, stage_pad = 0;
    int32_t *d_stage_pairs = nullptr;
    float *d_stage = nullptr;
    float *d_zero = nullptr;  // kZeroBuf zeros
    int32_t *d_lx = nullptr, *d_ly = nullptr;
    double *d_r = nullptr;
    float *d_sq = nullptr;
    double *d_sq_plain = nullptr;
    void *d_xconv = nullptr, *d_yconv = nullptr;
    // WDTW weights sized by t_pad, owned by the plan (a captured graph keeps
    // these pointers, so they live exactly as long as the plan)
    double *d_w64 = nullptr;
    float *d_w32 = nullptr;
    // completion of the last enqueued run or replay: recorded after every
    // run outside stream capture, waited on before anything a run reads or
    // writes (graph, streams, buffers) is destroyed
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
// graph, or reading the plan's buffers from the forked streams).
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
    for (int a = 0; a < b->n_aux; ++a) {
        if (b->aux[a]) (void)hipStreamDestroy(b->aux[a]);
        if (b->ev_join[a]) (void)hipEventDestroy(b->ev_join[a]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    void *ptrs[] = {b->d_items, b->d_counters, b->d_plain_list, b->d_lx,
                    b->d_ly,    b->d_r,        b->d_sq,         b->d_sq_plain,
                    b->d_xconv, b->d_yconv,    b->d_stage_pairs, b->d_stage,
                    b->d_zero,  b->d_w64,      b->d_w32};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete b;
}

// Launch shape of the strip buckets: persistent waves per bucket (two pairs
// per wave in flight, at most waves_per_cu per CU), and the buckets spread
// over the first n_streams auxiliary streams longest-total-cost first.
static int bucket_waves_for(const BatchPlan *b, size_t k, int wpc) {
    const int64_t waves = (b->bucket_count[k] + 1) / 2;
    const int64_t cap = (int64_t)b->n_cu * wpc;
    return (int)(waves < cap ? waves : cap);
}

// Launch shape of the strip buckets: persistent waves per bucket (two pairs
// per wave in flight), and the buckets spread over the first n_streams
// auxiliary streams longest first (by measured time once autotuned, else by
// cost).
static void plan_assign(BatchPlan *b) {
    const size_t nb = b->bucket_R.size();
    b->bucket_waves.assign(nb, 0);
    const bool tuned = b->bucket_wpc.size() == nb && b->bucket_ms.size() == nb;
    for (size_t k = 0; k < nb; ++k)
        b->bucket_waves[k] = bucket_waves_for(b, k, tuned ? b->bucket_wpc[k] : b->waves_per_cu);
    b->bucket_stream.assign(nb, 0);
    const int ns = std::max(1, std::min(b->n_streams, b->n_aux));
    std::vector<double> w(nb);
    for (size_t k = 0; k < nb; ++k) w[k] = tuned ? (double)b->bucket_ms[k] : (double)b->bucket_cost[k];
    std::vector<int> by(nb);
    std::iota(by.begin(), by.end(), 0);
    std::sort(by.begin(), by.end(), [&](int u, int v) { return w[u] > w[v]; });
    std::vector<double> load((size_t)ns, 0.0);
    for (int k : by) {
        const int a = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->bucket_stream[k] = a;
        load[a] += w[k];
    }
}

static int plan_build(BatchPlan *b, const int32_t *lx, const int32_t *ly, const double *r) {
    const int64_t n = b->n_pairs;
    const bool fp32 = b->metric.precision == WBX_FP32;
    const bool dtw = b->metric.kind == WBX_DTW;
    std::vector<StripItem> items;
    std::vector<int> itemR;
    std::vector<int64_t> cost;
    std::vector<int32_t> plain, staged;
    int64_t stage_hi = 0;
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
        it.stage = -1;
        if (!strip_reads_inside(q, n, b->t_pad, rx, bd.off, R)) {
            it.stage = (int32_t)staged.size();
            staged.push_back((int32_t)q);
            stage_hi = std::max(stage_hi, strip_read_hi(rx, bd.off, R));
        }
        items.push_back(it);
        itemR.push_back(R);
        cost.push_back((int64_t)(rx + nl) * (int64_t)(R + 1));
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
    b->bucket_cost.assign(b->bucket_R.size(), 0);
    for (size_t k = 0; k < b->bucket_R.size(); ++k)
        for (int q = 0; q < b->bucket_count[k]; ++q)
            b->bucket_cost[k] += cost[order[b->bucket_begin[k] + q]];
    b->n_aux = (int)std::min<size_t>(BatchPlan::kAux, b->bucket_R.size());
    plan_assign(b);
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
    RC(dmalloc(&b->d_zero, kZeroBuf));
    WBX_HIP_CHECK(hipMemset(b->d_zero, 0, kZeroBuf * sizeof(float)));
    if (!staged.empty()) {
        // a staged row has t_pad + 64 floats before it (-64 - off >= -64 - t_pad)
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
        RC(dmalloc(&b->d_stage, total));
        WBX_HIP_CHECK(hipMemset(b->d_stage, 0, total * sizeof(float)));
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

// FP32 inputs of the FP32 path (float64 inputs converted into the plan's
// buffers).
template <typename T>
static int plan_f32_inputs(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                           hipStream_t st, const float **X32, int64_t *xs32, const float **Y32,
                           int64_t *ys32) {
    const int64_t n = b->n_pairs, t = b->t_pad;
    *X32 = (const float *)X;
    *Y32 = (const float *)Y;
    *xs32 = xs;
    *ys32 = ys;
    if (!std::is_same<T, float>::value) {
        RC((launch_convert_rows<T, float>(X, n, t, xs, (float *)b->d_xconv, t, st)));
        RC((launch_convert_rows<T, float>(Y, n, t, ys, (float *)b->d_yconv, t, st)));
        *X32 = (const float *)b->d_xconv;
        *Y32 = (const float *)b->d_yconv;
        *xs32 = *ys32 = t;
    }
    return WBX_OK;
}

// Arrays the strip kernel reads: 32-bit element offsets (dense copies when
// the caller's strides need more) and the staged rows.
static int plan_strip_arrays(BatchPlan *b, const float *X32, int64_t xs32, const float *Y32,
                             int64_t ys32, hipStream_t st, StripArrays *ar) {
    const int64_t n = b->n_pairs, t = b->t_pad;
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
    if (b->n_stage > 0) {
        const unsigned gx = (unsigned)std::min<int64_t>((t + 255) / 256, 64);
        hipLaunchKernelGGL(k_stage_rows, dim3(gx, (unsigned)(2 * b->n_stage)), dim3(256), 0, st, xb,
                           xst, yb, yst, b->d_stage_pairs, (int)t, b->d_stage, b->stage_stride,
                           b->stage_pad);
        WBX_HIP_CHECK(hipGetLastError());
    }
    *ar = StripArrays{xb,       yb,       b->d_stage,      b->d_zero,
                      (int)xst, (int)yst, b->stage_stride, b->stage_pad};
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
        RC(launch_plain_dtw<double>(pl, sx, sy, b->metric.r, w64, b->d_sq_plain, st, b->d_r,
                                    b->policy == 2));
        return launch_finish<double>(pl, b->d_sq_plain, out, 0, st);
    }
    const float *X32, *Y32;
    int64_t xs32, ys32;
    RC(plan_f32_inputs(b, X, xs, Y, ys, st, &X32, &xs32, &Y32, &ys32));
    Series<float> sx{X32, xs32, (int)t, b->d_lx}, sy{Y32, ys32, (int)t, b->d_ly};
    if (!b->bucket_R.empty()) {
        const int nb = (int)b->bucket_R.size();
        hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, st,
                           b->d_counters, nb);
        StripArrays ar;
        RC(plan_strip_arrays(b, X32, xs32, Y32, ys32, st, &ar));
        // the strip buckets start once the copies above are done
        WBX_HIP_CHECK(hipEventRecord(b->ev_fork, st));
        for (int a = 0; a < b->n_aux; ++a) WBX_HIP_CHECK(hipStreamWaitEvent(b->aux[a], b->ev_fork, 0));
        for (int k = 0; k < nb; ++k)
            RC(launch_strip_r(b->bucket_R[k], b->d_items + b->bucket_begin[k], b->bucket_count[k],
                              b->d_counters + k, ar, b->d_sq, b->bucket_waves[k],
                              b->aux[b->bucket_stream[k]]));
    }
    // the plain kernel overlaps the strip buckets on the caller's stream
    float *sqp = (float *)b->d_sq_plain;
    if (b->n_plain > 0)
        RC(launch_plain_dtw<float>(pl, sx, sy, b->metric.r, w32, sqp, st, b->d_r,
                                   b->policy == 2));
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

int wbx_batch_set_launch(void *plan, int waves_per_cu, int n_streams) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    if (waves_per_cu < 1 || waves_per_cu > 32)
        return fail(WBX_EINVAL, "waves_per_cu == %d, must be in [1, 32]", waves_per_cu);
    if (n_streams < 1 || n_streams > BatchPlan::kAux)
        return fail(WBX_EINVAL, "n_streams == %d, must be in [1, %d]", n_streams, BatchPlan::kAux);
    b->waves_per_cu = waves_per_cu;
    b->n_streams = n_streams;
    b->bucket_wpc.clear();
    b->bucket_ms.clear();
    plan_assign(b);
    plan_drop_graph(b);
    return WBX_OK;
}

}  // extern "C"

namespace wbx {

// Autotuner: for each strip bucket, the persistent waves per CU (pairs in
// flight) timed on the caller's data with the bucket's kernel alone on the
// caller's stream, the fastest kept; the measured times order the buckets
// over the streams (longest first).  Choices are cached for the
// process by the plan's bucket signature (strip sizes, counts and costs,
// precision, sizes, device), so plans of the same workload reuse them without
// timing.
static std::mutex g_tune_mu;
static std::map<uint64_t, std::pair<std::vector<int>, std::vector<float>>> g_tune_cache;

static uint64_t plan_signature(const BatchPlan *b, int device) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix((uint64_t)device);
    mix((uint64_t)b->metric.precision);
    mix((uint64_t)b->n_pairs);
    mix((uint64_t)b->t_pad);
    for (size_t k = 0; k < b->bucket_R.size(); ++k) {
        mix((uint64_t)b->bucket_R[k]);
        mix((uint64_t)b->bucket_count[k]);
        mix((uint64_t)b->bucket_cost[k]);
    }
    return h;
}

template <typename T>
static int plan_autotune(BatchPlan *b, const T *X, int64_t xs, const T *Y, int64_t ys,
                         hipStream_t st, int reps) {
    static const int kWpc[] = {4, 6, 8, 12};
    const float *X32, *Y32;
    int64_t xs32, ys32;
    RC(plan_f32_inputs(b, X, xs, Y, ys, st, &X32, &xs32, &Y32, &ys32));
    StripArrays ar;
    RC(plan_strip_arrays(b, X32, xs32, Y32, ys32, st, &ar));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(WBX_EDEVICE, "autotune: event creation failed");
    }
    const size_t nb = b->bucket_R.size();
    std::vector<int> wpc(nb, b->waves_per_cu);
    std::vector<float> ms(nb, 0.0f);
    int rc = WBX_OK;
    for (size_t k = 0; k < nb && !rc; ++k) {
        int last_waves = -1;
        float best = -1.0f;
        for (int w : kWpc) {
            const int waves = bucket_waves_for(b, k, w);
            if (waves == last_waves) continue;  // the bucket has fewer pairs than this fills
            last_waves = waves;
            auto launch = [&]() -> int {
                hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, b->d_counters + k, 1);
                return launch_strip_r(b->bucket_R[k], b->d_items + b->bucket_begin[k],
                                      b->bucket_count[k], b->d_counters + k, ar, b->d_sq, waves, st);
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
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            t /= (float)reps;
            if (best < 0.0f || t < best) {
                best = t;
                wpc[k] = w;
            }
        }
        ms[k] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->bucket_wpc = wpc;
    b->bucket_ms = ms;
    return WBX_OK;
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
    if (!b->bucket_R.empty()) {
        const uint64_t key = plan_signature(b, ctx->device);
        bool cached = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto f = g_tune_cache.find(key);
            if (f != g_tune_cache.end()) {
                b->bucket_wpc = f->second.first;
                b->bucket_ms = f->second.second;
                cached = true;
            }
        }
        if (!cached) {
            hipStream_t st = (hipStream_t)stream;
            const int rc = dtype == WBX_F64
                               ? plan_autotune<double>(b, (const double *)X, x_stride,
                                                       (const double *)Y, y_stride, st, reps)
                               : plan_autotune<float>(b, (const float *)X, x_stride,
                                                      (const float *)Y, y_stride, st, reps);
            if (rc) return rc;
            total = 0.0f;
            for (float t : b->bucket_ms) total += t;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[key] = {b->bucket_wpc, b->bucket_ms};
        }
        plan_assign(b);
        plan_drop_graph(b);
    }
    if (tuned_ms) *tuned_ms = total;
    return WBX_OK;
}

int wbx_batch_bucket_info(void *plan, int k, int *n_buckets, int *R, int *pairs, int *waves,
                          int *stream, float *ms) {
    BatchPlan *b = (BatchPlan *)plan;
    if (!b) return fail(WBX_EINVAL, "NULL plan");
    const int nb = (int)b->bucket_R.size();
    if (n_buckets) *n_buckets = nb;
    if (k < 0) return WBX_OK;
    if (k >= nb) return fail(WBX_EINVAL, "bucket %d of %d", k, nb);
    if (R) *R = b->bucket_R[k];
    if (pairs) *pairs = b->bucket_count[k];
    if (waves) *waves = b->bucket_waves[k];
    if (stream) *stream = b->bucket_stream[k];
    if (ms) *ms = k < (int)b->bucket_ms.size() ? b->bucket_ms[k] : -1.0f;
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
    const double *w64 = b->d_w64;
    const float *w32 = b->d_w32;
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
                                          y_stride, out, s, w64, w32)
                   : plan_enqueue<float>(b, (const float *)X, x_stride, (const float *)Y, y_stride,
                                         out, s, w64, w32);
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
