// scaled.hip -- z-normalised subsequence scans, float64.
//
// ScaledEuclideanSubsequenceMetric (distance/_metric.pyx:96-160,
// scaled_euclidean_distance :561-651) and ScaledDtwSubsequenceMetric
// (distance/_elastic.pyx:1928-2203, scaled_dtw_subsequence_distance
// :343-482, inner_scaled_dtw_subsequence_distance :263-339) as
// _pairwise_subsequence_distance drives them (_cdistance.pyx:1018-1037).
//
// Window statistics.  The reference normalises window I with the mean and
// std of a running scan, ex += x[k], ex2 += x[k]^2 and, after window I,
// ex -= x[I], ex2 -= x[I]^2; mean = ex / L, std = sqrt(ex2 / L - mean^2)
// (1.0 when that is <= 0).  The rounding of those sums depends on every
// earlier step, so k_window_stats_all replays the scan, one thread per
// (series, pattern length).
//
// Cascade.  Per window the reference computes, in this order, LB_Kim
// (constant_lower_bound :158-226, partial sums returned once they reach the
// best so far), LB_Keogh of the window against the pattern's envelope and
// of the pattern against the series' envelope (cumulative_bound :229-260,
// same early exit), then the DTW, abandoned at row i when the row minimum
// plus the remaining bound cb[i + r + 1] reaches the best.  A window is
// accepted (best := its DTW, index := I) iff every one of those values is
// below the best.  Each is a sum of non-negative terms, so its partial value
// reaches the best iff its full value does; hence the window is accepted iff
//     T(I) = max(K, LK, LK2, max_i(rowmin_i + cb[i + r + 1]), D) < best,
// with every term computed in full.  The kernels compute T and D for every
// window in parallel and replay the acceptance rule over the windows in
// increasing order: the reference's result bit for bit, including windows
// its LB_Kim prunes although their DTW is the smallest (its last-but-one
// term omits the (T[L-2], S[L-1]) cell, and below length 6 the first and
// last three points overlap).  Below length 3 the reference's LB_Kim reads
// outside both buffers; there it is not applied (K = 0).
#pragma clang fp contract(off)

#include <algorithm>
#include <vector>

#include "wbx_internal.h"

namespace wbx {

namespace {

__host__ __device__ inline int warp_width(int L, double r) {
    // _compute_warp_width, distance/_elastic.pyx:1917-1921
    if (r == 1.0) return L - 1;
    return (int)floor((double)L * r);
}

// min / max as v_min_f64 / v_max_f64: equal to the reference's
// b < a ? b : a here -- no operand is NaN (finite inputs, window std > 0,
// pattern std 0 read as 1.0) and none is -0 (costs and bounds are sums of
// squares).
__device__ __forceinline__ double dmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double sqd(double a, double b) {
    const double s = a - b;
    return s * s;
}

// One block per pattern: zs = (S - mean) / std (std 0 read as 1.0,
// persistent_distance, _cdistance.pyx:283-299) and, for DTW, the pattern's
// envelope (find_min_max with the warp width, min / max over [k - w, k + w])
// normalised the same way, as cumulative_bound normalises it.
__global__ __launch_bounds__(256) void k_scaled_patterns(
    const double *__restrict__ S, const int64_t *__restrict__ off, const int64_t *__restrict__ len,
    const double *__restrict__ s_mean, const double *__restrict__ s_std, double r, int with_env,
    double *__restrict__ zs, double *__restrict__ zl, double *__restrict__ zu) {
    const int64_t j = blockIdx.x;
    const int L = (int)len[j];
    const int64_t o = off[j];
    const double *__restrict__ s = S + o;
    // no statistics: the pattern itself ((s - 0) / 1 == s)
    const double mu = s_mean ? s_mean[j] : 0.0;
    const double sd = s_std ? (s_std[j] != 0.0 ? s_std[j] : 1.0) : 1.0;
    const int w = warp_width(L, r);
    for (int k = threadIdx.x; k < L; k += 256) {
        zs[o + k] = (s[k] - mu) / sd;
        if (with_env) {
            const int a = k - w > 0 ? k - w : 0;
            const int b = k + w < L - 1 ? k + w : L - 1;
            double lo = s[a], hi = s[a];
            for (int q = a + 1; q <= b; ++q) {
                const double v = s[q];
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
            zl[o + k] = (lo - mu) / sd;
            zu[o + k] = (hi - mu) / sd;
        }
    }
}

struct ScaledTask {
    const double *X;  // rows of the current chunk
    int64_t xstride;
    int t, L, rw, B;
    int64_t c0, cn;   // first series of the chunk, series in the chunk
    const double *wmean, *wstd;  // (cn, t - L + 1)
    const double *elo, *eup;     // series envelopes (cn, t), DTW only
    const double *zs, *zl, *zu;  // normalised patterns (packed as S)
    const int64_t *s_off;
    const int64_t *order;        // scans: pattern of each task column; profiles: of each
                                 // row (NULL: pattern 0)
    int64_t gc, n_s;
    double *dist;
    int64_t *idx;
    // window profiles (prof != NULL): task row ii writes prof[out_row[ii] * nwp + window]
    const int64_t *out_row;
    double th;                   // threshold^2 of the reference's matching
    int64_t nwp;
    double *prof;
    bool values_are_distances;   // accept: the per-window values are already sqrt'd
    int prof_mode;               // k_scaled_wrap profiles: PROF_DIST / PROF_WRAP_MATCH / PROF_COST
};

enum { PROF_DIST = 0, PROF_WRAP_MATCH = 1, PROF_COST = 2 };

__device__ __forceinline__ int64_t task_id() {
    return (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
}

// Keeps the first strict minimum of each lane, then the wave's smallest
// value, smallest window among equal values.
__device__ __forceinline__ void wave_first_min(double &best, int &best_s) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o);
        const int os = __shfl_xor(best_s, o);
        if (ob < best || (ob == best && os < best_s)) {
            best = ob;
            best_s = os;
        }
    }
}

// scaled_euclidean_distance: one wave per (series, pattern), one window per
// lane; the full sum of ((S[k] - s_mean) / s_std - (x - mean) / std)^2 in k
// order (the partial-sum abandon only drops windows whose full sum is not
// below the best).
__global__ __launch_bounds__(64) void k_scaled_euclid(ScaledTask a) {
    const int64_t q = task_id();
    if (q >= a.cn * a.gc) return;
    const int64_t ii = q / a.gc, jj = q - ii * a.gc;
    const int64_t pj = a.order[jj];
    const int L = a.L, nwin = a.t - L + 1;
    const double *__restrict__ x = a.X + ii * a.xstride;
    const double *__restrict__ z = a.zs + a.s_off[pj];
    const double *__restrict__ mrow = a.wmean + ii * nwin;
    const double *__restrict__ srow = a.wstd + ii * nwin;
    double best = WBX_INF_D;
    int best_s = nwin;
    for (int b0 = 0; b0 < nwin; b0 += 64) {
        const int s = b0 + (int)threadIdx.x;
        const int sc = s < nwin ? s : nwin - 1;
        const double mu = mrow[sc], sd = srow[sc];
        const double *__restrict__ xw = x + sc;
        double d = 0.0;
        for (int k = 0; k < L; ++k) {
            double v = z[k];
            v -= (xw[k] - mu) / sd;
            d += v * v;
        }
        if (s < nwin && d < best) {
            best = d;
            best_s = s;
        }
    }
    wave_first_min(best, best_s);
    if (threadIdx.x == 0) {
        const int64_t o = (a.c0 + ii) * a.n_s + pj;
        a.dist[o] = sqrt(best);
        a.idx[o] = best_s < nwin ? best_s : 0;
    }
}

// LB_Kim of the reference (constant_lower_bound), in full, on the
// normalised window zx[k * 64] (the same quotients the reference forms);
// L >= 3.
__device__ double scaled_kim(const double *__restrict__ z, const double *__restrict__ zx, int L) {
    const double tx0 = zx[0], ty0 = zx[(int64_t)(L - 1) * 64];
    const double sx0 = z[0], sy0 = z[L - 1];
    double d = sqd(tx0, sx0) + sqd(ty0, sy0);
    const double tx1 = zx[64], sx1 = z[1];
    d += dmin(dmin(sqd(tx1, sx0), sqd(tx0, sx1)), sqd(tx1, sx1));
    const double ty1 = zx[(int64_t)(L - 2) * 64], sy1 = z[L - 2];
    d += dmin(dmin(sqd(ty1, sy1), sqd(ty0, sy1)), sqd(ty1, sy1));
    const double tx2 = zx[128], sx2 = z[2];
    d += dmin(dmin(dmin(sqd(tx0, sx2), dmin(sqd(tx1, sx2), sqd(tx2, sx2))), sqd(tx2, sx1)),
              sqd(tx2, sx0));
    const double ty2 = zx[(int64_t)(L - 3) * 64], sy2 = z[L - 3];
    d += dmin(dmin(dmin(sqd(ty0, sy2), dmin(sqd(ty1, sy2), sqd(ty2, sy2))), sqd(ty2, sy1)),
              sqd(ty2, sy0));
    return d;
}

__device__ __forceinline__ double keogh_term(double x, double ls, double us) {
    if (x > us) return sqd(x, us);
    if (x < ls) return sqd(x, ls);
    return 0.0;
}

// The bounds of one window: returns max(K, LK, LK2) and leaves, at
// cb[m * 64] for m in [0, nc), nc = L - rw - 1, the suffix sum
// cb[rw + 1 + m] of the chosen LB_Keogh terms (cb_1 when LK > LK2, else
// cb_2), built from the last position backwards as the reference does.
// zx receives the normalised window at zx[k * 64].
__device__ double scaled_bounds(const ScaledTask &a, const double *__restrict__ z,
                                const double *__restrict__ zl, const double *__restrict__ zu,
                                const double *__restrict__ xw, const double *__restrict__ el,
                                const double *__restrict__ eu, double mu, double sd,
                                double *__restrict__ cb, double *__restrict__ zx) {
    const int L = a.L, rw = a.rw, nc = L - rw - 1;
    double lk = 0.0, lk2 = 0.0;
    for (int k = 0; k < L; ++k) {
        const double x = (xw[k] - mu) / sd;
        zx[(int64_t)k * 64] = x;
        const double d = keogh_term(x, zl[k], zu[k]);
        lk += d;
        if (k > rw) cb[(int64_t)(k - rw - 1) * 64] = d;
    }
    for (int k = 0; k < L; ++k) {
        const double us = (eu[k] - mu) / sd;
        const double ls = (el[k] - mu) / sd;
        const double d = keogh_term(z[k], ls, us);
        lk2 += d;
        if (k > rw) cb[(int64_t)(nc + k - rw - 1) * 64] = d;
    }
    const bool one = lk > lk2;
    const double *src = one ? cb : cb + (int64_t)nc * 64;  // may alias cb
    // suffix sums from the last position down, 8 loads issued before their
    // sums (each slot is read before it is rewritten)
    double acc = 0.0;
    constexpr int KB = 8;
    for (int m0 = nc - 1; m0 >= 0; m0 -= KB) {
        double v[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) v[u] = m0 - u >= 0 ? src[(int64_t)(m0 - u) * 64] : 0.0;
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            if (m0 - u < 0) break;
            acc = m0 - u == nc - 1 ? v[u] : acc + v[u];
            cb[(int64_t)(m0 - u) * 64] = acc;
        }
    }
    const double T = L >= 3 ? scaled_kim(z, zx, L) : 0.0;
    return dmax(dmax(T, lk), lk2);
}

// Replays the reference's acceptance over the 64 windows b0 + l in order.
__device__ __forceinline__ void accept_block(double T, double D, int b0, int nwin, double &best,
                                             int &bi) {
    for (int l = 0; l < 64; ++l) {
        const double Tl = __shfl(T, l), Dl = __shfl(D, l);
        if (b0 + l < nwin && Tl < best) {
            best = Dl;
            bi = b0 + l;
        }
    }
}

// Normalised window value of column j (+inf outside the window).
__device__ __forceinline__ double ldz(const double *__restrict__ zx, int j, int L) {
    const int jc = j < 0 ? 0 : (j >= L ? L - 1 : j);
    const double v = zx[(int64_t)jc * 64];
    return (unsigned)j < (unsigned)L ? v : WBX_INF_D;
}

// One row of the register band: slot k holds column j = i + k - rw.  Slots
// >= B are reset to +inf after the row, so the row minimum over all BP slots
// is the minimum over the band cells (cells outside the window are +inf).
template <int BP>
__device__ __forceinline__ double srow(double (&row)[BP], const double *yw, double xi,
                                       const double (&pen)[8]) {
    double left = WBX_INF_D;
#pragma unroll
    for (int k = 0; k < BP; ++k) {
        const double up = (k + 1 < BP) ? row[k + 1] : WBX_INF_D;
        const double m = dmin(dmin(up, left), row[k]);
        const double v = xi - yw[k];
        const double c = m + v * v;
        row[k] = c;
        left = c;
    }
#pragma unroll
    for (int q = 1; q < 8; ++q) row[BP - 8 + q] = fmax(row[BP - 8 + q], pen[q]);
    double rm = row[0];
#pragma unroll
    for (int k = 1; k < BP; ++k) rm = dmin(rm, row[k]);
    return rm;
}

// Banded DTW (|i - j| <= rw) of z (rows) against the normalised window
// (columns) in registers; D = cell (L-1, L-1), M = max over rows i with
// i + rw < L - 1 of v_i = rowmin_i + cb[i * 64], A = the first v_i >= th
// (the value the reference's abandoned DTW returns), else D.
template <int BP, int U>
__device__ __forceinline__ void scaled_cost(const double *__restrict__ z, int L,
                                            const double *__restrict__ zx, int rw, int B,
                                            const double *__restrict__ cb, double th, double &D,
                                            double &M, double &A) {
    constexpr int W = BP + U - 1;
    const int off = rw;
    double row[BP];
#pragma unroll
    for (int k = 0; k < BP; ++k) row[k] = (k == off) ? 0.0 : WBX_INF_D;
    double pen[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) pen[q] = (BP - 8 + q >= B) ? WBX_INF_D : -WBX_INF_D;
    double yw[W];
#pragma unroll
    for (int m = 0; m < W; ++m) yw[m] = ldz(zx, m - off, L);
    const int last = L - 1 - rw;  // rows i < last are checked against cb
    M = -WBX_INF_D;
    bool ab = false;
    A = WBX_INF_D;
    auto check = [&](double v) {
        M = dmax(M, v);
        if (!ab && v >= th) {
            ab = true;
            A = v;
        }
    };
    double rm = srow<BP>(row, yw, z[0], pen);
    if (0 < last) check(rm + cb[0]);
#pragma unroll
    for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
    yw[W - 1] = ldz(zx, 1 - off + W - 1, L);
    const int rem = (L - 1) % U;
    for (int i = 1; i <= rem; ++i) {
        rm = srow<BP>(row, yw, z[i], pen);
        if (i < last) check(rm + cb[(int64_t)i * 64]);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldz(zx, i + 1 - off + W - 1, L);
    }
    for (int i0 = 1 + rem; i0 < L; i0 += U) {
        double zv[U], yn[U], cv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            zv[u] = z[i0 + u];
            yn[u] = ldz(zx, i0 + U - off + W - U + u, L);
            cv[u] = i0 + u < last ? cb[(int64_t)(i0 + u) * 64] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rm = srow<BP>(row, yw + u, zv[u], pen);
            if (i0 + u < last) check(rm + cv[u]);
        }
#pragma unroll
        for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
        for (int u = 0; u < U; ++u) yw[W - U + u] = yn[u];
    }
    double res = WBX_INF_D;
#pragma unroll
    for (int k = 0; k < BP; ++k)
        if (k == off) res = row[k];
    D = res;
    if (!ab) A = res;
}

// Scaled DTW work items: (task q, block of 64 windows) flattened over a
// grid of persistent waves; block b owns the scratch columns
// scratch[b * per_lane * 64 + m * 64 + lane].  Each lane writes T and D of
// its window to Tb / Db [q * nwin + window]; k_scaled_accept replays the
// acceptance rule per task afterwards.
struct ScaledItems {
    int64_t n_items, nblk;
    double *Tb, *Db;
};

__device__ __forceinline__ void item_task(const ScaledTask &a, const ScaledItems &it, int64_t item,
                                          int64_t &q, int &b0, int64_t &ii, int64_t &pj) {
    q = item / it.nblk;
    b0 = (int)(item - q * it.nblk) * 64;
    ii = q / a.gc;
    if (a.prof)  // profiles: one task per row, the row's own pattern
        pj = a.order ? a.order[ii] : 0;
    else
        pj = a.order[q - ii * a.gc];
}

// The reference's matching records a window when every bound is below th
// and the (possibly abandoned) DTW value A is <= th; the profile holds A
// then, +inf otherwise.
__device__ __forceinline__ void scaled_store(const ScaledTask &a, const ScaledItems &it,
                                             int64_t q, int64_t ii, int s, int nwin, double T,
                                             double M, double D, double A, double Q) {
    if (s >= nwin) return;
    if (a.prof) {
        a.prof[a.out_row[ii] * a.nwp + s] = (Q < a.th && A <= a.th) ? A : WBX_INF_D;
    } else {
        it.Tb[q * nwin + s] = dmax(dmax(T, M), D);
        it.Db[q * nwin + s] = D;
    }
}

// Band in registers (2 rw + 1 <= BP).  Scratch per lane: the suffix bounds
// (2 nc values) and the normalised window (L values).
template <int BP, int U>
__global__ __launch_bounds__(64) void k_scaled_dtw(ScaledTask a, ScaledItems it,
                                                   double *__restrict__ scratch, int64_t per_lane) {
    double *__restrict__ cb = scratch + (int64_t)blockIdx.x * per_lane * 64 + threadIdx.x;
    const int L = a.L, nwin = a.t - L + 1;
    double *__restrict__ zx = cb + (int64_t)2 * (L - a.rw - 1) * 64;
    for (int64_t item = blockIdx.x; item < it.n_items; item += gridDim.x) {
        int64_t q, ii, pj;
        int b0;
        item_task(a, it, item, q, b0, ii, pj);
        const int64_t so = a.s_off[pj];
        const int s = b0 + (int)threadIdx.x;
        const int sc = s < nwin ? s : nwin - 1;
        const double mu = a.wmean[ii * nwin + sc], sd = a.wstd[ii * nwin + sc];
        const double T = scaled_bounds(a, a.zs + so, a.zl + so, a.zu + so,
                                       a.X + ii * a.xstride + sc, a.elo + ii * a.t + sc,
                                       a.eup + ii * a.t + sc, mu, sd, cb, zx);
        double D, M, A;
        scaled_cost<BP, U>(a.zs + so, L, zx, a.rw, a.B, cb, a.th, D, M, A);
        scaled_store(a, it, q, ii, s, nwin, T, M, D, A, T);
    }
}

// Any band: the normalised window and one cost row per lane live in scratch
// (columns j at zx[j * 64], row[j * 64]); the row is updated in place left
// to right with the previous row's value of column j - 1 kept as the
// diagonal.  Columns a row has not reached yet hold +inf.
__global__ __launch_bounds__(64) void k_scaled_dtw_wide(ScaledTask a, ScaledItems it,
                                                        double *__restrict__ scratch,
                                                        int64_t per_lane) {
    double *__restrict__ base = scratch + (int64_t)blockIdx.x * per_lane * 64 + threadIdx.x;
    const int L = a.L, rw = a.rw, nwin = a.t - L + 1;
    const int nc = L - rw - 1;
    double *__restrict__ cb = base;
    double *__restrict__ zx = base + (int64_t)2 * nc * 64;
    double *__restrict__ row = zx + (int64_t)L * 64;
    for (int64_t item = blockIdx.x; item < it.n_items; item += gridDim.x) {
        int64_t q, ii, pj;
        int b0;
        item_task(a, it, item, q, b0, ii, pj);
        const int64_t so = a.s_off[pj];
        const double *__restrict__ z = a.zs + so;
        const int s = b0 + (int)threadIdx.x;
        const int sc = s < nwin ? s : nwin - 1;
        const double mu = a.wmean[ii * nwin + sc], sd = a.wstd[ii * nwin + sc];
        const double T = scaled_bounds(a, z, a.zl + so, a.zu + so, a.X + ii * a.xstride + sc,
                                       a.elo + ii * a.t + sc, a.eup + ii * a.t + sc, mu, sd, cb,
                                       zx);
        for (int j = 0; j < L; ++j) row[(int64_t)j * 64] = WBX_INF_D;
        double M = -WBX_INF_D, A = WBX_INF_D;
        bool ab = false;
        for (int i = 0; i < L; ++i) {
            const double xi = z[i];
            const int j0 = i - rw > 0 ? i - rw : 0;
            const int j1 = i + rw < L - 1 ? i + rw : L - 1;
            double left = WBX_INF_D;
            double diag = i == 0 ? 0.0 : (j0 > 0 ? row[(int64_t)(j0 - 1) * 64] : WBX_INF_D);
            double rm = WBX_INF_D;
            for (int j = j0; j <= j1; ++j) {
                const double up = row[(int64_t)j * 64];
                const double v = xi - zx[(int64_t)j * 64];
                const double c = dmin(dmin(up, left), diag) + v * v;
                diag = up;
                row[(int64_t)j * 64] = c;
                left = c;
                rm = dmin(rm, c);
            }
            if (i < nc) {
                const double v = rm + cb[(int64_t)i * 64];
                M = dmax(M, v);
                if (!ab && v >= a.th) {
                    ab = true;
                    A = v;
                }
            }
        }
        const double D = row[(int64_t)(L - 1) * 64];
        if (!ab) A = D;
        scaled_store(a, it, q, ii, s, nwin, T, M, D, A, T);
    }
}

// One wave per task: the windows' (T, D) in increasing order, 64 at a time.
__global__ __launch_bounds__(64) void k_scaled_accept(ScaledTask a, ScaledItems it) {
    const int64_t q = task_id();
    if (q >= a.cn * a.gc) return;
    const int64_t ii = q / a.gc;
    const int64_t pj = a.order[q - ii * a.gc];
    const int nwin = a.t - a.L + 1;
    double best = WBX_INF_D;
    int bi = -1;
    for (int b0 = 0; b0 < nwin; b0 += 64) {
        const int s = b0 + (int)threadIdx.x;
        const int sc = s < nwin ? s : nwin - 1;
        const double T = it.Tb[q * nwin + sc], D = it.Db[q * nwin + sc];
        accept_block(T, D, b0, nwin, best, bi);
    }
    if (threadIdx.x == 0) {
        const int64_t o = (a.c0 + ii) * a.n_s + pj;
        a.dist[o] = a.values_are_distances ? best : sqrt(best);
        a.idx[o] = bi >= 0 ? bi : 0;
    }
}

// ------------------------------------------------ wrapped metrics
// ScaledSubsequenceMetricWrap (_cdistance.pyx:470-567) around WeightedDtw-,
// DerivativeDtw-, WeightedDerivativeDtw- and AmercingDtwMetric: every window
// is normalised with incremental statistics (IncStats, utils/_stats.pyx:
// 45-92: Welford add, the matching remove, population variance, < 1e-13 ->
// 0, then std 0 -> 1.0) and the wrapped metric's _eadistance compares
// sqrt(cost) with the best so far (strict <).  Its early abandon (a row
// minimum above best^2) only drops windows whose full sqrt(cost) is not
// below the best, so the result is the first window of smallest sqrt(cost).

struct WrapArgs {
    const double *weights;  // WDTW / WDDTW weights (NULL: none)
    double pen;             // ADTW penalty (0 otherwise)
    int deriv;              // average slopes of both buffers
    int band;               // _compute_r(Lw, r)
};

// The wrap's running statistics of one series for pattern length L: mean /
// std of windows [0, t - L].
__device__ void incstats_row(const double *__restrict__ x, int t, int L, double *__restrict__ mo,
                             double *__restrict__ so) {
    const int nwin = t - L + 1;
    double n = 0.0, m = 0.0, ss = 0.0;
    auto add = [&](double v) {
        n += 1.0;
        const double next_m = m + (v - m) / n;
        ss += (v - m) * (v - next_m);
        m = next_m;
    };
    for (int k = 0; k < L - 1; ++k) add(x[k]);
    for (int I = 0; I < nwin; ++I) {
        add(x[I + L - 1]);
        double var = 0.0;
        if (!(n <= 1.0)) {
            var = ss / n;
            if (var < 1e-13) var = 0.0;
        }
        so[I] = var == 0.0 ? 1.0 : sqrt(var);
        mo[I] = m;
        const double v = x[I];
        if (n == 1.0) {
            n = 0.0;
            m = 0.0;
            ss = 0.0;
        } else {
            const double old_m = (n * m - v) / (n - 1.0);
            ss -= (v - m) * (v - old_m);
            m = old_m;
            n -= 1.0;
        }
    }
}

// One thread per (group, series), laid out as k_window_stats_all.
__global__ __launch_bounds__(64) void k_incstats_all(const double *__restrict__ X, int64_t cn,
                                                     int64_t stride, int t, int n_groups,
                                                     const int *__restrict__ lens,
                                                     const int64_t *__restrict__ goff,
                                                     double *__restrict__ mean,
                                                     double *__restrict__ sdev) {
    const int64_t id = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (id >= cn * n_groups) return;
    const int64_t g = id / cn, i = id - g * cn;
    const int L = lens[g];
    const int nwin = t - L + 1;
    incstats_row(X + i * stride, t, L, mean + goff[g] + i * nwin, sdev + goff[g] + i * nwin);
}

// One thread per row, one length (rows of t - L + 1).
__global__ __launch_bounds__(64) void k_incstats_rows(const double *__restrict__ X, int64_t cn,
                                                      int64_t stride, int t, int L,
                                                      double *__restrict__ mean,
                                                      double *__restrict__ sdev) {
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= cn) return;
    const int nwin = t - L + 1;
    incstats_row(X + i * stride, t, L, mean + i * nwin, sdev + i * nwin);
}

// fast_mean_std (utils/_stats.pyx:22-42) of each pattern, one thread each;
// the scaled argmin then reads a 0 std as 1.0.
__global__ __launch_bounds__(64) void k_fast_mean_std(const double *__restrict__ S,
                                                      const int64_t *__restrict__ off,
                                                      const int64_t *__restrict__ len, int64_t n,
                                                      double *__restrict__ mean,
                                                      double *__restrict__ sdev) {
    const int64_t j = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (j >= n) return;
    const double *__restrict__ s = S + off[j];
    const int64_t L = len[j];
    double ex = 0.0, ex2 = 0.0;
    for (int64_t i = 0; i < L; ++i) {
        const double v = s[i];
        ex += v;
        ex2 += v * v;
    }
    const double mu = ex / (double)L;
    const double v2 = ex2 / (double)L - mu * mu;
    mean[j] = mu;
    sdev[j] = v2 > 1e-13 ? sqrt(v2) : 0.0;
}

// One lane per window: the normalised window (or its average slopes) and one
// cost row in scratch; the reference's dtw_distance / adtw_distance band and
// row 0 (weights w[j - 1] there, w[|i - j|] below) on Lw x Lw, updated in
// place.  Writes sqrt(cost) per window to it.Tb (= it.Db).
__global__ __launch_bounds__(64) void k_scaled_wrap(ScaledTask a, ScaledItems it, WrapArgs w,
                                                    double *__restrict__ scratch,
                                                    int64_t per_lane) {
    double *__restrict__ zx = scratch + (int64_t)blockIdx.x * per_lane * 64 + threadIdx.x;
    const int L = a.L, nwin = a.t - L + 1;
    const int Lw = w.deriv ? L - 2 : L;
    double *__restrict__ row = zx + (int64_t)L * 64;
    const int r = w.band;
    for (int64_t item = blockIdx.x; item < it.n_items; item += gridDim.x) {
        int64_t q, ii, pj;
        int b0;
        item_task(a, it, item, q, b0, ii, pj);
        const double *__restrict__ z = a.zs + a.s_off[pj];
        const int s = b0 + (int)threadIdx.x;
        const int sc = s < nwin ? s : nwin - 1;
        // no statistics: the window itself ((x - 0) / 1 == x)
        const double mu = a.wmean ? a.wmean[ii * nwin + sc] : 0.0;
        const double sd = a.wstd ? a.wstd[ii * nwin + sc] : 1.0;
        const double *__restrict__ xw = a.X + ii * a.xstride + sc;
        for (int k = 0; k < L; ++k) zx[(int64_t)k * 64] = (xw[k] - mu) / sd;
        if (w.deriv)  // average_slope (_elastic.pyx:3220-3225), in place
            for (int k = 0; k < Lw; ++k) {
                const double q0 = zx[(int64_t)k * 64], q1 = zx[(int64_t)(k + 1) * 64],
                             q2 = zx[(int64_t)(k + 2) * 64];
                zx[(int64_t)k * 64] = ((q1 - q0) + ((q2 - q0) / 2)) / 2;
            }
        auto xrow = [&](int i) {
            if (!w.deriv) return z[i];
            return ((z[i + 1] - z[i]) + ((z[i + 2] - z[i]) / 2)) / 2;
        };
        // row 0: prefix sum over [0, min(Lw, r))
        const int r0 = Lw < r ? Lw : r;
        {
            const double x0 = xrow(0);
            double acc = 0.0;
            for (int j = 0; j < Lw; ++j) {
                double c = WBX_INF_D;
                if (j < r0) {
                    const double v = x0 - zx[(int64_t)j * 64];
                    const double wt = w.weights ? w.weights[j > 0 ? j - 1 : 0] : 1.0;
                    c = j == 0 ? v * v * wt : acc + v * v * wt;
                    acc = c;
                }
                row[(int64_t)j * 64] = c;
            }
        }
        for (int i = 1; i < Lw; ++i) {
            const double xi = xrow(i);
            const int j0 = i - r + 1 > 0 ? i - r + 1 : 0;
            const int j1 = i + r < Lw ? i + r : Lw;
            double left = WBX_INF_D;
            double diag = j0 > 0 ? row[(int64_t)(j0 - 1) * 64] : WBX_INF_D;
            for (int j = j0; j < j1; ++j) {
                const double up = row[(int64_t)j * 64];
                const double v = xi - zx[(int64_t)j * 64];
                const double wt = w.weights ? w.weights[i > j ? i - j : j - i] : 1.0;
                const double c = dmin(dmin(up + w.pen, left + w.pen), diag) + v * v * wt;
                diag = up;
                row[(int64_t)j * 64] = c;
                left = c;
            }
        }
        if (s < nwin) {
            const double c = row[(int64_t)(Lw - 1) * 64];
            const double d = sqrt(c);
            if (a.prof) {
                double v = d;                                            // PROF_DIST
                if (a.prof_mode == PROF_WRAP_MATCH) v = d < a.th ? d * d : WBX_INF_D;
                if (a.prof_mode == PROF_COST) v = c;
                a.prof[a.out_row[ii] * a.nwp + s] = v;
            } else {
                it.Tb[q * nwin + s] = d;
            }
        }
    }
}

// Window statistics of every pattern length at once: thread (g, i) replays
// the scan of series i for length lens[g] into mean / sdev + goff[g]
// (rows of t - lens[g] + 1).
__global__ __launch_bounds__(64) void k_window_stats_all(const double *__restrict__ X, int64_t cn,
                                                         int64_t stride, int t, int n_groups,
                                                         const int *__restrict__ lens,
                                                         const int64_t *__restrict__ goff,
                                                         double *__restrict__ mean,
                                                         double *__restrict__ sdev) {
    const int64_t id = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (id >= cn * n_groups) return;
    const int64_t g = id / cn, i = id - g * cn;
    const int L = lens[g];
    const int nwin = t - L + 1;
    const double *__restrict__ x = X + i * stride;
    double *__restrict__ mo = mean + goff[g] + i * nwin;
    double *__restrict__ so = sdev + goff[g] + i * nwin;
    const double dl = (double)L;
    double ex = 0.0, ex2 = 0.0;
    for (int k = 0; k < t; ++k) {
        const double v = x[k];
        ex += v;
        ex2 += v * v;
        if (k >= L - 1) {
            const int I = k - (L - 1);
            const double mu = ex / dl;
            const double tmp = ex2 / dl - mu * mu;
            mo[I] = mu;
            so[I] = tmp > 0.0 ? sqrt(tmp) : 1.0;
            const double u = x[I];
            ex -= u;
            ex2 -= u * u;
        }
    }
}

// Scaled euclidean profile: one wave per (task row, 64 windows); the
// reference's partial sum stops once it reaches th (scaled_euclidean_distance_
// matches, _metric.pyx:798-866); the window is recorded when the sum is <= th.
__global__ __launch_bounds__(64) void k_scaled_euclid_prof(ScaledTask a, int64_t nblk) {
    const int64_t item = task_id();
    if (item >= a.cn * nblk) return;
    const int64_t ii = item / nblk;
    const int b0 = (int)(item - ii * nblk) * 64;
    const int64_t pj = a.order ? a.order[ii] : 0;
    const int L = a.L, nwin = a.t - L + 1;
    const int s = b0 + (int)threadIdx.x;
    if (s >= nwin) return;
    const double *__restrict__ z = a.zs + a.s_off[pj];
    const double *__restrict__ xw = a.X + ii * a.xstride + s;
    const double mu = a.wmean[ii * nwin + s], sd = a.wstd[ii * nwin + s];
    double d = 0.0;
    for (int k = 0; k < L; ++k) {
        if (d >= a.th) break;
        double v = z[k];
        v -= (xw[k] - mu) / sd;
        d += v * v;
    }
    a.prof[a.out_row[ii] * a.nwp + s] = d <= a.th ? d : WBX_INF_D;
}

// Window statistics of one pattern length for cn rows (rows of t - L + 1).
__global__ __launch_bounds__(64) void k_window_stats(const double *__restrict__ X, int64_t cn,
                                                     int64_t stride, int t, int L,
                                                     double *__restrict__ mean,
                                                     double *__restrict__ sdev) {
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= cn) return;
    const int nwin = t - L + 1;
    const double *__restrict__ x = X + i * stride;
    double *__restrict__ mo = mean + i * nwin;
    double *__restrict__ so = sdev + i * nwin;
    const double dl = (double)L;
    double ex = 0.0, ex2 = 0.0;
    for (int k = 0; k < t; ++k) {
        const double v = x[k];
        ex += v;
        ex2 += v * v;
        if (k >= L - 1) {
            const int I = k - (L - 1);
            const double mu = ex / dl;
            const double tmp = ex2 / dl - mu * mu;
            mo[I] = mu;
            so[I] = tmp > 0.0 ? sqrt(tmp) : 1.0;
            const double u = x[I];
            ex -= u;
            ex2 -= u * u;
        }
    }
}

// dst row i = X row sel[i] (t values).
__global__ __launch_bounds__(256) void k_gather_rows(const double *__restrict__ X, int64_t stride,
                                                     const int64_t *__restrict__ sel, int64_t cn,
                                                     int t, double *__restrict__ dst) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= cn * t) return;
    const int64_t i = g / t;
    const int k = (int)(g - i * t);
    dst[g] = X[sel[i] * stride + k];
}

constexpr int kMaxBPS = 56;

#define WBX_SCALED_CLASSES(X) X(8) X(16) X(24) X(32) X(40) X(48) X(56)

inline dim3 task_grid(int64_t nq) {
    const int64_t gx = nq < 65535 ? nq : 65535;
    return dim3((unsigned)gx, (unsigned)((nq + gx - 1) / gx));
}

// Scratch doubles per lane for a pattern length: 2 (L - w - 1) suffix-bound
// values and L normalised window values, plus L row values for the wide
// kernel.  At most 4 L.
inline int64_t per_lane_doubles(int64_t L, int rw, bool wide) {
    return 2 * (L - rw - 1) + L + (wide ? L : 0);
}
constexpr size_t kScratchCap = (size_t)512 << 20;  // DTW scratch rows
constexpr size_t kChunkCap = (size_t)512 << 20;    // window statistics + envelopes
constexpr size_t kTDCap = (size_t)256 << 20;       // per-window (T, D)
constexpr int64_t kMaxWaves = 8192;

size_t scratch_bytes(int64_t max_len) {
    const size_t one = (size_t)64 * 8 * (size_t)(4 * max_len);  // at least one wave
    const size_t want = (size_t)kMaxWaves * one;
    return want < kScratchCap ? want : (one > kScratchCap ? one : kScratchCap);
}

int64_t group_waves(int64_t n_items, size_t bytes, int64_t per_lane) {
    int64_t w = (int64_t)(bytes / ((size_t)per_lane * 64 * 8));
    w = w < 1 ? 1 : (w > kMaxWaves ? kMaxWaves : w);
    return n_items < w ? n_items : w;
}

struct Groups {
    std::vector<int64_t> order;             // pattern indices sorted by length
    std::vector<int64_t> start;             // group g = order[start[g] .. start[g + 1])
    std::vector<int> len;                   // length of group g
    int64_t win_total = 0;                  // sum over groups of t - len + 1
};

Groups make_groups(const int64_t *hl, int64_t n_s, int64_t t) {
    Groups G;
    G.order.resize((size_t)n_s);
    for (int64_t j = 0; j < n_s; ++j) G.order[j] = j;
    std::stable_sort(G.order.begin(), G.order.end(),
                     [&](int64_t a, int64_t b) { return hl[a] < hl[b]; });
    for (int64_t k = 0; k < n_s; ++k) {
        const int L = (int)hl[G.order[k]];
        if (k == 0 || L != G.len.back()) {
            G.start.push_back(k);
            G.len.push_back(L);
            G.win_total += t - L + 1;
        }
    }
    G.start.push_back(n_s);
    return G;
}

int64_t chunk_rows(int kind, int64_t n, int64_t t, int64_t win_total) {
    const int64_t per = win_total * 16 + (kind == WBX_DTW ? t * 16 : 0);
    int64_t c = (int64_t)(kChunkCap / (size_t)per);
    if (c < 1) c = 1;
    return c < n ? c : n;
}

}  // namespace

size_t scaled_ws_bytes(int kind, const int64_t *hl, int64_t n_s, int64_t total, int64_t n,
                       int64_t t) {
    const Groups G = make_groups(hl, n_s, t);
    int64_t max_len = G.len.back();
    const int64_t cn = chunk_rows(kind, n, t, G.win_total);
    size_t b = 0;
    auto take = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
    take((size_t)total * 8);                                // zs
    take((size_t)total * 8);                                // zl
    take((size_t)total * 8);                                // zu
    take((size_t)n_s * 8);                                  // order
    take(G.len.size() * 16);                                // group offsets, lengths
    take((size_t)cn * (size_t)G.win_total * 16);            // window mean / std
    if (kind == WBX_DTW) {
        take((size_t)cn * (size_t)t * 16);                  // series envelopes
        take(scratch_bytes(max_len));
        take(kTDCap > (size_t)cn * (size_t)t * 16 ? kTDCap : (size_t)cn * (size_t)t * 16);
    }
    return b;
}

int launch_scaled_subseq(int kind, double r, const double *S, const int64_t *s_off,
                         const int64_t *s_len, const int64_t *hl, int64_t n_s, int64_t total,
                         const double *s_mean, const double *s_std, const double *X,
                         int64_t xstride, int64_t n, int64_t t, double *dist, int64_t *idx,
                         char *ws, bool wide_only, hipStream_t st) {
    if (n <= 0 || n_s <= 0) return WBX_OK;
    if (t > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "scaled scan: series too long");
    const bool dtw = kind == WBX_DTW;
    const Groups G = make_groups(hl, n_s, t);
    const int n_groups = (int)G.len.size();
    const int64_t max_len = G.len.back();
    const int64_t cn_max = chunk_rows(kind, n, t, G.win_total);
    std::vector<int64_t> goff((size_t)n_groups);
    for (int64_t g = 0, w = 0; g < n_groups; ++g) {
        goff[g] = w * cn_max;
        w += t - G.len[g] + 1;
    }
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = ws + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    double *zs = (double *)take((size_t)total * 8);
    double *zl = (double *)take((size_t)total * 8);
    double *zu = (double *)take((size_t)total * 8);
    int64_t *d_order = (int64_t *)take((size_t)n_s * 8);
    char *gmeta = take((size_t)n_groups * 16);
    int64_t *d_goff = (int64_t *)gmeta;
    int *d_glen = (int *)(gmeta + (size_t)n_groups * 8);
    double *wmean = (double *)take((size_t)cn_max * (size_t)G.win_total * 16);
    double *wstd = wmean + cn_max * G.win_total;
    double *elo = nullptr, *eup = nullptr, *scratch = nullptr, *Tb = nullptr;
    const size_t sbytes = scratch_bytes(max_len);
    size_t td_bytes = 0;
    if (dtw) {
        elo = (double *)take((size_t)cn_max * (size_t)t * 16);
        eup = elo + cn_max * t;
        scratch = (double *)take(sbytes);
        td_bytes = kTDCap > (size_t)cn_max * (size_t)t * 16 ? kTDCap : (size_t)cn_max * (size_t)t * 16;
        Tb = (double *)take(td_bytes);
    }
    WBX_HIP_CHECK(hipMemcpyAsync(d_order, G.order.data(), sizeof(int64_t) * n_s,
                                 hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(d_glen, G.len.data(), sizeof(int) * n_groups,
                                 hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(d_goff, goff.data(), sizeof(int64_t) * n_groups,
                                 hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors end with this call
    hipLaunchKernelGGL(k_scaled_patterns, dim3((unsigned)n_s), dim3(256), 0, st, S, s_off, s_len,
                       s_mean, s_std, r, dtw ? 1 : 0, zs, zl, zu);
    WBX_HIP_CHECK(hipGetLastError());
    for (int64_t c0 = 0; c0 < n; c0 += cn_max) {
        const int64_t cn = n - c0 < cn_max ? n - c0 : cn_max;
        const double *xc = X + c0 * xstride;
        const int64_t nst = cn * n_groups;
        hipLaunchKernelGGL(k_window_stats_all, dim3((unsigned)((nst + 63) / 64)), dim3(64), 0, st,
                           xc, cn, xstride, (int)t, n_groups, d_glen, d_goff, wmean, wstd);
        WBX_HIP_CHECK(hipGetLastError());
        int env_w = -1;
        for (int g = 0; g < n_groups; ++g) {
            const int L = G.len[g];
            const int nwin = (int)(t - L + 1);
            ScaledTask a{};
            a.X = xc;
            a.xstride = xstride;
            a.t = (int)t;
            a.L = L;
            a.c0 = c0;
            a.cn = cn;
            a.wmean = wmean + goff[g];
            a.wstd = wstd + goff[g];
            a.zs = zs;
            a.zl = zl;
            a.zu = zu;
            a.s_off = s_off;
            a.n_s = n_s;
            a.dist = dist;
            a.idx = idx;
            a.th = WBX_INF_D;
            const int64_t g0 = G.start[g], g1 = G.start[g + 1];
            if (!dtw) {
                a.order = d_order + g0;
                a.gc = g1 - g0;
                hipLaunchKernelGGL(k_scaled_euclid, task_grid(cn * a.gc), dim3(64), 0, st, a);
                WBX_HIP_CHECK(hipGetLastError());
                continue;
            }
            const int rw = warp_width(L, r);
            if (rw != env_w) {
                const int rc = launch_envelope(xc, cn, xstride, (int)t, rw, elo, eup, st);
                if (rc) return rc;
                env_w = rw;
            }
            a.elo = elo;
            a.eup = eup;
            a.rw = rw;
            a.B = 2 * rw + 1;
            const int BP = (a.B + 7) / 8 * 8;
            const bool wide = wide_only || BP > kMaxBPS;
            const int64_t per_lane = per_lane_doubles(L, rw, wide);
            // patterns of the group in batches whose (T, D) fit the buffer
            int64_t bsz = (int64_t)(td_bytes / ((size_t)cn * (size_t)nwin * 16));
            if (bsz < 1) bsz = 1;
            for (int64_t b0 = g0; b0 < g1; b0 += bsz) {
                a.order = d_order + b0;
                a.gc = g1 - b0 < bsz ? g1 - b0 : bsz;
                ScaledItems it;
                it.nblk = (nwin + 63) / 64;
                it.n_items = cn * a.gc * it.nblk;
                it.Tb = Tb;
                it.Db = Tb + cn * a.gc * nwin;
                const dim3 grid((unsigned)group_waves(it.n_items, sbytes, per_lane));
                if (!wide) {
#define WBX_SCALED_LAUNCH(C)                                                              \
    case C:                                                                               \
        hipLaunchKernelGGL((k_scaled_dtw<C, 4>), grid, dim3(64), 0, st, a, it, scratch, \
                           per_lane);                                                     \
        break;
                    switch (BP) {
                        WBX_SCALED_CLASSES(WBX_SCALED_LAUNCH)
                        default: return fail(WBX_EUNSUPPORTED, "scaled dtw: band %d", a.B);
                    }
#undef WBX_SCALED_LAUNCH
                } else {
                    hipLaunchKernelGGL(k_scaled_dtw_wide, grid, dim3(64), 0, st, a, it, scratch,
                                       per_lane);
                }
                WBX_HIP_CHECK(hipGetLastError());
                hipLaunchKernelGGL(k_scaled_accept, task_grid(cn * a.gc), dim3(64), 0, st, a, it);
                WBX_HIP_CHECK(hipGetLastError());
            }
        }
    }
    return WBX_OK;
}

// ------------------------------------------------------- wrapped metrics
size_t scaled_wrap_ws_bytes(const int64_t *hl, int64_t n_s, int64_t total, int64_t n,
                            int64_t t) {
    const Groups G = make_groups(hl, n_s, t);
    const int64_t max_len = G.len.back();
    const int64_t cn = chunk_rows(WBX_EUCLIDEAN, n, t, G.win_total);
    size_t b = 0;
    auto take = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
    take((size_t)total * 8);                      // zs
    take((size_t)n_s * 8);                        // order
    take(G.len.size() * 16);                      // group offsets, lengths
    take((size_t)cn * (size_t)G.win_total * 16);  // window mean / std
    take(scratch_bytes(max_len));
    take(kTDCap > (size_t)cn * (size_t)t * 8 ? kTDCap : (size_t)cn * (size_t)t * 8);
    return b;
}

int launch_scaled_wrap(int kind, double r, const double *weights, double penalty,
                       const double *S, const int64_t *s_off, const int64_t *s_len,
                       const int64_t *hl, int64_t n_s, int64_t total, const double *s_mean,
                       const double *s_std, const double *X, int64_t xstride, int64_t n,
                       int64_t t, double *dist, int64_t *idx, char *ws, hipStream_t st) {
    if (n <= 0 || n_s <= 0) return WBX_OK;
    if (t > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "scaled scan: series too long");
    const bool deriv = kind == WBX_DDTW || kind == WBX_WDDTW;
    const Groups G = make_groups(hl, n_s, t);
    const int n_groups = (int)G.len.size();
    const int64_t max_len = G.len.back();
    const int64_t cn_max = chunk_rows(WBX_EUCLIDEAN, n, t, G.win_total);
    std::vector<int64_t> goff((size_t)n_groups);
    for (int64_t g = 0, w = 0; g < n_groups; ++g) {
        goff[g] = w * cn_max;
        w += t - G.len[g] + 1;
    }
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = ws + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    double *zs = (double *)take((size_t)total * 8);
    int64_t *d_order = (int64_t *)take((size_t)n_s * 8);
    char *gmeta = take((size_t)n_groups * 16);
    int64_t *d_goff = (int64_t *)gmeta;
    int *d_glen = (int *)(gmeta + (size_t)n_groups * 8);
    double *wmean = (double *)take((size_t)cn_max * (size_t)G.win_total * 16);
    double *wstd = wmean + cn_max * G.win_total;
    const size_t sbytes = scratch_bytes(max_len);
    double *scratch = (double *)take(sbytes);
    const size_t td_bytes =
        kTDCap > (size_t)cn_max * (size_t)t * 8 ? kTDCap : (size_t)cn_max * (size_t)t * 8;
    double *Tb = (double *)take(td_bytes);
    WBX_HIP_CHECK(hipMemcpyAsync(d_order, G.order.data(), sizeof(int64_t) * n_s,
                                 hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(d_glen, G.len.data(), sizeof(int) * n_groups,
                                 hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(d_goff, goff.data(), sizeof(int64_t) * n_groups,
                                 hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors end with this call
    hipLaunchKernelGGL(k_scaled_patterns, dim3((unsigned)n_s), dim3(256), 0, st, S, s_off, s_len,
                       s_mean, s_std, r, 0, zs, (double *)nullptr, (double *)nullptr);
    WBX_HIP_CHECK(hipGetLastError());
    for (int64_t c0 = 0; c0 < n; c0 += cn_max) {
        const int64_t cn = n - c0 < cn_max ? n - c0 : cn_max;
        const double *xc = X + c0 * xstride;
        const int64_t nst = cn * n_groups;
        hipLaunchKernelGGL(k_incstats_all, dim3((unsigned)((nst + 63) / 64)), dim3(64), 0, st, xc,
                           cn, xstride, (int)t, n_groups, d_glen, d_goff, wmean, wstd);
        WBX_HIP_CHECK(hipGetLastError());
        for (int g = 0; g < n_groups; ++g) {
            const int L = G.len[g];
            const int nwin = (int)(t - L + 1);
            const int64_t g0 = G.start[g], g1 = G.start[g + 1];
            ScaledTask a{};
            a.X = xc;
            a.xstride = xstride;
            a.t = (int)t;
            a.L = L;
            a.c0 = c0;
            a.cn = cn;
            a.wmean = wmean + goff[g];
            a.wstd = wstd + goff[g];
            a.zs = zs;
            a.s_off = s_off;
            a.n_s = n_s;
            a.dist = dist;
            a.idx = idx;
            a.th = WBX_INF_D;
            a.values_are_distances = true;
            const int Lw = deriv ? L - 2 : L;
            WrapArgs w;
            w.weights = weights;
            w.pen = penalty;
            w.deriv = deriv ? 1 : 0;
            {
                const double v = floor((double)(Lw > 0 ? Lw : 0) * r);
                w.band = (int)(v > 1.0 ? v : 1.0);  // _compute_r, _elastic.pyx:1924-1925
            }
            const int64_t per_lane = 2 * (int64_t)L;
            int64_t bsz = (int64_t)(td_bytes / ((size_t)cn * (size_t)nwin * 8));
            if (bsz < 1) bsz = 1;
            for (int64_t b0 = g0; b0 < g1; b0 += bsz) {
                a.order = d_order + b0;
                a.gc = g1 - b0 < bsz ? g1 - b0 : bsz;
                ScaledItems it;
                it.nblk = (nwin + 63) / 64;
                it.n_items = cn * a.gc * it.nblk;
                it.Tb = Tb;
                it.Db = Tb;
                if (deriv && L < 3) {
                    // DerivativeDtwMetric._eadistance never accepts below
                    // length 3 (_elastic.pyx:3291-3292): the wrap returns
                    // +inf with the index unset (0 here)
                    const int rc = launch_fill_f64(Tb, cn * a.gc * nwin, WBX_INF_D, st);
                    if (rc) return rc;
                } else {
                    const dim3 grid((unsigned)group_waves(it.n_items, sbytes, per_lane));
                    hipLaunchKernelGGL(k_scaled_wrap, grid, dim3(64), 0, st, a, it, w, scratch,
                                       per_lane);
                    WBX_HIP_CHECK(hipGetLastError());
                }
                hipLaunchKernelGGL(k_scaled_accept, task_grid(cn * a.gc), dim3(64), 0, st, a, it);
                WBX_HIP_CHECK(hipGetLastError());
            }
        }
    }
    return WBX_OK;
}

// Window profiles of the wrapped / plain metrics on normalised buffers, for
// series i against pattern (paired ? i : 0), +inf past the pattern's last
// window: argmin_subsequence_distance(scale=True) (match = false: the
// distance; patterns by fast_mean_std when s_mean is NULL) and the wrap's
// _matches (ScaledSubsequenceMetricWrap._matches, _cdistance.pyx:547-615;
// match = true: the squared distance where distance < threshold, else +inf).
// Kinds: euclidean (the diagonal, band 1), dtw, wdtw, ddtw, wddtw, adtw.
size_t scaled_wrap_profile_ws_bytes(const int64_t *hl, int64_t n_s, int64_t n, int64_t total,
                                    int64_t t) {
    int64_t max_len = 1;
    for (int64_t i = 0; i < n_s; ++i) max_len = hl[i] > max_len ? hl[i] : max_len;
    const int64_t per = t * 8 + t * 16;
    int64_t cn = (int64_t)(kChunkCap / (size_t)per);
    cn = cn < 1 ? 1 : (cn < n ? cn : n);
    size_t b = 0;
    auto take = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
    take((size_t)total * 8);                 // zs
    take((size_t)n_s * 16);                  // pattern mean / std
    take((size_t)n * 8);                     // sel
    take((size_t)cn * (size_t)t * 8);        // gathered rows
    take((size_t)cn * (size_t)t * 16);       // window mean / std
    take(scratch_bytes(max_len));
    return b;
}

int launch_scaled_wrap_profile(int kind, double r, const double *weights, double penalty,
                               const double *S, const int64_t *s_off, const int64_t *s_len,
                               const int64_t *hl, int64_t n_s, int64_t total, int norm,
                               const double *s_mean, const double *s_std, int paired,
                               const double *X, int64_t xstride, int64_t n, int64_t t,
                               int64_t nwp, int mode, int band_add, double threshold,
                               double *prof, char *ws, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    if (t > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "scaled profile: series too long");
    const bool deriv = kind == WBX_DDTW || kind == WBX_WDDTW;
    std::vector<int64_t> sel((size_t)n), start;
    std::vector<int> lens;
    auto len_of = [&](int64_t i) { return paired ? hl[i] : hl[0]; };
    for (int64_t i = 0; i < n; ++i) sel[i] = i;
    std::stable_sort(sel.begin(), sel.end(),
                     [&](int64_t a, int64_t b) { return len_of(a) < len_of(b); });
    for (int64_t k = 0; k < n; ++k) {
        const int L = (int)len_of(sel[k]);
        if (k == 0 || L != lens.back()) {
            start.push_back(k);
            lens.push_back(L);
        }
    }
    start.push_back(n);
    int64_t max_len = 1;
    for (int L : lens) max_len = L > max_len ? L : max_len;
    const int64_t per = t * 8 + t * 16;
    int64_t cn_max = (int64_t)(kChunkCap / (size_t)per);
    cn_max = cn_max < 1 ? 1 : (cn_max < n ? cn_max : n);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = ws + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    double *zs = (double *)take((size_t)total * 8);
    double *pmean = (double *)take((size_t)n_s * 16);
    double *pstd = pmean + n_s;
    int64_t *d_sel = (int64_t *)take((size_t)n * 8);
    double *rows = (double *)take((size_t)cn_max * (size_t)t * 8);
    double *wmean = (double *)take((size_t)cn_max * (size_t)t * 16);
    double *wstd = wmean + cn_max * t;
    const size_t sbytes = scratch_bytes(max_len);
    double *scratch = (double *)take(sbytes);
    WBX_HIP_CHECK(hipMemcpyAsync(d_sel, sel.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                                 st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));  // sel ends with this call
    int rc = launch_fill_f64(prof, n * nwp, WBX_INF_D, st);
    if (rc) return rc;
    if (norm == NORM_FAST) {
        // the scaled argmin normalises its patterns with fast_mean_std
        hipLaunchKernelGGL(k_fast_mean_std, dim3((unsigned)((n_s + 63) / 64)), dim3(64), 0, st, S,
                           s_off, s_len, n_s, pmean, pstd);
        WBX_HIP_CHECK(hipGetLastError());
        s_mean = pmean;
        s_std = pstd;
    } else if (norm == NORM_NONE) {
        s_mean = nullptr;
        s_std = nullptr;
    }
    hipLaunchKernelGGL(k_scaled_patterns, dim3((unsigned)n_s), dim3(256), 0, st, S, s_off, s_len,
                       s_mean, s_std, r, 0, zs, (double *)nullptr, (double *)nullptr);
    WBX_HIP_CHECK(hipGetLastError());
    for (size_t g = 0; g < lens.size(); ++g) {
        const int L = lens[g];
        const int nwin = (int)(t - L + 1);
        const int Lw = deriv ? L - 2 : L;
        if (deriv && L < 3) continue;  // never accepted: the profile stays +inf
        WrapArgs w;
        w.weights = weights;
        w.pen = penalty;
        w.deriv = deriv ? 1 : 0;
        if (kind == WBX_EUCLIDEAN) {
            w.band = 1;  // the diagonal: euclidean_distance's sequential sum
        } else {
            const double v = floor((double)(Lw + band_add) * r);
            w.band = (int)(v > 1.0 ? v : 1.0);  // _compute_r, _elastic.pyx:1924-1925
        }
        for (int64_t c0 = start[g]; c0 < start[g + 1]; c0 += cn_max) {
            const int64_t cn = start[g + 1] - c0 < cn_max ? start[g + 1] - c0 : cn_max;
            hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((cn * t + 255) / 256)), dim3(256), 0,
                               st, X, xstride, d_sel + c0, cn, (int)t, rows);
            WBX_HIP_CHECK(hipGetLastError());
            if (norm != NORM_NONE) {
                hipLaunchKernelGGL(k_incstats_rows, dim3((unsigned)((cn + 63) / 64)), dim3(64), 0,
                                   st, rows, cn, t, (int)t, L, wmean, wstd);
                WBX_HIP_CHECK(hipGetLastError());
            }
            ScaledTask a{};
            a.X = rows;
            a.xstride = t;
            a.t = (int)t;
            a.L = L;
            a.cn = cn;
            a.wmean = norm != NORM_NONE ? wmean : nullptr;
            a.wstd = norm != NORM_NONE ? wstd : nullptr;
            a.zs = zs;
            a.s_off = s_off;
            a.order = paired ? d_sel + c0 : nullptr;
            a.gc = 1;
            a.n_s = n_s;
            a.out_row = d_sel + c0;
            a.th = threshold;
            a.prof_mode = mode;
            a.nwp = nwp;
            a.prof = prof;
            ScaledItems it{};
            it.nblk = (nwin + 63) / 64;
            it.n_items = cn * it.nblk;
            const int64_t per_lane = 2 * (int64_t)L;
            const dim3 grid((unsigned)group_waves(it.n_items, sbytes, per_lane));
            hipLaunchKernelGGL(k_scaled_wrap, grid, dim3(64), 0, st, a, it, w, scratch, per_lane);
            WBX_HIP_CHECK(hipGetLastError());
        }
    }
    return WBX_OK;
}

// ------------------------------------------------------------ profiles
// Task rows are series; paired: pattern i for series i, else pattern 0.
// Rows are grouped by their pattern's length, gathered into a contiguous
// chunk, and scanned with the threshold of the reference's matching.
struct ProfPlan {
    std::vector<int64_t> sel;    // series, grouped by pattern length
    std::vector<int64_t> start;  // group g = sel[start[g] .. start[g + 1])
    std::vector<int> len;
    int64_t cn_max = 0;
};

static ProfPlan prof_plan(int kind, const int64_t *hl, int paired, int64_t n, int64_t t) {
    ProfPlan P;
    P.sel.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) P.sel[i] = i;
    auto len_of = [&](int64_t i) { return paired ? hl[i] : hl[0]; };
    std::stable_sort(P.sel.begin(), P.sel.end(),
                     [&](int64_t a, int64_t b) { return len_of(a) < len_of(b); });
    for (int64_t k = 0; k < n; ++k) {
        const int L = (int)len_of(P.sel[k]);
        if (k == 0 || L != P.len.back()) {
            P.start.push_back(k);
            P.len.push_back(L);
        }
    }
    P.start.push_back(n);
    // per row: the gathered series, its window statistics, its envelope
    const int64_t per = t * 8 + t * 16 + (kind == WBX_DTW ? t * 16 : 0);
    int64_t c = (int64_t)(kChunkCap / (size_t)per);
    c = c < 1 ? 1 : c;
    P.cn_max = c < n ? c : n;
    return P;
}

size_t scaled_profile_ws_bytes(int kind, const int64_t *hl, int64_t n_s, int64_t total,
                               int paired, int64_t n, int64_t t) {
    const ProfPlan P = prof_plan(kind, hl, paired, n, t);
    int64_t max_len = 1;
    for (int L : P.len) max_len = L > max_len ? L : max_len;
    size_t b = 0;
    auto take = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
    take((size_t)total * 8);
    take((size_t)total * 8);
    take((size_t)total * 8);
    take((size_t)n * 8);                              // sel
    take((size_t)P.cn_max * (size_t)t * 8);           // gathered rows
    take((size_t)P.cn_max * (size_t)t * 16);          // window mean / std
    if (kind == WBX_DTW) {
        take((size_t)P.cn_max * (size_t)t * 16);      // envelopes
        take(scratch_bytes(max_len));
    }
    (void)n_s;
    return b;
}

int launch_scaled_profile(int kind, double r, const double *S, const int64_t *s_off,
                          const int64_t *s_len, const int64_t *hl, int64_t n_s, int64_t total,
                          const double *s_mean, const double *s_std, int paired, const double *X,
                          int64_t xstride, int64_t n, int64_t t, int64_t nwp, double threshold,
                          double *prof, char *ws, bool wide_only, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    if (t > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "scaled profile: series too long");
    const bool dtw = kind == WBX_DTW;
    const ProfPlan P = prof_plan(kind, hl, paired, n, t);
    int64_t max_len = 1;
    for (int L : P.len) max_len = L > max_len ? L : max_len;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = ws + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    double *zs = (double *)take((size_t)total * 8);
    double *zl = (double *)take((size_t)total * 8);
    double *zu = (double *)take((size_t)total * 8);
    int64_t *d_sel = (int64_t *)take((size_t)n * 8);
    const int64_t cn_max = P.cn_max;
    double *rows = (double *)take((size_t)cn_max * (size_t)t * 8);
    double *wmean = (double *)take((size_t)cn_max * (size_t)t * 16);
    double *wstd = wmean + cn_max * t;
    double *elo = nullptr, *eup = nullptr, *scratch = nullptr;
    const size_t sbytes = scratch_bytes(max_len);
    if (dtw) {
        elo = (double *)take((size_t)cn_max * (size_t)t * 16);
        eup = elo + cn_max * t;
        scratch = (double *)take(sbytes);
    }
    WBX_HIP_CHECK(hipMemcpyAsync(d_sel, P.sel.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                                 st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));  // P.sel ends with this call
    int rc = launch_fill_f64(prof, n * nwp, WBX_INF_D, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scaled_patterns, dim3((unsigned)n_s), dim3(256), 0, st, S, s_off, s_len,
                       s_mean, s_std, r, dtw ? 1 : 0, zs, zl, zu);
    WBX_HIP_CHECK(hipGetLastError());
    const double th = threshold * threshold;
    for (size_t g = 0; g < P.len.size(); ++g) {
        const int L = P.len[g];
        const int nwin = (int)(t - L + 1);
        const int rw = warp_width(L, r);
        for (int64_t c0 = P.start[g]; c0 < P.start[g + 1]; c0 += cn_max) {
            const int64_t cn = P.start[g + 1] - c0 < cn_max ? P.start[g + 1] - c0 : cn_max;
            hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((cn * t + 255) / 256)), dim3(256), 0,
                               st, X, xstride, d_sel + c0, cn, (int)t, rows);
            WBX_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_window_stats, dim3((unsigned)((cn + 63) / 64)), dim3(64), 0, st,
                               rows, cn, t, (int)t, L, wmean, wstd);
            WBX_HIP_CHECK(hipGetLastError());
            ScaledTask a{};
            a.X = rows;
            a.xstride = t;
            a.t = (int)t;
            a.L = L;
            a.c0 = 0;
            a.cn = cn;
            a.wmean = wmean;
            a.wstd = wstd;
            a.zs = zs;
            a.zl = zl;
            a.zu = zu;
            a.s_off = s_off;
            a.order = paired ? d_sel + c0 : nullptr;
            a.gc = 1;
            a.n_s = n_s;
            a.out_row = d_sel + c0;
            a.th = th;
            a.nwp = nwp;
            a.prof = prof;
            const int64_t nblk = (nwin + 63) / 64;
            if (!dtw) {
                hipLaunchKernelGGL(k_scaled_euclid_prof, task_grid(cn * nblk), dim3(64), 0, st, a,
                                   nblk);
                WBX_HIP_CHECK(hipGetLastError());
                continue;
            }
            rc = launch_envelope(rows, cn, t, (int)t, rw, elo, eup, st);
            if (rc) return rc;
            a.elo = elo;
            a.eup = eup;
            a.rw = rw;
            a.B = 2 * rw + 1;
            const int BP = (a.B + 7) / 8 * 8;
            const bool wide = wide_only || BP > kMaxBPS;
            const int64_t per_lane = per_lane_doubles(L, rw, wide);
            ScaledItems it{};
            it.nblk = nblk;
            it.n_items = cn * nblk;
            const dim3 grid((unsigned)group_waves(it.n_items, sbytes, per_lane));
            if (!wide) {
#define WBX_SCALED_LAUNCH(C)                                                              \
    case C:                                                                               \
        hipLaunchKernelGGL((k_scaled_dtw<C, 4>), grid, dim3(64), 0, st, a, it, scratch, \
                           per_lane);                                                     \
        break;
                switch (BP) {
                    WBX_SCALED_CLASSES(WBX_SCALED_LAUNCH)
                    default: return fail(WBX_EUNSUPPORTED, "scaled dtw: band %d", a.B);
                }
#undef WBX_SCALED_LAUNCH
            } else {
                hipLaunchKernelGGL(k_scaled_dtw_wide, grid, dim3(64), 0, st, a, it, scratch,
                                   per_lane);
            }
            WBX_HIP_CHECK(hipGetLastError());
        }
    }
    return WBX_OK;
}

}  // namespace wbx
