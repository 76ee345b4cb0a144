// lane64.hip -- float64 DTW / WDTW with one window (or pair) per lane and the
// band row in registers, bit-identical to the reference.
//
// Same register-band layout as lane.hip (slot k of row i <-> column
// j = i + k - off; up = slot k+1 of the previous row, left = slot k-1 of this
// row, diag = slot k of the previous row), but with the reference's cell
// arithmetic (distance/_elastic.pyx:925-928, contraction off):
//     v = x_i - y_j ;  c = min(min(up, left), diag) + (v * v) * w
// where w = 1 for DTW, w[|i - j|] for WDTW and w[j - 1] in row 0
// (_elastic.pyx:898-903).  Every operation is exact-rounded IEEE float64 in
// the same order as the reference, min is exact, and (x - y)^2 == (y - x)^2,
// so each cell equals the reference's bit for bit.
//
// Columns outside [0, ly) read +inf, which makes their cells +inf (the
// weighted form needs every weight > 0 so that inf * w stays +inf; the host
// routes other weight vectors to the block kernels).  Slots past the band
// width B (padding up to BP) are reset to +inf after every row.
//
// Register classes BP = 8 .. 104 slots for DTW and ADTW (rows past 56 live
// in VGPRs plus AGPRs, 88..104 with a few scratch slots), 8 .. 56 for WDTW
// (the per-slot weights take registers too).
// Kernels:
//   k_lane64_pairs   -- pairwise / paired / argmin rows with one geometry
//   k_profile_lane64 -- window profiles: row q, windows 64 b .. 64 b + 63 of
//                       pattern (paired ? q : 0); tasks of another band class
//                       return at once (one launch per class present)
//   k_subseq_lane64  -- minimum over windows + first argmin per (sample,
//                       shapelet) (dtw_subsequence_distance,
//                       _elastic.pyx:622-655), one wave per task
#pragma clang fp contract(off)

#include "lane64_cost.h"

namespace wbx {

namespace {

template <int BP, int U, bool WT, bool AD>
__global__ __launch_bounds__(64) void k_lane64_pairs(PairSpec ps, Series<double> xs,
                                                     Series<double> ys, int off, int B, int kf,
                                                     const double *__restrict__ weights, int wn,
                                                     double pen, double *__restrict__ sq) {
    const int64_t p0 = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int64_t p = p0 < ps.n_pairs ? p0 : ps.n_pairs - 1;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const double c = lane64_cost<BP, U, WT, AD, lane64_prefetch<0, BP, WT, AD>()>(
        xs.row(a), xs.len, ys.row(b), ys.len, off, B, kf, weights, wn, pen);
    if (p0 < ps.n_pairs) sq[p0] = c;
}

// Band slots of a pattern of length L (both sides L): B = 2 w - 1.
__device__ __forceinline__ int pattern_B(int L, double r) { return 2 * band_w(L, L, r) - 1; }

template <int BP, int U, bool WT>
__global__ __launch_bounds__(64) void k_profile_lane64(Shapelets sh, int paired,
                                                       Series<double> xs, int64_t n, int64_t nwp,
                                                       int nblk, double r,
                                                       const double *__restrict__ weights,
                                                       double *__restrict__ prof) {
    const int64_t q = blockIdx.x / nblk;
    const int b = (int)(blockIdx.x - q * nblk);
    const int64_t j = paired ? q : 0;
    const int L = (int)sh.len[j];
    const int B = pattern_B(L, r);
    if ((B + 7) / 8 * 8 != BP) return;  // another class's launch
    const double *__restrict__ S = (const double *)sh.S + sh.off[j];
    const int t = xs.len;
    const int nwin = t - L + 1;
    const int s = b * 64 + threadIdx.x;
    const int sc = s < nwin ? s : nwin - 1;
    const int off = (B - 1) / 2;
    const double c = lane64_cost<BP, U, WT, false, lane64_prefetch<1, BP, WT, false>()>(
        S, L, xs.row(q) + sc, L, off, B, off, weights, t);
    if (s < nwp) prof[q * nwp + s] = s < nwin ? c : WBX_INF_D;
}

// band_add: the band comes from pattern length L + band_add (2 for the
// derivative metrics, whose patterns are average slopes of length L + 2 series,
// distance/_elastic.pyx:2490-2493).  AD: the amercing recursion with penalty
// pen (adtw_subsequence_distance, :701-734).
template <int BP, int U, bool WT, bool AD>
__global__ __launch_bounds__(64) void k_subseq_lane64(Shapelets sh, Series<double> xs, int64_t n,
                                                      double r, const double *__restrict__ weights,
                                                      int band_add, double pen,
                                                      double *__restrict__ dist,
                                                      int64_t *__restrict__ idx) {
    const int64_t q = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (q >= n * sh.n) return;
    const int64_t i = q / sh.n, j = q - i * sh.n;
    const int L = (int)sh.len[j];
    if (L < 1) return;  // handled by the caller
    const int B = pattern_B(L + band_add, r);
    if ((B + 7) / 8 * 8 != BP) return;
    const double *__restrict__ S = (const double *)sh.S + sh.off[j];
    const int t = xs.len;
    const int nwin = t - L + 1;
    const int off = (B - 1) / 2;
    const double *__restrict__ x = xs.row(i);
    // each lane keeps the first strict minimum over its windows (taken in
    // increasing order); the wave then takes the smallest value and, among
    // equal values, the smallest window
    double best = WBX_INF_D;
    int best_s = nwin;
    for (int b0 = 0; b0 < nwin; b0 += 64) {
        const int s = b0 + threadIdx.x;
        const int sc = s < nwin ? s : nwin - 1;
        const double c = lane64_cost<BP, U, WT, AD, lane64_prefetch<2, BP, WT, AD>()>(
            S, L, x + sc, L, off, B, off, weights, t, pen);
        if (s < nwin && c < best) {
            best = c;
            best_s = s;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o);
        const int os = __shfl_xor(best_s, o);
        if (ob < best || (ob == best && os < best_s)) {
            best = ob;
            best_s = os;
        }
    }
    if (threadIdx.x == 0) {
        // no window below +inf keeps the reference's initial index 0
        dist[q] = sqrt(best);
        idx[q] = best_s < nwin ? best_s : 0;
    }
}

constexpr int kMaxBPW = 56;   // WDTW (per-slot weights in registers too; 40..56 in AGPRs)
// Unweighted DTW and ADTW (pairs, windows, profiles): wider rows live in
// VGPRs + AGPRs (64..80 without scratch; 88..104 spill a few slots to
// scratch), one wave per SIMD like class 56 -- still far ahead of the block
// kernel.
constexpr int kMaxBPWide = 104;

#define WBX_LANE64_CLASSES(X) X(8) X(16) X(24) X(32) X(40) X(48) X(56)
#define WBX_LANE64_WIDE(X) X(64) X(72) X(80) X(88) X(96) X(104)

}  // namespace

int lane64_cap(bool weighted) { return weighted ? kMaxBPW : kMaxBPWide; }

int lane64_class(int B, bool weighted) {
    const int BP = (B + 7) / 8 * 8;
    return BP <= (weighted ? kMaxBPW : kMaxBPWide) ? BP : 0;
}

int lane64_pairs_class(int B, bool weighted) { return lane64_class(B, weighted); }

int launch_lane64_pairs(const PairSpec &ps, Series<double> xs, Series<double> ys, double r,
                        const double *weights, double *sq, hipStream_t st, int w_fixed,
                        const double *penalty) {
    if (ps.n_pairs <= 0) return WBX_OK;
    const Band bd = make_band(xs.len, ys.len, r, w_fixed);
    const int BP = lane64_pairs_class(bd.B, weights != nullptr);
    if (!BP || xs.lens || ys.lens) return fail(WBX_EUNSUPPORTED, "lane64: unsupported geometry");
    if (penalty && weights) return fail(WBX_EUNSUPPORTED, "adtw takes no weights");
    const int kf = bd.ly - bd.lx + bd.off;
    const int wn = xs.len > ys.len ? xs.len : ys.len;
    const double pen = penalty ? *penalty : 0.0;
    const dim3 grid((unsigned)((ps.n_pairs + 63) / 64));
#define WBX_L64_PAIRS(C)                                                                        \
    case C:                                                                                     \
        if (weights)                                                                            \
            hipLaunchKernelGGL((k_lane64_pairs<C, 4, true, false>), grid, dim3(64), 0, st, ps,  \
                               xs, ys, bd.off, bd.B, kf, weights, wn, 0.0, sq);                 \
        else if (penalty)                                                                       \
            hipLaunchKernelGGL((k_lane64_pairs<C, 4, false, true>), grid, dim3(64), 0, st, ps,  \
                               xs, ys, bd.off, bd.B, kf, weights, wn, pen, sq);                 \
        else                                                                                    \
            hipLaunchKernelGGL((k_lane64_pairs<C, 4, false, false>), grid, dim3(64), 0, st, ps, \
                               xs, ys, bd.off, bd.B, kf, weights, wn, 0.0, sq);                 \
        break;
#define WBX_L64_PAIRS_WIDE(C)                                                                   \
    case C:                                                                                     \
        if (penalty)                                                                            \
            hipLaunchKernelGGL((k_lane64_pairs<C, 4, false, true>), grid, dim3(64), 0, st, ps,  \
                               xs, ys, bd.off, bd.B, kf, weights, wn, pen, sq);                 \
        else                                                                                    \
            hipLaunchKernelGGL((k_lane64_pairs<C, 4, false, false>), grid, dim3(64), 0, st, ps, \
                               xs, ys, bd.off, bd.B, kf, weights, wn, 0.0, sq);                 \
        break;
    switch (BP) {
        WBX_LANE64_CLASSES(WBX_L64_PAIRS)
        WBX_LANE64_WIDE(WBX_L64_PAIRS_WIDE)
        default: return fail(WBX_EUNSUPPORTED, "lane64: band %d", bd.B);
    }
#undef WBX_L64_PAIRS_WIDE
#undef WBX_L64_PAIRS
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

// Classes from the shortest to the longest pattern (B grows with L).
static void class_range(int min_len, int max_len, double r, bool weighted, int *lo, int *hi) {
    *lo = lane64_class(2 * band_w(min_len, min_len, r) - 1, weighted);
    *hi = lane64_class(2 * band_w(max_len, max_len, r) - 1, weighted);
    if (*lo && !*hi) *hi = weighted ? kMaxBPW : kMaxBPWide;
}

int launch_profile_lane64(double r, Shapelets sh, int paired, int min_len, int max_len,
                          Series<double> xs, int64_t n, int64_t nwp, const double *weights,
                          double *prof, hipStream_t st) {
    int lo, hi;
    class_range(min_len, max_len, r, weights != nullptr, &lo, &hi);
    if (!lo || n * nwp <= 0) return WBX_OK;
    const int nblk = (int)((nwp + 63) / 64);
    const int64_t nb = n * nblk;
    if (nb > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "profile grid too large");
    const dim3 grid((unsigned)nb);
    for (int BP = lo; BP <= hi; BP += 8) {
#define WBX_L64_PROF(C)                                                                         \
    case C:                                                                                     \
        if (weights)                                                                            \
            hipLaunchKernelGGL((k_profile_lane64<C, 4, true>), grid, dim3(64), 0, st, sh,       \
                               paired, xs, n, nwp, nblk, r, weights, prof);                     \
        else                                                                                    \
            hipLaunchKernelGGL((k_profile_lane64<C, 4, false>), grid, dim3(64), 0, st, sh,      \
                               paired, xs, n, nwp, nblk, r, weights, prof);                     \
        break;
#define WBX_L64_PROF_WIDE(C)                                                                    \
    case C:                                                                                     \
        hipLaunchKernelGGL((k_profile_lane64<C, 4, false>), grid, dim3(64), 0, st, sh, paired,  \
                           xs, n, nwp, nblk, r, weights, prof);                                 \
        break;
        switch (BP) {
            WBX_LANE64_CLASSES(WBX_L64_PROF)
            WBX_LANE64_WIDE(WBX_L64_PROF_WIDE)
            default: return fail(WBX_EUNSUPPORTED, "lane64: class %d", BP);
        }
#undef WBX_L64_PROF_WIDE
#undef WBX_L64_PROF
        WBX_HIP_CHECK(hipGetLastError());
    }
    return WBX_OK;
}

int launch_subseq_lane64(double r, Shapelets sh, int min_len, int max_len, Series<double> xs,
                         int64_t n, const double *weights, double *dist, int64_t *idx,
                         hipStream_t st, int band_add, const double *penalty) {
    int lo, hi;
    class_range(min_len + band_add, max_len + band_add, r, weights != nullptr, &lo, &hi);
    const double pen = penalty ? *penalty : 0.0;
    const int64_t nq = n * sh.n;
    if (!lo || nq <= 0) return WBX_OK;
    const dim3 grid((unsigned)(nq < 65536 ? nq : 65536), (unsigned)((nq + 65535) / 65536));
    for (int BP = lo; BP <= hi; BP += 8) {
#define WBX_L64_SUB(C)                                                                          \
    case C:                                                                                     \
        if (weights)                                                                            \
            hipLaunchKernelGGL((k_subseq_lane64<C, 4, true, false>), grid, dim3(64), 0, st, sh, \
                               xs, n, r, weights, band_add, 0.0, dist, idx);                    \
        else if (penalty)                                                                       \
            hipLaunchKernelGGL((k_subseq_lane64<C, 4, false, true>), grid, dim3(64), 0, st, sh, \
                               xs, n, r, weights, band_add, pen, dist, idx);                    \
        else                                                                                    \
            hipLaunchKernelGGL((k_subseq_lane64<C, 4, false, false>), grid, dim3(64), 0, st,    \
                               sh, xs, n, r, weights, band_add, 0.0, dist, idx);                \
        break;
#define WBX_L64_SUB_WIDE(C)                                                                     \
    case C:                                                                                     \
        if (penalty)                                                                            \
            hipLaunchKernelGGL((k_subseq_lane64<C, 4, false, true>), grid, dim3(64), 0, st, sh, \
                               xs, n, r, weights, band_add, pen, dist, idx);                    \
        else                                                                                    \
            hipLaunchKernelGGL((k_subseq_lane64<C, 4, false, false>), grid, dim3(64), 0, st,    \
                               sh, xs, n, r, weights, band_add, 0.0, dist, idx);                \
        break;
        switch (BP) {
            WBX_LANE64_CLASSES(WBX_L64_SUB)
            WBX_LANE64_WIDE(WBX_L64_SUB_WIDE)
            default: return fail(WBX_EUNSUPPORTED, "lane64: class %d", BP);
        }
#undef WBX_L64_SUB_WIDE
#undef WBX_L64_SUB
        WBX_HIP_CHECK(hipGetLastError());
    }
    return WBX_OK;
}

}  // namespace wbx
