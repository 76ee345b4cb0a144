// strip.h -- the persistent strip kernel of the batched uneven pairs
// (strip.hip) and the plan-side facts it relies on.
#pragma once

#include <algorithm>

#include "wbx_internal.h"

namespace wbx {

// One pair of the queue.
struct StripItem {
    int32_t pair;    // output slot
    int32_t swap;    // 1: rows come from Y[pair], columns from X[pair] (DTW is symmetric)
    int32_t lx, ly;  // lengths along rows / columns
    int32_t off, B;  // band offset and width (slots)
    int32_t n, a;    // lanes used; the first a lanes own R + 1 slots, the rest R
    int32_t R;       // strip size
    int32_t stage;   // -1: read the caller's arrays; else the pair's padded staged rows
    int32_t pad_[2];
};

// Arrays of one launch (element offsets are 32-bit): X and Y with their row
// strides, the staged rows (x row of stage e at spad + 2 e sstride, y row
// sstride later), and the buffers read by the lanes past a pair's band: zeros
// for x and +inf for y (their cells are +inf without a select).
template <typename T>
struct StripArrays {
    const T *xb, *yb, *sb, *zb, *ib;
    int xstride, ystride, sstride, spad;
    int zpad;  // offset of the zero buffer's middle (it covers every read of a pair)
};

// Strip size R, lanes n and long lanes a for band width B (false: wider than
// the compiled sizes cover).
bool strip_split_t(int B, bool fp64, int *R_out, int *n_out, int *a_out);

// Reads of the kernel relative to a pair's row start.  Lane p (<= 63) reads x
// rows [-p, lx + n + 6] and y columns from -p - off up to
// lx + n + 6 + 63 (R + 1) + R + 7 - off (window of R + 4 plus the next four).
inline int64_t strip_read_lo(int off) { return -64 - (int64_t)off; }
inline int64_t strip_read_hi(int lx, int off, int R) {
    return std::max<int64_t>(lx + 80, lx + 80 + 64 * (int64_t)(R + 1) + R - off);
}
// Whether those reads stay inside n x t arrays of any row stride >= t for
// pair q (pairs at the two ends of the arrays can fail this: they are staged).
inline bool strip_reads_inside(int64_t q, int64_t n, int64_t t, int lx, int off, int R) {
    return q * t + strip_read_lo(off) >= 0 && strip_read_hi(lx, off, R) <= (n - 1 - q) * t + t - 1;
}

// Register tier of a strip size; each tier is one persistent launch over its
// part of the queue (one tier: every strip size shares the launch).
constexpr int kStripTiers = 1;
int strip_tier(int R, bool fp64);

template <typename T>
int launch_strip_batch(int tier, const StripItem *items, int n_items, int *counter,
                       StripArrays<T> ar, T *sq, int waves, hipStream_t st);

}  // namespace wbx
