// group.hip -- DTW / WDTW for calls whose pairs share one geometry (pairwise,
// singleton, paired, argmin candidates), with a group of n lanes per pair and
// 64 / n pairs per wave.
//
// The band of every pair (distance/_elastic.pyx:869-940; slot k of row i is
// column j = i + k - off, B slots) is split over the n lanes of a group: lane
// q owns slots [k0, k0 + R + 1) when q < a, else [k0, k0 + R) (plus one
// phantom register), so that n R <= B <= n (R + 1) covers the band exactly.
// Lane q works on row s - q at step s (one row behind its left neighbour):
//   - the left input of its first slot is lane q-1's last slot of the same
//     row, made one step earlier (DPP wave_shr:1);
//   - the up input of its last slot is lane q+1's first slot of the previous
//     row, made at the start of the same step (DPP wave_shl:1), which is why
//     every lane computes its first slot before the rest of its row;
//   - lane 0 of a group reads +inf on the left, lane n-1 +inf from the right.
// All groups of a wave run in lockstep: steps [0, n) fill the pipeline and
// steps [lx, lx + n - 1) drain it (a lane outside its rows keeps its state
// through a select per cell); the steps between run without selects.
//
// Cell arithmetic:
//   float32: c = fma(d, d, min3(up, left, diag))            (3 VALU)
//            WDTW: c = fma(d * d, w, min3(...))              (4 VALU)
//   float64: the reference's order, contraction off:
//            c = min(min(up, left), diag) + (d * d) * w      (bit-exact)
// with d = x_i - y_j, w = w[|i - j|] (WeightedDtwMetric, :3322-3341) fixed per
// slot, and the row-0 prefix sum weighted w[j - 1] (w[0] at j = 0, :893-906).
// Columns outside [0, ly) read +inf (every weight must be > 0 so inf * w
// stays +inf).
#pragma clang fp contract(off)

#include "wbx_internal.h"

namespace wbx {

namespace {

template <typename T>
__device__ __forceinline__ T inf_v();
template <>
__device__ __forceinline__ float inf_v<float>() {
    return WBX_INF_F;
}
template <>
__device__ __forceinline__ double inf_v<double>() {
    return WBX_INF_D;
}

// Value of lane - 1 (wave_shr:1) / lane + 1 (wave_shl:1).  The lanes
// without a source (0 and 63) are the first / last lane of a group, whose
// value every caller replaces by +inf, so float32 keeps v there (no register
// of +inf written before each shift).
__device__ __forceinline__ float shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138,
                                                      0xF, 0xF, false));
}
__device__ __forceinline__ float shl1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130,
                                                      0xF, 0xF, false));
}
__device__ __forceinline__ double shr1(double v) {
    const int2 b = *reinterpret_cast<const int2 *>(&v);
    const int2 inf = make_int2(0, 0x7ff00000);
    int2 o;
    o.x = __builtin_amdgcn_update_dpp(inf.x, b.x, 0x138, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(inf.y, b.y, 0x138, 0xF, 0xF, false);
    return *reinterpret_cast<const double *>(&o);
}
__device__ __forceinline__ double shl1(double v) {
    const int2 b = *reinterpret_cast<const int2 *>(&v);
    const int2 inf = make_int2(0, 0x7ff00000);
    int2 o;
    o.x = __builtin_amdgcn_update_dpp(inf.x, b.x, 0x130, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(inf.y, b.y, 0x130, 0xF, 0xF, false);
    return *reinterpret_cast<const double *>(&o);
}

template <bool WT>
__device__ __forceinline__ float cell(float up, float left, float diag, float x, float y, float w) {
    const float m = fminf(fminf(up, left), diag);
    const float d = x - y;
    if constexpr (WT)
        return __builtin_fmaf(d * d, w, m);
    else
        return __builtin_fmaf(d, d, m);
}

template <bool WT>
__device__ __forceinline__ double cell(double up, double left, double diag, double x, double y,
                                       double w) {
    // the reference's min(min(x, y), z) is exact in any order here (no NaN,
    // no -0 among costs); (v * v) * w then the add, unfused
    const double m = __builtin_fmin(__builtin_fmin(up, left), diag);
    const double d = x - y;
    double t = d * d;
    if constexpr (WT) t = t * w;
    return m + t;
}

template <typename T>
__device__ __forceinline__ T ldx(const T *__restrict__ X, int i, int lx) {
    return X[i < 0 ? 0 : (i >= lx ? lx - 1 : i)];
}
template <typename T>
__device__ __forceinline__ T ldyg(const T *__restrict__ Y, int j, int ly) {
    const int jc = j < 0 ? 0 : (j >= ly ? ly - 1 : j);
    const T v = Y[jc];
    return (unsigned)j < (unsigned)ly ? v : inf_v<T>();
}

// Four consecutive elements at p (16 bytes for float32, two 16-byte loads for
// float64; only 4-byte / 8-byte alignment is assumed).
__device__ __forceinline__ void ld4g(const float *__restrict__ p, float *v) {
    typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
    const f4u q = *(const f4u *)p;
    v[0] = q[0];
    v[1] = q[1];
    v[2] = q[2];
    v[3] = q[3];
}
__device__ __forceinline__ void ld4g(const double *__restrict__ p, double *v) {
    typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
    const d2u q0 = *(const d2u *)p;
    const d2u q1 = *(const d2u *)(p + 2);
    v[0] = q0[0];
    v[1] = q0[1];
    v[2] = q1[0];
    v[3] = q1[1];
}

// A steady group's U x values (rows xr ..) and U new window entries
// (columns yc ..).  When every lane of the wave reads inside its rows and
// columns the loads are plain vector loads; otherwise each value is clamped
// into the row (x) or replaced by +inf outside [0, ly) (y).
template <typename T, int U>
__device__ __forceinline__ void group_loads(const T *__restrict__ X, const T *__restrict__ Y,
                                            int xr, int yc, int lx, int ly, T (&xo)[U],
                                            T (&yo)[U]) {
    static_assert(U % 4 == 0, "groups of four");
    const bool in = xr >= 0 && xr + U <= lx && yc >= 0 && yc + U <= ly;
    if (__ballot(!in) == 0) {  // uniform
#pragma unroll
        for (int u = 0; u < U; u += 4) {
            ld4g(X + xr + u, xo + u);
            ld4g(Y + yc + u, yo + u);
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xo[u] = ldx(X, xr + u, lx);
            yo[u] = ldyg(Y, yc + u, ly);
        }
    }
}

// Per-lane state of a group member.
template <typename T, int R, int U, bool WT>
struct GLane {
    static constexpr int S = R + 1;
    static constexpr int W = S + U - 1;
    T row[S];
    T yw[W];   // yw[q] = y[yj + q]
    T wt[WT ? S : 1];
    T l;       // last real slot of the lane's last computed row
    int r;     // row of the lane at the current step
    int yj;    // column of yw[0]
    bool first, last, shortl;
};

// One step (one row of the lane): SEL keeps the state of lanes outside
// [0, lx); W0 uses the row-0 weights w0 for a lane at row 0.
template <typename T, int R, int U, bool WT, bool SEL, bool W0>
__device__ __forceinline__ void gstep(GLane<T, R, U, WT> &L, int u, T x, int lx, const T *w0) {
    constexpr int S = R + 1;
    const bool act = !SEL || (unsigned)(L.r + u) < (unsigned)lx;
    const bool row0 = W0 && L.r + u == 0;
    auto wgt = [&](int m) -> T {
        if constexpr (!WT)
            return (T)1;
        else if constexpr (W0)
            return row0 ? w0[m] : L.wt[m];
        else
            return L.wt[m];
    };
    T Lin = shr1(L.l);
    Lin = L.first ? inf_v<T>() : Lin;
    const T c0 = cell<WT>(L.row[1], Lin, L.row[0], x, L.yw[u], wgt(0));
    T Uin = shl1(SEL && !act ? L.row[0] : c0);
    Uin = L.last ? inf_v<T>() : Uin;
    T left = c0;
    const T r0 = SEL ? (act ? c0 : L.row[0]) : c0;
#pragma unroll
    for (int m = 1; m < S; ++m) {
        T up;
        if (m == S - 1)
            up = Uin;
        else if (m == S - 2)
            up = L.shortl ? Uin : L.row[m + 1];
        else
            up = L.row[m + 1];
        const T c = cell<WT>(up, left, L.row[m], x, L.yw[u + m], wgt(m));
        L.row[m] = SEL ? (act ? c : L.row[m]) : c;
        left = c;
    }
    L.row[0] = r0;
    L.l = L.shortl ? L.row[R - 1] : L.row[R];
}

// Single steps s in [s0, s1) (fill / drain / remainder), window shifted by
// one each step.
template <typename T, int R, int U, bool WT, bool SEL, bool W0>
__device__ __forceinline__ void gsingle(GLane<T, R, U, WT> &L, int steps, const T *__restrict__ X,
                                        const T *__restrict__ Y, int lx, int ly, const T *w0) {
    constexpr int W = GLane<T, R, U, WT>::W;
    for (int s = 0; s < steps; ++s) {
        const T x = ldx(X, L.r, lx);
        const T ynew = ldyg(Y, L.yj + W, ly);
        gstep<T, R, U, WT, SEL, W0>(L, 0, x, lx, w0);
#pragma unroll
        for (int q = 0; q < W - 1; ++q) L.yw[q] = L.yw[q + 1];
        L.yw[W - 1] = ynew;
        L.r += 1;
        L.yj += 1;
    }
}

template <typename T, int R, int U, bool WT>
__global__ __launch_bounds__(64) void k_group_dtw(PairSpec ps, Series<T> xs, Series<T> ys,
                                                  GroupGeo g, const T *__restrict__ wv,
                                                  T *__restrict__ sq) {
    using Ln = GLane<T, R, U, WT>;
    constexpr int S = Ln::S, W = Ln::W;
    const int lane = threadIdx.x;
    const int gi = lane / g.n;
    const int q = lane - gi * g.n;
    const int64_t p = (int64_t)blockIdx.x * g.P + gi;
    const bool valid = gi < g.P && p < ps.n_pairs;
    int64_t a = 0, b = 0;
    if (valid) decode_pair(ps, p, a, b);
    const T *__restrict__ X = xs.row(a);
    const T *__restrict__ Y = ys.row(b);
    const int lx = g.lx, ly = g.ly, off = g.off;

    Ln L;
    L.first = q == 0;
    L.last = q == g.n - 1;
    L.shortl = q >= g.a;
    const int k0 = q < g.a ? q * S : g.a * S + (q - g.a) * R;
#pragma unroll
    for (int m = 0; m < S; ++m) L.row[m] = (k0 + m == off) ? (T)0 : inf_v<T>();
    L.l = L.shortl ? L.row[R - 1] : L.row[R];
    L.r = -q;
    L.yj = L.r + k0 - off;
#pragma unroll
    for (int m = 0; m < W; ++m) L.yw[m] = ldyg(Y, L.yj + m, ly);
    T w0[WT ? S : 1];
    if constexpr (WT) {
        const int top = g.nw - 1;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const int k = k0 + m - off;  // i - j of the slot (any row); row-0 column
            const int dw = k < 0 ? -k : k;
            L.wt[m] = wv[dw < top ? dw : top];
            const int c0 = k >= 1 ? k - 1 : 0;
            w0[m] = wv[c0 < top ? c0 : top];
        }
    }
    // fill: steps [0, n); lanes before their row 0 keep their state
    const int fill = g.n;
    gsingle<T, R, U, WT, true, WT>(L, fill, X, Y, lx, ly, w0);
    // steady: steps [n, lx), every lane inside rows [1, lx)
    int steady = lx - fill;
    if (steady > 0) {
        const int rem = steady % U;
        gsingle<T, R, U, WT, false, false>(L, rem, X, Y, lx, ly, w0);
        T xv[U], yn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u] = ldx(X, L.r + u, lx);
            yn[u] = ldyg(Y, L.yj + W + u, ly);
        }
        // groups in pairs: the second group's x and y entries go to registers
        // the first does not read, and the next pair's to xv / yn, so no
        // group copies its successor's loads
        auto shift_in = [&](const T(&ynew)[U]) {
#pragma unroll
            for (int m = 0; m < W - U; ++m) L.yw[m] = L.yw[m + U];
#pragma unroll
            for (int u = 0; u < U; ++u) L.yw[W - U + u] = ynew[u];
            L.r += U;
            L.yj += U;
        };
        int s = rem;
        for (; s + 2 * U <= steady; s += 2 * U) {
            T xb[U], yb[U];
            group_loads<T, U>(X, Y, L.r + U, L.yj + W + U, lx, ly, xb, yb);
#pragma unroll
            for (int u = 0; u < U; ++u) gstep<T, R, U, WT, false, false>(L, u, xv[u], lx, w0);
            shift_in(yn);
            group_loads<T, U>(X, Y, L.r + U, L.yj + W + U, lx, ly, xv, yn);
#pragma unroll
            for (int u = 0; u < U; ++u) gstep<T, R, U, WT, false, false>(L, u, xb[u], lx, w0);
            shift_in(yb);
        }
        if (s < steady) {  // one group left
#pragma unroll
            for (int u = 0; u < U; ++u) gstep<T, R, U, WT, false, false>(L, u, xv[u], lx, w0);
            shift_in(yn);
        }
    } else {
        steady = 0;
    }
    // drain: steps [max(n, lx), lx + n - 1); lanes past row lx - 1 keep it
    const int done = fill + steady;
    const int drain = lx + g.n - 1 - done;
    gsingle<T, R, U, WT, true, false>(L, drain, X, Y, lx, ly, w0);

    const int nslots = L.shortl ? R : S;
    if (valid && g.kf >= k0 && g.kf < k0 + nslots) {
        T res = inf_v<T>();
#pragma unroll
        for (int m = 0; m < S; ++m)
            if (k0 + m == g.kf) res = L.row[m];
        sq[p] = res;
    }
}

// Compiled strip sizes (slots per lane R or R + 1).
constexpr int kGroupR32[] = {2, 3, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40};
constexpr int kGroupR64[] = {2, 3, 4, 6, 8, 10, 12, 14, 16, 20, 24};

template <typename T, int R, bool WT>
void launch_group_one(const PairSpec &ps, Series<T> xs, Series<T> ys, const GroupGeo &g,
                      const T *w, T *sq, hipStream_t st) {
    constexpr int U = 4;
    const dim3 grid((unsigned)((ps.n_pairs + g.P - 1) / g.P));
    hipLaunchKernelGGL((k_group_dtw<T, R, U, WT>), grid, dim3(64), 0, st, ps, xs, ys, g, w, sq);
}

template <typename T, bool WT>
int launch_group_r(int R, const PairSpec &ps, Series<T> xs, Series<T> ys, const GroupGeo &g,
                   const T *w, T *sq, hipStream_t st) {
    switch (R) {
#define WBX_G(RR)                                                      \
    case RR:                                                           \
        launch_group_one<T, RR, WT>(ps, xs, ys, g, w, sq, st);         \
        break;
        WBX_G(2) WBX_G(3) WBX_G(4) WBX_G(6) WBX_G(8) WBX_G(10) WBX_G(12) WBX_G(14) WBX_G(16)
        WBX_G(20) WBX_G(24)
#undef WBX_G
        default:
            if constexpr (std::is_same<T, float>::value) {
                switch (R) {
                    case 28: launch_group_one<T, 28, WT>(ps, xs, ys, g, w, sq, st); break;
                    case 32: launch_group_one<T, 32, WT>(ps, xs, ys, g, w, sq, st); break;
                    case 40: launch_group_one<T, 40, WT>(ps, xs, ys, g, w, sq, st); break;
                    default: return fail(WBX_EUNSUPPORTED, "group kernel: strip size %d", R);
                }
            } else {
                return fail(WBX_EUNSUPPORTED, "group kernel: strip size %d", R);
            }
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace

// Chooses the strip size R and lanes per pair n for band width B: among the
// compiled R with n R <= B <= n (R + 1), n <= 64, the one with the most
// band cells per VALU instruction of a wave step, counting the 64 / n pairs a
// wave holds, the per-row cost k (R + 1) + 8 (cells, two DPP moves and their
// group-edge selects, the short-lane select) and the fill and drain steps.
bool group_plan(int B, int lx, bool fp64, bool weighted, GroupGeo *g) {
    const int *tab = fp64 ? kGroupR64 : kGroupR32;
    const int nt = fp64 ? (int)(sizeof(kGroupR64) / sizeof(int))
                        : (int)(sizeof(kGroupR32) / sizeof(int));
    const double k = fp64 ? (weighted ? 6.0 : 5.0) : (weighted ? 4.0 : 3.0);
    // float64 weighted lanes hold two weight vectors: keep them to 17 slots
    const int rmax = fp64 && weighted ? 16 : 1 << 30;
    double best = -1.0;
    for (int t = 0; t < nt; ++t) {
        const int R = tab[t];
        if (R > rmax) continue;
        const int n = (B + R) / (R + 1);  // ceil(B / (R + 1))
        if (n < 1 || n > 64 || (int64_t)n * R > B) continue;
        const int P = 64 / n;
        const double score = (double)P * B / (k * (R + 1) + 8.0) * lx / (double)(lx + n - 1);
        if (score > best) {
            best = score;
            g->n = n;
            g->a = B - n * R;
            g->P = P;
            g->R = R;
        }
    }
    return best > 0.0;
}

template <typename T>
int launch_group_dtw(const PairSpec &ps, Series<T> xs, Series<T> ys, double r, const T *weights,
                     int n_weights, T *sq, hipStream_t st, int w_fixed) {
    if (ps.n_pairs <= 0) return WBX_OK;
    if (xs.lens || ys.lens) return fail(WBX_EUNSUPPORTED, "group kernel: needs one geometry");
    const Band bd = make_band(xs.len, ys.len, r, w_fixed);
    GroupGeo g{};
    g.lx = bd.lx;
    g.ly = bd.ly;
    g.off = bd.off;
    g.B = bd.B;
    g.kf = bd.ly - bd.lx + bd.off;
    g.nw = n_weights;
    const bool dbl = std::is_same<T, double>::value;
    if (!group_plan(bd.B, bd.lx, dbl, weights != nullptr, &g))
        return fail(WBX_EUNSUPPORTED, "group kernel: band %d", bd.B);
    return weights ? launch_group_r<T, true>(g.R, ps, xs, ys, g, weights, sq, st)
                   : launch_group_r<T, false>(g.R, ps, xs, ys, g, nullptr, sq, st);
}

template int launch_group_dtw<float>(const PairSpec &, Series<float>, Series<float>, double,
                                     const float *, int, float *, hipStream_t, int);
template int launch_group_dtw<double>(const PairSpec &, Series<double>, Series<double>, double,
                                      const double *, int, double *, hipStream_t, int);

}  // namespace wbx
