// full.hip -- DTW when the band covers the whole matrix (w >= min(lx, ly):
// the reference's default r = 1, DtwMetric, distance/_elastic.pyx:3126-3217),
// for calls whose pairs share one geometry, in matrix coordinates.
//
// With the whole matrix in the band there is no diagonal band to follow, so
// lanes own columns: lane q of a group owns columns [q S, q S + S) (the y
// values stay in registers for the whole pair; columns >= ly read +inf and
// never feed a column < ly) and works on row s - q at step s.  A cell needs
//   up    D(i-1, j)    -- the lane's own previous row,
//   left  D(i, j-1)    -- its previous column, or for its first column lane
//                         q-1's last column of the same row, made one step
//                         earlier (DPP wave_shr:1),
//   diag  D(i-1, j-1)  -- the previous row's left neighbour: for the first
//                         column the value lane q-1 sent one step before.
// Nothing flows right to left, so one DPP move per step suffices (the band-
// coordinate kernels need two, and half their slots lie outside the matrix
// at r = 1).  D(-1, -1) = 0 starts the path; row and column -1 are +inf.
// Groups run in lockstep; steps before a lane's row 0 and after its row
// lx - 1 keep its state (a select per cell), the steps between do not.
//
// Cell arithmetic: float32 c = fma(d, d, min3(up, left, diag)); float64 the
// reference's c = min(min(up, left), diag) + d * d, contraction off
// (bit-exact: _elastic.pyx:893-940 with w = 1).
#pragma clang fp contract(off)

#include "wbx_internal.h"

namespace wbx {

namespace {

template <typename T>
__device__ __forceinline__ T finf();
template <>
__device__ __forceinline__ float finf<float>() {
    return WBX_INF_F;
}
template <>
__device__ __forceinline__ double finf<double>() {
    return WBX_INF_D;
}

__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double from_left(double v) {
    const int2 b = *reinterpret_cast<const int2 *>(&v);
    int2 o;
    o.x = __builtin_amdgcn_update_dpp(0, b.x, 0x138, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0x7ff00000, b.y, 0x138, 0xF, 0xF, false);
    return *reinterpret_cast<const double *>(&o);
}

__device__ __forceinline__ float fcell(float up, float left, float diag, float x, float y) {
    const float d = x - y;
    return __builtin_fmaf(d, d, fminf(fminf(up, left), diag));
}
__device__ __forceinline__ double fcell(double up, double left, double diag, double x, double y) {
    const double d = x - y;
    return __builtin_fmin(__builtin_fmin(up, left), diag) + d * d;
}

template <typename T, int S>
struct FLane {
    T row[S];  // D(r - 1, c0 + m) before row r, D(r, c0 + m) after
    T yv[S];
    T dg;      // D(r - 1, c0 - 1): the left value received one step earlier
    T l;       // D(r, c0 + S - 1) of the lane's last row
    int r;
    bool first;
};

template <typename T, int S, bool SEL>
__device__ __forceinline__ void fstep(FLane<T, S> &L, T x, int lx) {
    T Lin = from_left(L.l);
    Lin = L.first ? finf<T>() : Lin;
    const bool act = !SEL || (unsigned)L.r < (unsigned)lx;
    T diag = L.dg;
    T left = Lin;
#pragma unroll
    for (int m = 0; m < S; ++m) {
        const T up = L.row[m];
        const T c = fcell(up, left, diag, x, L.yv[m]);
        diag = up;
        L.row[m] = SEL ? (act ? c : up) : c;
        left = c;
    }
    if (SEL) {
        L.dg = act ? Lin : L.dg;
    } else {
        L.dg = Lin;
    }
    L.l = L.row[S - 1];
    L.r += 1;
}

template <typename T>
__device__ __forceinline__ T fldx(const T *__restrict__ X, int i, int lx) {
    return X[i < 0 ? 0 : (i >= lx ? lx - 1 : i)];
}

// Waves per SIMD the register allocation is held to: float64 lanes of 24
// columns fit three waves' 168 registers without spills (184 otherwise).
template <typename T, int S>
constexpr int full_waves() {
    return sizeof(T) == 8 && S == 24 ? 3 : 1;
}

template <typename T, int S, int U>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(full_waves<T, S>()))) void
k_full_dtw(PairSpec ps, Series<T> xs, Series<T> ys, FullGeo g, T *__restrict__ sq) {
    const int lane = threadIdx.x;
    const int gi = lane / g.n;
    const int q = lane - gi * g.n;
    const int64_t p = (int64_t)blockIdx.x * g.P + gi;
    const bool valid = gi < g.P && p < ps.n_pairs;
    int64_t a = 0, b = 0;
    if (valid) decode_pair(ps, p, a, b);
    const T *__restrict__ X = xs.row(a);
    const T *__restrict__ Y = ys.row(b);
    const int lx = g.lx, ly = g.ly;
    const int c0 = q * S;

    FLane<T, S> L;
#pragma unroll
    for (int m = 0; m < S; ++m) {
        const int j = c0 + m;
        const T v = Y[j < ly ? j : ly - 1];
        L.yv[m] = j < ly ? v : finf<T>();
        L.row[m] = finf<T>();
    }
    L.first = q == 0;
    L.dg = q == 0 ? (T)0 : finf<T>();  // D(-1, -1) = 0
    L.l = finf<T>();
    L.r = -q;
    // fill: steps [0, n - 1) (lanes before row 0 keep their state)
    const int fill = g.n - 1;
    for (int s = 0; s < fill; ++s) fstep<T, S, true>(L, fldx(X, L.r, lx), lx);
    // steady: steps [n - 1, lx), every lane inside rows [0, lx)
    const int steady = lx - fill > 0 ? lx - fill : 0;
    const int rem = steady % U;
    for (int s = 0; s < rem; ++s) fstep<T, S, false>(L, fldx(X, L.r, lx), lx);
    T xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = fldx(X, L.r + u, lx);
    for (int s = rem; s < steady; s += U) {
        T xn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xn[u] = fldx(X, L.r + U + u, lx);
#pragma unroll
        for (int u = 0; u < U; ++u) fstep<T, S, false>(L, xv[u], lx);
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = xn[u];
    }
    // drain: up to step lx + n - 2 (lanes past row lx - 1 keep it)
    const int drain = lx + g.n - 1 - fill - steady;
    for (int s = 0; s < drain; ++s) fstep<T, S, true>(L, fldx(X, L.r, lx), lx);

    const int jf = ly - 1;
    if (valid && jf >= c0 && jf < c0 + S) {
        T res = finf<T>();
#pragma unroll
        for (int m = 0; m < S; ++m)
            if (c0 + m == jf) res = L.row[m];
        sq[p] = res;
    }
}

constexpr int kFullS32[] = {4, 8, 12, 16, 24, 32};
constexpr int kFullS64[] = {4, 8, 12, 16, 24};

template <typename T, int S>
void launch_full_one(const PairSpec &ps, Series<T> xs, Series<T> ys, const FullGeo &g, T *sq,
                     hipStream_t st) {
    const dim3 grid((unsigned)((ps.n_pairs + g.P - 1) / g.P));
    hipLaunchKernelGGL((k_full_dtw<T, S, 4>), grid, dim3(64), 0, st, ps, xs, ys, g, sq);
}

}  // namespace

// Columns per lane S and lanes per pair n = ceil(ly / S) <= 64: the compiled S
// with the most matrix cells per VALU instruction of a wave step, counting
// the 64 / n pairs a wave holds, k S + 4 instructions per lane step (cells,
// the DPP move, its group-edge select, loop) and the fill and drain steps.
bool full_plan(int lx, int ly, bool fp64, FullGeo *g) {
    const int *tab = fp64 ? kFullS64 : kFullS32;
    const int nt = fp64 ? (int)(sizeof(kFullS64) / sizeof(int))
                        : (int)(sizeof(kFullS32) / sizeof(int));
    const double k = fp64 ? 5.0 : 3.0;
    double best = -1.0;
    for (int t = 0; t < nt; ++t) {
        const int S = tab[t];
        const int n = (ly + S - 1) / S;
        if (n > 64) continue;
        const int P = 64 / n;
        const double score = (double)P * ly / (k * S + 4.0) * lx / (double)(lx + n - 1);
        if (score > best) {
            best = score;
            g->S = S;
            g->n = n;
            g->P = P;
        }
    }
    g->lx = lx;
    g->ly = ly;
    return best > 0.0;
}

template <typename T>
int launch_full_dtw(const PairSpec &ps, Series<T> xs, Series<T> ys, T *sq, hipStream_t st) {
    if (ps.n_pairs <= 0) return WBX_OK;
    if (xs.lens || ys.lens) return fail(WBX_EUNSUPPORTED, "full kernel: needs one geometry");
    FullGeo g{};
    const bool dbl = std::is_same<T, double>::value;
    if (!full_plan(xs.len, ys.len, dbl, &g))
        return fail(WBX_EUNSUPPORTED, "full kernel: %d columns", ys.len);
    switch (g.S) {
        case 4: launch_full_one<T, 4>(ps, xs, ys, g, sq, st); break;
        case 8: launch_full_one<T, 8>(ps, xs, ys, g, sq, st); break;
        case 12: launch_full_one<T, 12>(ps, xs, ys, g, sq, st); break;
        case 16: launch_full_one<T, 16>(ps, xs, ys, g, sq, st); break;
        case 24: launch_full_one<T, 24>(ps, xs, ys, g, sq, st); break;
        case 32:
            if constexpr (!std::is_same<T, double>::value) {
                launch_full_one<T, 32>(ps, xs, ys, g, sq, st);
                break;
            }
            [[fallthrough]];
        default: return fail(WBX_EUNSUPPORTED, "full kernel: %d columns per lane", g.S);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_full_dtw<float>(const PairSpec &, Series<float>, Series<float>, float *,
                                    hipStream_t);
template int launch_full_dtw<double>(const PairSpec &, Series<double>, Series<double>, double *,
                                     hipStream_t);

}  // namespace wbx
