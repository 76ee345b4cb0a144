// strip.hip -- the persistent kernel of the batched uneven pairs: one launch,
// one queue of pairs sorted by cost (longest first), any band width up to
// 64 (R + 1) slots with R <= 32 (2112 slots; wider bands take the plain
// kernel).
//
// A wave takes the next pair from the queue and runs it alone.  The pair's
// band (distance/_elastic.pyx:869-940, w = _compute_r(min(lx, ly), r_p);
// slot k of row i is column j = i + k - off, B slots) is split over n <= 64
// lanes: lane p owns slots [k0, k0 + R + 1) when p < a, else [k0, k0 + R)
// with one phantom register, so n R <= B <= n (R + 1).  Lane p works on row
// s - p at step s:
//   - the left input of its first slot is lane p-1's last slot of the same
//     row, made one step earlier (DPP wave_shr:1);
//   - the up input of its last slot is lane p+1's first slot of the previous
//     row, made at the start of the same step (DPP wave_shl:1), so every lane
//     computes its first slot before the rest of its row.
// Groups of U steps in which every lane is inside its rows run without
// selects; the fill and drain groups (about n + U steps per pair) keep the
// state of a lane outside its rows through a select per cell.  Lanes past n
// read no columns (+inf cells), so lanes 0 and n - 1 see +inf beyond the
// band without masking.
//
// Reads are unconditional 16-byte loads; the plan gives the kernel only pairs
// whose reads stay inside the arrays (the others are staged into padded rows
// first), and columns outside [0, ly) are replaced by +inf after the load.
//
// Cell arithmetic: float32 c = fma(d, d, min3(up, left, diag)); float64 the
// reference's c = min(min(up, left), diag) + d * d, contraction off (the
// FP64 and FP32_ACC64 modes are bit-identical to the reference).
#pragma clang fp contract(off)

#include "strip.h"

namespace wbx {

namespace {

template <typename T>
__device__ __forceinline__ T sinf_();
template <>
__device__ __forceinline__ float sinf_<float>() {
    return WBX_INF_F;
}
template <>
__device__ __forceinline__ double sinf_<double>() {
    return WBX_INF_D;
}

__device__ __forceinline__ float shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float shl1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double shr1(double v) {
    const int2 b = *reinterpret_cast<const int2 *>(&v);
    int2 o;
    o.x = __builtin_amdgcn_update_dpp(0, b.x, 0x138, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0x7ff00000, b.y, 0x138, 0xF, 0xF, false);
    return *reinterpret_cast<const double *>(&o);
}
__device__ __forceinline__ double shl1(double v) {
    const int2 b = *reinterpret_cast<const int2 *>(&v);
    int2 o;
    o.x = __builtin_amdgcn_update_dpp(0, b.x, 0x130, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0x7ff00000, b.y, 0x130, 0xF, 0xF, false);
    return *reinterpret_cast<const double *>(&o);
}

// c0 of a row: the squared difference added to the minimum
__device__ __forceinline__ float fma_or_add(float d, float m) { return __builtin_fmaf(d, d, m); }
__device__ __forceinline__ double fma_or_add(double d, double m) { return m + d * d; }

__device__ __forceinline__ float scell(float up, float left, float diag, float x, float y) {
    const float d = x - y;
    return __builtin_fmaf(d, d, fminf(fminf(up, left), diag));
}
__device__ __forceinline__ double scell(double up, double left, double diag, double x, double y) {
    const double d = x - y;
    return __builtin_fmin(__builtin_fmin(up, left), diag) + d * d;
}

// Four consecutive elements at a linear offset (the plan keeps every offset
// inside the arrays).
template <typename T>
struct V4 {
    T v[4];
};
__device__ __forceinline__ V4<float> ld4(const float *__restrict__ b, int o) {
    typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
    const f4u q = *(const f4u *)(b + o);
    return V4<float>{{q[0], q[1], q[2], q[3]}};
}
__device__ __forceinline__ V4<double> ld4(const double *__restrict__ b, int o) {
    typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
    const d2u q0 = *(const d2u *)(b + o);
    const d2u q1 = *(const d2u *)(b + o + 2);
    return V4<double>{{q0[0], q0[1], q1[0], q1[1]}};
}

// min(a, b, left neighbour's v): min(a, b) then, in every lane with a left
// neighbour, min with that lane's v (v_min_f32 with DPP wave_shr:1: lane 0
// has no source and keeps min(a, b), i.e. its left input is +inf).  The
// s_nop gives the two wait states a DPP read needs after v was written.
// Both minima are written in the asm: as fminf the compiler would quiet each
// operand first (two v_max per row); no operand is a NaN here.
__device__ __forceinline__ float min_left(float a, float b, float v) {
    float t;
    asm volatile(
        "v_min_f32 %0, %1, %2\n\ts_nop 0\n\t"
        "v_min_f32_dpp %0, %3, %0 wave_shr:1 row_mask:0xf bank_mask:0xf"
        : "=&v"(t)
        : "v"(a), "v"(b), "v"(v));
    return t;
}
__device__ __forceinline__ double min_left(double a, double b, double v) {
    return __builtin_fmin(__builtin_fmin(a, shr1(v)), b);
}

// The up input of a lane's last slot: lane p + 1's value (DPP wave_shl:1);
// lane 63 has no source and keeps the old value, +inf from the start, so the
// destination needs no +inf written before every shift.
__device__ __forceinline__ float up_in(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130,
                                                      0xF, 0xF, false));
}
__device__ __forceinline__ double up_in(double old, double v) {
    (void)old;
    return shl1(v);
}

template <typename T, int R, int U>
struct Strip {
    static constexpr int S = R + 1;
    static constexpr int W = S + U - 1;
    T row[S];
    T yw[W + U];  // yw[q] = y[yj + q]; the last U are the next group's in a group pair
    T xv[U];  // x of the group's U rows
    T l;      // last real slot of the lane's last row
    T uin;    // last up input (+inf in lane 63)
    int r;    // the lane's row at the start of the group
    int yj;
    int lyl;  // ly for the pair's lanes, 0 for lanes past the band (+inf columns)
    bool shortl;
};

// One group of U rows.  STEADY: every lane of the pair is inside its rows for
// all U rows (lanes past the band compute +inf), so no cell needs a select.
// The window starts at yw[YO] (YO = U for the second group of a pair) and the
// group's x values are xs.
template <typename T, int R, int U, bool STEADY, int YO = 0>
__device__ __forceinline__ void strip_rows(Strip<T, R, U> &L, int lx, const T (&xs)[U]) {
    constexpr int S = R + 1;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const bool act = STEADY || (unsigned)(L.r + YO + u) < (unsigned)lx;
        const T x = xs[u];
        // the select variant writes the difference as y - x ((y - x)^2 ==
        // (x - y)^2 exactly): otherwise the compiler computes both variants'
        // differences once above their branch and keeps them all live
        const T d = STEADY ? x - L.yw[YO + u] : L.yw[YO + u] - x;
        const T c0 = fma_or_add(d, min_left(L.row[1], L.row[0], L.l));
        const T Uin = up_in(L.uin, act ? c0 : L.row[0]);
        L.uin = Uin;
        T left = c0;
        const T r0 = act ? c0 : L.row[0];
#pragma unroll
        for (int m = 1; m < S; ++m) {
            T up;
            if (m == S - 1)
                up = Uin;
            else if (m == S - 2)
                up = L.shortl ? Uin : L.row[m + 1];
            else
                up = L.row[m + 1];
            const T c = STEADY ? scell(up, left, L.row[m], x, L.yw[YO + u + m])
                               : scell(up, left, L.row[m], L.yw[YO + u + m], x);
            L.row[m] = act ? c : L.row[m];  // the old slot m was needed only by cells m, m + 1
            left = c;
        }
        L.row[0] = r0;
        L.l = L.shortl ? L.row[R - 1] : L.row[R];
        // one row at a time: the differences of later rows are not hoisted
        // (they would all be live at once)
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One group: the next group's x and the window's U new entries are issued
// before the group's U rows; then the window moves on by U.
template <typename T, int R, int U, bool STEADY>
__device__ __forceinline__ void strip_group(Strip<T, R, U> &L, const T *__restrict__ Xa,
                                            const T *__restrict__ Ya, int xo, int yo, int lx,
                                            int ly, int s0, int c0l, int cnl) {
    constexpr int W = Strip<T, R, U>::W;
    const T INF = sinf_<T>();
    T xn[U], yn[U];
    const int jn = L.yj + W;
#pragma unroll
    for (int u = 0; u < U; u += 4) {
        const V4<T> a = ld4(Xa, xo + L.r + U + u);
        const V4<T> b = ld4(Ya, yo + jn + u);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xn[u + q] = a.v[q];
            yn[u + q] = b.v[q];
        }
    }
    strip_rows<T, R, U, STEADY>(L, lx, L.xv);
#pragma unroll
    for (int q = 0; q < W - U; ++q) L.yw[q] = L.yw[q + U];
    if (s0 + c0l >= 0 && s0 + cnl < ly) {
#pragma unroll
        for (int u = 0; u < U; ++u) L.yw[W - U + u] = yn[u];
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u)
            L.yw[W - U + u] = (unsigned)(jn + u) < (unsigned)L.lyl ? yn[u] : INF;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) L.xv[u] = xn[u];
    L.r += U;
    L.yj += U;
}

// Two steady groups, with the second group's x and new window entries
// loaded straight into registers the first group does not read (xb and
// yw[W, W + U)): no copies of them, and one window shift per two groups.
template <typename T, int R, int U>
__device__ __forceinline__ void strip_group_pair(Strip<T, R, U> &L, const T *__restrict__ Xa,
                                                 const T *__restrict__ Ya, int xo, int yo,
                                                 int lx, int ly, int s0, int c0l, int cnl) {
    constexpr int W = Strip<T, R, U>::W;
    const T INF = sinf_<T>();
    T xb[U], yn[U];
    int jn = L.yj + W;
#pragma unroll
    for (int u = 0; u < U; u += 4) {
        const V4<T> a = ld4(Xa, xo + L.r + U + u);
        const V4<T> b = ld4(Ya, yo + jn + u);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xb[u + q] = a.v[q];
            L.yw[W + u + q] = b.v[q];
        }
    }
    strip_rows<T, R, U, true, 0>(L, lx, L.xv);
    if (!(s0 + c0l >= 0 && s0 + cnl < ly)) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            L.yw[W + u] = (unsigned)(jn + u) < (unsigned)L.lyl ? L.yw[W + u] : INF;
    }
    jn += U;
#pragma unroll
    for (int u = 0; u < U; u += 4) {
        const V4<T> a = ld4(Xa, xo + L.r + 2 * U + u);
        const V4<T> b = ld4(Ya, yo + jn + u);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            L.xv[u + q] = a.v[q];
            yn[u + q] = b.v[q];
        }
    }
    strip_rows<T, R, U, true, U>(L, lx, xb);
#pragma unroll
    for (int q = 0; q < W - U; ++q) L.yw[q] = L.yw[q + 2 * U];
    if (s0 + U + c0l >= 0 && s0 + U + cnl < ly) {
#pragma unroll
        for (int u = 0; u < U; ++u) L.yw[W - U + u] = yn[u];
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u)
            L.yw[W - U + u] = (unsigned)(jn + u) < (unsigned)L.lyl ? yn[u] : INF;
    }
    L.r += 2 * U;
    L.yj += 2 * U;
}

// One pair on the wave.  R: slots of a short lane (long lanes R + 1).
template <typename T, int R>
__device__ __forceinline__ void strip_pair(const StripItem &it, const StripArrays<T> &ar,
                                           T *__restrict__ sq, int p) {
    constexpr int U = R <= 16 ? 8 : 4;
    using St = Strip<T, R, U>;
    constexpr int S = St::S, W = St::W;
    const T INF = sinf_<T>();
    const bool live = p < it.n;
    St L;
    L.shortl = p >= it.a;
    const int k0 = p < it.a ? p * S : it.a * S + (p - it.a) * R;
    const T *__restrict__ Xa;
    const T *__restrict__ Ya;
    int xo, yo;
    if (!live) {
        // lanes past the band: no columns (+inf cells); x from the zero
        // buffer and y from the +inf buffer, which cover every offset a pair
        // reads (d = 0 - inf, d * d = +inf: no select keeps them at +inf)
        Xa = ar.zb;
        Ya = ar.ib;
        xo = yo = ar.zpad;
    } else if (it.stage < 0) {
        Xa = it.swap ? ar.yb : ar.xb;
        Ya = it.swap ? ar.xb : ar.yb;
        xo = it.pair * (it.swap ? ar.ystride : ar.xstride);
        yo = it.pair * (it.swap ? ar.xstride : ar.ystride);
    } else {
        const int xr = ar.spad + 2 * it.stage * ar.sstride, yr = xr + ar.sstride;
        Xa = Ya = ar.sb;
        xo = it.swap ? yr : xr;
        yo = it.swap ? xr : yr;
    }
    const int lx = it.lx, off = it.off;
    L.lyl = live ? it.ly : 0;
    const int kf = it.ly - lx + off;  // slot of the final cell (lx-1, ly-1)
#pragma unroll
    for (int m = 0; m < S; ++m) L.row[m] = (live && k0 + m == off) ? (T)0 : INF;
    L.r = -p;
    L.yj = L.r + k0 - off;
#pragma unroll
    for (int q = 0; q < W; q += 4) {
        const V4<T> v = ld4(Ya, yo + L.yj + q);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q + u < W) L.yw[q + u] = (unsigned)(L.yj + q + u) < (unsigned)L.lyl ? v.v[u] : INF;
    }
#pragma unroll
    for (int u = 0; u < U; u += 4) {
        const V4<T> v = ld4(Xa, xo + L.r + u);
#pragma unroll
        for (int q = 0; q < 4; ++q) L.xv[u + q] = v.v[q];
    }
    L.l = L.shortl ? L.row[R - 1] : L.row[R];
    L.uin = INF;
    // The final cell (lx - 1, ly - 1) is slot kf of lane pk, which finishes
    // row lx - 1 at step lx - 1 + pk: the pair runs steps [0, lx + pk), in G
    // groups.  Lanes p < pk then run past row lx - 1 in the last steps; what
    // they compute there reaches only lanes that are past their last row too
    // (lane p + 1 takes lane p's row of one step earlier, lane p - 1 the row
    // lane p works on), so it never reaches lane pk's rows.
    const int pk = kf < it.a * S ? kf / S : it.a + (kf - it.a * S) / R;
    const int G = (lx + pk + U - 1) / U;
    // new window entries of group g: columns [jn, jn + U) with jn = g U + c_p,
    // c_p = W - p + k0 - off increasing in p; between lane 0's first and lane
    // n - 1's last inside [0, ly) no entry needs the +inf select
    const int kl = it.n - 1 < it.a ? (it.n - 1) * S : it.a * S + (it.n - 1 - it.a) * R;
    const int c0l = W - off, cnl = W - (it.n - 1) + kl - off + U - 1;
    // Fill groups [0, gs) keep a lane's state until it reaches row 0, with a
    // select per cell.  From gs on every lane has started: groups [gs, ge)
    // are inside every lane's rows, and [ge, G - 1) run the same select-free
    // rows (a lane past its last row computes values nothing reads); the last
    // group keeps each lane's state after its last row with selects, which
    // leaves lane pk at row lx - 1.  Each range is a loop of its own, so the
    // compiler keeps the band row in the same registers across a loop's
    // iterations instead of merging two variants at every group.
    const int gs = std::min((it.n - 1 + U - 1) / U, G);
    const int ge = std::max(gs, std::min(lx / U, G));
    int g = 0;
    for (; g < gs; ++g) strip_group<T, R, U, false>(L, Xa, Ya, xo, yo, lx, it.ly, g * U, c0l, cnl);
    // group pairs for the strips whose extra U registers keep the kernel
    // within its register budget (R >= 24 is at the limit)
    if constexpr (R <= 22)
        for (; g + 1 < ge; g += 2)
            strip_group_pair<T, R, U>(L, Xa, Ya, xo, yo, lx, it.ly, g * U, c0l, cnl);
    for (; g < std::max(ge, G - 1); ++g)
        strip_group<T, R, U, true>(L, Xa, Ya, xo, yo, lx, it.ly, g * U, c0l, cnl);
    if (g < G) strip_group<T, R, U, false>(L, Xa, Ya, xo, yo, lx, it.ly, g * U, c0l, cnl);
    const int nslots = L.shortl ? R : S;
    if (live && kf >= k0 && kf < k0 + nslots) {
        T res = INF;
#pragma unroll
        for (int m = 0; m < S; ++m)
            if (k0 + m == kf) res = L.row[m];
        sq[it.pair] = res;
    }
}

// Strip sizes compiled.  All of them share one launch (and its register
// budget, set by the widest): three launches by register tier were measured
// slower on the uneven-pairs workload (3.07 ms against 2.91 ms), the
// launches' tails costing more than the narrow strips gained in occupancy.
#define WBX_STRIP_R32_T0(X)                                                                   \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(18) \
        X(20) X(22) X(24) X(26) X(28) X(30) X(32)
#define WBX_STRIP_R64_T0(X) \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(24) X(28) X(32)

// Waves per SIMD the register allocation is held to: the float64 strips fit
// two waves' 256 registers without spills (the allocator takes 285, one wave,
// otherwise); the float32 ones sit at three waves by themselves.
template <typename T>
constexpr int strip_waves() {
    return sizeof(T) == 8 ? 2 : 1;
}

template <typename T, int TIER>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(strip_waves<T>()))) void
k_strip_batch(const StripItem *__restrict__ items, int n_items, int *__restrict__ counter,
              StripArrays<T> ar, T *__restrict__ sq) {
    const int p = threadIdx.x;
    while (true) {
        int k = 0;
        if (p == 0) k = atomicAdd(counter, 1);
        k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
        if (k >= n_items) break;  // every wave leaves once the queue is drained
        const StripItem it = items[k];
#define WBX_CASE(RR)                      \
    case RR:                              \
        strip_pair<T, RR>(it, ar, sq, p); \
        break;
        if constexpr (std::is_same<T, float>::value) {
            switch (it.R) { WBX_STRIP_R32_T0(WBX_CASE) default: break; }
        } else {
            switch (it.R) { WBX_STRIP_R64_T0(WBX_CASE) default: break; }
        }
#undef WBX_CASE
    }
}

#define WBX_LIST(RR) RR,
constexpr int kR32[] = {WBX_STRIP_R32_T0(WBX_LIST)};
constexpr int kR64[] = {WBX_STRIP_R64_T0(WBX_LIST)};
#undef WBX_LIST

}  // namespace

int strip_tier(int R, bool fp64) {
    (void)R;
    (void)fp64;
    return 0;
}

// The smallest compiled R whose 64 lanes cover the band, with a split
// n R <= B <= n (R + 1), n <= 64: every lane of the wave issues each row's
// instructions whatever n is, so the fewest slots per lane is the fastest.
bool strip_split_t(int B, bool fp64, int *R_out, int *n_out, int *a_out) {
    const int *tab = fp64 ? kR64 : kR32;
    const int nt = fp64 ? (int)(sizeof(kR64) / sizeof(int)) : (int)(sizeof(kR32) / sizeof(int));
    for (int q = 0; q < nt; ++q) {
        const int R = tab[q];
        if (64 * (R + 1) < B) continue;
        const int n = (B + R) / (R + 1);  // ceil(B / (R + 1))
        if (n * R > B || n > 64) continue;
        *R_out = R;
        *n_out = n;
        *a_out = B - n * R;
        return true;
    }
    return false;
}

template <typename T>
int launch_strip_batch(int tier, const StripItem *items, int n_items, int *counter,
                       StripArrays<T> ar, T *sq, int waves, hipStream_t st) {
    if (n_items <= 0) return WBX_OK;
    const dim3 g((unsigned)waves), b(64);
    if (tier != 0) return fail(WBX_EINVAL, "strip tier %d", tier);
    hipLaunchKernelGGL((k_strip_batch<T, 0>), g, b, 0, st, items, n_items, counter, ar, sq);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_strip_batch<float>(int, const StripItem *, int, int *, StripArrays<float>,
                                       float *, int, hipStream_t);
template int launch_strip_batch<double>(int, const StripItem *, int, int *, StripArrays<double>,
                                        double *, int, hipStream_t);

}  // namespace wbx
