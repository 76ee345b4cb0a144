// train_seg.hip -- differentiable DTW and WDTW for any band width, with the
// cost rows checkpointed along anti-diagonal bands in the forward and each
// band recomputed, with its backtrack decisions, in the backward.
//
// Layout (the group layout of group.hip): a pair's band (distance/_elastic.pyx
// :869-940; slot k of row i is column j = i + k - off, B slots) is split over
// n lanes, lane q owning slots [k0, k0 + R + 1) for q < a, else [k0, k0 + R)
// plus one phantom register; 64 / n pairs share a wave in lockstep.  Lane q
// works on row s - q at step s: its first slot's left input is lane q-1's last
// slot of the same row (made one step earlier), its last slot's up input lane
// q+1's first slot of the previous row (made in the same step), so the state
// of the whole pair at the start of step s -- every lane's band row -- fixes
// every later step.  The steps run from 0 to lx + n - 2.
//
// Forward.  At every step s that is a multiple of K the state is stored: n
// band rows per pair, one anti-diagonal band of the (row, slot) plane per K
// steps.  Storage per pair: ceil((lx + n - 1) / K) n (R + 1) floats, about
// B (lx + n) / K, against lx B cells of the cost matrix.
//
// Backward.  Bands are taken from the last to the first.  The wave restores
// the band's starting state, replays its K steps with the forward's own
// arithmetic, and records for every cell the two decisions "min == diag" and
// "min == up" (one word per plane, lane and step, in LDS).  Then one lane per
// pair walks dtw_mapping's backtrack (distance/dtw.py:395-423: from
// (lx-1, ly-1) to the first minimum of [diag, up, left]) through the band's
// cells -- a cell (i, k) of lane q belongs to step i + q, which never grows
// along the walk -- until it reaches a cell of an earlier band.  The path is
// monotone: the walk writes, once per row and once per column of the pair,
// the path's segment there (first index and cell count).  The walk touches
// only LDS and these stores, so no memory wait sits on its chain.
//
// Gradients.  One thread per row (and per column) of a pair sums its
// segment's cells:
//     dD/dx_i = G/D sum_j w_ij (x_i - y_j),   dD/dy_j = -G/D sum_i w_ij (x_i - y_j)
// (D = sqrt of the squared cost, 0 when D == 0; w_ij = 1 for DTW, the WDTW
// weight of the cell otherwise, w[j - 1] on row 0 as the forward weighs it):
// straight into dX / dY for paired calls, summed over pairs in a fixed order
// for all-pairs calls (the result is bitwise deterministic).
//
// Cell arithmetic (float32, as the group kernel): c = fma(d, d, min3(up, left,
// diag)), WDTW c = fma(d * d, w, min3(...)), d = x_i - y_j.
#include "wbx_internal.h"

namespace wbx {

namespace {

constexpr int kSegK = 32;  // steps per checkpoint band
constexpr int kSegU = 4;   // steps per group (one window shift and load batch)

struct SegGeo {
    int lx, ly, off, B, kf;
    int n, a, R, P;      // lanes per pair, long lanes, strip size, pairs per wave
    int nsteps, nseg;    // lx + n - 1 steps in ceil(nsteps / K) bands
    int nw;              // WDTW weight count (0: DTW)
};

// The path of pair p as segments: rows[p][i] = (j0 << 16) | n, the path's
// cells in row i being columns j0 .. j0 + n - 1; cols[p][j] = (i0 << 16) | n
// likewise for column j (lengths below 2^16).
struct SegPath {
    uint32_t *rows, *cols;
};

__device__ __forceinline__ float sg_shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float sg_shl1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(WBX_INF_F), __float_as_int(v),
                                                      0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ float sg_ldx(const float *__restrict__ X, int i, int lx) {
    return X[i < 0 ? 0 : (i >= lx ? lx - 1 : i)];
}
__device__ __forceinline__ float sg_ldy(const float *__restrict__ Y, int j, int ly) {
    const float v = Y[j < 0 ? 0 : (j >= ly ? ly - 1 : j)];
    return (unsigned)j < (unsigned)ly ? v : WBX_INF_F;
}

template <int R, bool WT>
struct SLane {
    static_assert(R + 1 <= 32, "one decision word per plane, lane and step");
    static constexpr int S = R + 1;
    float row[S];
    float yw[S + kSegU - 1];  // yw[m] = y[yj + m]: step u of a group reads yw[u + m]
    float wt[WT ? S : 1], w0[WT ? S : 1];
    float l;      // last real slot of the lane's previous row
    int r, yj, k0;
    bool first, last, shortl;
};

// Slot k's lane and first slot.
__device__ __forceinline__ void slot_lane(int k, int a, int R, int &q, int &k0) {
    const int S = R + 1;
    if (k < a * S) {
        q = k / S;
        k0 = q * S;
    } else {
        q = a + (k - a * S) / R;
        k0 = a * S + (q - a) * R;
    }
}

template <int R, bool WT>
__device__ __forceinline__ void seg_init(SLane<R, WT> &L, const SegGeo &g, int q,
                                         const float *__restrict__ wv) {
    constexpr int S = R + 1;
    L.first = q == 0;
    L.last = q == g.n - 1;
    L.shortl = q >= g.a;
    L.k0 = q < g.a ? q * S : g.a * S + (q - g.a) * R;
    if constexpr (WT) {
        const int top = g.nw - 1;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const int k = L.k0 + m - g.off;  // j - i of the slot; column j on row 0
            const int dw = k < 0 ? -k : k;
            L.wt[m] = wv[dw < top ? dw : top];
            const int c0 = k >= 1 ? k - 1 : 0;
            L.w0[m] = wv[c0 < top ? c0 : top];
        }
    }
}

// The lane's state at the start of step s: band row `rowv`, row s - q.
template <int R, bool WT>
__device__ __forceinline__ void seg_restore(SLane<R, WT> &L, const SegGeo &g, int q, int s,
                                            const float *__restrict__ Y) {
    constexpr int S = R + 1;
    L.l = L.shortl ? L.row[R - 1] : L.row[R];
    L.r = s - q;
    L.yj = L.r + L.k0 - g.off;
#pragma unroll
    for (int m = 0; m < S + kSegU - 1; ++m) L.yw[m] = sg_ldy(Y, L.yj + m, g.ly);
}

// A cell of the replay: the same float32 value as the forward's
// (fma(d, d, mn), WDTW fma(d * d, w, mn)) and two decision bits shifted into
// pd, pu: compare into an SGPR mask, then an add-with-carry doubles the word
// and adds the bit.  Each mask is read four or more VALU instructions after
// it is written, which covers the wait states between a VALU mask write and a
// VALU carry read.
template <bool WT>
__device__ __forceinline__ float rec_cell(float mn, float dg, float up, float x, float y, float w,
                                          uint32_t &pd, uint32_t &pu) {
    float c = mn, d;
    uint64_t sa, sb;
    if constexpr (WT)
        asm("v_cmp_eq_f32_e64 %[sa], %[c], %[dg]\n\t"
            "v_cmp_eq_f32_e64 %[sb], %[c], %[up]\n\t"
            "v_sub_f32_e32 %[d], %[x], %[y]\n\t"
            "v_mul_f32_e32 %[d], %[d], %[d]\n\t"
            "v_fmac_f32_e32 %[c], %[d], %[w]\n\t"
            "v_addc_co_u32_e64 %[pd], %[sa], %[pd], %[pd], %[sa]\n\t"
            "v_addc_co_u32_e64 %[pu], %[sb], %[pu], %[pu], %[sb]"
            : [c] "+v"(c), [pd] "+v"(pd), [pu] "+v"(pu), [d] "=&v"(d), [sa] "=&s"(sa),
              [sb] "=&s"(sb)
            : [dg] "v"(dg), [up] "v"(up), [x] "v"(x), [y] "v"(y), [w] "v"(w));
    else
        asm("v_cmp_eq_f32_e64 %[sa], %[c], %[dg]\n\t"
            "v_cmp_eq_f32_e64 %[sb], %[c], %[up]\n\t"
            "v_sub_f32_e32 %[d], %[x], %[y]\n\t"
            "v_fmac_f32_e32 %[c], %[d], %[d]\n\t"
            "s_nop 0\n\t"
            "v_addc_co_u32_e64 %[pd], %[sa], %[pd], %[pd], %[sa]\n\t"
            "v_addc_co_u32_e64 %[pu], %[sb], %[pu], %[pu], %[sb]"
            : [c] "+v"(c), [pd] "+v"(pd), [pu] "+v"(pu), [d] "=&v"(d), [sa] "=&s"(sa),
              [sb] "=&s"(sb)
            : [dg] "v"(dg), [up] "v"(up), [x] "v"(x), [y] "v"(y));
    (void)w;
    return c;
}

// One step: row L.r of the lane.  SEL keeps the state of a lane outside
// [0, lx); REC shifts into bd / bu, slot by slot, whether the slot's minimum
// equals its diag / up input (bd, bu start at 0).
template <int R, bool WT, bool SEL, bool REC, int YO = 0>
__device__ __forceinline__ void seg_step(SLane<R, WT> &L, float x, int lx, uint32_t &bd,
                                         uint32_t &bu) {
    constexpr int S = R + 1;
    const bool act = !SEL || (unsigned)L.r < (unsigned)lx;
    const bool row0 = WT && L.r == 0;
    auto wgt = [&](int m) -> float {
        if constexpr (WT)
            return row0 ? L.w0[m] : L.wt[m];
        else
            return 1.0f;
    };
    auto cell = [&](float up, float left, float diag, float y, int m) -> float {
        const float mn = fminf(fminf(up, left), diag);
        if constexpr (REC)  // shifted in from the low end: slot m ends at bit S - 1 - m
            return rec_cell<WT>(mn, diag, up, x, y, wgt(m), bd, bu);
        const float d = x - y;
        if constexpr (WT)
            return __builtin_fmaf(d * d, wgt(m), mn);
        else
            return __builtin_fmaf(d, d, mn);
    };
    float Lin = sg_shr1(L.l);
    Lin = L.first ? WBX_INF_F : Lin;
    const float c0 = cell(L.row[1], Lin, L.row[0], L.yw[YO], 0);
    float Uin = sg_shl1(SEL && !act ? L.row[0] : c0);
    Uin = L.last ? WBX_INF_F : Uin;
    float left = c0;
    const float r0 = SEL ? (act ? c0 : L.row[0]) : c0;
#pragma unroll
    for (int m = 1; m < S; ++m) {
        float up;
        if (m == S - 1)
            up = Uin;
        else if (m == S - 2)
            up = L.shortl ? Uin : L.row[m + 1];
        else
            up = L.row[m + 1];
        const float c = cell(up, left, L.row[m], L.yw[YO + m], m);
        L.row[m] = SEL ? (act ? c : L.row[m]) : c;
        left = c;
    }
    L.row[0] = r0;
    L.l = L.shortl ? L.row[R - 1] : L.row[R];
}

// One step without a group: the window moves on by one.
template <int R, bool WT>
__device__ __forceinline__ void seg_advance(SLane<R, WT> &L, const float *__restrict__ Y,
                                            int ly) {
    constexpr int W = R + kSegU;  // S + kSegU - 1 entries
#pragma unroll
    for (int m = 0; m < W - 1; ++m) L.yw[m] = L.yw[m + 1];
    L.yw[W - 1] = sg_ldy(Y, L.yj + W, ly);
    L.r += 1;
    L.yj += 1;
}

// kSegU steps: the next group's x and window entries are loaded first, the
// steps read the window at offsets 0 .. kSegU - 1, then it moves on by kSegU.
// xv holds this group's x and receives the next group's.
template <int R, bool WT, bool SEL, bool REC>
__device__ __forceinline__ void seg_group(SLane<R, WT> &L, float (&xv)[kSegU],
                                          const float *__restrict__ X, int lx,
                                          const float *__restrict__ Y, int ly,
                                          uint32_t (&bd)[kSegU], uint32_t (&bu)[kSegU]) {
    constexpr int W = R + kSegU;
    float xn[kSegU], yn[kSegU];
#pragma unroll
    for (int u = 0; u < kSegU; ++u) {
        xn[u] = sg_ldx(X, L.r + kSegU + u, lx);
        yn[u] = sg_ldy(Y, L.yj + W + u, ly);
    }
#define WBX_SEG_STEP(u)                                                  \
    seg_step<R, WT, SEL, REC, u>(L, xv[u], lx, bd[u], bu[u]);            \
    L.r += 1;
    WBX_SEG_STEP(0)
    WBX_SEG_STEP(1)
    WBX_SEG_STEP(2)
    WBX_SEG_STEP(3)
#undef WBX_SEG_STEP
    static_assert(kSegU == 4, "four steps per group");
#pragma unroll
    for (int m = 0; m < W - kSegU; ++m) L.yw[m] = L.yw[m + kSegU];
#pragma unroll
    for (int u = 0; u < kSegU; ++u) {
        L.yw[W - kSegU + u] = yn[u];
        xv[u] = xn[u];
    }
    L.yj += kSegU;
}

template <int R, bool WT>
__global__ __launch_bounds__(64) void k_seg_forward(PairSpec ps, Series<float> xs,
                                                    Series<float> ys, SegGeo g,
                                                    const float *__restrict__ wv,
                                                    float *__restrict__ sq,
                                                    float *__restrict__ ckpt) {
    constexpr int S = R + 1;
    const int lane = threadIdx.x;
    const int gi = lane / g.n;
    const int q = lane - gi * g.n;
    const int64_t p = (int64_t)blockIdx.x * g.P + gi;
    const bool valid = gi < g.P && p < ps.n_pairs;
    int64_t a = 0, b = 0;
    if (valid) decode_pair(ps, p, a, b);
    const float *__restrict__ X = xs.row(a);
    const float *__restrict__ Y = ys.row(b);
    SLane<R, WT> L;
    seg_init(L, g, q, wv);
#pragma unroll
    for (int m = 0; m < S; ++m) L.row[m] = (L.k0 + m == g.off) ? 0.0f : WBX_INF_F;
    seg_restore(L, g, q, 0, Y);
    float *__restrict__ ck = ckpt + ((p * g.nseg) * g.n + q) * S;
    const int64_t cstride = (int64_t)g.n * S;
    uint32_t nb[kSegU] = {};
    float xv[kSegU];
#pragma unroll
    for (int u = 0; u < kSegU; ++u) xv[u] = sg_ldx(X, L.r + u, g.lx);
    int s = 0;
    // WDTW's forward keeps single steps: its grouped form needs more
    // registers than two waves per SIMD allow
    for (; !WT && s + kSegU <= g.nsteps; s += kSegU) {  // kSegK is a multiple of kSegU
        if (s % kSegK == 0 && valid) {
            float *dst = ck + (s / kSegK) * cstride;
#pragma unroll
            for (int m = 0; m < S; ++m) dst[m] = L.row[m];
        }
        if (s < g.n || s + kSegU > g.lx)
            seg_group<R, WT, true, false>(L, xv, X, g.lx, Y, g.ly, nb, nb);
        else
            seg_group<R, WT, false, false>(L, xv, X, g.lx, Y, g.ly, nb, nb);
    }
    for (; s < g.nsteps; ++s) {  // the last steps one at a time (WDTW: every step)
        if (s % kSegK == 0 && valid) {
            float *dst = ck + (s / kSegK) * cstride;
#pragma unroll
            for (int m = 0; m < S; ++m) dst[m] = L.row[m];
        }
        seg_step<R, WT, true, false>(L, sg_ldx(X, L.r, g.lx), g.lx, nb[0], nb[0]);
        seg_advance(L, Y, g.ly);
    }
    const int nslots = L.shortl ? R : S;
    if (valid && g.kf >= L.k0 && g.kf < L.k0 + nslots) {
        float res = WBX_INF_F;
#pragma unroll
        for (int m = 0; m < S; ++m)
            if (L.k0 + m == g.kf) res = L.row[m];
        sq[p] = res;
    }
}

template <int R, bool WT>
__global__ __launch_bounds__(64) void k_seg_backward(PairSpec ps, Series<float> xs,
                                                     Series<float> ys, SegGeo g,
                                                     const float *__restrict__ wv,
                                                     const float *__restrict__ ckpt, SegPath out) {
    constexpr int S = R + 1;
    __shared__ uint32_t bits[kSegK][64][2];
    const int lane = threadIdx.x;
    const int gi = lane / g.n;
    const int q = lane - gi * g.n;
    const int64_t p = (int64_t)blockIdx.x * g.P + gi;
    const bool valid = gi < g.P && p < ps.n_pairs;
    int64_t a = 0, b = 0;
    if (valid) decode_pair(ps, p, a, b);
    const float *__restrict__ X = xs.row(a);
    const float *__restrict__ Y = ys.row(b);
    SLane<R, WT> L;
    seg_init(L, g, q, wv);
    const float *__restrict__ ck = ckpt + ((p * g.nseg) * g.n + q) * S;
    const int64_t cstride = (int64_t)g.n * S;
    const int base = gi * g.n;  // the group's first lane

    // Walk state (lane q == 0 of each valid group).  The walk reads only the
    // decisions in LDS and writes one word per finished row and column, so no
    // memory wait sits on its chain.
    const bool walker = valid && q == 0;
    int wi = g.lx - 1, wk = g.kf;
    bool done = !walker;
    int rn = 0, cn = 0;  // cells of the path so far in the current row / column
    uint32_t *__restrict__ prow = walker ? out.rows + p * g.lx : nullptr;
    uint32_t *__restrict__ pcol = walker ? out.cols + p * g.ly : nullptr;
    // slot wk's lane qq and its position bit in the lane (tracked as the walk
    // moves by one slot at most)
    int qq = 0, kk0 = 0;
    slot_lane(wk, g.a, g.R, qq, kk0);
    int bit = wk - kk0;

    for (int c = g.nseg - 1; c >= 0; --c) {
        const int s0 = c * kSegK;
        const int s1 = s0 + kSegK < g.nsteps ? s0 + kSegK : g.nsteps;
        // lanes of no pair read pair 0's first checkpoint (in bounds) and drop it
        const float *src = valid ? ck + c * cstride : ckpt;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const float v = src[m];
            L.row[m] = valid ? v : WBX_INF_F;
        }
        seg_restore(L, g, q, s0, Y);
        float xv[kSegU];
#pragma unroll
        for (int u = 0; u < kSegU; ++u) xv[u] = sg_ldx(X, L.r + u, g.lx);
        int s = s0;
        for (; s + kSegU <= s1; s += kSegU) {
            uint32_t bd[kSegU] = {}, bu[kSegU] = {};
            if (s < g.n || s + kSegU > g.lx)
                seg_group<R, WT, true, true>(L, xv, X, g.lx, Y, g.ly, bd, bu);
            else
                seg_group<R, WT, false, true>(L, xv, X, g.lx, Y, g.ly, bd, bu);
#pragma unroll
            for (int u = 0; u < kSegU; ++u) {
                bits[s + u - s0][lane][0] = bd[u];
                bits[s + u - s0][lane][1] = bu[u];
            }
        }
        for (; s < s1; ++s) {  // the pair's last steps one at a time
            uint32_t bd = 0, bu = 0;
            seg_step<R, WT, true, true>(L, sg_ldx(X, L.r, g.lx), g.lx, bd, bu);
            bits[s - s0][lane][0] = bd;
            bits[s - s0][lane][1] = bu;
            seg_advance(L, Y, g.ly);
        }
        __syncthreads();
        while (!done) {
            const int sc = wi + qq;  // step of cell (wi, wk)
            if (sc < s0) break;      // the cell is in an earlier band
            const int j = wi + wk - g.off;
            ++rn;
            ++cn;
            if (wi == 0 && j == 0) {
                prow[0] = (uint32_t)rn;
                pcol[0] = (uint32_t)cn;
                done = true;
                break;
            }
            const uint32_t bdw = bits[sc - s0][base + qq][0], buw = bits[sc - s0][base + qq][1];
            const int wb = g.R - bit;  // word bit of the slot (R + 1 bits per word)
            const bool mi = ((bdw | buw) >> wb) & 1u;       // the walk leaves row wi
            const bool mj = !(((buw & ~bdw) >> wb) & 1u);  // and column j (not "up")
            if (mi) {
                prow[wi] = ((uint32_t)j << 16) | (uint32_t)rn;
                rn = 0;
            }
            if (mj) {
                pcol[j] = ((uint32_t)wi << 16) | (uint32_t)cn;
                cn = 0;
            }
            // left: slot wk - 1; up: row wi - 1, slot wk + 1; diag: row wi - 1
            const int dk = !mi ? -1 : (!mj ? 1 : 0);
            wi -= mi ? 1 : 0;
            wk += dk;
            bit += dk;
            const bool below = bit < 0;
            const bool above = bit == (qq < g.a ? g.R + 1 : g.R);
            qq += below ? -1 : (above ? 1 : 0);
            bit = below ? (qq < g.a ? g.R + 1 : g.R) - 1 : (above ? 0 : bit);
            if (wi < 0 || wi + wk - g.off < 0) {  // not reached: every path ends at (0, 0)
                done = true;
                break;
            }
        }
        __syncthreads();
    }
}

// coef[p] = G[p] / D_p (0 when D_p == 0), D_p = sqrt of the squared cost.
__global__ __launch_bounds__(256) void k_seg_coef(const float *__restrict__ sq,
                                                  const double *__restrict__ G, int64_t n,
                                                  float *__restrict__ coef) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const double D = sqrt((double)sq[p]);
    coef[p] = D > 0.0 ? (float)(G[p] / D) : 0.0f;
}

// The weight of cell (i, j) as the forward applied it.
template <bool WT>
__device__ __forceinline__ float seg_w(const float *__restrict__ wv, int top, int i, int j) {
    if constexpr (!WT) {
        return 1.0f;
    } else {
        int dw;
        if (i == 0)
            dw = j > 0 ? j - 1 : 0;
        else
            dw = i > j ? i - j : j - i;
        return wv[dw < top ? dw : top];
    }
}

// Sum over row i's path cells of w (x_i - y_j) (columns j0 .. j0 + n - 1).
template <bool WT>
__device__ __forceinline__ float seg_row_sum(uint32_t seg, const float *__restrict__ x,
                                             const float *__restrict__ y,
                                             const float *__restrict__ wv, int top, int i) {
    const int j0 = (int)(seg >> 16), n = (int)(seg & 0xFFFFu);
    const float xi = x[i];
    float s = 0.0f;
    for (int j = j0; j < j0 + n; ++j) s += seg_w<WT>(wv, top, i, j) * (xi - y[j]);
    return s;
}

// Minus the sum over column j's path cells of w (x_i - y_j).
template <bool WT>
__device__ __forceinline__ float seg_col_sum(uint32_t seg, const float *__restrict__ x,
                                             const float *__restrict__ y,
                                             const float *__restrict__ wv, int top, int j) {
    const int i0 = (int)(seg >> 16), n = (int)(seg & 0xFFFFu);
    const float yj = y[j];
    float s = 0.0f;
    for (int i = i0; i < i0 + n; ++i) s -= seg_w<WT>(wv, top, i, j) * (x[i] - yj);
    return s;
}

// P[s * (len + 1) + i] = sum of series s's first i values in float64; one
// wave per series (a fixed sum order: deterministic).
__global__ __launch_bounds__(64) void k_seg_prefix(const float *__restrict__ X, int64_t stride,
                                                   int len, double *__restrict__ P) {
    const int64_t sidx = blockIdx.x;
    const int lane = threadIdx.x;
    const float *x = X + sidx * stride;
    double *out = P + sidx * (int64_t)(len + 1);
    if (lane == 0) out[0] = 0.0;
    double carry = 0.0;
    for (int i0 = 0; i0 < len; i0 += 64) {
        const int i = i0 + lane;
        double v = i < len ? (double)x[i] : 0.0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(v, d);
            if (lane >= d) v += o;
        }
        v += carry;
        if (i < len) out[i + 1] = v;
        carry = __shfl(v, 63);
    }
}

// Block (c, p - p0): rows and columns c * 256 + threadIdx.x of pair p (rows first,
// then columns; DTW sums a segment from the float64 prefix sums of the other
// series, WDTW cell by cell): the gradient of the row / column, into dX / dY rows p (paired
// calls: ld = the caller's strides), or, for all pairs, over the segment word
// itself (gx = path.rows, gy = path.cols as floats: each thread reads its own
// word before writing it).
template <bool WT>
__global__ __launch_bounds__(256) void k_seg_grad_pairs(
    SegPath path, PairSpec ps, int64_t p0, Series<float> xs, Series<float> ys,
    const float *__restrict__ wv, int nw, const float *__restrict__ coef,
    const double *__restrict__ PX, const double *__restrict__ PY, float *gx, int64_t ldx,
    float *gy, int64_t ldy) {
    const int64_t p = p0 + blockIdx.y;
    const int lx = xs.len, ly = ys.len;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= lx + ly) return;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const float cf = coef[p];
    if (t < lx) {
        const uint32_t seg = path.rows[p * lx + t];
        float s;
        if constexpr (WT) {
            s = seg_row_sum<WT>(seg, xs.row(a), ys.row(b), wv, nw - 1, t);
        } else {  // n x_i - (y_j0 + ... + y_(j0+n-1)) from the prefix sums of y_b
            const int j0 = (int)(seg >> 16), n = (int)(seg & 0xFFFFu);
            const double *py = PY + b * (int64_t)(ly + 1);
            s = (float)((double)n * (double)xs.row(a)[t] - (py[j0 + n] - py[j0]));
        }
        gx[p * ldx + t] = cf * s;
    } else {
        const int j = t - lx;
        const uint32_t seg = path.cols[p * ly + j];
        float s;
        if constexpr (WT) {
            s = seg_col_sum<WT>(seg, xs.row(a), ys.row(b), wv, nw - 1, j);
        } else {  // n y_j - (x_i0 + ... + x_(i0+n-1)) from the prefix sums of x_a
            const int i0 = (int)(seg >> 16), n = (int)(seg & 0xFFFFu);
            const double *px = PX + a * (int64_t)(lx + 1);
            s = (float)((double)n * (double)ys.row(b)[j] - (px[i0 + n] - px[i0]));
        }
        gy[p * ldy + j] = cf * s;
    }
}

// All pairs: dX[a][i] = sum over b (in order) of gx[a ny + b][i].
__global__ __launch_bounds__(256) void k_seg_sum_x(const float *__restrict__ gx, int64_t nx,
                                                   int64_t ny, int lx, float *__restrict__ dX,
                                                   int64_t ldx) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nx * lx) return;
    const int64_t a = t / lx;
    const int i = (int)(t - a * lx);
    const float *src = gx + a * ny * lx + i;
    float s = 0.0f;
    for (int64_t b = 0; b < ny; ++b) s += src[b * lx];
    dX[a * ldx + i] = s;
}

// All pairs, columns in two fixed-order passes: part[c][b][j] = sum over a in
// chunk c (kSegChunk series, in order) of gy[a ny + b][j], then dY[b][j] =
// sum of the parts in order of c.
constexpr int kSegChunk = 256;

__global__ __launch_bounds__(256) void k_seg_sum_y_part(const float *__restrict__ gy, int64_t nx,
                                                        int64_t ny, int ly,
                                                        float *__restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ny * ly) return;
    const int64_t c = blockIdx.y;
    const int64_t a0 = c * kSegChunk, a1 = a0 + kSegChunk < nx ? a0 + kSegChunk : nx;
    const int64_t step = ny * ly;
    const float *src = gy + t;
    float s = 0.0f;
    for (int64_t a = a0; a < a1; ++a) s += src[a * step];
    part[c * step + t] = s;
}

// DTW, all pairs: the two sums above with each pair's term formed from its
// segment word and the prefix sums on the fly (the same float arithmetic as
// k_seg_grad_pairs followed by k_seg_sum_x / k_seg_sum_y_part, without the
// per-pair terms in memory).
__global__ __launch_bounds__(256) void k_seg_sum_x_dtw(const uint32_t *__restrict__ rows,
                                                       const float *__restrict__ coef,
                                                       Series<float> xs, const double *__restrict__ PY,
                                                       int64_t nx, int64_t ny, int ly,
                                                       float *__restrict__ dX, int64_t ldx) {
    const int lx = xs.len;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nx * lx) return;
    const int64_t a = t / lx;
    const int i = (int)(t - a * lx);
    const double xi = (double)xs.row(a)[i];
    const uint32_t *src = rows + a * ny * lx + i;
    const float *cf = coef + a * ny;
    float s = 0.0f;
#pragma unroll 8
    for (int64_t b = 0; b < ny; ++b) {
        const uint32_t seg = src[b * lx];
        const int j0 = (int)(seg >> 16), n = (int)(seg & 0xFFFFu);
        const double *py = PY + b * (int64_t)(ly + 1);
        s += cf[b] * (float)((double)n * xi - (py[j0 + n] - py[j0]));
    }
    dX[a * ldx + i] = s;
}

__global__ __launch_bounds__(256) void k_seg_sum_y_part_dtw(const uint32_t *__restrict__ cols,
                                                            const float *__restrict__ coef,
                                                            Series<float> ys,
                                                            const double *__restrict__ PX, int64_t nx,
                                                            int64_t ny, int lx,
                                                            float *__restrict__ part) {
    const int ly = ys.len;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ny * ly) return;
    const int64_t b = t / ly;
    const int j = (int)(t - b * ly);
    const int64_t c = blockIdx.y;
    const int64_t a0 = c * kSegChunk, a1 = a0 + kSegChunk < nx ? a0 + kSegChunk : nx;
    const double yj = (double)ys.row(b)[j];
    const int64_t step = ny * ly;
    float s = 0.0f;
#pragma unroll 8
    for (int64_t a = a0; a < a1; ++a) {
        const uint32_t seg = cols[a * step + t];
        const int i0 = (int)(seg >> 16), n = (int)(seg & 0xFFFFu);
        const double *px = PX + a * (int64_t)(lx + 1);
        s += coef[a * ny + b] * (float)((double)n * yj - (px[i0 + n] - px[i0]));
    }
    part[c * step + t] = s;
}

__global__ __launch_bounds__(256) void k_seg_sum_y_final(const float *__restrict__ part, int nc,
                                                         int64_t ny, int ly,
                                                         float *__restrict__ dY, int64_t ldy) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ny * ly) return;
    const int64_t step = ny * ly;
    float s = 0.0f;
    for (int c = 0; c < nc; ++c) s += part[c * step + t];
    const int64_t b = t / ly;
    dY[b * ldy + (t - b * ly)] = s;
}

// Strip sizes of the checkpointed kernels: a lane's R + 1 decisions per
// plane and step fit one 32-bit word.
constexpr int kSegR[] = {2, 3, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28};

// Among the compiled R with n R <= B <= n (R + 1), n <= 64, the one with the
// most band cells per VALU instruction of a wave step (the 64 / n pairs of a
// wave, about k (R + 1) + 8 instructions per lane and step, the n - 1 fill and
// drain steps).
bool seg_plan(int B, int lx, bool weighted, GroupGeo *g) {
    const double k = weighted ? 4.0 : 3.0;
    double best = -1.0;
    for (int R : kSegR) {
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

int seg_geo(const wbx_metric *m, int64_t tx, int64_t ty, SegGeo *g) {
    // path segments pack a start and a count into 16 bits each
    if (tx >= (1 << 16) || ty >= (1 << 16))
        return fail(WBX_EUNSUPPORTED, "differentiable DTW: series of 65536 or more steps");
    const Band bd = make_band((int)tx, (int)ty, m->r);
    GroupGeo gg{};
    if (!seg_plan(bd.B, bd.lx, m->kind == WBX_WDTW, &gg))
        return fail(WBX_EUNSUPPORTED, "differentiable DTW: band of %d slots is too wide (at most %d)",
                    bd.B, 64 * 29);
    *g = SegGeo{};
    g->lx = bd.lx;
    g->ly = bd.ly;
    g->off = bd.off;
    g->B = bd.B;
    g->kf = bd.ly - bd.lx + bd.off;
    g->n = gg.n;
    g->a = gg.a;
    g->R = gg.R;
    g->P = gg.P;
    g->nsteps = bd.lx + gg.n - 1;
    g->nseg = (g->nsteps + kSegK - 1) / kSegK;
    g->nw = m->kind == WBX_WDTW ? (int)(tx > ty ? tx : ty) : 0;
    return WBX_OK;
}

struct SegState {
    int64_t sq, w, ck, total;  // byte offsets / size
};

SegState seg_state(const SegGeo &g, int64_t np) {
    auto al = [](int64_t v) { return (v + 255) / 256 * 256; };
    SegState s{};
    int64_t o = 0;
    s.sq = o;
    o += al(np * 4);
    s.w = o;
    o += al((int64_t)(g.nw > 0 ? g.nw : 1) * 4);
    s.ck = o;
    o += al(np * g.nseg * g.n * (g.R + 1) * 4);
    s.total = o;
    return s;
}

#define SEG_R_CASES(X) X(2) X(3) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(20) X(24) X(28)

template <bool WT>
int seg_launch_fwd(const PairSpec &ps, Series<float> xs, Series<float> ys, const SegGeo &g,
                   const float *w, float *sq, float *ck, hipStream_t st) {
    const dim3 grid((unsigned)((ps.n_pairs + g.P - 1) / g.P));
    switch (g.R) {
#define WBX_SF(RR)                                                                           \
    case RR:                                                                                 \
        hipLaunchKernelGGL((k_seg_forward<RR, WT>), grid, dim3(64), 0, st, ps, xs, ys, g, w, \
                           sq, ck);                                                          \
        break;
        SEG_R_CASES(WBX_SF)
#undef WBX_SF
        default: return fail(WBX_EUNSUPPORTED, "differentiable DTW: strip size %d", g.R);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template <bool WT>
int seg_launch_bwd(const PairSpec &ps, Series<float> xs, Series<float> ys, const SegGeo &g,
                   const float *w, const float *ck, SegPath o, hipStream_t st) {
    const dim3 grid((unsigned)((ps.n_pairs + g.P - 1) / g.P));
    switch (g.R) {
#define WBX_SB(RR)                                                                            \
    case RR:                                                                                  \
        hipLaunchKernelGGL((k_seg_backward<RR, WT>), grid, dim3(64), 0, st, ps, xs, ys, g, w, \
                           ck, o);                                                            \
        break;
        SEG_R_CASES(WBX_SB)
#undef WBX_SB
        default: return fail(WBX_EUNSUPPORTED, "differentiable DTW: strip size %d", g.R);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

// The gradients from the path segments (and, for all pairs, their fixed-order
// sums over pairs).
template <bool WT>
int seg_launch_grad(int mode, SegPath path, int64_t nx, int64_t ny, Series<float> xs,
                    Series<float> ys, const float *w, int nw, const float *sq, const double *G,
                    float *dX, int64_t ldx, float *dY, int64_t ldy, float *coef, float *part,
                    double *PX, double *PY, hipStream_t st) {
    const int64_t tx = xs.len, ty = ys.len;
    const int64_t np = mode == 1 ? nx : nx * ny;
    hipLaunchKernelGGL(k_seg_coef, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, sq, G, np,
                       coef);
    if (!WT) {
        hipLaunchKernelGGL(k_seg_prefix, dim3((unsigned)nx), dim3(64), 0, st, xs.base, xs.stride,
                           (int)tx, PX);
        hipLaunchKernelGGL(k_seg_prefix, dim3((unsigned)ny), dim3(64), 0, st, ys.base, ys.stride,
                           (int)ty, PY);
    }
    PairSpec ps{mode == 0 ? PAIR_GRID : PAIR_DIAG, np, mode == 0 ? ny : nx, nullptr, nullptr};
    // pairs along y (at most 65535 per launch), rows and columns along x
    const unsigned gxb = (unsigned)((tx + ty + 255) / 256);
    auto grads = [&](float *gx, int64_t lgx, float *gy, int64_t lgy) {
        for (int64_t p0 = 0; p0 < np; p0 += 65535) {
            const int64_t cnt = np - p0 < 65535 ? np - p0 : 65535;
            hipLaunchKernelGGL(k_seg_grad_pairs<WT>, dim3(gxb, (unsigned)cnt), dim3(256), 0, st,
                               path, ps, p0, xs, ys, w, nw, coef, PX, PY, gx, lgx, gy, lgy);
        }
    };
    if (mode == 1) {
        grads(dX, ldx, dY, ldy);
    } else if (!WT) {
        hipLaunchKernelGGL(k_seg_sum_x_dtw, dim3((unsigned)((nx * tx + 255) / 256)), dim3(256), 0,
                           st, path.rows, coef, xs, PY, nx, ny, (int)ty, dX, ldx);
        const int nc = (int)((nx + kSegChunk - 1) / kSegChunk);
        const unsigned yb = (unsigned)((ny * ty + 255) / 256);
        hipLaunchKernelGGL(k_seg_sum_y_part_dtw, dim3(yb, (unsigned)nc), dim3(256), 0, st,
                           path.cols, coef, ys, PX, nx, ny, (int)tx, part);
        hipLaunchKernelGGL(k_seg_sum_y_final, dim3(yb), dim3(256), 0, st, part, nc, ny, (int)ty,
                           dY, ldy);
    } else {
        float *gx = (float *)path.rows, *gy = (float *)path.cols;
        grads(gx, tx, gy, ty);
        hipLaunchKernelGGL(k_seg_sum_x, dim3((unsigned)((nx * tx + 255) / 256)), dim3(256), 0, st,
                           gx, nx, ny, (int)tx, dX, ldx);
        const int nc = (int)((nx + kSegChunk - 1) / kSegChunk);
        const unsigned yb = (unsigned)((ny * ty + 255) / 256);
        hipLaunchKernelGGL(k_seg_sum_y_part, dim3(yb, (unsigned)nc), dim3(256), 0, st, gy, nx, ny,
                           (int)ty, part);
        hipLaunchKernelGGL(k_seg_sum_y_final, dim3(yb), dim3(256), 0, st, part, nc, ny, (int)ty,
                           dY, ldy);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace

// Backward scratch: the path segments (one word per row and column of each
// pair), G / D per pair, for all pairs the column partials, and the float64
// prefix sums of X and Y (DTW segment sums).
struct SegScratch {
    int64_t rows, cols, coef, part, px, py, total;
};

static SegScratch seg_scratch(int mode, int64_t nx, int64_t ny, int64_t tx, int64_t ty) {
    auto al = [](int64_t v) { return (v + 255) / 256 * 256; };
    const int64_t np = mode == 0 ? nx * ny : nx;
    const int64_t nc = (nx + kSegChunk - 1) / kSegChunk;
    SegScratch s{};
    int64_t o = 0;
    s.rows = o;
    o += al(np * tx * 4);
    s.cols = o;
    o += al(np * ty * 4);
    s.coef = o;
    o += al(np * 4);
    s.part = o;
    o += al(mode == 0 ? nc * ny * ty * 4 : 0);
    s.px = o;
    o += al(nx * (tx + 1) * 8);
    s.py = o;
    o += al(ny * (ty + 1) * 8);
    s.total = o;
    return s;
}

int seg_train_sizes(const wbx_metric *m, int mode, int64_t nx, int64_t ny, int64_t tx, int64_t ty,
                    int64_t *state_bytes, int64_t *scratch_bytes) {
    SegGeo g;
    int rc = seg_geo(m, tx, ty, &g);
    if (rc) return rc;
    const int64_t np = mode == 0 ? nx * ny : nx;
    if (state_bytes) *state_bytes = seg_state(g, np).total;
    // the path segments (one word per row and column of each pair), G / D per
    // pair, for all pairs the column partials, and the prefix sums of X and Y
    if (scratch_bytes) *scratch_bytes = seg_scratch(mode, nx, ny, tx, ty).total;
    return WBX_OK;
}

int seg_train_forward(const wbx_metric *m, int mode, const float *X, int64_t nx, int64_t x_stride,
                      int64_t tx, const float *Y, int64_t ny, int64_t y_stride, int64_t ty,
                      double *D, void *state, hipStream_t st) {
    SegGeo g;
    int rc = seg_geo(m, tx, ty, &g);
    if (rc) return rc;
    PairSpec ps{mode == 0 ? PAIR_GRID : PAIR_DIAG, mode == 0 ? nx * ny : nx, mode == 0 ? ny : nx,
                nullptr, nullptr};
    const SegState ss = seg_state(g, ps.n_pairs);
    char *sb = (char *)state;
    float *sq = (float *)(sb + ss.sq);
    float *w = (float *)(sb + ss.w);
    float *ck = (float *)(sb + ss.ck);
    if (g.nw > 0) {
        DeviceCtx *ctx = ctx_current();
        if (!ctx) return WBX_EDEVICE;
        const double *w64 = nullptr;
        const float *w32 = nullptr;
        rc = cached_weights(ctx, g.nw, m->g, &w64, &w32);
        if (rc) return rc;
        // a zero float32 weight would turn the +inf of a column outside the
        // series into NaN
        if (!(cached_weights_min(ctx, w32) > 0.0))
            return fail(WBX_EUNSUPPORTED, "differentiable WDTW: a weight rounds to zero (g = %g)",
                        m->g);
        // the state keeps its own copy: the backward does not depend on the cache
        WBX_HIP_CHECK(hipMemcpyAsync(w, w32, sizeof(float) * g.nw, hipMemcpyDeviceToDevice, st));
    }
    Series<float> xs{X, x_stride, (int)tx, nullptr}, ys{Y, y_stride, (int)ty, nullptr};
    rc = g.nw > 0 ? seg_launch_fwd<true>(ps, xs, ys, g, w, sq, ck, st)
                  : seg_launch_fwd<false>(ps, xs, ys, g, nullptr, sq, ck, st);
    if (rc) return rc;
    return launch_finish<float>(ps, sq, D, ny, st);
}

int seg_train_backward(const wbx_metric *m, int mode, const float *X, int64_t nx,
                       int64_t x_stride, int64_t tx, const float *Y, int64_t ny, int64_t y_stride,
                       int64_t ty, const double *G, const void *state, float *dX, int64_t ldx,
                       float *dY, int64_t ldy, void *scratch, hipStream_t st) {
    SegGeo g;
    int rc = seg_geo(m, tx, ty, &g);
    if (rc) return rc;
    PairSpec ps{mode == 0 ? PAIR_GRID : PAIR_DIAG, mode == 0 ? nx * ny : nx, mode == 0 ? ny : nx,
                nullptr, nullptr};
    const SegState ss = seg_state(g, ps.n_pairs);
    const char *sb = (const char *)state;
    const float *sq = (const float *)(sb + ss.sq);
    const float *w = (const float *)(sb + ss.w);
    const float *ck = (const float *)(sb + ss.ck);
    const SegScratch sc = seg_scratch(mode, nx, ny, tx, ty);
    char *scb = (char *)scratch;
    SegPath path{(uint32_t *)(scb + sc.rows), (uint32_t *)(scb + sc.cols)};
    float *coef = (float *)(scb + sc.coef);
    float *part = (float *)(scb + sc.part);
    double *px = (double *)(scb + sc.px), *py = (double *)(scb + sc.py);
    Series<float> xs{X, x_stride, (int)tx, nullptr}, ys{Y, y_stride, (int)ty, nullptr};
    rc = g.nw > 0 ? seg_launch_bwd<true>(ps, xs, ys, g, w, ck, path, st)
                  : seg_launch_bwd<false>(ps, xs, ys, g, nullptr, ck, path, st);
    if (rc) return rc;
    return g.nw > 0 ? seg_launch_grad<true>(mode, path, nx, ny, xs, ys, w, g.nw, sq, G, dX, ldx,
                                            dY, ldy, coef, part, px, py, st)
                    : seg_launch_grad<false>(mode, path, nx, ny, xs, ys, nullptr, 1, sq, G, dX,
                                             ldx, dY, ldy, coef, part, px, py, st);
}

}  // namespace wbx
