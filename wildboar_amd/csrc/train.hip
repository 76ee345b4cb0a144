// train.hip -- differentiable DTW: forward that records backtrack decisions,
// row-lockstep backtrack, and gradient reductions.
//
// The gradient of D = sqrt(C) along the optimal path P (dtw_mapping's
// backtrack, distance/dtw.py:395-423: from (n-1, m-1) step to
// argmin([diag, up, left]), the first minimum winning) is
//     dD/dx_i = sum_{(i,j) in P} (x_i - y_j) / D,   dD/dy_j = -sum ... / D
// (0 when D == 0).  The step taken out of cell (i, j) by the backtrack is the
// argmin over the same three predecessor costs the forward DP minimised at
// (i, j), so the forward records it: for every cell two bits, "min came from
// diag" and "min came from up" (diag before up before left on ties).
//
//  k_lane_train      the lane-per-pair kernel of lane.hip plus the two bit
//                    planes, one 32-bit word per 32 band slots per row, stored
//                    [wave][row][plane word][lane] so every store and load of
//                    a wave is one contiguous 256-byte line.
//  k_lane_backtrack  one pair per lane, all lanes of a wave on the same row,
//                    rows from the last to the first.  Within row i the path
//                    enters at slot k and moves left until the first cell
//                    whose decision is diag or up; that exit is a bit scan of
//                    (diag | up) over the row's words, which sit at addresses
//                    independent of the path and are fetched one row ahead.
//                    The cells of the segment add (x_i - y_j) G / D to the
//                    row's and the column's sums.
//  k_sum_*           fixed-order sums of per-pair gradients into dX, dY
//                    (the backward is bitwise deterministic).
#include "lane_common.h"

namespace wbx {

// One cell after m = min3(up, left, dg):
//   bd = (m == dg), bu = (m == up)          two compares into SGPR masks
//   c  = m + (x - y)^2                      sub + fmac
//   pd = 2 pd + bd, pu = 2 pu + bu          add-with-carry shifts each bit in
// Each mask is read three VALU instructions after it is written, which covers
// the wait states gfx950 requires between a VALU mask write and a VALU carry
// read, so the block needs no s_nop.  Seven VALU instructions per cell.
__device__ __forceinline__ float train_cell(float m, float dg, float up, float x, float y,
                                            uint32_t &pd, uint32_t &pu) {
    float c = m, d;
    uint64_t sa, sb;
    asm("v_cmp_eq_f32_e64 %[sa], %[c], %[dg]\n\t"
        "v_cmp_eq_f32_e64 %[sb], %[c], %[up]\n\t"
        "v_sub_f32_e32 %[d], %[x], %[y]\n\t"
        "v_fmac_f32_e32 %[c], %[d], %[d]\n\t"
        "v_addc_co_u32_e64 %[pd], %[sa], %[pd], %[pd], %[sa]\n\t"
        "v_addc_co_u32_e64 %[pu], %[sb], %[pu], %[pu], %[sb]"
        : [c] "+v"(c), [pd] "+v"(pd), [pu] "+v"(pu), [d] "=&v"(d), [sa] "=&s"(sa),
          [sb] "=&s"(sb)
        : [dg] "v"(dg), [up] "v"(up), [x] "v"(x), [y] "v"(y));
    return c;
}

template <int BP>
__device__ __forceinline__ void train_row(float (&row)[BP], const float *yw, float xi,
                                          const float (&pen)[8], uint32_t *pd, uint32_t *pu) {
    float left = WBX_INF_F;
#pragma unroll
    for (int k = 0; k < BP; ++k) {
        const float up = (k + 1 < BP) ? row[k + 1] : WBX_INF_F;
        const float dg = row[k];
        const float m = fminf(fminf(up, left), dg);
        const float c = train_cell(m, dg, up, xi, yw[k], pd[k >> 5], pu[k >> 5]);
        row[k] = c;
        left = c;
    }
#pragma unroll
    for (int q = 1; q < 8; ++q) row[BP - 8 + q] = fmaxf(row[BP - 8 + q], pen[q]);
}

// Plane words of wave v, row i, lane l: 2 NW words starting at
// dirs[((v * lx + i) * 64 + l) * 2 NW], as (diag word w, up word w) pairs, so a
// lane's row is one contiguous piece and a wave's row is 64 of them in a row.
__device__ __forceinline__ int64_t plane_base(int64_t wave, int lx, int i, int nw2, int lane) {
    return (((int64_t)wave * lx + i) * 64 + lane) * nw2;
}

template <int BP>
__device__ __forceinline__ void store_planes(uint32_t *__restrict__ dirs, int64_t wave, int lane,
                                             int lx, int i, uint32_t (&pd)[(BP + 31) / 32],
                                             uint32_t (&pu)[(BP + 31) / 32]) {
    constexpr int NW = (BP + 31) / 32;
    uint2 *d = (uint2 *)(dirs + plane_base(wave, lx, i, 2 * NW, lane));
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        d[w] = make_uint2(pd[w], pu[w]);
        pd[w] = 0u;
        pu[w] = 0u;
    }
}

template <int BP, int U>
__global__ __launch_bounds__(64) void k_lane_train(PairSpec ps, Series<float> xs, Series<float> ys,
                                                   LaneGeo g, float *__restrict__ sq,
                                                   uint32_t *__restrict__ dirs) {
    constexpr int W = BP + U - 1;
    constexpr int NW = (BP + 31) / 32;
    const int lane = threadIdx.x;
    const int64_t wave = blockIdx.x;
    const int64_t p0 = wave * 64 + lane;
    const int64_t p = p0 < ps.n_pairs ? p0 : ps.n_pairs - 1;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const float *__restrict__ X = xs.row(a);
    const float *__restrict__ Y = ys.row(b);
    const int lx = g.lx, ly = g.ly, off = g.off, B = g.B;

    float row[BP];
#pragma unroll
    for (int k = 0; k < BP; ++k) row[k] = (k == off) ? 0.0f : WBX_INF_F;
    float pen[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) pen[q] = (BP - 8 + q == B) ? WBX_INF_F : -WBX_INF_F;
    uint32_t pd[NW], pu[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) pd[w] = pu[w] = 0u;
    float yw[W];
#pragma unroll
    for (int m = 0; m < W; ++m) yw[m] = ldy(Y, m - off, ly);

    const int rem = lx % U;
    for (int i = 0; i < rem; ++i) {
        train_row<BP>(row, yw, X[i], pen, pd, pu);
        store_planes<BP>(dirs, wave, lane, lx, i, pd, pu);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldy(Y, i + 1 - off + W - 1, ly);
    }
    float xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = X[rem + u < lx ? rem + u : lx - 1];
    for (int i0 = rem; i0 < lx; i0 += U) {
        float xn[U], yn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int inext = i0 + U + u;
            xn[u] = X[inext < lx ? inext : lx - 1];
            yn[u] = ldy(Y, i0 + U - off + W - U + u, ly);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            train_row<BP>(row, yw + u, xv[u], pen, pd, pu);
            store_planes<BP>(dirs, wave, lane, lx, i0 + u, pd, pu);
        }
#pragma unroll
        for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            yw[W - U + u] = yn[u];
            xv[u] = xn[u];
        }
    }
    float res = WBX_INF_F;
#pragma unroll
    for (int k = 0; k < BP; ++k)
        if (k == g.kf) res = row[k];
    if (p0 < ps.n_pairs) sq[p0] = res;
}

// Sum over the 64 lanes in a fixed order: a running sum within each row of 16
// lanes (DPP row_shr 1, 2, 4, 8), then rows folded in with row_bcast 15 and
// 31, which leaves the total in lane 63.  Every lane returns that total.
#define WBX_DPP(v, ctrl, rmask) \
    __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, rmask, 0xF, false))
__device__ __forceinline__ float wave_sum(float v) {
    v += WBX_DPP(v, 0x111, 0xF);  // row_shr:1
    v += WBX_DPP(v, 0x112, 0xF);  // row_shr:2
    v += WBX_DPP(v, 0x114, 0xF);  // row_shr:4
    v += WBX_DPP(v, 0x118, 0xF);  // row_shr:8
    v += WBX_DPP(v, 0x142, 0xA);  // row_bcast:15 into rows 1, 3
    v += WBX_DPP(v, 0x143, 0xC);  // row_bcast:31 into rows 2, 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
#undef WBX_DPP

// Where the per-pair gradients go.  For column j of pair (a, b) the walk keeps
// u = coef * (number of path cells in column j) and v = coef * (sum of x_i
// over those cells) and writes the one term dD/dy_j = y_j u - v when the
// column's 8-column window is complete; the window's 8 y values are fetched
// one window ahead from YP, so the walk never waits on them.
struct GradOut {
    int mode;          // 0: GRID pairs, 1: DIAG (paired) pairs
    int xred;          // GRID with ny % 64 == 0: a wave's lanes share one x row
    int64_t nx, ny, nbc;  // nbc = ny / 64 when xred
    float *dX;         // DIAG, and GRID xred with nbc == 1: written directly
    int64_t ldx;
    float *xpart;      // GRID xred, nbc > 1: [a][bchunk][i]
    float *gxp;        // GRID without lane sharing: per-pair rows [wave][i][lane]
    float *gw;         // per-pair column terms, pair-major rows of lyp floats
    int lyp;           // ly rounded up to a multiple of 8 (rows stay 32-byte aligned)
    const float *YP;   // y_b in rows of lyp floats, zero past ly
    // X transposed (XT[i * nx + a]) and prefix sums of Y transposed
    // (PT[j * ny + b] = y_b[0] + ... + y_b[j-1], j in [0, ly], float64): the
    // lanes of a wave hold neighbouring pairs, so their loads coalesce
    const float *XT;
    const double *PT;
};

// PT[j * ny + b] = sum of y_b[0 .. j) in float64; one wave per series, 64
// columns at a time with a shuffle scan (the sum order is fixed: the result
// is deterministic).
// YP[b * lyp + j] = y_b[j], zero for ly <= j < lyp.
__global__ __launch_bounds__(64) void k_prefix_t(const float *__restrict__ Y, int64_t ny, int ly,
                                                 int64_t stride, double *__restrict__ PT,
                                                 float *__restrict__ YP, int lyp) {
    const int64_t b = blockIdx.x;
    const int lane = threadIdx.x;
    const float *y = Y + b * stride;
    if (lane == 0) PT[b] = 0.0;
    for (int j = lane; j < lyp; j += 64) YP[b * lyp + j] = j < ly ? y[j] : 0.0f;
    double carry = 0.0;
    for (int j0 = 0; j0 < ly; j0 += 64) {
        const int j = j0 + lane;
        double s = j < ly ? (double)y[j] : 0.0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(s, d);
            if (lane >= d) s += o;
        }
        s += carry;
        if (j < ly) PT[(int64_t)(j + 1) * ny + b] = s;
        carry = __shfl(s, 63);
    }
}

// dst[c * n + r] = src[r * stride + c] through a 32 x 32 LDS tile.
__global__ __launch_bounds__(256) void k_transpose(const float *__restrict__ src, int64_t n,
                                                   int len, int64_t stride,
                                                   float *__restrict__ dst) {
    __shared__ float tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty in [0, 8)
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
        const int64_t r = r0 + ty + q, c = c0 + tx;
        if (r < n && c < len) tile[ty + q][tx] = src[r * stride + c];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
        const int64_t c = c0 + ty + q, r = r0 + tx;
        if (r < n && c < len) dst[c * n + r] = tile[tx][ty + q];
    }
}

// Bit of slot k in its plane word: slots are shifted in from the low end, so
// slot k of word w sits at bit (count_w - 1 - (k - 32 w)).
template <int BP>
__global__ __launch_bounds__(64) void k_lane_backtrack(PairSpec ps, LaneGeo g,
                                                       const float *__restrict__ sq,
                                                       const double *__restrict__ G,
                                                       const uint32_t *__restrict__ dirs,
                                                       GradOut go) {
    constexpr int NW = (BP + 31) / 32;
    constexpr int LASTC = BP - 32 * (NW - 1);
    const int lane = threadIdx.x;
    const int64_t wave = blockIdx.x;
    const int64_t p0 = wave * 64 + lane;
    const bool active = p0 < ps.n_pairs;
    const int64_t p = active ? p0 : ps.n_pairs - 1;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const float *__restrict__ XT = go.XT + a;   // x_i = XT[i * nx]
    const double *__restrict__ PT = go.PT + b;  // prefix of y_b up to j: PT[j * ny]
    const int64_t nxs = go.nx, nys = go.ny;
    const int lx = g.lx, ly = g.ly, off = g.off, B = g.B;
    const double D = sqrt((double)sq[p]);
    const double coefd = (active && D > 0.0) ? G[p] / D : 0.0;
    const float coef = (float)coefd;
    // Columns are finished in decreasing order, each exactly once: their
    // terms go to the lane's CW-column window in LDS (slot col % CW) and out
    // as one vector store per CW columns when column col % CW == 0 completes
    // it (columns >= ly of the top window stay the zeros written first).  A
    // column costs two LDS writes, not a shift of two register windows, in a
    // loop every lane of the wave runs as often as its busiest lane.
    constexpr int CW = 8;
    typedef float fcw __attribute__((ext_vector_type(CW)));
    __shared__ float wbuf[2][CW][64];  // u, v of column slot e, lane
#pragma unroll
    for (int e = 0; e < CW; ++e) wbuf[0][e][lane] = wbuf[1][e][lane] = 0.0f;
    fcw *__restrict__ gww = (fcw *)(go.gw + p * go.lyp);
    const fcw *__restrict__ ypw = (const fcw *)(go.YP + b * go.lyp);
    fcw yv_ = ypw[(ly - 1) / CW];  // y of the first window to complete
    auto finish = [&](int col, float u, float v) {
        const int e = col & (CW - 1);
        wbuf[0][e][lane] = u;
        wbuf[1][e][lane] = v;
        if (e == 0) {  // columns col .. col + CW - 1 are complete
            fcw wu, wv;
#pragma unroll
            for (int q = 0; q < CW; ++q) {
                wu[q] = wbuf[0][q][lane];
                wv[q] = wbuf[1][q][lane];
            }
            if (active) gww[col / CW] = wu * yv_ - wv;
            yv_ = ypw[col >= CW ? col / CW - 1 : 0];  // the next window down
        }
    };

    // Rows are taken RB at a time: first the path through the RB rows (bits
    // only, already in registers) while every x and prefix load those rows need
    // is issued, then the sums; the next RB rows' bits are fetched in between.
    // RB is as many rows as keep the kernel within 128 registers (four waves
    // per SIMD): more rows per batch are fewer memory round trips per pair.
    constexpr int RB = BP <= 32 ? 7 : BP <= 64 ? 6 : BP <= 96 ? 5 : 4;
    uint32_t wd[RB][NW], wu[RB][NW];
    auto load_bits = [&](int ih) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = ih - r < 0 ? 0 : ih - r;
            const uint2 *src = (const uint2 *)(dirs + plane_base(wave, lx, i, 2 * NW, lane));
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint2 v = src[w];
                wd[r][w] = v.x;
                wu[r][w] = v.y;
            }
        }
    };
    load_bits(lx - 1);
    int k = g.kf;
    int cj = ly - 1;   // column being accumulated
    float cnt = 0.0f;  // its number of path cells
    float sxc = 0.0f;  // and the sum of their x
    for (int ih = lx - 1; ih >= 0; ih -= RB) {
        int jer[RB], jxr[RB];
        float xr[RB];
        double ph[RB], pl[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = ih - r;
            if (i < 0) break;  // uniform
            const int klo = off - i > 0 ? off - i : 0;
            int khi = ly - 1 - i + off;
            if (khi > B - 1) khi = B - 1;
            if (k > khi) k = khi;
            // exit: the highest slot <= k whose decision is diag or up
            const int wk = k >> 5;
            int kx = -1;
            bool isd = false;
#pragma unroll
            for (int w = NW - 1; w >= 0; --w) {
                const int cnt_w = (w == NW - 1) ? LASTC : 32;
                uint32_t msk = wd[r][w] | wu[r][w];
                if (w > wk) msk = 0u;
                if (w == wk) msk &= ~0u << (cnt_w - 1 - (k - 32 * w));
                if (kx < 0 && msk != 0u) {
                    const int q = __builtin_ctz(msk);
                    kx = 32 * w + cnt_w - 1 - q;
                    isd = ((wd[r][w] >> q) & 1u) != 0u;
                }
            }
            // Row 0 always ends at column 0 (only left moves there, dtw.py:404-406);
            // kx < klo is unreachable on a finite path and keeps the walk in bounds.
            if (i == 0 || kx < klo) {
                kx = klo;
                isd = false;
            }
            if (k < kx) k = kx;  // likewise unreachable; keeps je >= jx
            const int je = i + k - off, jx = i + kx - off;  // in [0, ly)
            jer[r] = je;
            jxr[r] = jx;
            xr[r] = XT[(int64_t)i * nxs];
            ph[r] = PT[(int64_t)(je + 1) * nys];
            pl[r] = PT[(int64_t)jx * nys];
            k = isd ? kx : kx + 1;
        }
        if (ih - RB >= 0) load_bits(ih - RB);
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int i = ih - r;
            if (i < 0) break;  // uniform
            const float xi = xr[r];
            const int je = jer[r], jx = jxr[r];
            // The open column cj takes this row's cell when the row enters
            // there (an up step; a diagonal step enters at cj - 1).  Columns
            // cj down to jx + 1 are then finished: cj with its sums, the rest
            // with this row's single cell; jx becomes the open column.
            if (je == cj) {
                cnt += 1.0f;
                sxc += xi;
            }
            const int nfin = cj - jx;
            for (int t = 0; t < nfin; ++t)
                finish(cj - t, t == 0 ? coef * cnt : coef, t == 0 ? coef * sxc : coef * xi);
            if (nfin > 0) {
                cj = jx;
                cnt = 1.0f;
                sxc = xi;
            }
            // dD/dx_i: coef * sum over the row's cells of (x_i - y_j)
            const float gx = (float)(coefd * ((double)(je - jx + 1) * (double)xi - (ph[r] - pl[r])));
            if (go.mode == 1) {
                if (active) go.dX[p * go.ldx + i] = gx;
            } else if (go.xred) {
                const float tot = wave_sum(gx);
                if (lane == 0) {
                    const int64_t bc = (wave * 64 - a * go.ny) / 64;
                    if (go.nbc == 1)
                        go.dX[a * go.ldx + i] = tot;
                    else
                        go.xpart[(a * go.nbc + bc) * lx + i] = tot;
                }
            } else {
                go.gxp[(wave * lx + i) * 64 + lane] = gx;
            }
        }
    }
    finish(cj, coef * cnt, coef * sxc);  // column 0 completes the last window
}

__device__ __forceinline__ int64_t pair_slot(int64_t p, int len, int idx) {
    return ((p >> 6) * len + idx) * 64 + (p & 63);
}

// dX[a][i] = sum_b gxp[pair (a, b)][i], b in order (GRID without lane sharing)
__global__ __launch_bounds__(256) void k_sum_rows(const float *__restrict__ gxp, int64_t nx,
                                                  int64_t ny, int lx, float *__restrict__ dX,
                                                  int64_t ldx) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nx * lx) return;
    const int64_t a = q / lx;
    const int i = (int)(q - a * lx);
    float s = 0.0f;
    for (int64_t b = 0; b < ny; ++b) s += gxp[pair_slot(a * ny + b, lx, i)];
    dX[a * ldx + i] = s;
}

// dX[a][i] = sum over b-chunks c of xpart[a][c][i]
__global__ __launch_bounds__(256) void k_sum_xpart(const float *__restrict__ xpart, int64_t nx,
                                                   int64_t nbc, int lx, float *__restrict__ dX,
                                                   int64_t ldx) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nx * lx) return;
    const int64_t a = q / lx;
    const int i = (int)(q - a * lx);
    float s = 0.0f;
    for (int64_t c = 0; c < nbc; ++c) s += xpart[(a * nbc + c) * lx + i];
    dX[a * ldx + i] = s;
}

// part[c][b][j]: sums over a in chunk c of the column terms of pair (a, b),
// a in order
__global__ __launch_bounds__(256) void k_sum_cols_partial(const float *__restrict__ gw, int64_t nx,
                                                          int64_t ny, int ly, int lyp, int chunk,
                                                          float *__restrict__ part) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = ny * ly;
    const int c = blockIdx.y;
    if (q >= per) return;
    const int64_t b = q / ly;
    const int j = (int)(q - b * ly);
    const int64_t a0 = (int64_t)c * chunk;
    const int64_t a1 = a0 + chunk < nx ? a0 + chunk : nx;
    float s = 0.0f;
#pragma unroll 8
    for (int64_t a = a0; a < a1; ++a) s += gw[(a * ny + b) * lyp + j];
    part[(int64_t)c * per + q] = s;
}

// dY[b][j] = sum of the chunk partials, chunks in order
__global__ __launch_bounds__(256) void k_sum_cols_final(const float *__restrict__ part, int nc,
                                                        int64_t ny, int ly, float *__restrict__ dY,
                                                        int64_t ldy_) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = ny * ly;
    if (q >= per) return;
    float s = 0.0f;
    for (int c = 0; c < nc; ++c) s += part[(int64_t)c * per + q];
    const int64_t b = q / ly;
    const int j = (int)(q - b * ly);
    dY[b * ldy_ + j] = s;
}

// Paired pairs: dY[p][j] is pair p's column term.
__global__ __launch_bounds__(256) void k_dy_paired(const float *__restrict__ gw, int64_t n, int ly,
                                                   int lyp, float *__restrict__ dY, int64_t ldy_) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n * ly) return;
    const int64_t p = q / ly;
    const int j = (int)(q - p * ly);
    const int64_t o = p * lyp + j;
    dY[p * ldy_ + j] = gw[o];
}

constexpr int TRAIN_MAX_BP = 112;
constexpr int COL_CHUNK = 256;  // x series per column partial: 16 partials for 4096

static int train_bp(int B) { return (B + 7) / 8 * 8; }

template <int BP, int U>
static void train_fwd_one(const PairSpec &ps, Series<float> xs, Series<float> ys, LaneGeo g,
                          float *sq, uint32_t *dirs, hipStream_t st) {
    hipLaunchKernelGGL((k_lane_train<BP, U>), dim3((unsigned)((ps.n_pairs + 63) / 64)), dim3(64), 0,
                       st, ps, xs, ys, g, sq, dirs);
}

template <int BP>
static void train_bwd_one(const PairSpec &ps, LaneGeo g, const float *sq, const double *G,
                          const uint32_t *dirs, GradOut go, hipStream_t st) {
    hipLaunchKernelGGL((k_lane_backtrack<BP>), dim3((unsigned)((ps.n_pairs + 63) / 64)), dim3(64), 0,
                       st, ps, g, sq, G, dirs, go);
}

#define TRAIN_CASES(X)                                                                          \
    X(8, 4) X(16, 4) X(24, 4) X(32, 4) X(40, 4) X(48, 4) X(56, 4) X(64, 4) X(72, 4) X(80, 4)    \
        X(88, 4) X(96, 4) X(104, 4) X(112, 4)

struct TrainGeo {
    LaneGeo g;
    int BP, NW;
};

#define RCT(x)               \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

// Which form serves a call: the decision planes of the lane kernels (DTW,
// bands up to TRAIN_MAX_BP slots) or the band checkpoints of train_seg.hip
// (WDTW, wider bands, and every call with mode bit 1 set).
static int train_route(const wbx_metric *m, int mode, int64_t tx, int64_t ty, bool *seg) {
    if (mode < 0 || mode > 3) return fail(WBX_EINVAL, "mode %d (0..3)", mode);
    if (m->kind != WBX_DTW && m->kind != WBX_WDTW)
        return fail(WBX_EUNSUPPORTED, "differentiable DTW supports metrics dtw and wdtw");
    if (m->precision != WBX_FP32)
        return fail(WBX_EUNSUPPORTED, "differentiable DTW runs in FP32");
    if (tx < 1 || ty < 1 || tx > (1 << 24) || ty > (1 << 24))
        return fail(WBX_EINVAL, "series lengths (%lld, %lld)", (long long)tx, (long long)ty);
    const Band bd = make_band((int)tx, (int)ty, m->r);
    // the checkpointed form splits the band over lanes of at least two slots;
    // a one-slot DTW band (the diagonal) takes the planes
    if (bd.B < 2) {
        if (m->kind == WBX_WDTW)
            return fail(WBX_EUNSUPPORTED, "differentiable WDTW needs a band of 2 slots or more");
        *seg = false;
        return WBX_OK;
    }
    *seg = (mode & 2) != 0 || m->kind == WBX_WDTW || bd.B > TRAIN_MAX_BP;
    return WBX_OK;
}

static int train_geo(const wbx_metric *m, int64_t tx, int64_t ty, TrainGeo *out) {
    if (m->kind != WBX_DTW)
        return fail(WBX_EUNSUPPORTED, "differentiable DTW supports metric dtw only");
    if (m->precision != WBX_FP32)
        return fail(WBX_EUNSUPPORTED, "differentiable DTW runs in FP32");
    const Band bd = make_band((int)tx, (int)ty, m->r);
    if (bd.B > TRAIN_MAX_BP)
        return fail(WBX_EUNSUPPORTED, "differentiable DTW: band of %d slots exceeds %d", bd.B,
                    TRAIN_MAX_BP);
    out->g = LaneGeo{bd.lx, bd.ly, bd.off, bd.B, bd.ly - bd.lx + bd.off};
    out->BP = train_bp(bd.B);
    out->NW = (out->BP + 31) / 32;
    return WBX_OK;
}

// Backward scratch (GRID only): per-pair y rows, per-pair x rows when the
// lanes of a wave do not share an x row, x partials per 64-column chunk, and
// the column partials.
struct TrainScratch {
    bool xred;
    int64_t nbc, nc, xt, pt, gw, yp, gx, xpart, part, total;  // byte offsets / sizes
};

static TrainScratch train_scratch(int mode, int64_t nx, int64_t ny, int64_t tx, int64_t ty) {
    TrainScratch t{};
    auto al = [](int64_t v) { return (v + 255) / 256 * 256; };
    const int64_t np = mode == 0 ? nx * ny : nx, waves = (np + 63) / 64;
    int64_t o = 0;
    t.xt = o;
    o += al(nx * tx * 4);
    t.pt = o;
    o += al((ty + 1) * ny * 8);
    const int64_t lyp = (ty + 7) / 8 * 8;
    t.gw = o;
    o += al(np * lyp * 4);
    t.yp = o;
    o += al(ny * lyp * 4);
    t.gx = t.xpart = t.part = o;
    if (mode != 0) {
        t.total = o;
        return t;
    }
    t.xred = ny % 64 == 0;
    t.nbc = t.xred ? ny / 64 : 0;
    t.nc = (nx + COL_CHUNK - 1) / COL_CHUNK;
    t.gx = o;
    if (!t.xred) o += al(waves * tx * 64 * 4);
    t.xpart = o;
    if (t.xred && t.nbc > 1) o += al(nx * t.nbc * tx * 4);
    t.part = o;
    o += al(t.nc * ny * ty * 4);
    t.total = o;
    return t;
}

}  // namespace wbx

using namespace wbx;

extern "C" {

// Sizes of the state kept from forward to backward (decision planes + the
// squared costs) and of the backward's scratch (per-pair gradients).
int wbx_dtw_train_sizes(const wbx_metric *metric, int mode, int64_t nx, int64_t ny, int64_t tx,
                        int64_t ty, int64_t *state_bytes, int64_t *scratch_bytes) {
    if (!metric) return fail(WBX_EINVAL, "metric is NULL");
    bool seg;
    RCT(train_route(metric, mode, tx, ty, &seg));
    if (seg)
        return seg_train_sizes(metric, mode & 1, nx, ny, tx, ty, state_bytes, scratch_bytes);
    TrainGeo tg;
    int rc = train_geo(metric, tx, ty, &tg);
    if (rc) return rc;
    const int64_t np = mode == 0 ? nx * ny : nx;
    const int64_t waves = (np + 63) / 64;
    const int64_t planes = waves * tx * 2 * tg.NW * 64 * 4;
    if (state_bytes) *state_bytes = planes + ((np * 4 + 255) / 256) * 256;
    if (scratch_bytes) *scratch_bytes = train_scratch(mode, nx, ny, tx, ty).total;
    return WBX_OK;
}

// mode 0: all (x, y) pairs of X (nx) and Y (ny), D is (nx, ny);
// mode 1: D[p] = DTW(X[p], Y[p]) for p < nx (ny must equal nx).
int wbx_dev_dtw_train_forward(const wbx_metric *metric, int mode, const float *X, int64_t nx,
                              int64_t x_stride, int64_t tx, const float *Y, int64_t ny,
                              int64_t y_stride, int64_t ty, double *D, void *state, void *stream) {
    if (!metric || !X || !Y || !D || !state) return fail(WBX_EINVAL, "NULL argument");
    if (nx < 1 || ny < 1 || tx < 1 || ty < 1) return fail(WBX_EINVAL, "empty input");
    if ((mode & 1) && nx != ny) return fail(WBX_EINVAL, "paired mode needs nx == ny");
    bool seg;
    RCT(train_route(metric, mode, tx, ty, &seg));
    if (seg)
        return seg_train_forward(metric, mode & 1, X, nx, x_stride, tx, Y, ny, y_stride, ty, D,
                                 state, (hipStream_t)stream);
    TrainGeo tg;
    int rc = train_geo(metric, tx, ty, &tg);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    PairSpec ps{mode == 0 ? PAIR_GRID : PAIR_DIAG, mode == 0 ? nx * ny : nx, mode == 0 ? ny : nx,
                nullptr, nullptr};
    const int64_t waves = (ps.n_pairs + 63) / 64;
    uint32_t *dirs = (uint32_t *)state;
    float *sq = (float *)((char *)state + waves * tx * 2 * tg.NW * 64 * 4);
    Series<float> xs{X, x_stride, (int)tx, nullptr}, ys{Y, y_stride, (int)ty, nullptr};
    switch (tg.BP) {
#define FWD_CASE(BPV, UV) \
    case BPV: train_fwd_one<BPV, UV>(ps, xs, ys, tg.g, sq, dirs, st); break;
        TRAIN_CASES(FWD_CASE)
#undef FWD_CASE
        default: return fail(WBX_EUNSUPPORTED, "band %d", tg.g.B);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return launch_finish<float>(ps, sq, D, ny, st);
}

// G: upstream gradient dL/dD, same layout as D.  dX (nx, tx) and dY (ny, ty)
// float32 with row strides ldx / ldy are overwritten.
int wbx_dev_dtw_train_backward(const wbx_metric *metric, int mode, const float *X, int64_t nx,
                               int64_t x_stride, int64_t tx, const float *Y, int64_t ny,
                               int64_t y_stride, int64_t ty, const double *G, const void *state,
                               float *dX, int64_t ldx, float *dY, int64_t ldy_, void *scratch,
                               void *stream) {
    if (!metric || !X || !Y || !G || !state || !dX || !dY || !scratch)
        return fail(WBX_EINVAL, "NULL argument");
    if (nx < 1 || ny < 1 || tx < 1 || ty < 1) return fail(WBX_EINVAL, "empty input");
    if ((mode & 1) && nx != ny) return fail(WBX_EINVAL, "paired mode needs nx == ny");
    bool seg;
    RCT(train_route(metric, mode, tx, ty, &seg));
    if (seg)
        return seg_train_backward(metric, mode & 1, X, nx, x_stride, tx, Y, ny, y_stride, ty, G,
                                  state, dX, ldx, dY, ldy_, scratch, (hipStream_t)stream);
    TrainGeo tg;
    int rc = train_geo(metric, tx, ty, &tg);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    PairSpec ps{mode == 0 ? PAIR_GRID : PAIR_DIAG, mode == 0 ? nx * ny : nx, mode == 0 ? ny : nx,
                nullptr, nullptr};
    const int64_t waves = (ps.n_pairs + 63) / 64;
    const uint32_t *dirs = (const uint32_t *)state;
    const float *sq = (const float *)((const char *)state + waves * tx * 2 * tg.NW * 64 * 4);
    const TrainScratch ts = train_scratch(mode, nx, ny, tx, ty);
    char *sc = (char *)scratch;
    GradOut go{};
    go.mode = mode;
    go.xred = ts.xred ? 1 : 0;
    go.nx = nx;
    go.ny = ny;
    go.nbc = ts.nbc;
    go.dX = dX;
    go.ldx = ldx;
    go.gxp = (float *)(sc + ts.gx);
    go.xpart = (float *)(sc + ts.xpart);
    go.gw = (float *)(sc + ts.gw);
    go.lyp = (int)((ty + 7) / 8 * 8);
    float *yp = (float *)(sc + ts.yp);
    go.YP = yp;
    float *part = (float *)(sc + ts.part);
    float *xt = (float *)(sc + ts.xt);
    double *pt = (double *)(sc + ts.pt);
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((tx + 31) / 32), (unsigned)((nx + 31) / 32)),
                       dim3(256), 0, st, X, nx, (int)tx, x_stride, xt);
    hipLaunchKernelGGL(k_prefix_t, dim3((unsigned)ny), dim3(64), 0, st, Y, ny, (int)ty, y_stride,
                       pt, yp, go.lyp);
    go.XT = xt;
    go.PT = pt;
    switch (tg.BP) {
#define BWD_CASE(BPV, UV) \
    case BPV: train_bwd_one<BPV>(ps, tg.g, sq, G, dirs, go, st); break;
        TRAIN_CASES(BWD_CASE)
#undef BWD_CASE
        default: return fail(WBX_EUNSUPPORTED, "band %d", tg.g.B);
    }
    WBX_HIP_CHECK(hipGetLastError());
    const unsigned gy_blocks = (unsigned)((ny * ty + 255) / 256);
    if (mode == 0) {
        const unsigned gx_blocks = (unsigned)((nx * tx + 255) / 256);
        if (!ts.xred)
            hipLaunchKernelGGL(k_sum_rows, dim3(gx_blocks), dim3(256), 0, st, go.gxp, nx, ny,
                               (int)tx, dX, ldx);
        else if (ts.nbc > 1)
            hipLaunchKernelGGL(k_sum_xpart, dim3(gx_blocks), dim3(256), 0, st, go.xpart, nx, ts.nbc,
                               (int)tx, dX, ldx);
        hipLaunchKernelGGL(k_sum_cols_partial, dim3(gy_blocks, (unsigned)ts.nc), dim3(256), 0, st,
                           go.gw, nx, ny, (int)ty, go.lyp, COL_CHUNK, part);
        hipLaunchKernelGGL(k_sum_cols_final, dim3(gy_blocks), dim3(256), 0, st, part, (int)ts.nc,
                           ny, (int)ty, dY, ldy_);
    } else {
        hipLaunchKernelGGL(k_dy_paired, dim3(gy_blocks), dim3(256), 0, st, go.gw, ny, (int)ty,
                           go.lyp, dY, ldy_);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // extern "C"
