// train.hip -- differentiable DTW: forward that records backtrack decisions,
// per-pair backtrack, and gradient reductions.
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
//                    [pair][row][plane word] so a backtrack reads whole lines.
//  k_lane_backtrack  one pair per lane: walks the bits from the final cell,
//                    accumulates (x_i - y_j) * g / D per row and per column
//                    and writes the per-pair gradients (every row and column
//                    is on a monotone path, so every entry is written).
//  k_sum_rows/cols   fixed-order sums of the per-pair gradients into dX, dY.
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

// Planes of pair p, row i: dirs[(p * lx + i) * 2 NW + {0..NW-1: diag, NW..: up}],
// so the backtrack of one pair walks consecutive memory.
template <int BP>
__device__ __forceinline__ void store_planes(uint32_t *__restrict__ dirs, int64_t p, int lx, int i,
                                             uint32_t (&pd)[(BP + 31) / 32],
                                             uint32_t (&pu)[(BP + 31) / 32]) {
    constexpr int NW = (BP + 31) / 32;
    uint32_t *base = dirs + (p * lx + i) * (2 * NW);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        base[w] = pd[w];
        base[NW + w] = pu[w];
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
        store_planes<BP>(dirs, p, lx, i, pd, pu);
#pragma unroll
        for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
        yw[W - 1] = ldy(Y, i + 1 - off + W - 1, ly);
    }
    for (int i0 = rem; i0 < lx; i0 += U) {
        float xv[U], yn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u] = X[i0 + u];
            yn[u] = ldy(Y, i0 + U - off + W - U + u, ly);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            train_row<BP>(row, yw + u, xv[u], pen, pd, pu);
            store_planes<BP>(dirs, p, lx, i0 + u, pd, pu);
        }
#pragma unroll
        for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
        for (int u = 0; u < U; ++u) yw[W - U + u] = yn[u];
    }
    float res = WBX_INF_F;
#pragma unroll
    for (int k = 0; k < BP; ++k)
        if (k == g.kf) res = row[k];
    if (p0 < ps.n_pairs) sq[p0] = res;
}

// Bit of slot k in its plane word: slots are shifted in from the low end, so
// slot k of word w sits at bit (count_w - 1 - (k - 32 w)).
template <int BP>
__global__ __launch_bounds__(64) void k_lane_backtrack(PairSpec ps, Series<float> xs,
                                                       Series<float> ys, LaneGeo g,
                                                       const float *__restrict__ sq,
                                                       const double *__restrict__ G,
                                                       const uint32_t *__restrict__ dirs,
                                                       float *__restrict__ gxp,
                                                       float *__restrict__ gyp) {
    constexpr int NW = (BP + 31) / 32;
    constexpr int LASTC = BP - 32 * (NW - 1);
    const int lane = threadIdx.x;
    const int64_t wave = blockIdx.x;
    const int64_t p = wave * 64 + lane;
    if (p >= ps.n_pairs) return;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const float *__restrict__ X = xs.row(a);
    const float *__restrict__ Y = ys.row(b);
    const int lx = g.lx, ly = g.ly;
    const double D = sqrt((double)sq[p]);
    const float coef = D > 0.0 ? (float)(G[p] / D) : 0.0f;
    float *gx = gxp + p * lx;
    float *gy = gyp + p * ly;
    int i = lx - 1, j = ly - 1, k = g.kf;
    int ci = i, cj = j;
    float sx = 0.0f, sy = 0.0f;
    const uint32_t *dbase = dirs + p * (int64_t)lx * (2 * NW);
    while (true) {
        const float t = coef * (X[i] - Y[j]);
        if (i != ci) {
            gx[ci] = sx;
            ci = i;
            sx = 0.0f;
        }
        if (j != cj) {
            gy[cj] = sy;
            cj = j;
            sy = 0.0f;
        }
        sx += t;
        sy -= t;
        if (i == 0 && j == 0) break;
        if (i == 0) {  // only left is finite (dtw.py:404-406)
            --j;
            --k;
            continue;
        }
        if (j == 0) {  // only up is finite
            --i;
            ++k;
            continue;
        }
        const int w = k >> 5;
        const int cnt = (w == NW - 1) ? LASTC : 32;
        const int bit = cnt - 1 - (k - 32 * w);
        const uint32_t wd = dbase[(int64_t)i * (2 * NW) + w];
        const uint32_t wu = dbase[(int64_t)i * (2 * NW) + NW + w];
        if ((wd >> bit) & 1u) {  // diag
            --i;
            --j;
        } else if ((wu >> bit) & 1u) {  // up: (i-1, j) is slot k+1 of row i-1
            --i;
            ++k;
        } else {  // left: (i, j-1) is slot k-1
            --j;
            --k;
        }
    }
    gx[ci] = sx;
    gy[cj] = sy;
}

// dX[a][i] = sum_b gxp[a * ny + b][i]   (GRID: rows of one a are contiguous)
__global__ __launch_bounds__(256) void k_sum_rows(const float *__restrict__ gxp, int64_t nx,
                                                  int64_t ny, int lx, float *__restrict__ dX,
                                                  int64_t ldx) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nx * lx) return;
    const int64_t a = q / lx;
    const int i = (int)(q - a * lx);
    float s = 0.0f;
    for (int64_t b = 0; b < ny; ++b) s += gxp[(a * ny + b) * lx + i];
    dX[a * ldx + i] = s;
}

// partial[c][b][j] = sum over a in chunk c of gyp[a * ny + b][j]
__global__ __launch_bounds__(256) void k_sum_cols_partial(const float *__restrict__ gyp,
                                                          int64_t nx, int64_t ny, int ly,
                                                          int chunk, float *__restrict__ part) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = ny * ly;
    const int c = blockIdx.y;
    if (q >= per) return;
    const int64_t b = q / ly;
    const int j = (int)(q - b * ly);
    const int64_t a0 = (int64_t)c * chunk;
    const int64_t a1 = a0 + chunk < nx ? a0 + chunk : nx;
    float s = 0.0f;
    for (int64_t a = a0; a < a1; ++a) s += gyp[(a * ny + b) * ly + j];
    part[c * per + q] = s;
}

__global__ __launch_bounds__(256) void k_sum_cols_final(const float *__restrict__ part, int nc,
                                                        int64_t ny, int ly, float *__restrict__ dY,
                                                        int64_t ldy_) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = ny * ly;
    if (q >= per) return;
    float s = 0.0f;
    for (int c = 0; c < nc; ++c) s += part[c * per + q];
    const int64_t b = q / ly;
    dY[b * ldy_ + (q - b * ly)] = s;
}

__global__ __launch_bounds__(256) void k_copy_rows(const float *__restrict__ src, int64_t n,
                                                   int len, float *__restrict__ dst, int64_t ld) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n * len) return;
    const int64_t a = q / len;
    dst[a * ld + (q - a * len)] = src[q];
}

constexpr int TRAIN_MAX_BP = 112;
constexpr int COL_CHUNK = 64;

static int train_bp(int B) { return (B + 7) / 8 * 8; }

template <int BP, int U>
static void train_fwd_one(const PairSpec &ps, Series<float> xs, Series<float> ys, LaneGeo g,
                          float *sq, uint32_t *dirs, hipStream_t st) {
    hipLaunchKernelGGL((k_lane_train<BP, U>), dim3((unsigned)((ps.n_pairs + 63) / 64)), dim3(64), 0,
                       st, ps, xs, ys, g, sq, dirs);
}

template <int BP>
static void train_bwd_one(const PairSpec &ps, Series<float> xs, Series<float> ys, LaneGeo g,
                          const float *sq, const double *G, const uint32_t *dirs, float *gxp,
                          float *gyp, hipStream_t st) {
    hipLaunchKernelGGL((k_lane_backtrack<BP>), dim3((unsigned)((ps.n_pairs + 63) / 64)), dim3(64), 0,
                       st, ps, xs, ys, g, sq, G, dirs, gxp, gyp);
}

#define TRAIN_CASES(X)                                                                          \
    X(8, 8) X(16, 8) X(24, 8) X(32, 8) X(40, 8) X(48, 8) X(56, 8) X(64, 8) X(72, 4) X(80, 4)    \
        X(88, 4) X(96, 4) X(104, 4) X(112, 4)

struct TrainGeo {
    LaneGeo g;
    int BP, NW;
};

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

}  // namespace wbx

using namespace wbx;

extern "C" {

// Sizes of the state kept from forward to backward (decision planes + the
// squared costs) and of the backward's scratch (per-pair gradients).
int wbx_dtw_train_sizes(const wbx_metric *metric, int mode, int64_t nx, int64_t ny, int64_t tx,
                        int64_t ty, int64_t *state_bytes, int64_t *scratch_bytes) {
    if (!metric) return fail(WBX_EINVAL, "metric is NULL");
    TrainGeo tg;
    int rc = train_geo(metric, tx, ty, &tg);
    if (rc) return rc;
    const int64_t np = mode == 0 ? nx * ny : nx;
    const int64_t waves = (np + 63) / 64;
    const int64_t planes = waves * tx * 2 * tg.NW * 64 * 4;
    if (state_bytes) *state_bytes = planes + ((np * 4 + 255) / 256) * 256;
    const int64_t chunks = mode == 0 ? (nx + COL_CHUNK - 1) / COL_CHUNK : 0;
    if (scratch_bytes) *scratch_bytes = np * (tx + ty) * 4 + chunks * ny * ty * 4 + 1024;
    return WBX_OK;
}

// mode 0: all (x, y) pairs of X (nx) and Y (ny), D is (nx, ny);
// mode 1: D[p] = DTW(X[p], Y[p]) for p < nx (ny must equal nx).
int wbx_dev_dtw_train_forward(const wbx_metric *metric, int mode, const float *X, int64_t nx,
                              int64_t x_stride, int64_t tx, const float *Y, int64_t ny,
                              int64_t y_stride, int64_t ty, double *D, void *state, void *stream) {
    if (!metric || !X || !Y || !D || !state) return fail(WBX_EINVAL, "NULL argument");
    if (nx < 1 || ny < 1 || tx < 1 || ty < 1) return fail(WBX_EINVAL, "empty input");
    if (mode == 1 && nx != ny) return fail(WBX_EINVAL, "paired mode needs nx == ny");
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
    if (mode == 1 && nx != ny) return fail(WBX_EINVAL, "paired mode needs nx == ny");
    TrainGeo tg;
    int rc = train_geo(metric, tx, ty, &tg);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    PairSpec ps{mode == 0 ? PAIR_GRID : PAIR_DIAG, mode == 0 ? nx * ny : nx, mode == 0 ? ny : nx,
                nullptr, nullptr};
    const int64_t np = ps.n_pairs;
    const int64_t waves = (np + 63) / 64;
    const uint32_t *dirs = (const uint32_t *)state;
    const float *sq = (const float *)((const char *)state + waves * tx * 2 * tg.NW * 64 * 4);
    float *gxp = (float *)scratch;
    float *gyp = gxp + np * tx;
    float *part = gyp + np * ty;
    Series<float> xs{X, x_stride, (int)tx, nullptr}, ys{Y, y_stride, (int)ty, nullptr};
    switch (tg.BP) {
#define BWD_CASE(BPV, UV) \
    case BPV: train_bwd_one<BPV>(ps, xs, ys, tg.g, sq, G, dirs, gxp, gyp, st); break;
        TRAIN_CASES(BWD_CASE)
#undef BWD_CASE
        default: return fail(WBX_EUNSUPPORTED, "band %d", tg.g.B);
    }
    WBX_HIP_CHECK(hipGetLastError());
    if (mode == 0) {
        hipLaunchKernelGGL(k_sum_rows, dim3((unsigned)((nx * tx + 255) / 256)), dim3(256), 0, st, gxp,
                           nx, ny, (int)tx, dX, ldx);
        const int nc = (int)((nx + COL_CHUNK - 1) / COL_CHUNK);
        hipLaunchKernelGGL(k_sum_cols_partial, dim3((unsigned)((ny * ty + 255) / 256), (unsigned)nc),
                           dim3(256), 0, st, gyp, nx, ny, (int)ty, COL_CHUNK, part);
        hipLaunchKernelGGL(k_sum_cols_final, dim3((unsigned)((ny * ty + 255) / 256)), dim3(256), 0,
                           st, part, nc, ny, (int)ty, dY, ldy_);
    } else {
        hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((nx * tx + 255) / 256)), dim3(256), 0, st,
                           gxp, nx, (int)tx, dX, ldx);
        hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((ny * ty + 255) / 256)), dim3(256), 0, st,
                           gyp, ny, (int)ty, dY, ldy_);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // extern "C"
