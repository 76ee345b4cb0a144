// subseq_mfma.hip -- euclidean subsequence scan on the matrix cores.
//
// Semantics of pairwise_subsequence_distance(metric="euclidean")
// (distance/_cdistance.pyx:1001-1015 over euclidean_distance,
// distance/_metric.pyx:654-683): for series x and shapelet s of length L the
// first window w that minimises F_w = sum_j (x[w + j] - s[j])^2, the sum taken
// in float64 in the order j = 0 .. L - 1, and sqrt(F_w).
//
// Scan.  C_w = X2_w - 2 D_w + S2 with X2_w = sum_j x[w + j]^2 (a difference of
// float64 prefix sums of squares), S2 = sum_j s[j]^2 and the sliding dot
// products D_w = sum_j x[w + j] s[j].  Shapelets are sorted by length and cut
// into chunks of 32 (zero-padded to the chunk's length K, a multiple of 8);
// one 64-lane wave takes one (series, chunk) and forms D for 32 windows x 32
// shapelets at a time as the product of the Hankel tile H[i][j] = x[w0 + i + j]
// with the chunk, K / 2 v_mfma_f32_32x32x2f32 steps.  The K index is split in
// halves: lanes 32 .. 63 carry j = K / 2 + t where lanes 0 .. 31 carry j = t,
// so each lane walks consecutive elements (16-byte loads) for both operands.
//
// Certification.  Every window gets float32 bounds lb <= C <= ub with
//   |C~ - C| <= kappa (X2 + S2) + A_row + A_abs,
// kappa = gamma_{K+2}(2^-23) + 7 u (+ 4.1 u for inputs rounded from float64),
// u = 2^-24: the float32 dot products accumulate at most K / 2 + 3 roundings
// per product at unit 2^-23 (round-to-nearest or towards zero), with
// 2 |D| <= X2 + S2; 7 u covers the float32 roundings of X2, S2, their sum, C~
// and of the bounds themselves; A_row the float64 prefix sums (3 gamma P_T);
// A_abs flushed subnormals (values are at most 2^40 in magnitude, else the
// task goes to the exact rescan).  Rounding the inputs to float32 moves C by
// at most 2 u (2 + u) (X2 + S2) (1 + u)^2.  With U = min_w ub_w the window
// that holds the float64 minimum (and every window tied with it) satisfies
// lb_w <= U (1 + 2^-22): the reference's float64 sum is within
// gamma_{L+2}(2^-53) of C.  Windows are tested against the running U as the
// tiles arrive; a lane's list (kListCap entries in LDS) drops entries above
// the current threshold when it fills, and a task whose list is still full is
// rescanned exactly.
//
// Finish.  Each lane evaluates its surviving windows with the reference's
// float64 arithmetic (v = x; v -= s; d += v * v), the lane pair of a shapelet
// keeps the smallest d and, among equal d, the smallest window, and writes
// sqrt(d) and the window.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <vector>

#include "subseq_fast.h"
#include "subseq_mfma.h"

namespace wbx {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float f4a __attribute__((ext_vector_type(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double d4u __attribute__((ext_vector_type(4), aligned(8)));

constexpr int kListCap = 8;   // candidate list entries per lane
constexpr int kGroup = 8;     // tiles of 32 windows staged at a time
constexpr float kThr = 1.0f + 0x1p-22f;
constexpr float kRangeMax = 0x1p40f;  // larger magnitudes: exact rescan

// ------------------------------------------------------------------ rows
// One 64-lane wave per row of X: the zero-padded float32 copy (Tp columns),
// the float64 prefix sums of squares P[0 .. Tp] and the row's absolute error
// term (3 gamma_{Tp + 72}(2^-53) P[Tp], doubled), +inf when a value exceeds
// 2^40 in magnitude.
__global__ __launch_bounds__(64) void k_mfma_rows(const float *__restrict__ x32, int64_t xs, int T,
                                                  int Tp, float *__restrict__ Xp, int64_t xps,
                                                  double *__restrict__ P, int64_t ps,
                                                  float *__restrict__ arow) {
    const int64_t row = blockIdx.x;
    const int lane = threadIdx.x;
    const float *src = x32 + row * xs;
    float *dst = Xp + row * xps;
    double *pr = P + row * ps;
    if (lane == 0) pr[0] = 0.0;
    double carry = 0.0;
    float mx = 0.0f;
    for (int b = 0; b < Tp; b += 64) {
        const int k = b + lane;
        const float v = k < T ? src[k] : 0.0f;
        dst[k] = v;
        mx = fmaxf(mx, fabsf(v));
        double s = (double)v * (double)v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double y = __shfl_up(s, o);
            if (lane >= o) s += y;
        }
        pr[k + 1] = carry + s;
        carry += __shfl(s, 63);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) {
        const double g = (double)(Tp + 72) * 0x1p-53 * 1.01;
        arow[row] = mx <= kRangeMax ? (float)(6.0 * g * carry * (1.0 + 0x1p-20)) : WBX_INF_F;
    }
}

// Chunk rows: Sp[base + c K + j] = s[j] (j < L) else 0; per slot the float32
// sum of squares (+inf when a value exceeds 2^40 in magnitude).  One block per
// chunk.
__global__ __launch_bounds__(256) void k_mfma_pack(const float *__restrict__ S32,
                                                   const int64_t *__restrict__ soff,
                                                   const int32_t *__restrict__ slist, int n_list,
                                                   const MfmaChunk *__restrict__ chunks,
                                                   const int32_t *__restrict__ slen,
                                                   float *__restrict__ Sp, float *__restrict__ S2) {
    const MfmaChunk ch = chunks[blockIdx.x];
    const int pos0 = (int)blockIdx.x * 32;
    for (int q = threadIdx.x; q < 32 * ch.K; q += 256) {
        const int c = q / ch.K, j = q - c * ch.K;
        const int pos = pos0 + c;
        float v = 0.0f;
        if (pos < n_list && j < slen[pos]) v = S32[soff[slist[pos]] + j];
        Sp[ch.base + q] = v;
    }
    if (threadIdx.x < 32) {
        const int pos = pos0 + threadIdx.x;
        double s2 = 0.0;
        float mx = 0.0f;
        if (pos < n_list) {
            const float *s = S32 + soff[slist[pos]];
            for (int j = 0; j < slen[pos]; ++j) {
                s2 += (double)s[j] * (double)s[j];
                mx = fmaxf(mx, fabsf(s[j]));
            }
        }
        S2[pos] = mx <= kRangeMax ? (float)s2 : WBX_INF_F;
    }
}

// ------------------------------------------------------------------ scan
template <int KQ>
__global__ __launch_bounds__(64) void k_subseq_mfma(MfmaScan a) {
    constexpr int K = 8 * KQ, KH = K / 2;
    // x and its prefix sums for kGroup tiles at a time are staged in LDS; the
    // next group's values are loaded into registers while this one computes
    constexpr int GW = 32 * kGroup;               // windows per group
    constexpr int NX = (GW + K + 63) / 64;        // staged x: NX x 64 floats
    constexpr int NP = (GW + K + 1 + 63) / 64;    // staged P: NP x 64 doubles
    const int lane = threadIdx.x, col = lane & 31, h = lane >> 5;
    const int series = (int)(blockIdx.x / (unsigned)a.nc);
    const int ci = a.c0 + (int)blockIdx.x - series * a.nc;
    const MfmaChunk ch = a.chunks[ci];
    const int pos = ci * 32 + col;
    const int Ls = a.slen[pos];  // 0: empty slot
    __shared__ float s_x[NX * 64];
    __shared__ double s_p[NP * 64];
    __shared__ int s_w[kListCap * 64];
    __shared__ float s_lb[kListCap * 64];
    __shared__ float s_tile[16 * 64];

    float b[KH];
    {
        const float *sp = a.Sp + ch.base + col * K + h * KH;
#pragma unroll
        for (int t = 0; t < KH; t += 4) {
            const f4a v = *(const f4a *)(sp + t);
            b[t] = v[0];
            b[t + 1] = v[1];
            b[t + 2] = v[2];
            b[t + 3] = v[3];
        }
    }
    const float s2 = a.S2[pos];
    const float *__restrict__ xr = a.Xp + (int64_t)series * a.xps;
    const double *__restrict__ pr = a.P + (int64_t)series * a.ps;
    const int xlast = a.Tp - 1, plast = a.Tp;  // written columns of the padded rows
    const float arow = a.arow[series];
    const float aw = arow + a.aabs;
    const float kappa = a.kappa;
    const int nwin = Ls > 0 ? a.T - Ls + 1 : 0;
    const int ntiles = (a.T - ch.lmin + 1 + 31) / 32;
    const int ngroups = (ntiles + kGroup - 1) / kGroup;
    // out of range (or empty): no list, the task is rescanned (empty: ignored)
    bool ovf = !(s2 < WBX_INF_F) || !(arow < WBX_INF_F);
    int cnt = 0;
    float U = WBX_INF_F;

    float gx[NX];
    double gp[NP];
#pragma unroll
    for (int k = 0; k < NX; ++k) gx[k] = xr[min(k * 64 + lane, xlast)];
#pragma unroll
    for (int k = 0; k < NP; ++k) gp[k] = pr[min(k * 64 + lane, plast)];

    for (int gi = 0; gi < ngroups; ++gi) {
        __syncthreads();  // the previous group's reads are done
#pragma unroll
        for (int k = 0; k < NX; ++k) s_x[k * 64 + lane] = gx[k];
#pragma unroll
        for (int k = 0; k < NP; ++k) s_p[k * 64 + lane] = gp[k];
        __syncthreads();
        if (gi + 1 < ngroups) {
            const int g1 = (gi + 1) * GW;
#pragma unroll
            for (int k = 0; k < NX; ++k) gx[k] = xr[min(g1 + k * 64 + lane, xlast)];
#pragma unroll
            for (int k = 0; k < NP; ++k) gp[k] = pr[min(g1 + k * 64 + lane, plast)];
        }
        const int tend = min(kGroup, ntiles - gi * kGroup);
        for (int tl = 0; tl < tend; ++tl) {
            const int w0 = gi * GW + tl * 32;
            v16f acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
            const float *xa = s_x + tl * 32 + col + h * KH;
#pragma unroll
            for (int t = 0; t < KH; t += 4) {
                const f4u v = *(const f4u *)(xa + t);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v[0], b[t], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v[1], b[t + 1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v[2], b[t + 2], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v[3], b[t + 3], acc, 0, 0, 0);
            }
            // accumulator r of this lane: window w0 + 8 (r / 4) + 4 h + r % 4
            const int lim = nwin - 1 - w0;
            float lb[16];
            float mub = WBX_INF_F, mlb = WBX_INF_F;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int i0 = 8 * g + 4 * h;
                const double *p0 = s_p + tl * 32 + i0;
                const double *p1 = p0 + Ls;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float x2 = (float)(p1[q] - p0[q]);
                    const float qq = x2 + s2;
                    const float c = __builtin_fmaf(acc[4 * g + q], -2.0f, qq);
                    const float e = __builtin_fmaf(qq, kappa, aw);
                    const bool ok = i0 + q <= lim;
                    const float ub = ok ? c + e : WBX_INF_F;
                    const float l = ok ? c - e : WBX_INF_F;
                    lb[4 * g + q] = l;
                    mub = fminf(mub, ub);
                    mlb = fminf(mlb, l);
                }
            }
            U = fminf(U, mub);
            U = fminf(U, __shfl_xor(U, 32));
            const float thr = U * kThr;
            if (__ballot(!ovf && mlb <= thr) != 0ull) {
#pragma unroll
                for (int r = 0; r < 16; ++r) s_tile[r * 64 + lane] = lb[r];
                if (!ovf && mlb <= thr) {
                    for (int r = 0; r < 16; ++r) {
                        const float l = s_tile[r * 64 + lane];
                        if (!(l <= thr)) continue;
                        if (cnt == kListCap) {  // drop entries above the threshold
                            int k = 0;
                            for (int e = 0; e < kListCap; ++e) {
                                const float v = s_lb[e * 64 + lane];
                                if (v <= thr) {
                                    s_w[k * 64 + lane] = s_w[e * 64 + lane];
                                    s_lb[k * 64 + lane] = v;
                                    ++k;
                                }
                            }
                            cnt = k;
                            if (cnt == kListCap) {
                                ovf = true;
                                break;
                            }
                        }
                        s_w[cnt * 64 + lane] = w0 + 8 * (r >> 2) + 4 * h + (r & 3);
                        s_lb[cnt * 64 + lane] = l;
                        ++cnt;
                    }
                }
            }
        }
    }

    // float64 finish of the surviving windows
    const float thr = U * kThr;
    double best = WBX_INF_D;
    int bw = 0x7fffffff;
    const int sid = pos < a.n_list ? a.slist[pos] : 0;
    if (!ovf && Ls > 0) {
        const double *__restrict__ x64 = a.X64 + (int64_t)series * a.x64s;
        const double *__restrict__ s64 = a.S64 + a.soff[sid];
        for (int e = 0; e < cnt; ++e) {
            if (!(s_lb[e * 64 + lane] <= thr)) continue;
            const int w = s_w[e * 64 + lane];
            double d = 0.0;
            for (int j = 0; j < Ls; ++j) {
                double v = x64[w + j];
                v -= s64[j];
                d += v * v;
            }
            if (d < best || (d == best && w < bw)) {
                best = d;
                bw = w;
            }
        }
    }
    const double ob = __shfl_xor(best, 32);
    const int ow = __shfl_xor(bw, 32);
    const int oo = __shfl_xor((int)ovf, 32);
    if (ob < best || (ob == best && ow < bw)) {
        best = ob;
        bw = ow;
    }
    if (h == 0 && Ls > 0) {
        if (ovf || oo || !(best < WBX_INF_D)) {
            const int o = atomicAdd(a.n_ovf, 1);
            a.ovf[o] = (int32_t)((int64_t)series * a.n_list + pos);
        } else {
            const int64_t o = (int64_t)series * a.n_shapelets + sid;
            a.dist[o] = sqrt(best);
            a.idx[o] = bw;
        }
    }
}

// ------------------------------------------------------------------ host
#define MF_RC(x)             \
    do {                     \
        const int rc_ = (x); \
        if (rc_) return rc_; \
    } while (0)

void mfma_plan(const std::vector<int32_t> &ids, const std::vector<int64_t> &hl, int64_t T,
               MfmaPlan *p) {
    p->slist = ids;
    std::stable_sort(p->slist.begin(), p->slist.end(),
                     [&](int32_t x, int32_t y) { return hl[x] > hl[y]; });
    const int64_t nl = (int64_t)p->slist.size();
    const int64_t nc = (nl + 31) / 32;
    p->chunks.assign((size_t)nc, MfmaChunk{});
    p->slen.assign((size_t)nc * 32, 0);
    int64_t base = 0;
    for (int64_t c = 0; c < nc; ++c) {
        int lmax = 0, lmin = (int)T;
        for (int64_t q = c * 32; q < std::min<int64_t>(nl, c * 32 + 32); ++q) {
            const int L = (int)hl[p->slist[q]];
            p->slen[q] = L;
            p->max_len = std::max(p->max_len, L);
            lmax = std::max(lmax, L);
            lmin = std::min(lmin, L);
        }
        MfmaChunk &ch = p->chunks[c];
        ch.K = std::max(8, (lmax + 7) / 8 * 8);
        ch.lmin = lmin;
        ch.base = (int32_t)base;
        base += 32 * (int64_t)ch.K;
    }
    p->sp_floats = base;
    p->T = (int)T;
    // reads reach column 32 ntiles - 1 + K - 1 <= T + 61 + (K - lmin) <= T + 195
    p->Tp = (int)((T + 200 + 63) / 64 * 64);
    p->xps = p->Tp;
    p->ps = p->Tp + 8;
}

size_t mfma_ws_bytes(const MfmaPlan &p, int64_t n) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t nslot = p.slen.size();
    return al((size_t)n * p.xps * 4) + al((size_t)n * p.ps * 8) + al((size_t)n * 4) +
           al((size_t)p.sp_floats * 4) + al(nslot * 4) * 3 + al(p.chunks.size() * sizeof(MfmaChunk)) +
           al((size_t)n * p.slist.size() * 4) + al(64);
}

int mfma_scan(const MfmaPlan &p, char *ws, int inputs_exact, const float *x32, int64_t x32s,
              const float *S32, const double *X64, int64_t x64s, const double *S64,
              const int64_t *soff, const int64_t *slen_dev, int64_t n, int64_t n_shapelets,
              double *dist, int64_t *idx, hipStream_t st) {
    const int64_t nl = (int64_t)p.slist.size();
    const int64_t nc = (int64_t)p.chunks.size();
    if (nl == 0 || n == 0) return WBX_OK;
    if (n > 0x7fffffff || n * nc > 0x7fffffff || n * nl > 0x7fffffff || p.sp_floats > 0x7fffffff)
        return fail(WBX_EUNSUPPORTED, "subsequence scan too large");
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t o = 0;
    auto take = [&](size_t b) {
        char *q = ws + o;
        o += al(b);
        return q;
    };
    const size_t nslot = p.slen.size();
    float *Xp = (float *)take((size_t)n * p.xps * 4);
    double *P = (double *)take((size_t)n * p.ps * 8);
    float *arow = (float *)take((size_t)n * 4);
    float *Sp = (float *)take((size_t)p.sp_floats * 4);
    float *S2 = (float *)take(nslot * 4);
    int32_t *slen = (int32_t *)take(nslot * 4);
    int32_t *slist = (int32_t *)take(nslot * 4);
    MfmaChunk *chunks = (MfmaChunk *)take(p.chunks.size() * sizeof(MfmaChunk));
    int32_t *ovf = (int32_t *)take((size_t)n * nl * 4);
    int *n_ovf = (int *)take(64);
    WBX_HIP_CHECK(hipMemcpyAsync(slen, p.slen.data(), nslot * 4, hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(slist, p.slist.data(), nl * 4, hipMemcpyHostToDevice, st));
    WBX_HIP_CHECK(hipMemcpyAsync(chunks, p.chunks.data(), p.chunks.size() * sizeof(MfmaChunk),
                                 hipMemcpyHostToDevice, st));
    MF_RC(launch_fill_i32(n_ovf, 1, 0, st));
    hipLaunchKernelGGL(k_mfma_rows, dim3((unsigned)n), dim3(64), 0, st, x32, x32s, p.T, p.Tp, Xp,
                       (int64_t)p.xps, P, (int64_t)p.ps, arow);
    WBX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_mfma_pack, dim3((unsigned)nc), dim3(256), 0, st, S32, soff, slist,
                       (int)nl, chunks, slen, Sp, S2);
    WBX_HIP_CHECK(hipGetLastError());

    MfmaScan a;
    a.Xp = Xp;
    a.xps = p.xps;
    a.P = P;
    a.ps = p.ps;
    a.arow = arow;
    a.Sp = Sp;
    a.S2 = S2;
    a.slen = slen;
    a.slist = slist;
    a.chunks = chunks;
    a.X64 = X64;
    a.x64s = x64s;
    a.S64 = S64;
    a.soff = soff;
    a.T = p.T;
    a.Tp = p.Tp;
    a.n_list = (int)nl;
    a.n_shapelets = n_shapelets;
    a.dist = dist;
    a.idx = idx;
    a.ovf = ovf;
    a.n_ovf = n_ovf;
    const double u = 0x1p-24;
    // chunks are sorted by K (descending): one launch per run of equal K
    for (int64_t c0 = 0; c0 < nc;) {
        const int K = p.chunks[c0].K;
        int64_t c1 = c0;
        while (c1 < nc && p.chunks[c1].K == K) ++c1;
        const double nK = (double)K + 2.0;
        const double gam = nK * 0x1p-23 / (1.0 - nK * 0x1p-23);
        a.kappa = (float)((gam + 7.0 * u + (inputs_exact ? 0.0 : 4.1 * u)) * (1.0 + 0x1p-20));
        a.aabs = (float)((K + 8) * 0x1p-83);
        a.c0 = (int)c0;
        a.nc = (int)(c1 - c0);
        const dim3 grid((unsigned)(n * (c1 - c0)));
        switch (K / 8) {
#define MF_CASE(Q) \
    case Q: hipLaunchKernelGGL(k_subseq_mfma<Q>, grid, dim3(64), 0, st, a); break;
            MF_CASE(1) MF_CASE(2) MF_CASE(3) MF_CASE(4) MF_CASE(5) MF_CASE(6) MF_CASE(7) MF_CASE(8)
            MF_CASE(9) MF_CASE(10) MF_CASE(11) MF_CASE(12) MF_CASE(13) MF_CASE(14) MF_CASE(15)
            MF_CASE(16)
#undef MF_CASE
            default: return fail(WBX_EUNSUPPORTED, "subsequence chunk length %d", K);
        }
        WBX_HIP_CHECK(hipGetLastError());
        c0 = c1;
    }
    int novf = 0;
    WBX_HIP_CHECK(hipMemcpyAsync(&novf, n_ovf, 4, hipMemcpyDeviceToHost, st));
    WBX_HIP_CHECK(hipStreamSynchronize(st));
    if (novf > 0) {
        TaskMap tm{slist, (int32_t)nl};
        MF_RC(launch_subseq_rescan(tm, ovf, n_ovf, novf, WBX_EUCLIDEAN, S64, soff, slen_dev, X64, x64s,
                                p.T, 0.0, p.max_len, n_shapelets, dist, idx, st));
    }
    return WBX_OK;
}

}  // namespace wbx
