// subseq_fast.hip -- FP32 subsequence scans with a certified exact finish.
//
// Semantics are those of _PairwiseSubsequenceDistance (distance/_cdistance.pyx:
// 1001-1015): for sample i and shapelet s, the minimum over windows of the
// squared cost (dtw_subsequence_distance, distance/_elastic.pyx:622-655, or
// euclidean_distance, distance/_metric.pyx:654-683), the first window that
// attains it, and sqrt of that minimum.
//
// Pass 1 (FP32): one (sample, shapelet) task per 64-lane wave, lane l scans
//   window ws = 64 b + l of block b.  For DTW each lane runs the register-band
//   recursion of lane.hip on its window (all lanes share the shapelet, so the
//   geometry is uniform); for euclidean each lane sums (x - s)^2 over its
//   window.  A window is a candidate when its FP32 cost c satisfies
//   c <= tau(c*) = rho (c* + A) + A, c* the smallest FP32 cost so far: by the
//   bound in rerank.hip no other window can hold the float64 minimum.  The
//   wave keeps up to 64 candidates with their costs in LDS and drops those
//   that fail tau of the current best when the list fills; at the end the
//   ones within tau(final best) are stored (at most kMaxCand, else the task
//   is marked as overflowed).
// Pass 2 (FP64): one candidate per lane, evaluated with the reference's exact
//   arithmetic (DTW: min(min(up, left), diag) + v*v, no FMA; euclidean: the
//   sequential sum), then per task the first strictly smallest candidate wins.
// Overflowed tasks are rescanned exactly (subseq.hip).
#pragma clang fp contract(off)

#include "block_dtw.h"
#include "lane_common.h"
#include "lane64_cost.h"
#include "subseq_fast.h"

namespace wbx {

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

// Wave-level candidate list in LDS: (window, FP32 cost) pairs, up to kCap.
// A window enters when its cost is within tau(best so far); when the list
// would overflow, entries no longer within tau(best now) are dropped (the
// best only decreases, so a dropped entry could never pass the final test).
constexpr int kCap = 64;

struct CandList {
    int count;
    bool overflow;
};

__device__ __forceinline__ void cand_push(CandList &cl, int *s_win, float *s_cost, bool take,
                                          int ws, float c, float tau) {
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t mask = __ballot(take);
    const int npush = (int)__popcll(mask);
    if (npush == 0 || cl.overflow) return;  // uniform
    if (cl.count + npush > kCap) {
        __syncthreads();
        const bool has = lane < cl.count;
        const int w = has ? s_win[lane] : 0;
        const float cc = has ? s_cost[lane] : WBX_INF_F;
        const bool keep = has && cc <= tau;
        const uint64_t km = __ballot(keep);
        __syncthreads();
        if (keep) {
            const int pos = (int)__popcll(km & below);
            s_win[pos] = w;
            s_cost[pos] = cc;
        }
        cl.count = (int)__popcll(km);
        __syncthreads();
        if (cl.count + npush > kCap) {
            cl.overflow = true;
            return;
        }
    }
    if (take) {
        const int pos = cl.count + (int)__popcll(mask & below);
        s_win[pos] = ws;
        s_cost[pos] = c;
    }
    cl.count += npush;
}

// Final list: the entries within tau(final best), at most kMaxCand of them,
// else the task is marked for the exact rescan (count kMaxCand + 1).
__device__ __forceinline__ void cand_finish(const CandList &cl, const int *s_win,
                                            const float *s_cost, float tau, float best,
                                            SubCand *__restrict__ out) {
    const int lane = threadIdx.x;
    __syncthreads();
    const bool has = !cl.overflow && lane < cl.count;
    const int w = has ? s_win[lane] : 0;
    const float cc = has ? s_cost[lane] : WBX_INF_F;
    const bool keep = has && cc <= tau;
    const uint64_t km = __ballot(keep);
    const int n = (int)__popcll(km);
    const bool ovf = cl.overflow || n > kMaxCand;
    if (!ovf && keep) out->win[__popcll(km & ((1ull << lane) - 1ull))] = w;
    if (lane == 0) {
        out->count = ovf ? kMaxCand + 1 : n;
        out->best32 = best;
    }
}

// Waves per SIMD the register allocation is held to, by band class: left
// alone the allocator spreads the window-batch set-up over up to 230
// registers (two or three waves); 13 and 15 fit four waves and 17 .. 23
// three without spills (wider classes would spill).
constexpr int subseq_dtw32_waves(int BP) {
    return BP <= 11 ? 1 : BP <= 15 ? 4 : BP <= 23 ? 3 : 1;
}

template <int BP, int U>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(subseq_dtw32_waves(BP)))) void
k_subseq_dtw32(TaskMap tasks, SubGeo sg, SubCand *__restrict__ cands) {
    static_assert(U % 4 == 0, "groups of whole 16-byte loads");
    constexpr int W = BP + U - 1;
    const int lane = threadIdx.x;
    const SubTask tk = tasks.at(blockIdx.x);
    // the shapelet is the same for every lane and is not written by the
    // launch: read through the constant address space, its values are scalar
    // loads into SGPRs
    const float *__restrict__ S = sg.S + sg.off[tk.shapelet];
    const __attribute__((address_space(4))) float *Sc =
        (const __attribute__((address_space(4))) float *)S;
    const int L = (int)sg.len[tk.shapelet];
    const int off = band_w(L, L, sg.r) - 1;  // the band has 2 w - 1 == BP slots
    const int nwin = sg.T - L + 1;
    const float *__restrict__ Xrow = sg.X + (int64_t)tk.sample * sg.xstride;
    SubCand *cout = cands + blockIdx.x;
    __shared__ int s_win[kCap];
    __shared__ float s_cost[kCap];
    CandList cl{0, false};
    float best = WBX_INF_F;
    for (int b0 = 0; b0 < nwin; b0 += 64) {
        const int ws = b0 + lane;
        const bool active = ws < nwin;
        const float *__restrict__ Y = Xrow + (active ? ws : nwin - 1);
        float row[BP];
#pragma unroll
        for (int k = 0; k < BP; ++k) row[k] = (k == off) ? 0.0f : WBX_INF_F;
        float yw[W];
        if (b0 >= off && W - off <= L) {
            // uniform: every lane's columns -off .. W - off - 1 are inside the
            // series row (ws >= off) and end inside the window; the first off
            // are left of it (+inf)
            const float *__restrict__ Yb = Y - off;
#pragma unroll
            for (int q = 0; q < W; q += 4) {
                if (q + 4 <= W) {
                    const f4a v = *(const f4a *)(Yb + q);
#pragma unroll
                    for (int u = 0; u < 4; ++u) yw[q + u] = v[u];
                } else {
#pragma unroll
                    for (int u = 0; q + u < W; ++u) yw[q + u] = Yb[q + u];
                }
            }
#pragma unroll
            for (int m = 0; m < W; ++m) yw[m] = m < off ? WBX_INF_F : yw[m];
        } else {
#pragma unroll
            for (int m = 0; m < W; ++m) yw[m] = ldy(Y, m - off, L);
        }
        const int rem = L % U;
        for (int i = 0; i < rem; ++i) {
            const float xi = S[i];
            float left = WBX_INF_F;
#pragma unroll
            for (int k = 0; k < BP; ++k) {
                const float up = (k + 1 < BP) ? row[k + 1] : WBX_INF_F;
                const float mn = fminf(fminf(up, left), row[k]);
                const float d = xi - yw[k];
                const float c = __builtin_fmaf(d, d, mn);
                row[k] = c;
                left = c;
            }
#pragma unroll
            for (int m = 0; m < W - 1; ++m) yw[m] = yw[m + 1];
            yw[W - 1] = ldy(Y, i + 1 - off + W - 1, L);
        }
        for (int i0 = rem; i0 < L; i0 += U) {
            // This group's new window entries, columns jn .. jn + U - 1
            // (jn >= 0).  Every lane has the same shapelet and row, so when
            // they lie inside the window (the usual case) they are plain
            // vector loads, waited for only at the window shift.  The
            // shapelet values are scalar loads (the shapelet stays in the
            // scalar cache).
            float yn[U];
            const int jn = i0 - off + W;
            if (jn + U <= L) {
#pragma unroll
                for (int u = 0; u < U; u += 4) {
                    const f4a b = *(const f4a *)(Y + jn + u);
#pragma unroll
                    for (int q = 0; q < 4; ++q) yn[u + q] = b[q];
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) yn[u] = ldy(Y, jn + u, L);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float xi = Sc[i0 + u];
                float left = WBX_INF_F;
#pragma unroll
                for (int k = 0; k < BP; ++k) {
                    const float up = (k + 1 < BP) ? row[k + 1] : WBX_INF_F;
                    const float mn = fminf(fminf(up, left), row[k]);
                    const float d = xi - yw[k + u];
                    const float c = __builtin_fmaf(d, d, mn);
                    row[k] = c;
                    left = c;
                }
            }
#pragma unroll
            for (int m = 0; m < W - U; ++m) yw[m] = yw[m + U];
#pragma unroll
            for (int u = 0; u < U; ++u) yw[W - U + u] = yn[u];
        }
        float c = WBX_INF_F;
#pragma unroll
        for (int k = 0; k < BP; ++k)
            if (k == off) c = row[k];
        if (!active) c = WBX_INF_F;
        best = fminf(best, wave_min(c));
        const float tau = sg.rho * (best + sg.A) + sg.A;
        cand_push(cl, s_win, s_cost, active && c <= tau, ws, c, tau);
    }
    cand_finish(cl, s_win, s_cost, sg.rho * (best + sg.A) + sg.A, best, cout);
}

// Four consecutive floats of a row of length T starting at e (e >= 0): one
// 16-byte load when it lies inside the row, else each index clamped to T - 1
// (those values only feed windows past the last one).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ f4u row4(const float *__restrict__ row, int e, int T) {
    if (e + 3 < T) return *(const f4u *)(row + e);
    f4u v;
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = row[e + q < T ? e + q : T - 1];
    return v;
}

// Lane l owns the U = 4 consecutive windows w0 .. w0 + 3 (w0 = b0 + 4 l) and
// keeps x[w0 + j .. w0 + j + 8) in registers: four steps of j use static
// indices (window q at step u reads xw[q + u]), then the window moves by four
// with one 16-byte load.  Each window's sum runs over j in order.
template <int U>
__global__ __launch_bounds__(64) void k_subseq_euclid32(TaskMap tasks, SubGeo sg,
                                                        SubCand *__restrict__ cands) {
    static_assert(U == 4, "four windows per lane");
    const int lane = threadIdx.x;
    const SubTask tk = tasks.at(blockIdx.x);
    const float *__restrict__ S = sg.S + sg.off[tk.shapelet];
    const int L = (int)sg.len[tk.shapelet];
    const int T = sg.T;
    const int nwin = T - L + 1;
    const float *__restrict__ Xrow = sg.X + (int64_t)tk.sample * sg.xstride;
    SubCand *cout = cands + blockIdx.x;
    __shared__ int s_win[kCap];
    __shared__ float s_cost[kCap];
    CandList cl{0, false};
    float best = WBX_INF_F;
    for (int b0 = 0; b0 < nwin; b0 += 64 * U) {
        const int w0 = b0 + lane * U;
        float acc[U];
#pragma unroll
        for (int q = 0; q < U; ++q) acc[q] = 0.0f;
        float xw[2 * U];
        {
            const f4u lo = row4(Xrow, w0, T), hi = row4(Xrow, w0 + U, T);
#pragma unroll
            for (int q = 0; q < U; ++q) {
                xw[q] = lo[q];
                xw[U + q] = hi[q];
            }
        }
        int j0 = 0;
        for (; j0 + U <= L; j0 += U) {
            const f4u nx = row4(Xrow, w0 + j0 + 2 * U, T);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float sj = S[j0 + u];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const float d = xw[q + u] - sj;
                    acc[q] = __builtin_fmaf(d, d, acc[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                xw[q] = xw[U + q];
                xw[U + q] = nx[q];
            }
        }
        for (int u = 0; j0 + u < L; ++u) {  // the last L % U steps
            const float sj = S[j0 + u];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                float xv = xw[q];
#pragma unroll
                for (int t = 1; t < U; ++t)
                    if (u == t) xv = xw[q + t];
                const float d = xv - sj;
                acc[q] = __builtin_fmaf(d, d, acc[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int ws = w0 + q;
            const bool active = ws < nwin;
            const float c = active ? acc[q] : WBX_INF_F;
            best = fminf(best, wave_min(c));
            const float tau = sg.rho * (best + sg.A) + sg.A;
            cand_push(cl, s_win, s_cost, active && c <= tau, ws, c, tau);
        }
    }
    cand_finish(cl, s_win, s_cost, sg.rho * (best + sg.A) + sg.A, best, cout);
}

// ------------------------------------------------------------------ exact
// Float64 cost of one candidate window per lane; per-lane geometry (all lanes
// of a launch share BP).  Same arithmetic as dtw_distance: rows of S against
// the window, cost = min(min(up, left), diag) + v * v.
template <int BP>
__global__ __launch_bounds__(64) void k_subseq_cand64(TaskMap tasks,
                                                      const SubCand *__restrict__ cands,
                                                      int64_t n_tasks, const double *__restrict__ S64,
                                                      const int64_t *__restrict__ soff,
                                                      const int64_t *__restrict__ slen,
                                                      const double *__restrict__ X64,
                                                      int64_t xstride, int T, double r, int kind,
                                                      double *__restrict__ cost64) {
    const int64_t q = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int64_t t = q / kMaxCand;
    const int slot = (int)(q - t * kMaxCand);
    if (t >= n_tasks) return;
    const SubCand &cd = cands[t];
    if (cd.count > kMaxCand || slot >= cd.count) return;
    const SubTask tk = tasks.at(t);
    const double *S = S64 + soff[tk.shapelet];
    const int L = (int)slen[tk.shapelet];
    const double *Y = X64 + (int64_t)tk.sample * xstride + cd.win[slot];
    double res;
    if (kind == WBX_EUCLIDEAN) {
        double dist = 0.0;
        for (int j = 0; j < L; ++j) {
            double v = Y[j];
            v -= S[j];
            dist += v * v;
        }
        res = dist;
    } else {
        // the band row in registers with a sliding window of the candidate's
        // values (the float64 lane kernels' cell arithmetic, bit for bit);
        // a launch holds one band class, B in (BP - 8, BP), whose padding
        // slot B the routine resets every row
        const int w = band_w(L, L, r);
        res = lane64_cost<BP, 4, false, false, true>(S, L, Y, L, w - 1, 2 * w - 1, w - 1, nullptr,
                                                    L);
    }
    cost64[q] = res;
}

// Per task: first strictly smallest exact cost among the candidates (window
// order = index order; euclidean candidates are sorted here first).
__global__ __launch_bounds__(256) void k_subseq_select(TaskMap tasks,
                                                       const SubCand *__restrict__ cands,
                                                       int64_t n_tasks,
                                                       const double *__restrict__ cost64,
                                                       int64_t n_shapelets,
                                                       double *__restrict__ dist,
                                                       int64_t *__restrict__ idx,
                                                       int32_t *__restrict__ overflow,
                                                       int *__restrict__ n_overflow) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n_tasks) return;
    const SubCand &cd = cands[t];
    const SubTask tk = tasks.at(t);
    if (cd.count > kMaxCand || cd.count == 0) {
        const int o = atomicAdd(n_overflow, 1);
        overflow[o] = (int32_t)t;
        return;
    }
    double best = WBX_INF_D;
    int bw = 0x7fffffff;
    for (int s = 0; s < cd.count; ++s) {
        const double c = cost64[t * kMaxCand + s];
        const int wv = cd.win[s];
        if (c < best || (c == best && wv < bw)) {
            best = c;
            bw = wv;
        }
    }
    const int64_t o = (int64_t)tk.sample * n_shapelets + tk.shapelet;
    dist[o] = sqrt(best);
    idx[o] = bw;
}

// Exact full scan of the tasks whose candidate list overflowed: one task per
// 64-lane block, windows in order (the plain kernel's block recursion).
__global__ __launch_bounds__(64) void k_subseq_rescan(TaskMap tasks, const int32_t *__restrict__ list,
                                                      const int *__restrict__ n_list, int kind,
                                                      const double *__restrict__ S64,
                                                      const int64_t *__restrict__ soff,
                                                      const int64_t *__restrict__ slen,
                                                      const double *__restrict__ X64,
                                                      int64_t xstride, int T, double r, int cap,
                                                      int64_t n_shapelets, double *__restrict__ dist,
                                                      int64_t *__restrict__ idx) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *buf = reinterpret_cast<double *>(smem_raw);
    if ((int)blockIdx.x >= *n_list) return;
    const SubTask tk = tasks.at(list[blockIdx.x]);
    const double *S = S64 + soff[tk.shapelet];
    const int L = (int)slen[tk.shapelet];
    const double *x = X64 + (int64_t)tk.sample * xstride;
    double best = WBX_INF_D;
    int64_t bi = 0;
    if (kind == WBX_EUCLIDEAN) {
        if (threadIdx.x == 0) {
            for (int s = 0; s < T - L + 1; ++s) {
                double d = 0.0;
                for (int k = 0; k < L; ++k) {
                    if (d >= best) break;
                    double v = x[s + k];
                    v -= S[k];
                    d += v * v;
                }
                if (d < best) {
                    best = d;
                    bi = s;
                }
            }
        }
    } else {
        const int w = band_w(L, L, r);
        for (int s = 0; s <= T - L; ++s) {
            const double c = block_dtw<double>(S, L, x + s, L, w, nullptr, buf, buf + cap,
                                               buf + 2 * cap);
            if (c < best) {
                best = c;
                bi = s;
            }
        }
    }
    if (threadIdx.x == 0) {
        const int64_t o = (int64_t)tk.sample * n_shapelets + tk.shapelet;
        dist[o] = sqrt(best);
        idx[o] = bi;
    }
}

int launch_subseq_rescan(TaskMap tasks, const int32_t *list, const int *n_list_dev,
                         int n_list_host, int kind, const double *S64, const int64_t *soff,
                         const int64_t *slen, const double *X64, int64_t xstride, int T, double r,
                         int max_len, int64_t n_shapelets, double *dist, int64_t *idx,
                         hipStream_t st) {
    if (n_list_host <= 0) return WBX_OK;
    const size_t lds = (size_t)3 * (size_t)max_len * sizeof(double);
    if (lds > 64 * 1024) return fail(WBX_EUNSUPPORTED, "shapelet too long for the rescan");
    hipLaunchKernelGGL(k_subseq_rescan, dim3((unsigned)n_list_host), dim3(64), lds, st, tasks, list,
                       n_list_dev, kind, S64, soff, slen, X64, xstride, T, r, max_len, n_shapelets,
                       dist, idx);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

// ------------------------------------------------------------------ launch
// FP32 scans are compiled for every odd band width B = 2 w - 1 <= 63 (equal
// lengths), so no slot beyond the band is computed.
#define SUB_SCAN_CASES(X)                                                                        \
    X(1, 8) X(3, 8) X(5, 8) X(7, 8) X(9, 8) X(11, 8) X(13, 8) X(15, 8) X(17, 8) X(19, 8) X(21, 8) \
        X(23, 8) X(25, 8) X(27, 8) X(29, 8) X(31, 8) X(33, 4) X(35, 4) X(37, 4) X(39, 4)        \
            X(41, 4) X(43, 4) X(45, 4) X(47, 4) X(49, 4) X(51, 4) X(53, 4) X(55, 4) X(57, 4)     \
                X(59, 4) X(61, 4) X(63, 4)
#define SUB_DTW_CASES(X) X(8, 8) X(16, 8) X(24, 8) X(32, 8) X(40, 8) X(48, 8) X(56, 8) X(64, 8)

int launch_subseq_fast(int kind, int B, TaskMap tasks, int64_t n_tasks, SubGeo sg,
                       SubCand *cands, hipStream_t st) {
    if (n_tasks <= 0) return WBX_OK;
    if (n_tasks > 0x7fffffff) return fail(WBX_EUNSUPPORTED, "too many subsequence tasks");
    const dim3 grid((unsigned)n_tasks);
    if (kind == WBX_EUCLIDEAN) {
        hipLaunchKernelGGL(k_subseq_euclid32<4>, grid, dim3(64), 0, st, tasks, sg, cands);
    } else {
        switch (B) {
#define SD_CASE(B_, U_) \
    case B_: hipLaunchKernelGGL((k_subseq_dtw32<B_, U_>), grid, dim3(64), 0, st, tasks, sg, cands); break;
            SUB_SCAN_CASES(SD_CASE)
#undef SD_CASE
            default: return fail(WBX_EUNSUPPORTED, "subsequence band %d", B);
        }
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_subseq_exact_cands(int kind, int B, TaskMap tasks, const SubCand *cands,
                              int64_t n_tasks, const double *S64, const int64_t *soff,
                              const int64_t *slen, const double *X64, int64_t xstride, int T,
                              double r, double *cost64, hipStream_t st) {
    if (n_tasks <= 0) return WBX_OK;
    const int64_t nl = n_tasks * kMaxCand;
    const dim3 grid((unsigned)((nl + 63) / 64));
    switch (kind == WBX_EUCLIDEAN ? 8 : (B + 7) / 8 * 8) {
#define C_CASE(B_, U_)                                                                            \
    case B_:                                                                                      \
        hipLaunchKernelGGL(k_subseq_cand64<B_>, grid, dim3(64), 0, st, tasks, cands, n_tasks, S64, \
                           soff, slen, X64, xstride, T, r, kind, cost64);                         \
        break;
        SUB_DTW_CASES(C_CASE)
#undef C_CASE
        default: return fail(WBX_EUNSUPPORTED, "subsequence band %d", B);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

int launch_subseq_select(TaskMap tasks, const SubCand *cands, int64_t n_tasks,
                         const double *cost64, int64_t n_shapelets, double *dist, int64_t *idx,
                         int32_t *overflow, int *n_overflow, hipStream_t st) {
    if (n_tasks <= 0) return WBX_OK;
    hipLaunchKernelGGL(k_subseq_select, dim3((unsigned)((n_tasks + 255) / 256)), dim3(256), 0, st,
                       tasks, cands, n_tasks, cost64, n_shapelets, dist, idx, overflow, n_overflow);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
