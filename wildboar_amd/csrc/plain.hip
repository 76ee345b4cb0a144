// plain.hip -- exact (FP64 bit-identical) kernels and small helpers.
//
//  k_plain_dtw    one pair per 64-lane block, anti-diagonal sweep of the
//                 reference's band with three diagonals in LDS.  Any lengths,
//                 any band, WDTW weights with the row-0 index quirk.  This is
//                 the FP64 parity kernel and the "plain per-pair" baseline.
//  k_euclid_grid  64x64 pairs per block, LDS tiles, each lane sums 4x4 pairs,
//                 each sequentially in the reference's order (exact in FP64).
//  k_finish       sqrt + scatter into the driver's output layout.
//  k_argmin_replay_wave  _ArgMinBatch's heap acceptance rule, one wave per x row.
//
// Arithmetic order per cell is the reference's (distance/_elastic.pyx:925-928):
//     cost = min(min(up, left), diag) + (v * v) * w,   v = x_i - y_j
// and this file is compiled with contraction off, so no FMA appears.
#pragma clang fp contract(off)

#include "block_dtw.h"

namespace wbx {

// w_fixed > 0 replaces the band width of every pair (the derivative
// metrics band by the lengths of the series before the slope transform).
template <typename T, bool AD>
__global__ __launch_bounds__(64) void k_plain_dtw(PairSpec ps, Series<T> xs,
                                                  Series<T> ys, double r,
                                                  const double *__restrict__ rp,
                                                  const T *__restrict__ weights,
                                                  int cap, int w_fixed, T pen,
                                                  T *__restrict__ sq) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *buf = reinterpret_cast<T *>(smem_raw);
    const int64_t p = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (p >= ps.n_pairs) return;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const int lx = xs.length(a), ly = ys.length(b);
    const double rr = rp ? rp[a] : r;
    const int w = w_fixed > 0 ? w_fixed : band_w(lx, ly, rr);
    const T res = block_dtw<T, AD>(xs.row(a), lx, ys.row(b), ly, w, weights, buf, buf + cap,
                                   buf + 2 * cap, pen);
    if (threadIdx.x == 0) sq[p] = res;
}

// A measurement baseline, selected only by the batch plan's policy 2: the
// plain kernel as it runs when its loops do not use the per-pair lengths.
// Every pair walks all 2 t_pad - 1 anti-diagonals of the padded matrix with
// its own band offsets; cells outside its (lx, ly) matrix are masked to +inf.
// The results equal k_plain_dtw's.
template <typename T>
__global__ __launch_bounds__(64) void k_plain_dtw_padded(PairSpec ps, Series<T> xs,
                                                         Series<T> ys, double r,
                                                         const double *__restrict__ rp,
                                                         const T *__restrict__ weights,
                                                         int cap, T *__restrict__ sq) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *d0 = reinterpret_cast<T *>(smem_raw);
    T *d1 = d0 + cap, *d2 = d0 + 2 * cap;
    const int64_t p = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (p >= ps.n_pairs) return;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const T *__restrict__ X = xs.row(a);
    const T *__restrict__ Y = ys.row(b);
    const int lx = xs.length(a), ly = ys.length(b);
    const int tp = cap;  // the padded length of every row
    const int w = band_w(lx, ly, rp ? rp[a] : r);
    const int min_len = lx > ly ? lx - ly : 0;
    const int max_len = (ly > lx ? ly - lx : 0) + w;
    const T INF = inf_of<T>();
    const int tid = threadIdx.x;
    T res = INF;
    for (int d = 0; d <= 2 * tp - 2; ++d) {
        int ilo = d - tp + 1;
        if (ilo < 0) ilo = 0;
        const int ilo2 = floordiv2(d - max_len) + 1;
        if (ilo2 > ilo) ilo = ilo2;
        int ihi = d < tp - 1 ? d : tp - 1;
        const int ihi2 = floordiv2(d + min_len + w - 1);
        if (ihi2 < ihi) ihi = ihi2;
        const int s = ilo - 1 > 0 ? ilo - 1 : 0;
        const int e = ihi + 1 < tp - 1 ? ihi + 1 : tp - 1;
        for (int i = s + tid; i <= e; i += blockDim.x) {
            T c = INF;
            const int j = d - i;
            if (i >= ilo && i <= ihi && i < lx && j < ly) {
                const T v = X[i] - Y[j];
                T wt = (T)1.0;
                if (weights) wt = (i == 0 && j > 0) ? weights[j - 1] : weights[i > j ? i - j : j - i];
                if (i == 0 && j == 0) {
                    c = v * v * wt;
                } else {
                    const T up = i > 0 ? d1[i - 1] : INF;
                    const T left = j > 0 ? d1[i] : INF;
                    const T dg = (i > 0 && j > 0) ? d2[i - 1] : INF;
                    c = cmin(cmin(up, left), dg) + v * v * wt;
                }
            }
            d0[i] = c;
        }
        __syncthreads();
        if (d == lx + ly - 2) res = d0[lx - 1];
        T *t = d2;
        d2 = d1;
        d1 = d0;
        d0 = t;
    }
    if (tid == 0) sq[p] = res;
}

template <typename T>
static int set_lds(const void *fn, size_t lds) {
    if (lds > 64 * 1024) {
        if (lds > 160 * 1024)
            return fail(WBX_EUNSUPPORTED, "plain kernel needs %zu bytes of LDS", lds);
        WBX_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    return WBX_OK;
}

template <typename T>
int launch_plain_dtw(const PairSpec &ps, Series<T> xs, Series<T> ys, double r,
                     const T *weights, T *sq, hipStream_t st, const double *rp,
                     bool padded_iter, int w_fixed, const double *penalty) {
    if (ps.n_pairs <= 0) return WBX_OK;
    const int cap = xs.len;  // every row length of xs is <= xs.len
    const size_t lds = (size_t)3 * (size_t)cap * sizeof(T);
    const int64_t nb = ps.n_pairs;
    dim3 grid((unsigned)(nb < 65536 ? nb : 65536), (unsigned)((nb + 65535) / 65536));
    if (padded_iter) {
        if (ys.len != xs.len) return fail(WBX_EINVAL, "padded iteration needs one padded length");
        if (w_fixed > 0 || penalty) return fail(WBX_EUNSUPPORTED, "padded iteration: dtw only");
        int rc = set_lds<T>((const void *)k_plain_dtw_padded<T>, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(k_plain_dtw_padded<T>, grid, dim3(64), lds, st, ps, xs, ys, r, rp,
                           weights, cap, sq);
    } else if (penalty) {
        if (weights) return fail(WBX_EUNSUPPORTED, "adtw takes no weights");
        int rc = set_lds<T>((const void *)k_plain_dtw<T, true>, lds);
        if (rc) return rc;
        hipLaunchKernelGGL((k_plain_dtw<T, true>), grid, dim3(64), lds, st, ps, xs, ys, r, rp,
                           weights, cap, w_fixed, (T)*penalty, sq);
    } else {
        int rc = set_lds<T>((const void *)k_plain_dtw<T, false>, lds);
        if (rc) return rc;
        hipLaunchKernelGGL((k_plain_dtw<T, false>), grid, dim3(64), lds, st, ps, xs, ys, r, rp,
                           weights, cap, w_fixed, (T)0, sq);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_plain_dtw<float>(const PairSpec &, Series<float>, Series<float>, double,
                                     const float *, float *, hipStream_t, const double *, bool,
                                     int, const double *);
template int launch_plain_dtw<double>(const PairSpec &, Series<double>, Series<double>, double,
                                      const double *, double *, hipStream_t, const double *,
                                      bool, int, const double *);

// ------------------------------------------------------------- euclidean
// euclidean_distance (distance/_metric.pyx:654-683) with S = x, T = y: windows
// of y; per window dist = sum_j (y[i+j] - x[j])^2 accumulated in order.  Equal
// lengths give one window.  GRID mode: 64 x 64 pairs per block, 32 timesteps
// of the 64 x rows and 64 y rows staged in LDS per chunk, each lane summing
// 4 x 4 pairs (rows ty + 16 i, columns tx + 16 j: conflict-free LDS reads),
// so a timestep costs 8 LDS reads for 16 pairs.  Every pair's sum stays
// sequential in j (float64: multiply then add, the reference's arithmetic;
// float32: fused).  Other modes use one lane per pair straight from memory.
constexpr int EB = 64;   // tile edge (pairs)
constexpr int EK = 32;   // timesteps per LDS chunk

__device__ __forceinline__ void euclid_acc(float &acc, float v) { acc = __builtin_fmaf(v, v, acc); }
__device__ __forceinline__ void euclid_acc(double &acc, double v) { acc += v * v; }

template <typename T>
__global__ __launch_bounds__(256) void k_euclid_grid(Series<T> xs, int64_t nx,
                                                     Series<T> ys, int64_t ny,
                                                     int t, T *__restrict__ sq) {
    // rows padded to EK + 2: even, so a pair of timesteps is one aligned
    // 2-element LDS read, and 16 rows read together fall in distinct banks
    __shared__ __attribute__((aligned(16))) T sx[EB][EK + 2];
    __shared__ __attribute__((aligned(16))) T sy[EB][EK + 2];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int64_t a0 = (int64_t)blockIdx.y * EB, b0 = (int64_t)blockIdx.x * EB;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
    // chunk c0's values are loaded into registers while chunk c0 - EK is
    // summed, then stored to LDS
    constexpr int PER = EB * EK / 256;  // values of each tile per thread
    T px[PER], py[PER];
    auto fetch = [&](int c0) {
        const int cn = t - c0 < EK ? t - c0 : EK;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = threadIdx.x + 256 * u;
            const int rr = q / EK, cc = q % EK;
            T vx = (T)0, vy = (T)0;
            if (cc < cn) {
                if (a0 + rr < nx) vx = xs.row(a0 + rr)[c0 + cc];
                if (b0 + rr < ny) vy = ys.row(b0 + rr)[c0 + cc];
            }
            px[u] = vx;
            py[u] = vy;
        }
    };
    fetch(0);
    for (int c0 = 0; c0 < t; c0 += EK) {
        const int cn = t - c0 < EK ? t - c0 : EK;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = threadIdx.x + 256 * u;
            sx[q / EK][q % EK] = px[u];
            sy[q / EK][q % EK] = py[u];
        }
        __syncthreads();
        if (c0 + EK < t) fetch(c0 + EK);
        int cc = 0;
        for (; cc + 2 <= cn; cc += 2) {  // two timesteps per LDS read, in order
            typedef T t2 __attribute__((ext_vector_type(2)));
            t2 xv[4], yv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                xv[i] = *(const t2 *)&sx[ty + 16 * i][cc];
                yv[i] = *(const t2 *)&sy[tx + 16 * i][cc];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) euclid_acc(acc[i][j], yv[j][h] - xv[i][h]);
        }
        for (; cc < cn; ++cc) {
            T xv[4], yv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                xv[i] = sx[ty + 16 * i][cc];
                yv[i] = sy[tx + 16 * i][cc];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) euclid_acc(acc[i][j], yv[j] - xv[i]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t a = a0 + ty + 16 * i;
        if (a >= nx) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t b = b0 + tx + 16 * j;
            if (b < ny) sq[a * ny + b] = acc[i][j];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_euclid_pairs(PairSpec ps, Series<T> xs,
                                                      Series<T> ys, T *__restrict__ sq) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= ps.n_pairs) return;
    int64_t a, b;
    decode_pair(ps, p, a, b);
    const T *__restrict__ S = xs.row(a);
    const T *__restrict__ Tt = ys.row(b);
    const int sl = xs.length(a), tl = ys.length(b);
    T min_dist = inf_of<T>();
    for (int i = 0; i < tl - sl + 1; ++i) {
        T dist = (T)0;
        for (int j = 0; j < sl; ++j) {
            if (dist >= min_dist) break;
            T v = Tt[i + j];
            v -= S[j];
            dist += v * v;
        }
        if (dist < min_dist) min_dist = dist;
    }
    sq[p] = min_dist;
}

template <typename T>
int launch_euclid_pairs(const PairSpec &ps, Series<T> xs, Series<T> ys, T *sq,
                        hipStream_t st) {
    if (ps.n_pairs <= 0) return WBX_OK;
    const bool fixed = xs.lens == nullptr && ys.lens == nullptr && xs.len == ys.len;
    if (ps.mode == PAIR_GRID && fixed) {
        const int64_t nx = ps.n_pairs / ps.ny, ny = ps.ny;
        dim3 grid((unsigned)((ny + EB - 1) / EB), (unsigned)((nx + EB - 1) / EB));
        if (grid.y > 65535) return fail(WBX_EUNSUPPORTED, "too many rows for euclidean tiles");
        hipLaunchKernelGGL(k_euclid_grid<T>, grid, dim3(256), 0, st, xs, nx, ys, ny, xs.len, sq);
    } else {
        dim3 grid((unsigned)((ps.n_pairs + 255) / 256));
        hipLaunchKernelGGL(k_euclid_pairs<T>, grid, dim3(256), 0, st, ps, xs, ys, sq);
    }
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

template int launch_euclid_pairs<float>(const PairSpec &, Series<float>, Series<float>, float *,
                                        hipStream_t);
template int launch_euclid_pairs<double>(const PairSpec &, Series<double>, Series<double>,
                                         double *, hipStream_t);

// ------------------------------------------------------------- finish
// GRID: out[a * ld + b];  TRIANGLE: out[a*ld+b] = out[b*ld+a] (diagonal is
// zeroed by the caller);  DIAG/LIST: out[p].
template <typename T>
__global__ __launch_bounds__(256) void k_finish(PairSpec ps, const T *__restrict__ sq,
                                                double *__restrict__ out, int64_t ld) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= ps.n_pairs) return;
    const double v = sqrt((double)sq[p]);
    if (ps.mode == PAIR_GRID) {
        out[p] = v;
    } else if (ps.mode == PAIR_TRIANGLE) {
        int64_t a, b;
        decode_pair(ps, p, a, b);
        out[a * ld + b] = v;
        out[b * ld + a] = v;
    } else if (ps.mode == PAIR_LIST) {
        out[(int64_t)ps.pa[p] * ld + ps.pb[p]] = v;
    } else {
        out[p] = v;
    }
}

template <typename T>
int launch_finish(const PairSpec &ps, const T *sq, double *out, int64_t ld, hipStream_t st) {
    if (ps.n_pairs <= 0) return WBX_OK;
    dim3 grid((unsigned)((ps.n_pairs + 255) / 256));
    hipLaunchKernelGGL(k_finish<T>, grid, dim3(256), 0, st, ps, sq, out, ld);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}
template int launch_finish<float>(const PairSpec &, const float *, double *, int64_t, hipStream_t);
template int launch_finish<double>(const PairSpec &, const double *, double *, int64_t,
                                   hipStream_t);

// ------------------------------------------------------------- argmin
// _ArgMinBatch.__call__ (distance/_cdistance.pyx:1318-1345) on exact distances:
// j is pushed iff not skipped by the lower bound (lb >= threshold) and
// D[j] < threshold; threshold is +inf until the heap holds k elements and the
// heap maximum afterwards.  Heap = utils/_misc.pyx:18-107.  Early abandoning
// in the reference never changes which j are pushed (a pair is abandoned only
// when its cost exceeds the threshold), so replaying exact values reproduces
// the reference's heap, element order included.
struct HeapEl {
    double value;
    int64_t index;
};

__device__ void heap_shift_down(HeapEl *h, int64_t startpos, int64_t pos) {
    HeapEl ne = h[pos];
    while (pos > startpos) {
        int64_t parent = (pos - 1) >> 1;
        if (ne.value > h[parent].value) {
            h[pos] = h[parent];
            pos = parent;
            continue;
        }
        break;
    }
    h[pos] = ne;
}

__device__ void heap_shift_up(HeapEl *h, int64_t pos, int64_t endpos) {
    int64_t startpos = pos;
    HeapEl ne = h[pos];
    int64_t child = 2 * pos + 1;
    while (child < endpos) {
        int64_t right = child + 1;
        if (right < endpos && h[child].value < h[right].value) child = right;
        h[pos] = h[child];
        pos = child;
        child = 2 * pos + 1;
    }
    h[pos] = ne;
    heap_shift_down(h, startpos, pos);
}

// One wave per row: 64 values are read together and only those below the
// threshold at the start of the group (and not skipped by the bound) are
// replayed, in index order, by lane 0; the threshold only falls, so every
// value the sequential rule accepts is among them and the heap sees exactly
// the reference's pushes.
__global__ __launch_bounds__(64) void k_argmin_replay_wave(const double *__restrict__ D,
                                                           int64_t nx, int64_t ny, int64_t k,
                                                           const double *__restrict__ lb,
                                                           HeapEl *__restrict__ heaps,
                                                           int64_t *__restrict__ idx,
                                                           double *__restrict__ dist) {
    const int64_t i = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    if (i >= nx) return;  // uniform: one row per wave
    const int lane = threadIdx.x;
    HeapEl *h = heaps + i * k;
    const double *row = D + i * ny;
    double threshold = WBX_INF_D;
    int64_t n = 0;
    for (int64_t b0 = 0; b0 < ny; b0 += 64) {
        const int64_t j = b0 + lane;
        const bool in = j < ny;
        const double v = in ? row[j] : WBX_INF_D;
        const double l = (lb && in) ? lb[i * ny + j] : -WBX_INF_D;
        const bool cand = in && v < threshold && !(l >= threshold);
        uint64_t mask = __ballot(cand);
        while (mask) {
            const int q = __builtin_ctzll(mask);
            mask &= mask - 1;
            const double vq = __shfl(v, q), lq = __shfl(l, q);
            if (lq >= threshold || !(vq < threshold)) continue;
            double top = 0.0;
            if (lane == 0) {
                HeapEl e{vq, b0 + q};
                if (n == 0) {
                    h[0] = e;
                } else if (n < k) {
                    h[n] = e;
                    heap_shift_down(h, 0, n);
                } else if (h[0].value > vq) {
                    h[0] = e;
                    heap_shift_up(h, 0, k);
                }
                top = h[0].value;
            }
            n = n < k ? n + 1 : n;
            top = __shfl(top, 0);
            threshold = (n == k) ? top : WBX_INF_D;
        }
    }
    if (lane == 0)
        for (int64_t q = 0; q < k; ++q) {
            idx[i * k + q] = q < n ? h[q].index : 0;
            dist[i * k + q] = q < n ? h[q].value : 0.0;
        }
}

int launch_argmin_replay(const double *D, int64_t nx, int64_t ny, int64_t k,
                         const double *lower_bound, void *heaps, int64_t *idx, double *dist,
                         hipStream_t st) {
    if (nx <= 0) return WBX_OK;
    const int64_t gx = nx < 65535 ? nx : 65535;
    hipLaunchKernelGGL(k_argmin_replay_wave, dim3((unsigned)gx, (unsigned)((nx + gx - 1) / gx)),
                       dim3(64), 0, st, D, nx, ny, k, lower_bound, (HeapEl *)heaps, idx, dist);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

// ------------------------------------------------------------- conversions
__global__ void k_f64_to_f32(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        d[q] = (float)s[q];
}
__global__ void k_f32_to_f64(const float *__restrict__ s, double *__restrict__ d, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        d[q] = (double)s[q];
}

int launch_f64_to_f32(const double *src, float *dst, int64_t n, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)(nb < 8192 ? nb : 8192)), dim3(256), 0, st,
                       src, dst, n);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}
int launch_f32_to_f64(const float *src, double *dst, int64_t n, hipStream_t st) {
    if (n <= 0) return WBX_OK;
    int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)(nb < 8192 ? nb : 8192)), dim3(256), 0, st,
                       src, dst, n);
    WBX_HIP_CHECK(hipGetLastError());
    return WBX_OK;
}

}  // namespace wbx
