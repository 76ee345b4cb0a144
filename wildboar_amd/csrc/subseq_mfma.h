// subseq_mfma.h -- plan and launcher of the matrix-core euclidean subsequence
// scan (subseq_mfma.hip).
#pragma once

#include <vector>

#include "wbx_internal.h"

namespace wbx {

// 32 shapelets of similar length, zero-padded to K (a multiple of 8).
struct MfmaChunk {
    int32_t base;  // float offset of the chunk's 32 x K rows
    int32_t K;
    int32_t lmin;  // shortest shapelet of the chunk (sets the window count)
    int32_t pad_;
};

// Host plan: shapelets sorted by length (descending) and cut into chunks.
struct MfmaPlan {
    std::vector<int32_t> slist;  // list position -> shapelet id
    std::vector<int32_t> slen;   // per slot (chunks x 32), 0 = empty
    std::vector<MfmaChunk> chunks;
    int64_t sp_floats = 0;
    int T = 0, Tp = 0, max_len = 0;
    int64_t xps = 0, ps = 0;  // row strides of the padded rows and prefix sums
};

// Kernel arguments of one launch (chunks c0 .. c0 + nc - 1, all of one K).
struct MfmaScan {
    const float *Xp;
    int64_t xps;
    const double *P;
    int64_t ps;
    const float *arow;
    const float *Sp;
    const float *S2;
    const int32_t *slen;
    const int32_t *slist;
    const MfmaChunk *chunks;
    const double *X64;
    int64_t x64s;
    const double *S64;
    const int64_t *soff;
    int32_t c0, nc, T, Tp, n_list;
    int64_t n_shapelets;
    float kappa, aabs;
    double *dist;
    int64_t *idx;
    int32_t *ovf;
    int *n_ovf;
};

// Longest shapelet the matrix-core scan takes (longer ones use the vector scan).
constexpr int kMfmaMaxLen = 128;

void mfma_plan(const std::vector<int32_t> &ids, const std::vector<int64_t> &hl, int64_t T,
               MfmaPlan *p);
size_t mfma_ws_bytes(const MfmaPlan &p, int64_t n);
// Scan, finish and (after one read-back of the overflow count) exact rescan
// of the overflowed tasks; writes dist / idx of every listed shapelet.
int mfma_scan(const MfmaPlan &p, char *ws, int inputs_exact, const float *x32, int64_t x32s,
              const float *S32, const double *X64, int64_t x64s, const double *S64,
              const int64_t *soff, const int64_t *slen_dev, int64_t n, int64_t n_shapelets,
              double *dist, int64_t *idx, hipStream_t st);

}  // namespace wbx
