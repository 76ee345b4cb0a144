// subseq_fast.h -- types and launchers of the FP32 subsequence scan with the
// certified float64 finish (subseq_fast.hip).
#pragma once

#include "wbx_internal.h"

namespace wbx {

constexpr int kMaxCand = 8;

struct SubTask {
    int32_t sample;    // row of X
    int32_t shapelet;  // index into the shapelet arrays
};

// Tasks of one launch: every sample against the shapelets listed in `list`
// (shapelets of one band class); task t = (t / n_list, list[t % n_list]).
struct TaskMap {
    const int32_t *list;
    int32_t n_list;
    __device__ __forceinline__ SubTask at(int64_t t) const {
        SubTask k;
        k.sample = (int32_t)(t / n_list);
        k.shapelet = list[t - (int64_t)k.sample * n_list];
        return k;
    }
};

struct SubCand {
    int32_t count;             // number of candidates (> kMaxCand: overflow)
    int32_t win[kMaxCand];     // window starts, increasing
    float best32;
};

struct SubGeo {
    const float *S;     // packed shapelets (float32)
    const int64_t *off;
    const int64_t *len;
    const float *X;
    int64_t xstride;
    int T;
    double r;
    float rho, A;       // certification: tau(c) = rho (c + A) + A
};

int launch_subseq_fast(int kind, int B, TaskMap tasks, int64_t n_tasks, SubGeo sg,
                       SubCand *cands, hipStream_t st);
int launch_subseq_exact_cands(int kind, int B, TaskMap tasks, const SubCand *cands,
                              int64_t n_tasks, const double *S64, const int64_t *soff,
                              const int64_t *slen, const double *X64, int64_t xstride, int T,
                              double r, double *cost64, hipStream_t st);
int launch_subseq_select(TaskMap tasks, const SubCand *cands, int64_t n_tasks,
                         const double *cost64, int64_t n_shapelets, double *dist, int64_t *idx,
                         int32_t *overflow, int *n_overflow, hipStream_t st);
int launch_subseq_rescan(TaskMap tasks, const int32_t *list, const int *n_list_dev,
                         int n_list_host, int kind, const double *S64, const int64_t *soff,
                         const int64_t *slen, const double *X64, int64_t xstride, int T, double r,
                         int max_len, int64_t n_shapelets, double *dist, int64_t *idx,
                         hipStream_t st);

}  // namespace wbx
