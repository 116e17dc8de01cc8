// Split search of SplitMethodBasic (SplitMethodBasic.cu:14-50, predicate
// SplitMethodBasic.cuh:9-41), shared by the bisection and learned kernels.
#pragma once

#include "atg_device.h"

namespace atg {

// 0 FOUND, 1 continue left, 2 continue right.
__device__ __forceinline__ int check_split(int64_t i, int64_t j, double target, const double* Lp, int64_t numL,
                                           const double* Hp, int64_t numH, double& sigma) {
    if (j + 1 >= numH || i - 1 < 0) return 1;  // LEFT
    if (i >= numL || j < 0) return 2;          // RIGHT
    const double s = Lp[i] + Hp[j];
    const double sn = Lp[i - 1] + Hp[j + 1];
    sigma = s;
    if (s <= target && sn > target) return 0;  // FOUND
    if (s <= target) return 2;
    return 1;
}

// Bisection; 64 probes bound the loop (the reference spins forever when
// FOUND is unreachable), then a deterministic clamp is returned.
__device__ __forceinline__ bool bisect_split(int64_t k, int64_t p, int64_t nToSplit, double wToSplit,
                                             const double* Lp, int64_t numL, const double* Hp, const double* Hw,
                                             int64_t numH, Split& out) {
    const int64_t n_ = (int64_t)ceil(((double)nToSplit * (double)k) / (double)p);
    const int64_t b0 = min(n_, numH) - 1;
    const double target = (wToSplit * (double)n_) / (double)nToSplit;
    int64_t lo = 0, hi = b0;
    double sigma = 0.0;
    for (int it = 0; it < 64; ++it) {
        const int64_t j = (lo + hi) / 2;
        const int64_t i = min(n_ - j, numL - 1);
        const int stt = check_split(i, j, target, Lp, numL, Hp, numH, sigma);
        if (stt == 0) {
            out.i = (int32_t)i; out.j = (int32_t)j; out.spill = sigma + Hw[j] - target;
            return true;
        }
        if (stt == 2) lo = j + 1; else hi = j;
    }
    int64_t j = min(max<int64_t>(lo, 0), max<int64_t>(b0, 0));
    int64_t i = min(n_ - j, numL - 1);
    if (i < 0) i = 0;
    out.i = (int32_t)i; out.j = (int32_t)j; out.spill = (Lp[i] + Hp[j]) + Hw[j] - target;
    return false;
}

}  // namespace atg
