// Learned split index for gfx950: a piecewise-linear model of the split
// position j*(k) over K segments, fitted on the build's own prefix sums, with an
// exact bounded local search and the plain bisection as fallback and judge.
//
// Model: K + 1 knots k_m = floor(m p / K); the knot splits are computed exactly
// by bisection (they are splits of the table).  For k inside segment m the
// model predicts j by linear interpolation between the knot values and the
// search gallops outwards from the prediction inside [j_m, j_{m+1}].
//
// Bit identity with bisection: the partition certifies max dL < min dH over
// the stored prefix sums (dL, dH: consecutive increments, sentinels included).  Then the
// FOUND predicate of SplitMethodBasic (SplitMethodBasic.cuh:23-39) holds at
// exactly one j and bisection returns it, so any exact search returns the same
// (i, j, spill).  Without the certificate, or when the bracket check or the
// local search fails, the k is resolved by full bisection (and counted).
#include "atg_kernels.h"

namespace atg {


__global__ __launch_bounds__(256) void split_knots(BuildArgs a) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m > min<int64_t>(a.learned_segments, st->p)) return;
    if (knot_split(a, m)) atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
}

__global__ __launch_bounds__(256) void split_learned(BuildArgs a, int64_t blk0) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t k = (blk0 + (int64_t)blockIdx.x) * blockDim.x + threadIdx.x + 1;
    if (k >= st->p) return;
    const uint32_t f = learned_split(a, k);
    if (f & 1u) atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
    if (f & 2u) atomicAdd((unsigned long long*)&st->learned_fallbacks, 1ull);
}

hipError_t launch_split_knots(const BuildArgs& a, hipStream_t s) {
    const int64_t kmax = a.p_req;
    const int64_t K = a.learned_segments < kmax ? (int64_t)a.learned_segments : kmax;
    hipLaunchKernelGGL(split_knots, dim3((unsigned)((K + 1 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_split_learned_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1) {
    if (b1 > b0) hipLaunchKernelGGL(split_learned, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, a, b0);
    return hipGetLastError();
}

hipError_t launch_split_learned(const BuildArgs& a, hipStream_t s) {
    hipError_t e = launch_degenerate(a, s);
    if (e == hipSuccess) e = launch_split_knots(a, s);
    if (e == hipSuccess) e = launch_split_learned_blocks(a, s, 0, split_blocks(a));
    return e;
}

}  // namespace atg
