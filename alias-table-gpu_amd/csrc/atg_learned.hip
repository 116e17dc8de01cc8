// Learned split index for gfx950: a piecewise-linear model of the split
// position j*(k) over K segments, fitted on the build's own prefix sums, with an
// exact bounded local search and the plain bisection as fallback and judge.
//
// Model: K + 1 knots k_m = floor(m p / K); the knot splits are computed exactly
// by bisection (they are splits of the table).  For k inside segment m the
// model predicts j by linear interpolation between the knot values and the
// search gallops outwards from the prediction inside [j_m, j_{m+1}].
//
// Bit identity with bisection: the partition pass certifies max dL < min dH
// over the stored prefix sums (dL, dH: consecutive increments).  Then the
// FOUND predicate of SplitMethodBasic (SplitMethodBasic.cuh:23-39) holds at
// exactly one j and bisection returns it, so any exact search returns the same
// (i, j, spill).  Without the certificate, or when the bracket check or the
// local search fails, the k is resolved by full bisection (and counted).
#include "atg_build.h"
#include "atg_device.h"
#include "atg_split.h"

namespace atg {


__device__ __forceinline__ int64_t knot_k(int64_t m, int64_t p, int64_t K) { return (m * p) / K; }

__global__ __launch_bounds__(256) void split_knots(BuildArgs a) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const int64_t K = min<int64_t>(a.learned_segments, p);
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m > K) return;
    const int64_t k = knot_k(m, p, K);
    Split sp;
    if (k == 0) {
        sp = Split{0, 0, 0.0};
    } else if (k == p) {
        sp = Split{(int32_t)(numL - 1), (int32_t)(numH - 1), 0.0};
    } else if (!bisect_split(k, p, st->nToSplit, st->wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, sp)) {
        atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
    }
    a.splits[k] = sp;
    a.knot_j[m] = sp.j;
}

__global__ __launch_bounds__(256) void split_learned(BuildArgs a) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const int64_t K = min<int64_t>(a.learned_segments, p);
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k >= p) return;
    // Segment m with k_m <= k < k_{m+1}.
    int64_t m = (k * K) / p;
    while (m < K && knot_k(m + 1, p, K) <= k) ++m;
    while (m > 0 && knot_k(m, p, K) > k) --m;
    const int64_t km = knot_k(m, p, K), km1 = knot_k(m + 1, p, K);
    if (k == km) return;  // a knot: written by split_knots
    const int64_t nToSplit = st->nToSplit;
    const double wToSplit = st->wToSplit;
    double dl, dh;
    {
        const uint64_t bl = st->cert_max_dL, bh = st->cert_min_dH;
        dl = u2d(bl);
        dh = u2d(bh);
    }
    Split out;
    bool ok = false;
    if (dl < dh) {
        const int64_t n_ = (int64_t)ceil(((double)nToSplit * (double)k) / (double)p);
        const int64_t b0 = min(n_, numH) - 1;
        const double target = (wToSplit * (double)n_) / (double)nToSplit;
        int64_t lo = a.knot_j[m], hi = a.knot_j[m + 1];
        if (hi > b0) hi = b0;
        if (lo < 0) lo = 0;
        if (lo <= hi) {
            // Model prediction: linear interpolation inside the segment.
            int64_t j = lo + ((hi - lo) * (k - km)) / (km1 - km);
            double sigma = 0.0;
            int stt = check_split(min(n_ - j, numL - 1), j, target, a.Lp, numL, a.Hp, numH, sigma);
            // Gallop from the prediction until the answer is bracketed, then bisect.
            int64_t a0 = lo, b1 = hi;
            if (stt == 2) {  // RIGHT: answer above j
                a0 = j + 1;
                int64_t step = 1;
                while (true) {
                    const int64_t t = j + step;
                    if (t >= hi) { b1 = hi; break; }
                    const int s2 = check_split(min(n_ - t, numL - 1), t, target, a.Lp, numL, a.Hp, numH, sigma);
                    if (s2 == 0) { j = t; stt = 0; break; }
                    if (s2 == 1) { b1 = t; break; }
                    a0 = t + 1;
                    step <<= 1;
                }
            } else if (stt == 1) {  // LEFT: answer below j
                b1 = j;
                int64_t step = 1;
                while (true) {
                    const int64_t t = j - step;
                    if (t <= lo) { a0 = lo; break; }
                    const int s2 = check_split(min(n_ - t, numL - 1), t, target, a.Lp, numL, a.Hp, numH, sigma);
                    if (s2 == 0) { j = t; stt = 0; break; }
                    if (s2 == 2) { a0 = t + 1; break; }
                    b1 = t;
                    step <<= 1;
                }
            }
            for (int it = 0; stt != 0 && it < 64 && a0 <= b1; ++it) {
                j = (a0 + b1) / 2;
                stt = check_split(min(n_ - j, numL - 1), j, target, a.Lp, numL, a.Hp, numH, sigma);
                if (stt == 2) a0 = j + 1;
                else if (stt == 1) b1 = j - 1;
            }
            if (stt == 0) {
                out.i = (int32_t)min(n_ - j, numL - 1);
                out.j = (int32_t)j;
                out.spill = sigma + a.Hw[j] - target;
                ok = true;
            }
        }
    }
    if (!ok) {
        if (!bisect_split(k, p, nToSplit, wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, out))
            atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
        atomicAdd((unsigned long long*)&st->learned_fallbacks, 1ull);
    }
    a.splits[k] = out;
}

hipError_t launch_split_learned(const BuildArgs& a, hipStream_t s) {
    hipError_t e = launch_degenerate(a, s);
    if (e != hipSuccess) return e;
    const int64_t kmax = a.p_req;
    const int64_t K = a.learned_segments < kmax ? (int64_t)a.learned_segments : kmax;
    hipLaunchKernelGGL(split_knots, dim3((unsigned)((K + 1 + 255) / 256)), dim3(256), 0, s, a);
    if (kmax >= 2) hipLaunchKernelGGL(split_learned, dim3((unsigned)((kmax - 1 + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace atg
