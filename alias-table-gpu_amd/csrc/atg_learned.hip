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

// P-ary split search (SplitMethodParySearch.cu:9-145): a group of PARY_LANES
// lanes per split probes PARY_LANES positions of the bracket at once and
// keeps the part between the last RIGHT and the first LEFT probe (about
// PARY_LANES + 1 times smaller each step), then probes the last <= PARY_LANES
// positions together.  Under the split certificate the FOUND position is the
// only one and the statuses run RIGHT .. FOUND .. LEFT along j, so the search
// returns the bisection's (i, j, spill); without the certificate, or when no
// probe is FOUND, the split is resolved by bisection (and counted).
#define PARY_LANES 16
__global__ __launch_bounds__(256) void split_pary(BuildArgs a, int64_t blk0) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int g = threadIdx.x % PARY_LANES;
    const int shift = (threadIdx.x & 63) / PARY_LANES * PARY_LANES;  // the group's first lane in the wave
    const int64_t k = ((blk0 + (int64_t)blockIdx.x) * blockDim.x + threadIdx.x) / PARY_LANES + 1;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const bool active = k < p;  // whole groups
    if (k == 1 && g == 0) {
        a.splits[0] = Split{0, 0, 0.0};
        a.splits[p] = Split{(int32_t)(numL - 1), (int32_t)(numH - 1), 0.0};
    }
    const bool cert = u2d(st->cert_max_dL) < u2d(st->cert_min_dH);
    const int64_t nToSplit = st->nToSplit;
    const double wToSplit = st->wToSplit;
    const int64_t n_ = active ? (int64_t)ceil(((double)nToSplit * (double)k) / (double)p) : 0;
    const int64_t b0 = min(n_, numH) - 1;
    const double target = active ? (wToSplit * (double)n_) / (double)nToSplit : 0.0;
    int64_t lo = 0, hi = b0;
    bool found = false, live = active && cert;
    int found_lane = -1;  // the group's FOUND lane (the same value on every lane of the group)
    Split out{0, 0, 0.0};
    // Every lane of a wave runs the loop until all its groups are done (the
    // ballots are wave-wide); finished groups only idle.
    for (int level = 0; level < 64; ++level) {
        const bool wide = live && hi - lo + 1 > PARY_LANES;
        int64_t j = -1;
        if (live) j = wide ? lo + ((hi - lo) * (int64_t)(g + 1)) / (PARY_LANES + 1) : lo + g;
        int stt = 1;
        double sigma = 0.0;
        int64_t i = 0;
        if (live && j <= hi) {
            i = min(n_ - j, numL - 1);
            stt = check_split(i, j, target, a.Lp, numL, a.Hp, numH, sigma);
        }
        const uint64_t fnd = (__ballot(live && j <= hi && stt == 0) >> shift) & ((1ull << PARY_LANES) - 1);
        const uint64_t rgt = (__ballot(live && j <= hi && stt == 2) >> shift) & ((1ull << PARY_LANES) - 1);
        if (live && fnd) {
            found_lane = __ffsll((long long)fnd) - 1;
            if (g == found_lane) {
                out.i = (int32_t)i;
                out.j = (int32_t)j;
                out.spill = sigma + a.Hw[j] - target;
            }
            found = true;
            live = false;
        } else if (live) {
            if (!wide) {
                live = false;  // no FOUND among the last positions: bisection below
            } else {
                // RIGHT probes are a prefix of the group: the answer lies
                // after the last of them and before the next probe.
                const int nr = __popcll(rgt);
                const int64_t jl = nr > 0 ? __shfl(j, shift + nr - 1, 64) + 1 : lo;
                const int64_t jh = nr < PARY_LANES ? __shfl(j, shift + nr, 64) - 1 : hi;
                lo = jl;
                hi = jh;
                if (lo > hi) live = false;
            }
        }
        if (__ballot(live) == 0ull) break;
    }
    if (!active) return;
    if (found) {
        if (g == found_lane) a.splits[k] = out;
    } else if (g == 0) {
        Split sp;
        if (!bisect_split(k, p, nToSplit, wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, sp))
            atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
        a.splits[k] = sp;
        atomicAdd((unsigned long long*)&st->learned_fallbacks, 1ull);
    }
}

// Blocks of 256 splits as the other split kernels count them (blocks
// [b0, b1)), 256 / PARY_LANES splits per launched block.
hipError_t launch_split_pary_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1) {
    constexpr int64_t per = 256 / PARY_LANES;
    if (b1 > b0) hipLaunchKernelGGL(split_pary, dim3((unsigned)((b1 - b0) * per)), dim3(256), 0, s, a, b0 * per);
    return hipGetLastError();
}

hipError_t launch_split_pary(const BuildArgs& a, hipStream_t s) {
    hipError_t e = launch_degenerate(a, s);
    if (e == hipSuccess) e = launch_split_pary_blocks(a, s, 0, split_blocks(a));
    return e;
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
