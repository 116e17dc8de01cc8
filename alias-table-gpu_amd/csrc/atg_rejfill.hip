// Search fills of the rejection array A and the learned inverse-CDF index.
//
// The reference fills A (slot s holds the item whose run of the inclusive
// counts contains s: the first i with incl[i] > s) by a bisection per slot
// (fillABinarySearch, RejectionSamplingBinarySearch.cu:12-33), by a
// block-wide p-ary narrowing followed by that bisection (fillAPArySearch,
// RejectionSamplingPArySearch.cu:12-123), or by one thread per item writing
// its run (fillADirect, RejectionSamplingDirect.cu:12-28; the direct fill of
// this library is rej_fill, which also splits long runs).  The search fills
// here work on incl[i] = c_0 + ... + c_i stored as uint32 (|A| <= 2n + 4096
// < 2^32 for the n they accept), and a third one, the learned fill, predicts
// the item of each wave's first slot with a piecewise-linear model of the
// inverse CDF slot -> item fitted on incl, then finds it exactly inside the
// model's error bound.  All of them write the same A: the item of a slot is
// unique (every run is non-empty), and a slot's acceptance depends only on
// its item and on whether the slot starts the item's run.
//
// The model, with M = |A| - 1 and K segments (integer arithmetic throughout):
//   segment of slot s    seg(s) = 0 if M = 0, else min(K - 1, floor(s K / M))
//   knot slots           s_k = ceil(k M / K), k = 0..K, so seg(s) = k iff s_k <= s < s_{k+1}
//   knot items           q_k = item of slot s_k (a bisection over incl per knot)
//   prediction           pred(s) = q_k + floor((s - s_k)(q_{k+1} - q_k) / (s_{k+1} - s_k))
//   error bound          e_k = max over the slots of segment k of |item(s) - pred(s)|
// The item of every slot of segment k lies in [q_k, q_{k+1}] (items grow with
// slots) and within e_k of pred.  The bound is exact and costs one pass over
// the items: inside one item's run the item is constant and pred is
// non-decreasing, so the largest deviation sits at an end of the run's part
// in the segment, and a segment that lies inside one run has zero deviation
// (its q_k is the run's item and q_{k+1} at most the next item).
#include "atg_device.h"
#include "atg_sample.h"

namespace atg {

#define FILL_THREADS 256
// Narrowing rounds of the p-ary fill; six reach a 64-item interval from 2^31.
#define PARY_FILL_ROUNDS 8
// The p-ary fill stops narrowing once the wave's slots straddle more than
// this many of its 64 splitter gaps: another round would gain little.
#define PARY_FILL_SPREAD 16
// P-ary rounds of the learned fill inside the model's window (64^6 > 2^31).
#define LEARNED_FILL_ROUNDS 6

typedef unsigned int fill_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void fill_store(RejEntry* p, const RejEntry& e) {
    __builtin_nontemporal_store(*reinterpret_cast<const fill_u4*>(&e), reinterpret_cast<fill_u4*>(p));
}

// The reference's bisection (RejectionSamplingBinarySearch.cu:17-32) over
// incl[left, right], the item of slot s known to lie in [left, right]
// (right = n for the whole array, as in the reference).
__device__ __forceinline__ int32_t incl_bisect(const uint32_t* incl, int64_t left, int64_t right, uint64_t s) {
    while (right - left > 1) {
        const int64_t c = (left + right) >> 1;
        if ((uint64_t)incl[c] > s) right = c;
        else left = c;
    }
    return (int32_t)((uint64_t)incl[left] <= s ? right : left);
}

// floor(num / den) for den > 0.  Below 2^53 the FP64 quotient is exact
// enough: for a non-integer quotient q = num / den the distance to the next
// integer is at least 1 / den, and the rounding error at most q 2^-53 < 1 /
// den because q den = num < 2^53; an integer quotient is exact.  So the
// floor of the correctly rounded FP64 quotient is the integer quotient, at a
// fraction of the cost of the 64-bit integer division, which stays for
// larger numerators.
__device__ __forceinline__ uint64_t udiv(uint64_t num, uint64_t den) {
    if (num < (1ull << 53)) return (uint64_t)floor((double)num / (double)den);
    return num / den;
}
__device__ __forceinline__ uint64_t knot_slot(uint64_t k, uint64_t M, uint64_t K) { return udiv(k * M + K - 1, K); }
__device__ __forceinline__ uint64_t seg_of(uint64_t s, uint64_t M, uint64_t K) {
    return M == 0 ? 0 : min<uint64_t>(K - 1, udiv(s * K, M));
}
__device__ __forceinline__ int64_t seg_pred(uint64_t s, uint64_t sk, uint64_t sk1, int32_t qk, int32_t qk1) {
    return sk1 > sk ? (int64_t)qk + (int64_t)udiv((s - sk) * (uint64_t)(qk1 - qk), sk1 - sk) : (int64_t)qk;
}

// Slot s of A holds `item`; `first` is the first slot of the item's run.
// The acceptance is computed as rej_fill computes it.
template <typename T>
__device__ __forceinline__ void fill_emit(const RejArgs& a, double W, uint64_t s, int32_t item, uint64_t first) {
    RejEntry e;
    e.accept = 1.0;
    if (s == first) {
        const double x = ((double)a.n * (double)((const T*)a.weights)[item]) / W;
        e.accept = x - floor(x);
    }
    e.item = item;
    e.pad = 0;
    fill_store(&a.A[s], e);
}

// incl for the block's REJ_ITEMS items (the elementsPerItem scan of
// RejectionSampling.cu:38-46), from the block offsets of rej_scan_totals:
// wave v owns items [1024 v, 1024 v + 1024) of the block in 16 coalesced
// rounds of 64, as rej_fill.
template <typename T>
__global__ __launch_bounds__(REJ_THREADS) void rej_incl(RejArgs a) {
    constexpr int U = REJ_ITEMS / REJ_THREADS;
    __shared__ uint64_t wtot[REJ_THREADS / 64];
    if (a.st->error) return;
    const double W = a.st->W;
    const T* w = (const T*)a.weights;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * REJ_ITEMS + (int64_t)wave * 64 * U;
    const int cnt = (int)max<int64_t>(0, min<int64_t>(64 * U, a.n - base));
    T wr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = u * 64 + lane;
        wr[u] = (k < cnt) ? __builtin_nontemporal_load(&w[base + k]) : T(0);
    }
    uint32_t c[U];
    uint64_t total = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const double x = ((double)a.n * (double)wr[u]) / W;
        c[u] = (u * 64 + lane < cnt) ? (uint32_t)floor(x) + 1u : 0u;
        total += c[u];
    }
    total = wave_sum_int<uint64_t>(total);
    if (lane == 0) wtot[wave] = total;
    __syncthreads();
    uint64_t first = a.block_off[blockIdx.x];
    for (int q = 0; q < wave; ++q) first += wtot[q];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = u * 64 + lane;
        const uint32_t incl = wave_incl_scan_u32(c[u]);
        if (k < cnt) a.incl[base + k] = (uint32_t)(first + incl);
        first += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
}

// Fit, part 1: the K + 1 knot items, one bisection each; clears the bounds.
__global__ __launch_bounds__(256) void rej_model_knots(RejArgs a) {
    if (a.st->error) return;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k > a.K) return;
    const uint64_t M = a.st->rej_size - 1;
    a.knot_items[k] = incl_bisect(a.incl, 0, a.n, knot_slot((uint64_t)k, M, (uint64_t)a.K));
    if (k < a.K) a.knot_err[k] = 0;
}

// Fit, part 2: the error bounds.  Item i's run [f, l] meets segments seg(f)
// .. seg(l); only the two end segments can deviate, each at the ends of the
// run's part in it.  A block takes ERR_ITEMS consecutive items.  When they
// span at most ERR_LDS_SEGS segments, the block keeps the segments' maxima
// in LDS (a wave whose items share one segment reduces first) and adds them
// to the bounds with one atomic per segment: the grid walks the items in
// order, so atomics straight to the bounds would queue on a few addresses.
// Zero deviations are not added (the bounds start at zero).
#define ERR_ITEMS 4096
#define ERR_LDS_SEGS 512
__global__ __launch_bounds__(256) void rej_model_err(RejArgs a) {
    __shared__ uint32_t lmax[ERR_LDS_SEGS];
    if (a.st->error) return;
    const uint64_t M = a.st->rej_size - 1, K = (uint64_t)a.K;
    const int64_t i0 = (int64_t)blockIdx.x * ERR_ITEMS, i1 = min<int64_t>(a.n, i0 + ERR_ITEMS);
    const uint64_t kb = seg_of(i0 ? a.incl[i0 - 1] : 0, M, K);
    const uint64_t ke = seg_of((uint64_t)a.incl[i1 - 1] - 1, M, K);
    const bool local = ke - kb < ERR_LDS_SEGS;  // the same for the whole block
    if (local) {
        for (uint64_t j = threadIdx.x; j <= ke - kb; j += 256) lmax[j] = 0;
        __syncthreads();
    }
    auto add = [&](uint64_t k, uint32_t e) {
        if (e == 0) return;
        if (local) atomicMax(&lmax[k - kb], e);
        else atomicMax(&a.knot_err[k], e);
    };
    // Wave v takes items [iw0, iw1), 1024 of the block's, in rounds of 64.
    // When their slots all lie in one segment (the usual case), the
    // segment's knots are read once and every item costs two predictions.
    const int lane = threadIdx.x & 63;
    const int64_t iw0 = min<int64_t>(i1, i0 + (int64_t)(threadIdx.x >> 6) * (ERR_ITEMS / 4));
    const int64_t iw1 = min<int64_t>(i1, iw0 + ERR_ITEMS / 4);
    bool one_seg = false;
    if (iw0 < iw1) {
        const uint64_t kw = seg_of(iw0 ? a.incl[iw0 - 1] : 0, M, K);
        one_seg = kw == seg_of((uint64_t)a.incl[iw1 - 1] - 1, M, K);
        if (one_seg) {
            const uint64_t sk = knot_slot(kw, M, K), sk1 = knot_slot(kw + 1, M, K);
            const int32_t qk = a.knot_items[kw], qk1 = a.knot_items[kw + 1];
            uint32_t m = 0;
            for (int64_t ib = iw0; ib < iw1; ib += 64) {
                const int64_t i = ib + lane;
                if (i < iw1) {
                    const uint64_t f = i ? a.incl[i - 1] : 0, l = (uint64_t)a.incl[i] - 1;
                    const int64_t pf = seg_pred(f, sk, sk1, qk, qk1), pl = seg_pred(l, sk, sk1, qk, qk1);
                    m = max(m, (uint32_t)max(pf > i ? pf - i : i - pf, pl > i ? pl - i : i - pl));
                }
            }
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d));
            if (lane == 0) add(kw, m);
        }
    }
    for (int64_t ib = iw0; !one_seg && ib < iw1; ib += 64) {
        const int64_t i = ib + lane;
        const bool act = i < iw1;
        uint64_t k1 = 0, k2 = 0;
        uint32_t e1 = 0, e2 = 0;
        if (act) {
            const uint64_t f = i ? a.incl[i - 1] : 0, l = (uint64_t)a.incl[i] - 1;
            k1 = seg_of(f, M, K);
            k2 = seg_of(l, M, K);
            auto dev = [&](uint64_t k, uint64_t s) -> uint32_t {
                const int64_t p = seg_pred(s, knot_slot(k, M, K), knot_slot(k + 1, M, K), a.knot_items[k],
                                           a.knot_items[k + 1]);
                return (uint32_t)(p > i ? p - i : i - p);
            };
            const uint64_t end1 = (k1 == K - 1) ? M : knot_slot(k1 + 1, M, K) - 1;
            e1 = max(dev(k1, f), dev(k1, min(l, end1)));
            if (k2 != k1) e2 = max(dev(k2, knot_slot(k2, M, K)), dev(k2, l));
        }
        const uint64_t k0 = __shfl(k1, 0);
        const uint64_t mixed = __ballot(act && (k1 != k0 || k2 != k0));
        if (mixed == 0) {
            uint32_t m = max(e1, e2);
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d));
            if ((threadIdx.x & 63) == 0 && act) add(k0, m);
        } else if (act) {
            add(k1, e1);
            if (k2 != k1) add(k2, e2);
        }
    }
    if (local) {
        __syncthreads();
        for (uint64_t j = threadIdx.x; j <= ke - kb; j += 256)
            if (lmax[j]) atomicMax(&a.knot_err[kb + j], lmax[j]);
    }
}

// fillABinarySearch: a bisection over all of incl per slot.
template <typename T>
__global__ __launch_bounds__(FILL_THREADS) void rej_fill_bisect(RejArgs a) {
    if (a.st->error) return;
    const uint64_t s = (uint64_t)blockIdx.x * FILL_THREADS + threadIdx.x;
    if (s >= a.st->rej_size) return;
    const int32_t item = incl_bisect(a.incl, 0, a.n, s);
    fill_emit<T>(a, a.st->W, s, item, item ? a.incl[item - 1] : 0);
}

// fillAPArySearch with a wave of 64 consecutive slots [s0, s1] as the group
// and ballots in place of the LDS vote: lane g probes splitter j_g = lo +
// floor((hi - lo) g / 64); a probe with incl <= s0 puts every item of the
// group after it, one with incl > s1 puts them at or before it.  The interval
// [lo, hi] shrinks to the gap between the last of the first kind and the
// first of the second; the narrowing stops at 64 items, once the group
// straddles more than PARY_FILL_SPREAD gaps, or after PARY_FILL_ROUNDS rounds
// (the reference's loop has no exit when a block lies inside one run), then
// every lane bisects inside [lo, hi].
template <typename T>
__global__ __launch_bounds__(FILL_THREADS) void rej_fill_pary(RejArgs a) {
    if (a.st->error) return;
    const uint64_t A_size = a.st->rej_size;
    const int lane = threadIdx.x & 63;
    const uint64_t s0 = (uint64_t)blockIdx.x * FILL_THREADS + (uint64_t)__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    if (s0 >= A_size) return;
    const uint64_t s1 = min<uint64_t>(s0 + 63, A_size - 1);
    int32_t lo = 0, hi = (int32_t)(a.n - 1);
    for (int r = 0; r < PARY_FILL_ROUNDS && hi - lo >= 64; ++r) {
        const int32_t j = lo + (int32_t)(((int64_t)(hi - lo) * lane) >> 6);
        const uint64_t v = a.incl[j];
        const uint64_t after = __ballot(v <= s0);   // a prefix of the lanes
        const uint64_t before = __ballot(v > s1);   // a suffix
        const int na = __popcll(after);
        const int fb = before ? __ffsll((long long)before) - 1 : 64;
        const int32_t jl = __shfl(j, na > 0 ? na - 1 : 0), jh = __shfl(j, fb < 64 ? fb : 63);
        if (na > 0) lo = jl + 1;
        if (fb < 64) hi = jh;
        if (fb - na > PARY_FILL_SPREAD) break;
    }
    const uint64_t s = s0 + lane;
    if (s > s1) return;
    const int32_t item = incl_bisect(a.incl, lo, hi, s);
    fill_emit<T>(a, a.st->W, s, item, item ? a.incl[item - 1] : 0);
}

// The learned fill: per wave of 64 consecutive slots from s0, the model's
// window [max(q_k, pred - e_k), min(q_{k+1}, pred + e_k)] for the item of s0,
// p-ary rounds of 64 probes inside it down to 63 positions, then one
// verifying round over incl[lo - 1 .. lo + 62]: the item is the first
// position whose incl exceeds s0 with the one before at most s0, so a model
// that is wrong cannot produce a wrong item, only a miss.  A miss falls back
// to the bisection over all of incl (and is counted).  The 64 slots' items
// lie in [a0, a0 + 63] (runs are non-empty); the verifying round loads those
// incl too, and a search over the lanes by shuffles gives every lane its
// item.
template <typename T>
__global__ __launch_bounds__(FILL_THREADS) void rej_fill_learned(RejArgs a) {
    if (a.st->error) return;
    const uint64_t A_size = a.st->rej_size;
    const int lane = threadIdx.x & 63;
    const uint64_t s0 = (uint64_t)blockIdx.x * FILL_THREADS + (uint64_t)__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    if (s0 >= A_size) return;
    const int64_t n = a.n;
    const uint64_t M = A_size - 1, K = (uint64_t)a.K;
    const uint64_t k = seg_of(s0, M, K);
    const int32_t qk = a.knot_items[k], qk1 = a.knot_items[k + 1];
    const int64_t p = seg_pred(s0, knot_slot(k, M, K), knot_slot(k + 1, M, K), qk, qk1);
    const int64_t e = a.knot_err[k];
    int64_t lo = max<int64_t>(qk, p - e), hi = min<int64_t>(qk1, p + e);
    bool miss = false;
    for (int r = 0; r < LEARNED_FILL_ROUNDS && hi - lo >= 63; ++r) {
        // 64 runs of `stride` positions; probe g is the last position of run g.
        const int64_t stride = (hi - lo + 64) >> 6;
        const int64_t j = min<int64_t>(hi, lo + (lane + 1) * stride - 1);
        const uint64_t gt = __ballot((uint64_t)a.incl[j] > s0);
        if (gt == 0) {
            miss = true;  // the item lies after the window
            break;
        }
        const int m = __ffsll((long long)gt) - 1;
        const int64_t jm = __shfl(j, m), jp = __shfl(j, m > 0 ? m - 1 : 0);
        if (m > 0) lo = jp + 1;
        hi = jm;
    }
    // incl at position j, 0 before the first item, above every slot past the
    // last one.
    auto incl_at = [&](int64_t j) -> uint64_t { return j < 0 ? 0ull : (j >= n ? ~0ull : (uint64_t)a.incl[j]); };
    int64_t a0 = -1;
    uint64_t prev = 0;  // incl[a0 - 1], 0 for a0 = 0
    uint64_t v = 0;     // lane g: incl[a0 + g]
    if (!miss) {
        // incl[lo - 1 .. lo + 126]: the first half decides a0, and holds with
        // the second the 64 positions from a0 on that the lanes search.
        const int64_t j = lo - 1 + lane;
        const uint64_t v1 = incl_at(j), v2 = incl_at(j + 64);
        const uint64_t gt = __ballot(v1 > s0);
        const int m = gt ? __ffsll((long long)gt) - 1 : 0;
        if (m > 0) {
            a0 = lo - 1 + m;
            prev = __shfl(v1, m - 1);
            const int x = m + lane;
            const uint64_t t1 = __shfl(v1, x & 63), t2 = __shfl(v2, x & 63);
            v = x < 64 ? t1 : t2;
        }
    }
    if (a0 < 0) {
        a0 = incl_bisect(a.incl, 0, n, s0);
        prev = a0 ? a.incl[a0 - 1] : 0;
        v = incl_at(a0 + lane);
        if (lane == 0) atomicAdd((unsigned long long*)&a.st->learned_fallbacks, 1ull);
    }
    const uint64_t s = s0 + lane;
    // c = number of the window's incl values <= s (they increase; the last
    // exceeds s0 + 63), so the item of s is a0 + c.
    int c = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const uint64_t t = __shfl(v, c + step - 1);
        if (t <= s) c += step;
    }
    const uint64_t vc = __shfl(v, c > 0 ? c - 1 : 0);
    if (s < A_size) fill_emit<T>(a, a.st->W, s, (int32_t)(a0 + c), c > 0 ? vc : prev);
}

hipError_t launch_rejection_incl(const RejArgs& a, bool model, hipStream_t s) {
    const unsigned nb = (unsigned)a.nb;
    if (a.dtype_f32) hipLaunchKernelGGL(rej_incl<float>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    else hipLaunchKernelGGL(rej_incl<double>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    if (model) {
        hipLaunchKernelGGL(rej_model_knots, dim3((unsigned)((a.K + 1 + 255) / 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(rej_model_err, dim3((unsigned)((a.n + ERR_ITEMS - 1) / ERR_ITEMS)), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_rejection_search_fill(const RejArgs& a, int fill, hipStream_t s) {
    // Blocks for every slot A can hold; the blocks past |A| return at once.
    const dim3 g((unsigned)((a.A_cap + FILL_THREADS - 1) / FILL_THREADS)), b(FILL_THREADS);
    switch (fill) {
        case REJ_FILL_BISECT:
            if (a.dtype_f32) hipLaunchKernelGGL(rej_fill_bisect<float>, g, b, 0, s, a);
            else hipLaunchKernelGGL(rej_fill_bisect<double>, g, b, 0, s, a);
            break;
        case REJ_FILL_PARY:
            if (a.dtype_f32) hipLaunchKernelGGL(rej_fill_pary<float>, g, b, 0, s, a);
            else hipLaunchKernelGGL(rej_fill_pary<double>, g, b, 0, s, a);
            break;
        case REJ_FILL_LEARNED:
            if (a.dtype_f32) hipLaunchKernelGGL(rej_fill_learned<float>, g, b, 0, s, a);
            else hipLaunchKernelGGL(rej_fill_learned<double>, g, b, 0, s, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace atg
