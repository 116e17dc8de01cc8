// Alias-table construction kernels for gfx950 (the hot path of
// AliasTableSplitGpu::build, AliasTableSplitGpu.cu:388-468).
//
// Phases (one stream, no host round trips):
//   1. sum_sections   : per-section FP64 sums (one wave per section) + weight
//                       validation + reset of the look-back slots.
//   2. finalize_sum   : one workgroup reduces the section sums with the
//                       canonical 64-ary pairwise tree -> W, W/N.
//   3. psa_sections   : one wave per section, taken in ticket order: stable
//                       light/heavy partition in LDS, PSA+ greedy chains that
//                       write finished rows, leftover compaction, the canonical
//                       local scan of the leftover weights and a decoupled
//                       look-back that folds predecessors left to right (bit
//                       identical to the chained prefix), leftover lists +
//                       prefix sums, the split certificate and the sentinels.
//   4. degenerate_fill: only when one leftover class is empty.
//   5. split          : bisection per k (SplitMethodBasic.cu:14-50).
//   6. pack           : one chain per k (PackMethodBasic.cuh:59-126).
#include "atg_build.h"
#include "atg_device.h"

namespace atg {

// ---------------------------------------------------------------------------
// 1. Section sums.  Lane l of the wave that owns a section adds items
//    l, l+64, ... in order; the lane sums meet in the pairwise tree.
//    A NaN / infinite / negative weight turns the section sum into NaN.
template <typename T>
__global__ __launch_bounds__(256) void sum_sections(BuildArgs a) {
    const int lane = threadIdx.x & 63;
    const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= a.G) return;
    const T* w = (const T*)a.weights;
    const int64_t start = min(a.n, s * a.Ls), end = min(a.n, (s + 1) * a.Ls);
    double acc = 0.0;
    bool bad = false;
    int64_t k = start + lane;
    for (; k + 7 * 64 < end; k += 8 * 64) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = (double)w[k + u * 64];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            bad |= !(x[u] >= 0.0) || (x[u] == __builtin_inf());
            acc = acc + x[u];
        }
    }
    for (; k < end; k += 64) {
        const double x = (double)w[k];
        bad |= !(x >= 0.0) || (x == __builtin_inf());
        acc = acc + x;
    }
    double sum = wave_tree_sum(acc);
    if (__ballot(bad) != 0ull) sum = __builtin_nan("");
    if (lane == 0) a.section_sums[s] = sum;
    if (lane < 6) a.slots[(int64_t)lane * a.G + s] = ATG_NOT_READY;
}

// ---------------------------------------------------------------------------
// 2. W = recursive 64-ary pairwise tree over the section sums.
__global__ __launch_bounds__(1024) void finalize_sum(BuildArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const double* src = a.section_sums;
    double* bufs[2] = {a.tree_scratch, a.tree_scratch + a.tree_scratch_half};
    int flip = 0;
    int64_t count = a.G;
    while (count > 1) {
        const int64_t groups = (count + 63) / 64;
        double* dst = bufs[flip];
        for (int64_t g = wave; g < groups; g += nwaves) {
            const int64_t idx = g * 64 + lane;
            const double v = (idx < count) ? src[idx] : 0.0;
            const double r = wave_tree_sum(v);
            if (lane == 0) dst[g] = r;
        }
        __threadfence_block();
        __syncthreads();
        src = dst;
        flip ^= 1;
        count = groups;
    }
    if (threadIdx.x == 0) {
        const double W = src[0];
        DevState* st = a.st;
        uint32_t err = 0;
        if (W != W) err = ERR_BAD_WEIGHT;
        else if (!(W > 0.0) || W == __builtin_inf()) err = ERR_BAD_SUM;
        if (a.Ls > a.Ls_max) err |= ERR_SECTION_TOO_LONG;
        st->W = W;
        st->W_N = W / (double)a.n;
        st->error = err;
        st->ticket = 0;
        st->degenerate = 0;
        st->cert_max_dL = 0ull;                    // +0.0
        st->cert_min_dH = 0x7FF0000000000000ull;   // +inf
        st->split_fallbacks = 0;
        st->learned_fallbacks = 0;
    }
}

// ---------------------------------------------------------------------------
// 3. PSA+ / plain partition per section.
struct SecLDS {
    double* w;      // [Ls]: lights from the front, heavies from the back
    uint16_t* idx;  // [Ls]: local item index
    int len;
    __device__ double& Lw(int i) { return w[i]; }
    __device__ uint16_t& Li(int i) { return idx[i]; }
    __device__ double& Hw(int j) { return w[len - 1 - j]; }
    __device__ uint16_t& Hi(int j) { return idx[len - 1 - j]; }
};

template <bool AOS>
__device__ __forceinline__ void store_row(const BuildArgs& a, int64_t item, double rw, int32_t alias) {
    if (AOS) {
        Row r;
        r.weight = rw;
        r.alias = alias;
        r.pad = 0;
        *reinterpret_cast<uint4*>(&a.rows[item]) = *reinterpret_cast<const uint4*>(&r);
    } else {
        a.tw[item] = rw;
        a.ta[item] = alias;
    }
}

// Rank r of the section's leftover list -> LDS position, given per-chain
// (offset, first position) held by lanes t < nch.
__device__ __forceinline__ int leftover_pos(int r, int off, int first, int nch) {
    int t = 0;
    for (int tt = 1; tt < nch; ++tt)
        if (__shfl(off, tt, 64) <= r) t = tt;
    return __shfl(first, t, 64) + (r - __shfl(off, t, 64));
}

// Canonical local scan aggregate (chunks of 64, Kogge-Stone, carries chained).
template <bool HEAVY>
__device__ double local_aggregate(SecLDS& L, int m, int off, int first, int nch) {
    double carry = 0.0;
    for (int r0 = 0; r0 < m; r0 += 64) {
        const int r = r0 + threadIdx.x;
        double v = 0.0;
        const int pos = leftover_pos(r < m ? r : 0, off, first, nch);  // wave-uniform shuffles
        if (r < m) v = HEAVY ? L.Hw(pos) : L.Lw(pos);
        const double inc = wave_ks_inclusive(v);
        carry = carry + shfl_d(inc, 63);
    }
    return carry;
}

// Writes the section's leftover list with global prefixes P + local, and
// returns the lane-local extreme increment for the split certificate.
template <bool HEAVY>
__device__ double write_leftovers(const BuildArgs& a, SecLDS& L, int64_t start, int m, int off, int first,
                                  int nch, int64_t goff, double P, double agg) {
    int32_t* oi = HEAVY ? a.Hi : a.Li;
    double* ow = HEAVY ? a.Hw : a.Lw;
    double* op = HEAVY ? a.Hp : a.Lp;
    double ext = HEAVY ? __builtin_inf() : 0.0;
    double carry = 0.0;
    const int lane = threadIdx.x;
    for (int r0 = 0; r0 < m; r0 += 64) {
        const int r = r0 + lane;
        double v = 0.0;
        int pos = leftover_pos(r < m ? r : 0, off, first, nch);
        if (r < m) v = HEAVY ? L.Hw(pos) : L.Lw(pos);
        const double inc = wave_ks_inclusive(v);
        const double prev = shfl_up_d(inc, 1);
        const double local = carry + (lane == 0 ? 0.0 : prev);
        const double pref = P + local;
        const double total = shfl_d(inc, 63);
        const double carry_next = carry + total;
        double succ = __shfl_down(pref, 1, 64);
        if (lane == 63) succ = P + carry_next;
        if (r == m - 1) succ = P + agg;
        if (r < m) {
            const int item = (int)(HEAVY ? L.Hi(pos) : L.Li(pos));
            oi[goff + r] = (int32_t)(start + item);
            ow[goff + r] = v;
            op[goff + r] = pref;
            const double d = succ - pref;
            ext = HEAVY ? fmin(ext, d) : fmax(ext, d);
        }
        carry = carry_next;
    }
    return ext;
}

// Look-back: returns the exclusive prefix of section s (packed counts and the
// two FP64 prefixes), folding from the nearest inclusive predecessor left to
// right so the bits equal the chained definition P_{s+1} = P_s + agg_s.
struct Prefix {
    uint64_t c;
    double l, h;
};

__device__ Prefix look_back(const BuildArgs& a, int64_t s, bool& timeout) {
    const int lane = threadIdx.x;
    const uint64_t* aggC = a.slots + 0 * a.G;
    const uint64_t* aggL = a.slots + 1 * a.G;
    const uint64_t* aggH = a.slots + 2 * a.G;
    const uint64_t* incC = a.slots + 3 * a.G;
    const uint64_t* incL = a.slots + 4 * a.G;
    const uint64_t* incH = a.slots + 5 * a.G;
    int64_t hi = s - 1;  // newest predecessor of the current window
    uint32_t spins = 0;
    // Phase 1: find the nearest predecessor with an inclusive value such that
    // every later predecessor has its aggregate.
    int64_t qstar = -1;
    bool first_window = true;
    uint64_t vc = 0, vl = 0, vh = 0;  // this lane's values (window of hi)
    int h = 0;
    while (true) {
        const int64_t q = hi - 63 + lane;
        int status;  // 2 inclusive, 1 aggregate, 0 none
        if (q < 0) {
            status = 2; vc = 0; vl = d2u(0.0); vh = d2u(0.0);
        } else {
            vc = slot_load(incC + q); vl = slot_load(incL + q); vh = slot_load(incH + q);
            if (vc != ATG_NOT_READY && vl != ATG_NOT_READY && vh != ATG_NOT_READY) {
                status = 2;
            } else {
                vc = slot_load(aggC + q); vl = slot_load(aggL + q); vh = slot_load(aggH + q);
                status = (vc != ATG_NOT_READY && vl != ATG_NOT_READY && vh != ATG_NOT_READY) ? 1 : 0;
            }
        }
        const uint64_t incl_mask = __ballot(status == 2);
        const uint64_t none_mask = __ballot(status == 0);
        if (incl_mask != 0ull) {
            h = 63 - __clzll(incl_mask);
            const uint64_t above = (h == 63) ? 0ull : (none_mask >> (h + 1));
            if (above == 0ull) { qstar = hi - 63 + h; break; }
        } else if (none_mask == 0ull) {
            hi -= 64;
            first_window = false;
            continue;
        }
        if (++spins > (1u << 22)) { timeout = true; Prefix z{0, 0.0, 0.0}; return z; }
        __builtin_amdgcn_s_sleep(2);
    }
    // Phase 2: fold left to right from q*.
    Prefix acc;
    acc.c = __shfl(vc, h, 64);
    acc.l = u2d(__shfl(vl, h, 64));
    acc.h = u2d(__shfl(vh, h, 64));
    if (first_window) {
        for (int t = h + 1; t < 64; ++t) {
            acc.c += __shfl(vc, t, 64);
            acc.l = acc.l + u2d(__shfl(vl, t, 64));
            acc.h = acc.h + u2d(__shfl(vh, t, 64));
        }
        return acc;
    }
    // q* lies in an older window: re-read the (already published) aggregates.
    for (int64_t cur = qstar + 1; cur < s; cur += 64) {
        const int64_t q = cur + lane;
        uint64_t c = 0, l = 0, hh = 0;
        if (q < s) { c = slot_load(aggC + q); l = slot_load(aggL + q); hh = slot_load(aggH + q); }
        const int cnt = (int)min<int64_t>(64, s - cur);
        for (int t = 0; t < cnt; ++t) {
            acc.c += __shfl(c, t, 64);
            acc.l = acc.l + u2d(__shfl(l, t, 64));
            acc.h = acc.h + u2d(__shfl(hh, t, 64));
        }
    }
    return acc;
}

template <typename T, bool AOS>
__global__ __launch_bounds__(64) void psa_sections(BuildArgs a) {
    extern __shared__ double lds_raw[];
    DevState* st = a.st;
    if (st->error) return;
    const int lane = threadIdx.x;
    uint32_t sid = 0;
    if (lane == 0) sid = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int64_t s = (int64_t)__shfl(sid, 0, 64);
    if (s >= a.G) return;
    const T* w = (const T*)a.weights;
    const int64_t start = min(a.n, s * a.Ls), end = min(a.n, (s + 1) * a.Ls);
    const int len = (int)(end - start);
    const double W_N = st->W_N;
    SecLDS L;
    L.w = lds_raw;
    L.idx = reinterpret_cast<uint16_t*>(lds_raw + a.Ls);
    L.len = len;

    // Stable partition: lights to the front in index order, heavies to the
    // back (h[k] at len-1-k) in index order.
    int cntL = 0, cntH = 0;
    for (int c0 = 0; c0 < len; c0 += 4 * 64) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = c0 + u * 64 + lane;
            x[u] = (k < len) ? (double)w[start + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = c0 + u * 64 + lane;
            const bool valid = k < len;
            const bool light = valid && x[u] <= W_N;
            const bool heavy = valid && !(x[u] <= W_N);
            const uint64_t mL = __ballot(light), mH = __ballot(heavy);
            const uint64_t lt = lanemask_lt();
            if (light) {
                const int pos = cntL + __popcll(mL & lt);
                L.Lw(pos) = x[u];
                L.Li(pos) = (uint16_t)k;
            }
            if (heavy) {
                const int pos = cntH + __popcll(mH & lt);
                L.Hw(pos) = x[u];
                L.Hi(pos) = (uint16_t)k;
            }
            cntL += __popcll(mL);
            cntH += __popcll(mH);
        }
    }
    __syncthreads();
    const int nL = cntL, nH = cntH;

    // Greedy chains (PsaPlusNormal.cu:51-123) or one chain holding everything.
    int nch, iLo, iHi, jLo, jHi;
    if (a.psa_plus) {
        const int C = a.chains;
        nch = C;
        const int stepsL = nL / C, stepsH = nH / C;
        const int t = lane;
        iLo = t * stepsL; iHi = (t == C - 1) ? nL : (t + 1) * stepsL;
        jLo = t * stepsH; jHi = (t == C - 1) ? nH : (t + 1) * stepsH;
        if (t >= C) { iLo = iHi = jLo = jHi = 0; }
        const bool active = t < C && iHi - iLo > 0 && jHi - jLo > 0;
        if (active) {
            int i = iLo, j = jLo;
            double wc = L.Hw(j);
            while (true) {
                const bool ex = wc <= W_N + ATG_EPSILON;
                if ((ex && j >= jHi - 1) || (!ex && i >= iHi - 1)) break;
                int item, alias;
                double rw, next;
                if (ex) {
                    item = L.Hi(j); rw = wc; alias = L.Hi(j + 1); next = L.Hw(j + 1); ++j;
                } else {
                    item = L.Li(i); rw = L.Lw(i); alias = L.Hi(j); next = L.Lw(i); ++i;
                }
                wc += next - W_N;
                store_row<AOS>(a, start + item, rw, (int32_t)(start + alias));
            }
            // Remainder (PsaPlusNormal.cu:112-123): back to L when its residual
            // is light and the chain consumed a light slot to hold it, else
            // back to H in place.
            if (wc <= W_N + ATG_EPSILON && i > iLo) {
                L.Lw(i - 1) = wc;
                L.Li(i - 1) = L.Hi(j);
                iLo = i - 1;
                jLo = j + 1;
            } else {
                L.Hw(j) = wc;
                iLo = i;
                jLo = j;
            }
        }
    } else {
        nch = 1;
        iLo = 0; iHi = (lane == 0) ? nL : 0;
        jLo = 0; jHi = (lane == 0) ? nH : 0;
    }
    __syncthreads();

    const int cl = iHi - iLo, ch = jHi - jLo;
    const int offL = wave_excl_scan_int(cl), offH = wave_excl_scan_int(ch);
    const int mL = wave_sum_int(cl), mH = wave_sum_int(ch);

    const double aggL = local_aggregate<false>(L, mL, offL, iLo, nch);
    const double aggH = local_aggregate<true>(L, mH, offH, jLo, nch);
    const uint64_t aggC = (uint64_t)(uint32_t)mL | ((uint64_t)(uint32_t)mH << 32);

    uint64_t* incC = a.slots + 3 * a.G;
    uint64_t* incLs = a.slots + 4 * a.G;
    uint64_t* incHs = a.slots + 5 * a.G;
    Prefix P{0, 0.0, 0.0};
    if (s > 0) {
        if (lane == 0) {
            slot_store(a.slots + 0 * a.G + s, aggC);
            slot_store(a.slots + 1 * a.G + s, d2u(aggL));
            slot_store(a.slots + 2 * a.G + s, d2u(aggH));
        }
        bool timeout = false;
        P = look_back(a, s, timeout);
        if (timeout && lane == 0) atomicOr(&st->error, ERR_LOOKBACK_TIMEOUT);
    }
    const uint64_t inclC = P.c + aggC;
    const double inclL = P.l + aggL, inclH = P.h + aggH;
    if (lane == 0) {
        slot_store(incC + s, inclC);
        slot_store(incLs + s, d2u(inclL));
        slot_store(incHs + s, d2u(inclH));
    }

    const int64_t goffL = (int64_t)(uint32_t)P.c, goffH = (int64_t)(P.c >> 32);
    const double dL = write_leftovers<false>(a, L, start, mL, offL, iLo, nch, goffL, P.l, aggL);
    const double dH = write_leftovers<true>(a, L, start, mH, offH, jLo, nch, goffH, P.h, aggH);
    double mx = dL, mn = dH;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        mx = fmax(mx, shfl_xor_d(mx, m));
        mn = fmin(mn, shfl_xor_d(mn, m));
    }
    if (lane == 0) {
        if (mL > 0) atomicMax((unsigned long long*)&st->cert_max_dL, (unsigned long long)d2u(mx));
        if (mH > 0) atomicMin((unsigned long long*)&st->cert_min_dH, (unsigned long long)d2u(mn));
    }

    if (s == a.G - 1 && lane == 0) {
        // Sentinels (AliasTableSplitGpu.cu:366-385) and the split inputs.
        const int64_t nLr = (int64_t)(uint32_t)inclC, nHr = (int64_t)(inclC >> 32);
        a.Li[nLr] = (int32_t)(a.n + 1); a.Lw[nLr] = 0.0; a.Lp[nLr] = inclL;
        a.Hi[nHr] = (int32_t)a.n; a.Hw[nHr] = __builtin_inf(); a.Hp[nHr] = inclH;
        const int64_t numL = nLr + 1, numH = nHr + 1;
        st->numL = numL;
        st->numH = numH;
        st->degenerate = (nLr == 0 || nHr == 0) ? 1u : 0u;
        int64_t p = a.p_req;
        if (a.psa_plus) { p = min(p, (numL + numH) / 10); if (p < 2) p = 2; }
        st->p = st->degenerate ? 0 : p;
        st->nToSplit = numL + numH - 2;
        st->wToSplit = inclL + inclH;
    }
}

// ---------------------------------------------------------------------------
// 4. Degenerate tables: every leftover row carries its own mass.
template <bool AOS>
__global__ void degenerate_fill(BuildArgs a) {
    const DevState* st = a.st;
    if (st->error || !st->degenerate) return;
    const int64_t nLr = st->numL - 1, nHr = st->numH - 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nLr + nHr;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t item = (t < nLr) ? a.Li[t] : a.Hi[t - nLr];
        const double w = (t < nLr) ? a.Lw[t] : a.Hw[t - nLr];
        store_row<AOS>(a, item, w, item);
    }
}

// ---------------------------------------------------------------------------
// 5. Split: bisection per k with the reference's predicate; 64 probes bound
//    the loop (the reference spins forever when FOUND is unreachable).
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

__global__ __launch_bounds__(256) void split_bisect(BuildArgs a) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k == 1) {
        a.splits[0] = Split{0, 0, 0.0};
        a.splits[p] = Split{(int32_t)(numL - 1), (int32_t)(numH - 1), 0.0};
    }
    if (k >= p) return;
    Split sp;
    const bool ok = bisect_split(k, p, st->nToSplit, st->wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, sp);
    a.splits[k] = sp;
    if (!ok) atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
}

// ---------------------------------------------------------------------------
// 6. Pack one section per thread (PackMethodBasic.cuh:59-126).
template <bool AOS>
__global__ __launch_bounds__(256) void pack_basic(BuildArgs a) {
    const DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > p) return;
    const double W_N = st->W_N;
    const Split prev = a.splits[k - 1], cur = a.splits[k];
    const int iUpper = cur.i, jUpper = cur.j;
    int i = prev.i, j = prev.j;
    double w = prev.spill;
    if (prev.spill == 0) w = a.Hw[j];
    while (true) {
        bool ex = w <= W_N + ATG_EPSILON;
        if (ex && j >= jUpper) ex = false;
        else if (!ex && i >= iUpper) ex = true;
        if ((ex && j >= jUpper) || (!ex && i >= iUpper)) break;
        int32_t item, alias;
        double rw, next;
        if (ex) {
            item = a.Hi[j]; rw = w; alias = a.Hi[j + 1]; next = a.Hw[j + 1]; ++j;
        } else {
            item = a.Li[i]; rw = a.Lw[i]; alias = a.Hi[j]; next = a.Lw[i]; ++i;
        }
        w += next - W_N;
        if ((uint32_t)item < (uint32_t)a.n) store_row<AOS>(a, item, rw, alias);
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
static inline unsigned blocks_for(int64_t items, int per) { return (unsigned)((items + per - 1) / per); }

hipError_t launch_sum(const BuildArgs& a, hipStream_t s) {
    const unsigned g = blocks_for(a.G, 4);
    if (a.dtype_f32) hipLaunchKernelGGL(sum_sections<float>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(sum_sections<double>, dim3(g), dim3(256), 0, s, a);
    hipLaunchKernelGGL(finalize_sum, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

size_t psa_lds_bytes(int64_t Ls_max) { return (size_t)Ls_max * 8 + (size_t)Ls_max * 2 + 16; }

hipError_t launch_partition(const BuildArgs& a, hipStream_t s) {
    const size_t lds = psa_lds_bytes(a.Ls_max);
    const dim3 g((unsigned)a.G), b(64);
    if (a.dtype_f32) {
        if (a.rows) hipLaunchKernelGGL((psa_sections<float, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((psa_sections<float, false>), g, b, lds, s, a);
    } else {
        if (a.rows) hipLaunchKernelGGL((psa_sections<double, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((psa_sections<double, false>), g, b, lds, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_degenerate(const BuildArgs& a, hipStream_t s) {
    if (a.rows) hipLaunchKernelGGL(degenerate_fill<true>, dim3(1024), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(degenerate_fill<false>, dim3(1024), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_split(const BuildArgs& a, hipStream_t s) {
    hipError_t e = launch_degenerate(a, s);
    if (e != hipSuccess) return e;
    const int64_t kmax = a.p_req;  // effective p <= p_req
    if (kmax >= 2) hipLaunchKernelGGL(split_bisect, dim3(blocks_for(kmax - 1, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pack(const BuildArgs& a, hipStream_t s) {
    const int64_t kmax = a.p_req;
    if (a.rows) hipLaunchKernelGGL(pack_basic<true>, dim3(blocks_for(kmax, 256)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(pack_basic<false>, dim3(blocks_for(kmax, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t set_partition_lds_limit(int64_t Ls_max) {
    const int bytes = (int)psa_lds_bytes(Ls_max);
    hipError_t e = hipSuccess;
    e = hipFuncSetAttribute((const void*)psa_sections<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)psa_sections<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)psa_sections<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)psa_sections<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace atg
