// Alias-table construction kernels for gfx950 (the hot path of
// AliasTableSplitGpu::build, AliasTableSplitGpu.cu:388-468).
//
// Phases (one stream, no host round trips):
//   1. sum_sections : per-section FP64 sums (64 sections per block), weight
//                     validation, the first level of the canonical 64-ary tree;
//                     the last block to finish completes the tree -> W, W/N.
//   2. psa_wave     : one wave per section, sections taken in ticket order:
//                     all loads issued at once, stable light/heavy partition
//                     into LDS, PSA+ greedy chains writing finished rows,
//                     leftover maps, the canonical local scan of the leftover
//                     weights, the section prefix from dyadic block sums that
//                     the sections complete cooperatively (O(log G) depth, the
//                     association the oracle defines), leftover lists + prefix
//                     sums, the split certificate and the sentinels.
//   3. degenerate_fill (only when one leftover class is empty), split, pack.
#include "atg_build.h"
#include "atg_device.h"
#include "atg_split.h"

namespace atg {

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------------------
// 1. Section sums.  Lane l of the wave that owns a section adds items
//    l, l+64, ... in order; the lane sums meet in the pairwise tree.  A NaN,
//    infinite or negative weight turns the section sum into NaN.
template <typename T>
__device__ __forceinline__ double section_sum(const T* w, int64_t start, int64_t end, int lane) {
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
    return sum;
}

// Block b owns sections [64 b, 64 b + 64) (16 waves x 4 sections) and writes
// the tree of their sums (level 1 of the 64-ary tree).  The last block to
// finish reduces the level-1 values and publishes W.
template <typename T>
__global__ __launch_bounds__(1024) void sum_sections(BuildArgs a) {
    __shared__ double sh[64];
    __shared__ uint32_t sh_last;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const T* w = (const T*)a.weights;
    for (int q = 0; q < 4; ++q) {
        const int idx = wave * 4 + q;
        const int64_t s = (int64_t)blockIdx.x * 64 + idx;
        double sum = 0.0;
        if (s < a.G) sum = section_sum<T>(w, min(a.n, s * a.Ls), min(a.n, (s + 1) * a.Ls), lane);
        if (lane == 0) sh[idx] = sum;
    }
    // Reset the dyadic block sums (3 values per block) and their arrival
    // counters for the partition pass.
    const int64_t nvals = 3 * a.tree_total, ncnt = (a.tree_total + 1) / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvals + ncnt;
         i += (int64_t)gridDim.x * blockDim.x)
        a.slots[i] = (i < nvals) ? ATG_NOT_READY : 0ull;
    __syncthreads();
    if (wave == 0) {
        const double r = wave_tree_sum(sh[lane]);
        if (lane == 0) {
            a.tree_scratch[blockIdx.x] = r;
            // Release: the level-1 value is visible before the counter moves.
            const uint32_t done = __hip_atomic_fetch_add(&a.st->sum_done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            sh_last = (done == gridDim.x - 1) ? 1u : 0u;
        }
    }
    __syncthreads();
    if (!sh_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // Last block: the remaining levels of the tree over tree_scratch.
    double* bufs[2] = {a.tree_scratch, a.tree_scratch + a.tree_scratch_half};
    int cur = 0;
    int64_t count = gridDim.x;
    const int nwaves = blockDim.x >> 6;
    while (count > 1) {
        const int64_t groups = (count + 63) / 64;
        const double* src = bufs[cur];
        double* dst = bufs[cur ^ 1];
        for (int64_t g = wave; g < groups; g += nwaves) {
            const int64_t i = g * 64 + lane;
            const double r = wave_tree_sum((i < count) ? src[i] : 0.0);
            if (lane == 0) dst[g] = r;
        }
        __threadfence_block();
        __syncthreads();
        cur ^= 1;
        count = groups;
    }
    if (threadIdx.x == 0) {
        const double W = bufs[cur][0];
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
        st->sum_done = 0;
    }
}

// ---------------------------------------------------------------------------
// 2. PSA+ / plain partition: one wave per section, sections taken in ticket
//    order.  The section prefixes of the leftover sums use the dyadic-block
//    association of the oracle: each section publishes its aggregate as the
//    level-0 block, the child that completes a pair second publishes the
//    parent, and P(s) folds the blocks of s's binary digits from high to low.
#define PSA_MAX_LEN 2048

struct SecLDS {
    double* w;      // [Ls]: lights from the front, heavies from the back
    uint16_t* idx;  // [Ls]: local item index
    uint16_t* map;  // [Ls]: leftover rank -> position (L ranks, then H ranks)
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

struct Prefix {
    uint64_t c;  // packed counts: lights | heavies << 32
    double l, h;
};

// Dyadic block sums: level b holds G >> b blocks of 2^b sections.
__device__ __forceinline__ int64_t level_offset(int64_t G, int b) {
    int64_t off = 0;
    for (int q = 0; q < b; ++q) off += G >> q;
    return off;
}

__device__ __forceinline__ void block_load(const BuildArgs& a, int64_t idx, uint64_t& c, uint64_t& l, uint64_t& h,
                                           bool& timeout) {
    const uint64_t* V = a.slots;
    uint32_t spins = 0;
    while (true) {
        c = slot_load(V + idx);
        l = slot_load(V + a.tree_total + idx);
        h = slot_load(V + 2 * a.tree_total + idx);
        if (c != ATG_NOT_READY && l != ATG_NOT_READY && h != ATG_NOT_READY) return;
        if (++spins > (1u << 22)) { timeout = true; c = 0; l = d2u(0.0); h = d2u(0.0); return; }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ void block_store(const BuildArgs& a, int64_t idx, uint64_t c, double l, double h) {
    uint64_t* V = a.slots;
    slot_store(V + idx, c);
    slot_store(V + a.tree_total + idx, d2u(l));
    slot_store(V + 2 * a.tree_total + idx, d2u(h));
}

// Publish section s's aggregate and complete the parents for which it is
// the second child to arrive (single lane).
__device__ void block_publish(const BuildArgs& a, int64_t s, uint64_t c, double l, double h, bool& timeout) {
    const int64_t G = a.G;
    block_store(a, s, c, l, h);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(a.slots + 3 * a.tree_total);
    int64_t child = s, off_child = 0, off_parent = G;
    for (int b = 1; ((int64_t)1 << b) <= G; ++b) {
        const int64_t parent = child >> 1;
        if (((parent + 1) << b) > G) break;  // incomplete block: never needed
        const uint32_t old = __hip_atomic_fetch_add(cnt + off_parent + parent, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) break;  // the sibling completes this block
        uint64_t lc, ll, lh, rc, rl, rh;
        block_load(a, off_child + 2 * parent, lc, ll, lh, timeout);
        block_load(a, off_child + 2 * parent + 1, rc, rl, rh, timeout);
        c = lc + rc;
        l = u2d(ll) + u2d(rl);
        h = u2d(lh) + u2d(rh);
        block_store(a, off_parent + parent, c, l, h);
        child = parent;
        off_child = off_parent;
        off_parent += G >> b;
    }
}

// P(s): lane b loads the block of bit b of s, then the blocks are folded from
// the highest bit to the lowest.
__device__ Prefix block_prefix(const BuildArgs& a, int64_t s, int lane, bool& timeout) {
    uint64_t c = 0, l = d2u(0.0), h = d2u(0.0);
    const bool set = lane < 62 && ((s >> lane) & 1);
    if (set) {
        const int64_t pos = (s >> (lane + 1)) << (lane + 1);
        block_load(a, level_offset(a.G, lane) + (pos >> lane), c, l, h, timeout);
    }
    const uint64_t bits = __ballot(set);
    Prefix P{0, 0.0, 0.0};
    for (int b = 63; b >= 0; --b) {
        if (!((bits >> b) & 1)) continue;
        P.c += readlane_u64(c, b);
        P.l = P.l + u2d(readlane_u64(l, b));
        P.h = P.h + u2d(readlane_u64(h, b));
    }
    return P;
}

template <typename T, bool AOS>
__global__ __launch_bounds__(64) void psa_wave(BuildArgs a) {
    extern __shared__ double lds_raw[];
    DevState* st = a.st;
    if (st->error) return;
    const int lane = threadIdx.x;
    uint32_t sid = 0;
    if (lane == 0) sid = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int64_t s = (int64_t)__shfl(sid, 0, 64);
    if (s >= a.G) return;
    const T* w = (const T*)a.weights;
    const int64_t start = min(a.n, s * a.Ls);
    const int len = (int)(min(a.n, (s + 1) * a.Ls) - start);
    const double W_N = st->W_N;
    SecLDS L;
    L.w = lds_raw;
    L.idx = reinterpret_cast<uint16_t*>(lds_raw + a.Ls_max);
    L.map = L.idx + a.Ls_max;
    L.len = len;
    const uint64_t lt = lanemask_lt();

    // All loads of the section in flight at once, then the stable partition:
    // lights to the front in index order, heavies to the back (h[k] at
    // len-1-k) in index order.
    double x[PSA_MAX_LEN / 64];
#pragma unroll
    for (int u = 0; u < PSA_MAX_LEN / 64; ++u) {
        const int k = (u << 6) + lane;
        x[u] = (k < len) ? (double)w[start + k] : 0.0;
    }
    int nL = 0, nH = 0;
#pragma unroll
    for (int u = 0; u < PSA_MAX_LEN / 64; ++u) {
        const int k = (u << 6) + lane;
        if ((u << 6) < len) {
            const bool valid = k < len;
            const bool light = valid && x[u] <= W_N, heavy = valid && !(x[u] <= W_N);
            const uint64_t bl = __ballot(light), bh = __ballot(heavy);
            if (light) {
                const int pos = nL + __popcll(bl & lt);
                L.Lw(pos) = x[u];
                L.Li(pos) = (uint16_t)k;
            }
            if (heavy) {
                const int pos = nH + __popcll(bh & lt);
                L.Hw(pos) = x[u];
                L.Hi(pos) = (uint16_t)k;
            }
            nL += __popcll(bl);
            nH += __popcll(bh);
        }
    }
    wave_lds_sync();

    int iLo = 0, iHi = 0, jLo = 0, jHi = 0;
    if (a.psa_plus) {
        // Greedy chains (PsaPlusNormal.cu:51-123), lane t runs chain t.
        const int C = a.chains;
        const int t = lane;
        if (t < C) {
            const int stepsL = nL / C, stepsH = nH / C;
            iLo = t * stepsL; iHi = (t == C - 1) ? nL : (t + 1) * stepsL;
            jLo = t * stepsH; jHi = (t == C - 1) ? nH : (t + 1) * stepsH;
        }
        if (t < C && iHi - iLo > 0 && jHi - jLo > 0) {
            int i = iLo, j = jLo;
            double wc = L.Hw(j);
            int hj = L.Hi(j);
            double lw = L.Lw(i);
            int li = L.Li(i);
            const double thr = W_N + ATG_EPSILON;
            while (true) {
                const bool ex = wc <= thr;
                if ((ex && j >= jHi - 1) || (!ex && i >= iHi - 1)) break;
                const double hw1 = L.Hw(j + 1);
                const int hi1 = L.Hi(j + 1);
                const int item = ex ? hj : li;
                const double rw = ex ? wc : lw;
                const int alias = ex ? hi1 : hj;
                const double next = ex ? hw1 : lw;
                wc += next - W_N;
                store_row<AOS>(a, start + item, rw, (int32_t)(start + alias));
                j += ex ? 1 : 0;
                i += ex ? 0 : 1;
                hj = ex ? hi1 : hj;
                lw = L.Lw(i);
                li = L.Li(i);
            }
            // Remainder (PsaPlusNormal.cu:112-123): back to L when its
            // residual is light and a consumed light slot can hold it,
            // otherwise back to H in place.
            if (wc <= thr && i > iLo) {
                L.Lw(i - 1) = wc;
                L.Li(i - 1) = (uint16_t)hj;
                iLo = i - 1;
                jLo = j + 1;
            } else {
                L.Hw(j) = wc;
                iLo = i;
                jLo = j;
            }
        }
    } else if (lane == 0) {
        iHi = nL;
        jHi = nH;
    }
    const int cl = iHi - iLo, ch = jHi - jLo;
    const int offL = wave_excl_scan_int(cl), offH = wave_excl_scan_int(ch);
    const int mL = wave_sum_int(cl), mH = wave_sum_int(ch);
    const bool identity = !a.psa_plus;
    if (!identity) {
        for (int q = 0; q < cl; ++q) L.map[offL + q] = (uint16_t)(iLo + q);
        for (int q = 0; q < ch; ++q) L.map[mL + offH + q] = (uint16_t)(jLo + q);
    }
    wave_lds_sync();

    // Canonical local scans (chunks of 64, wave scan, carries chained): aggregates.
    double aggL = 0.0, aggH = 0.0;
    for (int r0 = 0; r0 < mL; r0 += 64) {
        const int r = r0 + lane;
        const double v = (r < mL) ? L.Lw(identity ? r : (int)L.map[r]) : 0.0;
        aggL = aggL + readlane_d(wave_scan(v), 63);
    }
    for (int r0 = 0; r0 < mH; r0 += 64) {
        const int r = r0 + lane;
        const double v = (r < mH) ? L.Hw(identity ? r : (int)L.map[mL + r]) : 0.0;
        aggH = aggH + readlane_d(wave_scan(v), 63);
    }
    const uint64_t aggC = (uint64_t)(uint32_t)mL | ((uint64_t)(uint32_t)mH << 32);
    bool timeout = false;
    if (lane == 0) block_publish(a, s, aggC, aggL, aggH, timeout);
    const Prefix P = block_prefix(a, s, lane, timeout);

    // Leftover lists with global prefixes P + local, and the split certificate.
    const int64_t goffL = (int64_t)(uint32_t)P.c, goffH = (int64_t)(P.c >> 32);
    double ext_l = 0.0, ext_h = __builtin_inf();
    for (int hv = 0; hv < 2; ++hv) {
        const bool heavy = hv == 1;
        const int m = heavy ? mH : mL;
        const double Pv = heavy ? P.h : P.l, agg = heavy ? aggH : aggL;
        int32_t* oi = heavy ? a.Hi : a.Li;
        double* ow = heavy ? a.Hw : a.Lw;
        double* op = heavy ? a.Hp : a.Lp;
        const int64_t goff = heavy ? goffH : goffL;
        double carry = 0.0;
        for (int r0 = 0; r0 < m; r0 += 64) {
            const int r = r0 + lane;
            double v = 0.0;
            int item = 0;
            if (r < m) {
                const int pos = identity ? r : (int)L.map[(heavy ? mL : 0) + r];
                v = heavy ? L.Hw(pos) : L.Lw(pos);
                item = heavy ? L.Hi(pos) : L.Li(pos);
            }
            const double inc = wave_scan(v);
            const double local = carry + wave_shr1(inc);
            const double pref = Pv + local;
            const double carry_next = carry + readlane_d(inc, 63);
            double succ = __shfl_down(pref, 1, 64);
            if (lane == 63) succ = Pv + carry_next;
            if (r == m - 1) succ = Pv + agg;
            if (r < m) {
                oi[goff + r] = (int32_t)(start + item);
                ow[goff + r] = v;
                op[goff + r] = pref;
                const double d = succ - pref;
                if (heavy) ext_h = fmin(ext_h, d); else ext_l = fmax(ext_l, d);
            }
            carry = carry_next;
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        ext_l = fmax(ext_l, shfl_xor_d(ext_l, m));
        ext_h = fmin(ext_h, shfl_xor_d(ext_h, m));
    }
    if (lane == 0) {
        if (mL > 0) atomicMax((unsigned long long*)&st->cert_max_dL, (unsigned long long)d2u(ext_l));
        if (mH > 0) atomicMin((unsigned long long*)&st->cert_min_dH, (unsigned long long)d2u(ext_h));
    }
    if (s == a.G - 1) {
        // Totals P(G), sentinels (AliasTableSplitGpu.cu:366-385) and the split inputs.
        const Prefix Tot = block_prefix(a, a.G, lane, timeout);
        if (lane == 0) {
            const int64_t nLr = (int64_t)(uint32_t)Tot.c, nHr = (int64_t)(Tot.c >> 32);
            a.Li[nLr] = (int32_t)(a.n + 1); a.Lw[nLr] = 0.0; a.Lp[nLr] = Tot.l;
            a.Hi[nHr] = (int32_t)a.n; a.Hw[nHr] = __builtin_inf(); a.Hp[nHr] = Tot.h;
            const int64_t numL = nLr + 1, numH = nHr + 1;
            st->numL = numL;
            st->numH = numH;
            st->degenerate = (nLr == 0 || nHr == 0) ? 1u : 0u;
            int64_t p = a.p_req;
            if (a.psa_plus) { p = min(p, (numL + numH) / 10); if (p < 2) p = 2; }
            st->p = st->degenerate ? 0 : p;
            st->nToSplit = numL + numH - 2;
            st->wToSplit = Tot.l + Tot.h;
        }
    }
    if (__ballot(timeout) != 0ull && lane == 0) atomicOr(&st->error, ERR_LOOKBACK_TIMEOUT);
}

// ---------------------------------------------------------------------------
// 3a. Degenerate tables: every leftover row carries its own mass.
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
// 3b. Split: bisection per k with the reference's predicate; 64 probes bound
//     the loop (the reference spins forever when FOUND is unreachable).
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
// 3c. Pack one section per thread (PackMethodBasic.cuh:59-126).
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
// 3d. Pack with LDS staging (the role of PackMethodBasicShared /
//     PackMethodChunkedShared, PackMethodBasicShared.cu:3-68): one wave packs
//     chains k0+1 .. k0+64.  Their light items [i_k0, i_kE) and heavy items
//     [j_k0, j_kE] are two contiguous ranges, copied to LDS with coalesced
//     loads when they fit; otherwise the wave reads the lists from HBM.
#define PACK_CAP 2048

struct GlobalLists {
    const int32_t *Li, *Hi;
    const double *Lw, *Hw;
    __device__ int li(int i) const { return Li[i]; }
    __device__ double lw(int i) const { return Lw[i]; }
    __device__ int hi(int j) const { return Hi[j]; }
    __device__ double hw(int j) const { return Hw[j]; }
};
struct StagedLists {
    const int32_t *Li, *Hi;
    const double *Lw, *Hw;
    int iA, jA;
    __device__ int li(int i) const { return Li[i - iA]; }
    __device__ double lw(int i) const { return Lw[i - iA]; }
    __device__ int hi(int j) const { return Hi[j - jA]; }
    __device__ double hw(int j) const { return Hw[j - jA]; }
};

// One pack chain (PackMethodBasic.cuh:59-126); hj caches the current heavy item.
template <bool AOS, typename Lists>
__device__ __forceinline__ void pack_chain(const BuildArgs& a, const Lists& l, Split prev, Split cur, double W_N) {
    const int iUpper = cur.i, jUpper = cur.j;
    int i = prev.i, j = prev.j;
    double w = prev.spill;
    if (prev.spill == 0) w = l.hw(j);
    int hj = l.hi(j);
    while (true) {
        bool ex = w <= W_N + ATG_EPSILON;
        if (ex && j >= jUpper) ex = false;
        else if (!ex && i >= iUpper) ex = true;
        if ((ex && j >= jUpper) || (!ex && i >= iUpper)) break;
        int32_t item, alias;
        double rw, next;
        if (ex) {
            const int hn = l.hi(j + 1);
            item = hj; rw = w; alias = hn; next = l.hw(j + 1); ++j; hj = hn;
        } else {
            item = l.li(i); rw = l.lw(i); alias = hj; next = rw; ++i;
        }
        w += next - W_N;
        if ((uint32_t)item < (uint32_t)a.n) store_row<AOS>(a, item, rw, alias);
    }
}

template <bool AOS>
__global__ __launch_bounds__(64) void pack_staged(BuildArgs a) {
    extern __shared__ double lds_raw[];
    const DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p;
    const int lane = threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.x * 64;
    if (k0 + 1 > p) return;
    const int64_t kE = min(k0 + 64, p);
    const Split lo = a.splits[k0], hi = a.splits[kE];
    const int64_t k = k0 + 1 + lane;
    Split prev{0, 0, 0.0}, cur{0, 0, 0.0};
    bool inside = true;
    if (k <= kE) {
        prev = a.splits[k - 1];
        cur = a.splits[k];
        inside = prev.i >= lo.i && prev.i <= cur.i && cur.i <= hi.i && prev.j >= lo.j && prev.j <= cur.j &&
                 cur.j <= hi.j;
    }
    const int nLu = hi.i - lo.i, nHu = hi.j - lo.j + 1;
    // Stage only when every chain of the wave stays inside the copied ranges
    // (always, unless a bisection fell back).
    const bool staged = __ballot(!inside) == 0ull && nLu >= 0 && nLu <= PACK_CAP && nHu >= 1 && nHu <= PACK_CAP;
    double* sLw = lds_raw;
    double* sHw = sLw + PACK_CAP;
    int32_t* sLi = reinterpret_cast<int32_t*>(sHw + PACK_CAP);
    int32_t* sHi = sLi + PACK_CAP;
    if (staged) {
        const int iA = lo.i, jA = lo.j;
        for (int t0 = 0; t0 < nLu; t0 += 8 * 64) {
            double wv[8];
            int32_t iv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u * 64 + lane;
                if (t < nLu) { wv[u] = a.Lw[iA + t]; iv[u] = a.Li[iA + t]; }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u * 64 + lane;
                if (t < nLu) { sLw[t] = wv[u]; sLi[t] = iv[u]; }
            }
        }
        for (int t0 = 0; t0 < nHu; t0 += 8 * 64) {
            double wv[8];
            int32_t iv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u * 64 + lane;
                if (t < nHu) { wv[u] = a.Hw[jA + t]; iv[u] = a.Hi[jA + t]; }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u * 64 + lane;
                if (t < nHu) { sHw[t] = wv[u]; sHi[t] = iv[u]; }
            }
        }
        wave_lds_sync();
    }
    if (k > kE) return;
    const double W_N = st->W_N;
    if (staged) {
        const StagedLists l{sLi, sHi, sLw, sHw, lo.i, lo.j};
        pack_chain<AOS>(a, l, prev, cur, W_N);
    } else {
        const GlobalLists l{a.Li, a.Hi, a.Lw, a.Hw};
        pack_chain<AOS>(a, l, prev, cur, W_N);
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
static inline unsigned blocks_for(int64_t items, int per) { return (unsigned)((items + per - 1) / per); }

hipError_t launch_sum(const BuildArgs& a, hipStream_t s) {
    const unsigned g = blocks_for(a.G, 64);
    if (a.dtype_f32) hipLaunchKernelGGL(sum_sections<float>, dim3(g), dim3(1024), 0, s, a);
    else hipLaunchKernelGGL(sum_sections<double>, dim3(g), dim3(1024), 0, s, a);
    return hipGetLastError();
}

size_t psa_lds_bytes(int64_t Ls_max) { return (((size_t)Ls_max * 12) + 15) & ~(size_t)15; }

int64_t block_tree_total(int64_t G) {
    int64_t total = 0;
    for (int b = 0; ((int64_t)1 << b) <= G; ++b) total += G >> b;
    return total;
}

hipError_t launch_partition(const BuildArgs& a, hipStream_t s) {
    const size_t lds = psa_lds_bytes(a.Ls_max);
    const dim3 g((unsigned)a.G), b(64);
    if (a.dtype_f32) {
        if (a.rows) hipLaunchKernelGGL((psa_wave<float, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((psa_wave<float, false>), g, b, lds, s, a);
    } else {
        if (a.rows) hipLaunchKernelGGL((psa_wave<double, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((psa_wave<double, false>), g, b, lds, s, a);
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
    const size_t lds = (size_t)PACK_CAP * 24;
    if (a.rows) hipLaunchKernelGGL(pack_staged<true>, dim3(blocks_for(kmax, 64)), dim3(64), lds, s, a);
    else hipLaunchKernelGGL(pack_staged<false>, dim3(blocks_for(kmax, 64)), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t set_partition_lds_limit(size_t bytes) {
    const void* fns[] = {(const void*)psa_wave<double, true>, (const void*)psa_wave<double, false>,
                         (const void*)psa_wave<float, true>, (const void*)psa_wave<float, false>};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace atg
