// Device building blocks of the alias-table construction, shared by the
// single-table kernels and the batched kernels (one launch per phase over all
// tables).  Each function computes exactly the canonical association, so both
// paths produce identical tables.
#pragma once

#include "atg_build.h"
#include "atg_device.h"
#include "atg_split.h"

namespace atg {

// Orders one wave's LDS writes before its later LDS reads by other lanes.  A
// wave's LDS operations execute in issue order, so wavefront-scope fences
// (compiler barriers, no wait instructions) suffice; a workgroup-scope fence
// would also wait for every outstanding global store of the wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// 1. Section sums.  Lane l of the wave that owns a section adds items
//    l, l+64, ... in order; the lane sums meet in the pairwise tree.  A NaN,
//    infinite or negative weight turns the section sum into NaN.  The loads
//    are non-temporal: the pass reads every weight once.
template <typename T>
__device__ __forceinline__ double section_sum(const T* w, int64_t start, int64_t end, int lane) {
    double acc = 0.0;
    bool bad = false;
    for (int64_t k0 = start; k0 < end; k0 += 32 * 64) {
        // 32 items per lane in flight, kept raw: a conversion next to each
        // load would serialise the loads.  start and end are scalars (the
        // caller's section index is), so the tests on k0 are scalar
        // branches.  Chunks past the end are skipped; in the last one, items
        // past the end read the last item (no exec masks around the loads)
        // and are not added.
        T xr[32];
        if (k0 + 32 * 64 <= end) {
#pragma unroll
            for (int u = 0; u < 32; ++u) xr[u] = __builtin_nontemporal_load(&w[k0 + u * 64 + lane]);
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const double x = (double)xr[u];
                bad |= !(x >= 0.0) || (x == __builtin_inf());
                acc = acc + x;
            }
        } else {
            // The section's tail: a scalar base and 32-bit lane offsets.
            const char* wb = reinterpret_cast<const char*>(w + k0);
            const int lim = (int)(end - 1 - k0);
#pragma unroll
            for (int u = 0; u < 32; ++u)
                xr[u] = (u * 64 <= lim)
                            ? __builtin_nontemporal_load(reinterpret_cast<const T*>(
                                  wb + (uint32_t)min(u * 64 + lane, lim) * (uint32_t)sizeof(T)))
                            : T(0);
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                if (k0 + u * 64 + lane < end) {
                    const double x = (double)xr[u];
                    bad |= !(x >= 0.0) || (x == __builtin_inf());
                    acc = acc + x;
                }
            }
        }
    }
    double sum = wave_tree_sum(acc);
    if (__ballot(bad) != 0ull) sum = __builtin_nan("");
    return sum;
}


#define PSA_MAX_LEN 2048
// Sections of at most this many items run the partition kernel that holds
// half the weight registers.
constexpr int64_t kShortSectionMax = 1024;
// Chunks of 64 leftovers a wave keeps in flight at once when staging or
// placing them.
#define PLACE_GROUP 8
// Chunks of 64 leftovers the placement pass keeps in flight (a 2048-item
// section has at most 32): 4 with lean light lists, 8 when it writes every
// list (fewer registers and more waves against more loads per wave; each the
// faster of 4, 8 and 16 for its pass).
template <bool LEAN>
constexpr int place_chunks() { return LEAN ? 4 : 8; }

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

// Level-b block c (all levels in one array per quantity) at slot
// level_offset(G, b) + c.
__device__ __forceinline__ void block_store(const BuildArgs& a, int64_t idx, uint64_t c, double l, double h) {
    uint64_t* V = a.slots;
    slot_store(V + idx, c);
    slot_store(V + a.tree_total + idx, d2u(l));
    slot_store(V + 2 * a.tree_total + idx, d2u(h));
}

// P(s): lane b loads the block of bit b of s (all complete, written by
// earlier launches), then the blocks are folded from the highest bit to the
// lowest.
__device__ __forceinline__ Prefix block_prefix_done(const BuildArgs& a, int64_t s, int lane) {
    uint64_t c = 0, l = d2u(0.0), h = d2u(0.0);
    const bool set = lane < 62 && ((s >> lane) & 1);
    if (set) {
        const int64_t idx = level_offset(a.G, lane) + (((s >> (lane + 1)) << (lane + 1)) >> lane);
        c = a.slots[idx];
        l = a.slots[a.tree_total + idx];
        h = a.slots[2 * a.tree_total + idx];
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

// Levels b0 + 1 .. b0 + 9 of the dyadic block tree, T(b, c) = T(b - 1, 2c) +
// T(b - 1, 2c + 1), over the level-b0 blocks [512 q, 512 q + 512) (one
// block of 256 threads; only complete blocks, c < G >> b, exist).
__device__ __forceinline__ void tree_block(const BuildArgs& a, int64_t q, int b0) {
    __shared__ uint64_t sc[512];
    __shared__ double sl[512], sh[512];
    const int64_t G = a.G, T = a.tree_total;
    uint64_t* V = a.slots;
    const int64_t cnt0 = G >> b0, first = q * 512;
    if (first >= cnt0) return;
    int64_t off = level_offset(G, b0);
    for (int i = threadIdx.x; i < 512; i += blockDim.x) {
        const int64_t c = first + i;
        if (c < cnt0) {
            sc[i] = V[off + c];
            sl[i] = u2d(V[T + off + c]);
            sh[i] = u2d(V[2 * T + off + c]);
        }
    }
    __syncthreads();
    for (int d = 1; d <= 9; ++d) {
        const int b = b0 + d;
        const int64_t cntb = G >> b;
        if (cntb == 0) break;
        off += G >> (b - 1);
        const int nodes = 512 >> d, i = threadIdx.x;
        const int64_t c = q * nodes + i;
        const bool valid = i < nodes && c < cntb;
        uint64_t vc = 0;
        double vl = 0.0, vh = 0.0;
        if (valid) {
            vc = sc[2 * i] + sc[2 * i + 1];
            vl = sl[2 * i] + sl[2 * i + 1];
            vh = sh[2 * i] + sh[2 * i + 1];
        }
        __syncthreads();
        if (valid) {
            sc[i] = vc;
            sl[i] = vl;
            sh[i] = vh;
            block_store(a, off + c, vc, vl, vh);
        }
        __syncthreads();
    }
}

// Positions of leftover ranks r0 + 64 u (u < U) of one class: chain t (lane
// t) holds its leftovers at [lo_t, lo_t + count_t) and its ranks start at
// off_t (exclusive scan of the counts, non-decreasing in t; lanes without a
// chain hold the total).  The chain of rank r is the largest t with
// off_t <= r, found by a binary search over the lanes; the U searches run
// interleaved, so each step issues U lane reads before it waits.  All lanes
// take part.  first: the search's first step, the largest power of two
// below the chain count (lanes from the chain count on hold the total, > r).
template <int U>
__device__ __forceinline__ void chain_positions(int r0, int off, int lo, int* pos, int first) {
    int t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        if (step > first) continue;
        int o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) o[u] = __shfl(off, t[u] + step, 64);
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] += (o[u] <= r0 + 64 * u) ? step : 0;
    }
    int b[U], c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        b[u] = __shfl(lo, t[u], 64);
        c[u] = __shfl(off, t[u], 64);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pos[u] = b[u] + (r0 + 64 * u - c[u]);
}

// The weights of section s, lane-strided (item 64 u + lane in x[u]), in
// their own type (converted to FP64 where they are used, so that no
// conversion waits for its load); all loads in flight at once.
template <typename T, int XU>
__device__ __forceinline__ void section_load(const BuildArgs& a, int64_t s, int lane, T (&x)[XU]) {
    const T* w = (const T*)a.weights + min(a.n, s * a.Ls);
    const int len = (int)(min(a.n, (s + 1) * a.Ls) - min(a.n, s * a.Ls));
    if (len == 64 * XU) {
#pragma unroll
        for (int u = 0; u < XU; ++u) x[u] = w[(u << 6) + lane];
    } else if (len > 0) {
        // A short section: lanes past its end read its last item (never
        // used), so no load needs its own exec mask.
#pragma unroll
        for (int u = 0; u < XU; ++u) x[u] = w[min((u << 6) + lane, len - 1)];
    }
}

// Leftover ranks g0 + 64 u + lane (u < U) of one class staged at base + rank
// (weight, item) and their chunk sums added to agg left to right.
template <int U>
__device__ __forceinline__ void stage_chunks(const BuildArgs& a, SecLDS& L, bool heavy, bool identity, int first,
                                             int g0, int m, int off, int lo, int64_t start, int64_t base, int lane,
                                             double& agg) {
    int pos[U];
    if (identity) {
#pragma unroll
        for (int u = 0; u < U; ++u) pos[u] = g0 + 64 * u + lane;
    } else {
        chain_positions<U>(g0 + lane, off, lo, pos, first);
    }
    double v[PLACE_GROUP];
#pragma unroll
    for (int u = 0; u < PLACE_GROUP; ++u) v[u] = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int r = g0 + 64 * u + lane;
        if (r < m) {
            v[u] = heavy ? L.Hw(pos[u]) : L.Lw(pos[u]);
            const int it = heavy ? L.Hi(pos[u]) : L.Li(pos[u]);
            // Streaming stores: the staging is read once, by the
            // placement pass (non-temporal row or list stores were
            // slower: those lines are merged or re-read in L2).
            __builtin_nontemporal_store(v[u], &a.stage_w[base + r]);
            __builtin_nontemporal_store((int32_t)(start + it), &a.stage_i[base + r]);
        }
    }
    // The chunk sums of the group (lane u: chunk u), added left to right.
    static_assert(PLACE_GROUP == 8, "wave_tree_sum8 sums eight chunks");
    const double t8 = wave_tree_sum8(v, lane);
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (g0 + 64 * u < m) agg = agg + readlane_d(t8, u);
}

// Partition pass 1, section s by one wave, its weights in x (section_load):
// stable light/heavy partition into LDS, then after_partition() (x is free
// from there on: the caller loads its next section into it), PSA+ greedy
// chains writing finished rows, the leftover order, the canonical local-scan
// aggregates (stored as the level-0 block of the dyadic section tree) and the
// leftovers staged in rank order at the section's own range (lights, then
// heavies).  Nothing here waits for another section's work.
template <typename T, bool AOS, int XU, typename AfterPartition>
__device__ __forceinline__ void psa_section(const BuildArgs& a, int64_t s, T (&x)[XU],
                                            double* lds_raw, int lane, AfterPartition after_partition) {
    DevState* st = a.st;
    const int64_t start = min(a.n, s * a.Ls);
    const int len = (int)(min(a.n, (s + 1) * a.Ls) - start);
    const double W_N = st->W_N;
    SecLDS L;
    L.w = lds_raw;
    L.idx = reinterpret_cast<uint16_t*>(lds_raw + a.Ls_max);
    L.len = len;
    const uint64_t lt = lanemask_lt();

    // The stable partition: lights to the front in index order, heavies to
    // the back (h[k] at len-1-k) in index order.  vmcnt(0) on every path:
    // without it the compiler keeps the loads pending on the skipped blocks
    // below and waits later, when the registers are reused, for every row
    // store the chains issued by then.
    __builtin_amdgcn_s_waitcnt(0x0F70);
    int nL = 0, nH = 0;
#pragma unroll
    for (int u = 0; u < XU; ++u) {
        const int k = (u << 6) + lane;
        if ((u << 6) + 64 <= len) {
            // A full chunk, without branches: a light's slot is nL + (lights
            // before the lane), a heavy's len - 1 - nH - (heavies before it).
            const double xv = (double)x[u];
            const bool light = xv <= W_N;
            const uint64_t bl = __ballot(light);
            const int before = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bl >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)bl, 0u));
            const int slot = before + (light ? nL : len - 1 - nH - lane);
            L.w[slot] = xv;
            L.idx[slot] = (uint16_t)k;
            const int c = __popcll(bl);
            nL += c;
            nH += 64 - c;
        } else if ((u << 6) < len) {
            const bool valid = k < len;
            const double xv = (double)x[u];
            const bool light = valid && xv <= W_N, heavy = valid && !(xv <= W_N);
            const uint64_t bl = __ballot(light), bh = __ballot(heavy);
            if (light) {
                const int pos = nL + __popcll(bl & lt);
                L.Lw(pos) = xv;
                L.Li(pos) = (uint16_t)k;
            }
            if (heavy) {
                const int pos = nH + __popcll(bh & lt);
                L.Hw(pos) = xv;
                L.Hi(pos) = (uint16_t)k;
            }
            nL += __popcll(bl);
            nH += __popcll(bh);
        }
    }
    wave_lds_sync();
    after_partition();

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
            // Registers hold the chain's current heavy (residual wc, item hj),
            // the next heavy (hw1, hi1) and the current light (lw, li); a step
            // stores one row and reads the light and the next heavy at its
            // new indices for the following step.
            // The loop is one straight run of instructions besides its exit.
            // Lights and heavies share one LDS array (heavy j at len-1-j).
            // The next heavy j + 1 <= jHi <= nH sits at slot hn = len - 2 - j
            // >= nL - 1 >= 0 (the chain has a light); past the chain's range
            // it is read but never used.  The exit test is lane-mask logic.
            int i = iLo, hn = len - 2 - jLo;
            double wc = L.Hw(jLo);
            int hj = L.Hi(jLo);
            double lw = L.Lw(i);
            int li = L.Li(i);
            double hw1 = L.w[hn];
            int hi1 = L.idx[hn];
            const double thr = W_N + ATG_EPSILON;
            const int iEnd = iHi - 1, hnEnd = len - 1 - jHi;
            while (true) {
                // Full bucket: the heavy's row, then the next heavy; else the
                // light's row.  A step needs a further element on its side.
                const bool ex = wc <= thr;
                const uint64_t mEx = __ballot(ex), mH = __ballot(hn <= hnEnd), mL = __ballot(i >= iEnd);
                if (__builtin_amdgcn_inverse_ballot_w64((mEx & mH) | (~mEx & mL))) break;
                store_row<AOS>(a, start + (ex ? hj : li), ex ? wc : lw, (int32_t)(start + (ex ? hi1 : hj)));
                wc += (ex ? hw1 : lw) - W_N;
                hj = ex ? hi1 : hj;
                hn -= (int)ex;
                i += (int)!ex;
                lw = L.Lw(i);
                li = L.Li(i);
                hw1 = L.w[hn];
                hi1 = L.idx[hn];
            }
            const int j = len - 2 - hn;
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
    // Both classes' leftover counts in one scan (each total is at most
    // PSA_MAX_LEN < 2^16): offsets from the exclusive scan, totals from the
    // last lane.
    static_assert(PSA_MAX_LEN < 65536, "16-bit count fields");
    const int cl = iHi - iLo, ch = jHi - jLo;
    const uint32_t cp = (uint32_t)cl | ((uint32_t)ch << 16);
    const uint32_t incl = wave_incl_scan_u32(cp), excl = incl - cp;
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const int offL = (int)(excl & 0xFFFFu), offH = (int)(excl >> 16);
    const int mL = (int)(tot & 0xFFFFu), mH = (int)(tot >> 16);
    const bool identity = !a.psa_plus;
    wave_lds_sync();

    // One pass per list: stage the leftovers (weight, item) in rank order
    // (lights at [start, start + mL), heavies at [start + mL, start + mL + mH))
    // and form the canonical local-scan aggregate (chunks of 64, each summed
    // by the pairwise tree of the wave scan's last lane, chunk sums added left
    // to right), up to PLACE_GROUP chunks at a time (the last group of a list
    // takes 2, 4 or 8, as many as it holds).
    // The rank search's first step: the largest power of two below the
    // chain count (one chain: step 1, a search over lanes 0 and 1).
    int first = 1;
    while (2 * first < a.chains) first *= 2;
    double agg[2] = {0.0, 0.0};
    for (int hv = 0; hv < 2; ++hv) {
        const bool heavy = hv == 1;
        const int m = heavy ? mH : mL;
        const int64_t base = start + (heavy ? mL : 0);
        const int off = heavy ? offH : offL, lo = heavy ? jLo : iLo;
        for (int g0 = 0; g0 < m; g0 += 64 * PLACE_GROUP) {
            const int nc = (m - g0 + 63) >> 6;
            if (nc > 4)
                stage_chunks<PLACE_GROUP>(a, L, heavy, identity, first, g0, m, off, lo, start, base, lane, agg[hv]);
            else if (nc > 2)
                stage_chunks<4>(a, L, heavy, identity, first, g0, m, off, lo, start, base, lane, agg[hv]);
            else
                stage_chunks<2>(a, L, heavy, identity, first, g0, m, off, lo, start, base, lane, agg[hv]);
        }
    }
    const double aggL = agg[0], aggH = agg[1];
    const uint64_t aggC = (uint64_t)(uint32_t)mL | ((uint64_t)(uint32_t)mH << 32);
    if (lane == 0) block_store(a, s, aggC, aggL, aggH);
}

// Partition pass 2, section s by one wave: the section prefix P(s) from the
// completed dyadic blocks, then the staged leftovers with prefixes
// P(s) + local (the canonical chunked scan recomputed in the same order),
// and for the last section the totals, sentinels and split inputs.
// LEAN: the lean light lists (a.lean_lights), fixed at compile time so that
// each kernel carries one path.
template <bool LEAN>
__device__ __forceinline__ void place_section(const BuildArgs& a, int64_t s, int lane) {
    DevState* st = a.st;
    const int64_t start = min(a.n, s * a.Ls);
    // The block sums were completed by the previous launch.
    const uint64_t c0 = a.slots[s];  // this section's own counts
    const int mL = (int)(uint32_t)c0, mH = (int)(c0 >> 32);
    const Prefix P = block_prefix_done(a, s, lane);
    const int64_t goffL = (int64_t)(uint32_t)P.c, goffH = (int64_t)(P.c >> 32);
    // Split certificate (learned split only): the extreme increments between
    // consecutive stored prefixes inside the section, and the section's first
    // and last prefix of each class; the pairs across sections are closed by
    // cert_reduce_block.
    double ext_l = 0.0, ext_h = __builtin_inf();
#pragma unroll 1
    for (int hv = 0; hv < 2; ++hv) {
        const bool heavy = hv == 1;
        const int m = heavy ? mH : mL;
        if (m == 0) continue;
        const int64_t src = start + (heavy ? mL : 0);
        const double Pv = heavy ? P.h : P.l;
        int32_t* oi = heavy ? a.Hi : a.Li;
        double* ow = heavy ? a.Hw : a.Lw;
        double* op = heavy ? a.Hp : a.Lp;
        const int64_t goff = heavy ? goffH : goffL;
        // Lean lists: the lights keep their items and weights in the staging.
        const bool full = heavy || !LEAN;
        double ext = heavy ? __builtin_inf() : 0.0;
        constexpr int PC = place_chunks<LEAN>();
        // Groups of PC chunks of 64: all loads of a group in flight
        // together, the chunk scans independent, then the carries chained left
        // to right (the same operations in the same order as one chunk at a
        // time).
        double carry = 0.0;
        for (int g0 = 0; g0 < m; g0 += 64 * PC) {
            double v[PC], inc[PC];
            int32_t item[PC];
            // Chunks past the list are skipped; in the last one, ranks past
            // the list read its last entry (no exec masks around the loads)
            // and count as 0.
            if (g0 + 64 * PC <= m) {
#pragma unroll
                for (int u = 0; u < PC; ++u) {
                    v[u] = __builtin_nontemporal_load(&a.stage_w[src + g0 + 64 * u + lane]);
                    item[u] = full ? __builtin_nontemporal_load(&a.stage_i[src + g0 + 64 * u + lane]) : 0;
                }
            } else {
                // The list's last group: a scalar base and 32-bit lane offsets.
                const char* bw = reinterpret_cast<const char*>(a.stage_w + src + g0);
                const char* bi = reinterpret_cast<const char*>(a.stage_i + src + g0);
                const int lim = m - 1 - g0;
#pragma unroll
                for (int u = 0; u < PC; ++u) {
                    const uint32_t rc = (uint32_t)min(64 * u + lane, lim);
                    v[u] = 0.0;
                    item[u] = 0;
                    if (64 * u <= lim) {
                        v[u] = __builtin_nontemporal_load(reinterpret_cast<const double*>(bw + rc * 8u));
                        if (full) item[u] = __builtin_nontemporal_load(reinterpret_cast<const int32_t*>(bi + rc * 4u));
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < PC; ++u)
                if (g0 + 64 * u + lane >= m) v[u] = 0.0;
#pragma unroll
            for (int u = 0; u < PC; ++u) inc[u] = wave_scan(v[u]);
#pragma unroll
            for (int u = 0; u < PC; ++u) {
                const int r0 = g0 + 64 * u;
                if (r0 < m) {
                    const int r = r0 + lane;
                    const double local = carry + wave_shr1(inc[u]);
                    const double pref = Pv + local;
                    const double carry_next = carry + readlane_d(inc[u], 63);
                    if (r < m) {
                        if (full) {
                            oi[goff + r] = item[u];
                            ow[goff + r] = v[u];
                        }
                        op[goff + r] = pref;
                    }
                    if (a.want_cert) {
                        // The next chunk's first prefix is Pv + (carry_next + 0).
                        double succ = __shfl_down(pref, 1, 64);
                        if (lane == 63) succ = Pv + carry_next;
                        if (r < m - 1) {
                            const double d = succ - pref;
                            ext = heavy ? fmin(ext, d) : fmax(ext, d);
                        }
                        if (r == m - 1) a.sec_cert[6 * s + (heavy ? 5 : 3)] = d2u(pref);
                    }
                    carry = carry_next;
                }
            }
        }
        if (heavy) ext_h = ext; else ext_l = ext;
    }
    if (a.want_cert) {
#pragma unroll
        for (int q = 1; q < 64; q <<= 1) {
            ext_l = fmax(ext_l, shfl_xor_d(ext_l, q));
            ext_h = fmin(ext_h, shfl_xor_d(ext_h, q));
        }
        if (lane == 0) {
            // [0] max dL, [1] min dH inside; [2] first L, [3] last L, [4] first H, [5] last H.
            a.sec_cert[6 * s] = d2u(ext_l);
            a.sec_cert[6 * s + 1] = d2u(ext_h);
            a.sec_cert[6 * s + 2] = d2u(P.l);
            a.sec_cert[6 * s + 4] = d2u(P.h);
        }
    }
    if (LEAN) {
        if (lane == 0) a.sec_loff[s] = (uint32_t)goffL;
        // The directory entries of the lights 1024 q this section holds.
        const int64_t q0 = (goffL + (1 << kLightDirShift) - 1) >> kLightDirShift;
        const int64_t q1 = mL > 0 ? (goffL + mL - 1) >> kLightDirShift : q0 - 1;
        for (int64_t q = q0 + lane; q <= q1; q += 64) a.sec_ldir[q] = (uint32_t)s;
    }
    if (s == a.G - 1) {
        // Totals P(G), sentinels (AliasTableSplitGpu.cu:366-385) and the split inputs.
        const Prefix Tot = block_prefix_done(a, a.G, lane);
        if (lane == 0) {
            const int64_t nLr = (int64_t)(uint32_t)Tot.c, nHr = (int64_t)(Tot.c >> 32);
            if (LEAN) a.sec_loff[a.G] = (uint32_t)nLr;
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
}


// ---------------------------------------------------------------------------
// W, pass 1: section sums, one wave per section (block blk of nblk holds
// sections 4 blk .. 4 blk + 3), and the reset of the dyadic block sums and
// their arrival counters for the partition.
template <typename T>
__device__ __forceinline__ void sum_sections_block(const BuildArgs& a, int64_t blk, int64_t nblk) {
    const int lane = threadIdx.x & 63, wave = wave_in_block();
    const T* w = (const T*)a.weights;
    const int64_t s = blk * (blockDim.x >> 6) + wave;
    if (s < a.G) {
        const double sum = section_sum<T>(w, min(a.n, s * a.Ls), min(a.n, (s + 1) * a.Ls), lane);
        if (lane == 0) a.section_sums[s] = sum;
    }
    const int64_t nvals = 3 * a.tree_total, ncnt = (a.tree_total + 1) / 2;
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < nvals + ncnt; i += nblk * blockDim.x)
        a.slots[i] = (i < nvals) ? ATG_NOT_READY : 0ull;
}

// W, pass 2 (one block): the 64-ary tree over the section sums, 64 values per
// wave and level, zero-padded; W, W/N, the error bits and the state resets.
__device__ __forceinline__ void sum_finish_block(const BuildArgs& a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double* bufs[2] = {a.tree_scratch, a.tree_scratch + a.tree_scratch_half};
    const double* src = a.section_sums;
    int cur = 0;
    int64_t count = a.G;
    while (true) {
        const int64_t groups = (count + 63) / 64;
        // Wave w reduces groups w, w + nw, ...: eight groups' loads in
        // flight, their pairwise-tree sums at once (lane u < 8: group u).
        for (int64_t g0 = wave; g0 < groups; g0 += 8 * nw) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t i = (g0 + u * nw) * 64 + lane;
                v[u] = (i < count) ? src[i] : 0.0;
            }
            const double r = wave_tree_sum8(v, lane);
            const int64_t g = g0 + lane * nw;
            if (lane < 8 && g < groups) bufs[cur][g] = r;
        }
        __syncthreads();
        src = bufs[cur];
        cur ^= 1;
        count = groups;
        if (count == 1) break;
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
        st->degenerate = 0;
        st->cert_max_dL = 0ull;                    // +0.0
        st->cert_min_dH = 0x7FF0000000000000ull;   // +inf
        st->cert_done = 0;
        st->split_fallbacks = 0;
        st->learned_fallbacks = 0;
    }
}


// ---------------------------------------------------------------------------
// Pack helpers.

struct GlobalLists {
    const int32_t *Li, *Hi;
    const double *Lw, *Hw;
    __device__ int li(int i) const { return Li[i]; }
    __device__ double lw(int i) const { return Lw[i]; }
    __device__ int hi(int j) const { return Hi[j]; }
    __device__ double hw(int j) const { return Hw[j]; }
};
// The lists without their weight arrays: each weight read through its item
// from the input weights (PackMethodNoWeights.cu:3-60 reads them from a table
// prefilled with the weights).  Without PSA+ the lists' weights are the input
// weights, so the chains see the same values; the sentinels keep theirs (the
// light one, item n + 1, weighs 0; the heavy one, item n, +inf).
template <typename T>
struct NoWeightsLists {
    const int32_t *Li, *Hi;
    const T* w;
    int64_t n;
    __device__ int li(int i) const { return Li[i]; }
    __device__ double lw(int i) const {
        const int it = Li[i];
        return (uint32_t)it < (uint32_t)n ? (double)w[it] : 0.0;
    }
    __device__ int hi(int j) const { return Hi[j]; }
    __device__ double hw(int j) const {
        const int it = Hi[j];
        return (uint32_t)it < (uint32_t)n ? (double)w[it] : __builtin_inf();
    }
};
// Lean light lists: light i lives in the staging of section s, the last
// section with sec_loff[s] <= i (the non-empty one), at position
// min(n, s Ls) + i - sec_loff[s].

// The directory sec_ldir[q] holds the section of light q << kLightDirShift, a
// section whose offset is at most any light of that block of lights.
// Staging positions of lights g0 + 64 u + lane (u < U; those at or past m
// from g0 are left alone), starting from a section s0 whose offset is at most
// g0.  A group that spans a few sections walks them (their bounds uniform,
// so scalar loads), each lane taking the position of its light from the
// section that holds it; past kLightWalk sections the rest is found from
// windows of 63 section offsets, each lane halving over the window.  Returns
// a section whose offset is at most the group's last light, the starting
// point for the next group.
constexpr int kLightWalk = 8;
template <int U>
__device__ __forceinline__ int64_t light_positions(const BuildArgs& a, int64_t s0, uint32_t g0, int m, int lane,
                                                   int32_t (&pos)[U]) {
    const uint32_t gEnd = g0 + (uint32_t)min(m, 64 * U);  // the group's lights are [g0, gEnd)
    int64_t s = __builtin_amdgcn_readfirstlane((int)s0);
    uint32_t lo = a.sec_loff[s], hi = a.sec_loff[s + 1];
    for (int walk = 0;; ++walk) {
        const int32_t base = (int32_t)min(a.n, s * a.Ls) - (int32_t)lo;  // position of light g: base + g
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t g = g0 + (uint32_t)(64 * u + lane);
            if (g >= lo && g < hi) pos[u] = base + (int32_t)g;
        }
        if (hi >= gEnd) return s;
        ++s;
        lo = hi;
        if (walk + 1 == kLightWalk) break;
        hi = a.sec_loff[s + 1];
    }
    // Lights [lo, gEnd) are left; sec_loff[s] = lo.
    bool done[U];
#pragma unroll
    for (int u = 0; u < U; ++u) done[u] = 64 * u + lane >= m || g0 + (uint32_t)(64 * u + lane) < lo;
    for (;; s += 63) {
        const uint32_t w = a.sec_loff[min(s + lane, a.G)];
        bool pending = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t g = g0 + (uint32_t)(64 * u + lane);
            int d = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) {
                const uint32_t v = (uint32_t)__shfl((int)w, min(d + step, 63), 64);
                if (d + step <= 63 && v <= g) d += step;
            }
            const uint32_t wd = (uint32_t)__shfl((int)w, d, 64);
            if (!done[u]) {
                if (d < 63) {
                    pos[u] = (int32_t)(min(a.n, (s + d) * a.Ls) + (int64_t)(g - wd));  // < n < 2^31
                    done[u] = true;
                } else {
                    pending = true;
                }
            }
        }
        if (__ballot(pending) == 0ull) return s;
    }
}

// Per-lane access to lean lists (the pack's fallback for non-monotone
// splits): a bisection over sec_loff per light.
struct LeanLists {
    const int32_t* Hi;
    const double* Hw;
    const uint32_t* loff;
    const double* sw;
    const int32_t* si;
    int64_t G, Ls, n;
    __device__ int64_t lpos(int i) const {
        int64_t lo = 0, hi = G - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (loff[mid] <= (uint32_t)i) lo = mid;
            else hi = mid - 1;
        }
        return min(n, lo * Ls) + (int64_t)((uint32_t)i - loff[lo]);
    }
    __device__ bool sentinel(int i) const { return (uint32_t)i >= loff[G]; }
    __device__ int li(int i) const { return sentinel(i) ? (int)(n + 1) : si[lpos(i)]; }
    __device__ double lw(int i) const { return sentinel(i) ? 0.0 : sw[lpos(i)]; }
    __device__ int hi(int j) const { return Hi[j]; }
    __device__ double hw(int j) const { return Hw[j]; }
};

// One pack chain (PackMethodBasic.cuh:59-126).  Registers hold the current
// bucket weight w, the current heavy item hj, the next heavy (hw1, hi1) and
// the current light (lw, li); every step issues the loads of the light after
// the current one and of the heavy after the next one, which the following
// step consumes at most one of, so list reads stay off the chain's critical
// path.  Read indices are clamped to the chain's own ranges (lights up to
// max(iUpper - 1, prev.i), heavies up to jUpper); a clamped value is never
// consumed.
template <bool AOS, typename Lists>
__device__ __forceinline__ void pack_chain(const BuildArgs& a, const Lists& l, Split prev, Split cur, double W_N) {
    const int iUpper = cur.i, jUpper = cur.j;
    int i = prev.i, j = prev.j;
    const int iLast = max(iUpper - 1, i);
    double w = prev.spill;
    if (prev.spill == 0) w = l.hw(j);
    int hj = l.hi(j);
    double lw = l.lw(min(i, iLast));
    int li = l.li(min(i, iLast));
    double hw1 = l.hw(min(j + 1, jUpper));
    int hi1 = l.hi(min(j + 1, jUpper));
    double nlw = l.lw(min(i + 1, iLast));
    int nli = l.li(min(i + 1, iLast));
    double nhw = l.hw(min(j + 2, jUpper));
    int nhi = l.hi(min(j + 2, jUpper));
    const double thr = W_N + ATG_EPSILON;
    while (true) {
        const bool jdone = j >= jUpper, idone = i >= iUpper;
        // A side that ran out hands the step to the other one.
        const bool ex = (w <= thr) ? !jdone : idone;
        if (ex ? jdone : idone) break;
        const int item = ex ? hj : li;
        const double rw = ex ? w : lw;
        const int alias = ex ? hi1 : hj;
        const double next = ex ? hw1 : lw;
        w += next - W_N;
        if ((uint32_t)item < (uint32_t)a.n) store_row<AOS>(a, item, rw, alias);
        j += ex ? 1 : 0;
        i += ex ? 0 : 1;
        hj = ex ? hi1 : hj;
        hw1 = ex ? nhw : hw1;
        hi1 = ex ? nhi : hi1;
        lw = ex ? lw : nlw;
        li = ex ? li : nli;
        const int qi = min(i + 1, iLast), qj = min(j + 2, jUpper);
        nlw = l.lw(qi);
        nli = l.li(qi);
        nhw = l.hw(qj);
        nhi = l.hi(qj);
    }
}

// One pack chain over its round's LDS pool (the steps of pack_chain, with
// pool indices: light i at i - Ia, heavy j at nLu + j - Ja).  Rows are not
// stored here but recorded in the pool: each step writes its row's alias to
// sA[slot] and, for a heavy, the residual weight to sW[slot] (a light's weight
// is already there), so that the caller writes all rows in pool order.  A
// light step writes its weight to the spare slot sW[spare] instead.  The
// chain's last heavy (jUpper) is the next chain's first, whose slot that
// chain overwrites: its weight is read once, before any chain steps, and
// taken from a register.  The loop's only branch is its exit.
__device__ __forceinline__ void pack_chain_pool(double* sW, const int32_t* sI, int32_t* sA, int spare, int nLu,
                                                int Ia, int Ja, Split prev, Split cur, double W_N) {
    const int iUpper = cur.i - Ia, jUpper = nLu + cur.j - Ja;
    int i = prev.i - Ia, j = nLu + prev.j - Ja;
    const int iLast = max(iUpper - 1, i);  // light reads are clamped to the chain's range
    const double hwLast = sW[jUpper];
    double w = prev.spill;
    if (prev.spill == 0) w = sW[j];
    int hj = sI[j];
    double lw = sW[min(i, iLast)];
    int q = min(j + 1, jUpper);
    double hw1 = sW[q];  // the next heavy's weight, unless it is the last (q == jUpper)
    int hi1 = sI[q];
    wave_lds_sync();  // every chain holds its last heavy's weight before any slot is overwritten
    const double thr = W_N + ATG_EPSILON;
    while (true) {
        const bool idone = i >= iUpper, jdone = j >= jUpper;
        if (idone && jdone) break;
        // A heavy step when the bucket is full (or no light is left) and a
        // heavy remains; a light step otherwise (pack_chain's rule).
        const bool ex = idone || (w <= thr && !jdone);
        sA[ex ? j : i] = ex ? hi1 : hj;
        sW[ex ? j : spare] = w;
        w += (ex ? (q == jUpper ? hwLast : hw1) : lw) - W_N;
        hj = ex ? hi1 : hj;
        j += ex ? 1 : 0;
        i += ex ? 0 : 1;
        lw = sW[min(i, iLast)];
        q = min(j + 1, jUpper);
        hw1 = sW[q];
        hi1 = sI[q];
    }
}


// ---------------------------------------------------------------------------
// Split certificate: max light increment and min heavy increment over all
// consecutive stored prefix pairs (sentinels included).  Per section s: the
// extremes inside it, and the pair from its last prefix of a class to the
// first prefix of the next section holding that class (that section's P) or
// to the class total P(G) of the sentinel.  max dL < min dH makes the FOUND
// predicate monotone in j.  Block blk of nblk folds sections blk, blk + nblk,
// ... with atomics on the state (max/min are exact in any order).
__device__ __forceinline__ void cert_reduce_block(const BuildArgs& a, int64_t blk, int64_t nblk) {
    __shared__ uint64_t sh_l[16], sh_h[16];
    DevState* st = a.st;
    const int64_t G = a.G;
    const double totL = a.Lp[st->numL - 1], totH = a.Hp[st->numH - 1];
    double l = 0.0, h = __builtin_inf();
    for (int64_t s = blk * blockDim.x + threadIdx.x; s < G; s += nblk * blockDim.x) {
        const uint64_t* c = a.sec_cert + 6 * s;
        const uint64_t cnt = a.slots[s];
        l = fmax(l, u2d(c[0]));
        h = fmin(h, u2d(c[1]));
        for (int hv = 0; hv < 2; ++hv) {
            if (((hv ? cnt >> 32 : cnt & 0xFFFFFFFFull)) == 0) continue;
            int64_t s2 = s + 1;
            while (s2 < G && ((hv ? a.slots[s2] >> 32 : a.slots[s2] & 0xFFFFFFFFull)) == 0) ++s2;
            const double next = s2 < G ? u2d(a.sec_cert[6 * s2 + (hv ? 4 : 2)]) : (hv ? totH : totL);
            const double d = next - u2d(c[hv ? 5 : 3]);
            if (hv) h = fmin(h, d); else l = fmax(l, d);
        }
    }
#pragma unroll
    for (int q = 1; q < 64; q <<= 1) {
        l = fmax(l, shfl_xor_d(l, q));
        h = fmin(h, shfl_xor_d(h, q));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh_l[wave] = d2u(l);
        sh_h[wave] = d2u(h);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t bl = sh_l[0], bh = sh_h[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            bl = max(bl, sh_l[q]);
            bh = min(bh, sh_h[q]);
        }
        atomicMax((unsigned long long*)&st->cert_max_dL, (unsigned long long)bl);
        atomicMin((unsigned long long*)&st->cert_min_dH, (unsigned long long)bh);
        if (blk == 0) st->cert_done = 1u;
    }
}

// ---------------------------------------------------------------------------
// Learned split index helpers.
__device__ __forceinline__ int64_t knot_k(int64_t m, int64_t p, int64_t K) { return (m * p) / K; }

// Knot m (0 <= m <= K) of the learned split index: the exact split at
// k_m = floor(m p / K), by bisection.
// Returns 1 when the bisection fell back to its clamp (the caller counts it).
__device__ __forceinline__ uint32_t knot_split(const BuildArgs& a, int64_t m) {
    DevState* st = a.st;
    uint32_t fb = 0;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const int64_t K = min<int64_t>(a.learned_segments, p);
    const int64_t k = knot_k(m, p, K);
    Split sp;
    if (k == 0) {
        sp = Split{0, 0, 0.0};
    } else if (k == p) {
        sp = Split{(int32_t)(numL - 1), (int32_t)(numH - 1), 0.0};
    } else if (!bisect_split(k, p, st->nToSplit, st->wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, sp)) {
        fb = 1;
    }
    a.splits[k] = sp;
    a.knot_j[m] = sp.j;
    return fb;
}

// Split k (0 < k < p, not a knot) from the model: interpolate between the
// knots, gallop, bisect inside the bracket; full bisection when the
// certificate is missing or the bracket search fails.
// Returns bit 0: the bisection fell back to its clamp; bit 1: the learned
// search fell back to full bisection (the caller counts them).
__device__ __forceinline__ uint32_t learned_split(const BuildArgs& a, int64_t k) {
    DevState* st = a.st;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const int64_t K = min<int64_t>(a.learned_segments, p);
    // Segment m with k_m <= k < k_{m+1}.
    int64_t m = (k * K) / p;
    while (m < K && knot_k(m + 1, p, K) <= k) ++m;
    while (m > 0 && knot_k(m, p, K) > k) --m;
    const int64_t km = knot_k(m, p, K), km1 = knot_k(m + 1, p, K);
    if (k == km) return 0;  // a knot: written by knot_split
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
    uint32_t flags = 0;
    if (!ok) {
        if (!bisect_split(k, p, nToSplit, wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, out)) flags |= 1u;
        flags |= 2u;
    }
    a.splits[k] = out;
    return flags;
}

// ---------------------------------------------------------------------------
// Degenerate tables (one leftover class empty): every leftover row carries
// its own mass.  Block blk of nblk.
template <bool AOS>
__device__ __forceinline__ void degenerate_block(const BuildArgs& a, int64_t blk, int64_t nblk) {
    const DevState* st = a.st;
    if (st->error || !st->degenerate) return;
    const int64_t nLr = st->numL - 1, nHr = st->numH - 1;
    if (a.lean_lights) {
        // The lights from the staging, one wave per section; then the heavies.
        const int lane = threadIdx.x & 63;
        const int64_t wpb = blockDim.x >> 6;
        for (int64_t s = blk * wpb + wave_in_block(); s < a.G; s += nblk * wpb) {
            const int64_t base = min(a.n, s * a.Ls);
            const int64_t m = (int64_t)a.sec_loff[s + 1] - (int64_t)a.sec_loff[s];
            for (int64_t r = lane; r < m; r += 64) {
                const int32_t item = a.stage_i[base + r];
                store_row<AOS>(a, item, a.stage_w[base + r], item);
            }
        }
        for (int64_t t = blk * blockDim.x + threadIdx.x; t < nHr; t += nblk * blockDim.x)
            store_row<AOS>(a, a.Hi[t], a.Hw[t], a.Hi[t]);
        return;
    }
    for (int64_t t = blk * blockDim.x + threadIdx.x; t < nLr + nHr; t += nblk * blockDim.x) {
        const int32_t item = (t < nLr) ? a.Li[t] : a.Hi[t - nLr];
        const double w = (t < nLr) ? a.Lw[t] : a.Hw[t - nLr];
        store_row<AOS>(a, item, w, item);
    }
}

// The light items and weights of the lean lists written out as lists (for a
// copy-out of the lists), one wave per section.
__device__ __forceinline__ void lean_lights_fill_block(const BuildArgs& a, int64_t blk, int64_t nblk) {
    if (a.st->error) return;
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    for (int64_t s = blk * wpb + wave_in_block(); s < a.G; s += nblk * wpb) {
        const int64_t base = min(a.n, s * a.Ls), l0 = a.sec_loff[s];
        const int64_t m = (int64_t)a.sec_loff[s + 1] - l0;
        for (int64_t r = lane; r < m; r += 64) {
            a.Li[l0 + r] = a.stage_i[base + r];
            a.Lw[l0 + r] = a.stage_w[base + r];
        }
    }
}

// Bisection splits k = blk * blockDim + tid + 1.
__device__ __forceinline__ void split_bisect_block(const BuildArgs& a, int64_t blk) {
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    const int64_t k = blk * blockDim.x + threadIdx.x + 1;
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

// Pack of chains k0+1 .. k0+64 (k0 = 64 blk) by one wave, in rounds.  A
// round takes the next chains whose lights [i_a, i_b) and heavies [j_a, j_b]
// (contiguous list ranges) fit the LDS pool of a.pack_cap items, copies them
// there with coalesced loads, runs the chains (heavy rows stored at once,
// light aliases recorded in the pool) and then stores the light rows in list
// order, where items rise, so that a store instruction covers neighbouring
// rows.  Chains that do not fit alone, and waves with a non-monotone split
// (a bisection fell back), read the lists from HBM.
template <bool AOS>
__device__ __forceinline__ void pack_staged_block(const BuildArgs& a, int64_t blk, double* lds_raw) {
    const DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p;
    const int lane = threadIdx.x;
    const int64_t k0 = blk * 64;
    if (k0 + 1 > p) return;
    const int nch = (int)(min(k0 + 64, p) - k0);
    const bool mine = lane < nch;
    Split prev{0, 0, 0.0}, cur{0, 0, 0.0};
    if (mine) {
        prev = a.splits[k0 + lane];
        cur = a.splits[k0 + lane + 1];
    }
    const double W_N = st->W_N;
    const GlobalLists gl{a.Li, a.Hi, a.Lw, a.Hw};
    const LeanLists ll{a.Hi, a.Hw, a.sec_loff, a.stage_w, a.stage_i, a.G, a.Ls, a.n};
    if (__ballot(mine && !(prev.i <= cur.i && prev.j <= cur.j)) != 0ull) {
        if (mine) {
            if (a.lean_lights) pack_chain<AOS>(a, ll, prev, cur, W_N);
            else pack_chain<AOS>(a, gl, prev, cur, W_N);
        }
        return;
    }
    // The pool: weights (and a spare slot), items, aliases.
    const int cap = a.pack_cap;
    double* sW = lds_raw;
    int32_t* sI = reinterpret_cast<int32_t*>(sW + cap + 1);
    int32_t* sA = sI + cap;
    int ta = 0;
    int64_t lsec = -1;  // lean lists: the section of the next light to load
    while (ta < nch) {
        const int Ia = __shfl(prev.i, ta, 64), Ja = __shfl(prev.j, ta, 64);
        const bool fits = mine && lane >= ta && (cur.i - Ia) + (cur.j - Ja + 1) <= cap;
        const int tb = ta + __popcll(__ballot(fits));
        if (tb - ta < kPackMinStagedChains && !a.lean_lights) {
            // Long chains: so few fit the pool at once that the rounds would
            // run the wave's chains nearly one after another; every lane
            // runs its own chain from the lists in HBM instead, all at once.
            if (mine && lane >= ta) pack_chain<AOS>(a, gl, prev, cur, W_N);
            break;
        }
        if (tb == ta) {
            if (lane == ta) {
                if (a.lean_lights) pack_chain<AOS>(a, ll, prev, cur, W_N);
                else pack_chain<AOS>(a, gl, prev, cur, W_N);
            }
            ta += 1;
            continue;
        }
        const int Ib = __shfl(cur.i, tb - 1, 64), Jb = __shfl(cur.j, tb - 1, 64);
        const int nLu = Ib - Ia, nHu = Jb - Ja + 1;
        for (int hv = 0; hv < 2; ++hv) {
            const int m = hv ? nHu : nLu, base = hv ? nLu : 0;
            const double* gw = hv ? a.Hw + Ja : a.Lw + Ia;
            const int32_t* gi = hv ? a.Hi + Ja : a.Li + Ia;
            const bool lean = hv == 0 && a.lean_lights;
            for (int t0 = 0; t0 < m; t0 += 8 * 64) {
                double wv[8];
                int32_t iv[8];
                if (lean) {
                    // The lights from the staging of their sections.
                    const uint32_t g0 = (uint32_t)(Ia + t0);
                    if (lsec < 0) lsec = a.sec_ldir[g0 >> kLightDirShift];
                    int32_t pos[8];
                    lsec = light_positions<8>(a, lsec, g0, m - t0, lane, pos);
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int t = t0 + u * 64 + lane;
                        if (t < m) { wv[u] = a.stage_w[pos[u]]; iv[u] = a.stage_i[pos[u]]; }
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int t = t0 + u * 64 + lane;
                        if (t < m) { wv[u] = gw[t]; iv[u] = gi[t]; }
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int t = t0 + u * 64 + lane;
                    if (t < m) { sW[base + t] = wv[u]; sI[base + t] = iv[u]; }
                }
            }
        }
        wave_lds_sync();
        if (lane >= ta && lane < tb) pack_chain_pool(sW, sI, sA, cap, nLu, Ia, Ja, prev, cur, W_N);
        wave_lds_sync();
        // The rows of the round in pool order (lights, then heavies but the
        // last, which the next round's first chain starts from), four per
        // lane at once; items rise along each list.
        const int nRows = nLu + nHu - 1;
        for (int r0 = 0; r0 < nRows; r0 += 4 * 64) {
            int it[4], al[4];
            double wv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = min(r0 + 64 * u + lane, nRows - 1);
                it[u] = sI[r];
                wv[u] = sW[r];
                al[u] = sA[r];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (r0 + 64 * u + lane < nRows && (uint32_t)it[u] < (uint32_t)a.n) store_row<AOS>(a, it[u], wv[u], al[u]);
        }
        wave_lds_sync();
        ta = tb;
    }
}


}  // namespace atg
