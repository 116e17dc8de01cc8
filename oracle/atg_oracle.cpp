/*
 * CPU oracle: sequential (optionally OpenMP-parallel, bit-identical) restatement
 * of the reference alias-table construction and samplers under the canonical
 * determinism contract.  TEST INFRASTRUCTURE ONLY - never linked into the product.
 *
 * Build: make -C oracle   (g++ -O2 -ffp-contract=off, no fast-math)
 */
#include "atg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#if defined(__FP_FAST_FMA) || defined(__FAST_MATH__)
#error "the oracle must be compiled without fast-math and without FP contraction"
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11; the Random123 reference method).  Replaces the reference's cuRAND
 * XORWOW/MTGP32 streams (utils/Prng.cuh:12-94), which are seeded from clock64()
 * and therefore not reproducible. */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

extern "C" void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

static inline void philox_u64(uint64_t idx, uint32_t purpose, uint32_t extra, uint64_t seed, uint32_t out[4]) {
    const uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), purpose, extra};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, out);
}

static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}
static inline double coin53(const uint32_t x[4]) {
    const uint64_t r = ((uint64_t)x[3] << 32) | x[2];
    return (double)(r >> 11) * 0x1.0p-53;
}
static inline uint64_t word64(const uint32_t x[4]) { return ((uint64_t)x[1] << 32) | x[0]; }

/* ------------------------------------------------------------------------ */
/* Canonical sums.  Replaces CUB DeviceReduce::Sum (utils/Adder.cu:144-157) and
 * the sequential sumWeightsCpu (construction/SamplingAlgorithm.cu:14-19) with a
 * fixed association: per section, lane l of a 64-lane wave accumulates items
 * l, l+64, l+128, ... sequentially; the 64 lane sums meet in a pairwise tree;
 * the section sums meet in a recursive 64-ary pairwise tree. */
static double tree64(const double* v, int64_t count) {
    double t[64];
    for (int l = 0; l < 64; ++l) t[l] = (l < count) ? v[l] : 0.0;
    for (int width = 1; width < 64; width *= 2)
        for (int m = 0; m < 64; m += 2 * width) t[m] = t[m] + t[m + width];
    return t[0];
}

extern "C" double oracle_tree64_reduce(const double* v, int64_t count) {
    if (count <= 0) return 0.0;
    std::vector<double> cur(v, v + count);
    while (cur.size() > 1) {
        std::vector<double> nxt((cur.size() + 63) / 64);
        for (size_t g = 0; g < nxt.size(); ++g)
            nxt[g] = tree64(cur.data() + 64 * g, (int64_t)std::min<size_t>(64, cur.size() - 64 * g));
        cur.swap(nxt);
    }
    return cur[0];
}

extern "C" double oracle_section_sum(const double* x, int64_t len) {
    double acc[64];
    for (int l = 0; l < 64; ++l) {
        double a = 0.0;
        for (int64_t k = l; k < len; k += 64) a = a + x[k];
        acc[l] = a;
    }
    return tree64(acc, 64);
}

/* Section mapping of PSA+: gridDim = N / sectionSize + 1 blocks
 * (AliasTableSplitGpu.cu:356-357), each covering N / gridDim + 1 items
 * (PsaPlusNormal.cu:27-29).  The same sections define the canonical sum and the
 * leftover prefix association. */
extern "C" void oracle_section_layout(int64_t n, int64_t section_param, int64_t* num_sections,
                                      int64_t* section_len) {
    if (section_param < 1) section_param = 1;
    const int64_t G = n / section_param + 1;
    *num_sections = G;
    *section_len = n / G + 1;
}

extern "C" double oracle_weight_sum(const double* w, int64_t n, int64_t section_param,
                                   double* section_sums) {
    int64_t G, Ls;
    oracle_section_layout(n, section_param, &G, &Ls);
    std::vector<double> sums(G);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < G; ++s) {
        const int64_t a = std::min(n, s * Ls), b = std::min(n, (s + 1) * Ls);
        sums[s] = oracle_section_sum(w + a, b - a);
    }
    if (section_sums) std::memcpy(section_sums, sums.data(), sizeof(double) * G);
    return oracle_tree64_reduce(sums.data(), G);
}

/* Canonical exclusive scan of one section's list (replaces CUB
 * DeviceScan::ExclusiveSum at AliasTableSplitGpu.cu:406-430; the CPU
 * exclusivePrefixSum is utils/PrefixSum.cu:3-11).  Chunks of 64 values: an
 * inclusive scan in the order of a 64-lane wave with four rows of 16 lanes -
 * Kogge-Stone offsets 1, 2, 4, 8 inside each row, then rows 1 and 3 add the
 * last lane of the row before them, then rows 2 and 3 add lane 31 - and chunk
 * carries chained left to right.  Lane 63 of that scan is the pairwise tree of
 * the chunk.  Returns the list aggregate (the final carry). */
static void wave_scan64(double s[64]) {
    double t[64];
    for (int off = 1; off < 16; off *= 2) {
        std::memcpy(t, s, sizeof(t));
        for (int l = 0; l < 64; ++l)
            if ((l & 15) >= off) s[l] = t[l - off] + t[l];
    }
    std::memcpy(t, s, sizeof(t));
    for (int l = 16; l < 32; ++l) s[l] = t[15] + t[l];
    for (int l = 48; l < 64; ++l) s[l] = t[47] + t[l];
    std::memcpy(t, s, sizeof(t));
    for (int l = 32; l < 64; ++l) s[l] = t[31] + t[l];
}

extern "C" double oracle_local_scan(const double* v, int64_t m, double* excl) {
    double carry = 0.0;
    for (int64_t c0 = 0; c0 < m; c0 += 64) {
        double s[64];
        for (int l = 0; l < 64; ++l) s[l] = (c0 + l < m) ? v[c0 + l] : 0.0;
        wave_scan64(s);
        if (excl) {
            for (int l = 0; l < 64 && c0 + l < m; ++l) excl[c0 + l] = carry + (l == 0 ? 0.0 : s[l - 1]);
        }
        carry = carry + s[63];
    }
    return carry;
}

/* ------------------------------------------------------------------------ */
/* Split (SplitMethodBasic.cu:22-49, predicate SplitMethodBasic.cuh:9-41).
 * The reference loops until FOUND and never terminates when the FP predicate
 * has no FOUND point; this restatement follows the identical bisection path for
 * 64 probes (enough for any range below 2^31 to converge) and then falls back
 * to a deterministic clamp, reporting the miss. */
enum { FOUND = 0, LEFT = 1, RIGHT = 2 };

static inline int check_split(int64_t i, int64_t j, double target, const double* Lp, int64_t numL,
                              const double* Hp, int64_t numH, double* sigma) {
    if (j + 1 >= numH || i - 1 < 0) return LEFT;
    if (i >= numL || j < 0) return RIGHT;
    const double s = Lp[i] + Hp[j];
    const double sn = Lp[i - 1] + Hp[j + 1];
    *sigma = s;
    if (s <= target && sn > target) return FOUND;
    if (s <= target) return RIGHT;
    return LEFT;
}

extern "C" int oracle_split_one(int64_t k, int64_t p, int64_t nToSplit, double wToSplit,
                                const double* Lp, int64_t numL, const double* Hp, const double* Hw,
                                int64_t numH, int32_t* out_i, int32_t* out_j, double* out_spill) {
    const int64_t n_ = (int64_t)std::ceil(((double)nToSplit * (double)k) / (double)p);
    const int64_t b0 = std::min(n_, numH) - 1;
    const double target = (wToSplit * (double)n_) / (double)nToSplit;
    int64_t a = 0, b = b0;
    double sigma = 0.0;
    for (int it = 0; it < 64; ++it) {
        const int64_t j = (a + b) / 2;
        const int64_t i = std::min(n_ - j, numL - 1);
        const int st = check_split(i, j, target, Lp, numL, Hp, numH, &sigma);
        if (st == FOUND) {
            *out_i = (int32_t)i; *out_j = (int32_t)j;
            *out_spill = sigma + Hw[j] - target;
            return 1;
        }
        if (st == RIGHT) a = j + 1; else b = j;
    }
    /* Fallback: deterministic clamp (never reached when FOUND exists). */
    int64_t j = std::min(std::max<int64_t>(a, 0), std::max<int64_t>(b0, 0));
    int64_t i = std::min(n_ - j, numL - 1);
    if (i < 0) i = 0;
    *out_i = (int32_t)i; *out_j = (int32_t)j;
    *out_spill = (Lp[i] + Hp[j]) + Hw[j] - target;
    return 0;
}

/* Pack one section (PackMethodBasic.cuh:59-126, packOptimized with PackAll). */
static void pack_section(int32_t iLower, int32_t iUpper, int32_t jLower, int32_t jUpper, double spill,
                         double W_N, const int32_t* Li, const double* Lw, const int32_t* Hi,
                         const double* Hw, int64_t n, double* row_w, int32_t* row_alias) {
    int64_t i = iLower, j = jLower;
    double w = spill;
    if (spill == 0) w = Hw[j];
    while (true) {
        bool ex = w <= W_N + ORACLE_EPSILON;
        if (ex && j >= jUpper) ex = false;
        else if (!ex && i >= iUpper) ex = true;
        if ((ex && j >= jUpper) || (!ex && i >= iUpper)) return;
        int32_t item, alias;
        double rw, next;
        if (ex) {
            item = Hi[j]; rw = w; alias = Hi[j + 1]; next = Hw[j + 1]; ++j;
        } else {
            item = Li[i]; rw = Lw[i]; alias = Hi[j]; next = Lw[i]; ++i;
        }
        w += next - W_N;
        if (item >= 0 && item < n) { row_w[item] = rw; row_alias[item] = alias; }
    }
}

/* ------------------------------------------------------------------------ */
/* PSA+ section work (PsaPlusNormal.cu:12-152) made deterministic: stable
 * in-section L/H order (the reference fills them with shared atomics,
 * :31-44), count-sliced greedy chains (:51-111), remainder rule (:112-123),
 * leftovers concatenated in chain order (:129-150; the reference appends them
 * with atomicAdd_system in arbitrary order).  One divergence: when the
 * remainder belongs to L but its chain consumed no light item, the reference
 * writes l[iMin-1] (out of bounds for chain 0, the neighbour's leftover
 * otherwise); here it stays in H. */
struct SectionOut {
    int64_t nL, nH;  /* leftover counts, stored at temp[start, start+n) */
};

static SectionOut process_section(const double* w, int64_t start, int64_t len, double W_N,
                                  int32_t psa_plus, int32_t C, double* row_w, int32_t* row_alias,
                                  int32_t* tLi, double* tLw, int32_t* tHi, double* tHw) {
    std::vector<int32_t> li, hi;
    std::vector<double> lw, hw;
    li.reserve(len); lw.reserve(len); hi.reserve(len); hw.reserve(len);
    for (int64_t k = 0; k < len; ++k) {
        const double x = w[start + k];
        if (x <= W_N) { li.push_back((int32_t)(start + k)); lw.push_back(x); }
        else { hi.push_back((int32_t)(start + k)); hw.push_back(x); }
    }
    const int64_t nL = (int64_t)li.size(), nH = (int64_t)hi.size();
    SectionOut out{0, 0};
    auto emitL = [&](int64_t a, int64_t b) {
        for (int64_t t = a; t < b; ++t) { tLi[start + out.nL] = li[t]; tLw[start + out.nL] = lw[t]; ++out.nL; }
    };
    auto emitH = [&](int64_t a, int64_t b) {
        for (int64_t t = a; t < b; ++t) { tHi[start + out.nH] = hi[t]; tHw[start + out.nH] = hw[t]; ++out.nH; }
    };
    if (!psa_plus || C <= 0) {
        emitL(0, nL); emitH(0, nH);
        return out;
    }
    const int64_t stepsL = nL / C, stepsH = nH / C;
    std::vector<int64_t> iMinV(C), iMaxV(C), jMinV(C), jMaxV(C);
    for (int t = 0; t < C; ++t) {
        iMinV[t] = t * stepsL; iMaxV[t] = (t == C - 1) ? nL : (t + 1) * stepsL;
        jMinV[t] = t * stepsH; jMaxV[t] = (t == C - 1) ? nH : (t + 1) * stepsH;
    }
    for (int t = 0; t < C; ++t) {
        const int64_t iMin = iMinV[t], iMax = iMaxV[t], jMin = jMinV[t], jMax = jMaxV[t];
        if (!(iMax - iMin > 0 && jMax - jMin > 0)) continue;
        int64_t i = iMin, j = jMin;
        double wc = hw[j];
        while (true) {
            const bool ex = wc <= W_N + ORACLE_EPSILON;
            if ((ex && j >= jMax - 1) || (!ex && i >= iMax - 1)) break;
            int32_t item, alias;
            double rw, next;
            if (ex) { item = hi[j]; rw = wc; alias = hi[j + 1]; next = hw[j + 1]; ++j; }
            else { item = li[i]; rw = lw[i]; alias = hi[j]; next = lw[i]; ++i; }
            wc += next - W_N;
            row_w[item] = rw; row_alias[item] = alias;
        }
        int64_t jm = j + 1, im = i;
        if (wc <= W_N + ORACLE_EPSILON && i > iMin) {
            --im; li[im] = hi[j]; lw[im] = wc;
        } else {
            --jm; hw[jm] = wc; /* hi[jm] == hi[j] already */
        }
        iMinV[t] = im; jMinV[t] = jm;
    }
    for (int t = 0; t < C; ++t) emitL(iMinV[t], iMaxV[t]);
    for (int t = 0; t < C; ++t) emitH(jMinV[t], jMaxV[t]);
    return out;
}

/* Exclusive section prefixes of the leftover sums (replaces the CUB scan
 * association, AliasTableSplitGpu.cu:406-430, which is unpinned).  Block sums
 * over aligned dyadic ranges of sections: T(0,c) = agg_c and
 * T(b,c) = T(b-1,2c) + T(b-1,2c+1) for complete ranges [c 2^b, (c+1) 2^b) in
 * [0, G).  P(s) folds, from the highest set bit of s to the lowest, the block
 * that starts where the previous one ended: P = 0; P = P + T(b, pos >> b);
 * pos += 2^b.  P(G) is the total.  Depth O(log G), and every P(s) depends only
 * on sections before s. */
static void section_prefixes(const std::vector<double>& agg, std::vector<double>& P) {
    const int64_t G = (int64_t)agg.size();
    std::vector<std::vector<double>> T;
    T.push_back(agg);
    for (int b = 1; ((int64_t)1 << b) <= G; ++b) {
        const std::vector<double>& prev = T.back();
        std::vector<double> cur(G >> b);
        for (int64_t c = 0; c < (int64_t)cur.size(); ++c) cur[c] = prev[2 * c] + prev[2 * c + 1];
        T.push_back(cur);
    }
    P.assign(G + 1, 0.0);
    for (int64_t s = 0; s <= G; ++s) {
        double acc = 0.0;
        int64_t pos = 0;
        for (int b = 62; b >= 0; --b) {
            if (!((s >> b) & 1)) continue;
            acc = acc + T[b][pos >> b];
            pos += (int64_t)1 << b;
        }
        P[s] = acc;
    }
}

extern "C" int oracle_build(const double* w, int64_t n, const oracle_opts* o, double* row_w,
                            int32_t* row_alias, int32_t* L_item, double* L_w, double* L_pref,
                            int32_t* H_item, double* H_w, double* H_pref, int32_t* split_i,
                            int32_t* split_j, double* split_spill, oracle_stats* st) {
    std::memset(st, 0, sizeof(*st));
    if (n < 1 || n > 0x7FFFFFF0LL) return 1;
#ifdef _OPENMP
    if (o->threads > 0) omp_set_num_threads(o->threads);
#endif
    int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double x = w[i];
        if (!(x >= 0.0) || !std::isfinite(x)) bad |= 1;
    }
    if (bad) return 1;

    int64_t G, Ls;
    oracle_section_layout(n, o->section_size, &G, &Ls);
    st->num_sections = G; st->section_len = Ls;
    const double W = oracle_weight_sum(w, n, o->section_size, nullptr);
    if (!(W > 0.0) || !std::isfinite(W)) return 1;
    const double W_N = W / (double)n;
    st->W = W; st->W_N = W_N;

    /* Per-section partition / PSA+ into temporaries at the section's offset. */
    std::vector<int32_t> tLi(n), tHi(n);
    std::vector<double> tLw(n), tHw(n);
    std::vector<int64_t> cntL(G), cntH(G);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < G; ++s) {
        const int64_t a = std::min(n, s * Ls), b = std::min(n, (s + 1) * Ls);
        SectionOut so = process_section(w, a, b - a, W_N, o->psa_plus, o->chains, row_w, row_alias,
                                        tLi.data(), tLw.data(), tHi.data(), tHw.data());
        cntL[s] = so.nL; cntH[s] = so.nH;
    }
    /* Local scans (per section), then chained section prefixes
     * P_{s+1} = P_s + agg_s (what decoupled look-back reproduces bit for bit). */
    std::vector<double> exL(n), exH(n), aggL(G), aggH(G);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < G; ++s) {
        const int64_t a = std::min(n, s * Ls);
        aggL[s] = oracle_local_scan(tLw.data() + a, cntL[s], exL.data() + a);
        aggH[s] = oracle_local_scan(tHw.data() + a, cntH[s], exH.data() + a);
    }
    std::vector<int64_t> offL(G + 1), offH(G + 1);
    std::vector<double> PL(G + 1), PH(G + 1);
    offL[0] = offH[0] = 0;
    for (int64_t s = 0; s < G; ++s) { offL[s + 1] = offL[s] + cntL[s]; offH[s + 1] = offH[s] + cntH[s]; }
    section_prefixes(aggL, PL);
    section_prefixes(aggH, PH);
    const int64_t nLr = offL[G], nHr = offH[G];
    std::vector<int32_t> Li(nLr + 1), Hi(nHr + 1);
    std::vector<double> Lw(nLr + 1), Hw(nHr + 1), Lp(nLr + 1), Hp(nHr + 1);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < G; ++s) {
        const int64_t a = std::min(n, s * Ls);
        for (int64_t r = 0; r < cntL[s]; ++r) {
            Li[offL[s] + r] = tLi[a + r]; Lw[offL[s] + r] = tLw[a + r]; Lp[offL[s] + r] = PL[s] + exL[a + r];
        }
        for (int64_t r = 0; r < cntH[s]; ++r) {
            Hi[offH[s] + r] = tHi[a + r]; Hw[offH[s] + r] = tHw[a + r]; Hp[offH[s] + r] = PH[s] + exH[a + r];
        }
    }
    /* Sentinels: light {N+1, 0} ends L, heavy {N, +inf} ends H
     * (AliasTableSplitGpu.cu:366-385; AliasTableSplit.cu:26-27). */
    Li[nLr] = (int32_t)(n + 1); Lw[nLr] = 0.0; Lp[nLr] = PL[G];
    Hi[nHr] = (int32_t)n; Hw[nHr] = std::numeric_limits<double>::infinity(); Hp[nHr] = PH[G];
    const int64_t numL = nLr + 1, numH = nHr + 1;
    st->numL = numL; st->numH = numH; st->R = nLr + nHr;

    /* Split certificate: max fl(dL) < min fl(dH) makes the split predicate
     * monotone, so every exact search finds the bisection's j. */
    double maxdL = 0.0, mindH = std::numeric_limits<double>::infinity();
    for (int64_t t = 0; t + 1 < numL; ++t) maxdL = std::max(maxdL, Lp[t + 1] - Lp[t]);
    for (int64_t t = 0; t + 1 < numH; ++t) mindH = std::min(mindH, Hp[t + 1] - Hp[t]);
    st->max_dL = maxdL; st->min_dH = mindH; st->monotone = maxdL < mindH ? 1 : 0;

    if (L_item) { std::memcpy(L_item, Li.data(), 4 * numL); std::memcpy(L_w, Lw.data(), 8 * numL); std::memcpy(L_pref, Lp.data(), 8 * numL); }
    if (H_item) { std::memcpy(H_item, Hi.data(), 4 * numH); std::memcpy(H_w, Hw.data(), 8 * numH); std::memcpy(H_pref, Hp.data(), 8 * numH); }

    /* Degenerate: one class of leftovers is empty (all-equal weights make the
     * reference assert numH > 1, AliasTableSplit.cu:43).  Every leftover item
     * then carries its own mass, within rounding of W/N: row = {w, self}. */
    if (nLr == 0 || nHr == 0) {
        st->degenerate = 1; st->p = 0;
        for (int64_t t = 0; t < nLr; ++t) { row_w[Li[t]] = Lw[t]; row_alias[Li[t]] = Li[t]; }
        for (int64_t t = 0; t < nHr; ++t) { row_w[Hi[t]] = Hw[t]; row_alias[Hi[t]] = Hi[t]; }
        return 0;
    }
    /* p clamps: AliasTableSplitGpu.cu:68-74 (min with N/10, at least 2) and,
     * after PSA+, :382-385 (min with (numL+numH)/10, at least 2). */
    int64_t p = std::min<int64_t>(o->num_splits, n / 10);
    if (p < 2) p = 2;
    if (o->psa_plus) { p = std::min<int64_t>(p, (numL + numH) / 10); if (p < 2) p = 2; }
    st->p = p;
    const int64_t nToSplit = numL + numH - 2;
    const double wToSplit = Lp[numL - 1] + Hp[numH - 1];
    st->nToSplit = nToSplit; st->wToSplit = wToSplit;

    std::vector<int32_t> si(p + 1), sj(p + 1);
    std::vector<double> ss(p + 1);
    si[0] = 0; sj[0] = 0; ss[0] = 0.0;
    si[p] = (int32_t)(numL - 1); sj[p] = (int32_t)(numH - 1); ss[p] = 0.0;
    int64_t fallbacks = 0;
#pragma omp parallel for reduction(+ : fallbacks) schedule(static)
    for (int64_t k = 1; k < p; ++k) {
        if (!oracle_split_one(k, p, nToSplit, wToSplit, Lp.data(), numL, Hp.data(), Hw.data(), numH,
                              &si[k], &sj[k], &ss[k]))
            ++fallbacks;
    }
    st->split_fallbacks = fallbacks;
    if (split_i) { std::memcpy(split_i, si.data(), 4 * (p + 1)); std::memcpy(split_j, sj.data(), 4 * (p + 1)); std::memcpy(split_spill, ss.data(), 8 * (p + 1)); }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 1; k <= p; ++k)
        pack_section(si[k - 1], si[k], sj[k - 1], sj[k], ss[k - 1], W_N, Li.data(), Lw.data(), Hi.data(),
                     Hw.data(), n, row_w, row_alias);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Implied probability (AliasTable.cu:24-65), accumulated exactly in binary128.
 * Mass follows the sampler's rule: a row keeps min(w, W/N); when w < W/N the
 * rest goes to the alias, or stays on the row when the alias is not an item
 * (the reference's heavy sentinel N can appear as an alias). */
extern "C" void oracle_verify_table(const double* w, int64_t n, double W, double W_N, const double* row_w,
                                   const int32_t* row_alias, double* max_rel_err, double* max_abs_mass_err,
                                   int64_t* rows_out_of_range, int32_t* reference_check_ok) {
    std::vector<__float128> m(n, 0);
    int64_t oor = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double rw = row_w[i];
        const int32_t a = row_alias[i];
        if (rw < W_N) {
            m[i] += (__float128)rw;
            if (a >= 0 && a < n) m[a] += (__float128)W_N - (__float128)rw;
            else { m[i] += (__float128)W_N - (__float128)rw; ++oor; }
        } else {
            m[i] += (__float128)W_N;
        }
    }
    const __float128 total = (__float128)W_N * (__float128)n;
    double mrel = 0.0, mabs = 0.0;
    const double maxError = std::max((double)n * 1e-8, 1e-3);
    int ok = 1;
    for (int64_t i = 0; i < n; ++i) {
        const __float128 pi = m[i] / total;
        const __float128 pt = (__float128)w[i] / (__float128)W;
        __float128 d = pi - pt;
        if (d < 0) d = -d;
        double rel = (w[i] > 0.0) ? (double)(d / pt) : (double)d;
        mrel = std::max(mrel, rel);
        /* Reference check: absolute mass delta against max(N*1e-8, 1e-3) and
         * row weight within [0, W/N (1 + maxError)]. */
        __float128 dm = m[i] - (__float128)w[i];
        if (dm < 0) dm = -dm;
        mabs = std::max(mabs, (double)dm);
        if ((double)dm > maxError) ok = 0;
        if (row_w[i] < 0 || row_w[i] > W_N * (1.0 + maxError)) ok = 0;
    }
    *max_rel_err = mrel;
    *max_abs_mass_err = mabs;
    *rows_out_of_range = oor;
    *reference_check_ok = ok;
}

/* ------------------------------------------------------------------------ */
/* Samplers.  Decision rule of doSample (SamplerGpuBasic.cu:148-166): take the
 * alias iff row.w < W/N and coin * W/N >= row.w.  Row and coin come from one
 * Philox4x32-10 block: row = floor(x64 * len / 2^64) from words 0-1 (a 128-bit
 * product, so rows above 2^24 are reachable, unlike the reference's float CPU
 * uniform, Prng.cuh:180-188), coin = 53-bit double from words 2-3.  An alias
 * that is not an item (the heavy sentinel) keeps the row. */
static inline int32_t decide(double rw, int32_t alias, int64_t row, double u, double W_N, int64_t n) {
    if (rw < W_N && u * W_N >= rw && alias >= 0 && alias < n) return alias;
    return (int32_t)row;
}

static void sample_rows(const double* row_w, const int32_t* row_alias, int64_t n, double W_N, uint64_t seed,
                        uint32_t purpose, uint64_t first, uint64_t count, int32_t* out) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)count; ++t) {
        uint32_t x[4];
        philox_u64(first + (uint64_t)t, purpose, 0, seed, x);
        const int64_t row = (int64_t)mulhi64(word64(x), (uint64_t)n);
        out[t] = decide(row_w[row], row_alias[row], row, coin53(x), W_N, n);
    }
}

extern "C" void oracle_sample_basic(const double* row_w, const int32_t* row_alias, int64_t n, double W_N,
                                   uint64_t seed, uint64_t first, uint64_t count, int32_t* out) {
    sample_rows(row_w, row_alias, n, W_N, seed, ORACLE_PURPOSE_BASIC, first, count, out);
}

/* Samples of one table of a batch: the Philox counters are the global output
 * positions first .. first + count - 1 under the batch purpose. */
extern "C" void oracle_sample_batch_table(const double* row_w, const int32_t* row_alias, int64_t n, double W_N,
                                          uint64_t seed, uint64_t first, uint64_t count, int32_t* out) {
    sample_rows(row_w, row_alias, n, W_N, seed, ORACLE_PURPOSE_BATCH, first, count, out);
}

/* Binomial draw for the sectioned sample counts.  The reference halves the
 * block range with a float normal approximation (SamplerGpuSectioned.cu:58-96,
 * Binomial::draw at Prng.cuh:190-199).  Restated with only IEEE + - * / sqrt
 * floor so CPU and GPU agree bit for bit: n <= 32 -> n Bernoulli(q) trials on
 * 32-bit uniforms; otherwise round(n q + sqrt(n q (1-q)) z) clamped to [0, n],
 * z an Irwin-Hall(12) normal (sum of 12 uniforms - 6).  Totals are exact. */
static uint64_t binomial_purpose(uint64_t n, double q, uint64_t seed, uint64_t node, uint32_t purpose) {
    if (n == 0) return 0;
    if (n <= 32) {
        uint64_t cnt = 0;
        uint32_t x[4];
        for (uint64_t t = 0; t < n; ++t) {
            if ((t & 3) == 0) philox_u64(node, purpose, (uint32_t)(t >> 2), seed, x);
            const double u = (double)x[t & 3] * 0x1.0p-32;
            if (u < q) ++cnt;
        }
        return cnt;
    }
    double z = 0.0;
    for (uint32_t d = 0; d < 3; ++d) {
        uint32_t x[4];
        philox_u64(node, purpose, d, seed, x);
        for (int k = 0; k < 4; ++k) z = z + (double)x[k] * 0x1.0p-32;
    }
    z = z - 6.0;
    const double mean = (double)n * q;
    const double var = mean * (1.0 - q);
    const double sd = std::sqrt(var);
    double r = std::floor((mean + sd * z) + 0.5);
    if (r < 0.0) r = 0.0;
    if (r > (double)n) r = (double)n;
    return (uint64_t)r;
}

extern "C" uint64_t oracle_binomial(uint64_t n, double q, uint64_t seed, uint64_t node) {
    return binomial_purpose(n, q, seed, node, ORACLE_PURPOSE_SECTION_COUNTS);
}

static void counts_rec(int64_t lo, int64_t hi, uint64_t cnt, uint64_t node, int64_t n, int64_t SS,
                       uint64_t seed, uint64_t* counts) {
    if (hi - lo == 1) { counts[lo] = cnt; return; }
    const int64_t mid = lo + (hi - lo) / 2;
    const int64_t rl = std::min(n, mid * SS) - lo * SS;
    const int64_t rt = std::min(n, hi * SS) - lo * SS;
    const double q = (double)rl / (double)rt;
    const uint64_t x = oracle_binomial(cnt, q, seed, node);
    counts_rec(lo, mid, x, 2 * node, n, SS, seed, counts);
    counts_rec(mid, hi, cnt - x, 2 * node + 1, n, SS, seed, counts);
}

extern "C" void oracle_section_counts(int64_t n, int64_t section_rows, uint64_t total, uint64_t seed,
                                     uint64_t* counts) {
    const int64_t nsec = (n + section_rows - 1) / section_rows;
    counts_rec(0, nsec, total, 1, n, section_rows, seed, counts);
}

/* Sectioned sampling (SamplerGpuSectioned.cu:98-136): section s holds rows
 * [s*SS, min(N,(s+1)*SS)) and draws counts[s] samples written contiguously
 * in section order; sample g uses Philox counter g. */
extern "C" void oracle_sample_sectioned(const double* row_w, const int32_t* row_alias, int64_t n,
                                       double W_N, int64_t SS, uint64_t total, uint64_t seed, int32_t* out) {
    const int64_t nsec = (n + SS - 1) / SS;
    std::vector<uint64_t> counts(nsec);
    oracle_section_counts(n, SS, total, seed, counts.data());
    std::vector<uint64_t> off(nsec + 1, 0);
    for (int64_t s = 0; s < nsec; ++s) off[s + 1] = off[s] + counts[s];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsec; ++s) {
        const int64_t start = s * SS, len = std::min(n, (s + 1) * SS) - start;
        for (uint64_t t = 0; t < counts[s]; ++t) {
            const uint64_t g = off[s] + t;
            uint32_t x[4];
            philox_u64(g, ORACLE_PURPOSE_SECTIONED, 0, seed, x);
            const int64_t row = start + (int64_t)mulhi64(word64(x), (uint64_t)len);
            out[g] = decide(row_w[row], row_alias[row], row, coin53(x), W_N, n);
        }
    }
}

/* Rejection build (RejectionSampling.cu:26-56): c_i = floor(N w_i / W) + 1,
 * inclusive scan, A[s] = item whose run holds slot s (fillA,
 * RejectionSamplingBinarySearch.cu:12-33; here filled directly). */
extern "C" int oracle_rejection_build(const double* w, int64_t n, double W, uint64_t* incl, int32_t* A,
                                     uint64_t* A_size) {
    uint64_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double x = ((double)n * w[i]) / W;
        acc += (uint64_t)std::floor(x) + 1;
        incl[i] = acc;
    }
    *A_size = acc;
    if (A) {
        uint64_t s = 0;
        for (int64_t i = 0; i < n; ++i)
            for (; s < incl[i]; ++s) A[s] = (int32_t)i;
    }
    return 0;
}

/* Rejection sampler (SamplerRejection.cu:17-35): slot uniform over A; a slot
 * that starts its item's run is accepted with probability frac(N w / W).  The
 * reference evaluates the fraction in float; restated in double (the float form
 * caps the per-item probability accuracy near 1e-7).  Trial t of sample g uses
 * Philox counter (g, purpose, t); the loop is bounded (acceptance per trial is
 * at least 1/2, so the bound is never reached in practice). */
extern "C" void oracle_sample_rejection(const int32_t* A, uint64_t A_size, const double* w, int64_t n,
                                       double W, uint64_t seed, uint64_t first, uint64_t count,
                                       int32_t* out, uint64_t* total_trials) {
    uint64_t trials = 0;
#pragma omp parallel for reduction(+ : trials) schedule(static)
    for (int64_t t = 0; t < (int64_t)count; ++t) {
        const uint64_t g = first + (uint64_t)t;
        int32_t item = 0;
        for (uint32_t tr = 0; tr < ORACLE_REJECTION_MAX_TRIALS; ++tr) {
            uint32_t x[4];
            philox_u64(g, ORACLE_PURPOSE_REJECTION, tr, seed, x);
            ++trials;
            const uint64_t slot = mulhi64(word64(x), A_size);
            item = A[slot];
            if (slot == 0 || A[slot - 1] != item) {
                const double xv = ((double)n * w[item]) / W;
                const double frac = xv - std::floor(xv);
                if (coin53(x) > frac) continue;
            }
            break;
        }
        out[t] = item;
    }
    if (total_trials) *total_trials = trials;
}

/* ------------------------------------------------------------------------ */
/* Faithful reference CPU builds (baselines; sequential sums as the reference). */

/* AliasTableSplit::build (AliasTableSplit.cu:24-80): sequential W, stable
 * partition over N+2 items with sentinels, sequential exclusive scans
 * (PrefixSum.cu:3-11), SplitMethodBasic::split for k in [1,p) and
 * PackMethodSweep::pack (PackMethodSweep.cu:3-65) for k in [1,p]. */
extern "C" void oracle_ref_split_build(double* wt, int64_t N, int64_t p, double* W_out, double* row_w,
                                      int32_t* row_alias) {
    double W = 0;
    for (int64_t i = 0; i < N; ++i) W += wt[i];
    wt[N] = std::numeric_limits<double>::infinity();
    wt[N + 1] = 0;
    for (int64_t i = 0; i < N; ++i) row_w[i] = wt[i];
    if (p > N / 2) p = N / 10;
    if (p < 2) p = 2;
    const double W_N = W / N;
    std::vector<int64_t> pre(N + 2);
    int64_t cnt = 0;
    for (int64_t i = 0; i < N + 2; ++i) { pre[i] = cnt; cnt += (wt[i] > W / N) ? 1 : 0; }
    const int64_t numH = cnt, numL = (N + 2) - numH;
    std::vector<int32_t> l(numL), h(numH);
    std::vector<double> lw(numL), hw(numH), pL(numL), pH(numH);
    for (int64_t i = 0; i < N + 2; ++i) {
        if (wt[i] > W / N) { h[pre[i]] = (int32_t)i; hw[pre[i]] = wt[i]; }
        else { l[i - pre[i]] = (int32_t)i; lw[i - pre[i]] = wt[i]; }
    }
    double s = 0;
    for (int64_t i = 0; i < numL; ++i) { pL[i] = s; s += lw[i]; }
    s = 0;
    for (int64_t i = 0; i < numH; ++i) { pH[i] = s; s += hw[i]; }
    std::vector<int32_t> si(p + 1), sj(p + 1);
    std::vector<double> ss(p + 1);
    si[0] = 0; sj[0] = 0; ss[0] = 0;
    si[p] = (int32_t)(numL - 1); sj[p] = (int32_t)(numH - 1); ss[p] = 0;
    for (int64_t k = 1; k < p; ++k)
        oracle_split_one(k, p, N, W, pL.data(), numL, pH.data(), hw.data(), numH, &si[k], &sj[k], &ss[k]);
    /* PackMethodSweep::pack walks the raw weight array between the item indices
     * of the section bounds, skipping the other class. */
    for (int64_t k = 1; k <= p; ++k) {
        int64_t i = l[si[k - 1]], j = h[sj[k - 1]];
        const int64_t iMax = l[si[k]], jMax = h[sj[k]];
        double w = ss[k - 1];
        if (ss[k - 1] == 0) w = hw[sj[k - 1]];
        while (true) {
            bool ex = w <= W_N + ORACLE_EPSILON;
            if (ex && j >= jMax) ex = false;
            else if (!ex && i >= iMax) ex = true;
            if (ex) {
                if (j >= jMax) break;
                int64_t j2 = j;
                do { ++j2; } while (j < jMax && wt[j2] <= W_N);
                row_w[j] = w; row_alias[j] = (int32_t)j2;
                w += wt[j2] - W_N;
                j = j2;
            } else {
                if (i >= iMax) break;
                row_w[i] = wt[i]; row_alias[i] = (int32_t)j;
                w += row_w[i] - W_N;
                do { ++i; } while (i < iMax && wt[i] > W_N);
            }
        }
    }
    *W_out = W;
}

/* AliasTableSweep::build (AliasTableSweep.cu:12-60). */
extern "C" void oracle_ref_sweep_build(double* wt, int64_t N, double* W_out, double* row_w,
                                      int32_t* row_alias) {
    double W = 0;
    for (int64_t i = 0; i < N; ++i) W += wt[i];
    wt[N] = std::numeric_limits<double>::infinity();
    wt[N + 1] = 0;
    int64_t i = 0;
    while (wt[i] > W / N) ++i;
    int64_t j = 0;
    while (wt[j] <= W / N) ++j;
    double w = wt[j];
    while (true) {
        bool ex = w <= W / N + ORACLE_EPSILON;
        if (ex && j >= N) ex = false;
        else if (!ex && i >= N) ex = true;
        if (ex) {
            if (j >= N) break;
            row_w[j] = w;
            int64_t j2 = j;
            do { ++j2; } while (j < N + 2 && wt[j2] <= W / N);
            row_alias[j] = (int32_t)j2;
            w += wt[j2] - W / N;
            j = j2;
        } else {
            if (i >= N) break;
            row_w[i] = wt[i]; row_alias[i] = (int32_t)j;
            w += wt[i] - W / N;
            do { ++i; } while (i < N + 2 && wt[i] > W / N);
        }
    }
    *W_out = W;
}

/* AliasTableStack::build (AliasTableStack.cu:11-39), Vose-style stacks, in FP64
 * (the reference compares with (float)W/N at :31). */
extern "C" void oracle_ref_stack_build(const double* wt, int64_t N, double* W_out, double* row_w,
                                      int32_t* row_alias) {
    double W = 0;
    for (int64_t i = 0; i < N; ++i) W += wt[i];
    std::vector<int32_t> l, h;
    l.reserve(N); h.reserve(N);
    for (int64_t i = 0; i < N; ++i) {
        if (wt[i] > W / N) h.push_back((int32_t)i); else l.push_back((int32_t)i);
        row_w[i] = wt[i];
        row_alias[i] = (int32_t)i;
    }
    while (!h.empty() && !l.empty()) {
        const int32_t j = h.back(); h.pop_back();
        while (row_w[j] > W / N && !l.empty()) {
            const int32_t i = l.back(); l.pop_back();
            row_alias[i] = j;
            row_w[j] += row_w[i] - W / N;
        }
        l.push_back(j);
    }
    *W_out = W;
}

/* ------------------------------------------------------------------------ */
/* Batches of independent tables (test infrastructure for the engine's
 * atg_batch_*, which has no counterpart in the reference).  Table t owns
 * w[offsets[t], offsets[t+1]) and is built exactly as oracle_build builds it
 * alone, with the per-table options the batch derives: section parameter
 * o->section_size, or min(1000, n/10) when it is <= 0 (AliasTableSplitGpu.cu:356);
 * splits o->num_splits, one per -num_splits items when it is negative.  One
 * table per thread. */
static void batch_table_opts(const oracle_opts* o, int64_t n, oracle_opts* t) {
    *t = *o;
    t->threads = 1;
    if (t->section_size <= 0) t->section_size = std::max<int64_t>(1, std::min<int64_t>(1000, n / 10));
    if (t->num_splits < 0) t->num_splits = std::max<int64_t>(2, n / -o->num_splits);
}

static void batch_weights(const void* w, int32_t w_f32, int64_t off, int64_t n, std::vector<double>& out) {
    out.resize((size_t)n);
    if (w_f32) {
        const float* f = (const float*)w + off;
        for (int64_t i = 0; i < n; ++i) out[(size_t)i] = (double)f[i];
    } else {
        std::memcpy(out.data(), (const double*)w + off, sizeof(double) * (size_t)n);
    }
}

extern "C" int oracle_build_batch(const void* w, int32_t w_f32, const int64_t* offsets, int64_t ntables,
                                  const int64_t* which, int64_t nwhich, const oracle_opts* o, double* row_w,
                                  int32_t* row_alias, double* W_out, int32_t threads) {
    const int64_t cnt = which ? nwhich : ntables;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(| : bad)
    for (int64_t q = 0; q < cnt; ++q) {
        const int64_t t = which ? which[q] : q;
        const int64_t off = offsets[t], n = offsets[t + 1] - off;
        std::vector<double> x;
        batch_weights(w, w_f32, off, n, x);
        oracle_opts to;
        batch_table_opts(o, n, &to);
        oracle_stats st;
        if (oracle_build(x.data(), n, &to, row_w + off, row_alias + off, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr, nullptr, &st) != 0) {
            bad |= 1;
            continue;
        }
        if (W_out) W_out[t] = st.W;
    }
    return bad;
}

/* Per table: the canonical W (section parameter as oracle_build_batch derives
 * it) and the reference's table check of AliasTable.cu:24-65 in double
 * precision, as the reference computes it: implied mass min(row, W/N) plus
 * W/N - row for every row aliased to the item, against the weight, within
 * max(n 1e-8, 1e-3), and row weights within [0, W/N (1 + that bound)].
 * W_table is the W the table was built with. */
extern "C" void oracle_check_batch(const void* w, int32_t w_f32, const int64_t* offsets, int64_t ntables,
                                   const oracle_opts* o, const double* row_w, const int32_t* row_alias,
                                   const double* W_table, double* W_canonical, int32_t* ok, double* max_abs,
                                   int32_t threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t t = 0; t < ntables; ++t) {
        const int64_t off = offsets[t], n = offsets[t + 1] - off;
        std::vector<double> x;
        batch_weights(w, w_f32, off, n, x);
        oracle_opts to;
        batch_table_opts(o, n, &to);
        int64_t G, Ls;
        oracle_section_layout(n, to.section_size, &G, &Ls);
        std::vector<double> sums((size_t)G);
        for (int64_t s = 0; s < G; ++s) {
            const int64_t a = std::min(n, s * Ls), b = std::min(n, (s + 1) * Ls);
            sums[(size_t)s] = oracle_section_sum(x.data() + a, b - a);
        }
        W_canonical[t] = oracle_tree64_reduce(sums.data(), G);
        const double W_N = W_table[t] / (double)n;
        std::vector<double> m((size_t)n, 0.0);
        const double* rw = row_w + off;
        const int32_t* ra = row_alias + off;
        int good = 1;
        const double maxError = std::max((double)n * 1e-8, 1e-3);
        for (int64_t i = 0; i < n; ++i) {
            if (rw[i] < W_N) {
                m[(size_t)i] += rw[i];
                /* an alias that is not an item keeps the mass on the row, as
                 * the samplers do (a row a rounding step below W/N may carry
                 * the heavy sentinel) */
                if (ra[i] >= 0 && ra[i] < n) m[(size_t)ra[i]] += W_N - rw[i];
                else m[(size_t)i] += W_N - rw[i];
            } else {
                m[(size_t)i] += W_N;
            }
            if (!(rw[i] >= 0.0) || rw[i] > W_N * (1.0 + maxError)) good = 0;
        }
        double mabs = 0.0;
        for (int64_t i = 0; i < n; ++i) mabs = std::max(mabs, std::fabs(m[(size_t)i] - x[(size_t)i]));
        if (!(mabs <= maxError)) good = 0;
        ok[t] = good;
        max_abs[t] = mabs;
    }
}

/* Sectioned sampling restricted to some sections: the samples that
 * oracle_sample_sectioned writes for sections sections[0..nsel) (counts from
 * the whole halving tree), concatenated in the given order; first[q] is the
 * output position of section sections[q]'s first sample in the full run and
 * cnt[q] its count. */
extern "C" void oracle_sample_sectioned_subset(const double* row_w, const int32_t* row_alias, int64_t n,
                                              double W_N, int64_t SS, uint64_t total, uint64_t seed,
                                              const int64_t* sections, int64_t nsel, int32_t* out,
                                              uint64_t* first, uint64_t* cnt) {
    const int64_t nsec = (n + SS - 1) / SS;
    std::vector<uint64_t> counts((size_t)nsec);
    oracle_section_counts(n, SS, total, seed, counts.data());
    std::vector<uint64_t> off((size_t)nsec + 1, 0);
    for (int64_t s = 0; s < nsec; ++s) off[(size_t)s + 1] = off[(size_t)s] + counts[(size_t)s];
    uint64_t pos = 0;
    for (int64_t q = 0; q < nsel; ++q) {
        const int64_t s = sections[q];
        first[q] = off[(size_t)s];
        cnt[q] = counts[(size_t)s];
        const int64_t start = s * SS, len = std::min(n, (s + 1) * SS) - start;
        for (uint64_t t = 0; t < counts[(size_t)s]; ++t) {
            const uint64_t g = off[(size_t)s] + t;
            uint32_t x[4];
            philox_u64(g, ORACLE_PURPOSE_SECTIONED, 0, seed, x);
            const int64_t row = start + (int64_t)mulhi64(word64(x), (uint64_t)len);
            out[pos++] = decide(row_w[row], row_alias[row], row, coin53(x), W_N, n);
        }
    }
}

/* Threads of the oracle's OpenMP loops (the CPU baselines time 1 thread and
 * all host cores). */
extern "C" void oracle_set_threads(int32_t threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

/* ------------------------------------------------------------------------ */
/* Sectioned rejection sampling (no counterpart in the reference; it combines
 * SamplerRejection.cu:17-35 with the sections of SamplerGpuSectioned.cu:58-136).
 * A is cut into sections of SS slots.  A trial on slot s accepts iff
 * coin53 <= accept(s), accept(s) = frac(N w/W) on the first slot of an item's
 * run and 1 elsewhere, so its acceptance probability is exactly
 * P(s) = min(floor(accept(s) 2^53) + 1, 2^53) 2^-53.  An accepted sample lies
 * in section k with probability M_k / sum M, M_k = sum of P(s) over the
 * section, so the sample count of every section comes from the halving tree
 * of the sectioned alias sampler with q = (Pre[mid] - Pre[lo]) /
 * (Pre[hi] - Pre[lo]) on the canonical exclusive prefix Pre of the M_k; inside
 * a section a sample repeats trials on uniform slots of the section.  M_k is
 * the canonical section sum (lane-strided, pairwise tree). */
static inline double rej_accept_prob(double accept) {
    if (accept >= 1.0) return 1.0;
    return (std::floor(accept * 0x1.0p53) + 1.0) * 0x1.0p-53;
}

static inline double rej_slot_accept(const int32_t* A, uint64_t s, const double* w, int64_t n, double W) {
    const int32_t item = A[s];
    if (s != 0 && A[s - 1] == item) return 1.0;
    const double xv = ((double)n * w[item]) / W;
    return xv - std::floor(xv);
}

/* Exclusive prefix Pre[0..nsec] (Pre[nsec] = total) of the section masses. */
extern "C" void oracle_rejection_section_prefix(const int32_t* A, uint64_t A_size, const double* w, int64_t n,
                                               double W, int64_t SS, double* pre) {
    const int64_t nsec = (int64_t)((A_size + (uint64_t)SS - 1) / (uint64_t)SS);
    std::vector<double> mass((size_t)nsec);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nsec; ++k) {
        const uint64_t a = (uint64_t)k * SS, b = std::min<uint64_t>(A_size, a + SS);
        double acc[64];
        for (int l = 0; l < 64; ++l) {
            double x = 0.0;
            for (uint64_t s = a + l; s < b; s += 64) x = x + rej_accept_prob(rej_slot_accept(A, s, w, n, W));
            acc[l] = x;
        }
        mass[(size_t)k] = tree64(acc, 64);
    }
    pre[nsec] = oracle_local_scan(mass.data(), nsec, pre);
}

static void rej_counts_rec(int64_t lo, int64_t hi, uint64_t cnt, uint64_t node, const double* pre, uint64_t seed,
                           uint64_t* counts) {
    if (hi - lo == 1) { counts[lo] = cnt; return; }
    const int64_t mid = lo + (hi - lo) / 2;
    /* a range whose mass vanishes in the prefix's rounding splits by section count */
    const double den = pre[hi] - pre[lo];
    const double q = den > 0.0 ? (pre[mid] - pre[lo]) / den : (double)(mid - lo) / (double)(hi - lo);
    const uint64_t x = binomial_purpose(cnt, q, seed, node, ORACLE_PURPOSE_REJ_SECTION_COUNTS);
    rej_counts_rec(lo, mid, x, 2 * node, pre, seed, counts);
    rej_counts_rec(mid, hi, cnt - x, 2 * node + 1, pre, seed, counts);
}

extern "C" void oracle_rejection_section_counts(const double* pre, int64_t nsec, uint64_t total, uint64_t seed,
                                               uint64_t* counts) {
    rej_counts_rec(0, nsec, total, 1, pre, seed, counts);
}

/* Samples of the sectioned rejection sampler: section k draws counts[k]
 * samples written contiguously in section order; sample g makes trials
 * t = 0, 1, ... with Philox counter (g, purpose, t), a uniform slot of the
 * section from words 0-1 and coin53 from words 2-3, until one accepts (at
 * most ORACLE_REJ_SECTIONED_MAX_TRIALS; then the last trial's item).
 * Only the sections listed (all when sections == NULL) are drawn; their
 * samples land at their output positions of the full run. */
extern "C" void oracle_sample_rejection_sectioned(const int32_t* A, uint64_t A_size, const double* w, int64_t n,
                                                 double W, int64_t SS, uint64_t total, uint64_t seed,
                                                 const int64_t* sections, int64_t nsel, int32_t* out,
                                                 uint64_t* first_out, uint64_t* cnt_out) {
    const int64_t nsec = (int64_t)((A_size + (uint64_t)SS - 1) / (uint64_t)SS);
    std::vector<double> pre((size_t)nsec + 1);
    oracle_rejection_section_prefix(A, A_size, w, n, W, SS, pre.data());
    std::vector<uint64_t> counts((size_t)nsec);
    oracle_rejection_section_counts(pre.data(), nsec, total, seed, counts.data());
    std::vector<uint64_t> off((size_t)nsec + 1, 0);
    for (int64_t k = 0; k < nsec; ++k) off[(size_t)k + 1] = off[(size_t)k] + counts[(size_t)k];
    const int64_t cnt = sections ? nsel : nsec;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < cnt; ++q) {
        const int64_t k = sections ? sections[q] : q;
        if (first_out) first_out[q] = off[(size_t)k];
        if (cnt_out) cnt_out[q] = counts[(size_t)k];
        const uint64_t a = (uint64_t)k * SS, len = std::min<uint64_t>(A_size, a + SS) - a;
        for (uint64_t t = 0; t < counts[(size_t)k]; ++t) {
            const uint64_t g = off[(size_t)k] + t;
            int32_t item = 0;
            for (uint32_t tr = 0; tr < ORACLE_REJ_SECTIONED_MAX_TRIALS; ++tr) {
                uint32_t x[4];
                philox_u64(g, ORACLE_PURPOSE_REJ_SECTIONED, tr, seed, x);
                const uint64_t s = a + mulhi64(word64(x), len);
                item = A[s];
                if (!(coin53(x) > rej_slot_accept(A, s, w, n, W))) break;
            }
            out[g] = item;
        }
    }
}

/* SplitMethodInverse (SplitMethodInverse.cu:3-31), a test-only second
 * brute force of the split: every heavy position j visits every split k
 * with j < n'(k) and records (i, j, spill) wherever the FOUND predicate
 * holds (sequentially, so the largest such j of a k is kept; the reference
 * runs the j in parallel).  Inputs as oracle_split_one; splits[0] and
 * splits[p] are the fixed ends; a k that no j satisfies keeps i = j = -1. */
extern "C" void oracle_split_inverse(int64_t p, int64_t nToSplit, double wToSplit, const double* Lp, int64_t numL,
                                    const double* Hp, const double* Hw, int64_t numH, int32_t* si, int32_t* sj,
                                    double* sspill) {
    for (int64_t k = 0; k <= p; ++k) { si[k] = -1; sj[k] = -1; sspill[k] = 0.0; }
    si[0] = 0; sj[0] = 0;
    si[p] = (int32_t)(numL - 1); sj[p] = (int32_t)(numH - 1);
    for (int64_t j = 0; j < numH - 1; ++j) {
        for (int64_t k = 1; k < p; ++k) {
            const int64_t n_ = (int64_t)std::ceil(((double)nToSplit * (double)k) / (double)p);
            if (j >= n_) continue;
            const int64_t i = std::min(n_ - j, numL - 1);
            const double target = (wToSplit * (double)n_) / (double)nToSplit;
            if (i < 1) continue;
            const double sigma = Lp[i] + Hp[j];
            const double sigmaNext = Lp[i - 1] + Hp[j + 1];
            if (sigma <= target && sigmaNext > target) {
                si[k] = (int32_t)i; sj[k] = (int32_t)j; sspill[k] = sigma + Hw[j] - target;
            }
        }
    }
}

/* fillABinarySearch (RejectionSamplingBinarySearch.cu:12-33): slot s holds the
 * item found by bisection over the inclusive counts, the first i with
 * incl[i] > s.  A second restatement of the fill beside the run-writing one
 * of oracle_rejection_build. */
extern "C" void oracle_rejection_fill_bisect(const uint64_t* incl, int64_t n, uint64_t A_size, int32_t* A) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)A_size; ++s) {
        int64_t left = 0, right = n;
        while (right - left > 1) {
            const int64_t c = (left + right) / 2;
            if (incl[c] > (uint64_t)s) right = c;
            else left = c;
        }
        A[s] = (int32_t)(incl[left] <= (uint64_t)s ? right : left);
    }
}

/* The learned inverse-CDF model of A over the inclusive counts (no reference
 * code; it predicts the item of a slot, the quantity fillABinarySearch
 * searches for).  With M = |A| - 1 and K segments:
 *   segment of slot s      seg(s) = 0 if M = 0, else min(K - 1, floor(s K / M))
 *   knot slot k (0..K)     s_k = ceil(k M / K)   (so seg(s) = k iff s_k <= s < s_{k+1})
 *   knot item              q_k = item of slot s_k
 *   prediction in seg k    pred(s) = q_k + floor((s - s_k)(q_{k+1} - q_k) / (s_{k+1} - s_k))
 *                          (q_k when s_{k+1} = s_k)
 *   error bound            e_k = max over the slots s of segment k of |item(s) - pred(s)|
 * all in exact integer arithmetic.  Restated here by brute force over every
 * slot; the library computes e_k from the two ends of each item's run. */
extern "C" void oracle_rejection_model(const uint64_t* incl, int64_t n, uint64_t A_size, int32_t K,
                                       int32_t* q, uint32_t* err) {
    const uint64_t M = A_size - 1;
    auto kslot = [&](uint64_t k) { return (k * M + (uint64_t)K - 1) / (uint64_t)K; };
    for (int32_t k = 0; k <= K; ++k) {
        const uint64_t s = kslot((uint64_t)k);
        q[k] = (int32_t)(std::upper_bound(incl, incl + n, s) - incl);
    }
    std::fill(err, err + K, 0u);
    int64_t item = 0;
    for (uint64_t s = 0; s < A_size; ++s) {
        while (incl[item] <= s) ++item;
        const uint64_t k = M == 0 ? 0 : std::min<uint64_t>((uint64_t)K - 1, s * (uint64_t)K / M);
        const uint64_t sk = kslot(k), sk1 = kslot(k + 1);
        const int64_t pred = sk1 > sk ? q[k] + (int64_t)((s - sk) * (uint64_t)(q[k + 1] - q[k]) / (sk1 - sk)) : q[k];
        const uint32_t e = (uint32_t)(item > pred ? item - pred : pred - item);
        err[k] = std::max(err[k], e);
    }
}
