// Alias-table construction kernels for gfx950 (the hot path of
// AliasTableSplitGpu::build, AliasTableSplitGpu.cu:388-468).
//
// Phases (one stream, no host round trips):
//   1. sum_sections : per-section FP64 sums (64 sections per block), weight
//                     validation, the first level of the canonical 64-ary tree;
//                     the last block to finish completes the tree -> W, W/N.
//   2. psa_tiles    : a 4-wave workgroup per tile of T sections, tiles taken
//                     in ticket order; per section: all loads issued at once,
//                     stable light/heavy partition into LDS, PSA+ greedy
//                     chains (spread over the 4 waves) writing finished rows,
//                     leftover maps, the canonical local scan of the leftover
//                     weights, a decoupled look-back over tiles that folds the
//                     predecessors' section aggregates left to right (bit
//                     identical to the chained prefix P_{s+1} = P_s + agg_s),
//                     leftover lists + prefix sums, the split certificate and
//                     the sentinels.
//   3. degenerate_fill (only when one leftover class is empty), split, pack.
#include "atg_build.h"
#include "atg_device.h"

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
    // Reset the look-back slots of the partition pass.
    const int64_t nslots = 3 * a.G + 3 * a.Gt;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (int64_t)gridDim.x * blockDim.x)
        a.slots[i] = ATG_NOT_READY;
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
// 2. PSA+ / plain partition.  A workgroup of 4 waves owns a tile of T
//    consecutive sections, processes them one after another with all its
//    lanes, then takes one look-back for the tile and writes the leftovers.
#define PSA_THREADS 256
#define PSA_MAX_LEN 4096
#define PSA_MAX_TILE 8

struct SecLDS {
    double* w;      // [Ls]: lights from the front, heavies from the back
    double* carL;   // [Ls/64 + 2]: chunk carries of the light leftover list
    double* carH;   // [Ls/64 + 2]
    uint16_t* idx;  // [Ls]: local item index
    uint16_t* map;  // [Ls]: leftover rank -> position (L ranks, then H ranks)
    int len;
    __device__ double& Lw(int i) { return w[i]; }
    __device__ uint16_t& Li(int i) { return idx[i]; }
    __device__ double& Hw(int j) { return w[len - 1 - j]; }
    __device__ uint16_t& Hi(int j) { return idx[len - 1 - j]; }
};

__host__ __device__ inline int64_t psa_carry_slots(int64_t Ls_max) { return Ls_max / 64 + 2; }

__device__ __forceinline__ SecLDS section_lds(const BuildArgs& a, double* lds_raw, int q) {
    char* base = reinterpret_cast<char*>(lds_raw) + (size_t)q * a.lds_region;
    const int64_t nc = psa_carry_slots(a.Ls_max);
    SecLDS L;
    L.w = reinterpret_cast<double*>(base);
    L.carL = L.w + a.Ls_max;
    L.carH = L.carL + nc;
    L.idx = reinterpret_cast<uint16_t*>(L.carH + nc);
    L.map = L.idx + a.Ls_max;
    L.len = 0;
    return L;
}

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

// Look-back over tiles: the exclusive prefix of the tile's first section.
// Phase 1 finds the nearest predecessor tile with a published inclusive value
// (moving back window by window; tile -1 is a virtual inclusive 0).  Phase 2
// folds the section aggregates of the tiles after it, left to right.
__device__ Prefix look_back_tile(const BuildArgs& a, int64_t tile, int T, int lane, bool& timeout) {
    const uint64_t* aggC = a.slots;
    const uint64_t* aggL = a.slots + a.G;
    const uint64_t* aggH = a.slots + 2 * a.G;
    const uint64_t* tC = a.slots + 3 * a.G;
    const uint64_t* tL = tC + a.Gt;
    const uint64_t* tH = tL + a.Gt;
    int64_t hi = tile - 1;
    Prefix acc{0, 0.0, 0.0};
    int64_t qstar;
    while (true) {
        const int64_t q = hi - 63 + lane;
        uint64_t vc = 0, vl = 0, vh = 0;
        bool ready;
        if (q < 0) {
            ready = true;
        } else {
            vc = slot_load(tC + q); vl = slot_load(tL + q); vh = slot_load(tH + q);
            ready = vc != ATG_NOT_READY && vl != ATG_NOT_READY && vh != ATG_NOT_READY;
        }
        const uint64_t mask = __ballot(ready);
        if (mask != 0ull) {
            const int h = 63 - __clzll(mask);
            qstar = hi - 63 + h;
            if (qstar >= 0) {
                acc.c = readlane_u64(vc, h);
                acc.l = u2d(readlane_u64(vl, h));
                acc.h = u2d(readlane_u64(vh, h));
            }
            break;
        }
        hi -= 64;
    }
    const int64_t first = (qstar + 1) * T, last = tile * T;  // sections to fold: [first, last)
    for (int64_t base = first; base < last; base += 64) {
        const int64_t sec = base + lane;
        const int cnt = (int)min<int64_t>(64, last - base);
        uint64_t c = 0, l = 0, h = 0;
        if (lane < cnt) {
            uint32_t spins = 0;
            while (true) {
                c = slot_load(aggC + sec); l = slot_load(aggL + sec); h = slot_load(aggH + sec);
                if (c != ATG_NOT_READY && l != ATG_NOT_READY && h != ATG_NOT_READY) break;
                if (++spins > (1u << 22)) { timeout = true; c = 0; l = d2u(0.0); h = d2u(0.0); break; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        for (int t = 0; t < cnt; ++t) {
            acc.c += readlane_u64(c, t);
            acc.l = acc.l + u2d(readlane_u64(l, t));
            acc.h = acc.h + u2d(readlane_u64(h, t));
        }
    }
    return acc;
}

// Shared scratch of one workgroup.
struct PsaShared {
    uint32_t tile;
    int nL, nH, mL, mH;
    int a[64], b[64];                  // per (chunk of 256, wave): counts -> bases
    int cl[64], ch[64], il[64], jl[64]; // per chain: leftover counts and first positions
    int ol[64], oh[64];                // per chain: leftover offsets
    int len[PSA_MAX_TILE], secL[PSA_MAX_TILE], secH[PSA_MAX_TILE];
    double aggL[PSA_MAX_TILE], aggH[PSA_MAX_TILE];
    uint64_t pc[PSA_MAX_TILE];
    double pl[PSA_MAX_TILE], ph[PSA_MAX_TILE];
    double red[8];
};

// Partition, chains, leftover maps and chunk scans of one section.
template <typename T, bool AOS>
__device__ void psa_section(const BuildArgs& a, PsaShared& sh, SecLDS& L, int64_t start, int len, double W_N,
                            int tid, int lane, int wave) {
    const T* w = (const T*)a.weights;
    const int U = (len + 255) >> 8;
    const uint64_t lt = lanemask_lt();
    double x[PSA_MAX_LEN / 256];
#pragma unroll
    for (int u = 0; u < PSA_MAX_LEN / 256; ++u) {
        const int k = (u << 8) + tid;
        x[u] = (u < U && k < len) ? (double)w[start + k] : 0.0;
    }
    // Stable partition: chunk u of 256 items, wave w covers its 64-item slice.
#pragma unroll
    for (int u = 0; u < PSA_MAX_LEN / 256; ++u) {
        if (u < U) {
            const int k = (u << 8) + tid;
            const bool valid = k < len;
            const uint64_t bl = __ballot(valid && x[u] <= W_N);
            const uint64_t bh = __ballot(valid && !(x[u] <= W_N));
            if (lane == 0) { sh.a[u * 4 + wave] = __popcll(bl); sh.b[u * 4 + wave] = __popcll(bh); }
        }
    }
    __syncthreads();
    if (wave == 0) {
        const int va = (lane < 4 * U) ? sh.a[lane] : 0, vb = (lane < 4 * U) ? sh.b[lane] : 0;
        const int ea = wave_excl_scan_int(va), eb = wave_excl_scan_int(vb);
        const int ta = wave_sum_int(va), tb = wave_sum_int(vb);
        sh.a[lane] = ea;
        sh.b[lane] = eb;
        if (lane == 0) { sh.nL = ta; sh.nH = tb; }
    }
    __syncthreads();
    L.len = len;
#pragma unroll
    for (int u = 0; u < PSA_MAX_LEN / 256; ++u) {
        if (u < U) {
            const int k = (u << 8) + tid;
            const bool valid = k < len;
            const bool light = valid && x[u] <= W_N, heavy = valid && !(x[u] <= W_N);
            const uint64_t bl = __ballot(light), bh = __ballot(heavy);
            if (light) {
                const int pos = sh.a[u * 4 + wave] + __popcll(bl & lt);
                L.Lw(pos) = x[u];
                L.Li(pos) = (uint16_t)k;
            }
            if (heavy) {
                const int pos = sh.b[u * 4 + wave] + __popcll(bh & lt);
                L.Hw(pos) = x[u];
                L.Hi(pos) = (uint16_t)k;
            }
        }
    }
    __syncthreads();
    const int nL = sh.nL, nH = sh.nH;
    const bool identity = !a.psa_plus;
    int nch = 1;
    if (a.psa_plus) {
        // Greedy chains (PsaPlusNormal.cu:51-123); chain t runs on wave t % 4.
        const int C = a.chains;
        nch = C;
        const int t = lane * 4 + wave;
        if (t < C) {
            const int stepsL = nL / C, stepsH = nH / C;
            int iLo = t * stepsL, iHi = (t == C - 1) ? nL : (t + 1) * stepsL;
            int jLo = t * stepsH, jHi = (t == C - 1) ? nH : (t + 1) * stepsH;
            if (iHi - iLo > 0 && jHi - jLo > 0) {
                int i = iLo, j = jLo;
                double wc = L.Hw(j);
                int hj = L.Hi(j);
                while (true) {
                    const bool ex = wc <= W_N + ATG_EPSILON;
                    if ((ex && j >= jHi - 1) || (!ex && i >= iHi - 1)) break;
                    int item, alias;
                    double rw, next;
                    if (ex) {
                        const int hn = L.Hi(j + 1);
                        item = hj; rw = wc; alias = hn; next = L.Hw(j + 1); ++j; hj = hn;
                    } else {
                        item = L.Li(i); rw = L.Lw(i); alias = hj; next = rw; ++i;
                    }
                    wc += next - W_N;
                    store_row<AOS>(a, start + item, rw, (int32_t)(start + alias));
                }
                // Remainder (PsaPlusNormal.cu:112-123): back to L when its
                // residual is light and a consumed light slot can hold it,
                // otherwise back to H in place.
                if (wc <= W_N + ATG_EPSILON && i > iLo) {
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
            sh.cl[t] = iHi - iLo; sh.ch[t] = jHi - jLo; sh.il[t] = iLo; sh.jl[t] = jLo;
        }
    } else if (tid == 0) {
        sh.cl[0] = nL; sh.ch[0] = nH; sh.il[0] = 0; sh.jl[0] = 0;
    }
    __syncthreads();
    if (wave == 0) {
        const int vl = (lane < nch) ? sh.cl[lane] : 0, vh = (lane < nch) ? sh.ch[lane] : 0;
        const int ol = wave_excl_scan_int(vl), oh = wave_excl_scan_int(vh);
        const int ml = wave_sum_int(vl), mh = wave_sum_int(vh);
        sh.ol[lane] = ol;
        sh.oh[lane] = oh;
        if (lane == 0) { sh.mL = ml; sh.mH = mh; }
    }
    __syncthreads();
    const int mL = sh.mL, mH = sh.mH;
    if (!identity) {
        const int t = lane * 4 + wave;
        if (t < nch) {
            const int cl = sh.cl[t], ch = sh.ch[t], il = sh.il[t], jl = sh.jl[t], ol = sh.ol[t], oh = sh.oh[t];
            for (int q = 0; q < cl; ++q) L.map[ol + q] = (uint16_t)(il + q);
            for (int q = 0; q < ch; ++q) L.map[mL + oh + q] = (uint16_t)(jl + q);
        }
    }
    __syncthreads();
    // Chunk scans of the two leftover lists: chunk totals in parallel.
    const int nchL = (mL + 63) >> 6, nchH = (mH + 63) >> 6;
    for (int c = wave; c < nchL + nchH; c += 4) {
        const bool heavy = c >= nchL;
        const int cc = heavy ? c - nchL : c;
        const int m = heavy ? mH : mL;
        const int r = (cc << 6) + lane;
        double v = 0.0;
        if (r < m) {
            const int pos = identity ? r : (int)L.map[(heavy ? mL : 0) + r];
            v = heavy ? L.Hw(pos) : L.Lw(pos);
        }
        const double inc = wave_scan(v);
        if (lane == 63) (heavy ? L.carH : L.carL)[cc] = inc;
    }
    __syncthreads();
    // Carries chained left to right (oracle_local_scan).
    if (tid == 0 || tid == 64) {
        double* car = (tid == 0) ? L.carL : L.carH;
        const int nc = (tid == 0) ? nchL : nchH;
        double carry = 0.0;
        for (int c = 0; c < nc; ++c) {
            const double t = car[c];
            car[c] = carry;
            carry = carry + t;
        }
        car[nc] = carry;
    }
    __syncthreads();
}

// Writes one section's leftover lists with global prefixes and folds the
// split certificate (max light increment, min heavy increment).
__device__ void psa_write(const BuildArgs& a, PsaShared& sh, SecLDS& L, int64_t start, int mL, int mH,
                          uint64_t Pc, double Pl, double Ph, int lane, int wave, double& ext_l, double& ext_h) {
    const bool identity = !a.psa_plus;
    const int nchL = (mL + 63) >> 6, nchH = (mH + 63) >> 6;
    const int64_t goffL = (int64_t)(uint32_t)Pc, goffH = (int64_t)(Pc >> 32);
    for (int c = wave; c < nchL + nchH; c += 4) {
        const bool heavy = c >= nchL;
        const int cc = heavy ? c - nchL : c;
        const int m = heavy ? mH : mL;
        const double* car = heavy ? L.carH : L.carL;
        const double P = heavy ? Ph : Pl;
        const int r = (cc << 6) + lane;
        double v = 0.0;
        int item = 0;
        if (r < m) {
            const int pos = identity ? r : (int)L.map[(heavy ? mL : 0) + r];
            v = heavy ? L.Hw(pos) : L.Lw(pos);
            item = heavy ? L.Hi(pos) : L.Li(pos);
        }
        const double inc = wave_scan(v);
        const double local = car[cc] + wave_shr1(inc);
        const double pref = P + local;
        double succ = __shfl_down(pref, 1, 64);
        if (lane == 63) succ = P + car[cc + 1];
        if (r == m - 1) succ = P + car[(m + 63) >> 6];
        if (r < m) {
            const int64_t g = (heavy ? goffH : goffL) + r;
            (heavy ? a.Hi : a.Li)[g] = (int32_t)(start + item);
            (heavy ? a.Hw : a.Lw)[g] = v;
            (heavy ? a.Hp : a.Lp)[g] = pref;
            const double d = succ - pref;
            if (heavy) ext_h = fmin(ext_h, d); else ext_l = fmax(ext_l, d);
        }
    }
}

template <typename T, bool AOS>
__global__ __launch_bounds__(PSA_THREADS) void psa_tiles(BuildArgs a) {
    extern __shared__ double lds_raw[];
    __shared__ PsaShared sh;
    DevState* st = a.st;
    if (st->error) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int TW = a.tile_sections;
    if (tid == 0) sh.tile = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int64_t tile = sh.tile;
    const double W_N = st->W_N;
    for (int q = 0; q < TW; ++q) {
        const int64_t s = tile * TW + q;
        if (s >= a.G) {
            if (tid == 0) sh.len[q] = -1;
            continue;
        }
        const int64_t start = min(a.n, s * a.Ls);
        const int len = (int)(min(a.n, (s + 1) * a.Ls) - start);
        SecLDS L = section_lds(a, lds_raw, q);
        psa_section<T, AOS>(a, sh, L, start, len, W_N, tid, lane, wave);
        if (tid == 0) {
            const int mL = sh.mL, mH = sh.mH;
            const double aggL = L.carL[(mL + 63) >> 6], aggH = L.carH[(mH + 63) >> 6];
            const uint64_t aggC = (uint64_t)(uint32_t)mL | ((uint64_t)(uint32_t)mH << 32);
            slot_store(a.slots + s, aggC);
            slot_store(a.slots + a.G + s, d2u(aggL));
            slot_store(a.slots + 2 * a.G + s, d2u(aggH));
            sh.len[q] = len;
            sh.secL[q] = mL;
            sh.secH[q] = mH;
            sh.aggL[q] = aggL;
            sh.aggH[q] = aggH;
        }
        __syncthreads();
    }
    if (wave == 0) {
        bool timeout = false;
        Prefix P{0, 0.0, 0.0};
        if (tile > 0) P = look_back_tile(a, tile, TW, lane, timeout);
        if (lane == 0) {
            if (timeout) atomicOr(&st->error, ERR_LOOKBACK_TIMEOUT);
            for (int q = 0; q < TW; ++q) {
                sh.pc[q] = P.c; sh.pl[q] = P.l; sh.ph[q] = P.h;
                if (sh.len[q] >= 0) {
                    P.c += (uint64_t)(uint32_t)sh.secL[q] | ((uint64_t)(uint32_t)sh.secH[q] << 32);
                    P.l = P.l + sh.aggL[q];
                    P.h = P.h + sh.aggH[q];
                }
            }
            uint64_t* tC = a.slots + 3 * a.G;
            slot_store(tC + tile, P.c);
            slot_store(tC + a.Gt + tile, d2u(P.l));
            slot_store(tC + 2 * a.Gt + tile, d2u(P.h));
        }
    }
    __syncthreads();
    double ext_l = 0.0, ext_h = __builtin_inf();
    bool anyL = false, anyH = false;
    for (int q = 0; q < TW; ++q) {
        if (sh.len[q] < 0) continue;
        const int64_t s = tile * TW + q;
        const int64_t start = min(a.n, s * a.Ls);
        SecLDS L = section_lds(a, lds_raw, q);
        L.len = sh.len[q];
        const int mL = sh.secL[q], mH = sh.secH[q];
        anyL |= mL > 0;
        anyH |= mH > 0;
        psa_write(a, sh, L, start, mL, mH, sh.pc[q], sh.pl[q], sh.ph[q], lane, wave, ext_l, ext_h);
        if (s == a.G - 1 && tid == 0) {
            // Sentinels (AliasTableSplitGpu.cu:366-385) and the split inputs.
            const uint64_t inclC = sh.pc[q] + ((uint64_t)(uint32_t)mL | ((uint64_t)(uint32_t)mH << 32));
            const double inclL = sh.pl[q] + sh.aggL[q], inclH = sh.ph[q] + sh.aggH[q];
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
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        ext_l = fmax(ext_l, shfl_xor_d(ext_l, m));
        ext_h = fmin(ext_h, shfl_xor_d(ext_h, m));
    }
    if (lane == 0) {
        if (anyL) atomicMax((unsigned long long*)&st->cert_max_dL, (unsigned long long)d2u(ext_l));
        if (anyH) atomicMin((unsigned long long*)&st->cert_min_dH, (unsigned long long)d2u(ext_h));
    }
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

size_t psa_region_bytes(int64_t Ls_max) {
    return (((size_t)Ls_max * 12 + 16 * (size_t)psa_carry_slots(Ls_max)) + 15) & ~(size_t)15;
}
size_t psa_lds_bytes(int64_t Ls_max, int tile_sections) { return psa_region_bytes(Ls_max) * tile_sections; }

hipError_t launch_partition(const BuildArgs& a, hipStream_t s) {
    const size_t lds = psa_lds_bytes(a.Ls_max, a.tile_sections);
    const dim3 g((unsigned)a.Gt), b(PSA_THREADS);
    if (a.dtype_f32) {
        if (a.rows) hipLaunchKernelGGL((psa_tiles<float, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((psa_tiles<float, false>), g, b, lds, s, a);
    } else {
        if (a.rows) hipLaunchKernelGGL((psa_tiles<double, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((psa_tiles<double, false>), g, b, lds, s, a);
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
    const void* fns[] = {(const void*)psa_tiles<double, true>, (const void*)psa_tiles<double, false>,
                         (const void*)psa_tiles<float, true>, (const void*)psa_tiles<float, false>};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace atg
