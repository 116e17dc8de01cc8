// Sampling kernels for gfx950: basic and sectioned alias sampling
// (sampling/aliasTable/SamplerGpuBasic.cu, SamplerGpuSectioned.cu,
// SamplerGpuSectionedShared.cu) and rejection sampling
// (construction/rejectionSampling/*, sampling/rejectionSampling/SamplerRejection.cu).
//
// Randomness: Philox4x32-10, key = seed, counter = (output position, purpose,
// extra).  Every sample is a pure function of (table, seed, position), so any
// launch shape yields the same output.
#include "atg_device.h"
#include "atg_sample.h"

namespace atg {

template <bool AOS>
__device__ __forceinline__ void load_row(const SampleArgs& a, int64_t row, double& w, int32_t& alias) {
    if (AOS) {
        const uint4 v = *reinterpret_cast<const uint4*>(&a.rows[row]);
        w = __longlong_as_double(((long long)v.y << 32) | v.x);
        alias = (int32_t)v.z;
    } else {
        w = a.tw[row];
        alias = a.ta[row];
    }
}

__device__ __forceinline__ void emit(const SampleArgs& a, uint64_t g, int32_t item, uint64_t& csum) {
    if (a.out) a.out[g] = item;
    if (a.hist) atomicAdd(&a.hist[item], 1u);
    csum += (uint64_t)(uint32_t)item;
}

__device__ __forceinline__ void flush_checksum(const SampleArgs& a, uint64_t csum) {
    if (!a.checksum) return;
    csum = wave_sum_int<uint64_t>(csum);
    if ((threadIdx.x & 63) == 0 && csum) atomicAdd(a.checksum, (unsigned long long)csum);
}

// ---------------------------------------------------------------------------
// Basic sampling: one uniform row per sample over the whole table.
#define BASIC_UNROLL 4
template <bool AOS>
__global__ __launch_bounds__(256) void sample_basic(SampleArgs a) {
    const double W_N = a.st->W_N;
    uint64_t csum = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // BASIC_UNROLL samples per lane and iteration: their row loads are issued
    // together, so more random reads are in flight (one at a time: 32 ms for
    // 1e9 samples from 1e8 rows; four: 22 ms; eight: no further gain).
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g0 < a.count; g0 += BASIC_UNROLL * stride) {
        u32x4 r[BASIC_UNROLL];
        int64_t row[BASIC_UNROLL];
        double w[BASIC_UNROLL];
        int32_t al[BASIC_UNROLL];
#pragma unroll
        for (int q = 0; q < BASIC_UNROLL; ++q) {
            const uint64_t g = g0 + (uint64_t)q * stride;
            r[q] = philox_at(g, PURPOSE_BASIC, 0, a.seed);
            row[q] = (int64_t)scale64(word64(r[q]), (uint64_t)a.n);
            if (g < a.count) load_row<AOS>(a, row[q], w[q], al[q]);
        }
#pragma unroll
        for (int q = 0; q < BASIC_UNROLL; ++q) {
            const uint64_t g = g0 + (uint64_t)q * stride;
            if (g < a.count) emit(a, g, decide(w[q], al[q], row[q], coin53(r[q]), W_N, a.n), csum);
        }
    }
    flush_checksum(a, csum);
}

// ---------------------------------------------------------------------------
// Sectioned sampling.  Section counts: recursive halving of the section range
// with a binomial draw per node (SamplerGpuSectioned.cu:58-96).  The reference
// uses a float normal approximation; restated with IEEE + - * / sqrt floor only.
__device__ uint64_t binomial(uint64_t n, double q, uint64_t seed, uint64_t node,
                             uint32_t purpose = PURPOSE_SECTION_COUNTS) {
    if (n == 0) return 0;
    if (n <= 32) {
        uint64_t cnt = 0;
        u32x4 x{0, 0, 0, 0};
        for (uint64_t t = 0; t < n; ++t) {
            if ((t & 3) == 0) x = philox_at(node, purpose, (uint32_t)(t >> 2), seed);
            const uint32_t word = ((t & 3) == 0) ? x.x : ((t & 3) == 1) ? x.y : ((t & 3) == 2) ? x.z : x.w;
            const double u = (double)word * 0x1.0p-32;
            if (u < q) ++cnt;
        }
        return cnt;
    }
    double z = 0.0;
    for (uint32_t d = 0; d < 3; ++d) {
        const u32x4 x = philox_at(node, purpose, d, seed);
        z = z + (double)x.x * 0x1.0p-32;
        z = z + (double)x.y * 0x1.0p-32;
        z = z + (double)x.z * 0x1.0p-32;
        z = z + (double)x.w * 0x1.0p-32;
    }
    z = z - 6.0;
    const double mean = (double)n * q;
    const double var = mean * (1.0 - q);
    const double sd = __builtin_sqrt(var);
    double r = floor((mean + sd * z) + 0.5);
    if (r < 0.0) r = 0.0;
    if (r > (double)n) r = (double)n;
    return (uint64_t)r;
}

// Count and output offset of section s: descend the halving tree from the root.
__device__ void section_count(const SampleArgs& a, int64_t s, uint64_t& cnt, uint64_t& off) {
    int64_t lo = 0, hi = a.nsec;
    uint64_t c = a.count, o = 0, node = 1;
    while (hi - lo > 1) {
        const int64_t mid = lo + (hi - lo) / 2;
        const int64_t rl = min(a.n, mid * a.SS) - lo * a.SS;
        const int64_t rt = min(a.n, hi * a.SS) - lo * a.SS;
        const double q = (double)rl / (double)rt;
        const uint64_t x = binomial(c, q, a.seed, node);
        if (s < mid) {
            hi = mid; c = x; node = 2 * node;
        } else {
            lo = mid; o += x; c = c - x; node = 2 * node + 1;
        }
    }
    cnt = c;
    off = o;
}

// A section of up to 8192 rows (64 KB at the default 4096) fills a large part
// of a CU's LDS, so few blocks are resident: 1024 threads per block keep
// enough waves per SIMD to hide the latency of the Philox chains (512 threads
// were 7% slower, 256 32%, at 1e9 samples from 1e8 rows).
#define SECTIONED_THREADS 1024

// Count and first output of every section, one thread per section: each
// descent is a serial chain of binomial draws, which would otherwise hold up
// a whole sampling block at its start.
__global__ __launch_bounds__(64) void section_counts(SampleArgs a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nsec) return;
    uint64_t c, o;
    section_count(a, s, c, o);
    a.sec_counts[2 * s] = c;
    a.sec_counts[2 * s + 1] = o;
}

template <bool AOS>
__global__ __launch_bounds__(SECTIONED_THREADS) void sample_sectioned(SampleArgs a) {
    extern __shared__ Row sect[];
    const int64_t s = blockIdx.x;
    const int64_t start = s * a.SS;
    const int64_t len = min(a.n, start + a.SS) - start;
    // Stage the section's rows in LDS (SamplerGpuSectionedShared.cu:58-63).
    for (int64_t k = threadIdx.x; k < len; k += blockDim.x) {
        if (AOS) {
            *reinterpret_cast<uint4*>(&sect[k]) = *reinterpret_cast<const uint4*>(&a.rows[start + k]);
        } else {
            sect[k].weight = a.tw[start + k];
            sect[k].alias = a.ta[start + k];
            sect[k].pad = 0;
        }
    }
    __syncthreads();
    const uint64_t cnt = a.sec_counts[2 * s], off = a.sec_counts[2 * s + 1];
    const double W_N = a.st->W_N;
    uint64_t csum = 0;
    for (uint64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
        const uint64_t g = off + t;
        const u32x4 r = philox_at(g, PURPOSE_SECTIONED, 0, a.seed);
        const int64_t row = (int64_t)scale64(word64(r), (uint64_t)len);
        const uint4 v = *reinterpret_cast<const uint4*>(&sect[row]);
        const double w = __longlong_as_double(((long long)v.y << 32) | v.x);
        emit(a, g, decide(w, (int32_t)v.z, start + row, coin53(r), W_N, a.n), csum);
    }
    flush_checksum(a, csum);
}

// ---------------------------------------------------------------------------
// Rejection structure (RejectionSampling.cu:26-56).  c_i = floor(N w_i / W) + 1
// is scanned with integers (exact in any order): per-block totals, one block
// scanning the totals, then each block writes its items' slots of A; the
// slots of long runs are cut into tasks that the whole grid writes.
#define REJ_WAVE 64     // runs up to this are written by their wave, longer ones as tasks
#define REJ_WAVE_BUDGET 8192  // ... in waves with more slots than this; other waves write all their runs
#define REJ_TASK 4096   // slots per task
#define REJ_LONG_BLOCKS 1024

// A is written once and read only by a later sampling launch: streaming
// (non-temporal) 16-byte stores.
typedef unsigned int rej_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void rej_store(RejEntry* p, const RejEntry& e) {
    __builtin_nontemporal_store(*reinterpret_cast<const rej_u4*>(&e), reinterpret_cast<rej_u4*>(p));
}

// Per-block totals of the counts.  The block's 16 loads per thread are issued
// together, raw: a conversion next to each load would serialise the loads.
template <typename T>
__global__ __launch_bounds__(REJ_THREADS) void rej_block_totals(RejArgs a) {
    constexpr int U = REJ_ITEMS / REJ_THREADS;
    const double W = a.st->W;
    const T* w = (const T*)a.weights;
    const int64_t base = (int64_t)blockIdx.x * REJ_ITEMS;
    const int cnt = (int)min<int64_t>(REJ_ITEMS, a.n - base);
    T wr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = u * REJ_THREADS + threadIdx.x;
        wr[u] = (k < cnt) ? __builtin_nontemporal_load(&w[base + k]) : T(0);
    }
    uint64_t acc = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const double x = ((double)a.n * (double)wr[u]) / W;
        acc += (u * REJ_THREADS + (int)threadIdx.x < cnt) ? (uint64_t)((uint32_t)floor(x) + 1u) : 0ull;
    }
    acc = wave_sum_int<uint64_t>(acc);
    __shared__ uint64_t part[REJ_THREADS / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int k = 0; k < REJ_THREADS / 64; ++k) t += part[k];
        a.block_tot[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of the nb block totals (integers, exact in any
// order).  Thread t owns the consecutive run [t q, t q + q) with q = ceil(nb /
// 1024): a sequential pass over its run, one block scan of the run totals,
// a second pass writing the offsets.
__global__ __launch_bounds__(1024) void rej_scan_totals(RejArgs a) {
    __shared__ uint64_t wsum[16];
    const int64_t q = (a.nb + 1023) / 1024;
    const int64_t b0 = min<int64_t>(a.nb, (int64_t)threadIdx.x * q), b1 = min<int64_t>(a.nb, b0 + q);
    // Batches of 16 loads in flight: the run's loads would otherwise wait one
    // after another.
    uint64_t run = 0;
    for (int64_t i0 = b0; i0 < b1; i0 += 16) {
        uint64_t v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = (i0 + j < b1) ? a.block_tot[i0 + j] : 0ull;
#pragma unroll
        for (int j = 0; j < 16; ++j) run += v[j];
    }
    const uint64_t ex = wave_excl_scan_int<uint64_t>(run);
    const uint64_t wt = wave_sum_int<uint64_t>(run);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = wt;
    __syncthreads();
    uint64_t off = ex;
    for (int k = 0; k < (int)(threadIdx.x >> 6); ++k) off += wsum[k];
    for (int64_t i0 = b0; i0 < b1; i0 += 16) {
        uint64_t v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = (i0 + j < b1) ? a.block_tot[i0 + j] : 0ull;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (i0 + j < b1) a.block_off[i0 + j] = off;
            off += v[j];
        }
    }
    if (threadIdx.x == 1023) {
        // The last thread's running offset is the total: its run ends at nb
        // (or is empty, past nb, after every non-empty run).
        a.st->rej_size = off;
        a.st->rej_tasks = 0;
        a.st->learned_fallbacks = 0;  // waves of the learned fill that fell back to bisection
        if (off > a.A_cap) atomicOr(&a.st->error, ERR_BAD_SUM);
    }
}

// A for the block's REJ_ITEMS items.  Wave v owns items [1024 v, 1024 v +
// 1024) of the block and reads them in 16 rounds of 64 (coalesced); the
// counts stay in registers.  One exchange of the waves' totals gives each
// wave its base; then every wave scans its rounds (integers: exact in any
// order) and writes each item's run [first, first + c).  The first slot
// carries the item's acceptance fraction (FP64; the reference computes it in
// float, RejectionSampling.cu:26-28), the others accept always.  The wave
// writes the runs of a round together, consecutive slots by consecutive
// lanes.  In a wave with more than REJ_WAVE_BUDGET slots, the rest of a run
// longer than REJ_WAVE slots becomes tasks of REJ_TASK slots for
// rej_fill_long, so that heavy items, however clustered, hold up no wave for
// long; a wave within the budget writes every run itself.
template <typename T>
__global__ __launch_bounds__(REJ_THREADS) void rej_fill(RejArgs a) {
    constexpr int U = REJ_ITEMS / REJ_THREADS;
    __shared__ uint64_t wtot[REJ_THREADS / 64];
    __shared__ uint32_t wtask[REJ_THREADS / 64];
    __shared__ uint32_t task_base;
    if (a.st->error) return;
    const double W = a.st->W;
    const T* w = (const T*)a.weights;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * REJ_ITEMS + (int64_t)wave * 64 * U;
    const int cnt = (int)max<int64_t>(0, min<int64_t>(64 * U, a.n - base));
    T wr[U];  // raw values: a conversion next to each load would serialise the loads
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
    // A wave with at most REJ_WAVE_BUDGET slots writes all its runs; a wave
    // with more hands the runs longer than REJ_WAVE to tasks.  The tasks are
    // placed with one atomic per block (a counter bumped per round would
    // serialise a grid of waves on it).
    const uint32_t long_min = total > REJ_WAVE_BUDGET ? (uint32_t)REJ_WAVE : 0xFFFFFFFFu;
    uint32_t ntask = 0;
#pragma unroll
    for (int u = 0; u < U; ++u)
        ntask += (c[u] > long_min) ? (c[u] - 1 + REJ_TASK - 1) / REJ_TASK : 0u;
    ntask = wave_sum_int<uint32_t>(ntask);
    if (lane == 0) {
        wtot[wave] = total;
        wtask[wave] = ntask;
    }
    __syncthreads();
    uint64_t first0 = a.block_off[blockIdx.x];
    for (int q = 0; q < wave; ++q) first0 += wtot[q];
    uint32_t block_tasks = 0;
    for (int q = 0; q < REJ_THREADS / 64; ++q) block_tasks += wtask[q];
    uint32_t t_run = 0;  // the wave's next task
    if (block_tasks != 0) {
        if (threadIdx.x == 0) task_base = atomicAdd(&a.st->rej_tasks, block_tasks);
        __syncthreads();
        t_run = task_base;
        for (int q = 0; q < wave; ++q) t_run += wtask[q];
    }
    for (int u = 0; u < U; ++u) {
        const int k = u * 64 + lane;
        // A round's counts sum to at most |A| <= 2n < 2^32: 32-bit scans.
        const uint32_t cu = c[u];
        const uint32_t incl = wave_incl_scan_u32(cu);
        const uint64_t first = first0 + (incl - cu);
        first0 += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const int32_t item = (int32_t)(base + k);
        const double x = ((double)a.n * (double)wr[u]) / W;
        const double frac = x - floor(x);
        // The round's slots in the wave, long runs counted by their first slot
        // only: lane v of each pass writes the v-th of them, owned by the
        // last item whose local start is <= v (a search over the 64 starts).
        const bool long_run = k < cnt && cu > long_min;
        const uint64_t any_long = __ballot(long_run);
        uint32_t lex = incl - cu, L = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        if (any_long != 0ull) {
            const uint32_t ell = long_run ? 1u : cu;
            const uint32_t il = wave_incl_scan_u32(ell);
            lex = il - ell;
            L = (uint32_t)__builtin_amdgcn_readlane((int)il, 63);
        }
        for (uint32_t v0 = 0; v0 < L; v0 += 64) {
            const uint32_t v = v0 + lane;
            int o = 0;
            uint32_t lo_start = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) {
                const uint32_t st = (uint32_t)__shfl((int)lex, o + step, 64);
                if (st <= v) {
                    o += step;
                    lo_start = st;
                }
            }
            const uint64_t f = (uint64_t)__shfl((unsigned long long)first, o, 64);
            const double fr = __shfl(frac, o, 64);
            if (v < L) {
                RejEntry e;
                e.accept = (v == lo_start) ? fr : 1.0;
                e.item = (int32_t)(base + u * 64 + o);
                e.pad = 0;
                rej_store(&a.A[f + (v - lo_start)], e);
            }
        }
        if (any_long != 0ull) {
            const uint32_t nt = long_run ? (cu - 1 + REJ_TASK - 1) / REJ_TASK : 0u;
            const uint32_t ti = wave_incl_scan_u32(nt);
            const uint32_t tex = ti - nt;
            const uint32_t tsum = (uint32_t)__builtin_amdgcn_readlane((int)ti, 63);
            const uint32_t t0 = t_run + tex;
            t_run += tsum;
            if (long_run) {
                if ((uint64_t)t0 + nt > a.tasks_cap) {
                    atomicOr(&a.st->error, ERR_BAD_SUM);
                } else {
                    for (uint32_t j = 0; j < nt; ++j) {
                        RejTask t;
                        t.first = first + 1 + (uint64_t)j * REJ_TASK;
                        t.len = min<uint32_t>(REJ_TASK, cu - 1 - j * REJ_TASK);
                        t.item = item;
                        *reinterpret_cast<uint4*>(&a.tasks[t0 + j]) = *reinterpret_cast<const uint4*>(&t);
                    }
                }
            }
        }
    }
}

// The tasks of rej_fill: one wave per task, REJ_TASK slots with acceptance 1.0.
__global__ __launch_bounds__(256) void rej_fill_long(RejArgs a) {
    if (a.st->error) return;
    const uint32_t nt = min(a.st->rej_tasks, a.tasks_cap);
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t t = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); t < nt; t += waves) {
        const RejTask k = a.tasks[t];
        RejEntry e;
        e.accept = 1.0;
        e.item = k.item;
        e.pad = 0;
        for (uint32_t q = lane; q < k.len; q += 64)
            rej_store(&a.A[k.first + q], e);
    }
}

// Rejection sampler (SamplerRejection.cu:17-35) with the acceptance fraction in
// FP64; bounded trials.
#define REJ_INTERLEAVE 2
template <typename T>
__global__ __launch_bounds__(256) void rej_sample(RejArgs a, SampleArgs sa) {
    if (a.st->error) return;
    const uint64_t Asz = a.st->rej_size;
    if (Asz == 0 || Asz > a.A_cap) return;
    uint64_t csum = 0, trials = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // REJ_INTERLEAVE samples per lane run their trial loops together, so that
    // each lane keeps that many random reads of A in flight (one sample at a
    // time: 4.9 ms per 1e8 samples; two: 4.0 ms; four: no further gain).  Each
    // sample's trials are those of the one-at-a-time loop.
    for (uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g0 < sa.count; g0 += REJ_INTERLEAVE * stride) {
        int32_t it[REJ_INTERLEAVE];
        uint32_t tq[REJ_INTERLEAVE];
        bool d[REJ_INTERLEAVE];
#pragma unroll
        for (int q = 0; q < REJ_INTERLEAVE; ++q) {
            it[q] = 0;
            tq[q] = 0;
            d[q] = g0 + (uint64_t)q * stride >= sa.count;
        }
        while (true) {
            bool all = true;
#pragma unroll
            for (int q = 0; q < REJ_INTERLEAVE; ++q) all = all && d[q];
            if (all) break;
            u32x4 r[REJ_INTERLEAVE];
            uint4 v[REJ_INTERLEAVE];
#pragma unroll
            for (int q = 0; q < REJ_INTERLEAVE; ++q) {
                if (!d[q]) {
                    r[q] = philox_at(g0 + (uint64_t)q * stride, PURPOSE_REJECTION, tq[q], sa.seed);
                    v[q] = *reinterpret_cast<const uint4*>(&a.A[scale64(word64(r[q]), Asz)]);
                }
            }
#pragma unroll
            for (int q = 0; q < REJ_INTERLEAVE; ++q) {
                if (!d[q]) {
                    ++trials;
                    it[q] = (int32_t)v[q].z;
                    const double acc = __longlong_as_double(((long long)v[q].y << 32) | v[q].x);
                    ++tq[q];
                    d[q] = !(coin53(r[q]) > acc) || tq[q] >= REJECTION_MAX_TRIALS;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < REJ_INTERLEAVE; ++q)
            if (g0 + (uint64_t)q * stride < sa.count) emit(sa, g0 + (uint64_t)q * stride, it[q], csum);
    }
    flush_checksum(sa, csum);
    if (a.trials) {
        trials = wave_sum_int<uint64_t>(trials);
        if ((threadIdx.x & 63) == 0) atomicAdd(a.trials, (unsigned long long)trials);
    }
}

// ---------------------------------------------------------------------------
// Sectioned rejection sampling: the trials of SamplerRejection.cu:17-35 inside
// sections of A, with the section counts of SamplerGpuSectioned.cu:58-136.  A
// trial on slot s accepts iff coin53 <= accept(s), so its acceptance
// probability is exactly P(s) = min(floor(accept(s) 2^53) + 1, 2^53) 2^-53.
// An accepted sample lies in section k with probability M_k / sum M, M_k the
// sum of P(s) over the section: the sample count of each section comes from
// the halving tree with q from the exclusive prefix of the M_k, and inside a
// section a sample repeats trials on uniform slots of the LDS-staged section.
#define REJ_SEC_THREADS 1024

__device__ __forceinline__ double rej_accept_prob(double accept) {
    return accept >= 1.0 ? 1.0 : (floor(accept * 0x1.0p53) + 1.0) * 0x1.0p-53;
}

// Section masses, one wave per section: lane l adds slots l, l + 64, ... in
// order, then the pairwise tree (the canonical section sum); 16 loads per
// lane in flight.
__global__ __launch_bounds__(256) void rej_section_mass(RejSecArgs a) {
    const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= a.nsec) return;  // whole waves
    const uint64_t s0 = (uint64_t)k * (uint64_t)a.SS, s1 = min(a.A_size, s0 + (uint64_t)a.SS);
    double acc = 0.0;
    for (uint64_t b = s0; b < s1; b += 16 * 64) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint64_t s = b + (uint64_t)u * 64 + lane;
            v[u] = s < s1 ? a.A[s].accept : 1.0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (b + (uint64_t)u * 64 + lane < s1) acc = acc + rej_accept_prob(v[u]);
    }
    const double m = wave_tree_sum(acc);
    if (lane == 0) a.mass[k] = m;
}

// Exclusive prefix pre[0..nsec] of the section masses by one block, in the
// canonical local-scan association: chunks of 64 masses scanned by the wave
// scan, chunk carries added left to right (carry_{c+1} = carry_c + the
// chunk's total), pre[64 c + l] = carry_c + the chunk's exclusive value.  In
// tiles of REJ_SCAN_TILE chunks: the waves scan the tile's chunks and leave
// their totals in LDS, one thread turns them into carries in order, then the
// waves scan the chunks again and write the prefixes.
#define REJ_SCAN_TILE 4096
__global__ __launch_bounds__(1024) void rej_mass_scan(RejSecArgs a) {
    __shared__ double tot[REJ_SCAN_TILE];
    __shared__ double carry_in;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int64_t nch = (a.nsec + 63) / 64;
    if (threadIdx.x == 0) carry_in = 0.0;
    for (int64_t t0 = 0; t0 < nch; t0 += REJ_SCAN_TILE) {
        const int cnt = (int)min<int64_t>(REJ_SCAN_TILE, nch - t0);
        for (int c = wave; c < cnt; c += nw) {
            const int64_t k = (t0 + c) * 64 + lane;
            const double s = wave_scan(k < a.nsec ? a.mass[k] : 0.0);
            if (lane == 63) tot[c] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double carry = carry_in;
            for (int c = 0; c < cnt; ++c) {
                const double x = tot[c];
                tot[c] = carry;
                carry = carry + x;
            }
            carry_in = carry;
        }
        __syncthreads();
        for (int c = wave; c < cnt; c += nw) {
            const int64_t k = (t0 + c) * 64 + lane;
            const double s = wave_scan(k < a.nsec ? a.mass[k] : 0.0);
            const double ex = tot[c] + wave_shr1(s);
            if (k < a.nsec) a.pre[k] = ex;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) a.pre[a.nsec] = carry_in;
}

// Count and first output of every section, one thread per section.
__global__ __launch_bounds__(64) void rej_section_counts(RejSecArgs a, uint64_t count, uint64_t seed) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nsec) return;
    int64_t lo = 0, hi = a.nsec;
    uint64_t c = count, o = 0, node = 1;
    while (hi - lo > 1) {
        const int64_t mid = lo + (hi - lo) / 2;
        // A range whose mass vanishes in the prefix's rounding splits by
        // section count.
        const double den = a.pre[hi] - a.pre[lo];
        const double q = den > 0.0 ? (a.pre[mid] - a.pre[lo]) / den : (double)(mid - lo) / (double)(hi - lo);
        const uint64_t x = binomial(c, q, seed, node, PURPOSE_REJ_SECTION_COUNTS);
        if (k < mid) {
            hi = mid; c = x; node = 2 * node;
        } else {
            lo = mid; o += x; c = c - x; node = 2 * node + 1;
        }
    }
    a.sec_counts[2 * k] = c;
    a.sec_counts[2 * k + 1] = o;
}

// One block per section: the section's slots staged in LDS, then every
// sample of the section makes trials t = 0, 1, ... with the Philox block at
// counter (output position, PURPOSE_REJ_SECTIONED, t): a uniform slot of the
// section from words 0-1 and the coin from words 2-3.
__global__ __launch_bounds__(REJ_SEC_THREADS) void rej_sample_sectioned(RejSecArgs a, SampleArgs sa) {
    extern __shared__ RejEntry rsect[];
    const int64_t k = blockIdx.x;
    const uint64_t start = (uint64_t)k * (uint64_t)a.SS;
    const uint64_t len = min(a.A_size, start + (uint64_t)a.SS) - start;
    for (uint64_t s = threadIdx.x; s < len; s += blockDim.x)
        *reinterpret_cast<uint4*>(&rsect[s]) = *reinterpret_cast<const uint4*>(&a.A[start + s]);
    __syncthreads();
    const uint64_t cnt = a.sec_counts[2 * k], off = a.sec_counts[2 * k + 1];
    uint64_t csum = 0;
    // One trial per iteration; a lane whose trial accepts goes on to its next
    // sample at once, so no lane waits for another lane's rejected trials
    // (the trials of every sample are those of the one-sample-at-a-time loop).
    uint64_t t = threadIdx.x;
    uint32_t tr = 0;
    while (t < cnt) {
        const uint64_t g = off + t;
        const u32x4 r = philox_at(g, PURPOSE_REJ_SECTIONED, tr, sa.seed);
        const uint4 v = *reinterpret_cast<const uint4*>(&rsect[scale64(word64(r), len)]);
        const double acc = __longlong_as_double(((long long)v.y << 32) | v.x);
        ++tr;
        if (!(coin53(r) > acc) || tr >= REJ_SECTIONED_MAX_TRIALS) {
            emit(sa, g, (int32_t)v.z, csum);
            t += blockDim.x;
            tr = 0;
        }
    }
    flush_checksum(sa, csum);
}

// ---------------------------------------------------------------------------
static inline unsigned grid_for(uint64_t count, unsigned block, unsigned cap) {
    uint64_t g = (count + block - 1) / block;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

hipError_t launch_sample_basic(const SampleArgs& a, hipStream_t s) {
    const unsigned g = grid_for(a.count, 256, 256 * 64);
    if (a.rows) hipLaunchKernelGGL(sample_basic<true>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(sample_basic<false>, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sample_sectioned(const SampleArgs& a, hipStream_t s) {
    const size_t lds = (size_t)a.SS * sizeof(Row);
    hipLaunchKernelGGL(section_counts, dim3((unsigned)((a.nsec + 63) / 64)), dim3(64), 0, s, a);
    if (a.rows) hipLaunchKernelGGL(sample_sectioned<true>, dim3((unsigned)a.nsec), dim3(SECTIONED_THREADS), lds, s, a);
    else hipLaunchKernelGGL(sample_sectioned<false>, dim3((unsigned)a.nsec), dim3(SECTIONED_THREADS), lds, s, a);
    return hipGetLastError();
}

hipError_t set_sample_lds_limit(int bytes) {
    hipError_t e = hipFuncSetAttribute((const void*)sample_sectioned<true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)sample_sectioned<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               bytes);
}

// rej_fill adds the long runs' tasks to a counter that the scan of the block
// totals clears; a repeated fill clears it itself.
__global__ void rej_clear_tasks(DevState* st) { st->rej_tasks = 0; }

static hipError_t launch_direct_fill(const RejArgs& a, hipStream_t s) {
    const unsigned nb = (unsigned)a.nb;
    if (a.dtype_f32) hipLaunchKernelGGL(rej_fill<float>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    else hipLaunchKernelGGL(rej_fill<double>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    hipLaunchKernelGGL(rej_fill_long, dim3(REJ_LONG_BLOCKS), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rejection_build(const RejArgs& a, int fill, hipStream_t s) {
    const unsigned nb = (unsigned)a.nb;
    if (a.dtype_f32) hipLaunchKernelGGL(rej_block_totals<float>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    else hipLaunchKernelGGL(rej_block_totals<double>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    hipLaunchKernelGGL(rej_scan_totals, dim3(1), dim3(1024), 0, s, a);
    if (fill == REJ_FILL_DIRECT) return launch_direct_fill(a, s);
    hipError_t e = launch_rejection_incl(a, fill == REJ_FILL_LEARNED, s);
    if (e != hipSuccess) return e;
    return launch_rejection_search_fill(a, fill, s);
}

hipError_t launch_rejection_fill(const RejArgs& a, int fill, hipStream_t s) {
    if (fill == REJ_FILL_DIRECT) {
        hipLaunchKernelGGL(rej_clear_tasks, dim3(1), dim3(1), 0, s, a.st);
        return launch_direct_fill(a, s);
    }
    return launch_rejection_search_fill(a, fill, s);
}

int64_t rejection_blocks(int64_t n) { return (n + REJ_ITEMS - 1) / REJ_ITEMS; }

// An item with a run of len > REJ_WAVE slots makes ceil((len - 1) / REJ_TASK)
// <= (len - 1) / REJ_TASK + (len - 1) / REJ_WAVE tasks, so |A| <= A_cap slots
// make at most A_cap (1 / REJ_TASK + 1 / REJ_WAVE) tasks.
uint32_t rejection_task_cap(uint64_t A_cap) { return (uint32_t)(A_cap / REJ_TASK + A_cap / REJ_WAVE + 1); }

hipError_t launch_rejection_sample_sectioned(const RejSecArgs& a, const SampleArgs& sa, bool masses, hipStream_t s) {
    if (masses) {
        hipLaunchKernelGGL(rej_section_mass, dim3((unsigned)((a.nsec + 3) / 4)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(rej_mass_scan, dim3(1), dim3(1024), 0, s, a);
    }
    hipLaunchKernelGGL(rej_section_counts, dim3((unsigned)((a.nsec + 63) / 64)), dim3(64), 0, s, a, sa.count, sa.seed);
    const size_t lds = (size_t)a.SS * sizeof(RejEntry);
    hipLaunchKernelGGL(rej_sample_sectioned, dim3((unsigned)a.nsec), dim3(REJ_SEC_THREADS), lds, s, a, sa);
    return hipGetLastError();
}

hipError_t set_rej_sample_lds_limit(int bytes) {
    return hipFuncSetAttribute((const void*)rej_sample_sectioned, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

hipError_t launch_rejection_sample(const RejArgs& a, const SampleArgs& sa, hipStream_t s) {
    const unsigned g = grid_for(sa.count, 256, 256 * 64);
    if (a.dtype_f32) hipLaunchKernelGGL(rej_sample<float>, dim3(g), dim3(256), 0, s, a, sa);
    else hipLaunchKernelGGL(rej_sample<double>, dim3(g), dim3(256), 0, s, a, sa);
    return hipGetLastError();
}

}  // namespace atg
