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
template <bool AOS>
__global__ __launch_bounds__(256) void sample_basic(SampleArgs a) {
    const double W_N = a.st->W_N;
    uint64_t csum = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.count; g += stride) {
        const u32x4 r = philox_at(g, PURPOSE_BASIC, 0, a.seed);
        const int64_t row = (int64_t)scale64(word64(r), (uint64_t)a.n);
        double w;
        int32_t al;
        load_row<AOS>(a, row, w, al);
        emit(a, g, decide(w, al, row, coin53(r), W_N, a.n), csum);
    }
    flush_checksum(a, csum);
}

// ---------------------------------------------------------------------------
// Sectioned sampling.  Section counts: recursive halving of the section range
// with a binomial draw per node (SamplerGpuSectioned.cu:58-96).  The reference
// uses a float normal approximation; restated with IEEE + - * / sqrt floor only.
__device__ uint64_t binomial(uint64_t n, double q, uint64_t seed, uint64_t node) {
    if (n == 0) return 0;
    if (n <= 32) {
        uint64_t cnt = 0;
        u32x4 x{0, 0, 0, 0};
        for (uint64_t t = 0; t < n; ++t) {
            if ((t & 3) == 0) x = philox_at(node, PURPOSE_SECTION_COUNTS, (uint32_t)(t >> 2), seed);
            const uint32_t word = ((t & 3) == 0) ? x.x : ((t & 3) == 1) ? x.y : ((t & 3) == 2) ? x.z : x.w;
            const double u = (double)word * 0x1.0p-32;
            if (u < q) ++cnt;
        }
        return cnt;
    }
    double z = 0.0;
    for (uint32_t d = 0; d < 3; ++d) {
        const u32x4 x = philox_at(node, PURPOSE_SECTION_COUNTS, d, seed);
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

template <bool AOS>
__global__ __launch_bounds__(256) void sample_sectioned(SampleArgs a) {
    extern __shared__ Row sect[];
    __shared__ uint64_t sh_cnt, sh_off;
    const int64_t s = blockIdx.x;
    if (threadIdx.x == 0) {
        uint64_t c, o;
        section_count(a, s, c, o);
        sh_cnt = c;
        sh_off = o;
    }
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
    const uint64_t cnt = sh_cnt, off = sh_off;
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
// scanning the totals, then each block writes its items' slots of A.
#define REJ_ITEMS 4096
#define REJ_THREADS 256
#define REJ_SHORT 8  // runs longer than this are written by the whole block

template <typename T>
__device__ __forceinline__ uint32_t rej_count(const RejArgs& a, int64_t i, double W) {
    const double x = ((double)a.n * (double)((const T*)a.weights)[i]) / W;
    return (uint32_t)floor(x) + 1u;
}

template <typename T>
__global__ __launch_bounds__(REJ_THREADS) void rej_block_totals(RejArgs a) {
    const double W = a.st->W;
    const int64_t base = (int64_t)blockIdx.x * REJ_ITEMS;
    uint64_t acc = 0;
    for (int64_t i = base + threadIdx.x; i < min(a.n, base + REJ_ITEMS); i += REJ_THREADS)
        acc += rej_count<T>(a, i, W);
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

__global__ __launch_bounds__(1024) void rej_scan_totals(RejArgs a) {
    // Single block: exclusive scan of nb block totals (integers).
    __shared__ uint64_t carry_sh;
    if (threadIdx.x == 0) carry_sh = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < a.nb; c0 += 1024) {
        const int64_t i = c0 + threadIdx.x;
        const uint64_t v = (i < a.nb) ? a.block_tot[i] : 0;
        const uint64_t ex = wave_excl_scan_int<uint64_t>(v);
        const uint64_t wt = wave_sum_int<uint64_t>(v);
        __shared__ uint64_t wsum[16];
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = wt;
        __syncthreads();
        uint64_t wbase = 0;
        for (int k = 0; k < (int)(threadIdx.x >> 6); ++k) wbase += wsum[k];
        const uint64_t carry = carry_sh;
        if (i < a.nb) a.block_off[i] = carry + wbase + ex;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = 0;
            for (int k = 0; k < 16; ++k) t += wsum[k];
            carry_sh = carry + t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t total = carry_sh;
        a.st->rej_size = total;
        if (total > a.A_cap) atomicOr(&a.st->error, ERR_BAD_SUM);
    }
}

template <typename T>
__global__ __launch_bounds__(REJ_THREADS) void rej_fill(RejArgs a) {
    __shared__ uint64_t incl[REJ_ITEMS];
    __shared__ uint64_t wsum[REJ_THREADS / 64];
    __shared__ int32_t big[REJ_ITEMS];
    __shared__ uint32_t nbig;
    if (a.st->error) return;
    const double W = a.st->W;
    const T* w = (const T*)a.weights;
    const int64_t base = (int64_t)blockIdx.x * REJ_ITEMS;
    const int cnt = (int)min<int64_t>(REJ_ITEMS, a.n - base);
    if (threadIdx.x == 0) nbig = 0;
    // Per-thread blocked items (16 consecutive), inclusive offsets in LDS.
    const int per = REJ_ITEMS / REJ_THREADS;
    uint64_t loc[REJ_ITEMS / REJ_THREADS];
    uint64_t run = 0;
    for (int u = 0; u < per; ++u) {
        const int k = threadIdx.x * per + u;
        run += (k < cnt) ? rej_count<T>(a, base + k, W) : 0u;
        loc[u] = run;
    }
    const uint64_t ex = wave_excl_scan_int<uint64_t>(run);
    const uint64_t wt = wave_sum_int<uint64_t>(run);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = wt;
    __syncthreads();
    uint64_t wbase = 0;
    for (int k = 0; k < (int)(threadIdx.x >> 6); ++k) wbase += wsum[k];
    const uint64_t boff = a.block_off[blockIdx.x];
    for (int u = 0; u < per; ++u) incl[threadIdx.x * per + u] = boff + wbase + ex + loc[u];
    if (a.incl_out)
        for (int u = 0; u < per; ++u) {
            const int k = threadIdx.x * per + u;
            if (k < cnt) a.incl_out[base + k] = incl[k];
        }
    __syncthreads();
    // Each item's run [incl[k-1], incl[k]) of A.  The first slot carries the
    // item's acceptance fraction (FP64; the reference computes it in float,
    // RejectionSampling.cu:26-28), the others accept always.  Runs of up to
    // REJ_SHORT slots are written by one thread, items striped over the
    // threads so that a wave's stores are nearly contiguous; longer runs are
    // listed and written by the whole block.
    for (int k = threadIdx.x; k < cnt; k += REJ_THREADS) {
        const uint64_t first = k == 0 ? boff : incl[k - 1];
        const uint64_t len = incl[k] - first;
        const double x = ((double)a.n * (double)w[base + k]) / W;
        RejEntry e;
        e.accept = x - floor(x);
        e.item = (int32_t)(base + k);
        e.pad = 0;
        *reinterpret_cast<uint4*>(&a.A[first]) = *reinterpret_cast<const uint4*>(&e);
        if (len > REJ_SHORT) {
            big[atomicAdd(&nbig, 1u)] = k;
        } else {
            e.accept = 1.0;
            for (uint64_t q = 1; q < len; ++q) *reinterpret_cast<uint4*>(&a.A[first + q]) = *reinterpret_cast<const uint4*>(&e);
        }
    }
    __syncthreads();
    for (uint32_t b = 0; b < nbig; ++b) {
        const int k = big[b];
        const uint64_t first = k == 0 ? boff : incl[k - 1];
        const uint64_t len = incl[k] - first;
        RejEntry e;
        e.accept = 1.0;
        e.item = (int32_t)(base + k);
        e.pad = 0;
        for (uint64_t q = 1 + threadIdx.x; q < len; q += REJ_THREADS)
            *reinterpret_cast<uint4*>(&a.A[first + q]) = *reinterpret_cast<const uint4*>(&e);
    }
}

// Rejection sampler (SamplerRejection.cu:17-35) with the acceptance fraction in
// FP64; bounded trials.
template <typename T>
__global__ __launch_bounds__(256) void rej_sample(RejArgs a, SampleArgs sa) {
    if (a.st->error) return;
    const uint64_t Asz = a.st->rej_size;
    if (Asz == 0 || Asz > a.A_cap) return;
    uint64_t csum = 0, trials = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < sa.count; g += stride) {
        int32_t item = 0;
        for (uint32_t tr = 0; tr < REJECTION_MAX_TRIALS; ++tr) {
            const u32x4 r = philox_at(g, PURPOSE_REJECTION, tr, sa.seed);
            ++trials;
            const uint64_t slot = scale64(word64(r), Asz);
            // One 16-byte read per trial: the first slot of a run accepts with
            // probability frac(N w / W), every other slot always.
            const uint4 v = *reinterpret_cast<const uint4*>(&a.A[slot]);
            item = (int32_t)v.z;
            const double accept = __longlong_as_double(((long long)v.y << 32) | v.x);
            if (coin53(r) > accept) continue;
            break;
        }
        emit(sa, g, item, csum);
    }
    flush_checksum(sa, csum);
    if (a.trials) {
        trials = wave_sum_int<uint64_t>(trials);
        if ((threadIdx.x & 63) == 0) atomicAdd(a.trials, (unsigned long long)trials);
    }
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
    if (a.rows) hipLaunchKernelGGL(sample_sectioned<true>, dim3((unsigned)a.nsec), dim3(256), lds, s, a);
    else hipLaunchKernelGGL(sample_sectioned<false>, dim3((unsigned)a.nsec), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t set_sample_lds_limit(int bytes) {
    hipError_t e = hipFuncSetAttribute((const void*)sample_sectioned<true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)sample_sectioned<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               bytes);
}

hipError_t launch_rejection_build(const RejArgs& a, hipStream_t s) {
    const unsigned nb = (unsigned)a.nb;
    if (a.dtype_f32) {
        hipLaunchKernelGGL(rej_block_totals<float>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
        hipLaunchKernelGGL(rej_scan_totals, dim3(1), dim3(1024), 0, s, a);
        hipLaunchKernelGGL(rej_fill<float>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    } else {
        hipLaunchKernelGGL(rej_block_totals<double>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
        hipLaunchKernelGGL(rej_scan_totals, dim3(1), dim3(1024), 0, s, a);
        hipLaunchKernelGGL(rej_fill<double>, dim3(nb), dim3(REJ_THREADS), 0, s, a);
    }
    return hipGetLastError();
}

int64_t rejection_blocks(int64_t n) { return (n + REJ_ITEMS - 1) / REJ_ITEMS; }

hipError_t launch_rejection_sample(const RejArgs& a, const SampleArgs& sa, hipStream_t s) {
    const unsigned g = grid_for(sa.count, 256, 256 * 64);
    if (a.dtype_f32) hipLaunchKernelGGL(rej_sample<float>, dim3(g), dim3(256), 0, s, a, sa);
    else hipLaunchKernelGGL(rej_sample<double>, dim3(g), dim3(256), 0, s, a, sa);
    return hipGetLastError();
}

}  // namespace atg
