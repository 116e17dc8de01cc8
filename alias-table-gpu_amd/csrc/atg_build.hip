// Alias-table construction kernels for gfx950 (the hot path of
// AliasTableSplitGpu::build, AliasTableSplitGpu.cu:388-468).
//
// Phases (one stream, no host round trips):
//   1. sum_sections : per-section FP64 sums (64 sections per block), weight
//                     validation, the first level of the canonical 64-ary tree;
//                     the last block to finish completes the tree -> W, W/N.
//   2. partition_local : one wave per section: all loads issued at once,
//                     stable light/heavy partition into LDS, PSA+ greedy
//                     chains writing finished rows, leftover order, the
//                     canonical local-scan aggregates published into the
//                     dyadic block tree of section sums (completed by the
//                     second child to arrive), leftovers staged in place.
//      partition_place : per section, P(s) from O(log G) completed blocks
//                     (the association the oracle defines), leftover lists
//                     with prefix sums, the split certificate, the sentinels.
//                     No section ever waits for another to finish its work.
//   3. degenerate_fill (only when one leftover class is empty), split, pack.
#include <algorithm>

#include "atg_kernels.h"

namespace atg {

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
        st->degenerate = 0;
        st->cert_max_dL = 0ull;                    // +0.0
        st->cert_min_dH = 0x7FF0000000000000ull;   // +inf
        st->split_fallbacks = 0;
        st->learned_fallbacks = 0;
        st->sum_done = 0;
    }
}

// ---------------------------------------------------------------------------
// 2. PSA+ / plain partition in two passes.  The section prefixes of the
//    leftover sums use the dyadic-block association of the oracle: each
//    section publishes its aggregate as the level-0 block, the child that
//    completes a pair second publishes the parent (pass 1), and P(s) folds
//    the blocks of s's binary digits from high to low (pass 2).
template <typename T, bool AOS>
__global__ __launch_bounds__(64) void partition_local(BuildArgs a) {
    extern __shared__ double lds_raw[];
    if (a.st->error) return;
    psa_local<T, AOS, true>(a, (int64_t)blockIdx.x, lds_raw, (int)threadIdx.x);
}

// Four sections per block, one per wave.
__global__ __launch_bounds__(256) void partition_place(BuildArgs a) {
    if (a.st->error) return;
    const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= a.G) return;
    double unused_l = 0.0, unused_h = 0.0;
    place_section<false>(a, s, (int)(threadIdx.x & 63), unused_l, unused_h);
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
        if (a.rows) hipLaunchKernelGGL((partition_local<float, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((partition_local<float, false>), g, b, lds, s, a);
    } else {
        if (a.rows) hipLaunchKernelGGL((partition_local<double, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((partition_local<double, false>), g, b, lds, s, a);
    }
    hipLaunchKernelGGL(partition_place, dim3(blocks_for(a.G, 4)), dim3(256), 0, s, a);
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

// AoS rows -> separate weight and alias arrays (device to device copy-out).
__global__ __launch_bounds__(256) void split_rows(const Row* rows, double* w, int32_t* al, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 v = *reinterpret_cast<const uint4*>(&rows[i]);
        w[i] = __longlong_as_double(((long long)v.y << 32) | v.x);
        al[i] = (int32_t)v.z;
    }
}

hipError_t launch_split_rows(const Row* rows, double* w, int32_t* al, int64_t n, hipStream_t s) {
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(split_rows, dim3(g > 0 ? g : 1), dim3(256), 0, s, rows, w, al, n);
    return hipGetLastError();
}

hipError_t set_partition_lds_limit(size_t bytes) {
    const void* fns[] = {(const void*)partition_local<double, true>, (const void*)partition_local<double, false>,
                         (const void*)partition_local<float, true>, (const void*)partition_local<float, false>};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace atg
