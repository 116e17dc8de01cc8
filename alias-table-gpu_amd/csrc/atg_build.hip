// Alias-table construction kernels for gfx950 (the hot path of
// AliasTableSplitGpu::build, AliasTableSplitGpu.cu:388-468).
//
// Phases (one stream, no host round trips):
//   1. sum_sections : per-section FP64 sums (one wave per section), weight
//                     validation; sum_finish: the canonical 64-ary tree -> W, W/N.
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

template <typename T>
__global__ __launch_bounds__(256) void sum_sections(BuildArgs a) {
    sum_sections_block<T>(a, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(1024) void sum_finish(BuildArgs a) {
    sum_finish_block(a);
}

// ---------------------------------------------------------------------------
// 2. PSA+ / plain partition in two passes.  The section prefixes of the
//    leftover sums use the dyadic-block association of the oracle: each
//    section publishes its aggregate as the level-0 block, the child that
//    completes a pair second publishes the parent (pass 1), and P(s) folds
//    the blocks of s's binary digits from high to low (pass 2).
// As many waves as fit the chip at once; wave w runs sections w, w + grid,
// ... and loads each next section while the current one's chains run.
// XU: weights per lane held for a section (sections of up to 64 XU items);
// the short-section kernel keeps fewer registers, so more waves fit a SIMD.
template <typename T, bool AOS, int XU>
__global__ __launch_bounds__(64) void partition_local(BuildArgs a) {
    extern __shared__ double lds_raw[];
    if (a.st->error) return;
    const int lane = (int)threadIdx.x;
    T x[XU];
    int64_t s = blockIdx.x;
    if (s < a.G) section_load<T>(a, s, lane, x);
    for (; s < a.G; s += gridDim.x) {
        const int64_t sn = s + gridDim.x;
        psa_section<T, AOS>(a, s, x, lds_raw, lane, [&]() {
            if (sn < a.G) section_load<T>(a, sn, lane, x);
        });
    }
}

// Levels b0 + 1 .. b0 + 9 of the dyadic block tree.
__global__ __launch_bounds__(256) void block_tree(BuildArgs a, int b0) {
    if (a.st->error) return;
    tree_block(a, blockIdx.x, b0);
}

// The split certificate from the per-section data.
__global__ __launch_bounds__(256) void cert_reduce(BuildArgs a) {
    if (a.st->error) return;
    cert_reduce_block(a, blockIdx.x, gridDim.x);
}

// Four sections per block, one per wave.
template <bool LEAN>
__global__ __launch_bounds__(256) void partition_place(BuildArgs a) {
    if (a.st->error) return;
    const int64_t s = (int64_t)blockIdx.x * 4 + wave_in_block();
    if (s >= a.G) return;
    place_section<LEAN>(a, s, (int)(threadIdx.x & 63));
}

// ---------------------------------------------------------------------------
// 3a. Degenerate tables: every leftover row carries its own mass.
template <bool AOS>
__global__ void degenerate_fill(BuildArgs a) {
    degenerate_block<AOS>(a, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// 3b. Split: bisection per k with the reference's predicate; 64 probes bound
//     the loop (the reference spins forever when FOUND is unreachable).
__global__ __launch_bounds__(256) void split_bisect(BuildArgs a, int64_t blk0) {
    split_bisect_block(a, blk0 + blockIdx.x);
}

// ---------------------------------------------------------------------------
// 3c. Pack with LDS staging (the role of PackMethodBasicShared /
//     PackMethodChunkedShared, PackMethodBasicShared.cu:3-68).
template <bool AOS>
__global__ __launch_bounds__(64) void pack_staged(BuildArgs a, int64_t blk0) {
    extern __shared__ double lds_raw[];
    pack_staged_block<AOS>(a, blk0 + blockIdx.x, lds_raw);
}

// 3d. Pack with one lane per chain and the lists read from HBM
//     (PackMethodBasic::packKernel, PackMethodBasic.cu:60-73), or with the
//     weights read through the items (PackMethodNoWeights::packKernel,
//     PackMethodNoWeights.cu:136-145).  Blocks of 64 chains, as pack_staged.
template <bool AOS, bool NOW, typename T>
__global__ __launch_bounds__(64) void pack_basic(BuildArgs a, int64_t blk0) {
    const DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t k = (blk0 + blockIdx.x) * 64 + threadIdx.x;  // chain k + 1
    if (k + 1 > st->p) return;
    const Split prev = a.splits[k], cur = a.splits[k + 1];
    if (NOW) {
        const NoWeightsLists<T> l{a.Li, a.Hi, (const T*)a.weights, a.n};
        pack_chain<AOS>(a, l, prev, cur, st->W_N);
    } else {
        const GlobalLists l{a.Li, a.Hi, a.Lw, a.Hw};
        pack_chain<AOS>(a, l, prev, cur, st->W_N);
    }
}

// 3e. Pack of long chains (far more items per chain than a round of the
//     LDS pool holds): every lane runs its own chain (pack_chain's steps)
//     through windows of its lists staged in LDS.  A round reads the next
//     kWinL lights and the current and next kWinH - 1 heavies of each lane's
//     chain with all loads in flight at once; the lanes then step through
//     their windows until a chain needs an item past its window, recording
//     each row in the window (a light's alias; a heavy's alias and residual
//     weight), and finally the wave writes the round's rows.  Windows are
//     loaded and rows written a few chains per instruction, consecutive list
//     entries on consecutive lanes, so that an instruction touches few lines.
//     Blocks of 64 chains, as pack_staged.
constexpr int kWinL = 16, kWinH = 8;
template <bool AOS>
__global__ __launch_bounds__(64) void pack_windowed(BuildArgs a, int64_t blk0) {
    __shared__ double wl[kWinL][64], wh[kWinH][64];
    __shared__ int32_t il[kWinL][64], ih[kWinH][64], al[kWinL][64], ah[kWinH][64];
    const DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t p = st->p;
    const int lane = (int)threadIdx.x;
    const int64_t k0 = (blk0 + blockIdx.x) * 64;
    if (k0 + 1 > p) return;
    const bool mine = k0 + lane < p;
    Split prev{0, 0, 0.0}, cur{0, 0, 0.0};
    if (mine) {
        prev = a.splits[k0 + lane];
        cur = a.splits[k0 + lane + 1];
    }
    const double W_N = st->W_N;
    const GlobalLists gl{a.Li, a.Hi, a.Lw, a.Hw};
    if (__ballot(mine && !(prev.i <= cur.i && prev.j <= cur.j)) != 0ull) {
        if (mine) pack_chain<AOS>(a, gl, prev, cur, W_N);  // a bisection fell back: no windows
        return;
    }
    const int iUpper = cur.i, jUpper = cur.j;
    int i = prev.i, j = prev.j;
    const int iLast = max(iUpper - 1, i);  // list reads are clamped as in pack_chain
    double w = prev.spill;
    if (mine && prev.spill == 0) w = a.Hw[j];
    bool run = mine;
    const double thr = W_N + ATG_EPSILON;
    while (__ballot(run) != 0ull) {
        // The windows: lights i .. i + kWinL - 1, heavies j .. j + kWinH - 1
        // (slot 0: the current heavy, whose residual is w).
        // Staged a window at a time: one load instruction covers the window
        // of 64 / kWinL chains (heavies: 64 / kWinH), kWinL consecutive list
        // entries each, so that it touches few lines.
        const int i0 = i, j0 = j;
        {
            double vw[kWinL], hw[kWinH];
            int32_t vi[kWinL], hv[kWinH];
#pragma unroll
            for (int k = 0; k < kWinL; ++k) {
                const int c = k * (64 / kWinL) + lane / kWinL, t = lane % kWinL;
                const int q = min(__shfl(i0, c, 64) + t, __shfl(iLast, c, 64));
                const bool load = __shfl((int)run, c, 64) != 0;
                vw[k] = load ? a.Lw[q] : 0.0;
                vi[k] = load ? a.Li[q] : 0;
            }
#pragma unroll
            for (int k = 0; k < kWinH; ++k) {
                const int c = k * (64 / kWinH) + lane / kWinH, t = lane % kWinH;
                const int q = min(__shfl(j0, c, 64) + t, __shfl(jUpper, c, 64));
                const bool load = __shfl((int)run, c, 64) != 0;
                hw[k] = load ? a.Hw[q] : 0.0;
                hv[k] = load ? a.Hi[q] : 0;
            }
#pragma unroll
            for (int k = 0; k < kWinL; ++k) {
                const int c = k * (64 / kWinL) + lane / kWinL, t = lane % kWinL;
                wl[t][c] = vw[k];
                il[t][c] = vi[k];
            }
#pragma unroll
            for (int k = 0; k < kWinH; ++k) {
                const int c = k * (64 / kWinH) + lane / kWinH, t = lane % kWinH;
                wh[t][c] = hw[k];
                ih[t][c] = hv[k];
            }
        }
        wave_lds_sync();
        while (run) {
            const bool idone = i >= iUpper, jdone = j >= jUpper;
            if (idone && jdone) {
                run = false;
                break;
            }
            // A heavy step when the bucket is full (or no light is left) and
            // a heavy remains; a light step otherwise (pack_chain's rule).
            const bool ex = idone || (w <= thr && !jdone);
            const int sl = i - i0, sh = j - j0;
            if (ex ? sh + 1 >= kWinH : sl >= kWinL) break;  // past the window: restage
            // sh < kWinH always (a heavy step needs slot sh + 1); sl can
            // reach kWinL on a heavy step, whose light is not used.
            const int tl = min(sl, kWinL - 1), th1 = min(sh + 1, kWinH - 1);
            const double lw = wl[tl][lane], hw1 = wh[th1][lane];
            const int hj = ih[sh][lane], hi1 = ih[th1][lane];
            if (ex) {
                ah[sh][lane] = hi1;
                wh[sh][lane] = w;
            } else {
                al[tl][lane] = hj;
            }
            w += (ex ? hw1 : lw) - W_N;
            j += ex ? 1 : 0;
            i += ex ? 0 : 1;
        }
        wave_lds_sync();
        // The round's rows (lights i0 .. i - 1, heavies j0 .. j - 1 of each
        // chain), written a window at a time: one store instruction covers
        // the consecutive light rows of 64 / kWinL chains (heavy rows: of
        // 64 / kWinH chains), whose items rise along the lists, so that it
        // touches few table lines.
        const int nl = i - i0, nh = j - j0;
#pragma unroll
        for (int k = 0; k < kWinL; ++k) {
            const int c = k * (64 / kWinL) + lane / kWinL, t = lane % kWinL;
            const int it = il[t][c];
            if (t < __shfl(nl, c, 64) && (uint32_t)it < (uint32_t)a.n) store_row<AOS>(a, it, wl[t][c], al[t][c]);
        }
#pragma unroll
        for (int k = 0; k < kWinH; ++k) {
            const int c = k * (64 / kWinH) + lane / kWinH, t = lane % kWinH;
            const int it = ih[t][c];
            if (t < __shfl(nh, c, 64) && (uint32_t)it < (uint32_t)a.n) store_row<AOS>(a, it, wh[t][c], ah[t][c]);
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
static inline unsigned blocks_for(int64_t items, int per) { return (unsigned)((items + per - 1) / per); }

hipError_t launch_sum(const BuildArgs& a, hipStream_t s) {
    const unsigned g = blocks_for(a.G, 4);
    if (a.dtype_f32) hipLaunchKernelGGL(sum_sections<float>, dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(sum_sections<double>, dim3(g), dim3(256), 0, s, a);
    hipLaunchKernelGGL(sum_finish, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

// LDS of one partition wave: weights (8 B) and local item indices (2 B).
// Workgroups of `threads` threads and `lds` bytes of dynamic LDS of kernel f
// that the device holds at once.
int64_t resident_blocks(const void* f, int threads, size_t lds) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, threads, lds) != hipSuccess || per < 1) {
        (void)hipGetLastError();
        return int64_t(1) << 30;  // unknown: one block per section
    }
    return (int64_t)cus * per;
}

size_t psa_lds_bytes(int64_t Ls_max) { return (((size_t)Ls_max * 10) + 15) & ~(size_t)15; }

int64_t block_tree_total(int64_t G) {
    int64_t total = 0;
    for (int b = 0; ((int64_t)1 << b) <= G; ++b) total += G >> b;
    return total;
}

// The partition phase in its two passes: the local pass (partition, PSA+
// chains, staged leftovers, level-0 block sums), then the block tree and the
// placement of the leftover lists (and the split certificate).
static void (*partition_local_kernel(bool f32, bool aos, int64_t Ls_max))(BuildArgs) {
    constexpr int S = kShortSectionMax / 64, L = PSA_MAX_LEN / 64;
    if (Ls_max <= kShortSectionMax)
        return f32 ? (aos ? partition_local<float, true, S> : partition_local<float, false, S>)
                   : (aos ? partition_local<double, true, S> : partition_local<double, false, S>);
    return f32 ? (aos ? partition_local<float, true, L> : partition_local<float, false, L>)
               : (aos ? partition_local<double, true, L> : partition_local<double, false, L>);
}

hipError_t launch_partition_local(const BuildArgs& a, hipStream_t s) {
    const size_t lds = psa_lds_bytes(a.Ls_max);
    void (*k)(BuildArgs) = partition_local_kernel(a.dtype_f32 != 0, a.rows != nullptr, a.Ls_max);
    const dim3 g((unsigned)std::min<int64_t>(a.G, resident_blocks((const void*)k, 64, lds)));
    hipLaunchKernelGGL(k, g, dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_partition(const BuildArgs& a, hipStream_t s) {
    const hipError_t e = launch_partition_local(a, s);
    if (e != hipSuccess) return e;
    return launch_partition_place(a, s);
}

hipError_t launch_partition_place(const BuildArgs& a, hipStream_t s) {
    for (int b0 = 0; (a.G >> b0) > 1; b0 += 9)
        hipLaunchKernelGGL(block_tree, dim3(blocks_for(a.G >> b0, 512)), dim3(256), 0, s, a, b0);
    if (a.lean_lights) hipLaunchKernelGGL(partition_place<true>, dim3(blocks_for(a.G, 4)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(partition_place<false>, dim3(blocks_for(a.G, 4)), dim3(256), 0, s, a);
    if (a.want_cert)
        hipLaunchKernelGGL(cert_reduce, dim3(blocks_for(a.G, 256)), dim3(256), 0, s, a);  // a section per thread
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void lean_lights_fill(BuildArgs a) {
    lean_lights_fill_block(a, blockIdx.x, gridDim.x);
}

hipError_t launch_lean_lights_fill(const BuildArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(lean_lights_fill, dim3(1024), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_degenerate(const BuildArgs& a, hipStream_t s) {
    if (a.rows) hipLaunchKernelGGL(degenerate_fill<true>, dim3(1024), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(degenerate_fill<false>, dim3(1024), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Blocks of 256 splits (k = 1 .. p_req - 1) and of 64 pack chains (k = 1 .. p_req).
int64_t split_blocks(const BuildArgs& a) { return a.p_req >= 2 ? blocks_for(a.p_req - 1, 256) : 0; }
int64_t pack_blocks(const BuildArgs& a) { return blocks_for(a.p_req, 64); }

hipError_t launch_split_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1) {
    if (b1 > b0) hipLaunchKernelGGL(split_bisect, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, a, b0);
    return hipGetLastError();
}

hipError_t launch_split(const BuildArgs& a, hipStream_t s) {
    hipError_t e = launch_degenerate(a, s);
    if (e != hipSuccess) return e;
    return launch_split_blocks(a, s, 0, split_blocks(a));  // effective p <= p_req
}

hipError_t launch_pack_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1) {
    const size_t lds = pack_lds_bytes(a.pack_cap);
    if (b1 <= b0) return hipSuccess;
    const dim3 g((unsigned)(b1 - b0)), b(64);
    if (a.pack_method == PACK_BASIC) {
        if (a.rows) hipLaunchKernelGGL((pack_basic<true, false, double>), g, b, 0, s, a, b0);
        else hipLaunchKernelGGL((pack_basic<false, false, double>), g, b, 0, s, a, b0);
    } else if (a.pack_method == PACK_NO_WEIGHTS) {
        if (a.dtype_f32) {
            if (a.rows) hipLaunchKernelGGL((pack_basic<true, true, float>), g, b, 0, s, a, b0);
            else hipLaunchKernelGGL((pack_basic<false, true, float>), g, b, 0, s, a, b0);
        } else {
            if (a.rows) hipLaunchKernelGGL((pack_basic<true, true, double>), g, b, 0, s, a, b0);
            else hipLaunchKernelGGL((pack_basic<false, true, double>), g, b, 0, s, a, b0);
        }
    } else if (pack_windowed_ok(a.n, a.p_req, a.pack_method) && !a.lean_lights) {
        if (a.rows) hipLaunchKernelGGL(pack_windowed<true>, g, b, 0, s, a, b0);
        else hipLaunchKernelGGL(pack_windowed<false>, g, b, 0, s, a, b0);
    } else {
        if (a.rows) hipLaunchKernelGGL(pack_staged<true>, g, b, lds, s, a, b0);
        else hipLaunchKernelGGL(pack_staged<false>, g, b, lds, s, a, b0);
    }
    return hipGetLastError();
}

hipError_t launch_pack(const BuildArgs& a, hipStream_t s) { return launch_pack_blocks(a, s, 0, pack_blocks(a)); }

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

// Separate weight and alias arrays -> AoS rows (a table loaded for sampling).
__global__ __launch_bounds__(256) void join_rows(const double* w, const int32_t* al, Row* rows, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Row r;
        r.weight = w[i];
        r.alias = al[i];
        r.pad = 0;
        *reinterpret_cast<uint4*>(&rows[i]) = *reinterpret_cast<const uint4*>(&r);
    }
}

hipError_t launch_join_rows(const double* w, const int32_t* al, Row* rows, int64_t n, hipStream_t s) {
    const int64_t g = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(join_rows, dim3(g > 0 ? g : 1), dim3(256), 0, s, w, al, rows, n);
    return hipGetLastError();
}

hipError_t set_partition_lds_limit(size_t bytes) {
    constexpr int S = kShortSectionMax / 64, L = PSA_MAX_LEN / 64;
    const void* fns[] = {(const void*)partition_local<double, true, L>, (const void*)partition_local<double, false, L>,
                         (const void*)partition_local<float, true, L>, (const void*)partition_local<float, false, L>,
                         (const void*)partition_local<double, true, S>, (const void*)partition_local<double, false, S>,
                         (const void*)partition_local<float, true, S>, (const void*)partition_local<float, false, S>};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace atg
