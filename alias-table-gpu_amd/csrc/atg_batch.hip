// Batched construction of many independent alias tables in one persistent
// launch (no counterpart in the reference, which builds one table per object:
// AliasTableSplitGpu.cu:388-468).
//
// One workgroup builds one table from start to end; the workgroups take
// tables from a queue in longest-first order, so the largest tables start
// first and the grid drains when the queue is empty.  Inside a table the
// phases are the single-table phases with workgroup barriers instead of
// kernel boundaries, and every phase calls the same device functions with
// the same association, so a table built here is bit-identical to the same
// weights built alone with the same options:
//   W       : section sums by the waves, then the 64-ary tree level by level;
//   PSA+    : wave w takes sections w, w + nw, ...: pass 1 (partition,
//             chains, staged leftovers, level-0 block sums), a barrier, the
//             upper block sums level by level, then pass 2 (section
//             prefixes, leftover lists, certificate);
//   split   : bisection or the learned index per k;
//   pack    : one chain per k.
// Device timestamps (the 100 MHz wall clock) at the start and end of every
// table give per-table build latencies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "atg_host.h"
#include "atg_kernels.h"
#include "atg_sample.h"

namespace atg {

struct BatchQueue {
    uint32_t next;   // next table ticket
    uint32_t done;   // workgroups that found the queue empty
    uint32_t pad[62];
};

// Per-table build state on its own 256-byte line pair, so that no cache line
// is shared between tables that different workgroups build.
struct alignas(256) BatchState {
    DevState s;
};

struct BatchBuildArgs {
    const BuildArgs* tables;  // per-table arguments (device)
    const int32_t* order;     // longest-first table order
    int64_t ntables;
    BatchQueue* q;
    uint64_t* ticks;          // [2 t] start, [2 t + 1] end
    int64_t wave_lds_doubles; // LDS stride of one wave's section buffer
    int32_t learned;
};

struct BatchSampleArgs {
    const Row* rows;
    const int64_t* offsets;   // ntables + 1
    const BatchState* st;
    int64_t ntables;
    uint64_t per_table;
    uint64_t seed;
    int32_t* out;             // ntables * per_table
};

// Workgroup-shared scratch of one table build.
struct TableLDS {
    BuildArgs a;
    uint32_t idx;
    uint32_t flags;
    uint32_t split_fallbacks;
    uint32_t learned_fallbacks;
    uint64_t cert_l, cert_h;  // bit patterns (non-negative doubles order like integers)
};

// W of one table: section sums (lane l adds items l, l + 64, ...), then the
// 64-ary tree over the section sums, 64 values per wave and level, zero-padded
// (the association of sum_sections).  Returns the error bits.
template <typename T>
__device__ __forceinline__ uint32_t wg_weight_sum(const BuildArgs& a, int tid, int nthreads, TableLDS& sh) {
    const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    const T* w = (const T*)a.weights;
    for (int64_t s = wave; s < a.G; s += nw) {
        const double sum = section_sum<T>(w, min(a.n, s * a.Ls), min(a.n, (s + 1) * a.Ls), lane);
        if (lane == 0) a.section_sums[s] = sum;
    }
    __syncthreads();
    double* bufs[2] = {a.tree_scratch, a.tree_scratch + a.tree_scratch_half};
    const double* src = a.section_sums;
    int cur = 0;
    int64_t count = a.G;
    while (true) {
        const int64_t groups = (count + 63) / 64;
        for (int64_t g = wave; g < groups; g += nw) {
            const int64_t i = g * 64 + lane;
            const double r = wave_tree_sum((i < count) ? src[i] : 0.0);
            if (lane == 0) bufs[cur][g] = r;
        }
        __syncthreads();
        src = bufs[cur];
        cur ^= 1;
        count = groups;
        if (count == 1) break;
    }
    if (tid == 0) {
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
        st->p = 0;
        sh.flags = err;
        sh.split_fallbacks = 0;
        sh.learned_fallbacks = 0;
        sh.cert_l = 0ull;                   // +0.0
        sh.cert_h = 0x7FF0000000000000ull;  // +inf
    }
    __syncthreads();
    return sh.flags;
}

// One table by one workgroup.  Communication between the waves goes through
// plain global memory ordered by workgroup barriers and through LDS; there is
// no waiting on another wave's progress and no global atomic.
template <typename T>
__device__ __forceinline__ void wg_build_table(const BuildArgs& a, double* lds_raw, int64_t wave_lds_doubles,
                                               bool learned, int tid, int nthreads, TableLDS& sh) {
    const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    DevState* st = a.st;
    if (wg_weight_sum<T>(a, tid, nthreads, sh)) return;

    // Partition pass 1 per section, the dyadic block sums, pass 2 per section.
    double* mine = lds_raw + (int64_t)wave * wave_lds_doubles;
    for (int64_t s = wave; s < a.G; s += nw) psa_local<T, true, false>(a, s, mine, lane);
    __syncthreads();
    wg_block_tree(a, tid, nthreads);
    double ext_l = 0.0, ext_h = __builtin_inf();
    for (int64_t s = wave; s < a.G; s += nw) place_section<true>(a, s, lane, ext_l, ext_h);
    if (lane == 0) {
        atomicMax((unsigned long long*)&sh.cert_l, (unsigned long long)d2u(ext_l));
        atomicMin((unsigned long long*)&sh.cert_h, (unsigned long long)d2u(ext_h));
    }
    __syncthreads();
    if (tid == 0) {
        st->cert_max_dL = sh.cert_l;
        st->cert_min_dH = sh.cert_h;
        sh.flags = st->degenerate;
    }
    __syncthreads();
    if (sh.flags) {
        // One leftover class is empty: every leftover row carries its own mass.
        const int64_t nLr = st->numL - 1, nHr = st->numH - 1;
        for (int64_t t = tid; t < nLr + nHr; t += nthreads) {
            const int32_t item = (t < nLr) ? a.Li[t] : a.Hi[t - nLr];
            const double w = (t < nLr) ? a.Lw[t] : a.Hw[t - nLr];
            store_row<true>(a, item, w, item);
        }
        if (tid == 0) {
            st->split_fallbacks = 0;
            st->learned_fallbacks = 0;
        }
        return;
    }

    // Split.
    const int64_t p = st->p, numL = st->numL, numH = st->numH;
    uint32_t sfb = 0, lfb = 0;
    if (learned) {
        const int64_t K = min<int64_t>(a.learned_segments, p);
        for (int64_t m = tid; m <= K; m += nthreads) sfb += knot_split(a, m);
        __syncthreads();
        for (int64_t k = 1 + tid; k < p; k += nthreads) {
            const uint32_t f = learned_split(a, k);
            sfb += f & 1u;
            lfb += (f >> 1) & 1u;
        }
    } else {
        if (tid == 0) {
            a.splits[0] = Split{0, 0, 0.0};
            a.splits[p] = Split{(int32_t)(numL - 1), (int32_t)(numH - 1), 0.0};
        }
        for (int64_t k = 1 + tid; k < p; k += nthreads) {
            Split sp;
            if (!bisect_split(k, p, st->nToSplit, st->wToSplit, a.Lp, numL, a.Hp, a.Hw, numH, sp)) ++sfb;
            a.splits[k] = sp;
        }
    }
    if (sfb) atomicAdd(&sh.split_fallbacks, sfb);
    if (lfb) atomicAdd(&sh.learned_fallbacks, lfb);
    __syncthreads();
    if (tid == 0) {
        st->split_fallbacks = sh.split_fallbacks;
        st->learned_fallbacks = sh.learned_fallbacks;
    }

    // Pack: one chain per section of the split.
    const double W_N = st->W_N;
    const GlobalLists l{a.Li, a.Hi, a.Lw, a.Hw};
    for (int64_t k = 1 + tid; k <= p; k += nthreads) pack_chain<true>(a, l, a.splits[k - 1], a.splits[k], W_N);
}

// Persistent: every workgroup takes tables until the queue is empty; the last
// workgroup to find it empty resets the queue for the next launch (so a
// replayed graph needs no reset node).
template <typename T>
__global__ __launch_bounds__(1024) void batch_build(BatchBuildArgs b) {
    extern __shared__ double lds_raw[];
    __shared__ TableLDS sh;
    const int tid = threadIdx.x, nthreads = blockDim.x;
    while (true) {
        if (tid == 0) sh.idx = __hip_atomic_fetch_add(&b.q->next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t idx = sh.idx;
        if (idx >= (uint64_t)b.ntables) break;
        const int32_t t = b.order[idx];
        if (tid == 0) {
            sh.a = b.tables[t];
            b.ticks[2 * t] = wall_clock64();
        }
        __syncthreads();
        wg_build_table<T>(sh.a, lds_raw, b.wave_lds_doubles, b.learned != 0, tid, nthreads, sh);
        __syncthreads();
        if (tid == 0) b.ticks[2 * t + 1] = wall_clock64();
    }
    if (tid == 0) {
        const uint32_t d = __hip_atomic_fetch_add(&b.q->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == gridDim.x - 1) {
            // Every workgroup has taken its last ticket.
            __hip_atomic_store(&b.q->next, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&b.q->done, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Basic alias sampling of every table: sample k of table t uses the Philox
// block at counter (t * per_table + k, PURPOSE_BATCH); the result is an item
// index local to the table (-1 when the table failed to build).
__global__ __launch_bounds__(256) void batch_sample(BatchSampleArgs a) {
    const uint64_t total = (uint64_t)a.ntables * a.per_table;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const int64_t t = (int64_t)(g / a.per_table);
        const int64_t off = a.offsets[t], n = a.offsets[t + 1] - off;
        const DevState& st = a.st[t].s;
        if (st.error) {
            a.out[g] = -1;
            continue;
        }
        const u32x4 r = philox_at(g, PURPOSE_BATCH, 0, a.seed);
        const int64_t row = (int64_t)scale64(word64(r), (uint64_t)n);
        const uint4 v = *reinterpret_cast<const uint4*>(&a.rows[off + row]);
        const double w = __longlong_as_double(((long long)v.y << 32) | v.x);
        const int32_t al = (int32_t)v.z;
        a.out[g] = decide(w, al, row, coin53(r), st.W_N, n);
    }
}

}  // namespace atg

using namespace atg;

struct atg_batch_s {
    int64_t ntables = 0, total = 0;
    atg_dtype dtype = ATG_F64;
    atg_build_opts opts{};
    int device = 0;
    std::vector<int64_t> offsets;
    std::vector<BuildArgs> args;        // host copy (pointers into the pools below)
    int64_t Ls_max = 1;
    int nwaves = 4;
    int grid = 0;
    size_t lds_bytes = 0;
    double wall_khz = 100000.0;
    // device memory
    void* d_weights_owned = nullptr;
    const void* d_weights = nullptr;
    BuildArgs* d_args = nullptr;
    int32_t* d_order = nullptr;
    int64_t* d_offsets = nullptr;
    BatchQueue* d_queue = nullptr;
    uint64_t* d_ticks = nullptr;
    BatchState* d_state = nullptr;
    Row* d_rows = nullptr;
    double* d_dpool = nullptr;          // section sums, tree, L/H weights and prefixes
    uint64_t* d_slots = nullptr;
    int32_t* d_ipool = nullptr;         // L/H items, knots
    Split* d_splits = nullptr;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool weights_set = false, built = false;
};

namespace {

inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

void free_batch(atg_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    void* ptrs[] = {b->d_weights_owned, b->d_args, b->d_order, b->d_offsets, b->d_queue, b->d_ticks, b->d_state,
                    b->d_rows, b->d_dpool, b->d_slots, b->d_ipool, b->d_splits};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// Points every table's weights at its slice of the concatenated input.
void bind_weights(atg_batch* b) {
    const size_t es = b->dtype == ATG_F32 ? 4 : 8;
    for (int64_t t = 0; t < b->ntables; ++t)
        b->args[t].weights = (const char*)b->d_weights + (size_t)b->offsets[t] * es;
}

template <typename T>
hipError_t launch_batch(const atg_batch* b, hipStream_t s) {
    BatchBuildArgs k{};
    k.tables = b->d_args;
    k.order = b->d_order;
    k.ntables = b->ntables;
    k.q = b->d_queue;
    k.ticks = b->d_ticks;
    k.wave_lds_doubles = (int64_t)(psa_lds_bytes(b->Ls_max) / sizeof(double));
    k.learned = b->opts.split_method == ATG_SPLIT_LEARNED;
    hipLaunchKernelGGL(batch_build<T>, dim3(b->grid), dim3(64 * b->nwaves), b->lds_bytes, s, k);
    return hipGetLastError();
}

}  // namespace

extern "C" {

atg_status atg_batch_create(atg_batch** out, const int64_t* offsets, int64_t ntables, atg_dtype dtype,
                            const atg_build_opts* opts) {
    if (!out || !offsets) return fail(ATG_EINVAL, "atg_batch_create: NULL argument");
    *out = nullptr;
    if (ntables < 1 || ntables > 0x7FFFFFFF) return fail(ATG_EINVAL, "atg_batch_create: ntables=%lld", (long long)ntables);
    if (dtype != ATG_F64 && dtype != ATG_F32) return fail(ATG_EINVAL, "atg_batch_create: bad dtype");
    if (offsets[0] != 0) return fail(ATG_EINVAL, "atg_batch_create: offsets[0] must be 0");
    for (int64_t t = 0; t < ntables; ++t) {
        const int64_t n = offsets[t + 1] - offsets[t];
        if (n < 1 || n > kMaxN)
            return fail(ATG_EINVAL, "atg_batch_create: table %lld has %lld items", (long long)t, (long long)n);
    }
    atg_build_opts o;
    if (opts) o = *opts; else atg_default_opts(&o);
    if (o.chains < 1 || o.chains > 64) return fail(ATG_EINVAL, "atg_batch_create: chains=%d outside [1, 64]", o.chains);
    if (o.split_method != ATG_SPLIT_BISECT && o.split_method != ATG_SPLIT_LEARNED)
        return fail(ATG_EINVAL, "atg_batch_create: unknown split method %d", o.split_method);
    if (o.learned_segments <= 0) o.learned_segments = 4096;
    const int64_t p_auto = 32LL * device_cus() * 10;
    int nwaves = o.batch_waves > 0 ? o.batch_waves : 4;
    if (nwaves < 1 || nwaves > 16) return fail(ATG_EINVAL, "atg_batch_create: %d waves per table", nwaves);

    atg_batch* b = new atg_batch_s();
    b->ntables = ntables;
    b->total = offsets[ntables];
    b->dtype = dtype;
    b->opts = o;
    b->nwaves = nwaves;
    b->offsets.assign(offsets, offsets + ntables + 1);
    b->args.resize((size_t)ntables);
    (void)hipGetDevice(&b->device);
    auto cleanup = [&](atg_status st) { free_batch(b); delete b; return st; };

    // Per-table layout exactly as atg_create computes it for one table; the
    // scratch of each table starts on a 256-byte boundary.
    struct Sizes { int64_t G, Ls, p_req, half, tree_total; };
    std::vector<Sizes> sz((size_t)ntables);
    int64_t nd = 0, ns = 0, ni = 0, nsp = 0;
    std::vector<int64_t> od(ntables), os(ntables), oi(ntables), osp(ntables);
    for (int64_t t = 0; t < ntables; ++t) {
        const int64_t n = offsets[t + 1] - offsets[t];
        int64_t S = o.section_size > 0 ? o.section_size : std::min<int64_t>(1000, n / 10);
        if (S < 1) S = 1;
        Sizes& z = sz[t];
        section_layout(n, S, &z.G, &z.Ls);
        if (z.Ls > kMaxSectionLen)
            return cleanup(fail(ATG_EINVAL, "atg_batch_create: table %lld section length %lld exceeds %lld",
                                (long long)t, (long long)z.Ls, (long long)kMaxSectionLen));
        b->Ls_max = std::max(b->Ls_max, z.Ls);
        const int64_t nsplits = o.num_splits > 0 ? o.num_splits : p_auto;
        z.p_req = std::max<int64_t>(2, std::min<int64_t>(nsplits, n / 10));
        z.half = (z.G + 63) / 64 + 1;
        z.tree_total = block_tree_total(z.G);
        od[t] = nd;
        nd += align_up(z.G, 32) + align_up(2 * z.half, 32) + 5 * align_up(n + 2, 32);
        os[t] = ns;
        ns += align_up(3 * z.tree_total + (z.tree_total + 1) / 2 + 1, 32);
        oi[t] = ni;
        ni += 3 * align_up(n + 2, 64) + align_up(o.learned_segments + 2, 64);
        osp[t] = nsp;
        nsp += align_up(z.p_req + 1, 16);
    }
#define BALLOC(ptr, cnt)                                                                          \
    do {                                                                                          \
        if (dalloc(&(ptr), (size_t)(cnt)) != hipSuccess) {                                        \
            (void)hipGetLastError();                                                              \
            return cleanup(fail(ATG_ENOMEM, "atg_batch_create: hipMalloc of %s (%lld items) failed", #ptr, \
                                (long long)(cnt)));                                               \
        }                                                                                         \
    } while (0)
    BALLOC(b->d_args, ntables);
    BALLOC(b->d_order, ntables);
    BALLOC(b->d_offsets, ntables + 1);
    BALLOC(b->d_queue, 1);
    BALLOC(b->d_ticks, 2 * ntables);
    BALLOC(b->d_state, ntables);
    BALLOC(b->d_rows, b->total);
    BALLOC(b->d_dpool, nd);
    BALLOC(b->d_slots, ns);
    BALLOC(b->d_ipool, ni);
    BALLOC(b->d_splits, nsp);
#undef BALLOC
    for (int64_t t = 0; t < ntables; ++t) {
        const Sizes& z = sz[t];
        const int64_t n = offsets[t + 1] - offsets[t];
        BuildArgs& a = b->args[t];
        a = BuildArgs{};
        a.dtype_f32 = dtype == ATG_F32;
        a.n = n;
        a.G = z.G;
        a.Ls = z.Ls;
        a.Ls_max = b->Ls_max;  // the LDS stride of one wave is sized for the longest section
        a.tree_total = z.tree_total;
        a.st = &b->d_state[t].s;
        double* d = b->d_dpool + od[t];
        a.section_sums = d;
        d += align_up(z.G, 32);
        a.tree_scratch = d;
        a.tree_scratch_half = z.half;
        d += align_up(2 * z.half, 32);
        const int64_t cap = align_up(n + 2, 32);
        a.Lw = d; a.Lp = d + cap; a.Hw = d + 2 * cap; a.Hp = d + 3 * cap; a.stage_w = d + 4 * cap;
        a.slots = b->d_slots + os[t];
        int32_t* ip = b->d_ipool + oi[t];
        const int64_t icap = align_up(n + 2, 64);
        a.Li = ip; a.Hi = ip + icap; a.stage_i = ip + 2 * icap; a.knot_j = ip + 3 * icap;
        a.rows = b->d_rows + offsets[t];
        a.tw = nullptr;
        a.ta = nullptr;
        a.psa_plus = o.psa_plus;
        a.chains = o.chains;
        a.p_req = z.p_req;
        a.splits = b->d_splits + osp[t];
        a.learned_segments = o.learned_segments;
    }
    // Longest tables first (ties by index).
    std::vector<int32_t> order((size_t)ntables);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
    });
    if (hipMemcpy(b->d_order, order.data(), sizeof(int32_t) * ntables, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_offsets, offsets, sizeof(int64_t) * (ntables + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_state, 0, sizeof(BatchState) * ntables) != hipSuccess ||
        hipMemset(b->d_queue, 0, sizeof(BatchQueue)) != hipSuccess ||
        hipMemset(b->d_ticks, 0, sizeof(uint64_t) * 2 * ntables) != hipSuccess)
        return cleanup(fail(ATG_EDEVICE, "atg_batch_create: upload failed"));

    // Launch shape: as many resident workgroups as LDS and registers allow.
    b->lds_bytes = (size_t)b->nwaves * psa_lds_bytes(b->Ls_max);
    const void* fn = dtype == ATG_F32 ? (const void*)batch_build<float> : (const void*)batch_build<double>;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds_bytes) != hipSuccess)
        return cleanup(fail(ATG_EDEVICE, "atg_batch_create: LDS limit of %zu bytes", b->lds_bytes));
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * b->nwaves, b->lds_bytes) != hipSuccess ||
        per_cu < 1)
        return cleanup(fail(ATG_EDEVICE, "atg_batch_create: no resident workgroup fits (%zu B LDS)", b->lds_bytes));
    const int64_t g = (int64_t)per_cu * device_cus();
    b->grid = (int)std::min<int64_t>(g, ntables);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess && khz > 0)
        b->wall_khz = khz;
    if (hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(ATG_EDEVICE, "atg_batch_create: stream"));
    *out = b;
    return ATG_OK;
}

void atg_batch_destroy(atg_batch* b) {
    if (!b) return;
    (void)hipDeviceSynchronize();
    free_batch(b);
    delete b;
}

atg_status atg_batch_set_weights(atg_batch* b, const void* w, atg_mem mem, void* stream) {
    if (!b || !w) return fail(ATG_EINVAL, "atg_batch_set_weights: NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)b->total * (b->dtype == ATG_F32 ? 4 : 8);
    if (mem == ATG_MEM_DEVICE) {
        b->d_weights = w;
    } else if (mem == ATG_MEM_HOST) {
        if (!b->d_weights_owned) HIPCHK(hipMalloc(&b->d_weights_owned, bytes));
        HIPCHK(hipMemcpyAsync(b->d_weights_owned, w, bytes, hipMemcpyHostToDevice, s));
        b->d_weights = b->d_weights_owned;
    } else {
        return fail(ATG_EINVAL, "atg_batch_set_weights: bad mem kind");
    }
    bind_weights(b);
    HIPCHK(hipMemcpyAsync(b->d_args, b->args.data(), sizeof(BuildArgs) * b->ntables, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    b->weights_set = true;
    b->built = false;
    return ATG_OK;
}

// The build launch is captured once into a hipGraph and replayed on the
// caller's stream (the queue resets itself at the end of every launch).
atg_status atg_batch_build(atg_batch* b, void* stream) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_build: NULL batch");
    if (!b->weights_set) return fail(ATG_ESTATE, "atg_batch_build: weights not set");
    if (!b->exec) {
        HIPCHK(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
        const hipError_t e = b->dtype == ATG_F32 ? launch_batch<float>(b, b->capture) : launch_batch<double>(b, b->capture);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->capture, &g);
        HIPCHK(e);
        HIPCHK(e2);
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
    b->built = true;
    return ATG_OK;
}

atg_status atg_batch_sync(atg_batch* b, void* stream) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_sync: NULL batch");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    std::vector<BatchState> st((size_t)b->ntables);
    HIPCHK(hipMemcpy(st.data(), b->d_state, sizeof(BatchState) * b->ntables, hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < b->ntables; ++t) {
        if (st[t].s.error) {
            const atg_status rc = state_status(st[t].s);
            return fail(rc, "table %lld: %s", (long long)t, atg_last_error());
        }
    }
    return ATG_OK;
}

atg_status atg_batch_get_tables(atg_batch* b, double* weights, int32_t* aliases, double* W_per_table) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_get_tables: NULL batch");
    if (!b->built) return fail(ATG_ESTATE, "atg_batch_get_tables: not built");
    HIPCHK(hipDeviceSynchronize());
    const size_t n = (size_t)b->total;
    if (weights || aliases) {
        std::vector<Row> rows(n);
        HIPCHK(hipMemcpy(rows.data(), b->d_rows, n * sizeof(Row), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            if (weights) weights[i] = rows[i].weight;
            if (aliases) aliases[i] = rows[i].alias;
        }
    }
    if (W_per_table) {
        std::vector<BatchState> st((size_t)b->ntables);
        HIPCHK(hipMemcpy(st.data(), b->d_state, sizeof(BatchState) * b->ntables, hipMemcpyDeviceToHost));
        for (int64_t t = 0; t < b->ntables; ++t) W_per_table[t] = st[t].s.W;
    }
    return ATG_OK;
}

atg_status atg_batch_sample(atg_batch* b, uint64_t per_table, uint64_t seed, int32_t* out, void* stream) {
    if (!b || !out) return fail(ATG_EINVAL, "atg_batch_sample: NULL argument");
    if (!b->built) return fail(ATG_ESTATE, "atg_batch_sample: not built");
    if (per_table == 0) return ATG_OK;
    BatchSampleArgs a{};
    a.rows = b->d_rows;
    a.offsets = b->d_offsets;
    a.st = b->d_state;
    a.ntables = b->ntables;
    a.per_table = per_table;
    a.seed = seed;
    a.out = out;
    const uint64_t total = (uint64_t)b->ntables * per_table;
    const uint64_t blocks = std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(batch_sample, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, a);
    HIPCHK(hipGetLastError());
    return ATG_OK;
}

atg_status atg_batch_latencies(atg_batch* b, double* ns_per_table) {
    if (!b || !ns_per_table) return fail(ATG_EINVAL, "atg_batch_latencies: NULL argument");
    HIPCHK(hipDeviceSynchronize());
    std::vector<uint64_t> tk((size_t)(2 * b->ntables));
    HIPCHK(hipMemcpy(tk.data(), b->d_ticks, sizeof(uint64_t) * 2 * b->ntables, hipMemcpyDeviceToHost));
    const double ns_per_tick = 1e6 / b->wall_khz;
    for (int64_t t = 0; t < b->ntables; ++t) ns_per_table[t] = (double)(tk[2 * t + 1] - tk[2 * t]) * ns_per_tick;
    return ATG_OK;
}

}  // extern "C"
