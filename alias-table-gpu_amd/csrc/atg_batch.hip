// Batched construction of many independent alias tables (no counterpart in
// the reference, which builds one table per object: AliasTableSplitGpu.cu:388-468).
//
// The tables are dealt into groups (largest table first, each to the group
// with the fewest items so far; one group unless batch_group_items is set).
// Every phase of a group is one launch over the group's tables: a block finds
// its table and its block within the table from a block->table map built at
// creation, then runs exactly the device code of the single-table kernels
// (the same launch shape per table: one wave per section, one wave per 64
// pack chains).  A table built here is therefore bit-identical to the same
// weights built alone with the same options.  The groups run one after the
// other, so a table is done when its group is.  Per-table latencies come from
// the 100 MHz device clock: the first block of the table's sum phase to the
// last block of its pack.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "atg_host.h"
#include "atg_kernels.h"
#include "atg_sample.h"

namespace atg {

// Per-table build state on its own 256-byte line pair, so that no cache line
// is shared between tables.
struct alignas(256) BatchState {
    DevState s;
};

// Global block b of a phase belongs to table tab[b]; its block within the
// table is b - base[tab[b]].
struct SegMap {
    const int32_t* tab;
    const int64_t* base;
    int64_t blocks;
};

struct SegArgs {
    const BuildArgs* tables;
    int64_t ntables;
    uint64_t* ticks;  // [2 t] first start, [2 t + 1] last end (device clock)
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

// Blocks of the degenerate-table fill per table: one per 65,536 items, at
// most 8 (a grid-stride loop covers the rest).  Most tables are not
// degenerate and leave at once, so few blocks keep the launch short.
__host__ __device__ inline int64_t degenerate_blocks(int64_t n) {
    const int64_t b = (n + 2 + 65535) / 65536;
    return b < 8 ? b : 8;
}

__device__ __forceinline__ void seg_locate(const SegMap& m, int64_t& t, int64_t& blk) {
    t = m.tab[blockIdx.x];
    blk = (int64_t)blockIdx.x - m.base[t];
}

__device__ __forceinline__ void tick_end(const SegArgs& g, int64_t t) {
    if (threadIdx.x == 0)
        atomicMax((unsigned long long*)&g.ticks[2 * t + 1], (unsigned long long)wall_clock64());
}

__global__ __launch_bounds__(256) void seg_reset_ticks(SegArgs g) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.ntables; t += (int64_t)gridDim.x * blockDim.x) {
        g.ticks[2 * t] = ~0ull;
        g.ticks[2 * t + 1] = 0ull;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void seg_sum(SegArgs g, SegMap m) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    if (threadIdx.x == 0) atomicMin((unsigned long long*)&g.ticks[2 * t], (unsigned long long)wall_clock64());
    const BuildArgs& a = g.tables[t];
    sum_sections_block<T>(a, blk, (a.G + 3) / 4);
}

// One block per table of the group.
__global__ __launch_bounds__(256) void seg_sum_finish(SegArgs g, SegMap m) {
    sum_finish_block(g.tables[m.tab[blockIdx.x]]);
}

// Sections of all tables, grid-stride over the phase's blocks (one section
// each); a wave loads its next section while the current one's chains run.
template <typename T>
__global__ __launch_bounds__(64) void seg_local(SegArgs g, SegMap m) {
    extern __shared__ double lds_raw[];
    const int lane = (int)threadIdx.x;
    T x[PSA_MAX_LEN / 64];
    auto load = [&](int64_t v) {
        const int64_t t = m.tab[v];
        const BuildArgs& a = g.tables[t];
        if (!a.st->error) section_load<T>(a, v - m.base[t], lane, x);
    };
    int64_t v = blockIdx.x;
    if (v < m.blocks) load(v);
    for (; v < m.blocks; v += gridDim.x) {
        const int64_t t = m.tab[v], vn = v + gridDim.x;
        const BuildArgs& a = g.tables[t];
        auto next = [&]() {
            if (vn < m.blocks) load(vn);
        };
        if (a.st->error) {
            next();
            continue;
        }
        psa_section<T, true>(a, v - m.base[t], x, lds_raw, lane, next);
    }
}

// Levels b0 + 1 .. b0 + 9 of every table's dyadic block tree.
__global__ __launch_bounds__(256) void seg_tree(SegArgs g, SegMap m, int b0) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    const BuildArgs& a = g.tables[t];
    if (a.st->error) return;
    tree_block(a, blk, b0);
}

__global__ __launch_bounds__(256) void seg_place(SegArgs g, SegMap m) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    const BuildArgs& a = g.tables[t];
    if (a.st->error) return;
    const int64_t s = blk * 4 + wave_in_block();
    if (s >= a.G) return;
    place_section<false>(a, s, (int)(threadIdx.x & 63));
}

__global__ __launch_bounds__(256) void seg_degenerate(SegArgs g, SegMap m) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    // No latency tick here: every table has pack blocks, which run later.
    degenerate_block<true>(g.tables[t], blk, degenerate_blocks(g.tables[t].n));
}

__global__ __launch_bounds__(256) void seg_split(SegArgs g, SegMap m) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    split_bisect_block(g.tables[t], blk);
}

// The split certificate of every table of the group: one block per table.
__global__ __launch_bounds__(256) void seg_cert(SegArgs g, SegMap m) {
    const BuildArgs& a = g.tables[m.tab[blockIdx.x]];
    if (a.st->error) return;
    cert_reduce_block(a, 0, 1);
}

__global__ __launch_bounds__(256) void seg_knots(SegArgs g, SegMap m) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    const BuildArgs& a = g.tables[t];
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t mm = blk * blockDim.x + threadIdx.x;
    if (mm > min<int64_t>(a.learned_segments, st->p)) return;
    if (knot_split(a, mm)) atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
}

__global__ __launch_bounds__(256) void seg_learned(SegArgs g, SegMap m) {
    int64_t t, blk;
    seg_locate(m, t, blk);
    const BuildArgs& a = g.tables[t];
    DevState* st = a.st;
    if (st->error || st->degenerate) return;
    const int64_t k = blk * blockDim.x + threadIdx.x + 1;
    if (k >= st->p) return;
    const uint32_t f = learned_split(a, k);
    if (f & 1u) atomicAdd((unsigned long long*)&st->split_fallbacks, 1ull);
    if (f & 2u) atomicAdd((unsigned long long*)&st->learned_fallbacks, 1ull);
}

__global__ __launch_bounds__(64) void seg_pack(SegArgs g, SegMap m) {
    extern __shared__ double lds_raw[];
    int64_t t, blk;
    seg_locate(m, t, blk);
    tick_end(g, t);
    pack_staged_block<true>(g.tables[t], blk, lds_raw);
    tick_end(g, t);
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

enum { MAP_SUM, MAP_LOCAL, MAP_PLACE, MAP_DEG, MAP_SPLIT, MAP_KNOTS, MAP_PACK, MAP_TREE, MAP_ONE, MAP_COUNT };

// The launches of one group: a block->table map per phase (MAP_ONE: one block
// per table) and the group's most sections of a table (tree levels).
struct BatchGroup {
    SegMap maps[MAP_COUNT];
    int64_t G_max = 1;
    int64_t items = 0;
    int64_t ntables = 0;
};

struct atg_batch_s {
    int64_t ntables = 0, total = 0;
    atg_dtype dtype = ATG_F64;
    atg_build_opts opts{};
    int device = 0;
    std::vector<int64_t> offsets;
    std::vector<BuildArgs> args;        // host copy (pointers into the pools below)
    int64_t Ls_max = 1;
    int32_t pack_cap_max = 16;          // largest pack LDS pool of the tables (items)
    double wall_khz = 100000.0;
    std::vector<BatchGroup> groups;     // launched in order
    std::vector<int32_t> group_of;      // table -> group
    // device memory
    void* d_weights_owned = nullptr;
    const void* d_weights = nullptr;
    BuildArgs* d_args = nullptr;
    int64_t* d_offsets = nullptr;
    uint64_t* d_ticks = nullptr;
    BatchState* d_state = nullptr;
    Row* d_rows = nullptr;
    double* d_dpool = nullptr;          // section sums, tree, L/H weights and prefixes, staging
    uint64_t* d_slots = nullptr;
    int32_t* d_ipool = nullptr;         // L/H items, staging, knots
    Split* d_splits = nullptr;
    int32_t* d_maptab = nullptr;        // block->table maps of all phases and groups
    int64_t* d_mapbase = nullptr;       // first block of every table within its group's map, all phases
    bool weights_set = false, built = false;
    hipStream_t last_stream = nullptr;  // stream of the last build (the getters wait for it)
    // The build's launch sequence captured once and replayed (atg_batch_use_graph).
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

void drop_graph(atg_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    b->exec = nullptr;
    b->graph = nullptr;
    b->cap_stream = nullptr;
}

void free_batch(atg_batch* b) {
    drop_graph(b);
    void* ptrs[] = {b->d_weights_owned, b->d_args, b->d_offsets, b->d_ticks, b->d_state, b->d_rows,
                    b->d_dpool, b->d_slots, b->d_ipool, b->d_splits, b->d_maptab, b->d_mapbase};
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
    SegArgs g{b->d_args, b->ntables, b->d_ticks};
    auto grid = [](const SegMap& m) { return dim3((unsigned)m.blocks); };
    hipLaunchKernelGGL(seg_reset_ticks, dim3((unsigned)std::min<int64_t>((b->ntables + 255) / 256, 4096)), dim3(256),
                       0, s, g);
    const size_t lds = psa_lds_bytes(b->Ls_max);
    const int64_t resident = resident_blocks((const void*)seg_local<T>, 64, lds);
    const bool learned = b->opts.split_method == ATG_SPLIT_LEARNED;
    for (const BatchGroup& grp : b->groups) {
        const SegMap* M = grp.maps;
        // Every group holds at least one table, so each map below has blocks
        // (MAP_KNOTS and MAP_TREE are checked: a group of one-section tables
        // has no tree blocks).
        hipLaunchKernelGGL(seg_sum<T>, grid(M[MAP_SUM]), dim3(256), 0, s, g, M[MAP_SUM]);
        hipLaunchKernelGGL(seg_sum_finish, grid(M[MAP_ONE]), dim3(256), 0, s, g, M[MAP_ONE]);
        hipLaunchKernelGGL(seg_local<T>, dim3((unsigned)std::min<int64_t>(M[MAP_LOCAL].blocks, resident)), dim3(64),
                           lds, s, g, M[MAP_LOCAL]);
        if (M[MAP_TREE].blocks > 0)
            for (int b0 = 0; (grp.G_max >> b0) > 1; b0 += 9)
                hipLaunchKernelGGL(seg_tree, grid(M[MAP_TREE]), dim3(256), 0, s, g, M[MAP_TREE], b0);
        hipLaunchKernelGGL(seg_place, grid(M[MAP_PLACE]), dim3(256), 0, s, g, M[MAP_PLACE]);
        if (learned) hipLaunchKernelGGL(seg_cert, grid(M[MAP_ONE]), dim3(256), 0, s, g, M[MAP_ONE]);
        hipLaunchKernelGGL(seg_degenerate, grid(M[MAP_DEG]), dim3(256), 0, s, g, M[MAP_DEG]);
        if (learned) {
            if (M[MAP_KNOTS].blocks > 0)
                hipLaunchKernelGGL(seg_knots, grid(M[MAP_KNOTS]), dim3(256), 0, s, g, M[MAP_KNOTS]);
            hipLaunchKernelGGL(seg_learned, grid(M[MAP_SPLIT]), dim3(256), 0, s, g, M[MAP_SPLIT]);
        } else {
            hipLaunchKernelGGL(seg_split, grid(M[MAP_SPLIT]), dim3(256), 0, s, g, M[MAP_SPLIT]);
        }
        hipLaunchKernelGGL(seg_pack, grid(M[MAP_PACK]), dim3(64), pack_lds_bytes(b->pack_cap_max), s, g, M[MAP_PACK]);
    }
    return hipGetLastError();
}

// Deals the tables into groups (largest table first, each to the group with
// the fewest items so far, ties to the lower group; members keep their index
// order) and uploads the block->table maps of every group and phase.  Blocks
// per table as the single-table launchers use them.  Replaces any earlier
// grouping; the caller drops a captured graph.
atg_status make_groups(atg_batch* b, int64_t group_items) {
    const int64_t ntables = b->ntables;
    const std::vector<int64_t>& offsets = b->offsets;
    int64_t ngroups = 1;
    if (group_items > 0)
        ngroups = std::min<int64_t>(ntables, std::max<int64_t>(1, (b->total + group_items - 1) / group_items));
    std::vector<BatchGroup> groups((size_t)ngroups);
    std::vector<int32_t> group_of((size_t)ntables, 0);
    std::vector<std::vector<int64_t>> members((size_t)ngroups);
    {
        std::vector<int64_t> order((size_t)ntables);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
            return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
        });
        for (int64_t t : order) {
            int64_t g = 0;
            for (int64_t q = 1; q < ngroups; ++q)
                if (groups[q].items < groups[g].items) g = q;
            groups[g].items += offsets[t + 1] - offsets[t];
            groups[g].ntables += 1;
            group_of[t] = (int32_t)g;
        }
        for (int64_t t = 0; t < ntables; ++t) members[group_of[t]].push_back(t);
        for (int64_t g = 0; g < ngroups; ++g)
            for (int64_t t : members[g]) groups[g].G_max = std::max(groups[g].G_max, b->args[t].G);
    }
    auto blocks_of = [&](int ph, int64_t t) -> int64_t {
        const BuildArgs& a = b->args[t];
        const int64_t K = std::min<int64_t>(a.learned_segments, a.p_req);
        switch (ph) {
            case MAP_SUM: return (a.G + 3) / 4;
            case MAP_LOCAL: return a.G;
            case MAP_PLACE: return (a.G + 3) / 4;
            case MAP_TREE: return a.G > 1 ? (a.G + 511) / 512 : 0;  // one section: no tree
            case MAP_DEG: return degenerate_blocks(a.n);
            case MAP_SPLIT: return std::max<int64_t>(1, (a.p_req - 1 + 255) / 256);
            case MAP_KNOTS: return (K + 1 + 255) / 256;
            case MAP_ONE: return 1;
            default: return (a.p_req + 63) / 64;  // MAP_PACK
        }
    };
    // The maps of phase ph, group g follow each other (phase-major); base
    // holds each table's first block relative to its group's map.
    std::vector<int64_t> base((size_t)MAP_COUNT * ntables);
    std::vector<int32_t> tab;
    std::vector<int64_t> map_start((size_t)MAP_COUNT * ngroups);
    for (int ph = 0; ph < MAP_COUNT; ++ph)
        for (int64_t g = 0; g < ngroups; ++g) {
            map_start[(size_t)ph * ngroups + g] = (int64_t)tab.size();
            int64_t nblk = 0;
            for (int64_t t : members[g]) {
                base[(size_t)ph * ntables + t] = nblk;
                const int64_t nb = blocks_of(ph, t);
                tab.insert(tab.end(), (size_t)nb, (int32_t)t);
                nblk += nb;
            }
            if (nblk > 0x7FFFFFFF) return fail(ATG_EINVAL, "batch: %lld blocks in one launch", (long long)nblk);
            groups[g].maps[ph].blocks = nblk;
        }
    int32_t* d_tab = nullptr;
    int64_t* d_base = nullptr;
    if (dalloc(&d_tab, tab.size()) != hipSuccess || dalloc(&d_base, base.size()) != hipSuccess) {
        (void)hipGetLastError();
        if (d_tab) (void)hipFree(d_tab);
        return fail(ATG_ENOMEM, "batch: block maps");
    }
    if (hipMemcpy(d_tab, tab.data(), sizeof(int32_t) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_base, base.data(), sizeof(int64_t) * base.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d_tab);
        (void)hipFree(d_base);
        return fail(ATG_EDEVICE, "batch: block map upload");
    }
    for (int ph = 0; ph < MAP_COUNT; ++ph)
        for (int64_t g = 0; g < ngroups; ++g) {
            SegMap& m = groups[g].maps[ph];
            m.tab = d_tab + map_start[(size_t)ph * ngroups + g];
            m.base = d_base + (size_t)ph * ntables;
        }
    // The old maps may still be read by enqueued launches.
    if (b->d_maptab || b->d_mapbase) (void)hipDeviceSynchronize();
    if (b->d_maptab) (void)hipFree(b->d_maptab);
    if (b->d_mapbase) (void)hipFree(b->d_mapbase);
    b->d_maptab = d_tab;
    b->d_mapbase = d_base;
    b->groups.swap(groups);
    b->group_of.swap(group_of);
    b->opts.batch_group_items = group_items > 0 ? group_items : 0;
    return ATG_OK;
}

// Switches every table between bisection and the learned split (the same
// splits either way; the learned split also needs the certificate pass) and
// uploads the tables' arguments when the weights are bound.
atg_status batch_set_split(atg_batch* b, int32_t split) {
    b->opts.split_method = split;
    for (BuildArgs& a : b->args) a.want_cert = split == ATG_SPLIT_LEARNED;
    if (b->weights_set) {
        HIPCHK(hipStreamSynchronize(b->last_stream));
        HIPCHK(hipMemcpy(b->d_args, b->args.data(), sizeof(BuildArgs) * b->ntables, hipMemcpyHostToDevice));
    }
    return ATG_OK;
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
    if (o.split_method == ATG_SPLIT_PARY)
        return fail(ATG_ENOTSUP, "atg_batch_create: the P-ary split is built for single tables; use bisection or the learned split");
    if (o.split_method != ATG_SPLIT_BISECT && o.split_method != ATG_SPLIT_LEARNED)
        return fail(ATG_EINVAL, "atg_batch_create: unknown split method %d", o.split_method);
    if (o.pack_method != ATG_PACK_CHUNKED)
        return fail(ATG_ENOTSUP, "atg_batch_create: the batch packs with the staged (chunked) pack only");
    if (o.learned_segments <= 0) o.learned_segments = 4096;
    const int64_t p_auto = 32LL * device_cus() * 10;

    atg_batch* b = new atg_batch_s();
    b->ntables = ntables;
    b->total = offsets[ntables];
    b->dtype = dtype;
    b->opts = o;
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
        const int64_t nsplits = o.num_splits > 0 ? o.num_splits
                                : o.num_splits < 0 ? std::max<int64_t>(2, n / -o.num_splits) : p_auto;
        z.p_req = std::max<int64_t>(2, std::min<int64_t>(nsplits, n / 10));
        z.half = (z.G + 63) / 64 + 1;
        z.tree_total = block_tree_total(z.G);
        od[t] = nd;
        nd += align_up(z.G, 32) + align_up(2 * z.half, 32) + 5 * align_up(n + 2, 32);
        os[t] = ns;
        ns += align_up(3 * z.tree_total + (z.tree_total + 1) / 2 + 1, 32) + align_up(6 * z.G, 32);
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
    BALLOC(b->d_offsets, ntables + 1);
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
        a.sec_cert = a.slots + align_up(3 * z.tree_total + (z.tree_total + 1) / 2 + 1, 32);
        a.want_cert = o.split_method == ATG_SPLIT_LEARNED;
        int32_t* ip = b->d_ipool + oi[t];
        const int64_t icap = align_up(n + 2, 64);
        a.Li = ip; a.Hi = ip + icap; a.stage_i = ip + 2 * icap; a.knot_j = ip + 3 * icap;
        a.rows = b->d_rows + offsets[t];
        a.tw = nullptr;
        a.ta = nullptr;
        a.psa_plus = o.psa_plus;
        a.chains = o.chains;
        a.p_req = z.p_req;
        a.pack_cap = pack_cap_items(n, z.p_req);
        b->pack_cap_max = std::max(b->pack_cap_max, a.pack_cap);
        a.splits = b->d_splits + osp[t];
        a.learned_segments = learned_knots(o.learned_segments, z.p_req);
    }
    {
        const atg_status rc = make_groups(b, o.batch_group_items);
        if (rc != ATG_OK) return cleanup(rc);
    }
    if (hipMemcpy(b->d_offsets, offsets, sizeof(int64_t) * (ntables + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_state, 0, sizeof(BatchState) * ntables) != hipSuccess ||
        hipMemset(b->d_ticks, 0, sizeof(uint64_t) * 2 * ntables) != hipSuccess)
        return cleanup(fail(ATG_EDEVICE, "atg_batch_create: upload failed"));
    const size_t lds = psa_lds_bytes(b->Ls_max);
    const void* fn = dtype == ATG_F32 ? (const void*)seg_local<float> : (const void*)seg_local<double>;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return cleanup(fail(ATG_EDEVICE, "atg_batch_create: LDS limit of %zu bytes", lds));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess && khz > 0)
        b->wall_khz = khz;
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

// One launch per phase for all tables, on the caller's stream.
atg_status atg_batch_build(atg_batch* b, void* stream) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_build: NULL batch");
    if (!b->weights_set) return fail(ATG_ESTATE, "atg_batch_build: weights not set");
    hipStream_t s = (hipStream_t)stream;
    b->last_stream = s;
    if (b->exec) HIPCHK(hipGraphLaunch(b->exec, s));
    else HIPCHK(b->dtype == ATG_F32 ? launch_batch<float>(b, s) : launch_batch<double>(b, s));
    b->built = true;
    return ATG_OK;
}

// Captures the build's launch sequence (kernels only: no copies or memsets)
// on a private stream into a graph that atg_batch_build replays from then on;
// enable = 0 drops the graph.  The kernels read every table's arguments from
// device memory, which atg_batch_set_weights rewrites, so the graph stays
// valid across new weights.
atg_status atg_batch_use_graph(atg_batch* b, int32_t enable) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_use_graph: NULL batch");
    drop_graph(b);
    if (!enable) return ATG_OK;
    HIPCHK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    const hipError_t e = b->dtype == ATG_F32 ? launch_batch<float>(b, b->cap_stream) : launch_batch<double>(b, b->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->cap_stream, &g);
    b->graph = g;
    if (e != hipSuccess || e2 != hipSuccess) {
        drop_graph(b);
        return fail(ATG_EDEVICE, "atg_batch_use_graph: capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    const hipError_t e3 = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
        drop_graph(b);
        return fail(ATG_EDEVICE, "atg_batch_use_graph: instantiate failed: %s", hipGetErrorString(e3));
    }
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
    HIPCHK(hipStreamSynchronize(b->last_stream));
    const size_t n = (size_t)b->total;
    if (weights || aliases) {
        // In chunks of 4 Mi rows (64 MiB): the rows of a large batch do not
        // need a host copy of their own.
        const size_t chunk = (size_t)1 << 22;
        std::vector<Row> rows(std::min(n, chunk));
        for (size_t c0 = 0; c0 < n; c0 += chunk) {
            const size_t m = std::min(chunk, n - c0);
            HIPCHK(hipMemcpy(rows.data(), b->d_rows + c0, m * sizeof(Row), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < m; ++i) {
                if (weights) weights[c0 + i] = rows[i].weight;
                if (aliases) aliases[c0 + i] = rows[i].alias;
            }
        }
    }
    if (W_per_table) {
        std::vector<BatchState> st((size_t)b->ntables);
        HIPCHK(hipMemcpy(st.data(), b->d_state, sizeof(BatchState) * b->ntables, hipMemcpyDeviceToHost));
        for (int64_t t = 0; t < b->ntables; ++t) W_per_table[t] = st[t].s.W;
    }
    return ATG_OK;
}

atg_status atg_batch_get_table(atg_batch* b, int64_t t, double* weights, int32_t* aliases, double* W) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_get_table: NULL batch");
    if (t < 0 || t >= b->ntables) return fail(ATG_EINVAL, "atg_batch_get_table: table %lld of %lld", (long long)t, (long long)b->ntables);
    if (!b->built) return fail(ATG_ESTATE, "atg_batch_get_table: not built");
    HIPCHK(hipStreamSynchronize(b->last_stream));
    const int64_t off = b->offsets[t], n = b->offsets[t + 1] - off;
    if (weights || aliases) {
        std::vector<Row> rows((size_t)n);
        HIPCHK(hipMemcpy(rows.data(), b->d_rows + off, (size_t)n * sizeof(Row), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            if (weights) weights[i] = rows[i].weight;
            if (aliases) aliases[i] = rows[i].alias;
        }
    }
    if (W) {
        BatchState st;
        HIPCHK(hipMemcpy(&st, b->d_state + t, sizeof(BatchState), hipMemcpyDeviceToHost));
        *W = st.s.W;
    }
    return ATG_OK;
}

atg_status atg_batch_groups(atg_batch* b, int32_t* group_of_table, int64_t* ngroups) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_groups: NULL batch");
    if (ngroups) *ngroups = (int64_t)b->groups.size();
    if (group_of_table) std::copy(b->group_of.begin(), b->group_of.end(), group_of_table);
    return ATG_OK;
}

// Times the table-neutral knobs of a batch in place: the split method
// (bisection or the learned index; with tune_split) and the launch groups (one group, or
// groups of total/2^k items down to about 2^26 items).  Each candidate is the
// median makespan of `reps` builds after a warm-up build; among the
// candidates whose makespan is within (1 + makespan_slack) of the fastest, the
// one with the lowest 99th-percentile table latency wins.  The batch keeps the
// winning configuration (any captured graph is dropped).
atg_status atg_batch_autotune(atg_batch* b, int32_t reps, int32_t tune_split, double makespan_slack,
                              atg_build_opts* chosen, double* makespan_ms, double* p99_latency_ms) {
    if (!b) return fail(ATG_EINVAL, "atg_batch_autotune: NULL batch");
    if (!b->weights_set) return fail(ATG_ESTATE, "atg_batch_autotune: weights not set");
    if (reps < 1) reps = 3;
    if (!(makespan_slack >= 0.0)) makespan_slack = 0.0;
    drop_graph(b);
    std::vector<int64_t> gsizes = {0};
    for (int64_t gi = b->total / 2; gi >= ((int64_t)1 << 26) && (int64_t)gsizes.size() < 8; gi /= 2)
        gsizes.push_back(gi);
    struct Res { int32_t split; int64_t group_items; double ms, p99; };
    std::vector<Res> res;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    atg_status rc = ATG_OK;
    std::vector<double> lat((size_t)b->ntables);
    std::vector<int32_t> splits = {b->opts.split_method};
    if (tune_split) splits = {ATG_SPLIT_BISECT, ATG_SPLIT_LEARNED};
    for (int32_t split : splits) {
        if (rc != ATG_OK) break;
        rc = batch_set_split(b, split);
        for (int64_t gi : gsizes) {
            if (rc != ATG_OK) break;
            rc = make_groups(b, gi);
            if (rc == ATG_OK) rc = atg_batch_build(b, nullptr);
            if (rc == ATG_OK) rc = atg_batch_sync(b, nullptr);
            std::vector<float> ts;
            std::vector<double> p99s;
            for (int r = 0; r < reps && rc == ATG_OK; ++r) {
                (void)hipEventRecord(e0, nullptr);
                rc = atg_batch_build(b, nullptr);
                (void)hipEventRecord(e1, nullptr);
                if (rc == ATG_OK) rc = atg_batch_sync(b, nullptr);
                float ms = 0.f;
                if (rc == ATG_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(ATG_EDEVICE, "events");
                if (rc == ATG_OK) rc = atg_batch_latencies(b, lat.data());
                if (rc != ATG_OK) break;
                ts.push_back(ms);
                std::vector<double> l = lat;
                const size_t k = (size_t)std::min<double>((double)(l.size() - 1), std::ceil(0.99 * (double)l.size()) - 1);
                std::nth_element(l.begin(), l.begin() + (long)k, l.end());
                p99s.push_back(l[k] * 1e-6);
            }
            if (rc != ATG_OK) break;
            std::sort(ts.begin(), ts.end());
            std::sort(p99s.begin(), p99s.end());
            res.push_back(Res{split, gi, ts[ts.size() / 2], p99s[p99s.size() / 2]});
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != ATG_OK) return rc;
    double fastest = 1e300;
    for (const Res& r : res) fastest = std::min(fastest, r.ms);
    const Res* win = nullptr;
    for (const Res& r : res)
        if (r.ms <= fastest * (1.0 + makespan_slack) && (!win || r.p99 < win->p99)) win = &r;
    rc = batch_set_split(b, win->split);
    if (rc == ATG_OK) rc = make_groups(b, win->group_items);
    if (rc != ATG_OK) return rc;
    if (chosen) *chosen = b->opts;
    if (makespan_ms) *makespan_ms = win->ms;
    if (p99_latency_ms) *p99_latency_ms = win->p99;
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
    HIPCHK(hipStreamSynchronize(b->last_stream));
    std::vector<uint64_t> tk((size_t)(2 * b->ntables));
    HIPCHK(hipMemcpy(tk.data(), b->d_ticks, sizeof(uint64_t) * 2 * b->ntables, hipMemcpyDeviceToHost));
    const double ns_per_tick = 1e6 / b->wall_khz;
    for (int64_t t = 0; t < b->ntables; ++t) ns_per_table[t] = (double)(tk[2 * t + 1] - tk[2 * t]) * ns_per_tick;
    return ATG_OK;
}

}  // extern "C"
