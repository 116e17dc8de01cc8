// C ABI of the engine (include/atg.h): handle lifetime, device memory, the
// build/sample orchestration on the caller's stream, copy-out and the table
// check of postBuild/verifyTableCorrectness.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/atg.h"
#include "atg_build.h"
#include "atg_host.h"
#include "atg_sample.h"

using namespace atg;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace atg {

atg_status fail(atg_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 256;
    return cus;
}

atg_status state_status(const DevState& st) {
    if (st.error & ERR_BAD_WEIGHT) return fail(ATG_EINVAL, "weights contain NaN, infinity or negative values");
    if (st.error & ERR_BAD_SUM) return fail(ATG_EINVAL, "sum of weights is not positive and finite");
    if (st.error & ERR_SECTION_TOO_LONG) return fail(ATG_EINVAL, "section too long for LDS");
    if (st.error & ERR_LOOKBACK_TIMEOUT) return fail(ATG_ETIMEOUT, "look-back wait expired");
    return ATG_OK;
}

}  // namespace atg

struct atg_table_s {
    int64_t n = 0;
    atg_dtype dtype = ATG_F64;
    atg_layout layout = ATG_AOS;
    atg_build_opts opts{};
    int64_t G = 0, Ls = 0, tree_total = 0, p_req = 0, tree_half = 0;
    int64_t p_cap = 0;  // capacity of the splits array minus one
    int32_t knots_cap = 0;  // learned-index knots the knot array holds
    int device = 0;
    // device memory
    void* d_weights_owned = nullptr;
    const void* d_weights = nullptr;
    DevState* d_state = nullptr;
    double* d_section_sums = nullptr;
    double* d_tree = nullptr;
    uint64_t* d_slots = nullptr;
    uint64_t* d_sec_cert = nullptr;
    uint32_t* d_sec_loff = nullptr;  // lean light lists: lights before each section
    uint32_t* d_sec_ldir = nullptr;  // lean light lists: the section of every 1024th light
    bool placed = false;             // the staging and sec_loff of the last partition are current
    double* d_stage_w = nullptr;
    int32_t* d_stage_i = nullptr;
    int32_t *d_Li = nullptr, *d_Hi = nullptr;
    double *d_Lw = nullptr, *d_Lp = nullptr, *d_Hw = nullptr, *d_Hp = nullptr;
    Row* d_rows = nullptr;
    double* d_tw = nullptr;
    int32_t* d_ta = nullptr;
    Split* d_splits = nullptr;
    int32_t* d_knots = nullptr;
    unsigned long long* d_checksum = nullptr;
    // rejection: its own W and |A| (the alias build's state stays untouched)
    DevState* d_rej_state = nullptr;
    RejEntry* d_A = nullptr;
    uint64_t A_cap = 0;
    uint64_t *d_rej_tot = nullptr, *d_rej_off = nullptr;
    RejTask* d_rej_tasks = nullptr;
    // search fills: the inclusive counts; the learned fill: its model for
    // rej_knots segments (buffers for rej_knots_cap)
    uint32_t* d_rej_incl = nullptr;
    int32_t* d_rej_knot_items = nullptr;
    uint32_t* d_rej_knot_err = nullptr;
    int32_t rej_knots_cap = 0, rej_knots = 0;
    int32_t rej_fill = ATG_FILL_DIRECT;
    // sectioned rejection sampling: |A| read back once per rejection build;
    // section masses, their prefix and the section counts for rej_sec_cap
    // sections, the masses valid for rej_sec_SS slots per section
    uint64_t rej_size_host = 0;
    bool rej_size_known = false;
    double *d_rej_mass = nullptr, *d_rej_pre = nullptr;
    uint64_t* d_rej_sec_counts = nullptr;
    int64_t rej_sec_cap = 0, rej_sec_SS = 0;
    bool rej_sec_valid = false;
    // sectioned sampling: count and first output of each section
    uint64_t* d_sec_counts = nullptr;
    int64_t sec_counts_cap = 0;
    bool rej_built = false;
    // host side
    std::vector<double> h_weights;  // host copy when weights were given on the host
    bool weights_set = false, built = false;
    // Stream of the last work enqueued on this table; the host-side getters
    // wait for it (not for the whole device) before reading device state.
    hipStream_t last_stream = nullptr;
    hipEvent_t ev[6] = {};
    bool events = false;
    // Split/pack interleave: the splits run in chunks on sp_stream, each pack
    // chunk waits for its splits.
    hipStream_t sp_stream = nullptr;
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_chunk[kInterleaveChunks] = {};

    BuildArgs args() const {
        BuildArgs a{};
        a.weights = d_weights;
        a.dtype_f32 = dtype == ATG_F32;
        a.n = n;
        a.G = G;
        a.Ls = Ls;
        a.Ls_max = Ls;
        a.tree_total = tree_total;
        a.st = d_state;
        a.section_sums = d_section_sums;
        a.tree_scratch = d_tree;
        a.tree_scratch_half = tree_half;
        a.slots = d_slots;
        a.sec_cert = d_sec_cert;
        a.want_cert = opts.split_method != ATG_SPLIT_BISECT;  // the learned and P-ary searches need it
        a.stage_w = d_stage_w;
        a.stage_i = d_stage_i;
        a.Li = d_Li; a.Lw = d_Lw; a.Lp = d_Lp;
        a.Hi = d_Hi; a.Hw = d_Hw; a.Hp = d_Hp;
        a.rows = d_rows;
        a.tw = d_tw;
        a.ta = d_ta;
        a.psa_plus = opts.psa_plus;
        a.chains = opts.chains;
        a.p_req = p_req;
        a.splits = d_splits;
        a.pack_cap = pack_cap_items(n, p_req);
        a.learned_segments = learned_knots(opts.learned_segments, p_req);
        a.knot_j = d_knots;
        a.pack_method = opts.pack_method;
        a.sec_loff = d_sec_loff;
        a.sec_ldir = d_sec_ldir;
        a.lean_lights = d_sec_loff && d_sec_ldir && lean_lights_ok(n, p_req, opts.pack_method) ? 1 : 0;
        return a;
    }
};

static_assert((int)PACK_CHUNKED == (int)ATG_PACK_CHUNKED && (int)PACK_BASIC == (int)ATG_PACK_BASIC &&
                  (int)PACK_NO_WEIGHTS == (int)ATG_PACK_NO_WEIGHTS,
              "pack method values");

extern "C" {

void atg_default_opts(atg_build_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->psa_plus = 1;
    o->chains = 16;         // NUM_LOCAL_SPLITS (PsaPlusNormal.cu:3)
    o->section_size = 0;    // min(1000, N/10) (AliasTableSplitGpu.cu:356)
    o->num_splits = 0;      // AUTO (AliasTableSplitGpu.cu:58-70)
    o->split_method = ATG_SPLIT_BISECT;
    o->learned_segments = 4096;
    o->record_phases = 0;
    o->interleave = 0;
}

const char* atg_last_error(void) { return g_last_error.c_str(); }
const char* atg_version(void) { return "atg 0.1 (gfx950)"; }
int32_t atg_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

static void free_table(atg_table* t) {
    if (!t) return;
    void* ptrs[] = {t->d_weights_owned, t->d_state, t->d_section_sums, t->d_tree, t->d_slots, t->d_sec_cert, t->d_sec_loff, t->d_sec_ldir, t->d_stage_w,
                    t->d_stage_i, t->d_Li, t->d_Hi,
                    t->d_Lw, t->d_Lp, t->d_Hw, t->d_Hp, t->d_rows, t->d_tw, t->d_ta, t->d_splits, t->d_knots,
                    t->d_checksum, t->d_rej_state, t->d_A, t->d_rej_tot, t->d_rej_off, t->d_rej_tasks,
                    t->d_sec_counts, t->d_rej_mass, t->d_rej_pre, t->d_rej_sec_counts, t->d_rej_incl,
                    t->d_rej_knot_items, t->d_rej_knot_err};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (t->events)
        for (auto& e : t->ev) (void)hipEventDestroy(e);
    if (t->sp_stream) {
        (void)hipStreamSynchronize(t->sp_stream);
        (void)hipStreamDestroy(t->sp_stream);
        (void)hipEventDestroy(t->ev_fork);
        for (auto& e : t->ev_chunk) (void)hipEventDestroy(e);
    }
}

void atg_destroy(atg_table* t) {
    free_table(t);
    delete t;
}

atg_status atg_create(atg_table** out, int64_t n, atg_dtype dtype, atg_layout layout, const atg_build_opts* opts) {
    if (!out) return fail(ATG_EINVAL, "atg_create: out is NULL");
    *out = nullptr;
    if (n < 1 || n > kMaxN) return fail(ATG_EINVAL, "atg_create: n=%lld outside [1, %lld]", (long long)n, (long long)kMaxN);
    if (dtype != ATG_F64 && dtype != ATG_F32) return fail(ATG_EINVAL, "atg_create: bad dtype");
    if (layout != ATG_AOS && layout != ATG_SOA) return fail(ATG_EINVAL, "atg_create: bad layout");
    atg_build_opts o;
    if (opts) o = *opts; else atg_default_opts(&o);
    if (o.chains < 1 || o.chains > 64) return fail(ATG_EINVAL, "atg_create: chains=%d outside [1, 64]", o.chains);
    if (o.split_method != ATG_SPLIT_BISECT && o.split_method != ATG_SPLIT_LEARNED && o.split_method != ATG_SPLIT_PARY)
        return fail(ATG_EINVAL, "atg_create: unknown split method %d", o.split_method);
    if (o.pack_method != ATG_PACK_CHUNKED && o.pack_method != ATG_PACK_BASIC && o.pack_method != ATG_PACK_NO_WEIGHTS)
        return fail(ATG_EINVAL, "atg_create: unknown pack method %d", o.pack_method);
    if (o.pack_method == ATG_PACK_NO_WEIGHTS && o.psa_plus)
        return fail(ATG_EINVAL, "atg_create: the NO_WEIGHTS pack needs psa_plus = 0 (PSA+ leaves heavy residues "
                                "that differ from the input weights)");
    if (o.learned_segments <= 0) o.learned_segments = 4096;
    if (o.section_size <= 0) o.section_size = std::min<int64_t>(1000, n / 10);
    if (o.section_size < 1) o.section_size = 1;
    if (o.num_splits < 0) o.num_splits = std::max<int64_t>(2, n / -o.num_splits);  // one split per |num_splits| items
    if (o.num_splits == 0) o.num_splits = 32LL * device_cus() * 10;

    atg_table* t = new atg_table_s();
    t->n = n;
    t->dtype = dtype;
    t->layout = layout;
    t->opts = o;
    section_layout(n, o.section_size, &t->G, &t->Ls);
    if (t->Ls > kMaxSectionLen) {
        delete t;
        return fail(ATG_EINVAL, "atg_create: section length %lld exceeds %lld", (long long)t->Ls, (long long)kMaxSectionLen);
    }
    t->p_req = std::max<int64_t>(2, std::min<int64_t>(o.num_splits, n / 10));
    t->tree_half = (t->G + 63) / 64 + 1;
    t->tree_total = block_tree_total(t->G);
    (void)hipGetDevice(&t->device);

    auto cleanup = [&](atg_status st) { atg_destroy(t); return st; };
#define ALLOC(ptr, cnt)                                                                          \
    do {                                                                                         \
        if (dalloc(&(ptr), (size_t)(cnt)) != hipSuccess) {                                       \
            (void)hipGetLastError();                                                             \
            return cleanup(fail(ATG_ENOMEM, "atg_create: hipMalloc of %s (%lld items) failed", #ptr, (long long)(cnt))); \
        }                                                                                        \
    } while (0)
    ALLOC(t->d_state, 1);
    ALLOC(t->d_section_sums, t->G);
    ALLOC(t->d_tree, 2 * t->tree_half);
    ALLOC(t->d_slots, 3 * t->tree_total + (t->tree_total + 1) / 2 + 1);
    ALLOC(t->d_Li, n + 2);
    ALLOC(t->d_Hi, n + 2);
    ALLOC(t->d_Lw, n + 2);
    ALLOC(t->d_Lp, n + 2);
    ALLOC(t->d_Hw, n + 2);
    ALLOC(t->d_Hp, n + 2);
    ALLOC(t->d_stage_w, n);
    ALLOC(t->d_sec_cert, 6 * t->G);
    ALLOC(t->d_sec_loff, t->G + 1);
    ALLOC(t->d_sec_ldir, light_dir_size(n));
    ALLOC(t->d_stage_i, n);
    if (layout == ATG_AOS) {
        ALLOC(t->d_rows, n);
    } else {
        ALLOC(t->d_tw, n);
        ALLOC(t->d_ta, n);
    }
    t->p_cap = t->p_req;
    ALLOC(t->d_splits, t->p_req + 1);
    ALLOC(t->d_knots, o.learned_segments + 2);
    t->knots_cap = o.learned_segments;
    ALLOC(t->d_checksum, 1);
    // Section counts of the sectioned sampler at its default 4096-row sections
    // (smaller sections grow the buffer on first use).
    t->sec_counts_cap = (n + 4095) / 4096;
    ALLOC(t->d_sec_counts, 2 * t->sec_counts_cap);
#undef ALLOC
    if (hipMemset(t->d_state, 0, sizeof(DevState)) != hipSuccess) return cleanup(fail(ATG_EDEVICE, "memset"));
    if (set_partition_lds_limit(psa_lds_bytes(t->Ls)) != hipSuccess)
        return cleanup(fail(ATG_EDEVICE, "partition LDS limit"));
    if (o.record_phases) {
        for (auto& e : t->ev)
            if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(ATG_EDEVICE, "hipEventCreate"));
        t->events = true;
    }
    if (o.interleave) {
        if (hipStreamCreateWithFlags(&t->sp_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&t->ev_fork, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(ATG_EDEVICE, "atg_create: split stream"));
        for (auto& e : t->ev_chunk)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return cleanup(fail(ATG_EDEVICE, "atg_create: split events"));
    }
    *out = t;
    return ATG_OK;
}

atg_status atg_set_weights(atg_table* t, const void* w, atg_mem mem, void* stream) {
    if (!t || !w) return fail(ATG_EINVAL, "atg_set_weights: NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)t->n * (t->dtype == ATG_F32 ? 4 : 8);
    if (mem == ATG_MEM_DEVICE) {
        t->d_weights = w;
        t->h_weights.clear();
    } else if (mem == ATG_MEM_HOST) {
        if (!t->d_weights_owned) HIPCHK(hipMalloc(&t->d_weights_owned, bytes));
        HIPCHK(hipMemcpyAsync(t->d_weights_owned, w, bytes, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        t->d_weights = t->d_weights_owned;
        t->h_weights.resize(t->n);
        if (t->dtype == ATG_F32) {
            const float* f = (const float*)w;
            for (int64_t i = 0; i < t->n; ++i) t->h_weights[i] = (double)f[i];
        } else {
            std::memcpy(t->h_weights.data(), w, bytes);
        }
    } else {
        return fail(ATG_EINVAL, "atg_set_weights: bad mem kind");
    }
    t->weights_set = true;
    t->built = false;
    t->placed = false;
    t->rej_built = false;
    t->last_stream = s;
    return ATG_OK;
}

atg_status atg_build_phase(atg_table* t, int32_t phase, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_build_phase: NULL table");
    if (!t->weights_set) return fail(ATG_ESTATE, "atg_build_phase: weights not set");
    hipStream_t s = (hipStream_t)stream;
    t->last_stream = s;
    const BuildArgs a = t->args();
    switch (phase) {
        case ATG_PHASE_SUM: HIPCHK(launch_sum(a, s)); break;
        case ATG_PHASE_PARTITION:
            HIPCHK(launch_partition(a, s));
            t->placed = true;
            break;
        case ATG_PHASE_PARTITION_LOCAL:
            t->placed = false;
            HIPCHK(launch_partition_local(a, s));
            break;
        case ATG_PHASE_PARTITION_PLACE:
            HIPCHK(launch_partition_place(a, s));
            t->placed = true;
            break;
        case ATG_PHASE_SPLIT:
            if (t->opts.split_method == ATG_SPLIT_LEARNED) HIPCHK(launch_split_learned(a, s));
            else if (t->opts.split_method == ATG_SPLIT_PARY) HIPCHK(launch_split_pary(a, s));
            else HIPCHK(launch_split(a, s));
            break;
        case ATG_PHASE_PACK:
            HIPCHK(launch_pack(a, s));
            t->built = true;
            break;
        default: return fail(ATG_EINVAL, "atg_build_phase: unknown phase %d", phase);
    }
    return ATG_OK;
}

// Split and pack interleaved (AliasTableSplitGpu.cu:220-262): the splits run
// in kInterleaveChunks chunks of whole 256-split blocks on the table's own
// stream while the caller's stream packs the chains of each finished chunk
// (pack chunk c needs splits up to 256 * its last split block).  Returns
// false when p is too small to chunk.
static bool build_split_pack_interleaved(atg_table* t, const BuildArgs& a, hipStream_t s, hipError_t& err) {
    const int64_t SB = split_blocks(a), PB = pack_blocks(a);
    const int C = (int)std::min<int64_t>(kInterleaveChunks, SB);
    if (C < 2) return false;
    const bool learned = t->opts.split_method == ATG_SPLIT_LEARNED, pary = t->opts.split_method == ATG_SPLIT_PARY;
    hipStream_t sp = t->sp_stream;
    err = hipEventRecord(t->ev_fork, s);
    if (err == hipSuccess) err = hipStreamWaitEvent(sp, t->ev_fork, 0);
    if (err == hipSuccess) err = launch_degenerate(a, sp);
    if (err == hipSuccess && learned) err = launch_split_knots(a, sp);
    for (int c = 0; c < C && err == hipSuccess; ++c) {
        const int64_t b0 = SB * c / C, b1 = SB * (c + 1) / C;
        err = learned ? launch_split_learned_blocks(a, sp, b0, b1)
              : pary  ? launch_split_pary_blocks(a, sp, b0, b1)
                      : launch_split_blocks(a, sp, b0, b1);
        if (err == hipSuccess) err = hipEventRecord(t->ev_chunk[c], sp);
    }
    for (int c = 0; c < C && err == hipSuccess; ++c) {
        const int64_t b0 = 4 * (SB * c / C);
        const int64_t b1 = c == C - 1 ? PB : std::min<int64_t>(PB, 4 * (SB * (c + 1) / C));
        err = hipStreamWaitEvent(s, t->ev_chunk[c], 0);
        if (err == hipSuccess) err = launch_pack_blocks(a, s, std::min(b0, PB), b1);
    }
    return true;
}

atg_status atg_build(atg_table* t, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_build: NULL table");
    if (!t->weights_set) return fail(ATG_ESTATE, "atg_build: weights not set");
    hipStream_t s = (hipStream_t)stream;
    t->last_stream = s;
    if (t->opts.interleave && t->sp_stream && !t->events) {
        const BuildArgs a = t->args();
        HIPCHK(launch_sum(a, s));
        HIPCHK(launch_partition(a, s));
        t->placed = true;
        hipError_t err = hipSuccess;
        if (build_split_pack_interleaved(t, a, s, err)) {
            HIPCHK(err);
            t->built = true;
            return ATG_OK;
        }
        atg_status rc = atg_build_phase(t, ATG_PHASE_SPLIT, stream);
        if (rc == ATG_OK) rc = atg_build_phase(t, ATG_PHASE_PACK, stream);
        return rc;
    }
    // Events follow TableSplitTimer's phases (utils/Timer.cu:62-77); the
    // prefix sums are fused into the partition, so that phase reads 0.
    static const int before[4] = {0, 1, 3, 4};  // event recorded before each phase
    for (int ph = 0; ph < 4; ++ph) {
        if (t->events && ph == 2) HIPCHK(hipEventRecord(t->ev[2], s));
        if (t->events) HIPCHK(hipEventRecord(t->ev[before[ph]], s));
        atg_status rc = atg_build_phase(t, ph, stream);
        if (rc != ATG_OK) return rc;
    }
    if (t->events) HIPCHK(hipEventRecord(t->ev[5], s));
    return ATG_OK;
}

atg_status atg_sync(atg_table* t, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_sync: NULL table");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    DevState st;
    HIPCHK(hipMemcpy(&st, t->d_state, sizeof(st), hipMemcpyDeviceToHost));
    return state_status(st);
}

atg_status atg_get_stats(atg_table* t, atg_stats* out) {
    if (!t || !out) return fail(ATG_EINVAL, "atg_get_stats: NULL argument");
    HIPCHK(hipStreamSynchronize(t->last_stream));
    DevState st;
    HIPCHK(hipMemcpy(&st, t->d_state, sizeof(st), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->W = st.W;
    out->W_N = st.W_N;
    out->n = t->n;
    out->numL = st.numL;
    out->numH = st.numH;
    out->R = st.numL + st.numH - 2;
    out->p = st.p;
    out->num_sections = t->G;
    out->section_len = t->Ls;
    out->split_fallbacks = (int64_t)st.split_fallbacks;
    out->learned_fallbacks = (int64_t)st.learned_fallbacks;
    out->degenerate = (int32_t)st.degenerate;
    double mdl, mdh;
    std::memcpy(&mdl, &st.cert_max_dL, 8);
    std::memcpy(&mdh, &st.cert_min_dH, 8);
    out->monotone = st.cert_done ? (mdl < mdh ? 1 : 0) : -1;
    if (t->events && t->built) {
        float ms;
        const int map[5][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}};
        for (int k = 0; k < 5; ++k) {
            if (hipEventElapsedTime(&ms, t->ev[map[k][0]], t->ev[map[k][1]]) == hipSuccess) out->phase_ms[k] = ms;
        }
    }
    return state_status(st);
}

atg_status atg_get_table(atg_table* t, double* w, int32_t* al, atg_mem mem, void* stream) {
    if (!t || !w || !al) return fail(ATG_EINVAL, "atg_get_table: NULL argument");
    if (!t->built) return fail(ATG_ESTATE, "atg_get_table: not built");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)t->n;
    if (t->layout == ATG_SOA) {
        const hipMemcpyKind k = mem == ATG_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        HIPCHK(hipMemcpyAsync(w, t->d_tw, n * 8, k, s));
        HIPCHK(hipMemcpyAsync(al, t->d_ta, n * 4, k, s));
        HIPCHK(hipStreamSynchronize(s));
        return ATG_OK;
    }
    // AoS: the 16-byte rows are split by a kernel (device destination) or
    // copied whole and split on the host.
    if (mem == ATG_MEM_DEVICE) {
        HIPCHK(launch_split_rows(t->d_rows, w, al, t->n, s));
        HIPCHK(hipStreamSynchronize(s));
        return ATG_OK;
    }
    std::vector<Row> rows(n);
    HIPCHK(hipMemcpyAsync(rows.data(), t->d_rows, n * sizeof(Row), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) {
        w[i] = rows[i].weight;
        al[i] = rows[i].alias;
    }
    return ATG_OK;
}

atg_status atg_save_table(atg_table* t, const char* path, uint64_t seed) {
    if (!t || !path) return fail(ATG_EINVAL, "atg_save_table: NULL argument");
    atg_stats st{};
    atg_status r = atg_get_stats(t, &st);
    if (r != ATG_OK) return r;
    std::vector<double> w((size_t)t->n);
    std::vector<int32_t> al((size_t)t->n);
    r = atg_get_table(t, w.data(), al.data(), ATG_MEM_HOST, nullptr);
    if (r != ATG_OK) return r;
    atg_table_file_info info{};
    info.n = t->n;
    info.W = st.W;
    info.layout = t->layout;
    info.source_dtype = t->dtype;
    info.psa_plus = t->opts.psa_plus;
    info.chains = t->opts.chains;
    info.split_method = t->opts.split_method;
    info.section_size = t->opts.section_size;
    info.num_splits = st.p;
    info.seed = seed;
    return atg_table_file_write(path, &info, w.data(), al.data());
}

atg_status atg_set_table(atg_table* t, const double* w, const int32_t* al, double W, atg_mem mem, void* stream) {
    if (!t || !w || !al) return fail(ATG_EINVAL, "atg_set_table: NULL argument");
    if (!(W > 0.0) || !std::isfinite(W)) return fail(ATG_EINVAL, "atg_set_table: W=%g", W);
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)t->n;
    if (t->layout == ATG_SOA) {
        const hipMemcpyKind k = mem == ATG_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        HIPCHK(hipMemcpyAsync(t->d_tw, w, n * 8, k, s));
        HIPCHK(hipMemcpyAsync(t->d_ta, al, n * 4, k, s));
    } else if (mem == ATG_MEM_DEVICE) {
        HIPCHK(launch_join_rows(w, al, t->d_rows, t->n, s));
    } else {
        std::vector<Row> rows(n);
        for (size_t i = 0; i < n; ++i) rows[i] = Row{w[i], al[i], 0};
        HIPCHK(hipMemcpyAsync(t->d_rows, rows.data(), n * sizeof(Row), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    // Sampling reads W/N from the device state; the rest describes no build.
    DevState st{};
    st.W = W;
    st.W_N = W / (double)t->n;
    st.cert_max_dL = 0ull;
    st.cert_min_dH = 0x7FF0000000000000ull;
    HIPCHK(hipMemcpyAsync(t->d_state, &st, sizeof(DevState), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    t->built = true;
    t->placed = false;
    t->rej_built = false;
    t->last_stream = s;
    return ATG_OK;
}

atg_status atg_table_device_ptr(atg_table* t, void** a, void** b) {
    if (!t || !a) return fail(ATG_EINVAL, "atg_table_device_ptr: NULL argument");
    if (t->layout == ATG_AOS) {
        *a = t->d_rows;
        if (b) *b = nullptr;
    } else {
        *a = t->d_tw;
        if (b) *b = t->d_ta;
    }
    return ATG_OK;
}

atg_status atg_get_splits(atg_table* t, int32_t* i, int32_t* j, double* spill, int64_t cap) {
    if (!t || !i || !j || !spill) return fail(ATG_EINVAL, "atg_get_splits: NULL argument");
    DevState st;
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(&st, t->d_state, sizeof(st), hipMemcpyDeviceToHost));
    if (st.p + 1 > cap) return fail(ATG_EINVAL, "atg_get_splits: capacity %lld < p+1=%lld", (long long)cap, (long long)(st.p + 1));
    std::vector<Split> sp(st.p + 1);
    if (st.p > 0) HIPCHK(hipMemcpy(sp.data(), t->d_splits, sizeof(Split) * (st.p + 1), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k <= st.p && st.p > 0; ++k) {
        i[k] = sp[k].i;
        j[k] = sp[k].j;
        spill[k] = sp[k].spill;
    }
    return ATG_OK;
}

atg_status atg_get_lh(atg_table* t, int32_t* Li, double* Lw, double* Lp, int32_t* Hi, double* Hw, double* Hp) {
    if (!t) return fail(ATG_EINVAL, "atg_get_lh: NULL table");
    DevState st;
    const BuildArgs a = t->args();
    if (a.lean_lights && t->placed && (Li || Lw)) {
        // The light items and weights stay in the staging: write them out as lists first.
        HIPCHK(launch_lean_lights_fill(a, t->last_stream));
    }
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(&st, t->d_state, sizeof(st), hipMemcpyDeviceToHost));
    if (Li) HIPCHK(hipMemcpy(Li, t->d_Li, 4 * st.numL, hipMemcpyDeviceToHost));
    if (Lw) HIPCHK(hipMemcpy(Lw, t->d_Lw, 8 * st.numL, hipMemcpyDeviceToHost));
    if (Lp) HIPCHK(hipMemcpy(Lp, t->d_Lp, 8 * st.numL, hipMemcpyDeviceToHost));
    if (Hi) HIPCHK(hipMemcpy(Hi, t->d_Hi, 4 * st.numH, hipMemcpyDeviceToHost));
    if (Hw) HIPCHK(hipMemcpy(Hw, t->d_Hw, 8 * st.numH, hipMemcpyDeviceToHost));
    if (Hp) HIPCHK(hipMemcpy(Hp, t->d_Hp, 8 * st.numH, hipMemcpyDeviceToHost));
    return ATG_OK;
}

// postBuild + verifyTableCorrectness (AliasTable.cu:24-65).  The implied mass
// of item j is min(w_j, W/N) + sum over rows aliased to j of (W/N - row.w),
// accumulated in binary128; compared as probabilities with the weights.
atg_status atg_post_build(atg_table* t, int32_t verify, double* max_rel_err, int32_t* ok) {
    if (!t) return fail(ATG_EINVAL, "atg_post_build: NULL table");
    atg_status rc = atg_sync(t, nullptr);
    if (rc != ATG_OK) return rc;
    if (!verify) {
        if (ok) *ok = 1;
        return ATG_OK;
    }
    const int64_t n = t->n;
    std::vector<double> rw(n), wt;
    std::vector<int32_t> ra(n);
    rc = atg_get_table(t, rw.data(), ra.data(), ATG_MEM_HOST, nullptr);
    if (rc != ATG_OK) return rc;
    const double* w = t->h_weights.data();
    if (t->h_weights.empty()) {
        wt.resize(n);
        if (t->dtype == ATG_F32) {
            std::vector<float> f(n);
            HIPCHK(hipMemcpy(f.data(), t->d_weights, 4 * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) wt[i] = f[i];
        } else {
            HIPCHK(hipMemcpy(wt.data(), t->d_weights, 8 * n, hipMemcpyDeviceToHost));
        }
        w = wt.data();
    }
    DevState st;
    HIPCHK(hipMemcpy(&st, t->d_state, sizeof(st), hipMemcpyDeviceToHost));
    std::vector<__float128> m(n, 0);
    const double W_N = st.W_N;
    for (int64_t i = 0; i < n; ++i) {
        if (rw[i] < W_N) {
            m[i] += (__float128)rw[i];
            const int32_t a = ra[i];
            const __float128 rest = (__float128)W_N - (__float128)rw[i];
            if (a >= 0 && a < n) m[a] += rest; else m[i] += rest;
        } else {
            m[i] += (__float128)W_N;
        }
    }
    const __float128 total = (__float128)W_N * (__float128)n;
    double mrel = 0;
    int good = 1;
    const double maxError = std::max((double)n * 1e-8, 1e-3);
    for (int64_t i = 0; i < n; ++i) {
        const __float128 pi = m[i] / total, pt = (__float128)w[i] / (__float128)st.W;
        __float128 d = pi - pt;
        if (d < 0) d = -d;
        const double rel = w[i] > 0 ? (double)(d / pt) : (double)d;
        if (rel > mrel) mrel = rel;
        __float128 dm = m[i] - (__float128)w[i];
        if (dm < 0) dm = -dm;
        if ((double)dm > maxError || rw[i] < 0 || rw[i] > W_N * (1.0 + maxError)) good = 0;
    }
    if (max_rel_err) *max_rel_err = mrel;
    if (ok) *ok = good;
    return ATG_OK;
}

atg_status atg_sample(atg_table* t, atg_sample_kind kind, uint64_t count, uint64_t seed, int64_t section_rows,
                      int32_t* out, uint32_t* hist, uint64_t* checksum, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_sample: NULL table");
    if (!t->built) return fail(ATG_ESTATE, "atg_sample: table not built");
    hipStream_t s = (hipStream_t)stream;
    t->last_stream = s;
    SampleArgs a{};
    a.rows = t->d_rows;
    a.tw = t->d_tw;
    a.ta = t->d_ta;
    a.st = t->d_state;
    a.n = t->n;
    a.count = count;
    a.seed = seed;
    a.out = out;
    a.hist = hist;
    a.checksum = checksum ? t->d_checksum : nullptr;
    if (checksum) HIPCHK(hipMemsetAsync(t->d_checksum, 0, 8, s));
    if (count > 0) {
        if (kind == ATG_SAMPLE_BASIC) {
            HIPCHK(launch_sample_basic(a, s));
        } else if (kind == ATG_SAMPLE_SECTIONED) {
            if (section_rows <= 0) section_rows = 4096;
            if (section_rows > 8192) return fail(ATG_EINVAL, "atg_sample: section_rows %lld > 8192", (long long)section_rows);
            a.SS = section_rows;
            a.nsec = (t->n + section_rows - 1) / section_rows;
            if (a.nsec > t->sec_counts_cap) {
                // Grows only; freed with the table.
                HIPCHK(hipStreamSynchronize(s));
                if (t->d_sec_counts) HIPCHK(hipFree(t->d_sec_counts));
                t->d_sec_counts = nullptr;
                t->sec_counts_cap = 0;
                if (dalloc(&t->d_sec_counts, (size_t)(2 * a.nsec)) != hipSuccess) {
                    (void)hipGetLastError();
                    return fail(ATG_ENOMEM, "atg_sample: section counts");
                }
                t->sec_counts_cap = a.nsec;
            }
            a.sec_counts = t->d_sec_counts;
            HIPCHK(set_sample_lds_limit((int)(section_rows * sizeof(Row))));
            HIPCHK(launch_sample_sectioned(a, s));
        } else {
            return fail(ATG_EINVAL, "atg_sample: unknown kind %d", (int)kind);
        }
    }
    if (checksum) {
        unsigned long long v = 0;
        HIPCHK(hipMemcpyAsync(&v, t->d_checksum, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *checksum = v;
    }
    return ATG_OK;
}

// The n the search fills accept: |A| <= A_cap = 2n + 4096 stays below 2^32,
// so the inclusive counts fit 32 bits.
static constexpr int64_t kMaxSearchFillN = 0x7FFFF000LL;
static constexpr int32_t kDefaultFillKnots = 4096, kMaxFillKnots = 1 << 22;

static RejArgs rej_args(atg_table* t) {
    RejArgs r{};
    r.weights = t->d_weights;
    r.dtype_f32 = t->dtype == ATG_F32;
    r.n = t->n;
    r.st = t->d_rej_state;
    r.nb = rejection_blocks(t->n);
    r.block_tot = t->d_rej_tot;
    r.block_off = t->d_rej_off;
    r.incl = t->d_rej_incl;
    r.knot_items = t->d_rej_knot_items;
    r.knot_err = t->d_rej_knot_err;
    r.K = t->rej_knots;
    r.A = t->d_A;
    r.A_cap = t->A_cap;
    r.tasks = t->d_rej_tasks;
    r.tasks_cap = rejection_task_cap(t->A_cap);
    return r;
}

atg_status atg_rejection_build(atg_table* t, void* stream) {
    return atg_rejection_build_fill(t, ATG_FILL_DIRECT, 0, stream);
}

atg_status atg_rejection_build_fill(atg_table* t, atg_fill_method fill, int32_t knots, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_rejection_build: NULL table");
    if (!t->weights_set) return fail(ATG_ESTATE, "atg_rejection_build: weights not set");
    if (fill < ATG_FILL_DIRECT || fill > ATG_FILL_LEARNED)
        return fail(ATG_EINVAL, "atg_rejection_build_fill: unknown fill %d", (int)fill);
    if (knots == 0) knots = kDefaultFillKnots;
    if (fill == ATG_FILL_LEARNED && (knots < 1 || knots > kMaxFillKnots))
        return fail(ATG_EINVAL, "atg_rejection_build_fill: %d knots outside [1, %d]", knots, kMaxFillKnots);
    if (fill != ATG_FILL_DIRECT && t->n > kMaxSearchFillN)
        return fail(ATG_EINVAL, "atg_rejection_build_fill: n=%lld exceeds %lld for the search fills",
                    (long long)t->n, (long long)kMaxSearchFillN);
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = rejection_blocks(t->n);
    if (!t->d_A) {
        // All five buffers or none: a partial set would leave later calls
        // launching on null pointers.
        const uint64_t cap = 2 * (uint64_t)t->n + 4096;
        DevState* st = nullptr;
        RejEntry* A = nullptr;
        uint64_t *tot = nullptr, *off = nullptr;
        RejTask* tasks = nullptr;
        if (dalloc(&st, 1) != hipSuccess || dalloc(&A, cap) != hipSuccess || dalloc(&tot, nb) != hipSuccess ||
            dalloc(&off, nb) != hipSuccess || dalloc(&tasks, rejection_task_cap(cap)) != hipSuccess ||
            hipMemsetAsync(st, 0, sizeof(DevState), s) != hipSuccess) {
            (void)hipGetLastError();
            void* ptrs[] = {st, A, tot, off, tasks};
            for (void* q : ptrs)
                if (q) (void)hipFree(q);
            return fail(ATG_ENOMEM, "atg_rejection_build: device allocation of the rejection structure failed");
        }
        t->d_rej_state = st;
        t->d_A = A;
        t->A_cap = cap;
        t->d_rej_tot = tot;
        t->d_rej_off = off;
        t->d_rej_tasks = tasks;
    }
    if (fill != ATG_FILL_DIRECT && !t->d_rej_incl && dalloc(&t->d_rej_incl, (size_t)t->n) != hipSuccess) {
        (void)hipGetLastError();
        t->d_rej_incl = nullptr;
        return fail(ATG_ENOMEM, "atg_rejection_build_fill: inclusive counts");
    }
    if (fill == ATG_FILL_LEARNED && knots > t->rej_knots_cap) {
        // Both model arrays or neither.
        if (t->d_rej_knot_items) (void)hipFree(t->d_rej_knot_items);
        if (t->d_rej_knot_err) (void)hipFree(t->d_rej_knot_err);
        t->d_rej_knot_items = nullptr;
        t->d_rej_knot_err = nullptr;
        t->rej_knots_cap = 0;
        if (dalloc(&t->d_rej_knot_items, (size_t)knots + 1) != hipSuccess ||
            dalloc(&t->d_rej_knot_err, (size_t)knots) != hipSuccess) {
            (void)hipGetLastError();
            if (t->d_rej_knot_items) (void)hipFree(t->d_rej_knot_items);
            t->d_rej_knot_items = nullptr;
            t->d_rej_knot_err = nullptr;
            return fail(ATG_ENOMEM, "atg_rejection_build_fill: model of %d knots", knots);
        }
        t->rej_knots_cap = knots;
    }
    t->last_stream = s;
    t->rej_built = false;
    t->rej_fill = fill;
    t->rej_knots = fill == ATG_FILL_LEARNED ? knots : 0;
    // W: the canonical sum (RejectionSampling.cu:36 uses the host sum), into
    // the rejection structure's own state.
    BuildArgs ba = t->args();
    ba.st = t->d_rej_state;
    HIPCHK(launch_sum(ba, s));
    HIPCHK(launch_rejection_build(rej_args(t), fill, s));
    t->rej_built = true;
    t->rej_size_known = false;
    t->rej_sec_valid = false;
    return ATG_OK;
}

atg_status atg_rejection_fill_a(atg_table* t, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_rejection_fill_a: NULL table");
    if (!t->rej_built) return fail(ATG_ESTATE, "atg_rejection_fill_a: rejection structure not built");
    hipStream_t s = (hipStream_t)stream;
    t->last_stream = s;
    HIPCHK(launch_rejection_fill(rej_args(t), t->rej_fill, s));
    return ATG_OK;
}

atg_status atg_rejection_fill_fallbacks(atg_table* t, uint64_t* fallbacks) {
    if (!t || !fallbacks) return fail(ATG_EINVAL, "atg_rejection_fill_fallbacks: NULL argument");
    if (!t->rej_built) return fail(ATG_ESTATE, "atg_rejection_fill_fallbacks: rejection structure not built");
    DevState st;
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(&st, t->d_rej_state, sizeof(st), hipMemcpyDeviceToHost));
    *fallbacks = st.learned_fallbacks;
    return state_status(st);
}

static atg_status check_model_args(atg_table* t, const void* q, const void* e, int32_t knots, const char* who) {
    if (!t || !q || !e) return fail(ATG_EINVAL, "%s: NULL argument", who);
    if (!t->rej_built || t->rej_fill != ATG_FILL_LEARNED)
        return fail(ATG_ESTATE, "%s: no learned fill built", who);
    if (knots != t->rej_knots) return fail(ATG_EINVAL, "%s: %d knots, the model has %d", who, knots, t->rej_knots);
    return ATG_OK;
}

atg_status atg_rejection_get_model(atg_table* t, int32_t* knot_items, uint32_t* err_bounds, int32_t knots) {
    atg_status rc = check_model_args(t, knot_items, err_bounds, knots, "atg_rejection_get_model");
    if (rc != ATG_OK) return rc;
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(knot_items, t->d_rej_knot_items, sizeof(int32_t) * ((size_t)knots + 1), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(err_bounds, t->d_rej_knot_err, sizeof(uint32_t) * (size_t)knots, hipMemcpyDeviceToHost));
    return ATG_OK;
}

atg_status atg_rejection_set_model(atg_table* t, const int32_t* knot_items, const uint32_t* err_bounds,
                                   int32_t knots) {
    atg_status rc = check_model_args(t, knot_items, err_bounds, knots, "atg_rejection_set_model");
    if (rc != ATG_OK) return rc;
    // The fill reads incl only between knot items: they must be positions,
    // in order.
    for (int32_t k = 0; k <= knots; ++k) {
        if (knot_items[k] < 0 || knot_items[k] >= t->n || (k > 0 && knot_items[k] < knot_items[k - 1]))
            return fail(ATG_EINVAL, "atg_rejection_set_model: knot item %d = %d out of order or range", k,
                        knot_items[k]);
    }
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(t->d_rej_knot_items, knot_items, sizeof(int32_t) * ((size_t)knots + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->d_rej_knot_err, err_bounds, sizeof(uint32_t) * (size_t)knots, hipMemcpyHostToDevice));
    return ATG_OK;
}

atg_status atg_rejection_sample_sectioned(atg_table* t, uint64_t count, uint64_t seed, int64_t section_slots,
                                          int32_t* out, uint32_t* hist, uint64_t* checksum, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_rejection_sample_sectioned: NULL table");
    if (!t->rej_built) return fail(ATG_ESTATE, "atg_rejection_sample_sectioned: rejection structure not built");
    if (section_slots <= 0) section_slots = 4096;
    if (section_slots > 8192)
        return fail(ATG_EINVAL, "atg_rejection_sample_sectioned: section_slots %lld > 8192", (long long)section_slots);
    hipStream_t s = (hipStream_t)stream;
    if (!t->rej_size_known) {
        // |A| sizes the launch: read back once per rejection build.
        HIPCHK(hipStreamSynchronize(t->last_stream));
        DevState st;
        HIPCHK(hipMemcpy(&st, t->d_rej_state, sizeof(st), hipMemcpyDeviceToHost));
        const atg_status rc = state_status(st);
        if (rc != ATG_OK) return rc;
        t->rej_size_host = st.rej_size;
        t->rej_size_known = true;
    }
    t->last_stream = s;
    const uint64_t Asz = t->rej_size_host;
    if (Asz == 0 || Asz > t->A_cap) return fail(ATG_ESTATE, "atg_rejection_sample_sectioned: |A| = %llu", (unsigned long long)Asz);
    const int64_t nsec = (int64_t)((Asz + (uint64_t)section_slots - 1) / (uint64_t)section_slots);
    if (nsec > t->rej_sec_cap) {
        // Grows only; freed with the table.
        HIPCHK(hipStreamSynchronize(s));
        void* ptrs[] = {t->d_rej_mass, t->d_rej_pre, t->d_rej_sec_counts};
        for (void* p : ptrs)
            if (p) HIPCHK(hipFree(p));
        t->d_rej_mass = t->d_rej_pre = nullptr;
        t->d_rej_sec_counts = nullptr;
        t->rej_sec_cap = 0;
        if (dalloc(&t->d_rej_mass, (size_t)nsec) != hipSuccess || dalloc(&t->d_rej_pre, (size_t)nsec + 1) != hipSuccess ||
            dalloc(&t->d_rej_sec_counts, 2 * (size_t)nsec) != hipSuccess) {
            (void)hipGetLastError();
            return fail(ATG_ENOMEM, "atg_rejection_sample_sectioned: section buffers");
        }
        t->rej_sec_cap = nsec;
        t->rej_sec_valid = false;
    }
    const bool masses = !t->rej_sec_valid || t->rej_sec_SS != section_slots;
    RejSecArgs r{};
    r.A = t->d_A;
    r.A_size = Asz;
    r.SS = section_slots;
    r.nsec = nsec;
    r.mass = t->d_rej_mass;
    r.pre = t->d_rej_pre;
    r.sec_counts = t->d_rej_sec_counts;
    SampleArgs a{};
    a.st = t->d_rej_state;
    a.n = t->n;
    a.count = count;
    a.seed = seed;
    a.out = out;
    a.hist = hist;
    a.checksum = checksum ? t->d_checksum : nullptr;
    if (checksum) HIPCHK(hipMemsetAsync(t->d_checksum, 0, 8, s));
    if (count > 0) {
        HIPCHK(set_rej_sample_lds_limit((int)(section_slots * sizeof(RejEntry))));
        HIPCHK(launch_rejection_sample_sectioned(r, a, masses, s));
        t->rej_sec_valid = true;
        t->rej_sec_SS = section_slots;
    }
    if (checksum) {
        unsigned long long v = 0;
        HIPCHK(hipMemcpyAsync(&v, t->d_checksum, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *checksum = v;
    }
    return ATG_OK;
}

atg_status atg_rejection_size(atg_table* t, uint64_t* a_size) {
    if (!t || !a_size) return fail(ATG_EINVAL, "atg_rejection_size: NULL argument");
    if (!t->rej_built) return fail(ATG_ESTATE, "atg_rejection_size: rejection structure not built");
    DevState st;
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(&st, t->d_rej_state, sizeof(st), hipMemcpyDeviceToHost));
    *a_size = st.rej_size;
    return state_status(st);
}

atg_status atg_rejection_weight_sum(atg_table* t, double* W) {
    if (!t || !W) return fail(ATG_EINVAL, "atg_rejection_weight_sum: NULL argument");
    if (!t->rej_built) return fail(ATG_ESTATE, "atg_rejection_weight_sum: rejection structure not built");
    DevState st;
    HIPCHK(hipStreamSynchronize(t->last_stream));
    HIPCHK(hipMemcpy(&st, t->d_rej_state, sizeof(st), hipMemcpyDeviceToHost));
    *W = st.W;
    return state_status(st);
}

atg_status atg_get_rejection_array(atg_table* t, int32_t* A, uint64_t cap) {
    uint64_t sz = 0;
    atg_status rc = atg_rejection_size(t, &sz);
    if (rc != ATG_OK) return rc;
    if (cap < sz) return fail(ATG_EINVAL, "atg_get_rejection_array: capacity %llu < %llu", (unsigned long long)cap, (unsigned long long)sz);
    std::vector<RejEntry> e((size_t)sz);
    if (sz) HIPCHK(hipMemcpy(e.data(), t->d_A, sizeof(RejEntry) * sz, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < sz; ++i) A[i] = e[i].item;
    return ATG_OK;
}

atg_status atg_get_rejection_accept(atg_table* t, double* accept, uint64_t cap) {
    if (!accept) return fail(ATG_EINVAL, "atg_get_rejection_accept: NULL argument");
    uint64_t sz = 0;
    atg_status rc = atg_rejection_size(t, &sz);
    if (rc != ATG_OK) return rc;
    if (cap < sz) return fail(ATG_EINVAL, "atg_get_rejection_accept: capacity %llu < %llu", (unsigned long long)cap, (unsigned long long)sz);
    std::vector<RejEntry> e((size_t)sz);
    if (sz) HIPCHK(hipMemcpy(e.data(), t->d_A, sizeof(RejEntry) * sz, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < sz; ++i) accept[i] = e[i].accept;
    return ATG_OK;
}

atg_status atg_rejection_sample(atg_table* t, uint64_t count, uint64_t seed, int32_t* out, uint32_t* hist,
                                uint64_t* checksum, void* stream) {
    if (!t) return fail(ATG_EINVAL, "atg_rejection_sample: NULL table");
    if (!t->rej_built) return fail(ATG_ESTATE, "atg_rejection_sample: rejection structure not built");
    hipStream_t s = (hipStream_t)stream;
    t->last_stream = s;
    RejArgs r{};
    r.weights = t->d_weights;
    r.dtype_f32 = t->dtype == ATG_F32;
    r.n = t->n;
    r.st = t->d_rej_state;
    r.A = t->d_A;
    r.A_cap = t->A_cap;
    SampleArgs a{};
    a.st = t->d_rej_state;
    a.n = t->n;
    a.count = count;
    a.seed = seed;
    a.out = out;
    a.hist = hist;
    a.checksum = checksum ? t->d_checksum : nullptr;
    if (checksum) HIPCHK(hipMemsetAsync(t->d_checksum, 0, 8, s));
    if (count > 0) HIPCHK(launch_rejection_sample(r, a, s));
    if (checksum) {
        unsigned long long v = 0;
        HIPCHK(hipMemcpyAsync(&v, t->d_checksum, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *checksum = v;
    }
    return ATG_OK;
}

}  // extern "C"

namespace atg {

// Changes the chain count and the requested splits of a created table without
// reallocating (the autotuner's inner loop); num_splits is clamped as
// atg_create clamps it and must fit the splits array.
atg_status table_set_shape(atg_table* t, int32_t chains, int64_t num_splits) {
    if (!t) return fail(ATG_EINVAL, "table_set_shape: NULL table");
    if (chains < 1 || chains > 64) return fail(ATG_EINVAL, "table_set_shape: chains=%d", chains);
    const int64_t p = std::max<int64_t>(2, std::min<int64_t>(num_splits, t->n / 10));
    if (p > t->p_cap) return fail(ATG_EINVAL, "table_set_shape: %lld splits exceed %lld", (long long)p, (long long)t->p_cap);
    t->opts.chains = chains;
    t->opts.num_splits = num_splits;
    t->p_req = p;
    t->built = false;
    return ATG_OK;
}

// Changes the knobs of a created table that leave its rows unchanged: the
// split method (the learned index returns the bisection's splits), the knot
// count (at most the count the table was created with) and the split/pack
// interleave (its second stream is created on first use).
atg_status table_set_method(atg_table* t, int32_t split_method, int32_t learned_segments, int32_t interleave) {
    if (!t) return fail(ATG_EINVAL, "table_set_method: NULL table");
    if (split_method != ATG_SPLIT_BISECT && split_method != ATG_SPLIT_LEARNED && split_method != ATG_SPLIT_PARY)
        return fail(ATG_EINVAL, "table_set_method: unknown split method %d", split_method);
    if (learned_segments < 1 || learned_segments > t->knots_cap)
        return fail(ATG_EINVAL, "table_set_method: %d knots outside [1, %d]", learned_segments, t->knots_cap);
    if (interleave && !t->sp_stream) {
        HIPCHK(hipStreamCreateWithFlags(&t->sp_stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&t->ev_fork, hipEventDisableTiming));
        for (auto& e : t->ev_chunk) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    t->opts.split_method = split_method;
    t->opts.learned_segments = learned_segments;
    t->opts.interleave = interleave ? 1 : 0;
    t->built = false;
    return ATG_OK;
}

}  // namespace atg
