// Autotuners (no counterpart in the reference, whose shape is compile-time:
// NUM_LOCAL_SPLITS, the 1000-item sections and the AUTO thread count,
// PsaPlusNormal.cu:3, AliasTableSplitGpu.cu:58-71,356).
//
// Two kinds of knobs:
//  * Knobs that leave the table unchanged: the split method (the learned
//    index returns the bisection's splits), the learned index's knot count,
//    the split/pack interleave, and for a batch the launch groups.
//    atg_autotune and atg_batch_autotune time these and nothing else, so the
//    table built from given weights and a given shape is the same whichever
//    candidate wins a timing race.
//  * The build shape: section size S, PSA+ chains C and splits p.  They
//    change the table (W's association, the leftovers and every row).
//    atg_autotune_shape searches them; its result is an explicit input that
//    the caller passes to atg_create and that atg_save_table records.
//
// Each candidate is timed as the median of `reps` builds (HIP events, after
// one warm-up build).  Results can be kept in a text cache (ATG_TUNE_CACHE),
// one line per key; the last line of a key wins.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "atg_host.h"

using namespace atg;

namespace {

bool cache_lookup(const std::string& key, std::vector<double>* vals) {
    const char* path = std::getenv("ATG_TUNE_CACHE");
    if (!path || !*path) return false;
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        const char* bar = std::strchr(line, '|');
        if (!bar || std::string(line, bar - line) != key) continue;
        std::vector<double> v;
        const char* p = bar + 1;
        char* end = nullptr;
        for (double x = std::strtod(p, &end); end != p; x = std::strtod(p, &end)) {
            v.push_back(x);
            p = end;
        }
        if (!v.empty()) {
            *vals = v;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const std::string& key, const std::vector<double>& vals) {
    const char* path = std::getenv("ATG_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s|", key.c_str());
    for (double v : vals) std::fprintf(f, " %.9g", v);
    std::fprintf(f, "\n");
    std::fclose(f);
}

// Median ms of `reps` builds of t (after a warm-up build) on the null stream.
atg_status time_builds(atg_table* t, int reps, hipEvent_t e0, hipEvent_t e1, double* med) {
    atg_status rc = atg_build(t, nullptr);
    if (rc == ATG_OK) rc = atg_sync(t, nullptr);
    std::vector<float> ts;
    for (int r = 0; r < reps && rc == ATG_OK; ++r) {
        if (hipEventRecord(e0, nullptr) != hipSuccess) return fail(ATG_EDEVICE, "hipEventRecord");
        rc = atg_build(t, nullptr);
        if (hipEventRecord(e1, nullptr) != hipSuccess) return fail(ATG_EDEVICE, "hipEventRecord");
        if (rc == ATG_OK) rc = atg_sync(t, nullptr);
        float ms = 0.f;
        if (rc == ATG_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return fail(ATG_EDEVICE, "hipEventElapsedTime");
        ts.push_back(ms);
    }
    if (rc != ATG_OK) return rc;
    std::sort(ts.begin(), ts.end());
    *med = ts[ts.size() / 2];
    return ATG_OK;
}

// A device copy of host weights (one for all candidates), or the caller's
// device pointer.
struct DeviceWeights {
    void* owned = nullptr;
    const void* ptr = nullptr;
    ~DeviceWeights() {
        if (owned) (void)hipFree(owned);
    }
    atg_status init(const void* w, atg_mem mem, size_t bytes) {
        if (mem == ATG_MEM_DEVICE) {
            ptr = w;
            return ATG_OK;
        }
        HIPCHK(hipMalloc(&owned, bytes));
        HIPCHK(hipMemcpy(owned, w, bytes, hipMemcpyHostToDevice));
        ptr = owned;
        return ATG_OK;
    }
};

struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    atg_status init() {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        return ATG_OK;
    }
};

atg_status check_common(const void* weights, atg_mem mem, int64_t n, const atg_build_opts* best, const char* who) {
    if (!weights || !best) return fail(ATG_EINVAL, "%s: NULL argument", who);
    if (n < 1 || n > kMaxN) return fail(ATG_EINVAL, "%s: n=%lld", who, (long long)n);
    if (mem != ATG_MEM_HOST && mem != ATG_MEM_DEVICE) return fail(ATG_EINVAL, "%s: bad mem kind", who);
    return ATG_OK;
}

}  // namespace

extern "C" atg_status atg_autotune(const void* weights, atg_mem mem, int64_t n, atg_dtype dtype, atg_layout layout,
                                   const atg_build_opts* base, int32_t reps, atg_build_opts* best, double* best_ms) {
    atg_status rc = check_common(weights, mem, n, best, "atg_autotune");
    if (rc != ATG_OK) return rc;
    if (reps < 1) reps = 3;
    atg_build_opts o;
    if (base) o = *base; else atg_default_opts(&o);
    o.record_phases = 0;
    char key[192];
    std::snprintf(key, sizeof(key), "neutral n2=%d dt=%d lay=%d psa=%d S=%lld C=%d p=%lld",
                  (int)std::lround(std::log2((double)n)), (int)dtype, (int)layout, (int)o.psa_plus,
                  (long long)o.section_size, (int)o.chains, (long long)o.num_splits);
    std::vector<double> cached;
    if (cache_lookup(key, &cached) && cached.size() == 4) {
        *best = o;
        best->split_method = (int32_t)cached[0];
        best->learned_segments = (int32_t)cached[1];
        best->interleave = (int32_t)cached[2];
        if (best_ms) *best_ms = -cached[3];  // negative: from the cache, not measured now
        return ATG_OK;
    }
    DeviceWeights dw;
    rc = dw.init(weights, mem, (size_t)n * (dtype == ATG_F32 ? 4 : 8));
    if (rc != ATG_OK) return rc;
    Events ev;
    if ((rc = ev.init()) != ATG_OK) return rc;
    // Candidates: bisection, the learned index with 1024 to 65536 knots (at
    // most one per 16 splits are used) and the P-ary search, each with and
    // without the split/pack interleave.
    struct Cand { int32_t split, knots, inter; };
    const Cand cands[] = {{ATG_SPLIT_BISECT, 4096, 0},     {ATG_SPLIT_BISECT, 4096, 1},
                          {ATG_SPLIT_LEARNED, 1024, 0},    {ATG_SPLIT_LEARNED, 4096, 0},
                          {ATG_SPLIT_LEARNED, 16384, 0},   {ATG_SPLIT_LEARNED, 65536, 0},
                          {ATG_SPLIT_LEARNED, 1024, 1},    {ATG_SPLIT_LEARNED, 4096, 1},
                          {ATG_SPLIT_LEARNED, 16384, 1},   {ATG_SPLIT_LEARNED, 65536, 1},
                          {ATG_SPLIT_PARY, 4096, 0},       {ATG_SPLIT_PARY, 4096, 1}};
    atg_build_opts co = o;
    co.learned_segments = 65536;
    co.interleave = 0;
    atg_table* t = nullptr;
    if ((rc = atg_create(&t, n, dtype, layout, &co)) != ATG_OK) return rc;
    rc = atg_set_weights(t, dw.ptr, ATG_MEM_DEVICE, nullptr);
    double bestv = 1e300;
    Cand bc = cands[0];
    for (const Cand& c : cands) {
        if (rc != ATG_OK) break;
        rc = table_set_method(t, c.split, c.knots, c.inter);
        double med = 0;
        if (rc == ATG_OK) rc = time_builds(t, reps, ev.e0, ev.e1, &med);
        if (rc == ATG_OK && med < bestv) {
            bestv = med;
            bc = c;
        }
    }
    atg_destroy(t);
    if (rc != ATG_OK) return rc;
    *best = o;
    best->split_method = bc.split;
    best->learned_segments = bc.knots;
    best->interleave = bc.inter;
    if (best_ms) *best_ms = bestv;
    cache_store(key, {(double)bc.split, (double)bc.knots, (double)bc.inter, bestv});
    return ATG_OK;
}

extern "C" atg_status atg_autotune_shape(const void* weights, atg_mem mem, int64_t n, atg_dtype dtype,
                                         atg_layout layout, const atg_build_opts* base, int32_t reps,
                                         atg_build_opts* best, double* best_ms) {
    atg_status rc = check_common(weights, mem, n, best, "atg_autotune_shape");
    if (rc != ATG_OK) return rc;
    if (reps < 1) reps = 3;
    atg_build_opts o;
    if (base) o = *base; else atg_default_opts(&o);
    o.record_phases = 0;
    char key[160];
    std::snprintf(key, sizeof(key), "shape n2=%d dt=%d lay=%d psa=%d split=%d", (int)std::lround(std::log2((double)n)),
                  (int)dtype, (int)layout, (int)o.psa_plus, (int)o.split_method);
    std::vector<double> cached;
    if (cache_lookup(key, &cached) && cached.size() == 4) {
        *best = o;
        best->section_size = std::min<int64_t>((int64_t)cached[0], std::max<int64_t>(1, n));
        best->chains = (int32_t)cached[1];
        best->num_splits = std::max<int64_t>(2, (int64_t)((double)n / cached[2]));
        if (best_ms) *best_ms = -cached[3];
        return ATG_OK;
    }
    DeviceWeights dw;
    rc = dw.init(weights, mem, (size_t)n * (dtype == ATG_F32 ? 4 : 8));
    if (rc != ATG_OK) return rc;
    Events ev;
    if ((rc = ev.init()) != ATG_OK) return rc;
    // The grid: sections near the LDS limit of a wave, chains from the
    // reference's 16 up to 32, and n/p around the pack's staging capacity.
    std::vector<int64_t> Ss;
    for (int64_t S : {1024, 1536, 2048})
        if (S <= std::max<int64_t>(1, n / 10) || Ss.empty()) Ss.push_back(std::min<int64_t>(S, std::max<int64_t>(1, n / 10)));
    const std::vector<int32_t> Cs = o.psa_plus ? std::vector<int32_t>{12, 16, 24, 32} : std::vector<int32_t>{o.chains};
    const double ratios[] = {24.0, 38.0, 64.0};
    double bestv = 1e300;
    int64_t bS = Ss[0];
    int32_t bC = Cs[0];
    double bR = 38.0;
    for (size_t si = 0; si < Ss.size() && rc == ATG_OK; ++si) {
        atg_build_opts co = o;
        co.section_size = Ss[si];
        co.num_splits = std::max<int64_t>(2, (int64_t)((double)n / ratios[0]));
        atg_table* t = nullptr;
        if ((rc = atg_create(&t, n, dtype, layout, &co)) != ATG_OK) break;
        rc = atg_set_weights(t, dw.ptr, ATG_MEM_DEVICE, nullptr);
        for (int32_t C : Cs)
            for (double ratio : ratios) {
                if (rc != ATG_OK) break;
                rc = table_set_shape(t, C, std::max<int64_t>(2, (int64_t)((double)n / ratio)));
                double med = 0;
                if (rc == ATG_OK) rc = time_builds(t, reps, ev.e0, ev.e1, &med);
                if (rc == ATG_OK && med < bestv) {
                    bestv = med;
                    bS = Ss[si];
                    bC = C;
                    bR = ratio;
                }
            }
        atg_destroy(t);
    }
    if (rc != ATG_OK) return rc;
    *best = o;
    best->section_size = bS;
    best->chains = bC;
    best->num_splits = std::max<int64_t>(2, (int64_t)((double)n / bR));
    if (best_ms) *best_ms = bestv;
    cache_store(key, {(double)bS, (double)bC, bR, bestv});
    return ATG_OK;
}
