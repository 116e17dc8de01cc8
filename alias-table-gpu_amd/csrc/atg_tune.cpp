// Autotuner of the build shape (no counterpart in the reference, whose shape
// is compile-time: NUM_LOCAL_SPLITS, the 1000-item sections and the AUTO
// thread count, PsaPlusNormal.cu:3, AliasTableSplitGpu.cu:58-71,356).
//
// The section size S, the PSA+ chains C and the splits p trade the in-section
// greedy pass against the leftover volume and the split/pack chain lengths;
// the best point depends on the device and on the weights.  The tuner builds
// the caller's weights with every candidate of a small grid, times `reps`
// builds of each with HIP events (after one warm-up) and returns the options
// of the fastest median.  Every candidate is an ordinary build, so the chosen
// options produce the table the oracle defines for them.  Results can be kept
// in a text cache (ATG_TUNE_CACHE) keyed by log2(n), dtype, layout, PSA+ and
// split method; the splits are stored as the ratio n/p.
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

std::string cache_key(int64_t n, atg_dtype dtype, atg_layout layout, const atg_build_opts& o) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "n2=%d dt=%d lay=%d psa=%d split=%d", (int)std::lround(std::log2((double)n)),
                  (int)dtype, (int)layout, (int)o.psa_plus, (int)o.split_method);
    return buf;
}

bool cache_lookup(const std::string& key, int64_t* S, int32_t* C, double* ratio, double* ms) {
    const char* path = std::getenv("ATG_TUNE_CACHE");
    if (!path || !*path) return false;
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        const char* bar = std::strchr(line, '|');
        if (!bar || std::string(line, bar - line) != key) continue;
        long long s;
        int c;
        double r, m;
        if (std::sscanf(bar + 1, "%lld %d %lf %lf", &s, &c, &r, &m) == 4) {
            *S = s; *C = c; *ratio = r; *ms = m;
            found = true;  // the last entry for a key wins
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const std::string& key, int64_t S, int32_t C, double ratio, double ms) {
    const char* path = std::getenv("ATG_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s|%lld %d %.6f %.6f\n", key.c_str(), (long long)S, C, ratio, ms);
    std::fclose(f);
}

}  // namespace

extern "C" atg_status atg_autotune(const void* weights, atg_mem mem, int64_t n, atg_dtype dtype, atg_layout layout,
                                   const atg_build_opts* base, int32_t reps, atg_build_opts* best, double* best_ms) {
    if (!weights || !best) return fail(ATG_EINVAL, "atg_autotune: NULL argument");
    if (n < 1 || n > kMaxN) return fail(ATG_EINVAL, "atg_autotune: n=%lld", (long long)n);
    if (mem != ATG_MEM_HOST && mem != ATG_MEM_DEVICE) return fail(ATG_EINVAL, "atg_autotune: bad mem kind");
    if (reps < 1) reps = 3;
    atg_build_opts o;
    if (base) o = *base; else atg_default_opts(&o);
    const std::string key = cache_key(n, dtype, layout, o);
    {
        int64_t S;
        int32_t C;
        double ratio, ms;
        if (cache_lookup(key, &S, &C, &ratio, &ms)) {
            *best = o;
            best->section_size = std::min<int64_t>(S, std::max<int64_t>(1, n));
            best->chains = C;
            best->num_splits = std::max<int64_t>(2, (int64_t)((double)n / ratio));
            if (best_ms) *best_ms = -ms;  // negative: from the cache, not measured now
            return ATG_OK;
        }
    }
    // One device copy of host weights for all candidates.
    const size_t bytes = (size_t)n * (dtype == ATG_F32 ? 4 : 8);
    void* dw_owned = nullptr;
    const void* dw = weights;
    if (mem == ATG_MEM_HOST) {
        HIPCHK(hipMalloc(&dw_owned, bytes));
        if (hipMemcpy(dw_owned, weights, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(dw_owned);
            return fail(ATG_EDEVICE, "atg_autotune: weight upload");
        }
        dw = dw_owned;
    }
    // The grid: sections near the LDS limit of a wave, chains from the
    // reference's 16 up to 32, and n/p around the pack's staging capacity.
    std::vector<int64_t> Ss;
    for (int64_t S : {1024, 1536, 2048})
        if (S <= std::max<int64_t>(1, n / 10) || Ss.empty()) Ss.push_back(std::min<int64_t>(S, std::max<int64_t>(1, n / 10)));
    std::vector<int32_t> Cs = o.psa_plus ? std::vector<int32_t>{12, 16, 24, 32} : std::vector<int32_t>{o.chains};
    const double ratios[] = {24.0, 38.0, 64.0};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    atg_status rc = ATG_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(ATG_EDEVICE, "events");
    double bestv = 1e300;
    int64_t bS = 0;
    int32_t bC = 0;
    double bR = 38.0;
    for (size_t si = 0; si < Ss.size() && rc == ATG_OK; ++si) {
        atg_build_opts co = o;
        co.section_size = Ss[si];
        co.num_splits = std::max<int64_t>(2, (int64_t)((double)n / ratios[0]));
        co.record_phases = 0;
        atg_table* t = nullptr;
        rc = atg_create(&t, n, dtype, layout, &co);
        if (rc != ATG_OK) break;
        rc = atg_set_weights(t, dw, ATG_MEM_DEVICE, nullptr);
        for (int32_t C : Cs) {
            for (double ratio : ratios) {
                if (rc != ATG_OK) break;
                rc = table_set_shape(t, C, std::max<int64_t>(2, (int64_t)((double)n / ratio)));
                if (rc != ATG_OK) break;
                rc = atg_build(t, nullptr);
                if (rc == ATG_OK) rc = atg_sync(t, nullptr);
                std::vector<float> ts;
                for (int r = 0; r < reps && rc == ATG_OK; ++r) {
                    (void)hipEventRecord(e0, nullptr);
                    rc = atg_build(t, nullptr);
                    (void)hipEventRecord(e1, nullptr);
                    if (rc == ATG_OK) rc = atg_sync(t, nullptr);
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    ts.push_back(ms);
                }
                if (rc != ATG_OK) break;
                std::sort(ts.begin(), ts.end());
                const double med = ts[ts.size() / 2];
                if (med < bestv) {
                    bestv = med;
                    bS = Ss[si];
                    bC = C;
                    bR = ratio;
                }
            }
        }
        atg_destroy(t);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dw_owned) (void)hipFree(dw_owned);
    if (rc != ATG_OK) return rc;
    *best = o;
    best->section_size = bS;
    best->chains = bC;
    best->num_splits = std::max<int64_t>(2, (int64_t)((double)n / bR));
    if (best_ms) *best_ms = bestv;
    cache_store(key, bS, bC, bR, bestv);
    return ATG_OK;
}
