// Input synthesis (SamplingAlgorithm::generateWeights, SamplingAlgorithm.cu:21-71)
// from a Philox4x32-10 stream so that every workload is reproducible (the
// reference draws from rand() after srand(time) and shuffles with a time seed).
// Host code, multi-threaded; item i uses counter (i, purpose, 0).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/atg.h"

namespace {

inline void philox(uint64_t idx, uint32_t purpose, uint32_t extra, uint64_t seed, uint32_t out[4]) {
    uint32_t c0 = (uint32_t)idx, c1 = (uint32_t)(idx >> 32), c2 = purpose, c3 = extra;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// u in [0, 1) with 53 bits from words 2-3; v from words 0-1.
inline double u53(const uint32_t x[4], int half) {
    const uint64_t r = half ? (((uint64_t)x[3] << 32) | x[2]) : (((uint64_t)x[1] << 32) | x[0]);
    return (double)(r >> 11) * 0x1.0p-53;
}

thread_local bool t_serial = false;  // set inside the per-table workers of a batch

template <typename F>
void parallel_for(int64_t n, F f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536 || t_serial) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=] {
            const int64_t a = n * t / nt, b = n * (t + 1) / nt;
            for (int64_t i = a; i < b; ++i) f(i);
        });
    for (auto& x : th) x.join();
}

// Fisher-Yates with Philox draws (shuffle of SamplingAlgorithm.cu:66-70).
void shuffle(double* w, int64_t n, uint64_t seed) {
    for (int64_t i = n - 1; i > 0; --i) {
        uint32_t x[4];
        philox((uint64_t)i, 0xA0F0u, 0, seed, x);
        const uint64_t r = ((uint64_t)x[1] << 32) | x[0];
        const int64_t j = (int64_t)(((unsigned __int128)r * (uint64_t)(i + 1)) >> 64);
        std::swap(w[i], w[j]);
    }
}

}  // namespace

extern "C" atg_status atg_generate_weights(int32_t dist, int64_t n, uint64_t seed, double param, double* w) {
    if (!w || n < 1) return ATG_EINVAL;
    const uint32_t purpose = 0xA000u + (uint32_t)dist;
    switch (dist) {
        case ATG_DIST_SINE:
            // 20.f * (1.1f + cos(0.008 * i)) evaluated as in the reference (float
            // constants promoted to double).
            parallel_for(n, [=](int64_t i) { w[i] = (double)20.f * ((double)1.1f + std::cos(0.008 * (double)i)); });
            break;
        case ATG_DIST_BEST_WORST:
            parallel_for(n, [=](int64_t i) { w[i] = (i == 0) ? 0.5 : (i == n - 1) ? 1.5 : 1.0; });
            break;
        case ATG_DIST_ONE_HEAVY:
            parallel_for(n, [=](int64_t i) { w[i] = (i == 0) ? 1.5 : 1.0; });
            break;
        case ATG_DIST_ONE_LIGHT:
            parallel_for(n, [=](int64_t i) { w[i] = (i == 0) ? 0.5 : 1.0; });
            break;
        case ATG_DIST_RAMP:
            parallel_for(n, [=](int64_t i) { w[i] = (double)i; });
            break;
        case ATG_DIST_UNIFORM:  // 10 * rand()/RAND_MAX, here 10 * u
            parallel_for(n, [=](int64_t i) { uint32_t x[4]; philox(i, purpose, 0, seed, x); w[i] = 10.0 * u53(x, 1); });
            break;
        case ATG_DIST_POWERLAW1:
        case ATG_DIST_POWERLAW1_SHUFFLED:
            parallel_for(n, [=](int64_t i) { w[i] = 1.0 / (double)(i + 1); });
            if (dist == ATG_DIST_POWERLAW1_SHUFFLED) shuffle(w, n, seed);
            break;
        case ATG_DIST_POWERLAW05_SHUFFLED:
            parallel_for(n, [=](int64_t i) { w[i] = 1.0 / std::sqrt((double)(i + 1)); });
            shuffle(w, n, seed);
            break;
        case ATG_DIST_POWERLAW2:
            parallel_for(n, [=](int64_t i) { const double x = (double)(i + 1); w[i] = 1.0 / (x * x); });
            break;
        case ATG_DIST_UNIT_UNIFORM:
            parallel_for(n, [=](int64_t i) { uint32_t x[4]; philox(i, purpose, 0, seed, x); w[i] = 1.0 - u53(x, 1); });
            break;
        case ATG_DIST_PARETO: {
            const double a = param > 0 ? param : 1.1;
            parallel_for(n, [=](int64_t i) {
                uint32_t x[4]; philox(i, purpose, 0, seed, x);
                w[i] = std::pow(1.0 - u53(x, 1), -1.0 / a);
            });
            break;
        }
        case ATG_DIST_EXPONENTIAL:
            parallel_for(n, [=](int64_t i) { uint32_t x[4]; philox(i, purpose, 0, seed, x); w[i] = -std::log1p(-u53(x, 1)); });
            break;
        case ATG_DIST_HEAVY_MIX: {
            // 1% heavy items at positions chosen by a partial Fisher-Yates; heavy
            // weights lognormal(10, 1) (Box-Muller), light U(0,1); heavy weights
            // rescaled so the heavy class carries `param` (default 0.9) of the mass.
            const double frac = (param > 0 && param < 1) ? param : 0.9;
            const int64_t nh = std::max<int64_t>(1, n / 100);
            std::vector<int64_t> pos(n);
            for (int64_t i = 0; i < n; ++i) pos[i] = i;
            std::vector<char> heavy(n, 0);
            for (int64_t i = 0; i < nh; ++i) {
                uint32_t x[4];
                philox((uint64_t)i, 0xA0F1u, 0, seed, x);
                const uint64_t r = ((uint64_t)x[1] << 32) | x[0];
                const int64_t j = i + (int64_t)(((unsigned __int128)r * (uint64_t)(n - i)) >> 64);
                std::swap(pos[i], pos[j]);
                heavy[pos[i]] = 1;
            }
            parallel_for(n, [=, &heavy](int64_t i) {
                uint32_t x[4]; philox(i, purpose, 0, seed, x);
                if (heavy[i]) {
                    const double u1 = 1.0 - u53(x, 0), u2 = u53(x, 1);
                    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    w[i] = std::exp(10.0 + z);
                } else {
                    w[i] = u53(x, 1);
                }
            });
            double sl = 0, sh = 0;
            for (int64_t i = 0; i < n; ++i) (heavy[i] ? sh : sl) += w[i];
            // (No light items, as for n = 1: the heavy weights stay as drawn.)
            if (sl > 0 && sh > 0) {
                const double s = (frac / (1.0 - frac)) * sl / sh;
                parallel_for(n, [=, &heavy](int64_t i) { if (heavy[i]) w[i] *= s; });
            }
            break;
        }
        case ATG_DIST_ZIPF_SHUFFLED: {
            const double s = param > 0 ? param : 1.2;
            parallel_for(n, [=](int64_t i) { w[i] = std::pow((double)(i + 1), -s); });
            shuffle(w, n, seed);
            break;
        }
        default:
            return ATG_EINVAL;
    }
    return ATG_OK;
}

// Many tables: sizes n_t = floor(exp(U(ln lo, ln hi))) (clamped to [lo, hi]),
// table t drawn with seed + t + 1 from the single-table generator.
extern "C" atg_status atg_generate_batch(int32_t dist, int64_t ntables, int64_t lo, int64_t hi, uint64_t seed,
                                         double param, int64_t* offsets, void* out, int32_t out_dtype) {
    if (!offsets || ntables < 1 || lo < 1 || hi < lo) return ATG_EINVAL;
    offsets[0] = 0;
    const double a = std::log((double)lo), b = std::log((double)hi);
    for (int64_t t = 0; t < ntables; ++t) {
        uint32_t x[4];
        philox((uint64_t)t, 0xA0F2u, 0, seed, x);
        int64_t n = (int64_t)std::floor(std::exp(a + (b - a) * u53(x, 1)));
        n = std::min(hi, std::max(lo, n));
        offsets[t + 1] = offsets[t] + n;
    }
    if (!out) return ATG_OK;
    if (out_dtype != ATG_F64 && out_dtype != ATG_F32) return ATG_EINVAL;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::vector<atg_status> rc(nt, ATG_OK);
    for (unsigned k = 0; k < nt; ++k)
        th.emplace_back([&, k] {
            t_serial = true;
            std::vector<double> tmp;
            for (int64_t t = k; t < ntables; t += nt) {
                const int64_t n = offsets[t + 1] - offsets[t];
                double* dst = out_dtype == ATG_F64 ? (double*)out + offsets[t] : nullptr;
                if (!dst) {
                    tmp.resize((size_t)n);
                    dst = tmp.data();
                }
                const atg_status r = atg_generate_weights(dist, n, seed + (uint64_t)t + 1, param, dst);
                if (r != ATG_OK) rc[k] = r;
                if (out_dtype == ATG_F32) {
                    float* f = (float*)out + offsets[t];
                    for (int64_t i = 0; i < n; ++i) f[i] = (float)dst[i];
                }
            }
        });
    for (auto& x : th) x.join();
    for (atg_status r : rc)
        if (r != ATG_OK) return r;
    return ATG_OK;
}
