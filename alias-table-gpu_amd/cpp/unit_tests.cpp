// The reference's self-tests (UnitTests.cu:3-145) restated on the C++ facade:
// table correctness for every weight distribution and build variant, the
// regression case N=505711 / p=16576 / OneLight, and the per-item sampling
// distribution bound |expected - sampled| < S/N for every sampler.
// Exit status 0 when everything passes.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>

#include "atg.hpp"

using namespace atg;

static int failures = 0;

template <Storage S>
static void expect_build(int64_t n, Variant v, int64_t threads, WeightDistribution d) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    AliasTableSplitGpu<S> t(n, v, threads, d);
    const auto t1 = clk::now();
    t.preBuild();
    t.build();
    const auto t2 = clk::now();
    const bool ok = t.postBuild();
    const auto t3 = clk::now();
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::printf("%s %s n=%lld dist=%d max_rel=%.3g (generate %.1f ms, build %.1f ms, verify %.1f ms)\n",
                ok ? "ok" : "FAILED:", t.name().c_str(), (long long)n, (int)d, t.maxRelError, ms(t0, t1), ms(t1, t2),
                ms(t2, t3));
    std::fflush(stdout);
    if (!ok) ++failures;
}

// UnitTests::testTableCorrectness (UnitTests.cu:10-94).
static void testTableCorrectness() {
    for (int d = 1; d <= MAX_WEIGHT_DISTRIBUTION; ++d) {
        for (bool psa : {false, true}) {
            for (atg_split_method sm : {ATG_SPLIT_BISECT, ATG_SPLIT_LEARNED}) {
                Variant v;
                v.psaPlus = psa;
                v.split = sm;
                expect_build<Storage::ArrayOfStructs>(10000, v, NUMBER_THREADS_AUTO, (WeightDistribution)d);
                expect_build<Storage::StructOfArrays>(10000, v, NUMBER_THREADS_AUTO, (WeightDistribution)d);
                if (sm == ATG_SPLIT_BISECT)
                    expect_build<Storage::ArrayOfStructs>(1000000, v, 1024, (WeightDistribution)d);
            }
        }
    }
    // The pack variants (Variant::pack) on every distribution.
    for (int d = 1; d <= MAX_WEIGHT_DISTRIBUTION; ++d) {
        for (atg_pack_method pm : {ATG_PACK_BASIC, ATG_PACK_NO_WEIGHTS}) {
            Variant v;
            v.pack = pm;
            expect_build<Storage::ArrayOfStructs>(100000, v, NUMBER_THREADS_AUTO, (WeightDistribution)d);
            expect_build<Storage::StructOfArrays>(10000, v, 1024, (WeightDistribution)d);
        }
    }
    // Regression: a section group with zero eligible pack lanes (UnitTests.cu:47-49).
    Variant v;
    v.psaPlus = false;
    expect_build<Storage::ArrayOfStructs>(505711, v, 16576, weightDistributionOneLight);
}

// The random rounds of UnitTests::testTableCorrectness (UnitTests.cu:51-93):
// 20 rounds, each a random variant, n = 100 + a random number below 10^e
// with e uniform in 3..7 (small tables as likely as large ones), a random
// weight distribution and a random PSA+ choice, built as AoS; then the same
// with PSA+ off as SoA.  The reference draws from an unseeded rand(); here
// a fixed seed makes a failure reproducible.
static void testRandomRounds(uint32_t seed) {
    std::mt19937 rng(seed);
    struct V { atg_split_method split; bool interleave; bool psa_ok; };
    const V variants[] = {{ATG_SPLIT_BISECT, false, true}, {ATG_SPLIT_LEARNED, false, true},
                          {ATG_SPLIT_BISECT, true, true}, {ATG_SPLIT_LEARNED, true, true},
                          {ATG_SPLIT_PARY, false, true}, {ATG_SPLIT_PARY, true, true}};
    const int numTests = 20;
    std::printf("random rounds (seed %u)\n", seed);
    for (int i = 0; i < numTests; ++i) {
        const V& pick = variants[rng() % (sizeof(variants) / sizeof(V))];
        const int64_t n_max = (int64_t)std::pow(10.0, 3 + (int)(rng() % 5));  // 1e3 .. 1e7
        const int64_t n = 100 + (int64_t)(rng() % (uint64_t)n_max);
        const WeightDistribution d = (WeightDistribution)(rng() % MAX_WEIGHT_DISTRIBUTION + 1);
        Variant v;
        v.split = pick.split;
        v.interleave = pick.interleave;
        v.psaPlus = pick.psa_ok && (rng() % 2) == 1;
        expect_build<Storage::ArrayOfStructs>(n, v, NUMBER_THREADS_AUTO, d);
        v.psaPlus = false;
        expect_build<Storage::StructOfArrays>(n, v, NUMBER_THREADS_AUTO, d);
    }
}

// Every fill of A writes the direct fill's A, on every weight distribution
// and on the outlier weights of RejectionSampling::benchmarkBuild
// (RejectionSampling.cu:58-70, PlotConstruction.cuh:18-27).
static void testRejectionFills() {
    for (int d = 1; d <= MAX_WEIGHT_DISTRIBUTION + 1; ++d) {
        const bool outlier = d > MAX_WEIGHT_DISTRIBUTION;
        const int64_t n = outlier ? 100000 : 30011;
        const WeightDistribution dist = outlier ? weightDistributionSine : (WeightDistribution)d;
        RejectionSampling direct(n, dist);
        if (outlier) direct.weights[0] = 2e7;
        direct.fullBuild();
        for (RejectionFill f : {RejectionFill::BinarySearch, RejectionFill::PArySearch, RejectionFill::Learned}) {
            RejectionSampling r(n, dist, f, 256);
            r.weights = direct.weights;
            r.fullBuild();
            const bool ok = r.ACpu == direct.ACpu && r.W == direct.W;
            std::printf("%s %s n=%lld dist=%d%s |A|=%zu\n", ok ? "ok" : "FAILED:", r.name().c_str(), (long long)n,
                        (int)dist, outlier ? " (outlier 2e7)" : "", r.ACpu.size());
            if (!ok) ++failures;
        }
    }
    for (RejectionFill f : {RejectionFill::Direct, RejectionFill::BinarySearch, RejectionFill::Learned}) {
        RejectionSampling r(1000000, weightDistributionSine, f);
        const double ms = r.benchmarkBuild(1e7);
        std::printf("%s: benchmarkBuild(1e7) at N=1e6: %.3f ms per fillA\n", r.name().c_str(), ms);
    }
}

// UnitTests::testSamplingDistribution (UnitTests.cu:96-145).
static void testSamplingDistribution() {
    const uint64_t numSamples = 2000000;
    const int64_t tableSize = 10000;
    AliasTableSplitGpu<Storage::ArrayOfStructs> aos(tableSize, Variant(), NUMBER_THREADS_AUTO, weightDistributionSine);
    AliasTableSplitGpu<Storage::StructOfArrays> soa(tableSize, Variant(), NUMBER_THREADS_AUTO, weightDistributionSine);
    RejectionSampling rej(tableSize, weightDistributionSine);
    aos.fullBuild();
    soa.fullBuild();
    rej.fullBuild();
    std::vector<double> expected(tableSize);
    for (int64_t k = 0; k < tableSize; ++k) expected[k] = aos.weights[k] / aos.W * (double)numSamples;  // SamplerExpected

    std::vector<std::unique_ptr<Sampler>> samplers;
    samplers.emplace_back(new SamplerGpuBasic<Storage::ArrayOfStructs>(aos));
    samplers.emplace_back(new SamplerGpuSectioned<Storage::ArrayOfStructs>(aos, tableSize / 4));
    samplers.emplace_back(new SamplerGpuBasic<Storage::StructOfArrays>(soa));
    samplers.emplace_back(new SamplerGpuSectioned<Storage::StructOfArrays>(soa, tableSize / 4));
    samplers.emplace_back(new SamplerRejection(rej));
    for (auto& s : samplers) {
        const std::vector<int> got = s->getSamplingDistribution(numSamples);
        long long total = 0;
        for (int64_t k = 0; k < tableSize; ++k) {
            total += got[k];
            if (std::fabs(expected[k] - got[k]) >= (double)(numSamples / tableSize)) {
                std::printf("FAILED: %s expected %.1f samples but got %d at %lld\n", s->name().c_str(), expected[k],
                            got[k], (long long)k);
                ++failures;
                break;
            }
        }
        if (total != (long long)numSamples) {
            std::printf("FAILED: %s drew %lld samples, asked %llu\n", s->name().c_str(), total,
                        (unsigned long long)numSamples);
            ++failures;
        }
        std::printf("%s: %.2f GSamples/s (benchmarkSampling of 1e8)\n", s->name().c_str(), s->benchmarkSampling(100000000));
    }
}

int main() {
    try {
        testTableCorrectness();
        testRandomRounds(20260101u);
        testRejectionFills();
        testSamplingDistribution();
    } catch (const Error& e) {
        std::printf("FAILED with status %d: %s\n", (int)e.status, e.what());
        return 2;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
