// C++ facade over the atg C ABI that mirrors the reference's class surface:
//   SamplingAlgorithm            construction/SamplingAlgorithm.cuh:30-48
//   AliasTableSplitGpu<Storage>  construction/aliasTable/buildMethod/AliasTableSplitGpu.cuh:96-144
//   RejectionSampling            construction/rejectionSampling/RejectionSampling.cuh:13-29 (and its
//                                Direct/BinarySearch/PArySearch subclasses as a fill choice)
//   Sampler / SamplerGpuBasic / SamplerGpuSectioned / SamplerRejection
//                                sampling/Sampler.cuh:10-31, sampling/aliasTable/*, sampling/rejectionSampling/*
// Errors are reported as atg::Error exceptions carrying the atg_status (the
// reference raises SIGTRAP, utils/Utils.cuh:13-26).
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/atg.h"

namespace atg {

struct Error : std::runtime_error {
    atg_status status;
    Error(atg_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(atg_status s) {
    if (s != ATG_OK) throw Error(s, atg_last_error());
}

// SamplingAlgorithm.cuh:16-28 (values kept) plus the workloads of the benchmarks.
enum WeightDistribution {
    weightDistributionSine = ATG_DIST_SINE,
    weightDistributionBestWorstCase = ATG_DIST_BEST_WORST,
    weightDistributionUniform = ATG_DIST_UNIFORM,
    weightDistributionRamp = ATG_DIST_RAMP,
    weightDistributionPowerLaw1 = ATG_DIST_POWERLAW1,
    weightDistributionPowerLaw2 = ATG_DIST_POWERLAW2,
    weightDistributionOneHeavy = ATG_DIST_ONE_HEAVY,
    weightDistributionOneLight = ATG_DIST_ONE_LIGHT,
    weightDistributionPowerLaw1Shuffled = ATG_DIST_POWERLAW1_SHUFFLED,
    weightDistributionPowerLaw05Shuffled = ATG_DIST_POWERLAW05_SHUFFLED,
    weightDistributionPareto = ATG_DIST_PARETO,
};
constexpr int MAX_WEIGHT_DISTRIBUTION = 10;
constexpr int64_t NUMBER_THREADS_AUTO = -4;  // AliasTableSplitGpu.cuh:47

class SamplingAlgorithm {
  public:
    SamplingAlgorithm(int64_t size, WeightDistribution dist, uint64_t seed = 42, double param = 1.1)
        : N(size), weights((size_t)size) {
        check(atg_generate_weights((int32_t)dist, size, seed, param, weights.data()));
    }
    virtual ~SamplingAlgorithm() = default;
    virtual void preBuild() = 0;
    virtual void build() = 0;
    virtual bool postBuild() = 0;
    virtual std::string name() = 0;
    bool fullBuild() {
        preBuild();
        build();
        return postBuild();
    }
    int64_t N;
    std::vector<double> weights;  // the public weights array (user weights may be written before preBuild)
    double W = 0;
};

enum class Storage { ArrayOfStructs, StructOfArrays };

// The options the reference's Variant and numThreads select
// (AliasTableSplitGpu.cuh:60-94): PSA+ on/off, split method, and the PSA+ shape.
struct Variant {
    bool psaPlus = true;
    atg_split_method split = ATG_SPLIT_BISECT;
    int chains = 16;
    int64_t sectionSize = 0;
    int learnedSegments = 4096;
    // performSplitAndPackInterleaved (AliasTableSplitGpu.cu:220-262): split
    // and pack on two streams; no per-phase timers then.
    bool interleave = false;
    // Variant::pack: CHUNKED (LDS-staged), BASIC or NO_WEIGHTS.  As in
    // Variant::setPsaPlusIfSupported (AliasTableSplitGpu.cuh:85-93), PSA+ is
    // off with NO_WEIGHTS.
    atg_pack_method pack = ATG_PACK_CHUNKED;
};

template <Storage S = Storage::ArrayOfStructs>
class AliasTableSplitGpu : public SamplingAlgorithm {
  public:
    AliasTableSplitGpu(int64_t size, Variant variant = Variant(), int64_t numThreads = NUMBER_THREADS_AUTO,
                       WeightDistribution dist = weightDistributionSine)
        : SamplingAlgorithm(size, dist), variant_(variant), numThreads_(numThreads) {}
    ~AliasTableSplitGpu() override {
        if (table_) atg_destroy(table_);
    }
    std::string name() override {
        std::string v = variant_.split == ATG_SPLIT_LEARNED ? "LearnedSplit"
                        : variant_.split == ATG_SPLIT_PARY  ? "ParySplit"
                                                            : "";
        if (variant_.pack == ATG_PACK_BASIC) v += "PackBasic";
        if (variant_.pack == ATG_PACK_NO_WEIGHTS) v += "PackWithoutWeights";
        if (variant_.interleave) v += "Interleaved";
        if (variant_.psaPlus && variant_.pack != ATG_PACK_NO_WEIGHTS) v += "PsaPlus";
        return std::string("AliasTableSplitGpu") +
               (S == Storage::ArrayOfStructs ? "ArrayOfStructs" : "StructOfArrays") + v;
    }
    // AliasTableSplitGpu::preBuild (AliasTableSplitGpu.cu:206-217): device
    // buffers and the weight upload.
    void preBuild() override {
        if (!table_) {
            atg_build_opts o;
            atg_default_opts(&o);
            o.psa_plus = variant_.psaPlus && variant_.pack != ATG_PACK_NO_WEIGHTS ? 1 : 0;
            o.pack_method = variant_.pack;
            o.chains = variant_.chains;
            o.section_size = variant_.sectionSize;
            o.num_splits = numThreads_ == NUMBER_THREADS_AUTO ? 0 : numThreads_;
            o.split_method = variant_.split;
            o.learned_segments = variant_.learnedSegments;
            o.interleave = variant_.interleave ? 1 : 0;
            o.record_phases = variant_.interleave ? 0 : 1;
            check(atg_create(&table_, N, ATG_F64, S == Storage::ArrayOfStructs ? ATG_AOS : ATG_SOA, &o));
        }
        check(atg_set_weights(table_, weights.data(), ATG_MEM_HOST, nullptr));
    }
    void build() override {
        check(atg_build(table_, nullptr));
        check(atg_sync(table_, nullptr));
        atg_stats st;
        check(atg_get_stats(table_, &st));
        W = st.W;
        numL = st.numL;
        numH = st.numH;
        for (int k = 0; k < 5; ++k) phaseMs[k] = st.phase_ms[k];
    }
    void performSplit() { check(atg_build_phase(table_, ATG_PHASE_SPLIT, nullptr)); }
    void performPack() { check(atg_build_phase(table_, ATG_PHASE_PACK, nullptr)); }
    // postBuild + verifyTableCorrectness (AliasTable.cu:24-65); the host table
    // copy is filled as aliasTable.
    bool postBuild() override {
        int32_t ok = 0;
        check(atg_post_build(table_, 1, &maxRelError, &ok));
        aliasWeights.resize((size_t)N);
        aliases.resize((size_t)N);
        check(atg_get_table(table_, aliasWeights.data(), aliases.data(), ATG_MEM_HOST, nullptr));
        return ok != 0;
    }
    atg_table* handle() { return table_; }

    int64_t numL = 0, numH = 0;
    double phaseMs[5] = {0, 0, 0, 0, 0};  // sum, partition, prefixsum, split, pack (TableSplitTimer)
    double maxRelError = 0;
    std::vector<double> aliasWeights;
    std::vector<int32_t> aliases;

  private:
    Variant variant_;
    int64_t numThreads_;
    atg_table* table_ = nullptr;
};

// The fill of A picks the reference's subclass (RejectionSamplingDirect,
// RejectionSamplingBinarySearch, RejectionSamplingPArySearch) or the learned
// inverse-CDF index.
enum class RejectionFill { Direct = ATG_FILL_DIRECT, BinarySearch = ATG_FILL_BISECT, PArySearch = ATG_FILL_PARY,
                           Learned = ATG_FILL_LEARNED };

class RejectionSampling : public SamplingAlgorithm {
  public:
    RejectionSampling(int64_t size, WeightDistribution dist = weightDistributionSine,
                      RejectionFill fill = RejectionFill::Direct, int32_t knots = 0)
        : SamplingAlgorithm(size, dist), fill_(fill), knots_(knots) {}
    ~RejectionSampling() override {
        if (table_) atg_destroy(table_);
    }
    std::string name() override {
        switch (fill_) {
            case RejectionFill::BinarySearch: return "RejectionSamplingBinarySearch";
            case RejectionFill::PArySearch: return "RejectionSamplingPArySearch";
            case RejectionFill::Learned: return "RejectionSamplingLearned";
            default: return "RejectionSamplingDirect";
        }
    }
    void preBuild() override {
        if (!table_) check(atg_create(&table_, N, ATG_F64, ATG_AOS, nullptr));
        check(atg_set_weights(table_, weights.data(), ATG_MEM_HOST, nullptr));
    }
    void build() override {
        check(atg_rejection_build_fill(table_, (atg_fill_method)fill_, knots_, nullptr));
        check(atg_sync(table_, nullptr));
        check(atg_rejection_weight_sum(table_, &W));
    }
    // RejectionSampling::fillA (RejectionSampling.cuh:22): the fill alone.
    void fillA() { check(atg_rejection_fill_a(table_, nullptr)); }
    // postBuild copies A to the host (RejectionSampling.cu:19-24).
    bool postBuild() override {
        uint64_t n = 0;
        check(atg_rejection_size(table_, &n));
        ACpu.resize((size_t)n);
        check(atg_get_rejection_array(table_, ACpu.data(), n));
        return true;
    }
    // RejectionSampling::benchmarkBuild (RejectionSampling.cu:58-70): item 0
    // raised to weightOfOutlier, one build, then the fill alone `iterations`
    // times; milliseconds per fill.
    double benchmarkBuild(double weightOfOutlier, int iterations = 10) {
        weights[0] = weightOfOutlier;
        preBuild();
        build();
        hipDeviceSynchronize();
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < iterations; ++i) fillA();
        check(atg_sync(table_, nullptr));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        postBuild();
        return ms / iterations;
    }
    atg_table* handle() { return table_; }
    std::vector<int32_t> ACpu;

  private:
    RejectionFill fill_;
    int32_t knots_;
    atg_table* table_ = nullptr;
};

// Sampler (sampling/Sampler.cuh:10-31).
class Sampler {
  public:
    virtual ~Sampler() = default;
    virtual std::string name() = 0;
    // Giga samples per second (no output materialised, checksum only).
    double benchmarkSampling(uint64_t numSamples) {
        hipDeviceSynchronize();
        const auto t0 = std::chrono::steady_clock::now();
        uint64_t cs = 0;
        sample(numSamples, seed_++, nullptr, nullptr, &cs);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return (double)numSamples / s / 1e9;
    }
    // How many times each item was sampled (SamplerGpu.cu:9-24).
    std::vector<int> getSamplingDistribution(uint64_t numSamples) {
        const int64_t n = size();
        uint32_t* hist = nullptr;
        if (hipMalloc((void**)&hist, sizeof(uint32_t) * (size_t)n) != hipSuccess) throw Error(ATG_ENOMEM, "hist");
        hipMemset(hist, 0, sizeof(uint32_t) * (size_t)n);
        sample(numSamples, seed_++, nullptr, hist, nullptr);
        std::vector<int> out((size_t)n);
        hipMemcpy(out.data(), hist, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost);
        hipFree(hist);
        return out;
    }

  protected:
    virtual void sample(uint64_t count, uint64_t seed, int32_t* out, uint32_t* hist, uint64_t* checksum) = 0;
    virtual int64_t size() = 0;
    uint64_t seed_ = 1;
};

template <Storage S = Storage::ArrayOfStructs>
class SamplerGpuBasic : public Sampler {
  public:
    explicit SamplerGpuBasic(AliasTableSplitGpu<S>& t) : t_(t) {}
    std::string name() override { return "SamplerGpuBasic"; }

  protected:
    void sample(uint64_t c, uint64_t seed, int32_t* out, uint32_t* hist, uint64_t* cs) override {
        check(atg_sample(t_.handle(), ATG_SAMPLE_BASIC, c, seed, 0, out, hist, cs, nullptr));
        hipDeviceSynchronize();
    }
    int64_t size() override { return t_.N; }
    AliasTableSplitGpu<S>& t_;
};

template <Storage S = Storage::ArrayOfStructs>
class SamplerGpuSectioned : public Sampler {
  public:
    SamplerGpuSectioned(AliasTableSplitGpu<S>& t, int64_t itemsPerSection = 4096) : t_(t), rows_(itemsPerSection) {}
    std::string name() override { return "SamplerGpuSectioned"; }

  protected:
    void sample(uint64_t c, uint64_t seed, int32_t* out, uint32_t* hist, uint64_t* cs) override {
        check(atg_sample(t_.handle(), ATG_SAMPLE_SECTIONED, c, seed, rows_, out, hist, cs, nullptr));
        hipDeviceSynchronize();
    }
    int64_t size() override { return t_.N; }
    AliasTableSplitGpu<S>& t_;
    int64_t rows_;
};

class SamplerRejection : public Sampler {
  public:
    explicit SamplerRejection(RejectionSampling& r) : r_(r) {}
    std::string name() override { return "SamplerRejection"; }

  protected:
    void sample(uint64_t c, uint64_t seed, int32_t* out, uint32_t* hist, uint64_t* cs) override {
        check(atg_rejection_sample(r_.handle(), c, seed, out, hist, cs, nullptr));
        hipDeviceSynchronize();
    }
    int64_t size() override { return r_.N; }
    RejectionSampling& r_;
};

}  // namespace atg
