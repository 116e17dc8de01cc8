/*
 * atg: MI355X (gfx950) weighted-sampling engine - C ABI.
 *
 * Drop-in boundary for the hot path of ByteHamster/alias-table-gpu: alias-table
 * construction (W sum, light/heavy partition, PSA+, prefix sums, split search,
 * pack), sectioned/basic alias sampling and rejection sampling.  Plain pointers
 * and sizes only; every GPU call takes the caller's hipStream_t (passed as
 * void*; NULL = the default stream).  All functions return an atg_status; the
 * message of the last failure on the calling thread is atg_last_error().
 *
 * Each declaration names the reference interface it replaces (file:line in the
 * reference repository).
 */
#ifndef ATG_H
#define ATG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATG_ABI_VERSION 1

typedef enum {
    ATG_OK = 0,
    ATG_EINVAL = 1,   /* invalid argument or weights (NaN, inf, negative, W <= 0) */
    ATG_ENOMEM = 2,   /* device or host allocation failed */
    ATG_EDEVICE = 3,  /* HIP runtime error */
    ATG_ESTATE = 4,   /* call out of order (e.g. sample before build) */
    ATG_ENOTSUP = 5,  /* option not supported */
    ATG_ETIMEOUT = 6, /* a device-side bounded wait expired (internal error) */
    ATG_EIO = 7       /* file I/O failed, or a table file is damaged (bad magic, checksum) */
} atg_status;

typedef enum { ATG_F64 = 0, ATG_F32 = 1 } atg_dtype;     /* weight type; sums are always FP64 */
typedef enum { ATG_AOS = 0, ATG_SOA = 1 } atg_layout;    /* TableStorage.cuh:6-128 */
typedef enum { ATG_MEM_HOST = 0, ATG_MEM_DEVICE = 1 } atg_mem;

typedef enum {
    ATG_SPLIT_BISECT = 0,   /* SplitMethodBasic (SplitMethodBasic.cu:14-50) */
    ATG_SPLIT_LEARNED = 1,  /* piecewise-linear index + exact local search, bisection fallback */
    ATG_SPLIT_PARY = 2      /* SplitMethodParySearch (SplitMethodParySearch.cu:9-145): 16 lanes per split
                               probe 16 positions per step; bisection fallback; single tables only */
} atg_split_method;

/* Pack of the leftover lists into rows (Variant::pack and sharedMemory,
 * AliasTableSplitGpu.cuh:38-94, dispatched in performPack,
 * AliasTableSplitGpu.cu:506-591).  Every method writes the same rows. */
typedef enum {
    ATG_PACK_CHUNKED = 0,    /* one wave per 64 chains, their lists staged in LDS in rounds
                                (the role of PackMethodBasicShared / PackMethodChunkedShared); default */
    ATG_PACK_BASIC = 1,      /* one lane per chain, lists read from HBM (PackMethodBasic::packKernel) */
    ATG_PACK_NO_WEIGHTS = 2  /* as BASIC, each weight read through its item from the input weights
                                instead of the lists (PackMethodNoWeights::packKernel); needs
                                psa_plus = 0 (the reference turns PSA+ off for it,
                                AliasTableSplitGpu.cuh:85-93); single tables only */
} atg_pack_method;

typedef enum {
    ATG_SAMPLE_BASIC = 0,      /* SamplerGpuBasic (SamplerGpuBasic.cu:77-133) */
    ATG_SAMPLE_SECTIONED = 1   /* SamplerGpuSectioned(Shared) (SamplerGpuSectioned.cu:98-166) */
} atg_sample_kind;

/* Build options.  Replaces the reference's Variant + numThreads
 * (AliasTableSplitGpu.cuh:60-100) and the compile-time PSA+ constants
 * (PsaPlusNormal.cu:3, AliasTableSplitGpu.cu:356).  Zero-initialise and call
 * atg_default_opts() for the reference shapes. */
typedef struct {
    int32_t psa_plus;        /* 1: PSA+ section pre-pack (Variant::psaPlus) */
    int32_t chains;          /* PSA+ greedy chains per section (NUM_LOCAL_SPLITS = 16), 1..64 */
    int64_t section_size;    /* section parameter; 0 = reference: min(1000, N/10) */
    int64_t num_splits;      /* p (numThreads); 0 = reference AUTO for the chunked pack: 320 * #CU;
                                negative: one split per -num_splits items (per table in a batch) */
    int32_t split_method;    /* atg_split_method */
    int32_t learned_segments;/* knots of the learned split index: at most this many (default 4096),
                                and at most one per 16 splits (at least 64) */
    int32_t record_phases;   /* 1: hipEvents around sum/partition/prefixsum/split/pack */
    int32_t interleave;      /* 1: atg_build runs split and pack interleaved on two streams
                                (performSplitAndPackInterleaved, AliasTableSplitGpu.cu:220-262);
                                not with record_phases.  Same table either way. */
    int64_t batch_group_items; /* batch only: the tables are dealt into groups of about this many
                                items (largest table first, each to the lightest group) and the
                                phases run group after group, so a table's build latency is its
                                group's, not the whole batch's; <= 0: one group.  Same tables
                                either way. */
    int32_t pack_method;     /* atg_pack_method (Variant::pack + sharedMemory); same table for each */
    int32_t reserved0;       /* 0 */
} atg_build_opts;

typedef struct {
    double W;                /* canonical FP64 sum of the weights */
    double W_N;              /* W / N */
    int64_t n;
    int64_t numL, numH;      /* leftover lists incl. sentinels (AliasTableSplitGpu.cuh:126-127) */
    int64_t R;               /* leftover items entering split/pack: numL + numH - 2 */
    int64_t p;               /* effective number of splits after the reference's clamps */
    int64_t num_sections;    /* PSA+/sum sections */
    int64_t section_len;     /* items per section */
    int64_t split_fallbacks; /* bisections that never reached FOUND (bounded; reported) */
    int64_t learned_fallbacks; /* learned-index predictions that fell back to bisection */
    int32_t degenerate;      /* one leftover class empty: every leftover row is {w, self} */
    int32_t monotone;        /* split certificate max dL < min dH over the stored prefix sums:
                                1 held, 0 not, -1 not computed (only the learned split needs it) */
    double phase_ms[5];      /* sum, partition(+PSA+ and prefix sums), prefixsum(0: fused), split, pack */
} atg_stats;

typedef struct atg_table_s atg_table;

/* ---- construction: AliasTableSplitGpu<TableStorage> -------------------- */

/* Reference defaults: PSA+ on, 16 chains, section min(1000, N/10), p AUTO. */
void atg_default_opts(atg_build_opts* opts);

/* AliasTableSplitGpu(int size, Variant, int numThreads)
 * (AliasTableSplitGpu.cuh:96-100) + the allocations of preBuild
 * (AliasTableSplitGpu.cu:206-217).  n in [1, 2^31 - 16). */
atg_status atg_create(atg_table** out, int64_t n, atg_dtype dtype, atg_layout layout,
                      const atg_build_opts* opts);
void atg_destroy(atg_table* t);

/* SamplingAlgorithm::weights + the H2D copy of preBuild
 * (SamplingAlgorithm.cuh:40, AliasTableSplitGpu.cu:206-217).  Device pointers
 * are used in place (must stay valid until the next set_weights/destroy). */
atg_status atg_set_weights(atg_table* t, const void* weights, atg_mem mem, void* stream);

/* AliasTableSplitGpu::build (AliasTableSplitGpu.cu:388-468).  Enqueues every
 * phase on the stream with no host synchronisation.  Invalid weights are
 * reported by atg_sync / atg_get_stats. */
atg_status atg_build(atg_table* t, void* stream);

/* One phase of build, in order SUM, PARTITION (PSA+ with the fused leftover
 * prefix sums), SPLIT, PACK: the reference exposes performSplit/performPack
 * the same way (AliasTableSplitGpu.cuh:107-115) and times the phases with
 * TableSplitTimer (utils/Timer.cuh:17-45).  atg_build == all four in order. */
typedef enum {
    ATG_PHASE_SUM = 0, ATG_PHASE_PARTITION = 1, ATG_PHASE_SPLIT = 2, ATG_PHASE_PACK = 3,
    /* ATG_PHASE_PARTITION in its two passes, in this order (for per-kernel
     * timing): the local pass (partition, PSA+ chains, staged leftovers), then
     * the section prefixes and the placement of the leftover lists. */
    ATG_PHASE_PARTITION_LOCAL = 4, ATG_PHASE_PARTITION_PLACE = 5
} atg_phase;
atg_status atg_build_phase(atg_table* t, int32_t phase, void* stream);

/* Waits for the stream and returns the build's device-side status. */
atg_status atg_sync(atg_table* t, void* stream);

/* postBuild + verifyTableCorrectness (AliasTableSplitGpu.cu:593-612,
 * AliasTable.cu:24-65): copies the table to the host and checks the implied
 * per-item probabilities against the weights (host weights required).
 * max_rel_err: max |p_table - p_weight| / p_weight. */
atg_status atg_post_build(atg_table* t, int32_t verify, double* max_rel_err, int32_t* ok);

atg_status atg_get_stats(atg_table* t, atg_stats* st);

/* AliasTable::aliasTable (AliasTable.cuh:23).  Copies row weights and aliases
 * (SoA form for either layout) to host or device buffers of n entries. */
atg_status atg_get_table(atg_table* t, double* weights, int32_t* aliases, atg_mem mem, void* stream);

/* Device pointer to the table: AoS rows {double weight; int32 alias; int32 0}
 * (16 B) or SoA {double weights[n]; int32 aliases[n]}. */
atg_status atg_table_device_ptr(atg_table* t, void** rows_or_weights, void** aliases_or_null);

/* The splits array (AliasTableSplitGpu.cuh:137, SplitConfig.cuh:4-8): p + 1
 * entries {i, j, spill}.  Host buffers. */
atg_status atg_get_splits(atg_table* t, int32_t* i, int32_t* j, double* spill, int64_t capacity);

/* The leftover lists l, h and prefixWeightL/H (AliasTableSplitGpu.cuh:133-136):
 * numL / numH entries each (host buffers, capacity n + 2). */
atg_status atg_get_lh(atg_table* t, int32_t* L_item, double* L_w, double* L_pref, int32_t* H_item,
                      double* H_w, double* H_pref);

/* A build that allocates nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Builds the table of
 * n device weights (dtype) into the caller's device rows, exactly as
 * atg_create + atg_set_weights(ATG_MEM_DEVICE) + atg_build with the same
 * options build it: AoS rows {double weight; int32 alias; int32 0} at
 * rows_or_weights, or SoA weights at rows_or_weights and aliases at aliases.
 * W_out (device, nullable) receives W; error_out (device, nullable) the
 * device-side error bits (0: valid weights).  Every scratch buffer lives in
 * one per-device array of ATG_NOALLOC_SCRATCH_BYTES, held until the process
 * ends; shapes that need more (about n > 2^20 at the default sections) are
 * rejected with ATG_EINVAL.  interleave and record_phases are not supported.
 * Calls that overlap on one device race on that array. */
#define ATG_NOALLOC_SCRATCH_BYTES (64u << 20)
atg_status atg_build_noalloc(const void* weights, atg_dtype dtype, int64_t n, atg_layout layout,
                             const atg_build_opts* opts, void* rows_or_weights, int32_t* aliases,
                             double* W_out, uint32_t* error_out, void* stream);

/* ---- table files ------------------------------------------------------ */

/* Header of a versioned table file (no counterpart in the reference, whose
 * tables live in memory only: AliasTable.cuh:13-24, postBuild at
 * AliasTableSplitGpu.cu:599-612).  The file holds the header, the n row
 * weights (FP64) and the n aliases (int32); checksums guard both. */
typedef struct {
    int64_t n;
    double W;                /* canonical FP64 sum of the weights */
    int32_t layout;          /* atg_layout the table was built with */
    int32_t source_dtype;    /* atg_dtype of the weights it was built from */
    int32_t psa_plus;        /* build shape: PSA+, chains, section size, split method, effective p */
    int32_t chains;
    int32_t split_method;
    int32_t reserved;        /* 0 */
    int64_t section_size;
    int64_t num_splits;
    uint64_t seed;           /* the caller's tag, e.g. the generator seed of the weights; 0 if none */
    uint64_t checksum;       /* read: the rows' checksum stored in the file; ignored on write */
} atg_table_file_info;

/* Host-only: writes a table (weights and aliases of info->n rows) to path. */
atg_status atg_table_file_write(const char* path, const atg_table_file_info* info, const double* weights,
                                const int32_t* aliases);
/* Host-only: reads the header into *info and, when weights and aliases are
 * not NULL (capacity >= n entries each), the rows; verifies both checksums. */
atg_status atg_table_file_read(const char* path, atg_table_file_info* info, double* weights, int32_t* aliases,
                               int64_t capacity);
/* Copies a built table out (atg_get_table) and writes it with its build
 * shape and W. */
atg_status atg_save_table(atg_table* t, const char* path, uint64_t seed);
/* Loads rows into a created table of the same n for sampling (for a table
 * read from a file); W is the weight sum the rows were built for.  The table
 * counts as built until the next atg_build or atg_set_weights. */
atg_status atg_set_table(atg_table* t, const double* weights, const int32_t* aliases, double W, atg_mem mem,
                         void* stream);

/* ---- sampling: Sampler::benchmarkSampling / getSamplingDistribution ---- */

/* SamplerGpuBasic / SamplerGpuSectioned executeBenchmarkSampling
 * (SamplerGpuBasic.cu:109-133, SamplerGpuSectioned.cu:138-166) and
 * getSamplingDistribution (SamplerGpu.cu:9-24).  Draws `count` samples from the
 * built table with Philox4x32-10 (key = seed, counter = output position).
 * out (device, count int32, nullable) receives the items; hist (device, n
 * uint32, nullable) is incremented; checksum (host, nullable) gets the sum of
 * all sampled items.  section_rows: sectioned sampler section (0 = 4096).  A
 * sectioned call with more sections than any before (rows below 4096) grows
 * the table's per-section count buffer (after waiting for the stream). */
atg_status atg_sample(atg_table* t, atg_sample_kind kind, uint64_t count, uint64_t seed,
                      int64_t section_rows, int32_t* out, uint32_t* hist, uint64_t* checksum,
                      void* stream);

/* ---- rejection sampling: RejectionSampling + SamplerRejection --------- */

/* RejectionSampling::build (RejectionSampling.cu:35-56): builds
 * A (item i repeated floor(N w_i / W) + 1 times) from the table's weights and
 * canonical W.  Requires atg_set_weights (a table build is not needed). */
atg_status atg_rejection_build(atg_table* t, void* stream);

/* How A is filled once the counts are scanned (RejectionSampling::fillA,
 * RejectionSampling.cuh:22; the reference picks it by subclass).  Every fill
 * writes the same A.
 *   DIRECT   each item's run written by its wave, long runs cut into tasks
 *            (RejectionSamplingDirect.cu:12-28); the default
 *   BISECT   a bisection over the inclusive counts per slot
 *            (RejectionSamplingBinarySearch.cu:12-33)
 *   PARY     a wave-wide p-ary narrowing, then a bisection per slot
 *            (RejectionSamplingPArySearch.cu:12-123)
 *   LEARNED  a piecewise-linear inverse-CDF model (slot -> item) fitted on the
 *            inclusive counts, with exact per-segment error bounds; each
 *            prediction is corrected by a verified search inside its bound,
 *            a miss falls back to the bisection (no reference code) */
typedef enum { ATG_FILL_DIRECT = 0, ATG_FILL_BISECT = 1, ATG_FILL_PARY = 2, ATG_FILL_LEARNED = 3 } atg_fill_method;

/* RejectionSampling::build (RejectionSampling.cu:35-56) with the chosen fill;
 * knots = the learned model's segments (0 = 4096, at most 2^22; ignored by
 * the other fills).  The search fills keep the n inclusive counts (4 B each)
 * and accept n <= 0x7FFFF000 (|A| < 2^32). */
atg_status atg_rejection_build_fill(atg_table* t, atg_fill_method fill, int32_t knots, void* stream);
/* RejectionSampling::fillA alone (the loop of benchmarkBuild,
 * RejectionSampling.cu:58-70): repeats the last build's fill over its counts
 * (and model); A comes out the same. */
atg_status atg_rejection_fill_a(atg_table* t, void* stream);
/* Waves of the learned fill whose prediction missed and fell back to the
 * bisection, summed over the fills since the last build. */
atg_status atg_rejection_fill_fallbacks(atg_table* t, uint64_t* fallbacks);
/* The learned model of the last build (its fill LEARNED with `knots`
 * segments): knot_items[knots + 1] and err_bounds[knots], host arrays. */
atg_status atg_rejection_get_model(atg_table* t, int32_t* knot_items, uint32_t* err_bounds, int32_t knots);
/* Replaces that model (same knots; knot items non-decreasing in [0, n-1],
 * any bounds) for the next atg_rejection_fill_a: a model with wrong bounds
 * changes the fallback count, never A.  For tests of the correction. */
atg_status atg_rejection_set_model(atg_table* t, const int32_t* knot_items, const uint32_t* err_bounds,
                                   int32_t knots);
/* The acceptance value of every slot of A (frac(N w / W) at the first slot
 * of a run, 1.0 elsewhere), host array of capacity >= |A|. */
atg_status atg_get_rejection_accept(atg_table* t, double* accept, uint64_t capacity);

atg_status atg_rejection_size(atg_table* t, uint64_t* a_size);
/* RejectionSampling::W (SamplingAlgorithm.cuh:44): the canonical weight sum
 * the rejection structure was built with (kept apart from the alias build's). */
atg_status atg_rejection_weight_sum(atg_table* t, double* W);
atg_status atg_get_rejection_array(atg_table* t, int32_t* A_host, uint64_t capacity);

/* SamplerRejection::executeBenchmarkSampling (SamplerRejection.cu:63-72). */
atg_status atg_rejection_sample(atg_table* t, uint64_t count, uint64_t seed, int32_t* out,
                                uint32_t* hist, uint64_t* checksum, void* stream);

/* Rejection sampling by sections of A (the trials of SamplerRejection.cu:17-35
 * with the sections of SamplerGpuSectioned.cu:58-136; the reference has no
 * sectioned rejection sampler).  A is cut into sections of section_slots
 * slots (0 = 4096, at most 8192), each staged in LDS; a section's sample count
 * follows its acceptance mass (halving tree of binomial draws, exact total);
 * inside a section a sample repeats trials on uniform slots with the
 * reference's acceptance rule.  Samples are written in section order.  The
 * first call after atg_rejection_build reads |A| back (a wait for the
 * stream); a call with more sections than any before grows the section
 * buffers. */
atg_status atg_rejection_sample_sectioned(atg_table* t, uint64_t count, uint64_t seed, int64_t section_slots,
                                          int32_t* out, uint32_t* hist, uint64_t* checksum, void* stream);

/* ---- input synthesis (SamplingAlgorithm::generateWeights, SamplingAlgorithm.cu:21-71) */

typedef enum {
    ATG_DIST_SINE = 1, ATG_DIST_BEST_WORST = 2, ATG_DIST_UNIFORM = 3, ATG_DIST_RAMP = 4,
    ATG_DIST_POWERLAW1 = 5, ATG_DIST_POWERLAW2 = 6, ATG_DIST_ONE_HEAVY = 7, ATG_DIST_ONE_LIGHT = 8,
    ATG_DIST_POWERLAW1_SHUFFLED = 9, ATG_DIST_POWERLAW05_SHUFFLED = 10,
    /* workloads of the benchmark configurations */
    ATG_DIST_UNIT_UNIFORM = 100,  /* w = 1 - u in (0, 1] */
    ATG_DIST_PARETO = 101,        /* w = (1 - u)^(-1/a), a = param (1.1) */
    ATG_DIST_EXPONENTIAL = 102,   /* w = -ln(1 - u) */
    ATG_DIST_HEAVY_MIX = 103,     /* 1% heavy lognormal(10,1) rescaled to carry param (0.9) of the mass, light U(0,1) */
    ATG_DIST_ZIPF_SHUFFLED = 104  /* w_r = r^-param shuffled */
} atg_distribution;

/* Host FP64 weights from a Philox4x32-10 stream (key = seed). */
atg_status atg_generate_weights(int32_t dist, int64_t n, uint64_t seed, double param, double* out);
/* Host weights of a batch of tables: sizes log-uniform over [lo, hi] into
 * offsets[0..ntables]; with out != NULL also the weights of every table
 * (table t from seed + t + 1), concatenated, as FP64 or FP32 (out_dtype).
 * Workload synthesis; no counterpart in the reference. */
atg_status atg_generate_batch(int32_t dist, int64_t ntables, int64_t lo, int64_t hi, uint64_t seed,
                              double param, int64_t* offsets, void* out, int32_t out_dtype);

/* ---- batched build: many independent tables ------------------------------ */

typedef struct atg_batch_s atg_batch;

/* Tables t = 0..ntables-1 own weights[offsets[t], offsets[t+1]).  Each table is
 * built exactly as atg_build would build it alone with the same options. */
atg_status atg_batch_create(atg_batch** out, const int64_t* offsets_host, int64_t ntables,
                            atg_dtype dtype, const atg_build_opts* opts);
void atg_batch_destroy(atg_batch* b);
atg_status atg_batch_set_weights(atg_batch* b, const void* weights, atg_mem mem, void* stream);
atg_status atg_batch_build(atg_batch* b, void* stream);
atg_status atg_batch_sync(atg_batch* b, void* stream);
/* Rows (AoS, 16 B) of all tables, concatenated; per-table W. */
atg_status atg_batch_get_tables(atg_batch* b, double* weights, int32_t* aliases, double* W_per_table);
/* Rows and W of table t alone (host buffers of offsets[t+1] - offsets[t]
 * entries; any pointer may be NULL): AliasTable::aliasTable of one table
 * (AliasTable.cuh:23) without copying the whole batch. */
atg_status atg_batch_get_table(atg_batch* b, int64_t t, double* weights, int32_t* aliases, double* W);
/* The number of groups the tables were dealt into (batch_group_items) and,
 * when group_of_table is not NULL, the group of every table (ntables entries).
 * Batch bookkeeping; no counterpart in the reference. */
atg_status atg_batch_groups(atg_batch* b, int32_t* group_of_table, int64_t* ngroups);
/* enable = 1: captures the build's launch sequence (kernels only) into a
 * hipGraph that atg_batch_build replays from then on; 0: back to direct
 * launches.  The graph stays valid across atg_batch_set_weights.  Replaces
 * the reference's per-phase launches of one table (AliasTableSplitGpu.cu:388-468)
 * for many tables at once. */
atg_status atg_batch_use_graph(atg_batch* b, int32_t enable);
/* `per_table` samples from every table; out[t * per_table + k] = item index
 * local to table t. */
atg_status atg_batch_sample(atg_batch* b, uint64_t per_table, uint64_t seed, int32_t* out,
                            void* stream);
/* Per-table build latency in ns from device timestamps (last build). */
atg_status atg_batch_latencies(atg_batch* b, double* ns_per_table);

/* ---- autotuner ----------------------------------------------------------- */

/* Chooses the knobs that leave the table unchanged for these weights and the
 * build shape in *base (section_size, chains, num_splits, PSA+; NULL: the
 * defaults): the split method, the learned index's knot count and the
 * split/pack interleave.  Times each candidate (median of `reps` builds after
 * a warm-up) and writes *base with the fastest knobs to *best, so every
 * candidate builds the same table.  *best_ms: the median build time, negative
 * when the choice came from the ATG_TUNE_CACHE file.  No counterpart in the
 * reference (its split and pack methods are compile-time choices,
 * AliasTableSplitGpu.cuh:60-100). */
atg_status atg_autotune(const void* weights, atg_mem mem, int64_t n, atg_dtype dtype, atg_layout layout,
                        const atg_build_opts* base, int32_t reps, atg_build_opts* best, double* best_ms);
/* Searches the build shape: section_size x chains x num_splits over a small
 * grid, the fastest median build wins.  The shape changes the table (W's
 * association, the leftovers and every row), so the result is meant as an
 * explicit input to atg_create, chosen once and recorded (atg_save_table
 * stores it); builds with one shape are deterministic.  No counterpart in the
 * reference (its shape is compile-time, PsaPlusNormal.cu:3,
 * AliasTableSplitGpu.cu:58-71,356). */
atg_status atg_autotune_shape(const void* weights, atg_mem mem, int64_t n, atg_dtype dtype, atg_layout layout,
                              const atg_build_opts* base, int32_t reps, atg_build_opts* best, double* best_ms);
/* Tunes a batch in place over the knobs that leave its tables unchanged: the
 * launch groups (batch_group_items) and, with tune_split = 1, the split
 * method (0 keeps the batch's).  Among the
 * candidates whose median makespan is within (1 + makespan_slack) of the
 * fastest, the one with the lowest 99th-percentile table latency wins and
 * stays configured; *chosen receives the batch's options.  Weights must be
 * set.  No counterpart in the reference. */
atg_status atg_batch_autotune(atg_batch* b, int32_t reps, int32_t tune_split, double makespan_slack,
                              atg_build_opts* chosen, double* makespan_ms, double* p99_latency_ms);

/* ---- misc --------------------------------------------------------------- */
const char* atg_last_error(void);
const char* atg_version(void);
int32_t atg_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* ATG_H */
