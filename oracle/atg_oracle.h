/*
 * CPU oracle for the alias-table hot path.  TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library; the product path never does.
 *
 * It restates, sequentially and in plain C++, the reference algorithms of
 * ByteHamster/alias-table-gpu under the canonical determinism contract that the
 * GPU kernels implement (fixed summation order, stable partitions, dyadic
 * section prefixes, bounded bisection, Philox4x32-10 streams).  Every function
 * cites the reference file:line it follows.
 *
 * Parity pins (each function is checked by tests/test_oracle.py against an
 * independent fact, not against a retyped copy of itself):
 *   philox          Random123 known-answer vectors + rocRAND's open-source
 *                   philox4x32_10 engine (tests/golden/philox_xcheck.cpp)
 *   weight_sum      closed form sum_{i<n} i (Adder.cu:159-171); integer weights
 *                   (every association exact) against a sequential sum
 *   local scan      `i % 7` closed form (PrefixSum.cu:265-296); integer weights
 *   build (PSA)     integer/dyadic weights: identical to the faithful restatement
 *                   of AliasTableSplit::build (AliasTableSplit.cu:24-80) whose sums
 *                   are then order-free; exact implied-mass check; closed-form
 *                   tables (all-equal, OneHeavy, OneLight, BestWorst);
 *                   brute-force split enumeration on tiny inputs
 *   PSA+            same table checks; PSA+ off reproduces the plain path
 *   samplers        chi-squared against the weights; basic == sectioned with a
 *                   single section over the same stream shape; rejection A
 *                   brute force
 * Functions with no such pin are marked "parity unpinned" below; there are none
 * on the build path.
 */
#ifndef ATG_ORACLE_H
#define ATG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference: EPSILON, construction/SamplingAlgorithm.cuh:14 */
#define ORACLE_EPSILON 1e-10

/* Philox stream purposes (word 2 of the counter). */
#define ORACLE_PURPOSE_BASIC 0x0000B001u
#define ORACLE_PURPOSE_SECTIONED 0x0000B002u
#define ORACLE_PURPOSE_SECTION_COUNTS 0x0000B003u
#define ORACLE_PURPOSE_REJECTION 0x0000B004u
#define ORACLE_PURPOSE_BATCH 0x0000B005u
#define ORACLE_PURPOSE_REJ_SECTIONED 0x0000B006u
#define ORACLE_PURPOSE_REJ_SECTION_COUNTS 0x0000B007u
#define ORACLE_REJ_SECTIONED_MAX_TRIALS (1u << 20)
#define ORACLE_REJECTION_MAX_TRIALS 256

typedef struct {
    int32_t psa_plus;      /* 1: PSA+ pre-pack (PsaPlusNormal.cu); 0: plain stable partition */
    int32_t chains;        /* PSA+ greedy chains per section (NUM_LOCAL_SPLITS, PsaPlusNormal.cu:3) */
    int64_t section_size;  /* section parameter (AliasTableSplitGpu.cu:356, sectionSize) */
    int64_t num_splits;    /* p requested (numThreads) before the reference's clamps */
    int32_t threads;       /* CPU threads for the parallel restatement (1 = sequential) */
} oracle_opts;

typedef struct {
    double W, W_N;
    int64_t num_sections, section_len;
    int64_t numL, numH;        /* leftover lists incl. sentinels */
    int64_t R;                 /* numL + numH - 2 (leftover items entering split/pack) */
    int64_t p;                 /* effective number of splits after clamps */
    double wToSplit;
    int64_t nToSplit;
    int64_t split_fallbacks;   /* bisections that never reached FOUND (bounded) */
    int32_t degenerate;        /* 1: one leftover class empty -> self rows */
    int32_t monotone;          /* 1: max dL < min dH (split certificate holds) */
    double max_dL, min_dH;
} oracle_stats;

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

void oracle_section_layout(int64_t n, int64_t section_param, int64_t* num_sections,
                           int64_t* section_len);
double oracle_section_sum(const double* x, int64_t len);
double oracle_tree64_reduce(const double* v, int64_t count);
double oracle_weight_sum(const double* w, int64_t n, int64_t section_param,
                         double* section_sums /* nullable, num_sections */);
/* Canonical exclusive scan of one section's list; returns the aggregate. */
double oracle_local_scan(const double* v, int64_t m, double* excl /* nullable */);

/* Full canonical build.  Table is written as SoA (row_w, row_alias, n rows).
 * The L/H/prefix outputs need capacity n + 2, splits capacity num_splits + 1;
 * all intermediate outputs are nullable. Returns 0 on success, 1 on invalid input. */
int oracle_build(const double* w, int64_t n, const oracle_opts* o,
                 double* row_w, int32_t* row_alias,
                 int32_t* L_item, double* L_w, double* L_pref,
                 int32_t* H_item, double* H_w, double* H_pref,
                 int32_t* split_i, int32_t* split_j, double* split_spill,
                 oracle_stats* st);

/* One bisection split (SplitMethodBasic.cu:22-49) on given prefix arrays. */
int oracle_split_one(int64_t k, int64_t p, int64_t nToSplit, double wToSplit,
                     const double* L_pref, int64_t numL, const double* H_pref,
                     const double* H_w, int64_t numH,
                     int32_t* out_i, int32_t* out_j, double* out_spill);

/* Implied-probability check (AliasTable.cu:24-65, exact accumulation). */
void oracle_verify_table(const double* w, int64_t n, double W, double W_N,
                         const double* row_w, const int32_t* row_alias,
                         double* max_rel_err, double* max_abs_mass_err,
                         int64_t* rows_out_of_range, int32_t* reference_check_ok);

/* Samplers (SamplerGpuBasic.cu:148-166, SamplerGpuSectioned.cu:19-136,
 * SamplerRejection.cu:17-35) on Philox streams. */
void oracle_sample_basic(const double* row_w, const int32_t* row_alias, int64_t n,
                         double W_N, uint64_t seed, uint64_t first, uint64_t count,
                         int32_t* out);
void oracle_sample_batch_table(const double* row_w, const int32_t* row_alias, int64_t n,
                               double W_N, uint64_t seed, uint64_t first, uint64_t count,
                               int32_t* out);
uint64_t oracle_binomial(uint64_t n, double q, uint64_t seed, uint64_t node);
void oracle_section_counts(int64_t n, int64_t section_rows, uint64_t total, uint64_t seed,
                           uint64_t* counts /* ceil(n/section_rows) */);
void oracle_sample_sectioned(const double* row_w, const int32_t* row_alias, int64_t n,
                             double W_N, int64_t section_rows, uint64_t total,
                             uint64_t seed, int32_t* out /* total */);
int oracle_rejection_build(const double* w, int64_t n, double W, uint64_t* incl_counts,
                           int32_t* A /* nullable */, uint64_t* A_size);
void oracle_sample_rejection(const int32_t* A, uint64_t A_size, const double* w, int64_t n,
                             double W, uint64_t seed, uint64_t first, uint64_t count,
                             int32_t* out, uint64_t* total_trials);

/* Faithful single-thread restatements of the reference CPU builds (baselines and
 * order-free parity pins).  They take an array of n + 2 weights (the two sentinel
 * slots are overwritten) and write AoS-free SoA rows. */
void oracle_ref_split_build(double* w_n2, int64_t n, int64_t p, double* W_out,
                            double* row_w, int32_t* row_alias);
void oracle_ref_sweep_build(double* w_n2, int64_t n, double* W_out, double* row_w,
                            int32_t* row_alias);
void oracle_ref_stack_build(const double* w, int64_t n, double* W_out, double* row_w,
                            int32_t* row_alias);

/* Batches of independent tables (the engine's atg_batch_*): see atg_oracle.cpp. */
int oracle_build_batch(const void* w, int32_t w_f32, const int64_t* offsets, int64_t ntables,
                       const int64_t* which /* nullable: all */, int64_t nwhich, const oracle_opts* o,
                       double* row_w, int32_t* row_alias, double* W_out /* nullable, ntables */,
                       int32_t threads);
void oracle_check_batch(const void* w, int32_t w_f32, const int64_t* offsets, int64_t ntables,
                        const oracle_opts* o, const double* row_w, const int32_t* row_alias,
                        const double* W_table, double* W_canonical, int32_t* ok, double* max_abs,
                        int32_t threads);
void oracle_sample_sectioned_subset(const double* row_w, const int32_t* row_alias, int64_t n, double W_N,
                                    int64_t section_rows, uint64_t total, uint64_t seed,
                                    const int64_t* sections, int64_t nsel, int32_t* out, uint64_t* first,
                                    uint64_t* cnt);
void oracle_set_threads(int32_t threads);
/* SplitMethodInverse (test-only second brute force of the split). */
void oracle_split_inverse(int64_t p, int64_t nToSplit, double wToSplit, const double* Lp, int64_t numL,
                          const double* Hp, const double* Hw, int64_t numH, int32_t* split_i, int32_t* split_j,
                          double* split_spill);
/* Sectioned rejection sampling (see atg_oracle.cpp). */
void oracle_rejection_section_prefix(const int32_t* A, uint64_t A_size, const double* w, int64_t n, double W,
                                     int64_t section_slots, double* prefix /* nsec + 1 */);
void oracle_rejection_section_counts(const double* prefix, int64_t nsec, uint64_t total, uint64_t seed,
                                     uint64_t* counts);
void oracle_sample_rejection_sectioned(const int32_t* A, uint64_t A_size, const double* w, int64_t n, double W,
                                       int64_t section_slots, uint64_t total, uint64_t seed,
                                       const int64_t* sections /* nullable: all */, int64_t nsel,
                                       int32_t* out /* total */, uint64_t* first_out, uint64_t* cnt_out);
/* Search fills of A and the learned inverse-CDF model (see atg_oracle.cpp). */
void oracle_rejection_fill_bisect(const uint64_t* incl, int64_t n, uint64_t A_size, int32_t* A);
void oracle_rejection_model(const uint64_t* incl, int64_t n, uint64_t A_size, int32_t K,
                            int32_t* knot_items /* K + 1 */, uint32_t* err_bounds /* K */);

#ifdef __cplusplus
}
#endif
#endif
