// Host-visible argument blocks and launchers of the sampling kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "atg_types.h"

namespace atg {

// Philox stream purposes (word 2 of the counter).
enum : uint32_t {
    PURPOSE_BASIC = 0x0000B001u,
    PURPOSE_SECTIONED = 0x0000B002u,
    PURPOSE_SECTION_COUNTS = 0x0000B003u,
    PURPOSE_REJECTION = 0x0000B004u,
    PURPOSE_BATCH = 0x0000B005u,
    PURPOSE_REJ_SECTIONED = 0x0000B006u,
    PURPOSE_REJ_SECTION_COUNTS = 0x0000B007u,
};
#define REJECTION_MAX_TRIALS 256
// Trials of one sample of the sectioned rejection sampler: a section of
// low-acceptance slots can need many, but it draws few samples.
#define REJ_SECTIONED_MAX_TRIALS (1u << 20)

struct SampleArgs {
    const Row* rows;       // AoS table or null
    const double* tw;      // SoA table or null
    const int32_t* ta;
    const DevState* st;    // W/N of the build
    int64_t n;
    uint64_t count;
    uint64_t seed;
    int32_t* out;          // nullable
    uint32_t* hist;        // nullable
    unsigned long long* checksum;  // nullable (device)
    int64_t SS;            // sectioned: rows per section
    int64_t nsec;
    uint64_t* sec_counts;  // sectioned: 2 nsec values, count and first output of each section
};

// The rejection structure is scanned in blocks of REJ_ITEMS items of
// REJ_THREADS threads (block totals, their scan, then the per-block fills).
#define REJ_ITEMS 4096
#define REJ_THREADS 256

// Fills of A (RejectionSampling::fillA): the item runs written directly, a
// bisection per slot, a wave-wide p-ary narrowing then a bisection per slot,
// and the learned inverse-CDF index.
enum RejFill : int { REJ_FILL_DIRECT = 0, REJ_FILL_BISECT = 1, REJ_FILL_PARY = 2, REJ_FILL_LEARNED = 3 };

struct RejArgs {
    const void* weights;
    int dtype_f32;
    int64_t n;
    DevState* st;          // W (input), rej_size (output)
    int64_t nb;            // blocks of 4096 items
    uint64_t* block_tot;   // nb
    uint64_t* block_off;   // nb
    uint32_t* incl;        // search fills: n inclusive counts (|A| < 2^32 for the n they accept)
    int32_t* knot_items;   // learned fill: K + 1 knot items of the inverse-CDF model
    uint32_t* knot_err;    // learned fill: K per-segment error bounds
    int32_t K;             // learned fill: segments
    RejEntry* A;          // |A| <= A_cap slots
    uint64_t A_cap;
    unsigned long long* trials;  // nullable
    RejTask* tasks;        // long runs cut into tasks (rejection_task_cap(A_cap) entries)
    uint32_t tasks_cap;
};

// Sectioned rejection sampling over A: sections of SS slots.
struct RejSecArgs {
    const RejEntry* A;
    uint64_t A_size;       // |A|
    int64_t SS;            // slots per section
    int64_t nsec;
    double* mass;          // nsec section acceptance masses
    double* pre;           // nsec + 1: their exclusive prefix
    uint64_t* sec_counts;  // 2 nsec: count and first output of each section
};

hipError_t launch_sample_basic(const SampleArgs& a, hipStream_t s);
// masses: recompute the section masses and their prefix (a new A or section size).
hipError_t launch_rejection_sample_sectioned(const RejSecArgs& a, const SampleArgs& sa, bool masses, hipStream_t s);
hipError_t set_rej_sample_lds_limit(int bytes);
hipError_t launch_sample_sectioned(const SampleArgs& a, hipStream_t s);
hipError_t set_sample_lds_limit(int bytes);
// The counts, their scan and the fill (with the incl pass and the model fit
// the search fills need); launch_rejection_fill repeats the fill alone.
hipError_t launch_rejection_build(const RejArgs& a, int fill, hipStream_t s);
hipError_t launch_rejection_fill(const RejArgs& a, int fill, hipStream_t s);
// The inclusive counts (and with model, the learned index fitted on them),
// then a search fill (atg_rejfill.hip).
hipError_t launch_rejection_incl(const RejArgs& a, bool model, hipStream_t s);
hipError_t launch_rejection_search_fill(const RejArgs& a, int fill, hipStream_t s);
hipError_t launch_rejection_sample(const RejArgs& a, const SampleArgs& sa, hipStream_t s);
int64_t rejection_blocks(int64_t n);
uint32_t rejection_task_cap(uint64_t A_cap);

}  // namespace atg
