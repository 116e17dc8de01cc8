// Types shared by the host orchestration and the gfx950 kernels.
#pragma once

#include <stdint.h>

namespace atg {

// Error bits reported through DevState::error.
enum : uint32_t {
    ERR_BAD_WEIGHT = 1u,      // NaN, inf or negative weight
    ERR_BAD_SUM = 2u,         // W not positive / not finite
    ERR_LOOKBACK_TIMEOUT = 4u,
    ERR_SECTION_TOO_LONG = 8u,
};

// Build state kept in device memory so that every phase runs without a host
// round trip (the reference copies numL/numH/W/totals to the host mid-build,
// AliasTableSplitGpu.cu:293,330,361-362,476-479).
struct DevState {
    double W;
    double W_N;
    uint32_t error;
    uint32_t pad2;
    uint32_t degenerate;
    uint32_t cert_done;       // the split certificate was computed (learned split)
    int64_t numL, numH;       // incl. sentinels
    int64_t p;                // effective splits
    int64_t nToSplit;
    double wToSplit;
    uint64_t cert_max_dL;     // bit pattern of max light increment (non-negative double)
    uint64_t cert_min_dH;     // bit pattern of min heavy increment
    uint64_t split_fallbacks;
    uint64_t learned_fallbacks;
    uint64_t rej_size;        // |A| of the rejection structure
    uint32_t rej_tasks;       // long-run tasks of the rejection fill
    uint32_t pad1;
};

// Rows of the AoS table: 16 bytes (TableStorage.cuh:8-11 TableRow plus explicit padding).
struct alignas(16) Row {
    double weight;
    int32_t alias;
    int32_t pad;
};

// One slot of the rejection array A (RejectionSampling.cu:26-56): the item and
// the acceptance threshold of the slot: frac(N w / W) for the first slot of
// the item's run, 1.0 (always accept) for the others.
// A run of A slots [first, first + len) that all hold `item` with acceptance
// 1.0: part of a long run, written by the whole grid.
struct alignas(16) RejTask {
    uint64_t first;
    uint32_t len;
    int32_t item;
};

struct alignas(16) RejEntry {
    double accept;
    int32_t item;
    int32_t pad;
};

struct Split {
    int32_t i, j;
    double spill;
};

}  // namespace atg
