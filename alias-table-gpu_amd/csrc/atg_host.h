// Host-side helpers shared by the C ABI translation units: error reporting,
// device queries and the limits of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/atg.h"
#include "atg_types.h"

namespace atg {

// Records a thread-local message (atg_last_error) and returns `st`.
atg_status fail(atg_status st, const char* fmt, ...);
// Compute units of the current device (256 on MI355X).
int device_cus();
// Maps DevState::error bits to a status and message.
atg_status state_status(const DevState& st);
// Sets the chain count and requested splits of a created table in place.
atg_status table_set_shape(atg_table* t, int32_t chains, int64_t num_splits);
// Sets the split method, knot count and split/pack interleave of a created
// table (knobs that leave the table's rows unchanged).
atg_status table_set_method(atg_table* t, int32_t split_method, int32_t learned_segments, int32_t interleave);

constexpr int64_t kMaxSectionLen = 2048;
constexpr int kInterleaveChunks = 8;  // split/pack interleave chunks
constexpr int64_t kMaxN = 0x7FFFFFF0LL;

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

}  // namespace atg

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::atg::fail(e_ == hipErrorOutOfMemory ? ATG_ENOMEM : ATG_EDEVICE, "%s: %s (%s:%d)", \
                               #call, hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)
