// Batched construction of many independent tables (placeholder until the
// persistent batched kernels land).
#include <hip/hip_runtime.h>

#include "../../include/atg.h"

extern "C" {
atg_status atg_batch_create(atg_batch** out, const int64_t*, int64_t, atg_dtype, const atg_build_opts*) {
    if (out) *out = nullptr;
    return ATG_ENOTSUP;
}
void atg_batch_destroy(atg_batch*) {}
atg_status atg_batch_set_weights(atg_batch*, const void*, atg_mem, void*) { return ATG_ENOTSUP; }
atg_status atg_batch_build(atg_batch*, void*) { return ATG_ENOTSUP; }
atg_status atg_batch_sync(atg_batch*, void*) { return ATG_ENOTSUP; }
atg_status atg_batch_get_tables(atg_batch*, double*, int32_t*, double*) { return ATG_ENOTSUP; }
atg_status atg_batch_sample(atg_batch*, uint64_t, uint64_t, int32_t*, void*) { return ATG_ENOTSUP; }
atg_status atg_batch_latencies(atg_batch*, double*) { return ATG_ENOTSUP; }
}
