// A table build that allocates nothing (atg_build_noalloc): the same kernels
// and launch sequence as atg_create + atg_set_weights + atg_build, with every
// scratch buffer (device state, section sums and tree, dyadic block sums,
// staged leftovers, leftover lists and prefixes, split certificate, splits,
// knots) carved out of one file-scope device array instead of hipMalloc'd
// memory.  For callers that cannot allocate: code under HIP graph capture,
// tight serving loops.  The caller passes device weights and the device rows
// to fill.
//
// The array is one per device for the life of the process; its device
// address comes from hipGetSymbolAddress, taken once per device on the first
// call with that device current.  Every call on a device shares the array:
// calls that overlap on one device (two streams, two threads) race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <mutex>

#include "atg_build.h"
#include "atg_host.h"

namespace atg {

static_assert(ATG_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) unsigned char scratch[ATG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (the host-side name
// `scratch` is only a stand-in and never used as an address).
atg_status scratch_base(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ATG_ENOTSUP, "atg_build_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ATG_EDEVICE, "atg_build_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ATG_OK;
}

// Zeroes the device state (a kernel rather than a memset, so that a captured
// graph holds no memset node).
__global__ __launch_bounds__(64) void clear_state(DevState* st) {
    uint32_t* p = reinterpret_cast<uint32_t*>(st);
    for (int i = threadIdx.x; i < (int)(sizeof(DevState) / 4); i += blockDim.x) p[i] = 0u;
}

// Bump allocation inside the scratch array, 256-byte aligned.
struct Carve {
    char* base;
    size_t used;
    template <typename T>
    T* take(int64_t count) {
        T* p = reinterpret_cast<T*>(base + used);
        used += ((size_t)std::max<int64_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

}  // namespace
}  // namespace atg

using namespace atg;

extern "C" atg_status atg_build_noalloc(const void* weights, atg_dtype dtype, int64_t n, atg_layout layout,
                                        const atg_build_opts* opts, void* rows_or_weights, int32_t* aliases,
                                        double* W_out, uint32_t* error_out, void* stream) {
    if (!weights || !rows_or_weights) return fail(ATG_EINVAL, "atg_build_noalloc: NULL weights or rows");
    if (n < 1 || n > kMaxN) return fail(ATG_EINVAL, "atg_build_noalloc: n=%lld outside [1, %lld]", (long long)n, (long long)kMaxN);
    if (dtype != ATG_F64 && dtype != ATG_F32) return fail(ATG_EINVAL, "atg_build_noalloc: bad dtype");
    if (layout != ATG_AOS && layout != ATG_SOA) return fail(ATG_EINVAL, "atg_build_noalloc: bad layout");
    if (layout == ATG_SOA && !aliases) return fail(ATG_EINVAL, "atg_build_noalloc: SoA needs the alias array");
    atg_build_opts o;
    if (opts) o = *opts; else atg_default_opts(&o);
    if (o.interleave || o.record_phases)
        return fail(ATG_ENOTSUP, "atg_build_noalloc: interleave and record_phases need streams and events of their own");
    if (o.pack_method != ATG_PACK_CHUNKED && o.pack_method != ATG_PACK_BASIC && o.pack_method != ATG_PACK_NO_WEIGHTS)
        return fail(ATG_EINVAL, "atg_build_noalloc: unknown pack method %d", o.pack_method);
    if (o.pack_method == ATG_PACK_NO_WEIGHTS && o.psa_plus)
        return fail(ATG_EINVAL, "atg_build_noalloc: the NO_WEIGHTS pack needs psa_plus = 0");
    if (o.chains < 1 || o.chains > 64) return fail(ATG_EINVAL, "atg_build_noalloc: chains=%d outside [1, 64]", o.chains);
    if (o.split_method != ATG_SPLIT_BISECT && o.split_method != ATG_SPLIT_LEARNED && o.split_method != ATG_SPLIT_PARY)
        return fail(ATG_EINVAL, "atg_build_noalloc: unknown split method %d", o.split_method);
    // The shape exactly as atg_create derives it.
    if (o.learned_segments <= 0) o.learned_segments = 4096;
    if (o.section_size <= 0) o.section_size = std::min<int64_t>(1000, n / 10);
    if (o.section_size < 1) o.section_size = 1;
    if (o.num_splits < 0) o.num_splits = std::max<int64_t>(2, n / -o.num_splits);
    if (o.num_splits == 0) o.num_splits = 32LL * device_cus() * 10;
    int64_t G = 0, Ls = 0;
    section_layout(n, o.section_size, &G, &Ls);
    if (Ls > kMaxSectionLen)
        return fail(ATG_EINVAL, "atg_build_noalloc: section length %lld exceeds %lld", (long long)Ls, (long long)kMaxSectionLen);
    const int64_t p_req = std::max<int64_t>(2, std::min<int64_t>(o.num_splits, n / 10));
    const int64_t tree_half = (G + 63) / 64 + 1, tree_total = block_tree_total(G);

    void* base = nullptr;
    atg_status rc = scratch_base(&base);
    if (rc != ATG_OK) return rc;
    Carve c{static_cast<char*>(base), 0};
    BuildArgs a{};
    a.weights = weights;
    a.dtype_f32 = dtype == ATG_F32;
    a.n = n;
    a.G = G;
    a.Ls = Ls;
    a.Ls_max = Ls;
    a.tree_total = tree_total;
    a.st = c.take<DevState>(1);
    a.section_sums = c.take<double>(G);
    a.tree_scratch = c.take<double>(2 * tree_half);
    a.tree_scratch_half = tree_half;
    a.slots = c.take<uint64_t>(3 * tree_total + (tree_total + 1) / 2 + 1);
    a.sec_cert = c.take<uint64_t>(6 * G);
    a.want_cert = o.split_method != ATG_SPLIT_BISECT;
    a.stage_w = c.take<double>(n);
    a.stage_i = c.take<int32_t>(n);
    a.Li = c.take<int32_t>(n + 2);
    a.Lw = c.take<double>(n + 2);
    a.Lp = c.take<double>(n + 2);
    a.Hi = c.take<int32_t>(n + 2);
    a.Hw = c.take<double>(n + 2);
    a.Hp = c.take<double>(n + 2);
    a.splits = c.take<Split>(p_req + 1);
    a.knot_j = c.take<int32_t>(o.learned_segments + 2);
    if (c.used > (size_t)ATG_NOALLOC_SCRATCH_BYTES)
        return fail(ATG_EINVAL, "atg_build_noalloc: n=%lld with sections of %lld items needs %zu bytes of scratch, "
                    "more than the %u-byte maximum", (long long)n, (long long)Ls, c.used,
                    (unsigned)ATG_NOALLOC_SCRATCH_BYTES);
    if (layout == ATG_AOS) {
        a.rows = static_cast<Row*>(rows_or_weights);
    } else {
        a.tw = static_cast<double*>(rows_or_weights);
        a.ta = aliases;
    }
    a.psa_plus = o.psa_plus;
    a.chains = o.chains;
    a.p_req = p_req;
    a.pack_cap = pack_cap_items(n, p_req);
    a.learned_segments = learned_knots(o.learned_segments, p_req);
    a.pack_method = o.pack_method;

    hipStream_t s = (hipStream_t)stream;
    HIPCHK(set_partition_lds_limit(psa_lds_bytes(Ls)));
    hipLaunchKernelGGL(clear_state, dim3(1), dim3(64), 0, s, a.st);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_sum(a, s));
    HIPCHK(launch_partition(a, s));
    HIPCHK(o.split_method == ATG_SPLIT_LEARNED ? launch_split_learned(a, s)
           : o.split_method == ATG_SPLIT_PARY  ? launch_split_pary(a, s)
                                               : launch_split(a, s));
    HIPCHK(launch_pack(a, s));
    const char* st = reinterpret_cast<const char*>(a.st);
    if (W_out) HIPCHK(hipMemcpyAsync(W_out, st + offsetof(DevState, W), sizeof(double), hipMemcpyDeviceToDevice, s));
    if (error_out)
        HIPCHK(hipMemcpyAsync(error_out, st + offsetof(DevState, error), sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    return ATG_OK;
}
