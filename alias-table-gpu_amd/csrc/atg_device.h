// Device-side building blocks shared by the gfx950 kernels: Philox4x32-10,
// 64-lane wavefront primitives for FP64, the canonical pairwise tree and
// Kogge-Stone scan, and agent-scope self-validating slots for look-back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "atg_types.h"

#define ATG_WAVE 64
#define ATG_EPSILON 1e-10  // construction/SamplingAlgorithm.cuh:14

// A published 64-bit slot holds this pattern until its owner writes it.  Sums
// of validated (finite, non-negative) weights and packed counts (< 2^31 per
// half) never produce it.
#define ATG_NOT_READY 0xFFFFFFFFFFFFFFFFull

namespace atg {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Random123 method).  Counter = (idx_lo, idx_hi, purpose, extra),
// key = (seed_lo, seed_hi).
struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // One 32x32 -> 64-bit product (v_mad_u64_u32) gives both halves:
        // half the multiply instructions of a separate mul_lo / mul_hi pair.
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

__device__ __forceinline__ u32x4 philox_at(uint64_t idx, uint32_t purpose, uint32_t extra, uint64_t seed) {
    return philox4x32_10(u32x4{(uint32_t)idx, (uint32_t)(idx >> 32), purpose, extra}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
}

__device__ __forceinline__ uint64_t word64(u32x4 r) { return ((uint64_t)r.y << 32) | r.x; }
__device__ __forceinline__ double coin53(u32x4 r) {
    const uint64_t v = ((uint64_t)r.w << 32) | r.z;
    return (double)(v >> 11) * 0x1.0p-53;
}
// floor(x * n / 2^64): a uniform row in [0, n) without the 2^-32 bias of a
// 32-bit multiply-shift.
__device__ __forceinline__ uint64_t scale64(uint64_t x, uint64_t n) { return __umul64hi(x, n); }

// ---------------------------------------------------------------------------
// Wavefront primitives (wave64).
__device__ __forceinline__ int lane_id() { return __lane_id(); }
// The wave's index in its block, as a scalar: values derived from it (a
// section per wave) stay uniform, so tests on them are scalar branches.
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, ATG_WAVE); }
__device__ __forceinline__ double shfl_up_d(double v, int d) { return __shfl_up(v, d, ATG_WAVE); }
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, ATG_WAVE); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

// DPP move of a double (both 32-bit halves with the same control).  Lanes the
// control leaves without a source (bound_ctrl) or masks out (row_mask) read 0.
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ __forceinline__ double dpp_d(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, ROW_MASK, 0xF, BOUND);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROW_MASK, 0xF, BOUND);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inclusive scan over the wave in the canonical order: Kogge-Stone offsets
// 1, 2, 4, 8 inside rows of 16 lanes (row_shr), rows 1 and 3 add the last lane
// of the row before (row_bcast:15), rows 2 and 3 add lane 31 (row_bcast:31).
// Lane 63 holds the pairwise tree of the 64 inputs.
__device__ __forceinline__ double wave_scan(double v) {
    v = v + dpp_d<0x111, 0xF, true>(v);
    v = v + dpp_d<0x112, 0xF, true>(v);
    v = v + dpp_d<0x114, 0xF, true>(v);
    v = v + dpp_d<0x118, 0xF, true>(v);
    v = v + dpp_d<0x142, 0xA, false>(v);
    v = v + dpp_d<0x143, 0xC, false>(v);
    return v;
}
// Lane l reads lane l ^ XOR (XOR < 32) of a double: ds_swizzle in bit mode
// (and-mask 0x1F, xor-mask XOR, lanes in groups of 32).
template <int XOR>
__device__ __forceinline__ double swizzle_xor_d(double v) {
    constexpr int kPattern = 0x1F | (XOR << 10);
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)u, kPattern);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(u >> 32), kPattern);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The pairwise-tree sums of eight vectors at once, each with the association
// of lane 63 of wave_scan (level b adds the two sibling blocks of 2^(b-1)
// lanes).  A butterfly over the lane bits: in levels 1-3 bit b-1 of the lane
// picks which of two vectors (or pairs, or quads) a lane keeps summing, so
// one register carries all eight from level 3 on.  Lane l returns the sum of
// vector l & 7.
__device__ __forceinline__ double wave_tree_sum8(const double (&v)[8], int lane) {
    const bool b0 = lane & 1, b1 = (lane >> 1) & 1, b2 = (lane >> 2) & 1;
    double r1[4], r2[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // lane l ^ 1 by quad_perm [1,0,3,2]
        const double x = b0 ? v[2 * q + 1] : v[2 * q], y = b0 ? v[2 * q] : v[2 * q + 1];
        r1[q] = x + dpp_d<0xB1, 0xF, false>(y);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {  // lane l ^ 2 by quad_perm [2,3,0,1]
        const double x = b1 ? r1[2 * q + 1] : r1[2 * q], y = b1 ? r1[2 * q] : r1[2 * q + 1];
        r2[q] = x + dpp_d<0x4E, 0xF, false>(y);
    }
    const double x = b2 ? r2[1] : r2[0], y = b2 ? r2[0] : r2[1];
    double r = x + swizzle_xor_d<4>(y);
    r = r + dpp_d<0x128, 0xF, false>(r);  // lane l ^ 8: row_ror:8
    r = r + swizzle_xor_d<16>(r);
    r = r + shfl_xor_d(r, 32);
    return r;
}

// Value of the previous lane (0 for lane 0): wave_shr:1.
__device__ __forceinline__ double wave_shr1(double v) { return dpp_d<0x138, 0xF, true>(v); }

__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Pairwise tree over the 64 lanes (t[m] = t[m] + t[m+w], w = 1, 2, ..., 32),
// returned to every lane.
__device__ __forceinline__ double wave_tree_sum(double v) { return readlane_d(wave_scan(v), 63); }

// Inclusive scan of 32-bit integers over the wave, with the DPP steps of
// wave_scan (integer sums are exact in any order).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum_int(T v) {
#pragma unroll
    for (int m = 1; m < ATG_WAVE; m <<= 1) v += __shfl_xor(v, m, ATG_WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_excl_scan_int(T v) {
    const int l = lane_id();
    T inc = v;
#pragma unroll
    for (int off = 1; off < ATG_WAVE; off <<= 1) {
        const T o = __shfl_up(inc, off, ATG_WAVE);
        if (l >= off) inc += o;
    }
    return inc - v;
}

// ---------------------------------------------------------------------------
// Agent-scope self-validating slots.  Each slot is written exactly once per
// build by its owner; readers poll until the value differs from ATG_NOT_READY.
// Relaxed agent-scope atomics compile to sc1 loads/stores on gfx950, which are
// coherent across the XCDs without an L2 writeback.
__device__ __forceinline__ void slot_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t slot_load(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t d2u(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ double u2d(uint64_t u) { return __longlong_as_double((long long)u); }

// ---------------------------------------------------------------------------
// Sampling decision of doSample (sampling/aliasTable/SamplerGpuBasic.cu:148-166):
// alias iff row.w < W/N and coin * W/N >= row.w.  An alias that is not an item
// (the pack's heavy sentinel N) keeps the row.
__device__ __forceinline__ int32_t decide(double rw, int32_t alias, int64_t row, double u, double W_N,
                                          int64_t n) {
    return (rw < W_N && u * W_N >= rw && alias >= 0 && alias < n) ? alias : (int32_t)row;
}

}  // namespace atg
