// Host-visible description of one build and the launchers of its phases.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "atg_types.h"

namespace atg {

// Pack methods (the values of atg_pack_method).
enum PackMethod : int32_t { PACK_CHUNKED = 0, PACK_BASIC = 1, PACK_NO_WEIGHTS = 2 };

// Everything a build kernel needs, passed by value (all device pointers).
struct BuildArgs {
    const void* weights;  // n weights (double or float)
    int dtype_f32;
    int64_t n;
    int64_t G;            // sections
    int64_t Ls;           // items per section
    int64_t Ls_max;       // LDS capacity of the partition kernel (items per section)
    int64_t tree_total;   // blocks of the dyadic section-prefix tree (sum over b of G >> b)
    DevState* st;
    double* section_sums;     // G
    double* tree_scratch;     // 2 * tree_scratch_half
    int64_t tree_scratch_half;
    uint64_t* slots;          // dyadic block sums (counts, light, heavy) + arrival counters
    double* stage_w;          // n: leftovers staged by partition pass 1 (weights)
    int32_t* stage_i;         // n: leftovers staged by partition pass 1 (items)
    uint64_t* sec_cert;       // 6 G: per-section split-certificate data (bit patterns)
    int32_t want_cert;        // compute the split certificate (learned split)
    int32_t* Li; double* Lw; double* Lp;  // leftover lights (capacity n + 2)
    int32_t* Hi; double* Hw; double* Hp;  // leftover heavies (capacity n + 2)
    Row* rows;                // AoS table or null
    double* tw; int32_t* ta;  // SoA table or null
    int32_t psa_plus;
    int32_t chains;
    int64_t p_req;            // p after the N/10 clamp (grid bound for split/pack)
    Split* splits;            // capacity p_req + 1
    int32_t pack_cap;         // list items one pack wave stages in LDS (pack_cap_items)
    // learned split index
    int32_t learned_segments;
    int32_t* knot_j;          // learned_segments + 1 exact split j at the knots
    int32_t pack_method;      // atg_pack_method
    // Lean light lists: the placement writes only the light prefixes; the
    // light items and weights stay in the staging, found through sec_loff
    // (G + 1 values: the lights before section s, then the light count).
    int32_t lean_lights;
    uint32_t* sec_loff;
    uint32_t* sec_ldir;       // light_dir_size(n): the section of light q << kLightDirShift
};

constexpr int kLightDirShift = 10;
inline int64_t light_dir_size(int64_t n) { return (n >> kLightDirShift) + 2; }

// The lean light lists need every pack chain to fit the staged pack's LDS
// pool (chains hold at most n / p + 3 items, the pool at least
// min(64 (n / p + 1), 1024)).
inline bool lean_lights_ok(int64_t n, int64_t p_req, int32_t pack_method) {
    return pack_method == PACK_CHUNKED && p_req > 0 && n / p_req <= 256;
}

// Section layout of PSA+ (AliasTableSplitGpu.cu:356-357, PsaPlusNormal.cu:27-29).
inline void section_layout(int64_t n, int64_t param, int64_t* G, int64_t* Ls) {
    if (param < 1) param = 1;
    *G = n / param + 1;
    *Ls = n / *G + 1;
}

// Items of the LDS pool of one pack wave: the items of all its 64 chains
// when they fit (splits are evenly spaced in list items, n' = ceil(nToSplit
// k / p), so 64 chains hold at most 64 ceil(nToSplit / p) + 1 items, with
// nToSplit <= n + 2 and chains near 10 items when p is clamped below p_req),
// at most kPackPoolItems; a wave whose chains need more runs them in rounds.
constexpr int32_t kPackPoolItems = 1024;
inline int32_t pack_cap_items(int64_t n, int64_t p_req) {
    if (p_req < 1) p_req = 1;
    int64_t per = (n + 2 + p_req - 1) / p_req;
    if (per < 11) per = 11;
    int64_t cap = (64 * (per + 1) + 2 + 15) / 16 * 16;
    return (int32_t)(cap < kPackPoolItems ? cap : kPackPoolItems);
}
// A pack round that fits fewer chains than this into the pool hands the
// wave's remaining chains to their lanes, each reading the lists from HBM.
constexpr int kPackMinStagedChains = 8;
// Chains of more than this many items (n / p) run the windowed pack: so few
// of them fit a pool round that the staged pack would hand nearly all of them
// to their lanes.
constexpr int64_t kPackWindowedItems = 256;
inline bool pack_windowed_ok(int64_t n, int64_t p_req, int32_t pack_method) {
    return pack_method == PACK_CHUNKED && p_req > 0 && n / p_req > kPackWindowedItems;
}
// Weight (8 B), item (4 B) and recorded alias (4 B) per pool item, and the
// spare weight slot the light steps record into (16 B keep the alignment).
inline size_t pack_lds_bytes(int32_t cap) { return (size_t)cap * 16 + 16; }

// Knots of the learned split index: at most `opt`, and at most one per 16
// requested splits (at least 64).  Each knot is a full bisection, so a table
// with few splits would otherwise bisect nearly every split twice; the
// splits found are the same for any knot count.
inline int32_t learned_knots(int32_t opt, int64_t p_req) {
    const int64_t k = p_req / 16 > 64 ? p_req / 16 : 64;
    return (int32_t)(k < opt ? k : opt);
}

size_t psa_lds_bytes(int64_t Ls_max);
int64_t resident_blocks(const void* f, int threads, size_t lds);
int64_t block_tree_total(int64_t G);
hipError_t set_partition_lds_limit(size_t bytes);
hipError_t launch_sum(const BuildArgs& a, hipStream_t s);
hipError_t launch_partition(const BuildArgs& a, hipStream_t s);
hipError_t launch_partition_local(const BuildArgs& a, hipStream_t s);
hipError_t launch_partition_place(const BuildArgs& a, hipStream_t s);
hipError_t launch_degenerate(const BuildArgs& a, hipStream_t s);
hipError_t launch_split(const BuildArgs& a, hipStream_t s);
hipError_t launch_split_learned(const BuildArgs& a, hipStream_t s);
hipError_t launch_split_pary(const BuildArgs& a, hipStream_t s);
hipError_t launch_pack(const BuildArgs& a, hipStream_t s);
// Lean light lists written out as lists (Li, Lw) from the staging.
hipError_t launch_lean_lights_fill(const BuildArgs& a, hipStream_t s);
// Chunked split and pack (split/pack interleave): blocks [b0, b1) of
// split_blocks (256 splits each) or pack_blocks (64 chains each).
int64_t split_blocks(const BuildArgs& a);
int64_t pack_blocks(const BuildArgs& a);
hipError_t launch_split_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1);
hipError_t launch_split_knots(const BuildArgs& a, hipStream_t s);
hipError_t launch_split_learned_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1);
hipError_t launch_split_pary_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1);
hipError_t launch_pack_blocks(const BuildArgs& a, hipStream_t s, int64_t b0, int64_t b1);
hipError_t launch_split_rows(const Row* rows, double* w, int32_t* al, int64_t n, hipStream_t s);
hipError_t launch_join_rows(const double* w, const int32_t* al, Row* rows, int64_t n, hipStream_t s);

}  // namespace atg
