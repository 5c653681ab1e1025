// Internal declarations shared by the HIP kernels (kernels.hip) and the host
// runtime (api.cpp, serve.cpp, tune.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ewfa.h"

namespace ewfa {

// Backtrace op codes of the reference (src/wavefront_structures.h:66-71).
enum : int { OP_NOOP = 0, OP_INS = 1, OP_SUB = 2, OP_DEL = 3 };

// Per-pair status values used only between the two kernels.
constexpr int32_t PAIR_DEFERRED = 100;  // distance beyond this pass's budget

// Backtrace storage: one "chunk" per (level d, 64 consecutive diagonals)
// holding two 64-bit ballots (op bit 0 and op bit 1).  Level d >= 1 has
// d/32 + 1 chunks (2d+1 diagonals); chunks of levels 1..d-1 come first.
__host__ __device__ inline uint32_t bt_level_base(uint32_t d) {
    const uint32_t q = d >> 5, r = d & 31u;
    return d + 16u * q * (q - (q ? 1u : 0u)) + q * r - 1u;
}
// Chunks needed for levels 1..dmax.
__host__ __device__ inline uint32_t bt_chunks(uint32_t dmax) { return bt_level_base(dmax + 1); }

// Words of the ops scratch per pair: op s (0-based) at bits 2(s%32) of word s/32.
__host__ __device__ inline uint32_t ops_words(uint32_t dmax) { return dmax / 32u + 1u; }

// How one length bucket of pairs is processed.
struct BucketPolicy {
    uint32_t len_hi;     // pairs with max(plen,tlen) <= len_hi (inclusive)
    int32_t waves;       // waves per pair: 1 = wave-per-pair, else block-per-pair
    int32_t dcap;        // wavefront levels held by this pass (deferred beyond)
    int32_t bt_lds;      // 1: backtrace in LDS, 0: in the HBM slot region
    int32_t grab;        // pairs taken per queue pull
};

constexpr int MAX_BUCKETS = 8;
// A dispatch's workgroups are dealt to the device's 8 XCDs in turn (block b
// runs on XCD b % 8).  Per-XCD work counters of a pass sit XQ_STRIDE words
// (128 bytes, a cache line of their own) apart.
constexpr int XCDS = 8;
constexpr int XQ_STRIDE = 32;

// Length limits of the buckets, by value into the bucketing kernel.
struct BucketTable {
    uint32_t hi[MAX_BUCKETS];
    int nb;
};

// Arguments of the align kernel (passed by value).
struct AlignArgs {
    const uint64_t* words;
    const ewfa_pair_t* pairs;
    ewfa_result_t* results;
    uint64_t* ops;            // ops scratch, ops_stride words per pair
    uint32_t ops_stride;
    // work list: list[0..count) pair indices, count read from *count_ptr when
    // count_ptr != null, else `count`
    const uint32_t* list;     // null = identity
    uint32_t count;
    const uint32_t* count_ptr;
    uint32_t* queue;          // persistent: queue head counter (null = plain)
    // persistent bucket passes: XCDS counters, XQ_STRIDE words apart, one
    // per contiguous eighth of the list (null: the one queue above)
    uint32_t* xqueue;
    uint32_t grab;
    // deferral output (pairs whose distance exceeds dcap)
    uint32_t* defer_list;
    uint32_t* defer_count;
    // per-block HBM backtrace slot region (bt_lds == 0)
    uint64_t* bt_global;
    uint64_t bt_slot_words;
    int32_t dcap;
    int32_t max_distance;     // reference cap, 0 = none
    int32_t seq_words;        // LDS words per sequence (pattern, text)
    int32_t bt_lds;
    uint32_t max_len;
    // emit fused into the register kernel: per-wave scratch of match-run
    // lengths (one byte per diagonal and level, saturating at 255), the CIGAR
    // base, and the list of pairs left to the emit kernel (a run of 255 or
    // more on their path)
    uint8_t* runs;            // null: no fused emit
    uint32_t runs_slot;       // bytes per wave
    int8_t* cigars;
    uint32_t* fallback_list;
    uint32_t* fallback_count;
    // per-pair upper bound on the distance (BOUND_NONE: none), written by
    // bound_kernel and read by the block kernel, which then leaves out the
    // cells no path of that cost can use; null: no bounds
    uint32_t* bound;
    // bounded passes with narrow slots: each offset generation holds the
    // band_cap + 6 cells of the diagonals a bound of at most band_cap leaves
    // (pairs with larger bounds are deferred); 0: generations of 2*dcap + 5
    int32_t band_cap;
    // 16-bit offsets in LDS (the pass's 2 * (max_len + dcap) fits)
    int32_t off16;
    // per-block slots in HBM (sequence copies and both offset generations,
    // gslot_words each) for passes whose wavefront LDS cannot hold; null: the
    // slot is in LDS.  The backtrace of such a pass is in HBM, stored through
    // 64-bit addresses.
    uint64_t* gslot;
    uint64_t gslot_words;
};

constexpr uint32_t BOUND_NONE = 0xFFFFFFFFu;

// Bytes of a wave's run scratch for levels 0..dmax: level d's 2d+1 diagonals
// at d*d, so a pair at distance D touches (D+1)^2 bytes; slots are rounded
// to 128-byte lines so no two waves share one.
__host__ __device__ inline uint32_t runs_slot_bytes(uint32_t dmax) { return ((dmax + 1) * (dmax + 1) + 127u) & ~127u; }

struct EmitArgs {
    const uint64_t* words;
    const ewfa_pair_t* pairs;
    ewfa_result_t* results;
    const uint64_t* ops;
    uint32_t ops_stride;
    int8_t* cigars;
    // pairs to emit: list[0..count) (list null = identity), count read from
    // *count_ptr when count_ptr != null
    const uint32_t* list;
    uint32_t count;
    const uint32_t* count_ptr;
    uint32_t per_wave;  // emit_kernel: pairs per wave (set by launch_emit)
};

// LDS bytes one pair slot needs for the given budget.
inline size_t lds_bytes(int seq_words, int dcap, int bt_lds, int off16 = 0) {
    size_t b = 16u * (size_t)seq_words;                   // pattern + text
    b += (off16 ? 4u : 8u) * (size_t)(2 * dcap + 5);      // two offset generations
    b = (b + 15) & ~(size_t)15;
    if (bt_lds) b += 16u * (size_t)bt_chunks((uint32_t)dcap);
    return b;
}

// LDS bytes of a bounded block pass's narrow slot (band_cap above): both
// sequences and two offset generations of band_cap + 6 32-bit cells.
inline size_t lds_bytes_narrow(int seq_words, int band_cap) {
    size_t b = 16u * (size_t)seq_words + 8u * (size_t)(band_cap + 6);
    return (b + 15) & ~(size_t)15;
}

// Guard words (64-bit) of a banded slot before the pattern copy and after
// the text copy: reads of lanes outside the wavefront, and of offsets past an
// end, land in them.
constexpr int BAND_HEAD_WORDS = 2;
constexpr int BAND_TAIL_WORDS = 4;
// The banded slot's run rows: a 64-byte row per level 0..63.
constexpr int BAND_RUN_BYTES = 64 * 64;

// LDS bytes of a register-wavefront slot.  One chunk (the banded wavefront)
// keeps its backtrace in registers: its slot is the two plain copies
// (seq_words each) between the guards, then the run rows.  More chunks: the
// overlapped copies and the backtrace.
inline size_t lds_bytes_reg(int seq_words, int dcap, int nc) {
    if (nc == 1) return 8u * (size_t)(2 * seq_words + BAND_HEAD_WORDS + BAND_TAIL_WORDS) + BAND_RUN_BYTES;
    return 16u * (size_t)seq_words + 16u * (size_t)bt_chunks((uint32_t)dcap);
}

// Register chunks the register wavefront needs for dcap levels (0: too many).
// Up to 63 levels the banded wavefront holds them in one chunk.
inline int reg_chunks_for(int dcap) {
    if (dcap <= 63) return 1;
    const int nc = (2 * dcap + 1 + 63) / 64;
    return nc <= 4 ? nc : 0;
}


// Bytes of a pair's CIGAR region the emit kernel wrote: whole dwords from
// cigar_off up to and including the terminator (or up to cigar_cap when the
// stream filled or overflowed it); none when the emit did not run.
__host__ __device__ inline uint32_t written_cigar_bytes(const ewfa_result_t& r, const ewfa_pair_t& p) {
    if (r.status != EWFA_PAIR_OK && r.status != EWFA_PAIR_CAPPED && r.status != EWFA_PAIR_UNDEFINED &&
        r.status != EWFA_PAIR_OVERFLOW)
        return 0u;
    if ((p.cigar_off & 3u) || (p.cigar_cap & 3u)) return 0u;
    const uint32_t len = r.cigar_len > 0 ? (uint32_t)r.cigar_len : 0u;
    const uint64_t w = (uint64_t)(len & ~3u) + 4u;
    return w < p.cigar_cap ? (uint32_t)w : p.cigar_cap;
}

// Pairs per block of the CIGAR compaction kernels.
constexpr uint32_t COMPACT_BLOCK = 1024;

// Kernel launchers (kernels.hip).  All asynchronous on `stream`.
// Distance bounds of list items from a 64-diagonal band: one wave per pair,
// reading the sequences in place, a.queue as its work counter.
hipError_t launch_bound(const AlignArgs& a, uint32_t grid, hipStream_t stream);
// Each XCD share of a bounded pass's list (list null: the identity; count
// from *count_ptr when not null) into `out`, by decreasing distance bound
// (ORDER_BINS bins of bound >> shift, none first).
constexpr int ORDER_BINS = 64;
hipError_t launch_order_by_bound(const uint32_t* list, const uint32_t* count_ptr, uint32_t count,
                                 const uint32_t* bound, uint32_t* out, int shift, hipStream_t stream);
// Bounded passes: those whose sequence copies (16 bytes per word pair) take
// at most 64 KiB, the block kernels' slots in LDS.
constexpr int BOUND_MAX_SEQ_WORDS = 4096;
hipError_t launch_align(const AlignArgs& a, int waves, bool persistent, uint32_t grid,
                        size_t lds, hipStream_t stream);
hipError_t launch_align_reg(const AlignArgs& a, int nc, bool persistent, uint32_t grid, size_t lds,
                            hipStream_t stream);
int align_reg_blocks_per_cu(int nc, size_t lds);
// One lane per pair, per_wave pairs per wave (1..64), grid-stride:
// grid_items bounds the launch (the count may be read on the device).
hipError_t launch_emit(const EmitArgs& a, uint32_t grid_items, uint32_t per_wave, hipStream_t stream);
// Pairs per wave of the one-lane-per-pair emit for large launches: a quarter
// of the lanes, so four times the waves are in flight to hide the walks'
// memory latency (on the mixed-length workload 1.6% faster than a full wave,
// and 4% faster than 8 pairs per wave).
// Smaller launches spread their pairs thinner, down to one per wave.
constexpr int EMIT_PER_WAVE = 16;
// Emit with one wave per pair and its sequences staged in LDS as overlapped
// copies of sw words (2 * sw * 8 bytes per block, at most EMIT_WAVE_MAX_LDS);
// at most max_grid blocks, grid-stride.
hipError_t launch_emit_wave(const EmitArgs& a, uint32_t grid_items, int sw, uint32_t max_grid, hipStream_t stream);
constexpr size_t EMIT_WAVE_MAX_LDS = 65536;
// Marks the pairs of list[0..*count_ptr) EWFA_PAIR_TOO_DISTANT.
hipError_t launch_mark_too_distant(const uint32_t* list, const uint32_t* count_ptr, ewfa_result_t* results,
                                   uint32_t max_items, hipStream_t stream);
// Emit with the sequences staged in LDS: sw words per sequence slot, 64 pairs
// per block (64 * 2 * sw * 8 bytes of LDS).
hipError_t launch_emit_lds(const EmitArgs& a, uint32_t grid_items, int sw, hipStream_t stream);
constexpr int EMIT_LDS_MAX_SW = 64;  // 64 KiB of LDS per block
hipError_t launch_pack(const char* ascii, size_t ascii_stride, const uint32_t* len, size_t nseq,
                       uint64_t* words, size_t words_stride, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream);
// zeroes p[i] for every bit i (< 64) set in mask
hipError_t launch_clear_words(uint32_t* p, uint64_t mask, hipStream_t stream);
hipError_t launch_bucketize(const ewfa_pair_t* pairs, uint32_t n, const BucketTable& t, uint32_t* lists,
                            uint32_t* counts, hipStream_t stream);
// Packs every pair's written CIGAR bytes, in pair order, densely into
// `dense` (offsets: the exclusive prefix sum of written_cigar_bytes).
// `scan` holds 2 * ceil(n / COMPACT_BLOCK) words.
hipError_t launch_compact(const ewfa_pair_t* pairs, const ewfa_result_t* results, const int8_t* cigars, size_t n,
                          uint64_t* scan, int8_t* dense, hipStream_t stream);
// Waves per pair of the passes with their slots in HBM.
constexpr int GSLOT_WAVES = 16;
// Resident blocks per CU of a block-kernel variant (0: backtrace in HBM, 1: in
// LDS, 2: with distance bounds, 3: with 16-bit offsets, 4: slots in HBM, 5:
// slots in HBM with distance bounds).
int align_blocks_per_cu(int waves, size_t lds, int variant = 1);
// Raises the dynamic LDS limit of every block-kernel variant to `lds` bytes
// (never lowers it).
hipError_t align_set_max_lds(size_t lds);

}  // namespace ewfa
