// runtime.h -- internal host interfaces shared by the library's translation
// units (the job server and the host pipeline use them; callers use ewfa.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ewfa.h"

namespace ewfa {

// Sets the calling thread's ewfa_last_error() message (printf format) and
// returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Device scratch of one persistent align call.  bt and runs share one region
// of `region` bytes (the passes of a call run one after another).
struct Scratch {
    uint64_t* ops;       // op words, ops_words(max_len + 1) per pair
    uint32_t* ctrl;      // control words (queue heads, counts)
    uint32_t* list_a;    // deferral lists (ping-pong) and the emit fallback list
    uint32_t* list_b;
    uint32_t* list_fb;
    uint32_t* buckets;   // per-bucket pair lists
    uint32_t* bound;     // per-pair distance bounds
    uint32_t* order;     // bounded bucket lists by decreasing bound (null: kept in list order)
    uint64_t* bt;        // HBM backtrace slots of the block passes
    uint8_t* runs;       // run rows of the multi-chunk register passes
    uint64_t* gslot;     // HBM slots of the passes beyond what LDS holds
    size_t region;
};

// ewfa_align_device on the persistent path with every scratch buffer carved
// out of `scratch` (`bytes` long, 256-byte aligned) and eager launches: no
// device or host memory is allocated and no graph is captured.  Pass grids
// are narrowed so that their backtrace and run-row slots fit the region left
// after the per-pair buffers; EWFA_ERR_RANGE when not even one slot does.
int align_device_in(ewfa_ctx_t* ctx, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n, uint32_t max_len,
                    ewfa_result_t* d_results, int8_t* d_cigars, void* stream, void* scratch, size_t bytes);

// The device a context runs on.
int context_device(ewfa_ctx_t* ctx);

// Op scratch bytes an align call over n pairs up to max_len bases needs.
size_t op_scratch_bytes(size_t n, uint32_t max_len);

// ewfa_align_device with the op scratch in caller memory (the job server's
// arena) instead of the context's own pool.
int align_device_with_scratch(ewfa_ctx_t* ctx, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n,
                              uint32_t max_len, ewfa_result_t* d_results, int8_t* d_cigars, void* stream,
                              uint64_t* d_ops, size_t ops_bytes);

// The context's own stream (its kernels run there).
hipStream_t context_stream(ewfa_ctx_t* ctx);

// Enqueues the persistent path (bucketing, align passes, deferral chain,
// emits) on the context's stream and returns without waiting.  The op
// scratch is caller memory of at least op_scratch_bytes(n, max_len) bytes.
// Work queued earlier on that stream may still be running: the context's
// buffers are only replaced after the device drains.
int enqueue_align(ewfa_ctx_t* ctx, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n, uint32_t max_len,
                  ewfa_result_t* d_results, int8_t* d_cigars, uint64_t* d_ops, size_t ops_bytes);

// CIGAR compaction after an align call: enqueues the packing of every pair's
// written CIGAR bytes into the context's dense buffer (`bound` >= the sum of
// the pairs' cigar_cap).
int enqueue_compact(ewfa_ctx_t* ctx, const ewfa_pair_t* d_pairs, const ewfa_result_t* d_results,
                    const int8_t* d_cigars, size_t n, size_t bound, hipStream_t stream);

// Waits for the stream, then copies the first `total` dense bytes into the
// context's pinned staging; *out points at them until the next call.
int fetch_compact(ewfa_ctx_t* ctx, size_t total, hipStream_t stream, const int8_t** out);

// Dense bytes of n pairs (host copies of descriptors and results).
size_t compact_total(const ewfa_pair_t* pairs, const ewfa_result_t* results, size_t n);

// Places each pair's dense bytes at its cigar_off in h_cigars and returns
// their total, or SIZE_MAX (placing nothing) when that exceeds dense_cap.
size_t scatter_streams(const int8_t* dense, size_t dense_cap, const ewfa_pair_t* pairs, const ewfa_result_t* results,
                       size_t n, int8_t* h_cigars, int threads);

}  // namespace ewfa
