// policy.h -- per-length-bucket launch policy of the persistent path.
//
// A policy is an ordered list of buckets; a pair (or a call) with longest
// sequence L uses the first bucket with L <= len_hi.  The autotuner
// (tune.cpp) writes policy files; without one the built-in table applies.
// A policy changes speed only: every policy gives identical results.
#pragma once

#include <vector>

#include "ewfa_internal.h"

namespace ewfa {

struct Policy {
    std::vector<BucketPolicy> buckets;
    const BucketPolicy& bucket_for(uint32_t len) const;
};

// Built-in table.
Policy default_policy();

// Loads `path` (NULL or empty: built-in).  File format, one bucket per line,
// '#' starts a comment:   len_hi waves dcap bt_lds grab
int load_policy(const char* path, Policy& out);

// Writes a policy file.
int save_policy(const char* path, const Policy& p);

// Replaces a context's policy (the tuner's candidates).
int set_context_policy(ewfa_ctx_t* ctx, const Policy& p);

// Op scratch bytes an align call over n pairs up to max_len bases needs.
size_t op_scratch_bytes(size_t n, uint32_t max_len);

// ewfa_align_device with the op scratch in caller memory (the job server's
// arena) instead of the context's own pool.
int align_device_with_scratch(ewfa_ctx_t* ctx, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n,
                              uint32_t max_len, ewfa_result_t* d_results, int8_t* d_cigars, void* stream,
                              uint64_t* d_ops, size_t ops_bytes);

}  // namespace ewfa
