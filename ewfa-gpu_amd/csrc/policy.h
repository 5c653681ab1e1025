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


}  // namespace ewfa
