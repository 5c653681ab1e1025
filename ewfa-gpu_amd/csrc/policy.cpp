// policy.cpp -- built-in launch policy and policy files.
#include "policy.h"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

namespace ewfa {

const BucketPolicy& Policy::bucket_for(uint32_t len) const {
    for (const BucketPolicy& b : buckets)
        if (len <= b.len_hi) return b;
    return buckets.back();
}

Policy default_policy() {
    Policy p;
    // tuned on 500,000 pairs of log-uniform length 100..10,000 bp at 1..10%
    // edits; longer pairs take an 8-wave block with as many levels as LDS
    // holds, bounded, in narrow slots (two blocks per CU: 20,000 pairs of
    // 10..50 kbp in 211 ms, against 290 ms for one 16-wave block per CU).
    // Those are two buckets of the same parameters, split at 25 kbp: their
    // passes run on separate streams, so that one's bound search (which
    // holds no LDS) runs beside the other's block pass (189 -> 184 ms)
    //              len_hi   waves dcap  bt_lds grab
    p.buckets.push_back({300u, 1, 63, 1, 16});
    p.buckets.push_back({700u, 1, 127, 1, 4});
    p.buckets.push_back({1200u, 1, 127, 1, 4});
    p.buckets.push_back({2500u, 2, 768, 0, 1});
    p.buckets.push_back({5000u, 2, 512, 0, 1});
    p.buckets.push_back({10500u, 2, 1024, 0, 1});
    p.buckets.push_back({25000u, 8, 1 << 20, 0, 1});
    p.buckets.push_back({0xFFFFFFFFu, 8, 1 << 20, 0, 1});
    return p;
}

int load_policy(const char* path, Policy& out) {
    out = default_policy();
    if (!path || !*path) return EWFA_OK;
    FILE* f = fopen(path, "r");
    if (!f) return EWFA_ERR_IO;
    Policy p;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char* hash = strchr(line, '#');
        if (hash) *hash = 0;
        unsigned long len_hi;
        int waves, dcap, bt, grab;
        if (sscanf(line, "%lu %d %d %d %d", &len_hi, &waves, &dcap, &bt, &grab) != 5) continue;
        if (waves != 1 && waves != 2 && waves != 4 && waves != 8 && waves != 16) continue;
        if (dcap < 1 || grab < 1) continue;
        p.buckets.push_back({(uint32_t)(len_hi > 0xFFFFFFFFul ? 0xFFFFFFFFul : len_hi), waves, dcap,
                             bt ? 1 : 0, grab});
    }
    fclose(f);
    // a plan holds at most MAX_BUCKETS passes; a longer file would leave its
    // last buckets (and the pairs bucketed into them) without a pass
    if (p.buckets.empty() || p.buckets.size() > (size_t)MAX_BUCKETS) return EWFA_ERR_FORMAT;
    p.buckets.back().len_hi = 0xFFFFFFFFu;  // the last bucket takes every longer pair
    out = p;
    return EWFA_OK;
}

int save_policy(const char* path, const Policy& p) {
    FILE* f = fopen(path, "w");
    if (!f) return EWFA_ERR_IO;
    fprintf(f, "# len_hi waves dcap bt_lds grab\n");
    for (const BucketPolicy& b : p.buckets)
        fprintf(f, "%u %d %d %d %d\n", b.len_hi, b.waves, b.dcap, b.bt_lds, b.grab);
    fclose(f);
    return EWFA_OK;
}

}  // namespace ewfa

extern "C" int ewfa_policy_read(const char* path, ewfa_bucket_policy_t* out, int cap) {
    ewfa::Policy p;
    const int rc = ewfa::load_policy(path, p);
    if (rc != EWFA_OK) return rc;
    const int n = (int)p.buckets.size();
    for (int i = 0; i < n && i < cap && out; ++i) {
        const ewfa::BucketPolicy& b = p.buckets[(size_t)i];
        out[i] = ewfa_bucket_policy_t{b.len_hi, b.waves, b.dcap, b.bt_lds, b.grab};
    }
    return n;
}
