// tune.cpp -- the autotuner: per length bucket, time candidate launch
// configurations on a sample and keep the fastest.  Every candidate must give
// identical outputs, so the policy it writes changes speed only.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "ewfa_internal.h"
#include "policy.h"

namespace {

struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) { return hipMalloc(&p, n ? n : 4) == hipSuccess ? EWFA_OK : EWFA_ERR_NOMEM; }
};

void appendf(std::string& out, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    out += buf;
}

std::vector<ewfa::BucketPolicy> candidates(uint32_t blen) {
    std::vector<ewfa::BucketPolicy> c;
    const int dmax = (int)blen + 1;
    // one wave per pair, wavefront in registers (levels up to 32*chunks-1)
    for (int dcap : {31, 63, 95, 127}) {
        if (dcap > dmax && dcap != 31) continue;
        for (int grab : {1, 4, 16, 64}) c.push_back({0xFFFFFFFFu, 1, dcap, 1, grab});
    }
    // one wave per pair, offsets in LDS
    if (blen >= 400) c.push_back({0xFFFFFFFFu, 1, std::min(255, dmax), 1, 4});
    // a block per pair, backtrace in HBM
    if (blen >= 400)
        for (int waves : {2, 4, 8, 16})
            for (int dcap : {384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 1 << 20}) {
                if (dcap > 4 * dmax && dcap != 512 && dcap != (1 << 20)) continue;
                if (waves == 16 && dcap < 1024) continue;
                if (dcap > 2048 && dcap < (1 << 20) && blen < 10000) continue;
                c.push_back({0xFFFFFFFFu, waves, dcap, 0, 1});
            }
    return c;
}

}  // namespace

extern "C" int ewfa_autotune(int device, const uint64_t* h_words, size_t n_words, const ewfa_pair_t* h_pairs,
                             size_t n, const uint32_t* bucket_hi, int nb, int reps, const char* out_path,
                             char* report, size_t report_cap) {
    if (!h_words || !h_pairs || !bucket_hi || nb < 1 || nb > ewfa::MAX_BUCKETS || !out_path) return EWFA_ERR_ARG;
    if (reps < 1) reps = 1;
    std::string rep;
    ewfa_opts_t o;
    ewfa_default_opts(&o);
    o.device = device;
    ewfa_ctx_t* ctx = nullptr;
    int rc = ewfa_create(&ctx, &o);
    if (rc != EWFA_OK) return rc;
    DevMem d_words;
    if ((rc = d_words.alloc(n_words * 8)) != EWFA_OK || hipMemcpy(d_words.p, h_words, n_words * 8, hipMemcpyHostToDevice) != hipSuccess) {
        ewfa_destroy(ctx);
        return rc != EWFA_OK ? rc : EWFA_ERR_HIP;
    }
    ewfa::Policy out;
    for (int b = 0; b < nb; ++b) {
        const uint32_t lo = b ? bucket_hi[b - 1] : 0u;
        const uint32_t hi = (b == nb - 1) ? 0xFFFFFFFFu : bucket_hi[b];
        std::vector<ewfa_pair_t> sub;
        uint32_t blen = 0;
        size_t cbytes = 0;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t L = std::max(h_pairs[i].plen, h_pairs[i].tlen);
            if ((b == 0 || L > lo) && L <= hi) {
                ewfa_pair_t p = h_pairs[i];
                p.cigar_off = cbytes;
                p.cigar_cap = (uint32_t)(((uint64_t)p.plen + p.tlen + 4 + 3) & ~3ull);
                cbytes += p.cigar_cap;
                sub.push_back(p);
                blen = std::max(blen, L);
            }
        }
        ewfa::BucketPolicy best{};
        if (sub.empty()) {
            best = ewfa::default_policy().bucket_for(hi == 0xFFFFFFFFu ? lo + 1 : hi);
            best.len_hi = hi;
            out.buckets.push_back(best);
            appendf(rep, "bucket %d (%u..%u]: no sample pairs, built-in entry\n", b, lo, hi);
            continue;
        }
        DevMem d_pairs, d_res, d_cig;
        if (d_pairs.alloc(sub.size() * sizeof(ewfa_pair_t)) != EWFA_OK || d_res.alloc(sub.size() * 16) != EWFA_OK ||
            d_cig.alloc(cbytes) != EWFA_OK ||
            hipMemcpy(d_pairs.p, sub.data(), sub.size() * sizeof(ewfa_pair_t), hipMemcpyHostToDevice) != hipSuccess) {
            ewfa_destroy(ctx);
            return EWFA_ERR_NOMEM;
        }
        std::vector<ewfa_result_t> r0, r1(sub.size());
        std::vector<int8_t> c0, c1(cbytes);
        double best_t = 1e30;
        appendf(rep, "bucket %d (%u..%u]: %zu pairs, longest %u\n", b, lo, hi, sub.size(), blen);
        for (const ewfa::BucketPolicy& cand : candidates(blen)) {
            ewfa::Policy pol;
            pol.buckets.push_back(cand);
            ewfa::set_context_policy(ctx, pol);
            const auto run = [&]() {
                return ewfa_align_device(ctx, (const uint64_t*)d_words.p, (const ewfa_pair_t*)d_pairs.p, sub.size(),
                                         blen, (ewfa_result_t*)d_res.p, (int8_t*)d_cig.p, nullptr);
            };
            if ((rc = run()) != EWFA_OK) {  // warm-up (also captures the graph)
                appendf(rep, "  waves %2d dcap %7d bt_lds %d grab %2d: error %d\n", cand.waves, cand.dcap,
                        cand.bt_lds, cand.grab, rc);
                continue;
            }
            const auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < reps && rc == EWFA_OK; ++r) rc = run();
            const double t =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
            if (rc != EWFA_OK) continue;
            if (hipMemcpy(r1.data(), d_res.p, sub.size() * 16, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(c1.data(), d_cig.p, cbytes, hipMemcpyDeviceToHost) != hipSuccess) {
                ewfa_destroy(ctx);
                return EWFA_ERR_HIP;
            }
            // streams are compared up to their lengths (bytes past a
            // stream's terminator are not written)
            for (size_t i = 0; i < sub.size(); ++i) {
                const int32_t len = std::min<int32_t>(r1[i].cigar_len, (int32_t)sub[i].cigar_cap);
                memset(c1.data() + sub[i].cigar_off + std::max(len, 0), 0,
                       sub[i].cigar_cap - (size_t)std::max(len, 0));
            }
            if (r0.empty()) {
                r0 = r1;
                c0 = c1;
            } else if (memcmp(r0.data(), r1.data(), r1.size() * 16) != 0 || c0 != c1) {
                appendf(rep, "  waves %d dcap %d bt_lds %d grab %d: OUTPUT DIFFERS\n", cand.waves, cand.dcap,
                        cand.bt_lds, cand.grab);
                if (report && report_cap) snprintf(report, report_cap, "%s", rep.c_str());
                ewfa_destroy(ctx);
                return EWFA_ERR_STATE;
            }
            appendf(rep, "  waves %2d dcap %7d bt_lds %d grab %2d: %9.3f ms, %.3g pairs/s\n", cand.waves,
                    cand.dcap, cand.bt_lds, cand.grab, t * 1e3, sub.size() / t);
            if (t < best_t) {
                best_t = t;
                best = cand;
            }
        }
        if (best_t >= 1e30) {
            ewfa_destroy(ctx);
            return rc != EWFA_OK ? rc : EWFA_ERR_STATE;
        }
        best.len_hi = hi;
        out.buckets.push_back(best);
        appendf(rep, "  chosen: waves %d dcap %d bt_lds %d grab %d\n", best.waves, best.dcap, best.bt_lds,
                best.grab);
    }
    ewfa_destroy(ctx);
    rc = ewfa::save_policy(out_path, out);
    if (report && report_cap) snprintf(report, report_cap, "%s", rep.c_str());
    return rc;
}
