// api.cpp -- libewfa host runtime: context, scratch pools, launch policy, the
// plain per-batch and persistent launch paths, deferral passes and hipGraph
// replay.  Replaces the reference's Sequences/Cigars host pipeline
// (src/wavefront.cuh:43-505, src/wavefront.cu:29-330).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "ewfa_internal.h"
#include "policy.h"
#include "runtime.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return fail(EWFA_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                             \
    } while (0)

#define CHK(x)                 \
    do {                       \
        int r_ = (x);          \
        if (r_ != EWFA_OK) return r_; \
    } while (0)

// Device buffer of a context that only grows, by at least half again, so
// that batches of varying sizes settle after a few growths.  Growing it never
// waits: work already queued may still use the old allocation, which the
// context keeps (its retired list) until it is destroyed, or until an
// allocation fails and the context's streams have drained.  (Freeing device
// memory waits for the whole device, other contexts' work included.)  It
// bumps the context's buffer generation: a captured graph that may hold the
// old address is captured again.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int reserve(ewfa_ctx* c, size_t n);
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// Control words in device memory.
enum : int {
    CTRL_QUEUE = 0,     // queue head of the plain/deferral passes
    CTRL_DEFER_A = 1,   // deferred-pair counts (ping-pong)
    CTRL_DEFER_B = 2,
    CTRL_FALLBACK = 3,  // pairs the fused emit left to the emit kernel
    CTRL_CHAIN_BQ = 4,  // work counter of a deferral pass's distance bounds
    CTRL_BQUEUE = 8,    // queue heads of the bucket passes (MAX_BUCKETS)
    CTRL_BCOUNT = 16,   // pairs per bucket (MAX_BUCKETS)
    CTRL_BOUNDQ = 24,   // work counters of the bucket passes' distance bounds (MAX_BUCKETS)
    CTRL_XQ = 64,       // per-XCD work counters of the bucket passes (MAX_BUCKETS x XCDS, strided)
    CTRL_WORDS = CTRL_XQ + ewfa::MAX_BUCKETS * ewfa::XCDS * ewfa::XQ_STRIDE
};

}  // namespace

struct ewfa_ctx {
    int device = 0;
    ewfa_opts_t opts{};
    std::string policy_path;
    ewfa::Policy policy;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_a0 = nullptr, ev_a1 = nullptr, ev_e1 = nullptr;
    // the persistent plan's bucket passes spread over `fanout` streams: the
    // context's own and aux[0..fanout-2], forked from and joined back into it
    // with ev_fork / ev_join, so that one pass's tail runs beside the next
    static constexpr int MAX_FANOUT = 3;
    int fanout = MAX_FANOUT;
    hipStream_t aux[MAX_FANOUT - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[MAX_FANOUT - 1] = {};
    int num_cus = 0;
    size_t lds_max = 0;
    DevBuf ops, bt, ctrl, list_a, list_b, buckets, runs, list_fb, bound, gslot, order;
    uint32_t* host_ctrl = nullptr;  // pinned mirror of the control words
    float align_ms = 0.f, emit_ms = 0.f;
    bool timing = false;
    // graph cache for the persistent path
    hipGraphExec_t graph = nullptr, graph2 = nullptr;  // graph2: emit part when timed
    struct Key {
        const void *words, *pairs, *results, *cigars, *ops;
        size_t n;
        uint32_t max_len;
        bool timing;
        bool operator==(const Key& o) const {
            return words == o.words && pairs == o.pairs && results == o.results &&
                   cigars == o.cigars && ops == o.ops && n == o.n && max_len == o.max_len &&
                   timing == o.timing;
        }
    } graph_key{};
    bool use_graph = true;
    int policy_gen = 0, graph_policy_gen = -1;
    uint64_t buffer_gen = 0;        // bumped whenever one of the buffers is replaced
    std::vector<void*> retired;       // replaced device buffers, freed when idle
    size_t retired_bytes = 0;
    std::vector<void*> retired_host;  // replaced pinned host buffers, likewise
    uint64_t graph_buffer_gen = 0;  // buffer_gen when the graph was captured
    // op scratch supplied by the caller for one call (the job server's arena)
    uint64_t* ops_ext = nullptr;
    size_t ops_ext_bytes = 0;
    // all of one call's scratch in caller memory (the allocation-free entry
    // point): when set, no context buffer is used or grown
    const ewfa::Scratch* ext = nullptr;
    // host pipeline buffers
    DevBuf h_words, h_pairs, h_results, h_cigars;
    // CIGAR compaction: dense device copy, block sums, pinned host staging
    DevBuf dense, scan;
    int8_t* h_stage = nullptr;
    size_t h_stage_cap = 0;
    // host pipeline: upload stream, per-chunk upload events, and two stages of
    // pinned memory the device writes compacted streams and results into
    hipStream_t up = nullptr;
    std::vector<hipEvent_t> ev_up;
    struct Stage {
        int8_t* dense = nullptr;
        int8_t* dense_dev = nullptr;
        size_t dense_cap = 0;
        ewfa_result_t* res = nullptr;
        size_t res_cap = 0;
        hipEvent_t done = nullptr;
    } stage[2];
    size_t resident_words = 0;
    // an ewfa_upload_words_async in progress: its pieces of WORDS_PIECE words
    // go up on `wup`, piece k ending with ev_words[k]
    hipStream_t wup = nullptr;
    std::vector<hipEvent_t> ev_words;
    size_t words_pieces = 0;
};

namespace {

// Waits for the context's streams, then frees its retired buffers.
void release_retired(ewfa_ctx* c) {
    if (c->retired.empty() && c->retired_host.empty()) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipStream_t a : c->aux)
        if (a) (void)hipStreamSynchronize(a);
    if (c->up) (void)hipStreamSynchronize(c->up);
    for (void* q : c->retired) (void)hipFree(q);
    for (void* q : c->retired_host) (void)hipHostFree(q);
    c->retired.clear();
    c->retired_host.clear();
    c->retired_bytes = 0;
}

int DevBuf::reserve(ewfa_ctx* c, size_t n) {
    if (n <= bytes) return EWFA_OK;
    size_t want = p ? std::max(n, bytes + bytes / 2) : n;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, want);
    if (e != hipSuccess && want > n) {
        (void)hipGetLastError();
        want = n;
        e = hipMalloc(&q, want);
    }
    if (e != hipSuccess && !c->retired.empty()) {
        (void)hipGetLastError();
        release_retired(c);
        e = hipMalloc(&q, want);
    }
    if (e != hipSuccess) return fail(EWFA_ERR_NOMEM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    if (p) {
        c->retired.push_back(p);
        c->retired_bytes += bytes;
    }
    p = q;
    bytes = want;
    ++c->buffer_gen;
    return EWFA_OK;
}

int set_device(ewfa_ctx* c) {
    HIPCHK(hipSetDevice(c->device));
    return EWFA_OK;
}

// Pass configuration: one launch of the align kernel.
struct Pass {
    int waves;
    int dcap;
    int bt_lds;
    int grab;
    int seq_words;
    int reg;  // register-wavefront chunks (0: offsets staged in LDS)
    int off16;  // 16-bit LDS offsets (block kernel, backtrace in HBM, no bounds)
    int band_cap;  // bounded pass with narrow slots: the largest bound they hold (0: none)
    size_t lds;
    uint32_t grid;
    bool persistent;
    size_t slot_off;  // byte offset of the pass's slots in the backtrace (block
                      // passes) or run-row (register passes) region
    int gslot;          // 1: the block's slot (sequence copies, offsets) in HBM
    size_t gslot_bytes; // bytes of one such slot
    size_t gslot_off;   // byte offset of the pass's HBM slots
};

// Passes whose wavefront LDS cannot hold keep their slots in HBM.  Their
// backtrace slot grows as the square of the budget, so the budget stops where
// one block's backtrace slot reaches GSLOT_BT_MAX bytes (about 131,000 levels),
// and a pass gets at most as many blocks as fit GSLOT_BUDGET.
constexpr size_t GSLOT_BT_MAX = (size_t)4 << 30;
constexpr size_t GSLOT_BUDGET = (size_t)2 << 30;

int gslot_max_dcap() {
    static const int most = [] {
        int lo = 1, hi = 1 << 20;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (16u * (size_t)ewfa::bt_chunks((uint32_t)mid) <= GSLOT_BT_MAX) lo = mid; else hi = mid - 1;
        }
        return lo;
    }();
    return most;
}

// 64-bit words of one sequence slot in LDS: the data plus two pad words.
int seq_words_for(uint32_t max_len) { return (int)(max_len / 32u) + 3; }

// 64-bit words of one overlapped sequence copy (register-wavefront passes).
int overlapped_words_for(uint32_t max_len) { return 2 * (int)(max_len / 32u + 1) + 2; }

// Dynamic LDS a pass may use: the device maximum less room for the kernel's
// static LDS.
size_t lds_budget(const ewfa_ctx* c) { return c->lds_max > 1024 ? c->lds_max - 256 : c->lds_max; }

// Largest dcap whose slot fits `budget` bytes of LDS.
int max_dcap_for(int seq_words, int bt_lds, size_t budget) {
    if (ewfa::lds_bytes(seq_words, 0, bt_lds) > budget) return -1;
    int lo = 0, hi = 1 << 16;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (ewfa::lds_bytes(seq_words, mid, bt_lds) <= budget) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Block passes with budgets of at least this many levels take distance
// bounds (enqueue_bound).
constexpr int BOUND_MIN_DCAP = 2048;

// 16-bit offsets for a block pass without bounds whose largest doubled
// offset, 2 * (max_len + dcap) (an offset runs at most one base a level past
// an end), fits: half the offsets' LDS, more pairs per CU.
void choose_off16(Pass& p, uint32_t max_len) {
    p.off16 = (!p.reg && !p.bt_lds && p.dcap < BOUND_MIN_DCAP && (size_t)max_len + (size_t)p.dcap <= 16383) ? 1 : 0;
}

// A bounded block pass (a budget of BOUND_MIN_DCAP levels or more, sequences
// of at most BOUND_MAX_SEQ_WORDS words) whose slot lets fewer blocks share a
// CU than its narrow slot would: narrow slots sized for one block more per
// CU.  Such a slot keeps only the diagonals within a pair's bound U, at most
// U + 1 (against 2 * dcap + 5 cells a generation), and holds bounds up to
// band_cap; pairs with larger bounds go to the deferral chain.  Only bucket
// passes take it: the chain's passes must hold every pair they get.
void choose_band_cap(ewfa_ctx* c, Pass& p) {
    p.band_cap = 0;
    if (p.reg || p.gslot || p.bt_lds || p.off16 || p.dcap < BOUND_MIN_DCAP) return;
    if (p.seq_words > ewfa::BOUND_MAX_SEQ_WORDS) return;  // no bounds for this pass
    const size_t budget = lds_budget(c);
    const size_t per_cu = budget / ewfa::lds_bytes(p.seq_words, p.dcap, 0, 0);
    const size_t room = budget / (per_cu + 1);
    const size_t seq = 16u * (size_t)p.seq_words;
    if (room < seq + 8u * (size_t)(BOUND_MIN_DCAP + 6)) return;
    const int cap = (int)(((room & ~(size_t)15) - seq) / 8u) - 6;
    p.band_cap = std::min(cap, p.dcap);
}

int finish_pass(ewfa_ctx* c, Pass& p, uint32_t items) {
    if (p.gslot) {
        // slot in HBM, no dynamic LDS; blocks bounded by occupancy and memory
        p.lds = 0;
        p.gslot_bytes = (ewfa::lds_bytes(p.seq_words, p.dcap, 0, 0) + 255) & ~(size_t)255;
        const int bpc = std::max(1, std::min(ewfa::align_blocks_per_cu(ewfa::GSLOT_WAVES, 0, 4),
                                             ewfa::align_blocks_per_cu(ewfa::GSLOT_WAVES, 0, 5)));
        uint64_t g = (uint64_t)bpc * (uint64_t)c->num_cus;
        if (items > 0 && g > items) g = items;
        const size_t per_block = p.gslot_bytes + 16u * (size_t)ewfa::bt_chunks((uint32_t)p.dcap);
        g = std::min<uint64_t>(g, std::max<size_t>(GSLOT_BUDGET / per_block, 1));
        p.grid = (uint32_t)std::max<uint64_t>(g, 1);
        return EWFA_OK;
    }
    p.lds = p.reg        ? ewfa::lds_bytes_reg(p.seq_words, p.dcap, p.reg)
            : p.band_cap ? ewfa::lds_bytes_narrow(p.seq_words, p.band_cap)
                         : ewfa::lds_bytes(p.seq_words, p.dcap, p.bt_lds, p.off16);
    if (p.lds > lds_budget(c))
        return fail(EWFA_ERR_RANGE, "pass needs %zu bytes of LDS (device has %zu)", p.lds, c->lds_max);
    // (the block kernels' LDS limit was raised to the budget at ewfa_create)
    if (p.lds > 65536 && p.reg) return fail(EWFA_ERR_RANGE, "register pass above 64 KiB of LDS");
    if (p.persistent) {
        int bpc = p.reg ? ewfa::align_reg_blocks_per_cu(p.reg, p.lds)
                        : ewfa::align_blocks_per_cu(p.waves, p.lds, p.off16 ? 3 : p.band_cap ? 2 : 1);
        if (bpc < 1) bpc = 1;
        uint64_t g = (uint64_t)bpc * (uint64_t)c->num_cus;
        if (items > 0 && g > items) g = items;
        p.grid = (uint32_t)std::max<uint64_t>(g, 1);
    } else {
        p.grid = items;
    }
    return EWFA_OK;
}

// Pass for pairs up to blen bases under one bucket's policy.
int pass_from_bucket(ewfa_ctx* c, const ewfa::BucketPolicy& b, uint32_t blen, uint32_t items, Pass& p) {
    p = Pass{};
    p.seq_words = seq_words_for(blen);
    p.persistent = true;
    p.waves = b.waves;
    p.bt_lds = b.bt_lds;
    p.grab = std::max(1, b.grab);
    p.dcap = b.dcap;
    // the bucket's dcap is a request; shrink it to what LDS holds
    const int most = max_dcap_for(p.seq_words, p.bt_lds, lds_budget(c));
    if (most < 0) {
        // sequences too long for LDS: the block's slot goes to HBM
        p.gslot = 1;
        p.waves = ewfa::GSLOT_WAVES;
        p.bt_lds = 0;
        p.dcap = std::max(1, std::min({p.dcap, (int)blen + 1, gslot_max_dcap()}));
        return finish_pass(c, p, items);
    }
    if (p.dcap > most) p.dcap = most;
    if (p.dcap < 1) p.dcap = 1;
    // one wave with the backtrace in LDS: offsets in registers when they fit
    if (p.waves == 1 && p.bt_lds) {
        const int nc = ewfa::reg_chunks_for(p.dcap);
        // the banded kernel (one chunk) keeps plain copies, the others
        // overlapped ones
        const int ow = nc == 1 ? seq_words_for(blen) : overlapped_words_for(blen);
        if (nc > 0 && ewfa::lds_bytes_reg(ow, p.dcap, nc) <= 65536) {
            p.reg = nc;
            p.seq_words = ow;
        }
    }
    choose_off16(p, blen);
    choose_band_cap(c, p);
    return finish_pass(c, p, items);
}

// Main pass of a plain (reference launch model) batch.
int main_pass(ewfa_ctx* c, uint32_t max_len, uint32_t items, bool persistent, Pass& p) {
    if (persistent) return pass_from_bucket(c, c->policy.bucket_for(max_len), max_len, items, p);
    p = Pass{};
    p.seq_words = seq_words_for(max_len);
    p.persistent = false;
    // the reference's launch model: one workgroup of one wave per pair, a
    // distance budget of 64 levels (src/wavefront.cuh:465)
    p.waves = 1;
    p.bt_lds = 1;
    p.dcap = 64;
    p.grab = 1;
    return finish_pass(c, p, items);
}

// The persistent path's plan: one pass per length bucket that pairs up to
// max_len can fall in, run longest bucket first; a single bucket runs over
// the identity list without the bucketing kernel.
// Deferral passes run on the device after the bucket passes: each goes over
// the pairs the one before deferred (ping-pong lists whose counts the kernels
// read), with a growing wavefront budget, so the whole plan needs no host
// round trip.  A deep pass's HBM backtrace slots are bounded by giving it a
// smaller grid: it only ever runs the few pairs that get that far.
constexpr int MAX_CHAIN = 8;
constexpr size_t CHAIN_BT_BUDGET = (size_t)1 << 30;

struct Chain {
    int np = 0;
    Pass pass[MAX_CHAIN];
};

struct Plan {
    int nb = 0;
    ewfa::BucketTable table{};
    Pass pass[ewfa::MAX_BUCKETS];
    uint32_t blen[ewfa::MAX_BUCKETS];
    Chain chain;
    // streams the bucket passes are spread over (1: one after another); with
    // more, each bucket pass has slots of its own
    int fanout = 1;
    size_t bt_bytes = 0, runs_bytes = 0;  // regions the plan's slots need
    size_t gslot_bytes = 0;               // HBM slots of the passes beyond LDS
};

// Bytes of one block's slot in the backtrace region (block passes with the
// backtrace in HBM) or the run-row region (multi-chunk register passes).
size_t slot_bytes(const Pass& p);

// Gives each pass its slots: bucket passes that may run side by side get
// disjoint ranges, the deferral chain (after them) starts at 0 again.
void assign_slots(Plan& pl) {
    size_t bt = 0, runs = 0;
    for (int b = 0; b < pl.nb; ++b) {
        Pass& p = pl.pass[b];
        size_t& at = p.reg ? runs : bt;
        p.slot_off = pl.fanout > 1 ? at : 0;
        const size_t need = p.slot_off + slot_bytes(p) * p.grid;
        if (pl.fanout > 1) at = need;
        else at = std::max(at, need);
    }
    size_t gs = 0;
    for (int b = 0; b < pl.nb; ++b) {
        Pass& p = pl.pass[b];
        if (!p.gslot) continue;
        p.gslot_off = pl.fanout > 1 ? gs : 0;
        gs = std::max(gs, p.gslot_off + p.gslot_bytes * p.grid);
    }
    for (int k = 0; k < pl.chain.np; ++k) {
        Pass& p = pl.chain.pass[k];
        p.slot_off = 0;
        p.gslot_off = 0;
        bt = std::max(bt, slot_bytes(p) * p.grid);
        if (p.gslot) gs = std::max(gs, p.gslot_bytes * p.grid);
    }
    pl.bt_bytes = bt;
    pl.runs_bytes = runs;
    pl.gslot_bytes = gs;
}

int next_pass(ewfa_ctx* c, uint32_t max_len, int prev_dcap, uint32_t items, Pass& p);

void plan_chain(ewfa_ctx* c, uint32_t max_len, int dcap_min, uint32_t n, Chain& ch) {
    ch = Chain{};
    int dcap = dcap_min;
    while (ch.np < MAX_CHAIN) {
        Pass p;
        if (next_pass(c, max_len, dcap, n, p) != EWFA_OK) break;  // no larger budget fits
        if (!p.bt_lds) {
            const size_t slot = 16u * (size_t)ewfa::bt_chunks((uint32_t)p.dcap);
            const size_t most = std::max<size_t>(CHAIN_BT_BUDGET / slot, 1);
            if (p.grid > most) p.grid = (uint32_t)most;
        }
        ch.pass[ch.np++] = p;
        dcap = p.dcap;
    }
    g_err.clear();
}

int plan_persistent(ewfa_ctx* c, uint32_t max_len, uint32_t n, Plan& pl) {
    pl = Plan{};
    for (const ewfa::BucketPolicy& b : c->policy.buckets) {
        if (pl.nb == ewfa::MAX_BUCKETS) break;
        pl.table.hi[pl.nb] = b.len_hi;
        pl.blen[pl.nb] = std::min(b.len_hi, max_len);
        CHK(pass_from_bucket(c, b, pl.blen[pl.nb], n, pl.pass[pl.nb]));
        ++pl.nb;
        if (max_len <= b.len_hi) break;
    }
    pl.table.nb = pl.nb;
    int dcap_min = pl.pass[0].dcap;
    for (int b = 1; b < pl.nb; ++b) dcap_min = std::min(dcap_min, pl.pass[b].dcap);
    plan_chain(c, max_len, dcap_min, n, pl.chain);
    pl.fanout = std::min(pl.nb, c->fanout);
    assign_slots(pl);
    return EWFA_OK;
}

// Deferral pass after one with budget prev_dcap.
int next_pass(ewfa_ctx* c, uint32_t max_len, int prev_dcap, uint32_t items, Pass& p) {
    p = Pass{};
    p.seq_words = seq_words_for(max_len);
    p.persistent = true;
    p.grab = 1;
    p.waves = 4;
    const int most_global = max_dcap_for(p.seq_words, 0, lds_budget(c));
    int want = std::max(prev_dcap * 4, 256);
    // no pair needs more than max(plen, tlen) levels; a budget whose next
    // step would reach that goes all the way (one pass fewer in the chain,
    // whose passes mostly find no pairs)
    const int dmax = std::min((int)max_len + 1, gslot_max_dcap());
    if (want > dmax || (int64_t)want * 4 >= dmax) want = dmax;
    if (most_global <= prev_dcap) {
        // LDS holds no larger wavefront: the block's slot goes to HBM
        if (want <= prev_dcap) return fail(EWFA_ERR_RANGE, "no larger wavefront budget available");
        p.gslot = 1;
        p.waves = ewfa::GSLOT_WAVES;
        p.dcap = want;
        return finish_pass(c, p, items);
    }
    if (want > most_global) want = most_global;
    if (want <= prev_dcap) return fail(EWFA_ERR_RANGE, "no larger wavefront budget available");
    p.dcap = want;
    const int most_lds = max_dcap_for(p.seq_words, 1, c->lds_max / 2);
    p.bt_lds = (most_lds >= want) ? 1 : 0;
    if (want > 2048) p.waves = 8;
    choose_off16(p, max_len);
    return finish_pass(c, p, items);
}

uint64_t* ops_base(ewfa_ctx* c) {
    if (c->ext) return c->ext->ops;
    return c->ops_ext ? c->ops_ext : c->ops.as<uint64_t>();
}

// The scratch a call's kernels use: the external one of an allocation-free
// call, else the context's own buffers.
ewfa::Scratch scratch_of(ewfa_ctx* c) {
    if (c->ext) return *c->ext;
    ewfa::Scratch x{};
    x.ops = ops_base(c);
    x.ctrl = c->ctrl.as<uint32_t>();
    x.list_a = c->list_a.as<uint32_t>();
    x.list_b = c->list_b.as<uint32_t>();
    x.list_fb = c->list_fb.as<uint32_t>();
    x.buckets = c->buckets.as<uint32_t>();
    x.bound = c->bound.as<uint32_t>();
    x.order = c->order.as<uint32_t>();
    x.bt = c->bt.as<uint64_t>();
    x.runs = c->runs.as<uint8_t>();
    x.gslot = c->gslot.as<uint64_t>();
    return x;
}

int reserve_scratch(ewfa_ctx* c, size_t n, uint32_t max_len) {
    const uint32_t ops_stride = ewfa::ops_words(max_len + 1);
    const size_t ops_bytes = (size_t)n * ops_stride * 8;
    if (c->ops_ext) {
        if (c->ops_ext_bytes < ops_bytes)
            return fail(EWFA_ERR_RANGE, "op scratch of %zu bytes, %zu needed", c->ops_ext_bytes, ops_bytes);
    } else {
        CHK(c->ops.reserve(c, ops_bytes));
    }
    CHK(c->list_a.reserve(c, (size_t)std::max<size_t>(n, 1) * 4));
    CHK(c->list_b.reserve(c, (size_t)std::max<size_t>(n, 1) * 4));
    CHK(c->list_fb.reserve(c, (size_t)std::max<size_t>(n, 1) * 4));
    return EWFA_OK;
}

// Run scratch of a register pass with the emit fused (bytes per wave).
// The banded kernel keeps its run rows in LDS.
uint32_t runs_slot_of(const Pass& p) {
    if (p.reg == 1) return 0u;
    return ewfa::runs_slot_bytes((uint32_t)std::min(p.dcap, 32 * p.reg - 1));
}

size_t slot_bytes(const Pass& p) {
    if (p.reg) return runs_slot_of(p);
    if (p.bt_lds) return 0;
    return 16u * (size_t)ewfa::bt_chunks((uint32_t)p.dcap);
}

int reserve_bt(ewfa_ctx* c, const Pass& p) {
    if (p.reg) return c->runs.reserve(c, p.slot_off + (size_t)runs_slot_of(p) * p.grid);
    if (p.bt_lds) return EWFA_OK;
    if (p.gslot) CHK(c->gslot.reserve(c, p.gslot_off + p.gslot_bytes * p.grid));
    return c->bt.reserve(c, p.slot_off + slot_bytes(p) * p.grid);
}

int reserve_plan(ewfa_ctx* c, const Plan& pl, size_t n) {
    CHK(c->runs.reserve(c, pl.runs_bytes));
    CHK(c->bt.reserve(c, pl.bt_bytes));
    CHK(c->gslot.reserve(c, pl.gslot_bytes));
    if (pl.nb > 1) CHK(c->buckets.reserve(c, (size_t)pl.nb * std::max<size_t>(n, 1) * 4));
    CHK(c->bound.reserve(c, std::max<size_t>(n, 1) * 4));
    bool narrow = false;
    for (int b = 0; b < pl.nb; ++b) narrow = narrow || pl.pass[b].band_cap > 0;
    if (narrow) CHK(c->order.reserve(c, (size_t)std::max(pl.nb, 1) * std::max<size_t>(n, 1) * 4));
    return EWFA_OK;
}

int launch_pass(const ewfa::AlignArgs& a, const Pass& p, hipStream_t s) {
    if (p.reg)
        HIPCHK(ewfa::launch_align_reg(a, p.reg, p.persistent, p.grid, p.lds, s));
    else
        HIPCHK(ewfa::launch_align(a, p.waves, p.persistent, p.grid, p.lds, s));
    return EWFA_OK;
}

// Ahead of a deep block pass: each of its pairs' distance bound from the
// banded search (same list and level limit), for the pass to leave
// out the cells no path within the bound uses.  The saving grows with the
// distance while the bound's cost grows with the levels only: passes with
// budgets under BOUND_MIN_DCAP levels (where the measured saving did not pay
// for the bound), register passes, and sequences of more than
// BOUND_MAX_SEQ_WORDS words go without.  `queue` must hold 0.
int enqueue_bound(ewfa_ctx* c, ewfa::AlignArgs& a, const Pass& p, uint32_t* queue, uint32_t n, hipStream_t s) {
    a.bound = nullptr;
    // (a pass with narrow slots always gets its bounds, as it must read them)
    if (p.reg || p.dcap < BOUND_MIN_DCAP || p.seq_words > ewfa::BOUND_MAX_SEQ_WORDS || (n == 0 && !p.band_cap))
        return EWFA_OK;
    ewfa::AlignArgs b = a;
    b.queue = queue;
    b.bound = scratch_of(c).bound;
    // one wave per block, 32 a CU (the search holds no LDS)
    const uint32_t grid = (uint32_t)std::min<size_t>((size_t)c->num_cus * 32u, n);
    HIPCHK(ewfa::launch_bound(b, grid, s));
    a.bound = b.bound;
    return EWFA_OK;
}

ewfa::AlignArgs base_args(ewfa_ctx* c, const uint64_t* words, const ewfa_pair_t* pairs,
                          ewfa_result_t* results, uint32_t max_len) {
    ewfa::AlignArgs a{};
    a.words = words;
    a.pairs = pairs;
    a.results = results;
    a.ops = ops_base(c);
    a.ops_stride = ewfa::ops_words(max_len + 1);
    a.max_distance = c->opts.max_distance;
    a.max_len = max_len;
    return a;
}

void fill_pass_args(ewfa::AlignArgs& a, const Pass& p, ewfa_ctx* c) {
    a.dcap = p.dcap;
    a.bt_lds = p.bt_lds;
    a.seq_words = p.seq_words;
    a.grab = (uint32_t)p.grab;
    const ewfa::Scratch sx = scratch_of(c);
    a.bt_global = sx.bt + p.slot_off / 8;
    a.bt_slot_words = 2u * (uint64_t)ewfa::bt_chunks((uint32_t)p.dcap);
    a.off16 = p.off16;
    a.band_cap = p.band_cap;
    a.gslot = p.gslot ? sx.gslot + p.gslot_off / 8 : nullptr;
    a.gslot_words = p.gslot ? p.gslot_bytes / 8 : 0;
}

// The emit kernel for a pass: sequences staged in LDS when they fit (the
// persistent path), read from global memory otherwise (and on the plain
// path, which keeps the reference's one-lane-per-pair global walk).
// Emit over a short list whose count is known only on the device (deferred
// pairs, paths with long runs): one wave per pair with its sequences in LDS
// when they fit, else one lane per pair.  For whole buckets one lane per pair
// keeps more pairs in flight, which wins on long sequences.
// Pairs per wave of the one-lane-per-pair emit: EMIT_PER_WAVE when the
// launch fills every CU with at least 16 such waves, fewer (down to one) for
// smaller launches, so that few long pairs still get many waves.
uint32_t emit_per_wave(const ewfa_ctx* c, uint32_t items) {
    const uint32_t fill = (uint32_t)c->num_cus * 16u;
    return std::min<uint32_t>(std::max<uint32_t>(items / fill, 1u), (uint32_t)ewfa::EMIT_PER_WAVE);
}

int emit_list(ewfa_ctx* c, const ewfa::EmitArgs& e, uint32_t items, uint32_t max_len, hipStream_t s) {
    const int ow = overlapped_words_for(max_len);
    if ((size_t)2 * ow * 8 <= ewfa::EMIT_WAVE_MAX_LDS)
        HIPCHK(ewfa::launch_emit_wave(e, items, ow, (uint32_t)c->num_cus * 8u, s));
    else
        HIPCHK(ewfa::launch_emit(e, items, emit_per_wave(c, items), s));
    return EWFA_OK;
}

int emit(ewfa_ctx* c, const ewfa::EmitArgs& e, uint32_t items, uint32_t max_len, bool staged, hipStream_t s) {
    const int sw = seq_words_for(max_len);
    if (staged && sw <= ewfa::EMIT_LDS_MAX_SW)
        HIPCHK(ewfa::launch_emit_lds(e, items, sw, s));
    else
        HIPCHK(ewfa::launch_emit(e, items, emit_per_wave(c, items), s));
    return EWFA_OK;
}

// Enqueues the main pass plus emit for items [first, first+count) of the
// identity list, deferrals collected in list_a / ctrl[CTRL_DEFER_A].
int enqueue_main(ewfa_ctx* c, const Pass& p, const uint64_t* words, const ewfa_pair_t* pairs,
                 size_t first, uint32_t count, uint32_t max_len, ewfa_result_t* results,
                 int8_t* cigars, hipStream_t s, bool timing) {
    const ewfa::Scratch sx = scratch_of(c);
    uint32_t* ctrl = sx.ctrl;
    ewfa::AlignArgs a = base_args(c, words, pairs, results, max_len);
    fill_pass_args(a, p, c);
    // batch-local view: shift the pair/result bases, the ops scratch is per pair
    a.pairs = pairs + first;
    a.results = results + first;
    a.ops = ops_base(c) + first * (size_t)a.ops_stride;
    a.list = nullptr;
    a.count = count;
    a.count_ptr = nullptr;
    a.queue = p.persistent ? ctrl + CTRL_QUEUE : nullptr;
    a.defer_list = sx.list_a;
    a.defer_count = ctrl + CTRL_DEFER_A;
    HIPCHK(ewfa::launch_fill_u32(ctrl, 0u, 3, s));
    if (timing) HIPCHK(hipEventRecord(c->ev_a0, s));
    if (p.reg)
        HIPCHK(ewfa::launch_align_reg(a, p.reg, p.persistent, p.grid, p.lds, s));
    else
        HIPCHK(ewfa::launch_align(a, p.waves, p.persistent, p.grid, p.lds, s));
    if (timing) HIPCHK(hipEventRecord(c->ev_a1, s));
    ewfa::EmitArgs e{};
    e.words = words;
    e.pairs = pairs + first;
    e.results = results + first;
    e.ops = a.ops;
    e.ops_stride = a.ops_stride;
    e.cigars = cigars;
    e.list = nullptr;
    e.count = count;
    e.count_ptr = nullptr;
    CHK(emit(c, e, count, max_len, p.persistent, s));
    if (timing) HIPCHK(hipEventRecord(c->ev_e1, s));
    HIPCHK(hipMemcpyAsync(c->host_ctrl + CTRL_DEFER_A, ctrl + CTRL_DEFER_A, 4,
                          hipMemcpyDeviceToHost, s));
    return EWFA_OK;
}

// Enqueues a persistent plan: bucketing, one align pass per bucket (longest
// first), the emit per bucket, and the deferred count's copy to the host.
// The deferral chain of a plan, then the marking of pairs no pass could
// resolve.
int enqueue_chain(ewfa_ctx* c, const Chain& ch, const uint64_t* words, const ewfa_pair_t* pairs, uint32_t n,
                  uint32_t max_len, ewfa_result_t* results, int8_t* cigars, hipStream_t s) {
    const ewfa::Scratch sx = scratch_of(c);
    uint32_t* ctrl = sx.ctrl;
    uint32_t* lists[2] = {sx.list_a, sx.list_b};
    const int ctr[2] = {CTRL_DEFER_A, CTRL_DEFER_B};
    int cur = 0;
    for (int k = 0; k < ch.np; ++k) {
        const Pass& p = ch.pass[k];
        // the pass's queue head, the count of the list it defers into and
        // its bounds' work counter start at zero (one launch)
        HIPCHK(ewfa::launch_clear_words(
            ctrl, (1ull << CTRL_QUEUE) | (1ull << ctr[1 - cur]) | (1ull << CTRL_CHAIN_BQ), s));
        ewfa::AlignArgs a = base_args(c, words, pairs, results, max_len);
        fill_pass_args(a, p, c);
        a.list = lists[cur];
        a.count = n;
        a.count_ptr = ctrl + ctr[cur];
        a.queue = ctrl + CTRL_QUEUE;
        a.defer_list = lists[1 - cur];
        a.defer_count = ctrl + ctr[1 - cur];
        CHK(enqueue_bound(c, a, p, ctrl + CTRL_CHAIN_BQ, n, s));
        CHK(launch_pass(a, p, s));
        ewfa::EmitArgs e{};
        e.words = words;
        e.pairs = pairs;
        e.results = results;
        e.ops = ops_base(c);
        e.ops_stride = ewfa::ops_words(max_len + 1);
        e.cigars = cigars;
        e.list = lists[cur];
        e.count = n;
        e.count_ptr = ctrl + ctr[cur];
        CHK(emit_list(c, e, n, max_len, s));
        cur = 1 - cur;
    }
    HIPCHK(ewfa::launch_mark_too_distant(lists[cur], ctrl + ctr[cur], results, n, s));
    return EWFA_OK;
}

// phase: PLAN_ALL, or PLAN_ALIGN (bucketing, the bucket passes and the emits
// of the block buckets) then PLAN_EMIT (the fused passes' fallback emit and
// the deferral chain) as two graphs timed between their launches.
enum { PLAN_ALL = 3, PLAN_ALIGN = 1, PLAN_EMIT = 2 };

// The streams (and join events) a plan's bucket passes fan out over beyond
// the context's own, created on the first plan that has more than one
// bucket: a context that only ever runs one bucket (a CLI run on one read
// length) never pays for them.
int ensure_aux(ewfa_ctx* c, int F) {
    for (int i = 0; i < F - 1; ++i) {
        if (!c->aux[i] && hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking) != hipSuccess) {
            c->aux[i] = nullptr;
            return fail(EWFA_ERR_HIP, "aux stream creation failed");
        }
        if (!c->ev_join[i] && hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming) != hipSuccess) {
            c->ev_join[i] = nullptr;
            return fail(EWFA_ERR_HIP, "join event creation failed");
        }
    }
    return EWFA_OK;
}

int enqueue_plan(ewfa_ctx* c, const Plan& pl, const uint64_t* words, const ewfa_pair_t* pairs, uint32_t n,
                 uint32_t max_len, ewfa_result_t* results, int8_t* cigars, hipStream_t s, bool timing,
                 int phase = PLAN_ALL) {
    const ewfa::Scratch sx = scratch_of(c);
    uint32_t* ctrl = sx.ctrl;
    uint32_t* lists = sx.buckets;
    const bool bucketed = pl.nb > 1;
    const int F = std::max(1, std::min(pl.fanout, ewfa_ctx::MAX_FANOUT));
    if (phase & PLAN_ALIGN) CHK(ensure_aux(c, F));
    hipStream_t st[ewfa_ctx::MAX_FANOUT] = {s};
    for (int i = 1; i < F; ++i) st[i] = c->aux[i - 1];
    bool fused = false;
    for (int b = 0; b < pl.nb; ++b) fused = fused || pl.pass[b].reg;
    if (phase & PLAN_ALIGN) {
        HIPCHK(ewfa::launch_fill_u32(ctrl, 0u, CTRL_XQ + (size_t)pl.nb * ewfa::XCDS * ewfa::XQ_STRIDE, s));
        if (bucketed) HIPCHK(ewfa::launch_bucketize(pairs, n, pl.table, lists, ctrl + CTRL_BCOUNT, s));
        if (timing) HIPCHK(hipEventRecord(c->ev_a0, s));
        if (F > 1) {
            HIPCHK(hipEventRecord(c->ev_fork, s));
            for (int i = 1; i < F; ++i) HIPCHK(hipStreamWaitEvent(st[i], c->ev_fork, 0));
        }
        // longest bucket first, the passes dealt over the streams in turn;
        // each block bucket's emit follows its pass on the same stream
        int k = 0;
        for (int b = pl.nb - 1; b >= 0; --b, ++k) {
            const Pass& p = pl.pass[b];
            hipStream_t q = st[k % F];
            ewfa::AlignArgs a = base_args(c, words, pairs, results, max_len);
            fill_pass_args(a, p, c);
            a.list = bucketed ? lists + (size_t)b * n : nullptr;
            a.count = n;
            a.count_ptr = bucketed ? ctrl + CTRL_BCOUNT + b : nullptr;
            a.queue = ctrl + CTRL_BQUEUE + b;
            a.xqueue = ctrl + CTRL_XQ + (size_t)b * ewfa::XCDS * ewfa::XQ_STRIDE;
            a.defer_list = sx.list_a;
            a.defer_count = ctrl + CTRL_DEFER_A;
            a.max_len = pl.blen[b];
            if (p.reg) {  // the register kernel writes its pairs' streams itself
                a.runs = sx.runs + p.slot_off;
                a.runs_slot = runs_slot_of(p);
                a.cigars = cigars;
                a.fallback_list = sx.list_fb;
                a.fallback_count = ctrl + CTRL_FALLBACK;
            }
            CHK(enqueue_bound(c, a, p, ctrl + CTRL_BOUNDQ + b, n, q));
            if (p.band_cap > 0 && a.bound && sx.order) {
                // the pass takes its pairs by decreasing bound, share by share
                uint32_t* ord = sx.order + (size_t)b * n;
                int shift = 0;
                while ((p.band_cap >> shift) >= ewfa::ORDER_BINS) ++shift;
                HIPCHK(ewfa::launch_order_by_bound(a.list, a.count_ptr, a.count, a.bound, ord, shift, q));
                a.list = ord;
            }
            CHK(launch_pass(a, p, q));
            if (!p.reg) {
                ewfa::EmitArgs e{};
                e.words = words;
                e.pairs = pairs;
                e.results = results;
                e.ops = ops_base(c);
                e.ops_stride = ewfa::ops_words(max_len + 1);
                e.cigars = cigars;
                e.list = bucketed ? lists + (size_t)b * n : nullptr;
                e.count = n;
                e.count_ptr = bucketed ? ctrl + CTRL_BCOUNT + b : nullptr;
                CHK(emit(c, e, n, pl.blen[b], true, q));
            }
        }
        for (int i = 1; i < F; ++i) {
            HIPCHK(hipEventRecord(c->ev_join[i - 1], st[i]));
            HIPCHK(hipStreamWaitEvent(s, c->ev_join[i - 1], 0));
        }
        if (timing) HIPCHK(hipEventRecord(c->ev_a1, s));
    }
    if (!(phase & PLAN_EMIT)) return EWFA_OK;
    if (fused) {
        // paths with a match run of 255 bases or more go through the emit kernel
        ewfa::EmitArgs e{};
        e.words = words;
        e.pairs = pairs;
        e.results = results;
        e.ops = ops_base(c);
        e.ops_stride = ewfa::ops_words(max_len + 1);
        e.cigars = cigars;
        e.list = sx.list_fb;
        e.count = n;
        e.count_ptr = ctrl + CTRL_FALLBACK;
        CHK(emit_list(c, e, n, max_len, s));
    }
    CHK(enqueue_chain(c, pl.chain, words, pairs, n, max_len, results, cigars, s));
    if (timing) HIPCHK(hipEventRecord(c->ev_e1, s));
    return EWFA_OK;
}

// Resolves pairs deferred by the main pass of a batch: repeated passes with a
// growing wavefront budget over the deferred list, then emit for them.
int resolve_deferred(ewfa_ctx* c, int prev_dcap, const uint64_t* words, const ewfa_pair_t* pairs,
                     size_t first, uint32_t max_len, ewfa_result_t* results, int8_t* cigars,
                     hipStream_t s) {
    uint32_t* ctrl = c->ctrl.as<uint32_t>();
    uint32_t pending = c->host_ctrl[CTRL_DEFER_A];
    uint32_t* in_list = c->list_a.as<uint32_t>();
    uint32_t* out_list = c->list_b.as<uint32_t>();
    int in_ctr = CTRL_DEFER_A, out_ctr = CTRL_DEFER_B;
    int dcap = prev_dcap;
    while (pending > 0) {
        Pass p;
        int rc = next_pass(c, max_len, dcap, pending, p);
        if (rc != EWFA_OK) {
            // no budget left: these pairs are beyond what the device can stage
            std::vector<uint32_t> ids(pending);
            HIPCHK(hipMemcpyAsync(ids.data(), in_list, pending * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (uint32_t id : ids) {
                ewfa_result_t r{0, 0, EWFA_PAIR_TOO_DISTANT, 0};
                HIPCHK(hipMemcpyAsync(results + first + id, &r, sizeof r, hipMemcpyHostToDevice, s));
                HIPCHK(hipStreamSynchronize(s));
            }
            g_err.clear();
            return EWFA_OK;
        }
        CHK(reserve_bt(c, p));
        ewfa::AlignArgs a = base_args(c, words, pairs, results, max_len);
        fill_pass_args(a, p, c);
        a.pairs = pairs + first;
        a.results = results + first;
        a.ops = ops_base(c) + first * (size_t)a.ops_stride;
        a.list = in_list;
        a.count = pending;
        a.count_ptr = nullptr;
        a.queue = ctrl + CTRL_QUEUE;
        a.defer_list = out_list;
        a.defer_count = ctrl + out_ctr;
        HIPCHK(ewfa::launch_clear_words(ctrl, (1ull << CTRL_QUEUE) | (1ull << out_ctr), s));
        HIPCHK(ewfa::launch_align(a, p.waves, true, p.grid, p.lds, s));
        ewfa::EmitArgs e{};
        e.words = words;
        e.pairs = pairs + first;
        e.results = results + first;
        e.ops = a.ops;
        e.ops_stride = a.ops_stride;
        e.cigars = cigars;
        e.list = in_list;
        e.count = pending;
        CHK(emit(c, e, pending, max_len, true, s));
        HIPCHK(hipMemcpyAsync(c->host_ctrl + out_ctr, ctrl + out_ctr, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        pending = c->host_ctrl[out_ctr];
        std::swap(in_list, out_list);
        std::swap(in_ctr, out_ctr);
        dcap = p.dcap;
    }
    return EWFA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host pipeline of the persistent path.  The pairs go in chunks (at most
// PIPE_CHUNK pairs and PIPE_CAPS bytes of CIGAR regions each).  All chunks'
// packed words and descriptors are queued on the upload stream at once, so
// chunk k computes as soon as its own uploads land.  Each chunk's written
// CIGAR bytes are packed by the device straight into pinned host memory,
// beside its results, and the host places them while the next chunk
// computes.  Two stages alternate.

constexpr size_t PIPE_CHUNK_MIN = (size_t)1 << 17;
constexpr size_t PIPE_CHUNKS = 4;
constexpr size_t PIPE_CAPS = (size_t)256 << 20;

struct Chunk {
    size_t lo, hi, w_lo, w_hi, caps;
    uint32_t max_len;
};

// Pinned host memory of a context.  coherent: memory the device writes with
// its own stores (they reach host memory without a cache flush); copies to
// and from the other kind run on the DMA engines instead of a copy kernel on
// the CUs.  A replaced buffer is retired, not freed (freeing pinned memory
// waits for the whole device).
int grow_pinned(ewfa_ctx* c, void** p, size_t* cap, size_t bytes, bool coherent) {
    if (bytes <= *cap && *p) return EWFA_OK;
    if (*p) c->retired_host.push_back(*p);
    *p = nullptr;
    *cap = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 20);
    if (hipHostMalloc(p, want, coherent ? hipHostMallocCoherent : hipHostMallocNonCoherent) != hipSuccess) {
        *p = nullptr;
        return fail(EWFA_ERR_NOMEM, "hipHostMalloc(%zu) failed", want);
    }
    *cap = want;
    return EWFA_OK;
}

int pipeline_setup(ewfa_ctx* c, size_t chunks) {
    if (!c->up) HIPCHK(hipStreamCreateWithFlags(&c->up, hipStreamNonBlocking));
    while (c->ev_up.size() < chunks) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_up.push_back(e);
    }
    for (ewfa_ctx::Stage& st : c->stage)
        if (!st.done) HIPCHK(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
    return EWFA_OK;
}

// Enqueues chunk k's uploads (words only when h_words is given) on the
// upload stream.  Every chunk's words and descriptors have their own place
// on the device, so all uploads go in at once.
int pipeline_upload(ewfa_ctx* c, const Chunk& ch, size_t k, const uint64_t* h_words, const ewfa_pair_t* h_pairs) {
    uint64_t* d_words = c->h_words.as<uint64_t>();
    ewfa_pair_t* d_pairs = c->h_pairs.as<ewfa_pair_t>();
    const size_t n = ch.hi - ch.lo;
    if (h_words && ch.w_hi > ch.w_lo)
        HIPCHK(hipMemcpyAsync(d_words + ch.w_lo, h_words + ch.w_lo, (ch.w_hi - ch.w_lo) * 8, hipMemcpyHostToDevice,
                              c->up));
    HIPCHK(hipMemcpyAsync(d_pairs + ch.lo, h_pairs + ch.lo, n * sizeof(ewfa_pair_t), hipMemcpyHostToDevice, c->up));
    HIPCHK(hipEventRecord(c->ev_up[k], c->up));
    return EWFA_OK;
}

// Enqueues chunk k on the compute stream, after its uploads: the persistent
// plan, the compaction into the stage's pinned buffer and the results copy.
// The stage must be free (its previous chunk placed by the host).
int pipeline_enqueue(ewfa_ctx* c, const Chunk& ch, size_t k) {
    uint64_t* d_words = c->h_words.as<uint64_t>();
    ewfa_pair_t* d_pairs = c->h_pairs.as<ewfa_pair_t>();
    ewfa_result_t* d_results = c->h_results.as<ewfa_result_t>();
    int8_t* d_cigars = c->h_cigars.as<int8_t>();
    const size_t n = ch.hi - ch.lo;
    HIPCHK(hipStreamWaitEvent(c->stream, c->ev_up[k], 0));
    ewfa_ctx::Stage& st = c->stage[k & 1];
    if (ch.caps > st.dense_cap || !st.dense) {
        CHK(grow_pinned(c, (void**)&st.dense, &st.dense_cap, ch.caps, true));
        void* dp = nullptr;
        HIPCHK(hipHostGetDevicePointer(&dp, st.dense, 0));
        st.dense_dev = static_cast<int8_t*>(dp);
    }
    size_t res_bytes = st.res_cap;
    CHK(grow_pinned(c, (void**)&st.res, &res_bytes, n * sizeof(ewfa_result_t), false));
    st.res_cap = res_bytes;
    Plan pl;
    CHK(reserve_scratch(c, n, ch.max_len));
    CHK(plan_persistent(c, ch.max_len, (uint32_t)n, pl));
    CHK(reserve_plan(c, pl, n));
    CHK(enqueue_plan(c, pl, d_words, d_pairs + ch.lo, (uint32_t)n, ch.max_len, d_results + ch.lo, d_cigars,
                     c->stream, false));
    CHK(c->scan.reserve(c, 16 * ((n + ewfa::COMPACT_BLOCK - 1) / ewfa::COMPACT_BLOCK)));
    HIPCHK(ewfa::launch_compact(d_pairs + ch.lo, d_results + ch.lo, d_cigars, n, c->scan.as<uint64_t>(), st.dense_dev,
                                c->stream));
    HIPCHK(hipMemcpyAsync(st.res, d_results + ch.lo, n * sizeof(ewfa_result_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(st.done, c->stream));
    return EWFA_OK;
}

int pipeline_complete(ewfa_ctx* c, const Chunk& ch, size_t k, const ewfa_pair_t* h_pairs, ewfa_result_t* h_results,
                      int8_t* h_cigars) {
    ewfa_ctx::Stage& st = c->stage[k & 1];
    HIPCHK(hipEventSynchronize(st.done));
    const size_t n = ch.hi - ch.lo;
    memcpy(h_results + ch.lo, st.res, n * sizeof(ewfa_result_t));
    if (ewfa::scatter_streams(st.dense, ch.caps, h_pairs + ch.lo, h_results + ch.lo, n, h_cigars, 16) == SIZE_MAX)
        return fail(EWFA_ERR_STATE, "compacted streams exceed their regions");
    return EWFA_OK;
}

// h_words null: the words of the last ewfa_upload_words are resident.  The
// descriptors are checked while the chunks are formed.
// Words per piece of an asynchronous upload (16 MiB).
constexpr size_t WORDS_PIECE = (size_t)2 << 20;

// Waits for the pieces of an asynchronous words upload still in flight
// (before anything else writes the context's words).
int words_settle(ewfa_ctx* c) {
    if (c->words_pieces) HIPCHK(hipStreamSynchronize(c->wup));
    c->words_pieces = 0;
    return EWFA_OK;
}

// Makes the context's stream wait for the pieces of an asynchronous upload
// that hold words below w_hi (the pieces land in order).
int words_wait(ewfa_ctx* c, size_t w_hi) {
    if (!c->words_pieces || w_hi == 0) return EWFA_OK;
    const size_t k = std::min((w_hi - 1) / WORDS_PIECE, c->words_pieces - 1);
    HIPCHK(hipStreamWaitEvent(c->stream, c->ev_words[k], 0));
    return EWFA_OK;
}

int run_pipeline(ewfa_ctx* c, const uint64_t* h_words, size_t n_words, const ewfa_pair_t* h_pairs, size_t n,
                 ewfa_result_t* h_results, int8_t* h_cigars, size_t cigar_bytes) {
    if (h_words) {
        CHK(words_settle(c));
        CHK(c->h_words.reserve(c, std::max<size_t>(n_words, 1) * 8));
        c->resident_words = 0;  // only the words these pairs read go up
    }
    CHK(c->h_pairs.reserve(c, n * sizeof(ewfa_pair_t)));
    CHK(c->h_results.reserve(c, n * sizeof(ewfa_result_t)));
    CHK(c->h_cigars.reserve(c, std::max<size_t>(cigar_bytes, 4)));
    // earlier work on the context's stream finishes before the uploads land
    HIPCHK(hipEventRecord(c->ev_in, c->stream));
    HIPCHK(hipStreamWaitEvent(c->up, c->ev_in, 0));
    // form the chunks, checking the descriptors; each chunk's uploads go in
    // as soon as it is formed, so the checks of the next overlap them
    std::vector<Chunk> chunks;
    uint32_t max_len = 0;
    size_t most = 0;
    const size_t per = std::max(PIPE_CHUNK_MIN, (n + PIPE_CHUNKS - 1) / PIPE_CHUNKS);
    int rc = EWFA_OK;
    for (size_t i = 0; i < n && rc == EWFA_OK;) {
        Chunk ch{i, i, SIZE_MAX, 0, 0, 0};
        while (ch.hi < n && ch.hi - ch.lo < per &&
               (ch.hi == ch.lo || ch.caps + h_pairs[ch.hi].cigar_cap <= PIPE_CAPS)) {
            const ewfa_pair_t& p = h_pairs[ch.hi];
            if (p.pattern_off + ewfa_packed_words(p.plen) - 1 > n_words ||
                p.text_off + ewfa_packed_words(p.tlen) - 1 > n_words) {
                rc = fail(EWFA_ERR_ARG, "pair %zu reads past the packed words", ch.hi);
                break;
            }
            if ((p.cigar_off & 3) || (p.cigar_cap & 3) || p.cigar_off + p.cigar_cap > cigar_bytes) {
                rc = fail(EWFA_ERR_ARG, "pair %zu has a bad CIGAR region", ch.hi);
                break;
            }
            ch.w_lo = std::min<size_t>(ch.w_lo, std::min(p.pattern_off, p.text_off));
            ch.w_hi = std::max<size_t>(ch.w_hi, std::max(p.pattern_off + ewfa_packed_words(p.plen),
                                                         p.text_off + ewfa_packed_words(p.tlen)));
            ch.caps += p.cigar_cap;
            ch.max_len = std::max(ch.max_len, std::max(p.plen, p.tlen));
            ++ch.hi;
        }
        if (rc != EWFA_OK) break;
        ch.w_hi = std::min(ch.w_hi, n_words);
        if (ch.w_lo > ch.w_hi) ch.w_lo = ch.w_hi;
        max_len = std::max(max_len, ch.max_len);
        most = std::max(most, ch.hi - ch.lo);
        chunks.push_back(ch);
        rc = pipeline_setup(c, chunks.size());
        if (rc == EWFA_OK) rc = pipeline_upload(c, ch, chunks.size() - 1, h_words, h_pairs);
        i = ch.hi;
    }
    // size the per-chunk scratch for the largest chunk before any chunk
    // computes, so that no buffer is replaced while chunks are in flight
    if (rc == EWFA_OK) rc = reserve_scratch(c, most, max_len);
    size_t enq = 0, done = 0;
    for (size_t k = 0; k < chunks.size() && rc == EWFA_OK; ++k) {
        if (!h_words) rc = words_wait(c, chunks[k].w_hi);
        if (rc == EWFA_OK) rc = pipeline_enqueue(c, chunks[k], k);
        if (rc == EWFA_OK) ++enq;
        if (rc == EWFA_OK && k >= 1) {
            rc = pipeline_complete(c, chunks[k - 1], k - 1, h_pairs, h_results, h_cigars);
            if (rc == EWFA_OK) ++done;
        }
    }
    if (rc == EWFA_OK && done < enq) rc = pipeline_complete(c, chunks[enq - 1], enq - 1, h_pairs, h_results, h_cigars);
    if (rc != EWFA_OK) {
        // leave nothing in flight that uses the context's buffers
        (void)hipStreamSynchronize(c->up);
        (void)hipStreamSynchronize(c->stream);
        return rc;
    }
    return EWFA_OK;
}

namespace ewfa {
int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

size_t op_scratch_bytes(size_t n, uint32_t max_len) { return n * (size_t)ops_words(max_len + 1) * 8; }

int align_device_with_scratch(ewfa_ctx_t* c, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n,
                              uint32_t max_len, ewfa_result_t* d_results, int8_t* d_cigars, void* stream,
                              uint64_t* d_ops, size_t ops_bytes) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (!d_ops) return fail(EWFA_ERR_ARG, "null op scratch");
    c->ops_ext = d_ops;
    c->ops_ext_bytes = ops_bytes;
    const int rc = ewfa_align_device(c, d_words, d_pairs, n, max_len, d_results, d_cigars, stream);
    c->ops_ext = nullptr;
    c->ops_ext_bytes = 0;
    return rc;
}

hipStream_t context_stream(ewfa_ctx_t* c) { return c ? c->stream : nullptr; }

int context_device(ewfa_ctx_t* c) { return c ? c->device : -1; }

int align_device_in(ewfa_ctx_t* c, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n, uint32_t max_len,
                    ewfa_result_t* d_results, int8_t* d_cigars, void* stream, void* scratch, size_t bytes) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (c->opts.mode != EWFA_MODE_PERSISTENT) return fail(EWFA_ERR_STATE, "the allocation-free call needs the persistent mode");
    if (n == 0) return EWFA_OK;
    if (!d_words || !d_pairs || !d_results || !d_cigars || !scratch) return fail(EWFA_ERR_ARG, "null buffer");
    if (n > 0xFFFFFFF0u) return fail(EWFA_ERR_RANGE, "too many pairs in one call");
    CHK(set_device(c));
    Plan pl;
    CHK(plan_persistent(c, max_len, (uint32_t)n, pl));
    // one stream: every pass's slots share the region.  The static scratch
    // has no room for slots in HBM: the deferral chain stops before the first
    // such pass (its pairs come back TOO_DISTANT), and sequences too long for
    // LDS are refused.
    for (int b = 0; b < pl.nb; ++b)
        if (pl.pass[b].gslot)
            return fail(EWFA_ERR_RANGE, "sequences of %u bases do not fit in LDS, which the allocation-free call "
                        "needs", max_len);
    for (int k = 0; k < pl.chain.np; ++k)
        if (pl.chain.pass[k].gslot) {
            pl.chain.np = k;
            break;
        }
    pl.fanout = 1;
    assign_slots(pl);
    // per-pair buffers first, then one region for the passes' slots
    const auto r256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t ops_bytes = op_scratch_bytes(n, max_len);
    const size_t list = r256(n * 4);
    const size_t fixed = r256(ops_bytes) + r256(CTRL_WORDS * 4) + 4 * list + (pl.nb > 1 ? r256((size_t)pl.nb * n * 4) : 0);
    if (fixed > bytes)
        return fail(EWFA_ERR_RANGE, "%zu pairs of up to %u bases need %zu bytes of scratch before any pass; the "
                    "allocation-free call holds %zu", n, max_len, fixed, bytes);
    char* b = static_cast<char*>(scratch);
    Scratch x{};
    size_t off = 0;
    const auto take = [&](size_t sz) { char* p = b + off; off += r256(sz); return p; };
    x.ops = reinterpret_cast<uint64_t*>(take(ops_bytes));
    x.ctrl = reinterpret_cast<uint32_t*>(take(CTRL_WORDS * 4));
    x.list_a = reinterpret_cast<uint32_t*>(take(n * 4));
    x.list_b = reinterpret_cast<uint32_t*>(take(n * 4));
    x.list_fb = reinterpret_cast<uint32_t*>(take(n * 4));
    x.bound = reinterpret_cast<uint32_t*>(take(n * 4));
    x.buckets = pl.nb > 1 ? reinterpret_cast<uint32_t*>(take((size_t)pl.nb * n * 4)) : nullptr;
    x.order = nullptr;  // bounded passes keep their lists in order
    x.region = bytes - off;
    x.bt = reinterpret_cast<uint64_t*>(b + off);
    x.runs = reinterpret_cast<uint8_t*>(b + off);
    // each pass gets at most as many persistent blocks as slots fit the region
    const auto fit = [&](Pass& p) -> int {
        size_t slot = 0;
        if (p.reg) slot = runs_slot_of(p);
        else if (!p.bt_lds) slot = 16u * (size_t)bt_chunks((uint32_t)p.dcap);
        if (slot == 0) return EWFA_OK;
        const size_t most = x.region / slot;
        if (most == 0)
            return fail(EWFA_ERR_RANGE, "a pass of %d levels needs a %zu-byte backtrace slot; %zu bytes of the "
                        "allocation-free call's scratch are left for it", p.dcap, slot, x.region);
        if (p.grid > most) p.grid = (uint32_t)most;
        return EWFA_OK;
    };
    for (int k = 0; k < pl.nb; ++k) CHK(fit(pl.pass[k]));
    for (int k = 0; k < pl.chain.np; ++k) CHK(fit(pl.chain.pass[k]));
    hipStream_t s = c->stream;
    HIPCHK(hipEventRecord(c->ev_in, (hipStream_t)stream));
    HIPCHK(hipStreamWaitEvent(s, c->ev_in, 0));
    c->align_ms = c->emit_ms = 0.f;
    c->ext = &x;
    const int rc = enqueue_plan(c, pl, d_words, d_pairs, (uint32_t)n, max_len, d_results, d_cigars, s, c->timing);
    c->ext = nullptr;
    CHK(rc);
    HIPCHK(hipStreamSynchronize(s));
    if (c->timing) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_a0, c->ev_a1) == hipSuccess) c->align_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev_a1, c->ev_e1) == hipSuccess) c->emit_ms = ms;
    }
    return EWFA_OK;
}

int enqueue_align(ewfa_ctx_t* c, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n, uint32_t max_len,
                  ewfa_result_t* d_results, int8_t* d_cigars, uint64_t* d_ops, size_t ops_bytes) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (c->opts.mode != EWFA_MODE_PERSISTENT) return fail(EWFA_ERR_STATE, "enqueue needs the persistent mode");
    if (n == 0) return EWFA_OK;
    if (!d_words || !d_pairs || !d_results || !d_cigars || !d_ops) return fail(EWFA_ERR_ARG, "null buffer");
    if (n > 0xFFFFFFF0u) return fail(EWFA_ERR_RANGE, "too many pairs in one call");
    CHK(set_device(c));
    c->ops_ext = d_ops;
    c->ops_ext_bytes = ops_bytes;
    Plan pl;
    int rc = reserve_scratch(c, n, max_len);
    if (rc == EWFA_OK) rc = plan_persistent(c, max_len, (uint32_t)n, pl);
    if (rc == EWFA_OK) rc = reserve_plan(c, pl, n);
    if (rc == EWFA_OK)
        rc = enqueue_plan(c, pl, d_words, d_pairs, (uint32_t)n, max_len, d_results, d_cigars, c->stream, false);
    c->ops_ext = nullptr;
    c->ops_ext_bytes = 0;
    return rc;
}

int enqueue_compact(ewfa_ctx_t* c, const ewfa_pair_t* d_pairs, const ewfa_result_t* d_results,
                    const int8_t* d_cigars, size_t n, size_t bound, hipStream_t s) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (n == 0) return EWFA_OK;
    CHK(set_device(c));
    const size_t nb = (n + COMPACT_BLOCK - 1) / COMPACT_BLOCK;
    CHK(c->dense.reserve(c, std::max<size_t>(bound, 4)));
    CHK(c->scan.reserve(c, nb * 16));
    HIPCHK(launch_compact(d_pairs, d_results, d_cigars, n, c->scan.as<uint64_t>(), c->dense.as<int8_t>(), s));
    return EWFA_OK;
}

int fetch_compact(ewfa_ctx_t* c, size_t total, hipStream_t s, const int8_t** out) {
    if (!c || !out) return fail(EWFA_ERR_ARG, "null argument");
    if (total > c->dense.bytes) return fail(EWFA_ERR_STATE, "%zu dense bytes, buffer of %zu", total, c->dense.bytes);
    if (total > c->h_stage_cap) {
        if (c->h_stage) (void)hipHostFree(c->h_stage);
        c->h_stage = nullptr;
        c->h_stage_cap = 0;
        const size_t cap = std::max<size_t>(total + total / 4, 1 << 20);
        if (hipHostMalloc((void**)&c->h_stage, cap, hipHostMallocDefault) != hipSuccess) {
            c->h_stage = nullptr;
            return fail(EWFA_ERR_NOMEM, "hipHostMalloc(%zu) failed", cap);
        }
        c->h_stage_cap = cap;
    }
    if (total) HIPCHK(hipMemcpyAsync(c->h_stage, c->dense.p, total, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *out = c->h_stage;
    return EWFA_OK;
}

int set_context_policy(ewfa_ctx_t* c, const Policy& p) {
    if (!c || p.buckets.empty()) return fail(EWFA_ERR_ARG, "empty policy");
    c->policy = p;
    c->policy.buckets.back().len_hi = 0xFFFFFFFFu;
    ++c->policy_gen;
    return EWFA_OK;
}
}  // namespace ewfa

extern "C" {

void ewfa_default_opts(ewfa_opts_t* o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->device = 0;
    o->mode = EWFA_MODE_PERSISTENT;
    o->max_distance = 0;
    o->batch_size = 0;
    o->policy_path = nullptr;
}

const char* ewfa_last_error(void) { return g_err.c_str(); }

int ewfa_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ewfa_device_info(int device, char* name, int name_cap, int* major, int* minor) {
    int n = ewfa_device_count();
    if (n <= 0) return fail(EWFA_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= n) return fail(EWFA_ERR_ARG, "device %d out of range", device);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (name && name_cap > 0) snprintf(name, (size_t)name_cap, "%s", prop.name);
    if (major) *major = prop.major;
    if (minor) *minor = prop.minor;
    return EWFA_OK;
}

int ewfa_create(ewfa_ctx_t** out, const ewfa_opts_t* opts) {
    if (!out) return fail(EWFA_ERR_ARG, "ctx is NULL");
    *out = nullptr;
    ewfa_opts_t o;
    ewfa_default_opts(&o);
    if (opts) o = *opts;
    if (o.mode != EWFA_MODE_PLAIN && o.mode != EWFA_MODE_PERSISTENT)
        return fail(EWFA_ERR_ARG, "unknown mode %d", o.mode);
    if (o.max_distance < 0 || o.batch_size < 0) return fail(EWFA_ERR_ARG, "negative option");
    int n = ewfa_device_count();
    if (n <= 0) return fail(EWFA_ERR_NODEV, "no HIP device");
    if (o.device < 0 || o.device >= n) return fail(EWFA_ERR_ARG, "device %d out of range", o.device);
    ewfa_ctx* c = new ewfa_ctx();
    c->device = o.device;
    c->opts = o;
    if (o.policy_path) {
        c->policy_path = o.policy_path;
        c->opts.policy_path = c->policy_path.c_str();
    }
    int rc = set_device(c);
    if (rc == EWFA_OK) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) rc = fail(EWFA_ERR_HIP, "props");
        else {
            c->num_cus = prop.multiProcessorCount;
            // per-block LDS a kernel may opt into (160 KiB on gfx950)
            c->lds_max = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
            if (prop.maxSharedMemoryPerMultiProcessor > 0)
                c->lds_max = std::min(c->lds_max, (size_t)prop.maxSharedMemoryPerMultiProcessor);
            if (c->lds_max == 0) c->lds_max = 65536;
        }
    }
    if (rc == EWFA_OK && lds_budget(c) > 65536 && ewfa::align_set_max_lds(lds_budget(c)) != hipSuccess)
        rc = fail(EWFA_ERR_HIP, "could not raise the block kernels' LDS limit to %zu bytes", lds_budget(c));
    if (rc == EWFA_OK) {
        rc = ewfa::load_policy(c->opts.policy_path, c->policy);
        if (rc == EWFA_ERR_IO) fail(rc, "could not open the policy file %s", c->opts.policy_path);
        if (rc == EWFA_ERR_FORMAT) fail(rc, "bad policy file %s", c->opts.policy_path);
    }
    // The streams a context's work overlaps on are created together, right
    // after its compute stream: the runtime deals streams to its (4) hardware
    // queues by use, so streams created one after another land on different
    // queues, while one created later, after other contexts' streams, can
    // share a queue with this context's compute stream, where its work waits
    // behind the compute (a host-pipeline call ran 51 instead of 68 M aln/s
    // that way).  With more than one length bucket those are the two streams
    // the bucket passes fan out over; with one bucket, the upload stream of
    // the host pipeline (a multi-bucket context creates it on first use).
    if (rc == EWFA_OK && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(EWFA_ERR_HIP, "hipStreamCreate failed");
    if (rc == EWFA_OK) {
        if (c->policy.buckets.size() > 1) rc = ensure_aux(c, ewfa_ctx::MAX_FANOUT);
        else if (hipStreamCreateWithFlags(&c->up, hipStreamNonBlocking) != hipSuccess)
            rc = fail(EWFA_ERR_HIP, "hipStreamCreate failed");
    }
    if (rc == EWFA_OK) {
        if (hipEventCreate(&c->ev_in) != hipSuccess || hipEventCreate(&c->ev_a0) != hipSuccess ||
            hipEventCreate(&c->ev_a1) != hipSuccess || hipEventCreate(&c->ev_e1) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
            rc = fail(EWFA_ERR_HIP, "hipEventCreate failed");
    }
    if (rc == EWFA_OK) rc = c->ctrl.reserve(c, CTRL_WORDS * 4);
    if (rc == EWFA_OK && hipHostMalloc((void**)&c->host_ctrl, CTRL_WORDS * 4, hipHostMallocDefault) != hipSuccess)
        rc = fail(EWFA_ERR_NOMEM, "hipHostMalloc failed");
    if (rc != EWFA_OK) {
        ewfa_destroy(c);
        return rc;
    }
    memset(c->host_ctrl, 0, CTRL_WORDS * 4);
    *out = c;
    return EWFA_OK;
}

int ewfa_destroy(ewfa_ctx_t* c) {
    if (!c) return EWFA_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->wup) {
        (void)hipStreamSynchronize(c->wup);
        (void)hipStreamDestroy(c->wup);
    }
    for (hipEvent_t e : c->ev_words) (void)hipEventDestroy(e);
    release_retired(c);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->graph2) (void)hipGraphExecDestroy(c->graph2);
    for (hipEvent_t e : {c->ev_in, c->ev_a0, c->ev_a1, c->ev_e1, c->ev_fork})
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < ewfa_ctx::MAX_FANOUT - 1; ++i) {
        if (c->aux[i]) {
            (void)hipStreamSynchronize(c->aux[i]);
            (void)hipStreamDestroy(c->aux[i]);
        }
        if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    for (DevBuf* b : {&c->ops, &c->bt, &c->ctrl, &c->list_a, &c->list_b, &c->buckets, &c->runs, &c->list_fb, &c->bound,
                      &c->gslot, &c->order, &c->h_words, &c->h_pairs, &c->h_results, &c->h_cigars, &c->dense, &c->scan})
        b->release();
    if (c->host_ctrl) (void)hipHostFree(c->host_ctrl);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->up) {
        (void)hipStreamSynchronize(c->up);
        (void)hipStreamDestroy(c->up);
    }
    for (hipEvent_t e : c->ev_up) (void)hipEventDestroy(e);
    for (ewfa_ctx::Stage& st : c->stage) {
        if (st.dense) (void)hipHostFree(st.dense);
        if (st.res) (void)hipHostFree(st.res);
        if (st.done) (void)hipEventDestroy(st.done);
    }
    delete c;
    return EWFA_OK;
}

int ewfa_device_alloc(int device, size_t bytes, void** ptr) {
    if (!ptr) return fail(EWFA_ERR_ARG, "ptr is NULL");
    *ptr = nullptr;
    if (ewfa_device_count() <= 0) return fail(EWFA_ERR_NODEV, "no HIP device");
    HIPCHK(hipSetDevice(device));
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 4);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(EWFA_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return EWFA_OK;
}

int ewfa_device_free(void* ptr) {
    if (ptr) HIPCHK(hipFree(ptr));
    return EWFA_OK;
}

int ewfa_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    if (bytes == 0) return EWFA_OK;
    if (!dst || !src) return fail(EWFA_ERR_ARG, "null pointer");
    hipMemcpyKind k;
    switch (kind) {
        case 0: k = hipMemcpyHostToDevice; break;
        case 1: k = hipMemcpyDeviceToHost; break;
        case 2: k = hipMemcpyDeviceToDevice; break;
        default: return fail(EWFA_ERR_ARG, "unknown copy kind %d", kind);
    }
    HIPCHK(hipMemcpy(dst, src, bytes, k));
    return EWFA_OK;
}

int ewfa_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(EWFA_ERR_ARG, "ptr is NULL");
    *ptr = nullptr;
    if (hipHostMalloc(ptr, bytes ? bytes : 4, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        return fail(EWFA_ERR_NOMEM, "hipHostMalloc(%zu) failed", bytes);
    }
    return EWFA_OK;
}

int ewfa_host_free(void* ptr) {
    if (ptr) HIPCHK(hipHostFree(ptr));
    return EWFA_OK;
}

int ewfa_device_synchronize(int device) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    return EWFA_OK;
}

int ewfa_pack_ascii(const char* d_ascii, size_t ascii_stride, const uint32_t* d_len, size_t nseq,
                    uint64_t* d_words, size_t words_stride, void* stream) {
    if (nseq == 0) return EWFA_OK;
    if (!d_ascii || !d_len || !d_words || words_stride == 0)
        return fail(EWFA_ERR_ARG, "null buffer or zero stride");
    if (words_stride < ewfa_packed_words(ascii_stride) && words_stride * 32 < ascii_stride)
        return fail(EWFA_ERR_ARG, "words_stride too small for ascii_stride");
    HIPCHK(ewfa::launch_pack(d_ascii, ascii_stride, d_len, nseq, d_words, words_stride,
                             (hipStream_t)stream));
    return EWFA_OK;
}

int ewfa_context_memory(ewfa_ctx_t* c, size_t* bytes) {
    if (!c || !bytes) return fail(EWFA_ERR_ARG, "null argument");
    size_t b = 0;
    for (const DevBuf* d : {&c->ops, &c->bt, &c->ctrl, &c->list_a, &c->list_b, &c->buckets, &c->runs,
                            &c->list_fb, &c->bound, &c->gslot, &c->order, &c->h_words, &c->h_pairs, &c->h_results,
                            &c->h_cigars, &c->dense, &c->scan})
        b += d->bytes;
    *bytes = b + c->retired_bytes;
    return EWFA_OK;
}

int ewfa_max_distance(ewfa_ctx_t* c, uint32_t max_len) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    // passes beyond what LDS holds keep their slots in HBM, up to the budget
    // where one backtrace slot reaches GSLOT_BT_MAX bytes
    return (int)std::min<int64_t>((int64_t)max_len + 1, gslot_max_dcap());
}

int ewfa_set_flags(ewfa_ctx_t* c, int flags) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (flags & ~(EWFA_FLAG_TIMING | EWFA_FLAG_NO_GRAPH | EWFA_FLAG_SERIAL)) return fail(EWFA_ERR_ARG, "unknown flag");
    c->timing = (flags & EWFA_FLAG_TIMING) != 0;
    c->use_graph = (flags & EWFA_FLAG_NO_GRAPH) == 0;
    const int fanout = (flags & EWFA_FLAG_SERIAL) ? 1 : ewfa_ctx::MAX_FANOUT;
    if (fanout != c->fanout) {
        c->fanout = fanout;
        ++c->policy_gen;  // a captured plan of the other shape is captured again
    }
    return EWFA_OK;
}

int ewfa_last_timing(ewfa_ctx_t* c, float* align_ms, float* emit_ms) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (align_ms) *align_ms = c->align_ms;
    if (emit_ms) *emit_ms = c->emit_ms;
    return EWFA_OK;
}

int ewfa_align_device(ewfa_ctx_t* c, const uint64_t* d_words, const ewfa_pair_t* d_pairs, size_t n,
                      uint32_t max_len, ewfa_result_t* d_results, int8_t* d_cigars, void* stream) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (n == 0) return EWFA_OK;
    if (!d_words || !d_pairs || !d_results || !d_cigars) return fail(EWFA_ERR_ARG, "null buffer");
    if (n > 0xFFFFFFF0u) return fail(EWFA_ERR_RANGE, "too many pairs in one call");
    CHK(set_device(c));
    hipStream_t s = c->stream;
    // order the work after what the caller queued on its stream
    HIPCHK(hipEventRecord(c->ev_in, (hipStream_t)stream));
    HIPCHK(hipStreamWaitEvent(s, c->ev_in, 0));
    CHK(reserve_scratch(c, n, max_len));
    c->align_ms = c->emit_ms = 0.f;

    if (c->opts.mode == EWFA_MODE_PERSISTENT) {
        Plan pl;
        CHK(plan_persistent(c, max_len, (uint32_t)n, pl));
        CHK(reserve_plan(c, pl, n));
        CHK(ensure_aux(c, pl.fanout));  // not while a graph is being captured
        const bool timing = c->timing;
        ewfa_ctx::Key key{d_words, d_pairs, d_results, d_cigars, ops_base(c), n, max_len, timing};
        if (c->use_graph) {
            // with timing: two graphs (align, emit) with events between them
            const int nparts = timing ? 2 : 1;
            const uint64_t gen = c->buffer_gen;
            if (!(c->graph && c->graph_key == key && c->graph_policy_gen == c->policy_gen &&
                  c->graph_buffer_gen == gen)) {
                for (hipGraphExec_t* ge : {&c->graph, &c->graph2})
                    if (*ge) {
                        (void)hipGraphExecDestroy(*ge);
                        *ge = nullptr;
                    }
                for (int part = 0; part < nparts; ++part) {
                    const int phase = timing ? (part == 0 ? PLAN_ALIGN : PLAN_EMIT) : PLAN_ALL;
                    hipGraph_t g = nullptr;
                    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    int rc = enqueue_plan(c, pl, d_words, d_pairs, (uint32_t)n, max_len, d_results, d_cigars, s,
                                          false, phase);
                    hipError_t ce = hipStreamEndCapture(s, &g);
                    if (rc != EWFA_OK) {
                        if (g) (void)hipGraphDestroy(g);
                        return rc;
                    }
                    if (ce != hipSuccess) return fail(EWFA_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
                    hipGraphExec_t* ge = part == 0 ? &c->graph : &c->graph2;
                    hipError_t ie = hipGraphInstantiate(ge, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    if (ie != hipSuccess) {
                        *ge = nullptr;
                        return fail(EWFA_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
                    }
                }
                c->graph_key = key;
                c->graph_policy_gen = c->policy_gen;
                c->graph_buffer_gen = gen;
            }
            if (timing) {
                HIPCHK(hipEventRecord(c->ev_a0, s));
                HIPCHK(hipGraphLaunch(c->graph, s));
                HIPCHK(hipEventRecord(c->ev_a1, s));
                HIPCHK(hipGraphLaunch(c->graph2, s));
                HIPCHK(hipEventRecord(c->ev_e1, s));
            } else {
                HIPCHK(hipGraphLaunch(c->graph, s));
            }
            HIPCHK(hipStreamSynchronize(s));
            if (timing) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, c->ev_a0, c->ev_a1) == hipSuccess) c->align_ms = ms;
                if (hipEventElapsedTime(&ms, c->ev_a1, c->ev_e1) == hipSuccess) c->emit_ms = ms;
            }
        } else {
            CHK(enqueue_plan(c, pl, d_words, d_pairs, (uint32_t)n, max_len, d_results, d_cigars, s, timing));
            HIPCHK(hipStreamSynchronize(s));
            if (timing) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, c->ev_a0, c->ev_a1) == hipSuccess) c->align_ms = ms;
                if (hipEventElapsedTime(&ms, c->ev_a1, c->ev_e1) == hipSuccess) c->emit_ms = ms;
            }
        }
        return EWFA_OK;
    }

    // plain: the reference's per-batch launch model (src/aligner.cu:62-66)
    const size_t batch = c->opts.batch_size > 0 ? (size_t)c->opts.batch_size : n;
    for (size_t first = 0; first < n; first += batch) {
        const uint32_t count = (uint32_t)std::min(batch, n - first);
        Pass p;
        CHK(main_pass(c, max_len, count, false, p));
        CHK(enqueue_main(c, p, d_words, d_pairs, first, count, max_len, d_results, d_cigars, s,
                         c->timing));
        HIPCHK(hipStreamSynchronize(s));
        if (c->timing) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, c->ev_a0, c->ev_a1) == hipSuccess) c->align_ms += ms;
            if (hipEventElapsedTime(&ms, c->ev_a1, c->ev_e1) == hipSuccess) c->emit_ms += ms;
        }
        CHK(resolve_deferred(c, p.dcap, d_words, d_pairs, first, max_len, d_results, d_cigars, s));
    }
    return EWFA_OK;
}

int ewfa_upload_words(ewfa_ctx_t* c, const uint64_t* h_words, size_t n_words) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (!h_words && n_words) return fail(EWFA_ERR_ARG, "null buffer");
    CHK(set_device(c));
    CHK(words_settle(c));
    CHK(c->h_words.reserve(c, std::max<size_t>(n_words, 1) * 8));
    if (n_words) {
        HIPCHK(hipMemcpyAsync(c->h_words.p, h_words, n_words * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->resident_words = n_words;
    return EWFA_OK;
}

int ewfa_upload_words_async(ewfa_ctx_t* c, const uint64_t* h_words, size_t n_words) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (!h_words && n_words) return fail(EWFA_ERR_ARG, "null buffer");
    CHK(set_device(c));
    CHK(words_settle(c));
    // work already queued on the context may still read the old words
    HIPCHK(hipStreamSynchronize(c->stream));
    CHK(c->h_words.reserve(c, std::max<size_t>(n_words, 1) * 8));
    c->resident_words = n_words;
    if (n_words == 0) return EWFA_OK;
    if (!c->wup) HIPCHK(hipStreamCreateWithFlags(&c->wup, hipStreamNonBlocking));
    const size_t pieces = (n_words + WORDS_PIECE - 1) / WORDS_PIECE;
    while (c->ev_words.size() < pieces) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_words.push_back(e);
    }
    uint64_t* d = c->h_words.as<uint64_t>();
    for (size_t k = 0; k < pieces; ++k) {
        const size_t lo = k * WORDS_PIECE, hi = std::min(n_words, lo + WORDS_PIECE);
        HIPCHK(hipMemcpyAsync(d + lo, h_words + lo, (hi - lo) * 8, hipMemcpyHostToDevice, c->wup));
        HIPCHK(hipEventRecord(c->ev_words[k], c->wup));
        c->words_pieces = k + 1;
    }
    return EWFA_OK;
}

int ewfa_align_resident(ewfa_ctx_t* c, const ewfa_pair_t* h_pairs, size_t n, ewfa_result_t* h_results,
                        int8_t* h_cigars, size_t cigar_bytes) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (n == 0) return EWFA_OK;
    if (!h_pairs || !h_results || !h_cigars) return fail(EWFA_ERR_ARG, "null buffer");
    if (!c->h_words.p) return fail(EWFA_ERR_STATE, "no words uploaded");
    const size_t n_words = c->resident_words;
    if (c->opts.mode == EWFA_MODE_PERSISTENT) {
        CHK(set_device(c));
        return run_pipeline(c, nullptr, n_words, h_pairs, n, h_results, h_cigars, cigar_bytes);
    }
    uint32_t max_len = 0;
    size_t caps = 0;
    for (size_t i = 0; i < n; ++i) {
        const ewfa_pair_t& p = h_pairs[i];
        max_len = std::max(max_len, std::max(p.plen, p.tlen));
        caps += p.cigar_cap;
        if (p.pattern_off + ewfa_packed_words(p.plen) - 1 > n_words ||
            p.text_off + ewfa_packed_words(p.tlen) - 1 > n_words)
            return fail(EWFA_ERR_ARG, "pair %zu reads past the packed words", i);
        if ((p.cigar_off & 3) || (p.cigar_cap & 3) || p.cigar_off + p.cigar_cap > cigar_bytes)
            return fail(EWFA_ERR_ARG, "pair %zu has a bad CIGAR region", i);
    }
    CHK(set_device(c));
    CHK(words_settle(c));  // the plain mode reads every word
    CHK(c->h_pairs.reserve(c, n * sizeof(ewfa_pair_t)));
    CHK(c->h_results.reserve(c, n * sizeof(ewfa_result_t)));
    CHK(c->h_cigars.reserve(c, std::max<size_t>(cigar_bytes, 4)));
    hipStream_t s = c->stream;
    HIPCHK(hipMemcpyAsync(c->h_pairs.p, h_pairs, n * sizeof(ewfa_pair_t), hipMemcpyHostToDevice, s));
    CHK(ewfa_align_device(c, c->h_words.as<uint64_t>(), c->h_pairs.as<ewfa_pair_t>(), n, max_len,
                          c->h_results.as<ewfa_result_t>(), c->h_cigars.as<int8_t>(), s));
    // only the written CIGAR bytes come back: packed densely on the device,
    // placed at each pair's offset on the host
    CHK(ewfa::enqueue_compact(c, c->h_pairs.as<ewfa_pair_t>(), c->h_results.as<ewfa_result_t>(),
                              c->h_cigars.as<int8_t>(), n, caps, s));
    HIPCHK(hipMemcpyAsync(h_results, c->h_results.p, n * sizeof(ewfa_result_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int8_t* dense = nullptr;
    const size_t total = ewfa::compact_total(h_pairs, h_results, n);
    CHK(ewfa::fetch_compact(c, total, s, &dense));
    if (ewfa::scatter_streams(dense, total, h_pairs, h_results, n, h_cigars, 16) == SIZE_MAX)
        return fail(EWFA_ERR_STATE, "compacted streams exceed their regions");
    return EWFA_OK;
}

int ewfa_align_host(ewfa_ctx_t* c, const uint64_t* h_words, size_t n_words, const ewfa_pair_t* h_pairs,
                    size_t n, ewfa_result_t* h_results, int8_t* h_cigars, size_t cigar_bytes) {
    if (!c) return fail(EWFA_ERR_ARG, "ctx is NULL");
    if (n == 0) return EWFA_OK;
    if (!h_words || !h_pairs || !h_results || !h_cigars) return fail(EWFA_ERR_ARG, "null buffer");
    if (c->opts.mode != EWFA_MODE_PERSISTENT) {
        CHK(ewfa_upload_words(c, h_words, n_words));
        return ewfa_align_resident(c, h_pairs, n, h_results, h_cigars, cigar_bytes);
    }
    CHK(set_device(c));
    return run_pipeline(c, h_words, n_words, h_pairs, n, h_results, h_cigars, cigar_bytes);
}

}  // extern "C"
