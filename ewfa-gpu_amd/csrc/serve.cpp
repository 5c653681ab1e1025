// serve.cpp -- the job server: continuous batching of alignment jobs.  Each
// slot (a worker thread with its own context and stream) takes every queued
// job that fits the device arena, merges them into one persistent launch and
// returns each job's results; jobs that arrive meanwhile form the next batch
// on the next free slot, so batches overlap on the device.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ewfa.h"
#include "policy.h"

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation.
class Arena {
public:
    static constexpr size_t kAlign = 256;

    int init(int device, size_t bytes) {
        if (hipSetDevice(device) != hipSuccess) return EWFA_ERR_HIP;
        if (hipMalloc(&base_, bytes) != hipSuccess) {
            base_ = nullptr;
            return EWFA_ERR_NOMEM;
        }
        cap_ = bytes;
        free_[0] = bytes;
        return EWFA_OK;
    }
    void release() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }
    static size_t round(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
    // offset of a free block of n bytes, or SIZE_MAX
    size_t alloc(size_t n) {
        n = round(std::max<size_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return SIZE_MAX;
    }
    void free(size_t off, size_t n) {
        n = round(std::max<size_t>(n, 1));
        used_ -= n;
        auto next = free_.lower_bound(off);
        size_t start = off, len = n;
        if (next != free_.end() && next->first == off + n) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }
    char* at(size_t off) const { return static_cast<char*>(base_) + off; }
    size_t capacity() const { return cap_; }
    size_t peak() const { return peak_; }

private:
    void* base_ = nullptr;
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

struct Job {
    uint64_t id;
    const uint64_t* words;
    size_t n_words;
    const ewfa_pair_t* pairs;
    size_t n;
    ewfa_result_t* results;
    int8_t* cigars;
    size_t cigar_bytes;
    uint32_t max_len;
    size_t data_bytes;  // resident block: packed words, then CIGAR streams
    size_t data_off = SIZE_MAX;
    Clock::time_point submitted, finished;
    int rc = EWFA_OK;
    bool done = false;
};

size_t words_block(const Job* j) { return Arena::round(j->n_words * 8); }

// Per-launch state of a batch of n pairs up to max_len bases: merged
// descriptors, results and the op scratch of the wavefront traceback.
size_t state_bytes(size_t n, uint32_t max_len) {
    return Arena::round(n * sizeof(ewfa_pair_t)) + Arena::round(n * sizeof(ewfa_result_t)) +
           Arena::round(ewfa::op_scratch_bytes(n, max_len));
}

// Jobs merged into one launch.
struct Batch {
    std::vector<Job*> jobs;
    size_t n = 0;
    uint32_t max_len = 0;
    size_t state_off = SIZE_MAX, state = 0;
};

}  // namespace

struct ewfa_server {
    ewfa_server_opts_t opts{};
    std::string policy;
    Arena arena;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::unordered_map<uint64_t, Job*> jobs;
    std::vector<std::thread> workers;
    uint64_t next_id = 1;
    bool stopping = false;
    int running = 0, peak_running = 0, peak_batch_jobs = 0;
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0;
};

namespace {

// Takes jobs from the head of the queue into b while their resident blocks
// and the batch's growing launch state fit the arena together (first come,
// first served: a head job that does not fit waits, and so does everything
// behind it).  Called with the server lock held; false if nothing fit.
bool form_batch(ewfa_server* s, Batch& b) {
    while (!s->queue.empty()) {
        Job* j = s->queue.front();
        if (!b.jobs.empty() && b.n + j->n > s->opts.max_batch_pairs) break;
        const uint32_t ml = std::max(b.max_len, j->max_len);
        const size_t st = state_bytes(b.n + j->n, ml);
        const size_t off = s->arena.alloc(j->data_bytes);
        if (off == SIZE_MAX) break;
        const size_t probe = s->arena.alloc(st);
        if (probe == SIZE_MAX) {
            s->arena.free(off, j->data_bytes);
            break;
        }
        s->arena.free(probe, st);
        j->data_off = off;
        b.jobs.push_back(j);
        b.n += j->n;
        b.max_len = ml;
        s->queue.pop_front();
    }
    if (b.jobs.empty()) return false;
    b.state = state_bytes(b.n, b.max_len);
    b.state_off = s->arena.alloc(b.state);  // the last probe of this size succeeded
    return b.state_off != SIZE_MAX;
}

// Pinned staging for a batch's merged descriptors.
struct Staging {
    ewfa_pair_t* p = nullptr;
    size_t cap = 0;
    ~Staging() {
        if (p) (void)hipHostFree(p);
    }
    bool reserve(size_t n) {
        if (n <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(ewfa_pair_t), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = n;
        return true;
    }
};

// Runs a batch as one launch: uploads each job's packed words into its
// block, rebases its descriptors onto the arena (sequence offsets in words,
// CIGAR offsets in bytes, both from the arena base), aligns all pairs
// together and copies each job's results and CIGAR streams back.
int run_batch(ewfa_server* s, ewfa_ctx_t* ctx, hipStream_t st, Staging& stage, const Batch& b) {
    char* base = s->arena.at(0);
    if (!stage.reserve(b.n)) return EWFA_ERR_NOMEM;
    size_t k = 0;
    for (const Job* j : b.jobs) {
        if (j->n_words &&
            hipMemcpyAsync(base + j->data_off, j->words, j->n_words * 8, hipMemcpyHostToDevice, st) != hipSuccess)
            return EWFA_ERR_HIP;
        const uint64_t woff = j->data_off / 8;
        const uint64_t coff = j->data_off + words_block(j);
        for (size_t i = 0; i < j->n; ++i) {
            ewfa_pair_t p = j->pairs[i];
            p.pattern_off += woff;
            p.text_off += woff;
            p.cigar_off += coff;
            stage.p[k++] = p;
        }
    }
    if (b.n == 0) return EWFA_OK;
    char* d_pairs = base + b.state_off;
    char* d_results = d_pairs + Arena::round(b.n * sizeof(ewfa_pair_t));
    char* d_ops = d_results + Arena::round(b.n * sizeof(ewfa_result_t));
    if (hipMemcpyAsync(d_pairs, stage.p, b.n * sizeof(ewfa_pair_t), hipMemcpyHostToDevice, st) != hipSuccess)
        return EWFA_ERR_HIP;
    const int rc = ewfa::align_device_with_scratch(
        ctx, reinterpret_cast<const uint64_t*>(base), reinterpret_cast<const ewfa_pair_t*>(d_pairs), b.n, b.max_len,
        reinterpret_cast<ewfa_result_t*>(d_results), reinterpret_cast<int8_t*>(base), st,
        reinterpret_cast<uint64_t*>(d_ops), ewfa::op_scratch_bytes(b.n, b.max_len));
    if (rc != EWFA_OK) return rc;
    k = 0;
    for (const Job* j : b.jobs) {
        if (j->n &&
            (hipMemcpyAsync(j->results, d_results + k * sizeof(ewfa_result_t), j->n * sizeof(ewfa_result_t),
                            hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipMemcpyAsync(j->cigars, base + j->data_off + words_block(j), j->cigar_bytes, hipMemcpyDeviceToHost,
                            st) != hipSuccess))
            return EWFA_ERR_HIP;
        k += j->n;
    }
    return hipStreamSynchronize(st) == hipSuccess ? EWFA_OK : EWFA_ERR_HIP;
}

void worker_main(ewfa_server* s) {
    ewfa_opts_t o;
    ewfa_default_opts(&o);
    o.device = s->opts.device;
    o.mode = EWFA_MODE_PERSISTENT;
    o.max_distance = s->opts.max_distance;
    o.policy_path = s->policy.empty() ? nullptr : s->policy.c_str();
    ewfa_ctx_t* ctx = nullptr;
    const int crc = ewfa_create(&ctx, &o);
    if (crc == EWFA_OK) ewfa_set_flags(ctx, EWFA_FLAG_NO_GRAPH);  // every batch has its own shape
    hipStream_t st = nullptr;
    (void)hipSetDevice(s->opts.device);
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
    Staging stage;
    for (;;) {
        Batch b;
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cv_work.wait(lk, [&] {
                if (s->queue.empty()) return s->stopping;
                return form_batch(s, b);
            });
            if (b.jobs.empty()) break;  // stopping, queue drained
            s->running += (int)b.jobs.size();
            s->peak_running = std::max(s->peak_running, s->running);
            s->peak_batch_jobs = std::max(s->peak_batch_jobs, (int)b.jobs.size());
            ++s->batches;
        }
        s->cv_work.notify_all();  // the next head may fit beside this batch
        const int rc = (crc != EWFA_OK || !st) ? EWFA_ERR_HIP : run_batch(s, ctx, st, stage, b);
        {
            std::lock_guard<std::mutex> lk(s->m);
            const auto now = Clock::now();
            for (Job* j : b.jobs) {
                s->arena.free(j->data_off, j->data_bytes);
                j->rc = rc;
                j->finished = now;
                j->done = true;
                ++s->completed;
                if (rc != EWFA_OK) ++s->failed;
            }
            s->arena.free(b.state_off, b.state);
            s->running -= (int)b.jobs.size();
        }
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
    if (st) (void)hipStreamDestroy(st);
    if (ctx) ewfa_destroy(ctx);
}

}  // namespace

extern "C" {

int ewfa_server_create(ewfa_server_t** out, const ewfa_server_opts_t* opts) {
    if (!out) return EWFA_ERR_ARG;
    *out = nullptr;
    if (ewfa_device_count() <= 0) return EWFA_ERR_NODEV;
    ewfa_server* s = new ewfa_server();
    if (opts) s->opts = *opts;
    if (s->opts.slots <= 0) s->opts.slots = 4;
    if (s->opts.slots > 64) s->opts.slots = 64;
    if (s->opts.arena_bytes == 0) s->opts.arena_bytes = (size_t)4 << 30;
    if (s->opts.max_batch_pairs == 0) s->opts.max_batch_pairs = (size_t)1 << 18;
    if (s->opts.policy_path) s->policy = s->opts.policy_path;
    s->opts.policy_path = nullptr;
    const int rc = s->arena.init(s->opts.device, s->opts.arena_bytes);
    if (rc != EWFA_OK) {
        delete s;
        return rc;
    }
    for (int i = 0; i < s->opts.slots; ++i) s->workers.emplace_back(worker_main, s);
    *out = s;
    return EWFA_OK;
}

int ewfa_server_destroy(ewfa_server_t* s) {
    if (!s) return EWFA_OK;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->jobs) delete kv.second;
    s->arena.release();
    delete s;
    return EWFA_OK;
}

int ewfa_server_submit(ewfa_server_t* s, const uint64_t* h_words, size_t n_words, const ewfa_pair_t* h_pairs,
                       size_t n, ewfa_result_t* h_results, int8_t* h_cigars, size_t cigar_bytes,
                       uint64_t* job_id) {
    if (!s || !job_id || (n && (!h_words || !h_pairs || !h_results || !h_cigars))) return EWFA_ERR_ARG;
    uint32_t max_len = 0;
    for (size_t i = 0; i < n; ++i) {
        const ewfa_pair_t& p = h_pairs[i];
        max_len = std::max(max_len, std::max(p.plen, p.tlen));
        if (p.pattern_off + ewfa_packed_words(p.plen) - 1 > n_words ||
            p.text_off + ewfa_packed_words(p.tlen) - 1 > n_words)
            return EWFA_ERR_ARG;
        if ((p.cigar_off & 3) || (p.cigar_cap & 3) || p.cigar_off + p.cigar_cap > cigar_bytes) return EWFA_ERR_ARG;
    }
    Job* j = new Job();
    j->words = h_words;
    j->n_words = n_words;
    j->pairs = h_pairs;
    j->n = n;
    j->results = h_results;
    j->cigars = h_cigars;
    j->cigar_bytes = cigar_bytes;
    j->max_len = max_len;
    j->data_bytes = Arena::round(n_words * 8) + Arena::round(cigar_bytes);
    if (j->data_bytes + state_bytes(n, max_len) > s->arena.capacity()) {
        delete j;
        return EWFA_ERR_RANGE;
    }
    j->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        j->id = s->next_id++;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        ++s->submitted;
        *job_id = j->id;
    }
    s->cv_work.notify_all();
    return EWFA_OK;
}

int ewfa_server_poll(ewfa_server_t* s, uint64_t id, int* done) {
    if (!s || !done) return EWFA_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return EWFA_ERR_ARG;
    *done = it->second->done ? 1 : 0;
    return EWFA_OK;
}

int ewfa_server_wait(ewfa_server_t* s, uint64_t id) {
    if (!s) return EWFA_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return EWFA_ERR_ARG;
    Job* j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    return j->rc;
}

int ewfa_server_job_latency(ewfa_server_t* s, uint64_t id, double* seconds) {
    if (!s || !seconds) return EWFA_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || !it->second->done) return EWFA_ERR_STATE;
    *seconds = std::chrono::duration<double>(it->second->finished - it->second->submitted).count();
    return EWFA_OK;
}

int ewfa_server_stats(ewfa_server_t* s, ewfa_server_stats_t* out) {
    if (!s || !out) return EWFA_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->arena_bytes = s->arena.capacity();
    out->arena_peak = s->arena.peak();
    out->slots = s->opts.slots;
    out->peak_running = s->peak_running;
    out->batches = s->batches;
    out->peak_batch_jobs = s->peak_batch_jobs;
    return EWFA_OK;
}

}  // extern "C"
