// serve.cpp -- the job server: continuous admission of alignment jobs into
// device slots by arena footprint, concurrent execution on per-slot streams,
// per-job latency.
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
    size_t footprint;
    Clock::time_point submitted, finished;
    int rc = EWFA_OK;
    bool done = false;
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
    int running = 0, peak_running = 0;
    uint64_t submitted = 0, completed = 0, failed = 0;
};

namespace {

size_t footprint_of(size_t n_words, size_t n, size_t cigar_bytes) {
    return Arena::round(n_words * 8) + Arena::round(n * sizeof(ewfa_pair_t)) +
           Arena::round(n * sizeof(ewfa_result_t)) + Arena::round(cigar_bytes);
}

// Runs one job on the slot's context: upload into its arena block, align,
// download.  Returns an EWFA status.
int run_job(ewfa_server* s, ewfa_ctx_t* ctx, hipStream_t st, Job* j, size_t base) {
    char* p = s->arena.at(base);
    char* d_words = p;
    char* d_pairs = d_words + Arena::round(j->n_words * 8);
    char* d_results = d_pairs + Arena::round(j->n * sizeof(ewfa_pair_t));
    char* d_cigars = d_results + Arena::round(j->n * sizeof(ewfa_result_t));
    if (hipMemcpyAsync(d_words, j->words, j->n_words * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_pairs, j->pairs, j->n * sizeof(ewfa_pair_t), hipMemcpyHostToDevice, st) != hipSuccess)
        return EWFA_ERR_HIP;
    int rc = ewfa_align_device(ctx, reinterpret_cast<const uint64_t*>(d_words),
                               reinterpret_cast<const ewfa_pair_t*>(d_pairs), j->n, j->max_len,
                               reinterpret_cast<ewfa_result_t*>(d_results), reinterpret_cast<int8_t*>(d_cigars), st);
    if (rc != EWFA_OK) return rc;
    if (hipMemcpyAsync(j->results, d_results, j->n * sizeof(ewfa_result_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(j->cigars, d_cigars, j->cigar_bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return EWFA_ERR_HIP;
    return EWFA_OK;
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
    if (crc == EWFA_OK) ewfa_set_flags(ctx, EWFA_FLAG_NO_GRAPH);  // each job has its own buffers
    hipStream_t st = nullptr;
    (void)hipSetDevice(s->opts.device);
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
    for (;;) {
        Job* j = nullptr;
        size_t base = SIZE_MAX;
        {
            std::unique_lock<std::mutex> lk(s->m);
            // first come, first served: the head job waits for arena room
            s->cv_work.wait(lk, [&] {
                if (s->queue.empty()) return s->stopping;
                base = s->arena.alloc(s->queue.front()->footprint);
                return base != SIZE_MAX;
            });
            if (s->queue.empty()) break;
            j = s->queue.front();
            s->queue.pop_front();
            ++s->running;
            s->peak_running = std::max(s->peak_running, s->running);
        }
        s->cv_work.notify_all();  // the next head may fit beside this one
        int rc = (crc != EWFA_OK || !st) ? EWFA_ERR_HIP : run_job(s, ctx, st, j, base);
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->arena.free(base, j->footprint);
            j->rc = rc;
            j->finished = Clock::now();
            j->done = true;
            --s->running;
            ++s->completed;
            if (rc != EWFA_OK) ++s->failed;
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
    j->footprint = footprint_of(n_words, n, cigar_bytes);
    if (j->footprint > s->arena.capacity()) {
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
    return EWFA_OK;
}

}  // extern "C"
