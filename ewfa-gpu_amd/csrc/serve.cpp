// serve.cpp -- the job server: continuous batching of alignment jobs.
//
// A staging thread admits submitted jobs in arrival order by their footprint
// in the device arena: each admitted job gets a block for its packed words and
// CIGAR regions, and its words go up at once on the server's upload stream
// (an event marks the end of each job's upload), so uploads run at the arrival
// rate, beside the device's compute.  Each slot owns a device context (its
// compute streams), a worker thread, a completion thread and two "flights"
// (batches in progress).  A worker with a free flight takes the staged jobs
// first-come first-served, merges their descriptors into one batch, and, on
// its compute stream, waits for the batch's last upload (jobs go up in the
// order they are batched), uploads the descriptors and enqueues the whole
// persistent path for the batch (bucketing, align passes,
// the device-driven deferral chain, emits), then the CIGAR compaction, which
// writes the written stream bytes straight into the flight's pinned host
// buffer, and the results copy, all without waiting.  The completion thread
// waits for a flight and hands each job its results and streams; a job
// completes as soon as its own outputs are in place.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ewfa.h"
#include "ewfa_internal.h"
#include "runtime.h"

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
    size_t caps;        // sum of the pairs' cigar_cap
    size_t data_bytes;  // resident block: packed words, then CIGAR streams
    size_t data_off = SIZE_MAX;  // its block in the arena, from admission
    hipEvent_t up_ev = nullptr;  // recorded after its upload, until its batch waits on it
    uint64_t stage_seq = 0;      // its place in the upload stream's order
    Clock::time_point submitted, finished;
    int rc = EWFA_OK;
    bool done = false;
};

// Host-to-device copies below this size run as a copy kernel on the CUs, which
// waits for CU slots behind the persistent passes (up to a pass's length);
// from this size on they run on a DMA engine.  A job's words smaller than
// this go up through a bounce slot of exactly this size.
constexpr size_t kDmaMin = (size_t)64 << 10;
constexpr int kBounceSlots = 64;

size_t words_bytes(size_t n_words) {
    const size_t b = n_words * 8;
    return Arena::round(b > 0 && b < kDmaMin ? kDmaMin : b);
}
size_t words_block(const Job* j) { return words_bytes(j->n_words); }

// Launch state of a batch of n pairs up to max_len bases: merged
// descriptors, results, the traceback op scratch and the compaction's block
// sums.  The descriptors and results regions hold at least kDmaMin bytes, the
// size their copies are padded to.
struct StateLayout {
    size_t pairs, results, ops, scan, total;
    StateLayout(size_t n, uint32_t max_len) {
        pairs = 0;
        results = pairs + Arena::round(std::max(n * sizeof(ewfa_pair_t), kDmaMin));
        ops = results + Arena::round(std::max(n * sizeof(ewfa_result_t), kDmaMin));
        scan = ops + Arena::round(ewfa::op_scratch_bytes(n, max_len));
        total = scan + Arena::round(16 * ((n + ewfa::COMPACT_BLOCK - 1) / ewfa::COMPACT_BLOCK));
    }
};

// Jobs merged into one launch.
struct Batch {
    std::vector<Job*> jobs;
    size_t n = 0, caps = 0;
    uint32_t max_len = 0;
    size_t state_off = SIZE_MAX, state = 0;
};

// Pinned host buffer that only grows (never while a copy may use it), by at
// least half again.  A replaced buffer is kept until the server is destroyed:
// freeing pinned memory waits for the whole device.
// COHERENT: for memory the device writes with its own stores (they reach host
// memory without a cache flush); copies to and from the other kind run on the
// DMA engines instead of a copy kernel on the CUs.
template <typename T, bool COHERENT = false>
struct Pinned {
    T* p = nullptr;
    size_t cap = 0;
    std::vector<T*> retired;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
        for (T* q : retired) (void)hipHostFree(q);
    }
    bool reserve(size_t n) {
        if (n <= cap) return true;
        const size_t want = std::max<size_t>(std::max(n, cap + cap / 2), 1024);
        T* q = nullptr;
        if (hipHostMalloc((void**)&q, want * sizeof(T), COHERENT ? hipHostMallocCoherent : hipHostMallocNonCoherent) !=
            hipSuccess)
            return false;
        if (p) retired.push_back(p);
        p = q;
        cap = want;
        return true;
    }
};

// A batch in progress on a slot.
struct Flight {
    Batch b;
    Pinned<ewfa_pair_t> pairs;      // merged, rebased descriptors
    Pinned<ewfa_result_t> results;  // the batch's results
    Pinned<int8_t, true> dense;     // the batch's compacted CIGAR bytes, written by the device
    int8_t* dense_dev = nullptr;    // the same memory as the device addresses it
    hipEvent_t done = nullptr;
    int rc = EWFA_OK;
};

struct Slot {
    ewfa_server* s = nullptr;  // declared in ewfa.h
    ewfa_ctx_t* ctx = nullptr;
    hipStream_t cs = nullptr;
    Flight flights[2];
    std::mutex m;
    std::condition_variable cv;
    std::deque<Flight*> idle, busy;
    bool worker_done = false;
    int init_rc = EWFA_OK;
    std::thread worker, completer;
};

}  // namespace

struct ewfa_server {
    ewfa_server_opts_t opts{};
    std::string policy;
    Arena arena;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    // submitted jobs waiting for admission, and admitted (staged) jobs whose
    // words are uploaded or on their way, waiting for a batch
    std::deque<Job*> incoming, ready;
    size_t staged_bytes = 0, stage_cap = 0;
    int staging = 0;  // jobs taken from `incoming` whose upload is being issued
    hipStream_t up = nullptr;  // the jobs' words
    char* bounce = nullptr;    // kBounceSlots pinned slots of kDmaMin bytes, for small jobs' words
    hipEvent_t bounce_ev[kBounceSlots] = {};
    bool bounce_used[kBounceSlots] = {};
    int bounce_next = 0;  // staging thread only
    std::vector<hipEvent_t> ev_pool;  // upload events not held by a job
    uint64_t stage_seq = 0;
    std::thread stager;
    std::unordered_map<uint64_t, Job*> jobs;
    std::vector<Slot*> slots;
    uint64_t next_id = 1;
    bool stopping = false;
    int running = 0, peak_running = 0, peak_batch_jobs = 0;
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0;
    double upload_s = 0, align_s = 0, handout_s = 0;
};

namespace {

// Returns a staged job's upload event to the pool.  Called with the server
// lock held.
void release_upload_event(ewfa_server* s, Job* j) {
    if (j->up_ev) s->ev_pool.push_back(j->up_ev);
    j->up_ev = nullptr;
}

// Takes staged jobs from the head of `ready` into b (first come, first
// served, up to the batch limits, a single larger job excepted) and places
// the batch's launch state in the arena; while the state does not fit, the
// last jobs go back to the head of `ready`.  Called with the server lock
// held; false when nothing could be taken.
bool form_batch(ewfa_server* s, Batch& b) {
    while (!s->ready.empty()) {
        Job* j = s->ready.front();
        if (!b.jobs.empty() &&
            (b.n + j->n > s->opts.max_batch_pairs || b.caps + j->caps > s->opts.max_batch_cigar_bytes))
            break;
        b.jobs.push_back(j);
        b.n += j->n;
        b.caps += j->caps;
        b.max_len = std::max(b.max_len, j->max_len);
        s->ready.pop_front();
    }
    while (!b.jobs.empty()) {
        b.state = StateLayout(b.n, b.max_len).total;
        b.state_off = s->arena.alloc(b.state);
        if (b.state_off != SIZE_MAX) break;
        Job* j = b.jobs.back();
        b.jobs.pop_back();
        s->ready.push_front(j);
        b.n -= j->n;
        b.caps -= j->caps;
        b.max_len = 0;
        for (const Job* k : b.jobs) b.max_len = std::max(b.max_len, k->max_len);
    }
    if (b.jobs.empty()) return false;
    for (const Job* j : b.jobs) s->staged_bytes -= j->data_bytes;
    return true;
}

// No batch is on the device and the head job's launch state does not fit
// beside the staged jobs: the staged jobs behind it give their blocks back
// (after their uploads land) and return to the head of `incoming`, in order.
// Called with the server lock held.
void unstage_behind_head(ewfa_server* s) {
    (void)hipStreamSynchronize(s->up);
    while (s->ready.size() > 1) {
        Job* j = s->ready.back();
        s->ready.pop_back();
        s->arena.free(j->data_off, j->data_bytes);
        s->staged_bytes -= j->data_bytes;
        j->data_off = SIZE_MAX;
        release_upload_event(s, j);
        s->incoming.push_front(j);
    }
}

double since(Clock::time_point& t) {
    const auto now = Clock::now();
    const double d = std::chrono::duration<double>(now - t).count();
    t = now;
    return d;
}

#define QCHK(x)                                 \
    do {                                        \
        if ((x) != hipSuccess) return EWFA_ERR_HIP; \
    } while (0)

// Runs fn(k) for k in [0, count) on up to `threads` threads (the calling one
// included), in contiguous ranges.
template <typename F>
void parallel_ranges(size_t count, int threads, F&& fn) {
    const size_t t = std::max<size_t>(1, std::min<size_t>((size_t)threads, count));
    if (t == 1) {
        for (size_t k = 0; k < count; ++k) fn(k);
        return;
    }
    std::vector<std::thread> pool;
    const size_t per = (count + t - 1) / t;
    for (size_t i = 1; i < t; ++i)
        pool.emplace_back([&fn, i, per, count] {
            for (size_t k = i * per; k < std::min(count, (i + 1) * per); ++k) fn(k);
        });
    for (size_t k = 0; k < std::min(count, per); ++k) fn(k);
    for (std::thread& th : pool) th.join();
}

// Threads that rebase or hand out one batch: one per 4,096 pairs, at most 8.
int handout_threads(size_t pairs) { return (int)std::min<size_t>(8, std::max<size_t>(1, pairs / 4096)); }

// Enqueues a batch on the slot's compute stream: a wait for the batch's last
// upload (the staging thread issues the jobs' words on one stream in order),
// the rebased descriptors (sequence offsets in words, CIGAR offsets in bytes,
// both from the arena base), the persistent path, the CIGAR compaction and
// the results copy.  Returns without waiting.
int enqueue_batch_on(Slot* sl, Flight* f) {
    ewfa_server* s = sl->s;
    const Batch& b = f->b;
    char* base = s->arena.at(0);
    // both copies padded to the DMA size (the regions hold it)
    const size_t pairs_copy = std::max(b.n * sizeof(ewfa_pair_t), kDmaMin);
    const size_t results_copy = std::max(b.n * sizeof(ewfa_result_t), kDmaMin);
    if (!f->pairs.reserve(pairs_copy / sizeof(ewfa_pair_t) + 1) ||
        !f->results.reserve(results_copy / sizeof(ewfa_result_t) + 1))
        return EWFA_ERR_NOMEM;
    if (b.caps > f->dense.cap || !f->dense_dev) {
        if (!f->dense.reserve(b.caps)) return EWFA_ERR_NOMEM;
        void* dp = nullptr;
        QCHK(hipHostGetDevicePointer(&dp, f->dense.p, 0));
        f->dense_dev = static_cast<int8_t*>(dp);
    }
    const size_t nj = b.jobs.size();
    std::vector<size_t> first(nj + 1, 0);
    for (size_t k = 0; k < nj; ++k) first[k + 1] = first[k] + b.jobs[k]->n;
    parallel_ranges(nj, handout_threads(b.n), [&](size_t k) {
        const Job* j = b.jobs[k];
        const uint64_t woff = j->data_off / 8;
        const uint64_t coff = j->data_off + words_block(j);
        ewfa_pair_t* out = f->pairs.p + first[k];
        for (size_t i = 0; i < j->n; ++i) {
            ewfa_pair_t p = j->pairs[i];
            p.pattern_off += woff;
            p.text_off += woff;
            p.cigar_off += coff;
            out[i] = p;
        }
    });
    const Job* last = nullptr;
    for (const Job* j : b.jobs)
        if (j->up_ev && (!last || j->stage_seq > last->stage_seq)) last = j;
    if (last) QCHK(hipStreamWaitEvent(sl->cs, last->up_ev, 0));
    if (b.n == 0) return EWFA_OK;
    const StateLayout L(b.n, b.max_len);
    char* st = base + b.state_off;
    QCHK(hipMemcpyAsync(st + L.pairs, f->pairs.p, pairs_copy, hipMemcpyHostToDevice, sl->cs));
    const ewfa_pair_t* d_pairs = reinterpret_cast<const ewfa_pair_t*>(st + L.pairs);
    ewfa_result_t* d_results = reinterpret_cast<ewfa_result_t*>(st + L.results);
    int rc = ewfa::enqueue_align(sl->ctx, reinterpret_cast<const uint64_t*>(base), d_pairs, b.n, b.max_len, d_results,
                                 reinterpret_cast<int8_t*>(base), reinterpret_cast<uint64_t*>(st + L.ops),
                                 ewfa::op_scratch_bytes(b.n, b.max_len));
    if (rc != EWFA_OK) return rc;
    QCHK(ewfa::launch_compact(d_pairs, d_results, reinterpret_cast<const int8_t*>(base), b.n,
                              reinterpret_cast<uint64_t*>(st + L.scan), f->dense_dev, sl->cs));
    QCHK(hipMemcpyAsync(f->results.p, d_results, results_copy, hipMemcpyDeviceToHost, sl->cs));
    return EWFA_OK;
}

// The batch's upload events go back to the pool once the compute stream's
// wait has been enqueued (a wait holds the record it was given), or when
// enqueueing failed.
int enqueue_batch(Slot* sl, Flight* f) {
    const int rc = enqueue_batch_on(sl, f);
    std::lock_guard<std::mutex> lk(sl->s->m);
    for (Job* j : f->b.jobs) release_upload_event(sl->s, j);
    return rc;
}

// After the batch's device work (results and compacted CIGAR bytes are in the
// flight's pinned buffers): each job gets its results and its streams placed
// at its own offsets, jobs in parallel, and completes (`done(job, rc)`) as
// soon as its own outputs are in place.
template <typename Done>
int deliver_batch(Flight* f, Done&& done) {
    const Batch& b = f->b;
    const size_t nj = b.jobs.size();
    // each job's first pair and first dense byte
    std::vector<size_t> first(nj + 1, 0), dense_at(nj + 1, 0);
    for (size_t k = 0; k < nj; ++k) first[k + 1] = first[k] + b.jobs[k]->n;
    parallel_ranges(nj, handout_threads(b.n), [&](size_t k) {
        dense_at[k + 1] = ewfa::compact_total(f->pairs.p + first[k], f->results.p + first[k], b.jobs[k]->n);
    });
    for (size_t k = 0; k < nj; ++k) dense_at[k + 1] += dense_at[k];
    if (dense_at[nj] > b.caps) {
        for (Job* j : b.jobs) done(j, EWFA_ERR_STATE);
        return EWFA_ERR_STATE;
    }
    std::atomic<int> bad{0};
    parallel_ranges(nj, handout_threads(b.n), [&](size_t k) {
        Job* j = b.jobs[k];
        memcpy(j->results, f->results.p + first[k], j->n * sizeof(ewfa_result_t));
        const size_t w = ewfa::scatter_streams(f->dense.p + dense_at[k], dense_at[k + 1] - dense_at[k], j->pairs,
                                               j->results, j->n, j->cigars, 1);
        const int rc = w == SIZE_MAX ? EWFA_ERR_STATE : EWFA_OK;
        if (rc != EWFA_OK) bad = 1;
        done(j, rc);
    });
    return bad ? EWFA_ERR_STATE : EWFA_OK;
}

void worker_main(Slot* sl) {
    ewfa_server* s = sl->s;
    for (;;) {
        Flight* f = nullptr;
        {
            std::unique_lock<std::mutex> lk(sl->m);
            sl->cv.wait(lk, [&] { return !sl->idle.empty(); });
            f = sl->idle.front();
            sl->idle.pop_front();
        }
        f->b = Batch{};
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cv_work.wait(lk, [&] {
                if (s->ready.empty()) return s->stopping && s->incoming.empty() && s->staging == 0;
                if (form_batch(s, f->b)) return true;
                if (s->running > 0) return false;  // a batch in flight will free memory
                // nothing on the device: make room for the head job's state
                unstage_behind_head(s);
                if (form_batch(s, f->b)) return true;
                // its words and its state do not fit the arena side by side
                Job* j = s->ready.front();
                s->ready.pop_front();
                s->arena.free(j->data_off, j->data_bytes);
                s->staged_bytes -= j->data_bytes;
                release_upload_event(s, j);
                j->rc = EWFA_ERR_NOMEM;
                j->finished = Clock::now();
                j->done = true;
                ++s->completed;
                ++s->failed;
                s->cv_done.notify_all();
                return false;
            });
            if (!f->b.jobs.empty()) {
                s->running += (int)f->b.jobs.size();
                s->peak_running = std::max(s->peak_running, s->running);
                s->peak_batch_jobs = std::max(s->peak_batch_jobs, (int)f->b.jobs.size());
                ++s->batches;
            }
        }
        if (f->b.jobs.empty()) {  // stopping, every job done
            std::lock_guard<std::mutex> lk(sl->m);
            sl->idle.push_back(f);
            break;
        }
        s->cv_work.notify_all();  // the next head may fit beside this batch
        auto t = Clock::now();
        if (sl->init_rc != EWFA_OK) {
            f->rc = sl->init_rc;
            std::lock_guard<std::mutex> lk(s->m);
            for (Job* j : f->b.jobs) release_upload_event(s, j);
        } else {
            f->rc = enqueue_batch(sl, f);
        }
        if (f->rc == EWFA_OK && hipEventRecord(f->done, sl->cs) != hipSuccess) f->rc = EWFA_ERR_HIP;
        const double up = since(t);
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->upload_s += up;
        }
        {
            std::lock_guard<std::mutex> lk(sl->m);
            sl->busy.push_back(f);
        }
        sl->cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(sl->m);
        sl->worker_done = true;
    }
    sl->cv.notify_all();
}

void completer_main(Slot* sl) {
    ewfa_server* s = sl->s;
    for (;;) {
        Flight* f = nullptr;
        {
            std::unique_lock<std::mutex> lk(sl->m);
            sl->cv.wait(lk, [&] { return !sl->busy.empty() || sl->worker_done; });
            if (sl->busy.empty()) break;
            f = sl->busy.front();
            sl->busy.pop_front();
        }
        auto t = Clock::now();
        int rc = f->rc;
        if (rc == EWFA_OK && hipEventSynchronize(f->done) != hipSuccess) rc = EWFA_ERR_HIP;
        if (rc != EWFA_OK) {
            // nothing may still use the batch's memory when it is freed
            (void)hipStreamSynchronize(s->up);
            (void)hipStreamSynchronize(sl->cs);
        }
        const double al = since(t);
        // a job completes as soon as its own results and streams are placed;
        // its resident block is freed at once
        const auto done = [&](Job* j, int jrc) {
            std::lock_guard<std::mutex> lk(s->m);
            s->arena.free(j->data_off, j->data_bytes);
            j->rc = jrc;
            j->finished = Clock::now();
            j->done = true;
            ++s->completed;
            if (jrc != EWFA_OK) ++s->failed;
            --s->running;
            s->cv_done.notify_all();
            s->cv_work.notify_all();  // freed arena space may admit the queue's head
        };
        if (rc == EWFA_OK) {
            (void)deliver_batch(f, done);
        } else {
            for (Job* j : f->b.jobs) done(j, rc);
        }
        const double ho = since(t);
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->arena.free(f->b.state_off, f->b.state);
            s->align_s += al;
            s->handout_s += ho;
        }
        f->b = Batch{};
        s->cv_work.notify_all();
        s->cv_done.notify_all();
        {
            std::lock_guard<std::mutex> lk(sl->m);
            sl->idle.push_back(f);
        }
        sl->cv.notify_all();
    }
}

// Issues a job's words on the upload stream: straight from the caller's
// memory, or, below the DMA size, copied into the next bounce slot (once its
// previous copy has left it) and uploaded whole into the job's words block.
bool upload_words(ewfa_server* s, const Job* j) {
    const size_t b = j->n_words * 8;
    if (b == 0) return true;
    char* dst = s->arena.at(j->data_off);
    if (b >= kDmaMin || !s->bounce)
        return hipMemcpyAsync(dst, j->words, b, hipMemcpyHostToDevice, s->up) == hipSuccess;
    const int k = s->bounce_next;
    s->bounce_next = (k + 1) % kBounceSlots;
    if (s->bounce_used[k] && hipEventSynchronize(s->bounce_ev[k]) != hipSuccess) return false;
    char* src = s->bounce + (size_t)k * kDmaMin;
    memcpy(src, j->words, b);
    if (hipMemcpyAsync(dst, src, kDmaMin, hipMemcpyHostToDevice, s->up) != hipSuccess) return false;
    s->bounce_used[k] = true;
    return hipEventRecord(s->bounce_ev[k], s->up) == hipSuccess;
}

// Admits submitted jobs in arrival order: a block in the arena for the head
// job's words and CIGAR regions (while the staged jobs' blocks stay within
// stage_cap, a lone job excepted), then its words up on the server's upload
// stream at once.
void stager_main(ewfa_server* s) {
    (void)hipSetDevice(s->opts.device);
    for (;;) {
        Job* j = nullptr;
        hipEvent_t ev = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cv_work.wait(lk, [&] {
                if (s->incoming.empty()) return s->stopping;
                Job* h = s->incoming.front();
                if (s->staged_bytes > 0 && s->staged_bytes + h->data_bytes > s->stage_cap) return false;
                h->data_off = s->arena.alloc(h->data_bytes);
                return h->data_off != SIZE_MAX;
            });
            if (s->incoming.empty()) break;  // stopping
            j = s->incoming.front();
            s->incoming.pop_front();
            s->staged_bytes += j->data_bytes;
            ++s->staging;
            if (!s->ev_pool.empty()) {
                ev = s->ev_pool.back();
                s->ev_pool.pop_back();
            }
        }
        bool ok = ev || hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
        if (!ok) ev = nullptr;
        ok = ok && upload_words(s, j);
        ok = ok && hipEventRecord(ev, s->up) == hipSuccess;
        {
            std::lock_guard<std::mutex> lk(s->m);
            --s->staging;
            if (ok) {
                j->up_ev = ev;
                j->stage_seq = ++s->stage_seq;
                s->ready.push_back(j);
            } else {
                if (ev) s->ev_pool.push_back(ev);
                (void)hipStreamSynchronize(s->up);
                s->arena.free(j->data_off, j->data_bytes);
                s->staged_bytes -= j->data_bytes;
                j->rc = EWFA_ERR_HIP;
                j->finished = Clock::now();
                j->done = true;
                ++s->completed;
                ++s->failed;
                s->cv_done.notify_all();
            }
        }
        s->cv_work.notify_all();
    }
}

int slot_init(ewfa_server* s, Slot* sl) {
    ewfa_opts_t o;
    ewfa_default_opts(&o);
    o.device = s->opts.device;
    o.mode = EWFA_MODE_PERSISTENT;
    o.max_distance = s->opts.max_distance;
    o.policy_path = s->policy.empty() ? nullptr : s->policy.c_str();
    int rc = ewfa_create(&sl->ctx, &o);
    if (rc != EWFA_OK) return rc;
    // eager launches (each batch differs)
    ewfa_set_flags(sl->ctx, EWFA_FLAG_NO_GRAPH | (s->opts.flags & EWFA_FLAG_SERIAL));
    sl->cs = ewfa::context_stream(sl->ctx);
    for (Flight& f : sl->flights) {
        if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess) return EWFA_ERR_HIP;
        sl->idle.push_back(&f);
    }
    return EWFA_OK;
}

void slot_release(Slot* sl) {
    for (Flight& f : sl->flights) {
        if (f.done) (void)hipEventDestroy(f.done);
    }
    if (sl->ctx) ewfa_destroy(sl->ctx);
}

}  // namespace

extern "C" {

int ewfa_server_create(ewfa_server_t** out, const ewfa_server_opts_t* opts) {
    if (!out) return EWFA_ERR_ARG;
    *out = nullptr;
    if (ewfa_device_count() <= 0) return EWFA_ERR_NODEV;
    ewfa_server* s = new ewfa_server();
    if (opts) s->opts = *opts;
    if (s->opts.slots <= 0) s->opts.slots = 2;
    if (s->opts.flags & ~EWFA_FLAG_SERIAL) {
        delete s;
        return EWFA_ERR_ARG;
    }
    if (s->opts.slots > 64) s->opts.slots = 64;
    if (s->opts.arena_bytes == 0) s->opts.arena_bytes = (size_t)4 << 30;
    if (s->opts.max_batch_pairs == 0) s->opts.max_batch_pairs = (size_t)3 << 17;
    if (s->opts.max_batch_cigar_bytes == 0) s->opts.max_batch_cigar_bytes = (size_t)512 << 20;
    if (s->opts.policy_path) s->policy = s->opts.policy_path;
    s->opts.policy_path = nullptr;
    const int rc = s->arena.init(s->opts.device, s->opts.arena_bytes);
    if (rc != EWFA_OK) {
        delete s;
        return rc;
    }
    // staged jobs (uploaded, waiting for a batch) hold at most half the arena;
    // the rest is for the batches on the device
    s->stage_cap = s->opts.arena_bytes / 2;
    if (hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess) {
        s->arena.release();
        delete s;
        return EWFA_ERR_HIP;
    }
    // bounce slots: without them small jobs go up directly (a copy kernel)
    if (hipHostMalloc((void**)&s->bounce, kBounceSlots * kDmaMin, hipHostMallocNonCoherent) != hipSuccess)
        s->bounce = nullptr;
    for (hipEvent_t& e : s->bounce_ev)
        if (s->bounce && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            (void)hipHostFree(s->bounce);
            s->bounce = nullptr;
        }
    for (int i = 0; i < s->opts.slots; ++i) {
        Slot* sl = new Slot();
        sl->s = s;
        sl->init_rc = slot_init(s, sl);  // a failed slot fails its batches
        s->slots.push_back(sl);
    }
    s->stager = std::thread(stager_main, s);
    for (Slot* sl : s->slots) {
        sl->worker = std::thread(worker_main, sl);
        sl->completer = std::thread(completer_main, sl);
    }
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
    s->stager.join();
    for (Slot* sl : s->slots) {
        sl->worker.join();
        sl->completer.join();
        slot_release(sl);
        delete sl;
    }
    for (auto& kv : s->jobs) delete kv.second;
    if (s->up) {
        (void)hipStreamSynchronize(s->up);
        (void)hipStreamDestroy(s->up);
    }
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->bounce_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->bounce) (void)hipHostFree(s->bounce);
    s->arena.release();
    delete s;
    return EWFA_OK;
}

int ewfa_server_submit(ewfa_server_t* s, const uint64_t* h_words, size_t n_words, const ewfa_pair_t* h_pairs,
                       size_t n, ewfa_result_t* h_results, int8_t* h_cigars, size_t cigar_bytes,
                       uint64_t* job_id) {
    if (!s || !job_id || (n && (!h_words || !h_pairs || !h_results || !h_cigars))) return EWFA_ERR_ARG;
    uint32_t max_len = 0;
    size_t caps = 0;
    for (size_t i = 0; i < n; ++i) {
        const ewfa_pair_t& p = h_pairs[i];
        max_len = std::max(max_len, std::max(p.plen, p.tlen));
        caps += p.cigar_cap;
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
    j->caps = caps;
    j->data_bytes = words_bytes(n_words) + Arena::round(cigar_bytes);
    if (j->data_bytes + StateLayout(n, max_len).total > s->arena.capacity()) {
        delete j;
        return EWFA_ERR_RANGE;
    }
    j->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        j->id = s->next_id++;
        s->jobs[j->id] = j;
        s->incoming.push_back(j);
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
    out->upload_s = s->upload_s;
    out->align_s = s->align_s;
    out->handout_s = s->handout_s;
    return EWFA_OK;
}

}  // extern "C"
