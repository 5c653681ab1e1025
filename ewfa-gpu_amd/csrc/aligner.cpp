// aligner.cpp -- wfe.aligner, the drop-in command line (src/aligner.cu:40-239).
//
// Same flags, same stdout lines in the same order:
//   Using device "<name>", with compute capability M.m
//   [-p] <index>: <CIGAR text>          one line per pair, batch by batch
//   Aligned executed. Wall time <s>s
// The wall time covers the same window as the reference's (after the file
// read, through the last printed line).  -t T splits the pairs into T
// contiguous ranges, one host thread and one device context each
// (src/aligner.cu:45-48); whole lines of different threads may interleave.
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ewfa.h"

namespace {

std::mutex g_out_mutex;

[[noreturn]] void fatal(const char* msg) {
    fputs(msg, stderr);
    fflush(stdout);
    fflush(stderr);
    exit(1);
}

struct Params {
    const ewfa_pairs_file_t* file;
    const uint64_t* words;  // the file's packed words, in pinned memory
    int64_t seq_len;
    size_t num_alignments;
    size_t batch_size;
    bool print;
    bool check;
    bool standard;
    bool plain;
    const char* policy;  // autotuned launch policy file, or NULL
    int max_distance;
    unsigned tid;
    unsigned total;
    int rc;
};

// The reference's per-pair CIGAR buffer (src/wavefront.cuh:48-50).
uint32_t cigar_capacity(int64_t seq_len) {
    int64_t c = seq_len + (4 - seq_len % 4);
    if (c < 4) c = 4;
    return (uint32_t)c;
}

uint32_t exact_capacity(const ewfa_pair_t& p) {
    uint64_t c = (uint64_t)p.plen + p.tlen + 4;
    return (uint32_t)((c + 3) & ~3ull);
}

// One batch worth of host buffers.
// A host buffer that only grows and is never zero-filled: the streams are
// read only up to the bytes the library wrote.
struct Bytes {
    std::unique_ptr<int8_t[]> p;
    size_t n = 0, cap = 0;
    int8_t* data() { return p.get(); }
    const int8_t* data() const { return p.get(); }
    size_t size() const { return n; }
    void resize(size_t m) {
        if (m > cap) {
            std::unique_ptr<int8_t[]> q(new int8_t[m]);
            if (n) memcpy(q.get(), p.get(), n);
            p = std::move(q);
            cap = m;
        }
        n = m;
    }
};

// A batch's descriptors in pinned host memory (the library copies them to
// the device by DMA; from ordinary memory the copy goes through a staging
// buffer in pieces).  Grows only.
struct PinnedPairs {
    ewfa_pair_t* p = nullptr;
    size_t n = 0, cap = 0;
    PinnedPairs() = default;
    PinnedPairs(const PinnedPairs&) = delete;
    PinnedPairs& operator=(const PinnedPairs&) = delete;
    ~PinnedPairs() {
        if (p) ewfa_host_free(p);
    }
    bool assign(const ewfa_pair_t* first, const ewfa_pair_t* last) {
        const size_t m = (size_t)(last - first);
        if (m > cap) {
            void* q = nullptr;
            if (ewfa_host_alloc(m * sizeof(ewfa_pair_t), &q) != EWFA_OK) return false;
            if (p) ewfa_host_free(p);
            p = static_cast<ewfa_pair_t*>(q);
            cap = m;
        }
        if (m) memcpy(p, first, m * sizeof(ewfa_pair_t));
        n = m;
        return true;
    }
    ewfa_pair_t* data() { return p; }
    const ewfa_pair_t* data() const { return p; }
    ewfa_pair_t& operator[](size_t i) { return p[i]; }
    const ewfa_pair_t& operator[](size_t i) const { return p[i]; }
};

struct Batch {
    size_t first = 0;   // global index of the batch's first pair
    size_t count = 0;
    std::vector<ewfa_result_t> results;
    Bytes cigars;
    PinnedPairs pairs;
};

// Host threads that format one batch's text (shared by the -t workers).
int g_format_threads = 1;

// Writes "<index>: " at out, returns its length.
size_t put_index(char* out, size_t idx) {
    char tmp[24];
    size_t n = 0;
    do {
        tmp[n++] = (char)('0' + idx % 10);
        idx /= 10;
    } while (idx);
    for (size_t i = 0; i < n; ++i) out[i] = tmp[n - 1 - i];
    out[n] = ':';
    out[n + 1] = ' ';
    return n + 2;
}

// Pairs [lo, hi) of a batch: their "<index>: <text>" lines appended to out
// (the reference's print loop, src/wavefront.cu:293-300) and the number whose
// stream passes the reference's check (src/wavefront.cu:302-320).
int format_range(const Params& p, const Batch& b, size_t lo, size_t hi, std::string& out) {
    int correct = 0;
    for (size_t i = lo; i < hi; ++i) {
        const ewfa_pair_t& pr = b.pairs[i];
        const ewfa_result_t& r = b.results[i];
        const int8_t* s = b.cigars.data() + pr.cigar_off;
        const size_t len = std::min<size_t>((size_t)std::max(r.cigar_len, 0), pr.cigar_cap);
        if (p.print) {
            // a stream byte prints as at most 4 characters ("68MD"); standard
            // runs as at most 11 ("4294967295=")
            const size_t room = 24 + (p.standard ? 12 : 4) * len + 2;
            const size_t at = out.size();
            out.resize(at + room);
            char* o = &out[at];
            size_t k = put_index(o, b.first + i);
            int n = p.standard ? ewfa_cigar_text_standard(s, len, o + k, room - k - 1)
                               : ewfa_cigar_text(s, len, o + k, room - k - 1);
            if (n < 0) n = 0;
            k += (size_t)n;
            o[k++] = '\n';
            out.resize(at + k);
        }
        if (p.check) {
            const size_t g = b.first + i;
            const char* pat = p.file->ascii + p.file->ascii_off[2 * g];
            const char* txt = p.file->ascii + p.file->ascii_off[2 * g + 1];
            if (ewfa_check_cigar(pat, pr.plen, txt, pr.tlen, s, len) == 1) ++correct;
        }
    }
    return correct;
}

// The batch's output: its lines formatted on g_format_threads threads over
// contiguous ranges, then written in index order as whole lines.
void format_batch(const Params& p, const Batch& b) {
    std::string head;
    if (p.max_distance > 0) {
        // the reference prints these from the device before the batch's lines
        for (size_t i = 0; i < b.count; ++i)
            if (b.results[i].status == EWFA_PAIR_CAPPED || b.results[i].status == EWFA_PAIR_UNDEFINED) {
                char line[96];
                snprintf(line, sizeof line, "Max distance reached at alignment %zu\n", i);
                head += line;
            }
    }
    const size_t nt = (p.print || p.check) ? std::max<size_t>(1, std::min<size_t>((size_t)g_format_threads,
                                                                                  b.count / 2048 + 1))
                                           : 0;
    std::vector<std::string> outs(nt);
    std::vector<int> correct(nt, 0);
    if (nt) {
        const size_t per = (b.count + nt - 1) / nt;
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t)
            pool.emplace_back([&, t] {
                correct[t] = format_range(p, b, std::min(b.count, t * per), std::min(b.count, (t + 1) * per), outs[t]);
            });
        correct[0] = format_range(p, b, 0, std::min(b.count, per), outs[0]);
        for (std::thread& th : pool) th.join();
    }
    std::lock_guard<std::mutex> lk(g_out_mutex);
    fwrite(head.data(), 1, head.size(), stdout);
    for (const std::string& o : outs) fwrite(o.data(), 1, o.size(), stdout);
    if (p.check) {
        int c = 0;
        for (int x : correct) c += x;
        fprintf(stderr, "Correct alignments: %d/%zu\n", c, b.count);
    }
}

void* worker(void* arg) {
    Params& p = *static_cast<Params*>(arg);
    // contiguous ranges; the last thread takes the remainder (src/aligner.cu:45-48)
    size_t todo = p.num_alignments / p.total;
    const size_t skip = todo * p.tid;
    if (p.tid == p.total - 1) todo += p.num_alignments % p.total;
    size_t batch = p.batch_size;
    if (batch > todo) {
        batch = todo;
        fprintf(stderr, "[Thread %d] Changing batch size to %d", (int)p.tid, (int)todo);
    }
    if (todo == 0) return nullptr;
    if (batch == 0) batch = todo;

    ewfa_opts_t o;
    ewfa_default_opts(&o);
    o.mode = p.plain ? EWFA_MODE_PLAIN : EWFA_MODE_PERSISTENT;
    o.policy_path = p.policy;
    o.max_distance = p.max_distance;
    ewfa_ctx_t* ctx = nullptr;
    if (ewfa_create(&ctx, &o) != EWFA_OK) {
        fprintf(stderr, "%s\n", ewfa_last_error());
        p.rc = 1;
        return nullptr;
    }
    const ewfa_pairs_file_t& f = *p.file;
    const uint32_t cap = cigar_capacity(p.seq_len);
    // this worker's pairs read one contiguous range of the packed words (the
    // reader lays them out pair by pair): it goes up once, in pieces that the
    // batches wait for only as far as they read, and every batch aligns
    // against it (the reference uploads each batch's sequences,
    // src/wavefront.cu:265-284)
    const uint64_t w0 = f.pairs[skip].pattern_off;
    const ewfa_pair_t& last = f.pairs[skip + todo - 1];
    const uint64_t w1 = last.text_off + ewfa_packed_words(last.tlen);
    if (ewfa_upload_words_async(ctx, p.words + w0, (size_t)(w1 - w0)) != EWFA_OK) {
        std::lock_guard<std::mutex> lk(g_out_mutex);
        fprintf(stderr, "%s\n", ewfa_last_error());
        p.rc = 1;
        ewfa_destroy(ctx);
        return nullptr;
    }

    // a printer thread formats batch b while the device aligns batch b+1
    std::mutex m;
    std::condition_variable cv;
    Batch* ready = nullptr;
    bool quit = false;
    std::thread printer([&] {
        for (;;) {
            Batch* b;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return ready || quit; });
                if (!ready) return;
                b = ready;
            }
            if (p.print || p.check || p.max_distance > 0) format_batch(p, *b);
            {
                std::lock_guard<std::mutex> lk(m);
                ready = nullptr;
            }
            cv.notify_all();
        }
    });

    Batch bufs[2];
    int cur = 0;
    for (size_t off = 0; off < todo && p.rc == 0; off += batch) {
        Batch& b = bufs[cur];
        b.first = skip + off;
        b.count = std::min(batch, todo - off);
        if (!b.pairs.assign(f.pairs + b.first, f.pairs + b.first + b.count)) {
            std::lock_guard<std::mutex> lk(g_out_mutex);
            fprintf(stderr, "%s\n", ewfa_last_error());
            p.rc = 1;
            break;
        }
        for (size_t i = 0; i < b.count; ++i) {
            b.pairs[i].pattern_off -= w0;
            b.pairs[i].text_off -= w0;
            b.pairs[i].cigar_off = (uint64_t)i * cap;
            b.pairs[i].cigar_cap = cap;
        }
        b.results.resize(b.count);
        b.cigars.resize((size_t)b.count * cap);
        int rc = ewfa_align_resident(ctx, b.pairs.data(), b.count, b.results.data(), b.cigars.data(),
                                     b.cigars.size());
        if (rc == EWFA_OK) {
            // streams longer than the reference's buffer: align again with
            // room for any stream of the pair
            std::vector<size_t> over;
            for (size_t i = 0; i < b.count; ++i)
                if (b.results[i].status == EWFA_PAIR_OVERFLOW) over.push_back(i);
            if (!over.empty()) {
                std::vector<ewfa_pair_t> op(over.size());
                std::vector<ewfa_result_t> orr(over.size());
                size_t total = 0;
                for (size_t j = 0; j < over.size(); ++j) {
                    op[j] = b.pairs[over[j]];
                    op[j].cigar_off = total;
                    op[j].cigar_cap = exact_capacity(op[j]);
                    total += op[j].cigar_cap;
                }
                std::vector<int8_t> oc(total);
                rc = ewfa_align_resident(ctx, op.data(), op.size(), orr.data(), oc.data(), oc.size());
                if (rc == EWFA_OK) {
                    // append the long streams after the batch's buffer
                    const size_t base = b.cigars.size();
                    b.cigars.resize(base + total);
                    memcpy(b.cigars.data() + base, oc.data(), total);
                    for (size_t j = 0; j < over.size(); ++j) {
                        b.pairs[over[j]].cigar_off = base + op[j].cigar_off;
                        b.pairs[over[j]].cigar_cap = op[j].cigar_cap;
                        b.results[over[j]] = orr[j];
                    }
                }
            }
        }
        if (rc != EWFA_OK) {
            std::lock_guard<std::mutex> lk(g_out_mutex);
            fprintf(stderr, "%s\n", ewfa_last_error());
            p.rc = 1;
            break;
        }
        for (size_t i = 0; i < b.count; ++i)
            if (b.results[i].status == EWFA_PAIR_TOO_DISTANT || b.results[i].status == EWFA_PAIR_TOO_LONG) {
                std::lock_guard<std::mutex> lk(g_out_mutex);
                fprintf(stderr, "alignment %zu: distance beyond what the device can hold\n", b.first + i);
            }
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return ready == nullptr; });
            ready = &b;
        }
        cv.notify_all();
        cur ^= 1;
        {
            // the other buffer must be free before it is refilled
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return ready != &bufs[cur]; });
        }
    }
    {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return ready == nullptr; });
        quit = true;
    }
    cv.notify_all();
    printer.join();
    ewfa_destroy(ctx);
    return nullptr;
}

}  // namespace

int main(int argc, char** argv) {
    // the device line comes before argument parsing (src/aligner.cu:141-151)
    char name[256];
    int major = 0, minor = 0;
    if (ewfa_device_count() <= 0) fatal("No available HIP devices detected.");
    if (ewfa_device_info(0, name, sizeof name, &major, &minor) != EWFA_OK) fatal(ewfa_last_error());
    printf("Using device \"%s\", with compute capability %d.%d\n", name, major, minor);

    ewfa_cli_args_t a;
    if (!ewfa_cli_parse(argc, argv, &a)) {
        ewfa_cli_usage();
        exit(1);
    }
    const size_t num_alignments = (size_t)a.num_alignments;
    size_t batch_size = num_alignments;
    if (a.batch_set) batch_size = (size_t)a.batch_size;
    unsigned num_threads = 1;
    if (a.threads_set) num_threads = (unsigned)a.cpu_threads;
    if (num_threads < 1) fatal("Minimum needed is 1 thread");
    if (batch_size > num_alignments) fatal("batch_size must be <= than the number of alignments.");
    if (num_threads > 4096) fatal("Could not allocate memory.\n");
    if (a.max_distance < 0 || a.max_distance > (1 << 30)) fatal("Invalid max distance.\n");
    if (a.policy) {
        // read before the timer, as every worker's context will
        ewfa_bucket_policy_t table[16];
        if (ewfa_policy_read(a.policy, table, 16) < 0) {
            fprintf(stderr, "Can not read the launch policy %s.\n", a.policy);
            exit(1);
        }
    }

    ewfa_pairs_file_t f;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int rc = ewfa_read_pairs(a.file, (size_t)a.seq_len, num_alignments, (int)std::min(hw, 16u), &f);
    // the reader's messages are the reference's (utils/sequence_reader.cu:124, :157-158)
    if (rc == EWFA_ERR_IO || rc == EWFA_ERR_FORMAT) fputs(ewfa_last_error(), stderr);
    if (rc == EWFA_ERR_NOMEM) fatal("Could not allocate memory.\n");
    if (rc != EWFA_OK) {
        fprintf(stderr, "Could not read the sequences from file %s\n", a.file);
        fflush(stdout);
        exit(1);
    }

    // the packed words in pinned memory (the reference reads into pinned
    // memory too, utils/sequence_reader.cu:80), copied before the timer starts
    uint64_t* pinned = nullptr;
    if (ewfa_host_alloc(std::max<size_t>(f.n_words, 1) * 8, (void**)&pinned) != EWFA_OK)
        fatal("Could not allocate memory.\n");
    {
        const size_t nt = std::min<size_t>(std::min(hw, 16u), std::max<size_t>(1, f.n_words >> 20));
        std::vector<std::thread> pool;
        const size_t per = (f.n_words + nt - 1) / nt;
        for (size_t t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                const size_t lo = std::min(f.n_words, t * per), hi = std::min(f.n_words, (t + 1) * per);
                if (hi > lo) memcpy(pinned + lo, f.words + lo, (hi - lo) * 8);
            });
        for (std::thread& th : pool) th.join();
    }

    // formatting threads per worker: the host's cores shared by the -t workers
    g_format_threads = (int)std::max(1u, std::min(hw, 16u) / std::max(1u, num_threads));

    struct timespec t0, t1;
    clock_gettime(CLOCK_REALTIME, &t0);
    std::vector<Params> params(num_threads);
    std::vector<pthread_t> threads(num_threads);
    for (unsigned i = 0; i < num_threads; ++i) {
        Params& p = params[i];
        p.file = &f;
        p.words = pinned;
        p.seq_len = a.seq_len;
        p.num_alignments = num_alignments;
        p.batch_size = batch_size;
        p.print = a.print_cigars != 0;
        p.check = a.check != 0;
        p.standard = a.standard_cigar != 0;
        p.plain = a.plain != 0;
        p.policy = a.policy;
        p.max_distance = (int)a.max_distance;
        p.tid = i;
        p.total = num_threads;
        p.rc = 0;
    }
    for (unsigned i = 1; i < num_threads; ++i)
        if (pthread_create(&threads[i], nullptr, worker, &params[i])) fatal("Can not create threads.");
    worker(&params[0]);
    for (unsigned i = 1; i < num_threads; ++i)
        if (pthread_join(threads[i], nullptr)) fatal("Can not join thread.");
    fflush(stdout);
    clock_gettime(CLOCK_REALTIME, &t1);
    const double seconds = (double)((t1.tv_sec + t1.tv_nsec * 1e-9) - (double)(t0.tv_sec + t0.tv_nsec * 1e-9));
    printf("Aligned executed. Wall time %fs\n", seconds);
    int status = 0;
    for (const Params& p : params) status |= p.rc;
    ewfa_host_free(pinned);
    ewfa_free_pairs(&f);
    return status;
}
