// host.cpp -- CPU-side parts of libewfa: packing, CIGAR text, the validity
// check, the pair-file reader and the synthetic workload generator.
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "ewfa.h"
#include "ewfa_internal.h"
#include "runtime.h"

namespace {

template <typename F>
void parallel_for(size_t n, int threads, F&& f) {
    if (threads <= 1 || n < 2) {
        f(0, n);
        return;
    }
    const size_t t = std::min<size_t>((size_t)threads, n);
    std::vector<std::thread> pool;
    const size_t per = n / t;
    for (size_t i = 0; i < t; ++i) {
        const size_t lo = per * i, hi = (i + 1 == t) ? n : per * (i + 1);
        pool.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto& th : pool) th.join();
}

double now_s() {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

}  // namespace

namespace ewfa {

size_t compact_total(const ewfa_pair_t* pairs, const ewfa_result_t* results, size_t n) {
    size_t t = 0;
    for (size_t i = 0; i < n; ++i) t += written_cigar_bytes(results[i], pairs[i]);
    return t;
}

size_t scatter_streams(const int8_t* dense, size_t dense_cap, const ewfa_pair_t* pairs, const ewfa_result_t* results,
                       size_t n, int8_t* h_cigars, int threads) {
    // pass 1: each block's byte count; pass 2: each block places its pairs'
    // streams from its own start in the dense buffer
    const size_t t = n >= 65536 ? std::max(1, threads) : 1;
    std::vector<size_t> start(t + 1, 0);
    const size_t per = (n + t - 1) / t;
    parallel_for(t, (int)t, [&](size_t lo, size_t hi) {
        for (size_t b = lo; b < hi; ++b) {
            size_t sum = 0;
            for (size_t i = b * per; i < std::min(n, (b + 1) * per); ++i) sum += written_cigar_bytes(results[i], pairs[i]);
            start[b + 1] = sum;
        }
    });
    for (size_t b = 0; b < t; ++b) start[b + 1] += start[b];
    if (start[t] > dense_cap) return SIZE_MAX;
    parallel_for(t, (int)t, [&](size_t lo, size_t hi) {
        for (size_t b = lo; b < hi; ++b) {
            size_t off = start[b];
            for (size_t i = b * per; i < std::min(n, (b + 1) * per); ++i) {
                const size_t w = written_cigar_bytes(results[i], pairs[i]);
                if (w) memcpy(h_cigars + pairs[i].cigar_off, dense + off, w);
                off += w;
            }
        }
    });
    return start[t];
}

}  // namespace ewfa

extern "C" {

size_t ewfa_packed_words(size_t len) { return (len + 31) / 32 + 1; }

// Base code (c & 6) >> 1 (src/kernels.cu:358), 32 bases per word, first base
// in the two high bits; one zero pad word after the data.
void ewfa_pack_host(const char* ascii, size_t len, uint64_t* words) {
    const unsigned char* s = reinterpret_cast<const unsigned char*>(ascii);
    const size_t full = len / 32;
    for (size_t w = 0; w < full; ++w) {
        uint64_t x = 0;
        const unsigned char* b = s + 32 * w;
        for (int i = 0; i < 32; ++i) x = (x << 2) | ((b[i] & 6u) >> 1);
        words[w] = x;
    }
    const size_t rem = len - 32 * full;
    size_t w = full;
    if (rem) {
        uint64_t x = 0;
        const unsigned char* b = s + 32 * full;
        for (size_t i = 0; i < rem; ++i) x |= (uint64_t)((b[i] & 6u) >> 1) << (62 - 2 * i);
        words[w++] = x;
    }
    words[w] = 0;
}

int ewfa_layout_pairs(const uint32_t* plen, const uint32_t* tlen, size_t n, uint32_t cigar_cap,
                      ewfa_pair_t* pairs, size_t* n_words, size_t* cigar_bytes) {
    if (n && (!plen || !tlen || !pairs)) return EWFA_ERR_ARG;
    size_t w = 0, c = 0;
    const uint32_t fixed = (cigar_cap + 3u) & ~3u;
    for (size_t i = 0; i < n; ++i) {
        ewfa_pair_t& p = pairs[i];
        p.plen = plen[i];
        p.tlen = tlen[i];
        p.pattern_off = w;
        w += ewfa_packed_words(p.plen);
        p.text_off = w;
        w += ewfa_packed_words(p.tlen);
        const uint64_t cap = fixed ? fixed : (((uint64_t)p.plen + p.tlen + 4 + 3) & ~3ull);
        if (cap > 0xFFFFFFFCull) return EWFA_ERR_RANGE;
        p.cigar_off = c;
        p.cigar_cap = (uint32_t)cap;
        p.flags = 0;
        c += cap;
    }
    if (n_words) *n_words = w;
    if (cigar_bytes) *cigar_bytes = c;
    return EWFA_OK;
}

int ewfa_pack_pairs_host(const char* pat, const char* txt, size_t stride, const ewfa_pair_t* pairs,
                         size_t n, uint64_t* words, int threads) {
    if (n && (!pat || !txt || !pairs || !words)) return EWFA_ERR_ARG;
    for (size_t i = 0; i < n; ++i)
        if (pairs[i].plen > stride || pairs[i].tlen > stride) return EWFA_ERR_ARG;
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            ewfa_pack_host(pat + i * stride, pairs[i].plen, words + pairs[i].pattern_off);
            ewfa_pack_host(txt + i * stride, pairs[i].tlen, words + pairs[i].text_off);
        }
    });
    return EWFA_OK;
}

// Cigars::generate_ascii_cigar (src/wavefront.cuh:146-184): negative bytes add
// to a pending match count; before an op byte a pending count is written as
// the op byte's own decimal value followed by 'M' (sprintf of curr_char at
// :170), then the op.  A pending count at the end is dropped.
int ewfa_cigar_text(const int8_t* stream, size_t len, char* out, size_t out_cap) {
    size_t n = 0;
    size_t pending = 0;
    for (size_t i = 0; i < len; ++i) {
        const int8_t b = stream[i];
        if (b == 0) break;
        if (b < 0) {
            pending += (size_t)(-(int)b);
            continue;
        }
        if (pending > 0) {
            char num[8];
            const int l = snprintf(num, sizeof num, "%d", (int)b);
            if (n + (size_t)l + 1 >= out_cap) return EWFA_ERR_RANGE;
            memcpy(out + n, num, (size_t)l);
            n += (size_t)l;
            out[n++] = 'M';
            pending = 0;
        }
        if (n + 1 >= out_cap) return EWFA_ERR_RANGE;
        out[n++] = (char)b;
    }
    if (n >= out_cap) return EWFA_ERR_RANGE;
    out[n] = 0;
    return (int)n;
}

// Extended CIGAR with merged runs: '=' match, 'X' mismatch, 'I' text-only,
// 'D' pattern-only.
int ewfa_cigar_text_standard(const int8_t* stream, size_t len, char* out, size_t out_cap) {
    size_t n = 0;
    char cur = 0;
    size_t run = 0;
    auto flush = [&]() -> bool {
        if (!run) return true;
        char num[24];
        const int l = snprintf(num, sizeof num, "%zu%c", run, cur);
        if (n + (size_t)l + 1 > out_cap) return false;
        memcpy(out + n, num, (size_t)l);
        n += (size_t)l;
        run = 0;
        return true;
    };
    for (size_t i = 0; i < len; ++i) {
        const int8_t b = stream[i];
        if (b == 0) break;
        const char op = b < 0 ? '=' : (char)b;
        const size_t cnt = b < 0 ? (size_t)(-(int)b) : 1;
        if (op != cur) {
            if (!flush()) return EWFA_ERR_RANGE;
            cur = op;
        }
        run += cnt;
    }
    if (!flush()) return EWFA_ERR_RANGE;
    if (n >= out_cap) return EWFA_ERR_RANGE;
    out[n] = 0;
    return (int)n;
}

// Cigars::check_cigar (src/wavefront.cuh:297-419) on raw ASCII bytes.
int ewfa_check_cigar(const char* pattern, size_t plen, const char* text, size_t tlen,
                     const int8_t* stream, size_t len) {
    size_t pp = 0, tp = 0;
    for (size_t i = 0; i < len; ++i) {
        const int8_t b = stream[i];
        if (b == 0) break;
        if (b < 0) {
            for (int j = 0; j < -(int)b; ++j) {
                if (pp >= plen || tp >= tlen || pattern[pp] != text[tp]) return 0;
                ++pp;
                ++tp;
            }
            continue;
        }
        switch (b) {
            case 'I': ++tp; break;
            case 'D': ++pp; break;
            case 'X':
                if (pp >= plen || tp >= tlen || pattern[pp] == text[tp]) return 0;
                ++pp;
                ++tp;
                break;
            default: return -1;
        }
    }
    return (pp == plen && tp == tlen) ? 1 : 0;
}

// SequenceReader::read_n_alignments (utils/sequence_reader.cu:108-201).
int ewfa_read_pairs(const char* path, size_t seq_len, size_t n, int threads, ewfa_pairs_file_t* out) {
    if (!out) return EWFA_ERR_ARG;
    memset(out, 0, sizeof *out);
    out->n = n;
    threads = std::max(1, threads);
    const size_t nn = std::max<size_t>(n, 1);
    out->pairs = (ewfa_pair_t*)calloc(nn, sizeof(ewfa_pair_t));
    out->ascii_off = (uint64_t*)calloc(2 * nn, sizeof(uint64_t));
    if (!out->pairs || !out->ascii_off) {
        ewfa_free_pairs(out);
        return ewfa::set_error(EWFA_ERR_NOMEM, "Could not allocate memory.\n");
    }
    const double t0 = now_s();
    const char* data = nullptr;
    size_t size = 0;
    if (seq_len != 0) {  // -l 0 reads nothing (utils/sequence_reader.cu:109-112)
        const int fd = open(path, O_RDONLY);
        struct stat st;
        if (fd < 0 || fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
            if (fd >= 0) close(fd);
            ewfa_free_pairs(out);
            return ewfa::set_error(EWFA_ERR_IO, "Could not open the file %s\n", path);
        }
        size = (size_t)st.st_size;
        if (size) {
            void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) {
                close(fd);
                ewfa_free_pairs(out);
                return ewfa::set_error(EWFA_ERR_IO, "Could not open the file %s\n", path);
            }
            out->map = m;
            out->map_len = size;
            data = static_cast<const char*>(m);
        }
        close(fd);
    }
    // The first 2n lines, [starts[i], ends[i]) each (a line includes its
    // '\n'; the last one of the file may lack it).  The file is scanned in
    // rounds of `threads` blocks, each block's newlines found by one thread,
    // until 2n lines are known or the file ends.
    std::vector<size_t> starts, ends;
    constexpr size_t BLOCK = (size_t)8 << 20;
    size_t scan = 0;  // first byte not yet scanned
    size_t line = 0;  // start of the next line
    while (starts.size() < 2 * n && line < size) {
        const size_t nb = std::min<size_t>((size_t)threads, (size - scan + BLOCK - 1) / BLOCK);
        std::vector<std::vector<size_t>> nl(nb);
        parallel_for(nb, (int)nb, [&](size_t lo, size_t hi) {
            for (size_t k = lo; k < hi; ++k) {
                const size_t b0 = scan + k * BLOCK, b1 = std::min(size, b0 + BLOCK);
                for (const char* q = data + b0; q < data + b1;) {
                    const void* f = memchr(q, '\n', (size_t)(data + b1 - q));
                    if (!f) break;
                    nl[k].push_back((size_t)((const char*)f - data));
                    q = (const char*)f + 1;
                }
            }
        });
        for (size_t k = 0; k < nb && starts.size() < 2 * n; ++k)
            for (size_t i = 0; i < nl[k].size() && starts.size() < 2 * n; ++i) {
                starts.push_back(line);
                ends.push_back(nl[k][i] + 1);
                line = nl[k][i] + 1;
            }
        scan = std::min(size, scan + nb * BLOCK);
        if (scan == size && starts.size() < 2 * n && line < size) {
            starts.push_back(line);  // the last line has no newline
            ends.push_back(size);
            line = size;
        }
    }
    // markers: '>' on pattern lines, '<' on text lines; the reference stops at
    // the first bad line (utils/sequence_reader.cu:154-176)
    size_t lines = starts.size();
    int rc = EWFA_OK;
    for (size_t i = 0; i < lines; ++i) {
        const char marker = (i % 2 == 0) ? '>' : '<';
        if (data[starts[i]] != marker) {
            const std::string text(data + starts[i], ends[i] - starts[i]);
            rc = ewfa::set_error(EWFA_ERR_FORMAT, "Invalid sequence format on line %d.\n    %s\n", (int)i,
                                 text.c_str());
            lines = i;
            break;
        }
    }
    out->scan_s = now_s() - t0;
    // sequences: bytes [1, len-1) of each line, i.e. the newline is assumed;
    // they stay in the mapping, where ascii_off points
    if (out->map) {
        out->ascii = static_cast<char*>(out->map);
    } else if (!(out->ascii = (char*)calloc(1, 1))) {
        ewfa_free_pairs(out);
        return ewfa::set_error(EWFA_ERR_NOMEM, "Could not allocate memory.\n");
    }
    std::vector<uint32_t> slen(2 * nn, 0);
    std::vector<uint64_t> woff(2 * nn, 0);
    size_t words_total = 0;
    uint32_t max_len = 0;
    for (size_t j = 0; j < 2 * n; ++j) {
        if (j < lines) {
            const size_t len = ends[j] - starts[j];
            slen[j] = len >= 2 ? (uint32_t)(len - 2) : 0u;
            out->ascii_off[j] = starts[j] + 1;
        }
        woff[j] = words_total;
        words_total += ewfa_packed_words(slen[j]);
        max_len = std::max(max_len, slen[j]);
    }
    out->words = (uint64_t*)malloc(std::max<size_t>(words_total, 1) * 8);
    if (!out->words) {
        ewfa_free_pairs(out);
        return ewfa::set_error(EWFA_ERR_NOMEM, "Could not allocate memory.\n");
    }
    out->n_words = words_total;
    out->max_len = max_len;
    const double t1 = now_s();
    parallel_for(2 * n, threads, [&](size_t lo, size_t hi) {
        for (size_t j = lo; j < hi; ++j)
            ewfa_pack_host(out->ascii + out->ascii_off[j], slen[j], out->words + woff[j]);
    });
    out->pack_s = now_s() - t1;
    for (size_t i = 0; i < n; ++i) {
        ewfa_pair_t& p = out->pairs[i];
        p.pattern_off = woff[2 * i];
        p.text_off = woff[2 * i + 1];
        p.plen = slen[2 * i];
        p.tlen = slen[2 * i + 1];
    }
    return rc;
}

void ewfa_free_pairs(ewfa_pairs_file_t* f) {
    if (!f) return;
    free(f->words);
    free(f->pairs);
    if (f->map) munmap(f->map, f->map_len);
    else free(f->ascii);
    free(f->ascii_off);
    memset(f, 0, sizeof *f);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// synthetic workloads

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded from (seed, pair index) through splitmix64.
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t idx) {
        uint64_t x = seed ^ (idx * 0xD1B54A32D192ED03ull);
        for (auto& w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t m) { return (uint64_t)(((unsigned __int128)next() * m) >> 64); }
};

enum EvType : int { EV_SUB = 0, EV_INS = 1, EV_DEL = 2 };
struct Event {
    uint32_t pos;
    int type;
    uint32_t len;
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

// Generates pair i into pat/txt; returns lengths.
void synth_one(const ewfa_synth_t& sp, uint64_t i, char* pat, char* txt, uint32_t* plen, uint32_t* tlen) {
    Rng r(sp.seed, i);
    uint32_t L;
    double err;
    if (sp.kind == 0) {
        L = sp.len;
        err = sp.err;
    } else if (sp.kind == 1) {
        const double lo = log((double)sp.len_lo), hi = log((double)sp.len_hi);
        L = (uint32_t)llround(exp(lo + (hi - lo) * r.uniform()));
        err = sp.err_lo + (sp.err_hi - sp.err_lo) * r.uniform();
    } else {
        L = sp.len_lo + (uint32_t)r.below((uint64_t)(sp.len_hi - sp.len_lo) + 1);
        err = sp.err_lo + (sp.err_hi - sp.err_lo) * r.uniform();
    }
    // pattern: iid uniform bases
    for (uint32_t j = 0; j < L; j += 32) {
        uint64_t x = r.next();
        for (uint32_t b = 0; b < 32 && j + b < L; ++b, x >>= 2) pat[j + b] = kBases[x & 3];
    }
    uint32_t E = (uint32_t)llround(err * (double)L);
    if (E > L) E = L;
    // edit events: single sub/ins/del, plus (kind 2) indel runs taking 30%
    // of the budget with geometric lengths of mean 5 (p = 0.2)
    std::vector<Event> ev;
    uint32_t singles = E;
    if (sp.kind == 2) {
        uint32_t budget = (uint32_t)llround(0.3 * (double)E);
        singles = E - budget;
        while (budget > 0) {
            uint32_t m = 1;
            while (r.uniform() >= 0.2) ++m;
            if (m > budget) m = budget;
            budget -= m;
            ev.push_back({0u, (r.next() & 1) ? EV_INS : EV_DEL, m});
        }
    }
    {
        const uint32_t third = singles / 3;
        for (uint32_t t = 0; t < 3; ++t)
            for (uint32_t j = 0; j < third; ++j) ev.push_back({0u, (int)t, 1u});
        for (uint32_t j = 0; j < singles - 3 * third; ++j) ev.push_back({0u, (int)r.below(3), 1u});
    }
    uint32_t ne = (uint32_t)ev.size();
    if (ne > L) ne = L;
    ev.resize(ne);
    // distinct uniform positions (Floyd's sampling), dealt to the events in
    // shuffled order
    std::vector<uint32_t> picks;
    picks.reserve(ne);
    {
        std::vector<uint8_t> taken(L, 0);
        for (uint32_t j = L - ne; j < L; ++j) {
            uint32_t t = (uint32_t)r.below((uint64_t)j + 1);
            if (taken[t]) t = j;
            taken[t] = 1;
            picks.push_back(t);
        }
    }
    for (uint32_t j = ne; j > 1; --j) std::swap(picks[j - 1], picks[r.below(j)]);
    for (uint32_t j = 0; j < ne; ++j) ev[j].pos = picks[j];
    std::sort(ev.begin(), ev.end(), [](const Event& a, const Event& b) { return a.pos < b.pos; });
    // apply in one forward pass (equivalent to applying in descending
    // position order on the original coordinates); a deletion run stops
    // before the next event and at the end
    uint32_t o = 0, src = 0;
    size_t e = 0;
    while (src < L) {
        if (e < ev.size() && ev[e].pos == src) {
            const Event& x = ev[e++];
            if (x.type == EV_SUB) {
                const int cur = (int)(strchr(kBases, pat[src]) - kBases);
                txt[o++] = kBases[(cur + 1 + (int)r.below(3)) & 3];
                ++src;
            } else if (x.type == EV_INS) {
                for (uint32_t m = 0; m < x.len; ++m) txt[o++] = kBases[r.next() & 3];
                txt[o++] = pat[src++];
            } else {
                uint32_t m = x.len;
                const uint32_t stop = (e < ev.size()) ? ev[e].pos : L;
                if (m > stop - src) m = stop - src;
                src += m;
            }
        } else {
            txt[o++] = pat[src++];
        }
    }
    *plen = L;
    *tlen = o;
}

}  // namespace

extern "C" {

uint32_t ewfa_synth_max_len(const ewfa_synth_t* sp) {
    if (!sp) return 0;
    double hi, err;
    if (sp->kind == 0) {
        hi = sp->len;
        err = sp->err;
    } else {
        hi = sp->len_hi;
        err = sp->err_hi;
    }
    if (err > 1.0) err = 1.0;
    return (uint32_t)(hi + ceil(err * hi) + 2);
}

int ewfa_synth_pairs(const ewfa_synth_t* sp, size_t first, size_t n, char* pat, char* txt, size_t stride,
                     uint32_t* plen, uint32_t* tlen, int threads) {
    if (!sp || (n && (!pat || !txt || !plen || !tlen))) return EWFA_ERR_ARG;
    if (sp->kind < 0 || sp->kind > 2) return EWFA_ERR_ARG;
    if (sp->kind == 0 && sp->len == 0) return EWFA_ERR_ARG;
    if (sp->kind != 0 && (sp->len_lo == 0 || sp->len_hi < sp->len_lo)) return EWFA_ERR_ARG;
    if (stride < ewfa_synth_max_len(sp)) return EWFA_ERR_ARG;
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i)
            synth_one(*sp, first + i, pat + i * stride, txt + i * stride, plen + i, tlen + i);
    });
    return EWFA_OK;
}

}  // extern "C"
