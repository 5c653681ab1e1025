/*
 * ewfa_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of quim0/eWFA-GPU's alignment path.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it, and only as the checker
 * or as the timed CPU baseline; the product never links it.
 *
 * How each function is pinned (checked against something that is not a
 * retyping of its own formula) -- the tests live in tests/test_oracle.py:
 *   or_code / or_pack_* : closed form (c&6)>>1 on all 256 bytes; round trip
 *                         through the reference LUT {A,C,T,G}; byte layout of
 *                         the reference's swizzle checked on hand words.
 *   or_extend           : brute-force byte-by-byte LCP on random inputs.
 *   or_align (distance) : O(nm) Levenshtein DP (or_levenshtein, itself pinned
 *                         by textbook cases and brute force on tiny inputs).
 *   or_align (ops)      : an independent formulation written in Python in the
 *                         tests: furthest-reaching offsets derived from the
 *                         full DP matrix plus a per-cell op table and an
 *                         explicit traceback; and the hand-derived known-answer
 *                         pairs (identity, single sub/ins/del, tie cases).
 *   or_cigar_stream     : validity (or_check_cigar), op count == distance,
 *                         127-splitting invariants, hand KATs.
 *   or_cigar_text       : hand KATs (e.g. [-1,'X',-2] -> "88MX").
 *   or_check_cigar      : planted mistakes (wrong op, short stream) fail.
 *   or_read_pairs       : hand-written files covering each quirk.
 */
#define _GNU_SOURCE
#include "ewfa_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int or_code(int c) { return (c & 6) >> 1; } /* src/kernels.cu:358 */

void or_codes(const char *ascii, size_t len, uint8_t *codes) {
    for (size_t i = 0; i < len; i++) codes[i] = (uint8_t)or_code((unsigned char)ascii[i]);
}

/* src/kernels.cu:351-392: four bases per byte, first base in the high bits of
 * the byte, bytes swizzled so that a little-endian u32 holds base 0 in bits
 * 31:30.  Equivalent closed form: base j at word j/16, bits 31-2(j%16). */
void or_pack_ref32(const char *ascii, size_t len, uint32_t *words) {
    size_t nw = (len + 15) / 16;
    memset(words, 0, nw * sizeof(uint32_t));
    for (size_t j = 0; j < len; j++) {
        uint32_t c = (uint32_t)or_code((unsigned char)ascii[j]);
        words[j / 16] |= c << (30 - 2 * (j % 16));
    }
}

/* src/wavefront.cuh:270-293, LUT {A,C,T,G}. */
void or_unpack_ref32(const uint32_t *words, size_t len, char *ascii) {
    static const char lut[4] = {'A', 'C', 'T', 'G'};
    for (size_t j = 0; j < len; j++) ascii[j] = lut[(words[j / 16] >> (30 - 2 * (j % 16))) & 3];
}

void or_pack_u64(const char *ascii, size_t len, uint64_t *words) {
    size_t nw = (len + 31) / 32 + 1;
    memset(words, 0, nw * sizeof(uint64_t));
    for (size_t j = 0; j < len; j++) {
        uint64_t c = (uint64_t)or_code((unsigned char)ascii[j]);
        words[j / 32] |= c << (62 - 2 * (j % 32));
    }
}

/* src/kernels.cu:27-109 (16 bases per step there; the step width does not
 * change the result): offset + number of equal codes from v=off-k, h=off,
 * stopping at either sequence end. */
int or_extend(const uint8_t *p, int plen, const uint8_t *t, int tlen, int k, int off) {
    int v = off - k, h = off, acc = 0;
    while (v < plen && h < tlen && p[v] == t[h]) { v++; h++; acc++; }
    return off + acc;
}

/* Lexicographic max of (value, op) with op priority DEL(3) > SUB(2) > INS(1):
 * the reference packs (value << 2) | op into an int32 and takes max
 * (src/kernels.cu:136-152); (x<<2)|op == 4x+op orders exactly like the pair. */
static void lexmax(int del, int sub, int ins, int *val, int *op) {
    int v = del, o = OR_OP_DEL;
    if (sub > v) { v = sub; o = OR_OP_SUB; }
    if (ins > v) { v = ins; o = OR_OP_INS; }
    *val = v; *op = o;
}

void or_alignment_free(or_alignment *a) {
    if (a && a->ops) { free(a->ops); a->ops = NULL; }
}

int or_op_at(const or_alignment *a, int s) {
    if (s < 0 || s / 32 >= a->nwords) return OR_OP_NOOP;
    return (int)((a->ops[s / 32] >> (2 * (s % 32))) & 3);
}

/* Wavefront state over k in [-(dlim+2), dlim+2]: two offset generations and
 * two history generations, W words of history per diagonal.  Grown (doubled)
 * when the search passes dlim, so memory follows the distance reached. */
typedef struct {
    int dlim, ctr, width, W;
    int *o[2];
    uint64_t *h[2];
} or_wf;

static void or_wf_free(or_wf *s) {
    for (int g = 0; g < 2; g++) { free(s->o[g]); free(s->h[g]); s->o[g] = NULL; s->h[g] = NULL; }
}

static int or_wf_alloc(or_wf *s, int dlim) {
    s->dlim = dlim;
    s->ctr = dlim + 2;
    s->width = 2 * dlim + 5;
    s->W = dlim / 32 + 1;
    for (int g = 0; g < 2; g++) {
        s->o[g] = (int *)malloc(sizeof(int) * (size_t)s->width);
        s->h[g] = (uint64_t *)calloc((size_t)s->width * s->W, sizeof(uint64_t));
        if (!s->o[g] || !s->h[g]) { or_wf_free(s); return -1; }
        for (int i = 0; i < s->width; i++) s->o[g][i] = -1;
    }
    return 0;
}

static int or_wf_grow(or_wf *s, int dlim) {
    or_wf n;
    memset(&n, 0, sizeof n);
    if (or_wf_alloc(&n, dlim) != 0) return -1;
    for (int g = 0; g < 2; g++) {
        for (int k = -(s->dlim + 2); k <= s->dlim + 2; k++) {
            n.o[g][k + n.ctr] = s->o[g][k + s->ctr];
            memcpy(n.h[g] + (size_t)(k + n.ctr) * n.W, s->h[g] + (size_t)(k + s->ctr) * s->W,
                   sizeof(uint64_t) * (size_t)s->W);
        }
    }
    or_wf_free(s);
    *s = n;
    return 0;
}

/* src/kernels.cu:182-310 with src/kernels.cu:111-180 inlined.
 * Offsets start at -1 outside the computed cone (src/kernels.cu:205-219);
 * each cell takes the lexicographic max of its three candidates, copies the
 * history of the chosen predecessor k+op-2 and ORs its op in at bit 2(d-1)
 * (src/kernels.cu:154-171); the search stops at the first d with |t|<=d and
 * O_d[t]==tlen (src/kernels.cu:243-270).  With a cap, d runs while d < cap
 * (src/kernels.cu:259) and the result is the level cap-1 history at the target
 * diagonal (src/kernels.cu:288-298). */
int or_align(const uint8_t *p, int plen, const uint8_t *t, int tlen,
             int max_distance, or_alignment *out) {
    memset(out, 0, sizeof(*out));
    const int maxlen = plen > tlen ? plen : tlen;
    or_wf s;
    memset(&s, 0, sizeof s);
    int dlim0 = max_distance > 0 ? max_distance : 64;
    if (or_wf_alloc(&s, dlim0) != 0) return -1;
    int gc = 0; /* generation holding the current level */
    const int tk = tlen - plen;       /* EWAVEFRONT_DIAGONAL(tlen, plen) */
    const int tk_abs = tk < 0 ? -tk : tk;
    const int toff = tlen;            /* EWAVEFRONT_OFFSET(tlen, plen) */

    s.o[gc][s.ctr] = or_extend(p, plen, t, tlen, 0, 0);
    int d = 0, done = (tk_abs <= 0 && s.o[gc][s.ctr + tk] == toff);
    int result_gen = 1 - gc; /* d = 0: the untouched (all-zero) history */
    if (!done) {
        for (d = 1; ; d++) {
            if (max_distance > 0 && d >= max_distance) break;
            if (d > maxlen + 1) { or_wf_free(&s); return -1; } /* D <= max(plen,tlen) */
            if (d > s.dlim && or_wf_grow(&s, 2 * s.dlim) != 0) return -1;
            int *cur = s.o[gc] + s.ctr, *nxt = s.o[1 - gc] + s.ctr;
            uint64_t *hc = s.h[gc], *hn = s.h[1 - gc];
            const int nw_used = (d - 1) / 32 + 1; /* words holding ops 0..d-1 */
            for (int k = -d; k <= d; k++) {
                int val, op;
                lexmax(cur[k + 1], cur[k] + 1, cur[k - 1] + 1, &val, &op);
                const uint64_t *src = hc + (size_t)(k + op - 2 + s.ctr) * s.W;
                uint64_t *dst = hn + (size_t)(k + s.ctr) * s.W;
                memcpy(dst, src, sizeof(uint64_t) * (size_t)nw_used);
                dst[(d - 1) / 32] |= (uint64_t)op << (2 * ((d - 1) % 32));
                nxt[k] = or_extend(p, plen, t, tlen, k, val);
            }
            if (tk_abs <= d && nxt[tk] == toff) { done = 1; result_gen = 1 - gc; break; }
            gc = 1 - gc;
        }
    }

    int nw;
    const uint64_t *hist;
    if (done) {
        out->distance = d;
        nw = d / 32 + 1;
        hist = s.h[result_gen] + (size_t)(tk + s.ctr) * s.W;
    } else {
        /* Cap reached.  The reference's history arrays span k in [-cap, cap]
         * (all zero where no level wrote them); beyond that its read is out
         * of bounds. */
        out->distance = max_distance;
        out->capped = 1;
        nw = (max_distance - 1) / 32 + 1;
        hist = s.h[gc] + (size_t)(tk + s.ctr) * s.W;
        if (tk_abs > max_distance) { out->undefined = 1; nw = 0; }
    }
    out->ops = (uint64_t *)calloc((size_t)(nw > 0 ? nw : 1), sizeof(uint64_t));
    if (!out->ops) { or_wf_free(&s); return -1; }
    out->nwords = nw;
    if (nw > 0) {
        /* a d=0 result or a cap with |t| == cap reads diagonals no level wrote */
        if (tk + s.ctr >= 0 && tk + s.ctr < s.width && nw <= s.W)
            memcpy(out->ops, hist, sizeof(uint64_t) * (size_t)nw);
        if (!done && tk_abs == max_distance) memset(out->ops, 0, sizeof(uint64_t) * (size_t)nw);
        if (done && d == 0) out->ops[0] = 0;
    }
    or_wf_free(&s);
    return 0;
}

/* src/kernels.cu:398-498: forward walk over the ops, re-extending before each
 * op and once more after the last, match runs emitted as -127 chunks then the
 * remainder.  An op 0 stops the walk after its preceding run (src/kernels.cu:471-478). */
int or_cigar_stream(const uint8_t *p, int plen, const uint8_t *t, int tlen,
                    const or_alignment *a, int8_t *out, int cap) {
    int n = 0, k = 0, off = 0;
#define EMIT(b) do { if (n < cap) out[n] = (int8_t)(b); n++; } while (0)
#define EMIT_RUN(r) do { int r_ = (r); for (int j_ = 0; j_ < r_ / 127; j_++) EMIT(-127); \
                         if (r_ % 127) EMIT(-(r_ % 127)); } while (0)
    if (a->undefined) return 0;
    for (int s = 0; s < a->distance; s++) {
        int prev = off;
        off = or_extend(p, plen, t, tlen, k, off);
        EMIT_RUN(off - prev);
        int op = or_op_at(a, s);
        switch (op) {
            case OR_OP_DEL: EMIT('D'); k--; break;
            case OR_OP_SUB: EMIT('X'); off++; break;
            case OR_OP_INS: EMIT('I'); k++; off++; break;
            default: return n; /* the reference prints an error and returns */
        }
    }
    {
        int prev = off;
        off = or_extend(p, plen, t, tlen, k, off);
        EMIT_RUN(off - prev);
    }
#undef EMIT
#undef EMIT_RUN
    return n;
}

/* src/wavefront.cuh:146-184.  Negative bytes accumulate a match count; before
 * each op byte, if the count is positive, the *op byte's* decimal value and
 * 'M' are written (sprintf("%d", curr_char) at :170), then the op itself.  A
 * trailing count is never written. */
int or_cigar_text(const int8_t *stream, int stream_len, char *out, int cap) {
    int n = 0;
    size_t m = 0;
    for (int i = 0; i < stream_len && stream[i] != 0; i++) {
        int8_t c = stream[i];
        if (c < 0) { m += (size_t)(-c); continue; }
        if (m > 0) {
            char buf[16];
            int l = snprintf(buf, sizeof buf, "%d", (int)c);
            if (n + l + 1 >= cap) return -1;
            memcpy(out + n, buf, (size_t)l); n += l;
            out[n++] = 'M';
            m = 0;
        }
        if (n + 1 >= cap) return -1;
        out[n++] = (char)c;
    }
    if (n >= cap) return -1;
    out[n] = 0;
    return n;
}

/* src/wavefront.cuh:297-419: runs must match on raw ASCII, 'X' must mismatch,
 * 'I' consumes text, 'D' consumes pattern, anything else is fatal; both
 * sequences must be consumed exactly. */
int or_check_cigar(const char *pattern, int plen, const char *text, int tlen,
                   const int8_t *stream, int stream_len) {
    int pp = 0, tp = 0;
    for (int i = 0; i < stream_len && stream[i] != 0; i++) {
        int8_t c = stream[i];
        if (c < 0) {
            for (int j = 0; j > c; j--) {
                if (pp >= plen || tp >= tlen) return 0; /* reads past the end */
                if (pattern[pp] != text[tp]) return 0;
                pp++; tp++;
            }
        } else if (c == 'I') {
            tp++;
        } else if (c == 'D') {
            pp++;
        } else if (c == 'X') {
            if (pp >= plen || tp >= tlen) return 0;
            if (pattern[pp] == text[tp]) return 0;
            pp++; tp++;
        } else {
            return -1; /* WF_FATAL("Invalid CIGAR generated.") */
        }
    }
    return (pp == plen && tp == tlen) ? 1 : 0;
}

int or_levenshtein(const uint8_t *p, int plen, const uint8_t *t, int tlen) {
    int *row = (int *)malloc(sizeof(int) * (size_t)(tlen + 1));
    if (!row) return -1;
    for (int j = 0; j <= tlen; j++) row[j] = j;
    for (int i = 1; i <= plen; i++) {
        int diag = row[0];
        row[0] = i;
        for (int j = 1; j <= tlen; j++) {
            int up = row[j];
            int best = diag + (p[i - 1] != t[j - 1]);
            if (up + 1 < best) best = up + 1;
            if (row[j - 1] + 1 < best) best = row[j - 1] + 1;
            row[j] = best;
            diag = up;
        }
    }
    int r = row[tlen];
    free(row);
    return r;
}

/* utils/sequence_reader.cu:108-201.  getline loop; line 2i must start with
 * '>' (pattern), line 2i+1 with '<' (text); bytes [1, len-1) are kept, which
 * assumes a trailing newline.  Stops after n pairs; missing pairs keep length 0. */
int or_read_pairs(const char *path, size_t seq_len, size_t n,
                  char *slab, size_t stride, int64_t *plen, int64_t *tlen) {
    for (size_t i = 0; i < n; i++) { plen[i] = 0; tlen[i] = 0; }
    if (seq_len == 0) return 1; /* :109-112 */
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char *buf = NULL;
    size_t cap = 0;
    ssize_t length;
    size_t idx = 0;
    int ret = 1;
    while (idx / 2 < n && (length = getline(&buf, &cap, fp)) > 0) {
        size_t pair = idx / 2;
        char *dst = slab + idx * stride;
        char marker = (idx % 2 == 0) ? '>' : '<';
        if (buf[0] != marker) { ret = 0; break; }
        int64_t l = (int64_t)length - 2;
        if (l < 0) { ret = 0; break; } /* the reference memcpy's (size_t)-1 here */
        memcpy(dst, buf + 1, (size_t)l < stride ? (size_t)l : stride);
        if (idx % 2 == 0) plen[pair] = l; else tlen[pair] = l;
        idx++;
    }
    free(buf);
    fclose(fp);
    return ret;
}

int or_pair_full(const char *pattern, int plen, const char *text, int tlen,
                 int max_distance, int *distance, int8_t *stream, int stream_cap,
                 int *stream_len, char *text_out, int text_cap) {
    uint8_t *pc = (uint8_t *)malloc((size_t)plen + 1);
    uint8_t *tc = (uint8_t *)malloc((size_t)tlen + 1);
    if (!pc || !tc) { free(pc); free(tc); return -2; }
    or_codes(pattern, (size_t)plen, pc);
    or_codes(text, (size_t)tlen, tc);
    or_alignment a;
    if (or_align(pc, plen, tc, tlen, max_distance, &a) != 0) { free(pc); free(tc); return -2; }
    int sl = or_cigar_stream(pc, plen, tc, tlen, &a, stream, stream_cap);
    if (distance) *distance = a.distance;
    if (stream_len) *stream_len = sl;
    int used = sl < stream_cap ? sl : stream_cap;
    if (text_out && text_cap > 0) or_cigar_text(stream, used, text_out, text_cap);
    int ok = or_check_cigar(pattern, plen, text, tlen, stream, used);
    or_alignment_free(&a);
    free(pc); free(tc);
    return ok;
}

typedef struct {
    const char *pat, *txt;
    const int64_t *pat_off, *txt_off;
    const int32_t *plen, *tlen;
    int64_t lo, hi;
    int max_distance;
    int32_t *dist, *slen;
    int8_t *streams;
    int64_t stride;
    int text;
    int64_t ok;
} or_batch_job;

static void *or_batch_worker(void *arg) {
    or_batch_job *j = (or_batch_job *)arg;
    int scap = 256, tcap = 1024;
    int8_t *stream = (int8_t *)malloc((size_t)scap);
    char *text = (char *)malloc((size_t)tcap);
    for (int64_t i = j->lo; i < j->hi; i++) {
        int pl = j->plen[i], tl = j->tlen[i];
        int need = pl + tl + 8;
        if (need > scap) { scap = need; stream = (int8_t *)realloc(stream, (size_t)scap); }
        if (4 * need > tcap) { tcap = 4 * need; text = (char *)realloc(text, (size_t)tcap); }
        int d = 0, sl = 0;
        int r = or_pair_full(j->pat + j->pat_off[i], pl, j->txt + j->txt_off[i], tl,
                             j->max_distance, &d, stream, scap, &sl, j->text ? text : NULL, tcap);
        if (r == 1) j->ok++;
        if (j->dist) j->dist[i] = d;
        if (j->slen) j->slen[i] = sl;
        if (j->streams) {
            int64_t c = sl < j->stride ? sl : j->stride;
            memcpy(j->streams + i * j->stride, stream, (size_t)c);
        }
    }
    free(stream); free(text);
    return NULL;
}

int64_t or_batch_streams(const char *pat, const int64_t *pat_off, const int32_t *plen,
                         const char *txt, const int64_t *txt_off, const int32_t *tlen,
                         int64_t n, int max_distance, int threads,
                         int32_t *dist, int32_t *slen, int8_t *streams, int64_t stride) {
    return or_batch_text(pat, pat_off, plen, txt, txt_off, tlen, n, max_distance, threads, 1,
                         dist, slen, streams, stride);
}

int64_t or_batch_text(const char *pat, const int64_t *pat_off, const int32_t *plen,
                      const char *txt, const int64_t *txt_off, const int32_t *tlen,
                      int64_t n, int max_distance, int threads, int text,
                      int32_t *dist, int32_t *slen, int8_t *streams, int64_t stride) {
    if (threads < 1) threads = 1;
    or_batch_job *jobs = (or_batch_job *)calloc((size_t)threads, sizeof(or_batch_job));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    int64_t per = n / threads;
    for (int i = 0; i < threads; i++) {
        or_batch_job *j = &jobs[i];
        j->pat = pat; j->txt = txt; j->pat_off = pat_off; j->txt_off = txt_off;
        j->plen = plen; j->tlen = tlen; j->max_distance = max_distance;
        j->dist = dist; j->slen = slen; j->streams = streams; j->stride = stride; j->text = text;
        j->lo = per * i;
        j->hi = (i == threads - 1) ? n : per * (i + 1); /* last takes the rest */
        if (i > 0) pthread_create(&th[i], NULL, or_batch_worker, j);
    }
    or_batch_worker(&jobs[0]);
    int64_t ok = jobs[0].ok;
    for (int i = 1; i < threads; i++) { pthread_join(th[i], NULL); ok += jobs[i].ok; }
    free(jobs); free(th);
    return ok;
}

int64_t or_batch(const char *pat, const int64_t *pat_off, const int32_t *plen,
                 const char *txt, const int64_t *txt_off, const int32_t *tlen,
                 int64_t n, int max_distance, int threads,
                 int32_t *dist, int32_t *slen) {
    return or_batch_streams(pat, pat_off, plen, txt, txt_off, tlen, n, max_distance, threads,
                            dist, slen, NULL, 0);
}
