/*
 * ewfa_oracle.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, CPU-only restatement of quim0/eWFA-GPU's alignment path, used by
 * tests/, the smoke check and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (libewfa, wfe.aligner) links or calls this code.
 *
 * Every function cites the reference file:line it restates.  Pins (what checks
 * each function against something other than itself) are listed in the header
 * comment of ewfa_oracle.c.
 */
#ifndef EWFA_ORACLE_H
#define EWFA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Backtrace op codes, src/wavefront_structures.h:66-71. */
enum { OR_OP_NOOP = 0, OR_OP_INS = 1, OR_OP_SUB = 2, OR_OP_DEL = 3 };

/* Result of one alignment, the CPU twin of WF_backtrace_result_t
 * (src/wavefront_structures.h:57-60) with an unbounded op history. */
typedef struct {
    int distance;     /* D, or max_distance when the cap was hit            */
    int capped;       /* 1 when max_distance stopped the search              */
    int undefined;    /* 1 when the reference's behaviour is undefined (the
                         capped target diagonal lies outside its arrays)     */
    int nwords;       /* number of valid words in ops                        */
    uint64_t *ops;    /* op s (0-based) at bits 2(s%32)..+1 of word s/32     */
} or_alignment;

/* Base code (c & 6) >> 1, src/kernels.cu:358. */
int or_code(int c);

/* ASCII -> codes, one byte per base. */
void or_codes(const char *ascii, size_t len, uint8_t *codes);

/* The reference's packed layout: 32-bit words, base j at word j/16, bits
 * 31-2(j%16)..30-2(j%16) (src/kernels.cu:313-394).  Writes ceil(len/16) words. */
void or_pack_ref32(const char *ascii, size_t len, uint32_t *words);

/* Inverse of or_pack_ref32 through the reference's LUT {A,C,T,G}
 * (src/wavefront.cuh:261-295). */
void or_unpack_ref32(const uint32_t *words, size_t len, char *ascii);

/* This build's packed layout: 64-bit words, base j at word j/32, bits
 * 63-2(j%32)..62-2(j%32).  Writes ceil(len/32)+1 words (one zero pad word). */
void or_pack_u64(const char *ascii, size_t len, uint64_t *words);

/* Extend along diagonal k from offset off: off + LCP (src/kernels.cu:27-109,
 * host twin src/wavefront.cuh:128-143).  Codes, one byte per base. */
int or_extend(const uint8_t *p, int plen, const uint8_t *t, int tlen, int k, int off);

/* WFA edit distance with path-copied op histories (src/kernels.cu:111-310).
 * max_distance > 0 reproduces the reference's cap (it searches d < cap);
 * max_distance == 0 extends the same rule without a cap.
 * Returns 0, or -1 on allocation failure.  Free with or_alignment_free. */
int or_align(const uint8_t *p, int plen, const uint8_t *t, int tlen,
             int max_distance, or_alignment *out);
void or_alignment_free(or_alignment *a);

/* Op s of an alignment (0 = NOOP when s is past the recorded history). */
int or_op_at(const or_alignment *a, int s);

/* Compressed CIGAR stream (src/kernels.cu:398-498).  Writes at most cap bytes
 * and returns the full stream length (which may exceed cap). */
int or_cigar_stream(const uint8_t *p, int plen, const uint8_t *t, int tlen,
                    const or_alignment *a, int8_t *out, int cap);

/* The reference's printed CIGAR text (src/wavefront.cuh:146-184), including
 * its two quirks.  stream_len bytes of stream are read, stopping at a zero
 * byte.  Writes a NUL-terminated string, returns its length (or -1 if cap is
 * too small). */
int or_cigar_text(const int8_t *stream, int stream_len, char *out, int cap);

/* The reference's validity check (src/wavefront.cuh:297-419) on raw ASCII.
 * Returns 1 valid, 0 invalid, -1 for a byte the reference treats as fatal. */
int or_check_cigar(const char *pattern, int plen, const char *text, int tlen,
                   const int8_t *stream, int stream_len);

/* Levenshtein distance on codes, O(plen*tlen) time, O(tlen) space. */
int or_levenshtein(const uint8_t *p, int plen, const uint8_t *t, int tlen);

/* Pair-file reader with the reference's quirks (utils/sequence_reader.cu:108-201).
 * seq_len is -l.  Fills n pairs: pattern i / text i copied into caller buffers
 * of stride `stride` bytes each (pattern i at slab + 2*i*stride, text at
 * slab + (2*i+1)*stride), lengths into plen/tlen.  Lines longer than stride
 * are truncated in the copy but their full length is reported (the reference
 * overflows its slot there; that case is undefined).  Pairs missing from the
 * file are left with length 0.  Returns 1 ok, 0 on a format error (the
 * reference's `return false`), -1 if the file cannot be opened. */
int or_read_pairs(const char *path, size_t seq_len, size_t n,
                  char *slab, size_t stride, int64_t *plen, int64_t *tlen);

/* One full pair through the reference path: align, stream, text and check.
 * Returns the check result (1/0/-1) and fills distance, stream and text. */
int or_pair_full(const char *pattern, int plen, const char *text, int tlen,
                 int max_distance, int *distance, int8_t *stream, int stream_cap,
                 int *stream_len, char *text_out, int text_cap);

/* Time-able batch: n pairs of ASCII (pattern i at pat + pat_off[i], text at
 * txt + txt_off[i]) through or_pair_full on `threads` pthreads over contiguous
 * ranges (the -t split of src/aligner.cu:45-48).  Distances go to dist (may be
 * NULL), stream lengths to slen (may be NULL).  Returns the number of pairs the
 * check accepted. */
int64_t or_batch(const char *pat, const int64_t *pat_off, const int32_t *plen,
                 const char *txt, const int64_t *txt_off, const int32_t *tlen,
                 int64_t n, int max_distance, int threads,
                 int32_t *dist, int32_t *slen);

/* or_batch that also copies each pair's stream into streams + i*stride (at
 * most stride bytes; streams may be NULL). */
int64_t or_batch_streams(const char *pat, const int64_t *pat_off, const int32_t *plen,
                         const char *txt, const int64_t *txt_off, const int32_t *tlen,
                         int64_t n, int max_distance, int threads,
                         int32_t *dist, int32_t *slen, int8_t *streams, int64_t stride);

/* or_batch_streams with the CIGAR text step on (text = 1, the reference's
 * -p path: Cigars::generate_ascii_cigar, src/wavefront.cuh:146-184, called at
 * src/wavefront.cu:293-300) or off (text = 0: align, stream and check only). */
int64_t or_batch_text(const char *pat, const int64_t *pat_off, const int32_t *plen,
                      const char *txt, const int64_t *txt_off, const int32_t *tlen,
                      int64_t n, int max_distance, int threads, int text,
                      int32_t *dist, int32_t *slen, int8_t *streams, int64_t stride);

#ifdef __cplusplus
}
#endif
#endif
