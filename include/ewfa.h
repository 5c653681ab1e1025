/*
 * ewfa.h -- C ABI of libewfa, the MI355X (gfx950) WFA edit-distance aligner.
 *
 * Drop-in for quim0/eWFA-GPU's alignment path.  Each entry point names the
 * reference interface it replaces (file:line in the reference tree).  All
 * functions return EWFA_OK (0) or a negative EWFA_ERR_* code and never exit
 * the process; ewfa_last_error() gives the message of the calling thread's
 * last failure.  Device pointers are plain HIP device addresses; `stream` is a
 * hipStream_t passed as void* (NULL = the library context's own stream).
 *
 * Packed sequence layout ("words"): 2 bits per base, code (c & 6) >> 1
 * (A/a=0, C/c=1, T/t/U=2, G/g/N=3; src/kernels.cu:358), base j of a sequence
 * at word j/32, bits 63-2(j%32)..62-2(j%32).  Each sequence occupies
 * ceil(len/32) words followed by one readable pad word.
 *
 * CIGAR stream layout: the reference's compressed stream (src/kernels.cu:398-498):
 * a match run of r bases is floor(r/127) bytes of -127 then -(r%127) if
 * nonzero; ops are the bytes 'X' (substitution), 'I' (consumes text) and 'D'
 * (consumes pattern).  Zero-terminated when shorter than its capacity.
 */
#ifndef EWFA_H
#define EWFA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define EWFA_OK            0
#define EWFA_ERR_ARG      -1   /* invalid argument */
#define EWFA_ERR_HIP      -2   /* a HIP runtime call failed */
#define EWFA_ERR_NOMEM    -3   /* host or device allocation failed */
#define EWFA_ERR_NODEV    -4   /* no HIP device */
#define EWFA_ERR_IO       -5   /* file could not be opened */
#define EWFA_ERR_FORMAT   -6   /* pair-file marker error */
#define EWFA_ERR_RANGE    -7   /* a pair exceeds what the library can hold */
#define EWFA_ERR_STATE    -8   /* call out of order */

/* ---- per-pair status (ewfa_result_t.status) ----------------------------- */
#define EWFA_PAIR_OK         0  /* distance and stream are final */
#define EWFA_PAIR_CAPPED     1  /* max_distance reached: the reference's
                                   capped output (src/kernels.cu:288-303) */
#define EWFA_PAIR_OVERFLOW   2  /* stream longer than cigar_cap: truncated,
                                   cigar_len holds the full length */
#define EWFA_PAIR_UNDEFINED  3  /* capped and |tlen-plen| > max_distance: the
                                   reference reads out of bounds here; the
                                   stream is left empty */
#define EWFA_PAIR_BADARG     4  /* descriptor invalid (alignment, capacity) */
#define EWFA_PAIR_TOO_LONG   5  /* plen/tlen above the call's max_len */
#define EWFA_PAIR_TOO_DISTANT 6 /* distance beyond ewfa_max_distance (or, in
                                   the allocation-free call, beyond what LDS
                                   holds for the pair's length) */

/* One pair: the replacement of WF_element (src/wavefront_structures.h:44-50)
 * plus the addresses PATTERN_PTR/TEXT_PTR and Cigars::get_cigar_ascii derive
 * (src/wavefront_structures.h:41-42, src/wavefront.cuh:85-87). */
typedef struct ewfa_pair {
    uint64_t pattern_off;  /* word index of the pattern's first packed word */
    uint64_t text_off;     /* word index of the text's first packed word */
    uint64_t cigar_off;    /* byte offset of the stream (multiple of 4) */
    uint32_t plen;         /* pattern length, bases */
    uint32_t tlen;         /* text length, bases */
    uint32_t cigar_cap;    /* stream capacity in bytes (multiple of 4) */
    uint32_t flags;        /* reserved, must be 0 */
} ewfa_pair_t;

/* Result of one pair: the replacement of WF_backtrace_result_t
 * (src/wavefront_structures.h:57-60). */
typedef struct ewfa_result {
    int32_t distance;      /* edit distance (the cap when CAPPED) */
    int32_t cigar_len;     /* stream bytes, excluding the terminator */
    int32_t status;        /* EWFA_PAIR_* */
    int32_t reserved;
} ewfa_result_t;

/* Launch model. */
#define EWFA_MODE_PLAIN       0  /* the reference's per-batch model: one
                                    workgroup per pair, batch by batch */
#define EWFA_MODE_PERSISTENT  1  /* persistent kernel over a device work queue
                                    ordered by length bucket, hipGraph-launched */

typedef struct ewfa_opts {
    int device;            /* HIP device ordinal */
    int mode;              /* EWFA_MODE_* */
    int max_distance;      /* 0: exact for every distance.  >0: the reference's
                              cap (it searches d < cap, src/kernels.cu:259);
                              the reference itself uses 64. */
    int batch_size;        /* PLAIN: pairs per batch (0 = all at once) */
    const char *policy_path; /* autotuner policy file; NULL = built-in */
} ewfa_opts_t;

typedef struct ewfa_ctx ewfa_ctx_t;

/* Fills opts with the defaults (device 0, PERSISTENT, no cap). */
void ewfa_default_opts(ewfa_opts_t *opts);

/* Context: device, streams, scratch pools and policy.  Replaces the Sequences
 * and Cigars constructors (src/wavefront.cuh:458-485, :45-78) and
 * Sequences::GPU_memory_init (src/wavefront.cu:29-108).  One context per host
 * thread; contexts may share a device. */
int ewfa_create(ewfa_ctx_t **ctx, const ewfa_opts_t *opts);

/* Replaces Sequences::GPU_memory_free (src/wavefront.cu:110-125). */
int ewfa_destroy(ewfa_ctx_t *ctx);

/* Message of the calling thread's last failure ("" if none). */
const char *ewfa_last_error(void);

/* Device name and gfx major/minor for the reference's device line
 * (src/aligner.cu:141-151).  Returns EWFA_ERR_NODEV when no device exists. */
int ewfa_device_info(int device, char *name, int name_cap, int *major, int *minor);

/* Number of HIP devices (0 when none). */
int ewfa_device_count(void);

/* Device memory plumbing for callers that hold no HIP runtime of their own
 * (e.g. Python through ctypes).  kind: 0 host->device, 1 device->host,
 * 2 device->device.  All synchronous. */
int ewfa_device_alloc(int device, size_t bytes, void **ptr);
int ewfa_device_free(void *ptr);
int ewfa_memcpy(void *dst, const void *src, size_t bytes, int kind);
int ewfa_device_synchronize(int device);

/* Pinned (page-locked) host memory for fast copies. */
int ewfa_host_alloc(size_t bytes, void **ptr);
int ewfa_host_free(void *ptr);

/* ASCII -> packed words on the device.  Replaces compact_sequences
 * (src/kernels.cu:313-394).  Sequence j is read from d_ascii + j*ascii_stride,
 * d_len[j] bases, and written to d_words + j*words_stride (words_stride >=
 * ceil(max len/32)+1; the tail of each slot is zero-filled). */
int ewfa_pack_ascii(const char *d_ascii, size_t ascii_stride, const uint32_t *d_len,
                    size_t nseq, uint64_t *d_words, size_t words_stride, void *stream);

/* Aligns n device-resident pairs: edit distance and CIGAR stream for each.
 * Replaces WF_edit_distance + generate_cigars[_sh_mem] and their launch
 * sites (src/kernels.cu:182-310, :398-625; src/wavefront.cu:173-260).
 * max_len must bound every plen and tlen.  Results are final on return
 * (the call synchronises `stream` once at the end). */
int ewfa_align_device(ewfa_ctx_t *ctx, const uint64_t *d_words, const ewfa_pair_t *d_pairs,
                      size_t n, uint32_t max_len, ewfa_result_t *d_results, int8_t *d_cigars,
                      void *stream);

/* Allocation-free variant of ewfa_align_device.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and streams as ewfa_align_device on a context in the persistent mode; its
 * kernels are launched eagerly, and all their scratch (op words, lists,
 * bounds, control words, backtrace slots) is carved from a device array of
 * EWFA_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends.  The call allocates nothing; a batch whose scratch
 * does not fit (the op words take ceil((max_len+1)/32)*8 bytes a pair, e.g.
 * 256 B at 1,000 bp) is refused with EWFA_ERR_RANGE.  Calls that overlap on
 * one GPU share its array and race: serialise them. */
#define EWFA_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ewfa_align_device_noalloc(ewfa_ctx_t *ctx, const uint64_t *d_words, const ewfa_pair_t *d_pairs,
                              size_t n, uint32_t max_len, ewfa_result_t *d_results, int8_t *d_cigars,
                              void *stream);

/* Host-buffer pipeline.  Replaces the Sequences batch loop
 * (src/wavefront.cu:173-330, src/aligner.cu:59-66).  Pairs index h_words /
 * h_cigars as in ewfa_align_device.  In the persistent mode the pairs go in
 * chunks: the words a chunk reads go up while the chunk before computes, and
 * only the written CIGAR bytes come back (bytes of a region past its stream
 * are left as they were), after which no words count as resident for
 * ewfa_align_resident.  In the plain mode: all words up (resident
 * afterwards), then batch by batch. */
int ewfa_align_host(ewfa_ctx_t *ctx, const uint64_t *h_words, size_t n_words,
                    const ewfa_pair_t *h_pairs, size_t n, ewfa_result_t *h_results,
                    int8_t *h_cigars, size_t cigar_bytes);

/* Uploads packed words into the context, where they stay resident for
 * ewfa_align_resident calls until the next upload or ewfa_destroy.  Replaces
 * the reference's per-batch sequence upload (src/wavefront.cu:66-74, :265-284). */
int ewfa_upload_words(ewfa_ctx_t *ctx, const uint64_t *h_words, size_t n_words);

/* As ewfa_upload_words, but returns once the copies are queued (an addition;
 * no reference counterpart, the reference uploads each batch's sequences,
 * src/wavefront.cu:265-284).  The words go up in pieces of 16 MiB on a
 * stream of their own, and each ewfa_align_resident call waits only for the
 * pieces that hold the words its pairs read, so the first batches align
 * while later words are still on their way.  h_words must stay unchanged
 * until the last piece has landed: until an ewfa_align_resident call that
 * reads the last words returns, or the context's next upload or
 * ewfa_align_host call, or ewfa_destroy.  From pinned memory
 * (ewfa_host_alloc) the copies are asynchronous; from ordinary memory each
 * piece is copied before the call returns. */
int ewfa_upload_words_async(ewfa_ctx_t *ctx, const uint64_t *h_words, size_t n_words);

/* ewfa_align_host against the words of the last ewfa_upload_words. */
int ewfa_align_resident(ewfa_ctx_t *ctx, const ewfa_pair_t *h_pairs, size_t n,
                        ewfa_result_t *h_results, int8_t *h_cigars, size_t cigar_bytes);

/* Kernel timing of the last ewfa_align_device call on this context, from HIP
 * events recorded on the launch stream: milliseconds in the align kernel(s)
 * and in the CIGAR-emit kernel. */
int ewfa_last_timing(ewfa_ctx_t *ctx, float *align_ms, float *emit_ms);

/* Context flags (ewfa_set_flags).  Default: 0 (hipGraph replay of the
 * persistent path, no per-kernel timing). */
#define EWFA_FLAG_TIMING    1  /* record HIP events around the align and
                                  emit kernels (inside the captured graph on
                                  the persistent path) */
#define EWFA_FLAG_NO_GRAPH  2  /* launch the persistent path eagerly */
#define EWFA_FLAG_SERIAL    4  /* run the persistent plan's length-bucket
                                  passes one after another on the context's
                                  stream (default: spread over three streams,
                                  so one pass's tail runs beside the next) */
int ewfa_set_flags(ewfa_ctx_t *ctx, int flags);

/* Device bytes the context holds now: op scratch, backtrace spill buffer,
 * bucket and deferral lists, control words and the host-pipeline buffers.
 * Buffers grow by at least half again and never wait for the device: one
 * that is replaced is kept (and counted here) until ewfa_destroy, or until an
 * allocation fails, since freeing device memory waits for the whole device. */
int ewfa_context_memory(ewfa_ctx_t *ctx, size_t *bytes);

/* Largest distance the device path resolves for sequences up to max_len:
 * max_len + 1 (every distance) up to about 131,000.  Wavefronts beyond what
 * LDS holds run in deferral passes whose slots live in HBM; their backtrace
 * takes (D+1)^2/4 bytes per pair in flight, which sets that limit. */
int ewfa_max_distance(ewfa_ctx_t *ctx, uint32_t max_len);

/* ---- serving -------------------------------------------------------------- */

/* A job server (an addition; the reference has no serving path) with
 * continuous batching.  A staging thread admits submitted jobs in arrival
 * order and uploads their words at once.  Each slot has its own device
 * context (compute streams) and two batches in flight.  A slot with a free
 * flight takes
 * the queued jobs first-come first-served, as many as fit the device arena
 * together (and at most max_batch_pairs pairs and max_batch_cigar_bytes of
 * CIGAR regions, a single larger job excepted), and enqueues them as ONE
 * persistent launch without waiting; jobs that arrive meanwhile form the next
 * batch, on another slot's stream (where it runs beside the first on the
 * device) or queued behind it.  Only the written CIGAR bytes cross PCIe,
 * straight into pinned host memory, and each job completes as soon as its own
 * results and streams are placed in its buffers.  A job's footprint in the arena is
 * its packed sequences and CIGAR regions (resident while it runs) plus its
 * share of the batch's launch state: descriptors, results and the traceback
 * op scratch for its pairs at the batch's longest sequence.  Every job's
 * results are identical to aligning it alone. */
typedef struct ewfa_server ewfa_server_t;

typedef struct ewfa_server_opts {
    int device;
    int slots;               /* device contexts, each with its own compute
                                stream and two batches in flight; batches of
                                different slots run side by side on the
                                device (0: 2) */
    size_t arena_bytes;      /* device arena for job data and launch state (0: 4 GiB) */
    int max_distance;        /* as ewfa_opts_t */
    const char *policy_path; /* as ewfa_opts_t */
    size_t max_batch_pairs;  /* pairs merged into one launch (0: 393216) */
    size_t max_batch_cigar_bytes; /* their CIGAR regions in bytes (0: 512 MiB);
                                     each slot keeps two pinned host buffers
                                     of this size for the compacted streams */
    int flags;               /* EWFA_FLAG_SERIAL: each batch's length-bucket
                                passes one after another on its slot's stream
                                (default: spread over the slot's three
                                streams, as ewfa_align_device does) */
} ewfa_server_opts_t;

typedef struct ewfa_server_stats {
    uint64_t submitted, completed, failed;
    size_t arena_bytes, arena_peak;   /* capacity and peak use */
    int slots, peak_running;          /* slots and peak jobs on the device */
    uint64_t batches;                 /* launches so far */
    int peak_batch_jobs;              /* most jobs merged into one launch */
    /* time summed over batches (seconds, host clock): staging and enqueueing
     * a batch, waiting for its device work, handing results and streams to
     * its jobs */
    double upload_s, align_s, handout_s;
} ewfa_server_stats_t;

int ewfa_server_create(ewfa_server_t **srv, const ewfa_server_opts_t *opts);
/* Waits for every submitted job, then frees the server. */
int ewfa_server_destroy(ewfa_server_t *srv);

/* Queues a job of n pairs laid out as for ewfa_align_host; the host buffers
 * must stay valid until the job completes.  Returns at once with its id.
 * EWFA_ERR_RANGE when the job can never fit the arena. */
int ewfa_server_submit(ewfa_server_t *srv, const uint64_t *h_words, size_t n_words,
                       const ewfa_pair_t *h_pairs, size_t n, ewfa_result_t *h_results,
                       int8_t *h_cigars, size_t cigar_bytes, uint64_t *job_id);

/* *done = 1 once the job's results are in its host buffers. */
int ewfa_server_poll(ewfa_server_t *srv, uint64_t job_id, int *done);
/* Blocks until the job completes; returns the job's own status. */
int ewfa_server_wait(ewfa_server_t *srv, uint64_t job_id);
/* Seconds from submit to completion of a finished job (host clock). */
int ewfa_server_job_latency(ewfa_server_t *srv, uint64_t job_id, double *seconds);
int ewfa_server_stats(ewfa_server_t *srv, ewfa_server_stats_t *out);

/* ---- autotuner ------------------------------------------------------------ */

/* Times candidate launch configurations for each length bucket on the pairs
 * of a sample (host buffers, laid out as for ewfa_align_host) and writes the
 * fastest per bucket to a policy file for ewfa_opts_t.policy_path.  Bucket b
 * takes pairs with max(plen, tlen) <= bucket_hi[b]; the last takes the rest.
 * Candidates: one wave per pair with the wavefront in registers or LDS, and
 * 4/8/16 waves per pair with the backtrace in HBM, over several wavefront
 * budgets and queue grab sizes.  Every candidate must give byte-identical
 * results and streams (a policy changes speed only): a difference fails the
 * call with EWFA_ERR_STATE.  A text report of every timing goes to report
 * (may be NULL).  An addition; the reference has no tuner. */
int ewfa_autotune(int device, const uint64_t *h_words, size_t n_words, const ewfa_pair_t *h_pairs,
                  size_t n, const uint32_t *bucket_hi, int nb, int reps, const char *out_path,
                  char *report, size_t report_cap);

/* One bucket of a launch policy: pairs with max(plen, tlen) <= len_hi run
 * with `waves` waves per pair, a wavefront budget of dcap levels (pairs
 * beyond it go to deferral passes), the backtrace in LDS (bt_lds 1) or HBM,
 * and `grab` pairs per queue pull. */
typedef struct ewfa_bucket_policy {
    uint32_t len_hi;
    int32_t waves, dcap, bt_lds, grab;
} ewfa_bucket_policy_t;

/* Reads a policy file as ewfa_create would (NULL or "": the built-in table)
 * into out[0..cap) and returns its bucket count, or EWFA_ERR_IO /
 * EWFA_ERR_FORMAT (no valid line, or more than the 8 buckets a plan holds).
 * CPU only; an addition (the reference has no tuner). */
int ewfa_policy_read(const char *path, ewfa_bucket_policy_t *out, int cap);

/* ---- host helpers (CPU, no device needed) ------------------------------- */

/* ASCII -> packed words on the host (same layout as ewfa_pack_ascii). */
void ewfa_pack_host(const char *ascii, size_t len, uint64_t *words);

/* Words a sequence of len bases occupies, pad word included. */
size_t ewfa_packed_words(size_t len);

/* Lays out n pairs for upload: sequences back to back in the packed words
 * (pattern i then text i), CIGAR regions back to back.  cigar_cap > 0 gives
 * every pair that capacity (rounded up to 4); 0 gives each pair room for any
 * stream it can produce (plen + tlen + 4, rounded up to 4).  Fills pairs and
 * the totals. */
int ewfa_layout_pairs(const uint32_t *plen, const uint32_t *tlen, size_t n, uint32_t cigar_cap,
                      ewfa_pair_t *pairs, size_t *n_words, size_t *cigar_bytes);

/* Packs the ASCII of n laid-out pairs (pattern i at pat + i*stride, text i at
 * txt + i*stride) into words, on `threads` host threads. */
int ewfa_pack_pairs_host(const char *pat, const char *txt, size_t stride, const ewfa_pair_t *pairs,
                         size_t n, uint64_t *words, int threads);

/* The reference's printed CIGAR text for one stream
 * (Cigars::generate_ascii_cigar, src/wavefront.cuh:146-184), quirks kept:
 * the number before 'M' is the op byte's decimal value and a trailing match
 * run is not printed.  Reads at most len bytes, stopping at a zero byte.
 * Returns the text length, or EWFA_ERR_RANGE if out_cap is too small. */
int ewfa_cigar_text(const int8_t *stream, size_t len, char *out, size_t out_cap);

/* Standard CIGAR text (runs as "<n>M", ops as "1X"/"1I"/"1D" merged into
 * runs) -- an addition, not in the reference.  Same return convention. */
int ewfa_cigar_text_standard(const int8_t *stream, size_t len, char *out, size_t out_cap);

/* Validity check of a stream against ASCII pattern/text (Cigars::check_cigar,
 * src/wavefront.cuh:297-419): 1 valid, 0 invalid, -1 for a byte the
 * reference treats as fatal. */
int ewfa_check_cigar(const char *pattern, size_t plen, const char *text, size_t tlen,
                     const int8_t *stream, size_t len);

/* ---- pair-file reader ---------------------------------------------------- */

/* Pairs read from a '>'/'<' file (SequenceReader, utils/sequence_reader.cu:108-201)
 * with its quirks: a line's bytes [1, len-1) are kept (a last line without a
 * newline loses a byte, '\r' stays), missing pairs have length 0, reading
 * stops after n pairs.  Sequences are stored packed, ready for upload.  One
 * deliberate difference: the reference copies each line into a fixed slot of
 * 2*l+4-(2*l mod 4) bytes (utils/sequence_reader.h:66-67) and a longer line
 * runs into the next slot; here every sequence keeps its own bytes. */
typedef struct ewfa_pairs_file {
    size_t n;              /* pairs (as requested by -n) */
    size_t n_words;        /* packed words used */
    uint64_t *words;       /* packed sequences (host) */
    ewfa_pair_t *pairs;    /* descriptors; cigar_off/cigar_cap left 0 */
    char *ascii;           /* raw bytes: sequence j at ascii + ascii_off[j]
                              (the file itself, mapped read-only) */
    uint64_t *ascii_off;   /* 2n byte offsets into ascii */
    uint32_t max_len;      /* longest sequence read */
    void *map;             /* the file mapping (owned; ewfa_free_pairs) */
    size_t map_len;
    double scan_s, pack_s; /* seconds spent finding lines and packing */
} ewfa_pairs_file_t;

/* seq_len is the reference's -l (it only gates reading: 0 reads nothing,
 * utils/sequence_reader.cu:109-112).  The file is mapped, its line starts are
 * found and its sequences packed on `threads` host threads.  Returns EWFA_OK,
 * EWFA_ERR_IO, EWFA_ERR_FORMAT (ewfa_last_error() then holds the reference's
 * message, "Invalid sequence format on line <i>." and the line, from
 * utils/sequence_reader.cu:157-158) or EWFA_ERR_NOMEM. */
int ewfa_read_pairs(const char *path, size_t seq_len, size_t n, int threads,
                    ewfa_pairs_file_t *out);
void ewfa_free_pairs(ewfa_pairs_file_t *f);

/* ---- synthetic workloads -------------------------------------------------- */

/* Deterministic pair generator (an addition for tests and benchmarks).
 * kind 0: fixed length len, edits = round(err*len) split equally between
 *         substitution, insertion and deletion (remainder at random);
 * kind 1: length round(exp(U[ln len_lo, ln len_hi])), err ~ U[err_lo, err_hi];
 * kind 2: length U[len_lo, len_hi], err ~ U[err_lo, err_hi], 30% of the edit
 *         budget as insertion/deletion runs of geometric length (mean 5).
 * Pair i depends only on (seed, i).  Writes ASCII into pat/txt (stride
 * `stride` bytes per sequence) and lengths into plen/tlen. */
typedef struct ewfa_synth {
    int kind;
    uint64_t seed;
    uint32_t len, len_lo, len_hi;
    double err, err_lo, err_hi;
} ewfa_synth_t;

/* Upper bound on any generated sequence length for a spec. */
uint32_t ewfa_synth_max_len(const ewfa_synth_t *spec);
int ewfa_synth_pairs(const ewfa_synth_t *spec, size_t first, size_t n, char *pat, char *txt,
                     size_t stride, uint32_t *plen, uint32_t *tlen, int threads);

/* ---- command-line contract ------------------------------------------------ */

/* The wfe.aligner arguments (src/aligner.cu:75-139), parsed like the
 * reference's parse_args (utils/arg_handler.c:88-138). */
typedef struct ewfa_cli_args {
    const char *file;
    int64_t num_alignments;
    int64_t seq_len;
    int64_t batch_size;    /* valid when batch_set */
    int64_t cpu_threads;   /* valid when threads_set */
    int batch_set, threads_set, print_cigars;
    int64_t max_distance;  /* --max-distance (addition); 0 if absent */
    int check;             /* --check (addition) */
    int standard_cigar;    /* --standard-cigar (addition) */
    int plain;             /* --plain (addition): reference launch model */
    const char *policy;    /* --policy <file> (addition): an autotuned launch
                              policy; NULL for the built-in table */
} ewfa_cli_args_t;

/* Returns 1 when every required option was given, 0 otherwise (the caller
 * then prints the usage and exits 1, src/aligner.cu:154-159). */
int ewfa_cli_parse(int argc, char **argv, ewfa_cli_args_t *out);

/* Writes the usage text to stderr (utils/arg_handler.c:140-187). */
void ewfa_cli_usage(void);

#ifdef __cplusplus
}
#endif
#endif /* EWFA_H */
