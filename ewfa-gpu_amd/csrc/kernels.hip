// kernels.hip -- gfx950 kernels of the WFA edit aligner.
//
//   align_reg_kernel<NC>     one wave per pair, the wavefront in registers:
//                            NC = 1 is the banded wavefront (one 64-diagonal
//                            window placed on the target diagonal, passes of
//                            up to 63 levels), NC = 2..4 the full wavefront in
//                            NC chunks.  Ops as two bits per lane and level in
//                            registers (banded) or two 64-bit ballots per level
//                            and chunk in LDS, traceback, and the CIGAR stream
//                            written by the same wave (fused emit).
//   align_kernel<G, BT_LDS>  wavefront search for one pair per group of G
//                            wave64s: lanes = diagonals, offsets staged in LDS,
//                            ops as two 64-bit ballots per (level, 64
//                            diagonals) in LDS or spilled to HBM, traceback at
//                            the end.
//                            Together these replace WF_edit_distance +
//                            WF_compute_kernel + WF_extend_kernel
//                            (src/kernels.cu:27-310).
//   emit_kernel (and the LDS-staged and wave-per-pair variants)
//                            forward walk over the ops with re-extension,
//                            writes the compressed CIGAR stream.  Replaces
//                            generate_cigars[_sh_mem] (src/kernels.cu:398-625).
//   pack_kernel              ASCII -> 2-bit words.  Replaces compact_sequences
//                            (src/kernels.cu:313-394).
//
// Integer work only; nothing here touches MFMA.
#include <mutex>
#include <type_traits>

#include "ewfa_internal.h"

namespace ewfa {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Word 3 of a raw byte buffer resource on gfx9 parts (32-bit data format, so
// untyped accesses are in range; bounds checked against the record count).
constexpr int BUF_DWORD3 = 0x00020000;

// x with lane `lane` set to the uniform value v (v_writelane, whatever the
// exec mask).  The lane select goes through M0: one instruction reads one
// scalar register besides M0.
__device__ __forceinline__ int writelane_i32(int x, int v, int lane) {
    asm("v_writelane_b32 %0, %1, m0" : "+v"(x) : "s"(v), "{m0}"(lane));
    return x;
}

// x with lane L (a constant) set to the uniform value v.
template <int L>
__device__ __forceinline__ int writelane_at(int x, int v) {
    asm("v_writelane_b32 %0, %1, %2" : "+v"(x) : "s"(v), "i"(L));
    return x;
}

// plane with its bits moved up one place and bit 0 set, per lane, where
// a >= b: a compare into VCC, then plane + plane + carry (v_addc), two VALU
// instructions.  After n calls bit i holds the compare of the call n-1-i.
__device__ __forceinline__ uint32_t shift_in_ge(uint32_t plane, int a, int b) {
    uint32_t r;
    asm("v_cmp_ge_i32 vcc, %1, %2\n\t"
        "v_addc_co_u32 %0, vcc, %3, %3, vcc"
        : "=v"(r)
        : "v"(a), "v"(b), "v"(plane)
        : "vcc");
    return r;
}

// Median of a, b and c in one instruction.
__device__ __forceinline__ int med3_i32(int a, int b, int c) {
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// ---------------------------------------------------------------------------
// sequence access

// Sequences are packed 2 bits per base in 64-bit words, first base in the
// high bits.

// ---- LDS copies as 32-bit words ------------------------------------------
// The LDS-staged emit kernel keeps its copies as 32-bit words in base
// order (word q = bases 16q..16q+15, first base in the high bits), written
// by swapping the halves of each 64-bit global word.  Two zero pad words
// (four 32-bit words) follow the data.

// Global 64-bit word -> the two 32-bit LDS words it holds, as one 64-bit store.
__device__ __forceinline__ uint64_t to_lds_order(uint64_t w) { return (w >> 32) | (w << 32); }

// 32 bases from base i (i >= 0 and at most 32 words from the copy's end).
__device__ __forceinline__ uint64_t win32(const uint32_t* S, int i) {
    const int w = i >> 4;
    const uint32_t s = (uint32_t)(i & 15) << 1;
    const uint32_t a = S[w], b = S[w + 1], c = S[w + 2];
    const uint64_t x1 = ((uint64_t)a << 32) | b;
    const uint64_t x2 = ((uint64_t)b << 32) | c;
    return ((x1 << s) & 0xFFFFFFFF00000000ull) | ((x2 << s) >> 32);
}

// Equal leading bases of two 32-base windows, 0..32.
__device__ __forceinline__ int lead_eq(uint64_t x) {
    return __clzll((long long)x) >> 1;  // __clzll(0) = 64: all 32 bases equal
}

// One 32-base step from pattern position v, text position h (clamped into
// the copies): equal bases before the first mismatch, capped at the bases
// left before either end (0..32).  Inactive lanes give 0.
__device__ __forceinline__ int step32w(const uint32_t* P, const uint32_t* T, int plen, int tlen, int v, int h,
                                       bool live) {
    const int rem = live ? max(min(plen - v, tlen - h), 0) : 0;
    const int vc = min(max(v, 0), plen), hc = min(max(h, 0), tlen);
    return min(lead_eq(win32(P, vc) ^ win32(T, hc)), rem);
}

__device__ __forceinline__ int extend_w(const uint32_t* P, const uint32_t* T, int plen, int tlen, int v, int h) {
    int acc = 0;
    for (;;) {
        const int e = step32w(P, T, plen, tlen, v + acc, h + acc, true);
        acc += e;
        if (e < 32) break;
    }
    return acc;
}

// ---- overlapped LDS copies (wave-per-pair emit) ---------------------------
// Word q of an overlapped copy holds 32-bit words q and q+1 of the sequence
// (bases 16q..16q+31, first base in the high bits).  From the global 64-bit
// words G: word 2m = G[m], word 2m+1 = (G[m] << 32) | (G[m+1] >> 32).  A
// copy of a len-base sequence takes 2*ceil(len/32) + 2 words (the last two
// zero), so any window starting at or before len is readable.

// 32 bases from base i: one ds_read2_b64, two 64-bit shifts; the high halves
// of the shifted words are the window.
__device__ __forceinline__ uint64_t win32o(const uint64_t* S, int i) {
    const int w = i >> 4;
    const uint32_t s = (uint32_t)(i & 15) << 1;
    const uint64_t x1 = S[w] << s;
    const uint64_t x2 = S[w + 1] << s;
    return (x1 & 0xFFFFFFFF00000000ull) | (x2 >> 32);
}

// Stages one sequence of len bases from global words g into an overlapped
// copy at S (2*ceil(len/32) + 2 words), cooperatively over 64 lanes.
__device__ __forceinline__ void stage_overlapped(uint64_t* S, const uint64_t* g, int len, int lane) {
    const int nw = (len + 31) >> 5;
    for (int m = lane; m <= nw; m += 64) {
        const uint64_t a = m < nw ? g[m] : 0ull;
        const uint64_t b = m + 1 < nw ? g[m + 1] : 0ull;
        S[2 * m] = a;
        S[2 * m + 1] = (a << 32) | (b >> 32);
    }
}

// A wave's next work item from a counter: lane 0 adds n, every lane gets the
// old value.  The wave barrier joins the lanes' paths first: without it the
// compiler may thread a branch on the lane elsewhere in the caller's loop
// (`if (lane == 0) store`) past the add, into a loop of lanes 1..63 alone
// that reads the counter's value from a lane that never adds.
__device__ __forceinline__ uint32_t wave_take(uint32_t* counter, uint32_t n, int lane) {
    __builtin_amdgcn_wave_barrier();
    uint32_t b = 0;
    if (lane == 0) b = atomicAdd(counter, n);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
}

// Synchronisation of the threads working on one pair.
template <int G>
__device__ __forceinline__ void group_sync() {
    if constexpr (G == 1) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// one pair

// Word offset of the backtrace inside an LDS slot.
__device__ __forceinline__ int slot_bt_word(int seq_words, int dcap) {
    return 2 * seq_words + ((2 * dcap + 5 + 1) & ~1);
}

// End of a pair's search: status, deferral, and the traceback into the ops
// scratch.  Called by every thread of the group with the same arguments; the
// traceback runs on the group's first wave.  For each run of 64 levels, lane
// j loads the backtrace chunks level top-j can need: the path diagonal moves
// by at most one per level, so there it lies within j of the diagonal at level
// top, i.e. in at most three consecutive 64-diagonal chunks.  The walk then
// takes each op bit out of lane j with readlane (scalar work), from LDS or
// HBM alike.
// Inclusive prefix sum over the 64 lanes with DPP only: row shifts by 1, 2,
// 4 and 8 within each row of 16 lanes, then lane 15 and lane 31 broadcast
// into the rows above.  Every lane of the wave must take part.
__device__ __forceinline__ int wave_incl_scan_i32(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
    return x;
}

// Maximum of x over the 64 lanes (every lane takes part; x >= -1), by the
// same DPP steps, read from lane 63.
__device__ __forceinline__ int wave_max_i32(int x) {
    x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x111, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x112, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x114, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x118, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x142, 0xA, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(x, 63);
}

// Stream bytes of a match run of r < 255 bases: floor(r/127) bytes of -127
// then -(r%127) when nonzero (src/kernels.cu:436-446).
__device__ __forceinline__ int short_run_bytes(int r) {
    const int q = (r >= 127) + (r >= 254);
    return q + (r - 127 * q != 0);
}

// Writes a short run and an op byte (op 0: none) at pos, bytes at or past
// cap dropped, as the emit kernel's dword writer leaves them.
__device__ __forceinline__ void put_run_op(int8_t* out, uint32_t cap, uint32_t pos, int r, int op) {
    const int q = (r >= 127) + (r >= 254);
    const int m = r - 127 * q;
    for (int i = 0; i < q; ++i, ++pos)
        if (pos < cap) out[pos] = (int8_t)-127;
    if (m) {
        if (pos < cap) out[pos] = (int8_t)-m;
        ++pos;
    }
    if (op && pos < cap) out[pos] = (int8_t)(op == OP_DEL ? 'D' : op == OP_SUB ? 'X' : 'I');
}

// Offset of (level, flattened diagonal 0..2*level) in a wave's run scratch.
__device__ __forceinline__ uint32_t run_at(int level, int idx) { return (uint32_t)(level * level + idx); }

// The CIGAR stream of one pair from its traced path, written by the wave
// that aligned it: step s (op at level s+1) sits in lane s%64 of round s/64;
// each lane finds its diagonal by a suffix sum of the ops, its match run in
// the run scratch, and its byte offset by a prefix sum, so the stream and its
// terminator come out as the emit kernel writes them (src/kernels.cu:398-498).
// A path with a run of 255 or more bases is left to the emit kernel.
// Run scratch byte of (level st, diagonal k).  Triangular layout: row st at
// st*st, diagonal k at k+st.  Band layout (band >= 0): row st at 64*st,
// diagonal k at lane k+band.  k is clamped into the level's diagonals, and
// the index into the row, so a read never leaves the row.
__device__ __forceinline__ uint32_t run_index(int st, int k, int band) {
    const int kc = min(max(k, -st), st);
    if (band < 0) return run_at(st, kc + st);
    return (uint32_t)(64 * st + min(max(kc + band, 0), 63));
}

// ROUNDS: 2 for paths of up to 128 steps, 1 when the pass has at most 64
// (distance at most 64).
// BAND (the banded pass, ROUNDS 1): the traced ops come in a register (the
// 32-bit half word q of the op words, steps 16q..16q+15, in lane q of opsv)
// rather than from the op scratch, and the stream is assembled in LDS (over
// the first 256 bytes of the run rows, all of whose runs are read by then)
// and written out as whole dwords, the terminator included.  The op scratch
// is written only for a pair left to the emit kernel.
template <int ROUNDS, bool BAND = false>
__device__ __forceinline__ void fused_emit(const AlignArgs& a, uint32_t pidx, const ewfa_pair_t& pr, uint8_t* R,
                                           int dist, int start, int status, int tk, int lane, int band,
                                           bool longrun, int opsv = 0) {
    static_assert(!BAND || ROUNDS == 1, "the LDS-staged stream holds paths of up to 64 steps");
    ewfa_result_t* res = a.results + pidx;
    if ((pr.cigar_off & 3u) || (pr.cigar_cap & 3u)) {
        if (lane == 0) *res = ewfa_result_t{dist, 0, EWFA_PAIR_BADARG, 0};
        return;
    }
    // the run scratch and the op words were written by other lanes of this wave
    __threadfence_block();
    // op of step s (level s+1) in lane s%64 of round s/64; steps up to dist-1
    // lie in the words the traceback wrote
    const uint64_t* ow = a.ops + (size_t)pidx * a.ops_stride;
    int opr0, opr1 = 0;
    if constexpr (BAND) {
        const uint32_t hw = (uint32_t)__builtin_amdgcn_ds_bpermute((lane >> 4) << 2, opsv);
        opr0 = lane < dist ? (int)((hw >> ((lane & 15) << 1)) & 3u) : 0;
    } else {
        opr0 = lane < dist ? (int)((ow[lane >> 5] >> ((lane & 31) << 1)) & 3u) : 0;
        opr1 = ROUNDS > 1 && 64 + lane < dist ? (int)((ow[(64 + lane) >> 5] >> ((lane & 31) << 1)) & 3u) : 0;
    }
    const int nsteps = status == EWFA_PAIR_UNDEFINED ? 0 : dist;
    // step s's diagonal: k_s = tk + (sum of op - 2 over steps s..start-1);
    // steps 64..127 first, their total carried into steps 0..63.  The walk
    // stops at the first op 0 (a capped path): at most one, in the last 64.
    const int x1 = 64 + lane < start ? opr1 - 2 : 0;
    const int x1_incl = ROUNDS > 1 ? wave_incl_scan_i32(x1) : 0;
    const int tot1 = ROUNDS > 1 ? (int)__builtin_amdgcn_readlane(x1_incl, 63) : 0;
    int stop = nsteps;
    {
        const uint64_t z0 = __ballot(lane < nsteps && opr0 == 0);
        const uint64_t z1 = ROUNDS > 1 ? __ballot(64 + lane < nsteps && opr1 == 0) : 0ull;
        if (z0) stop = __ffsll((long long)z0) - 1;
        else if (z1) stop = 64 + __ffsll((long long)z1) - 1;
    }
    // pass 1: runs and byte counts; pass 2 (no run of 255 or more): the bytes.
    // longrun: the scratch may hold a run's low byte only, so the pair goes to
    // the emit kernel
    bool big = longrun;
    int fin = 0;
    if (status != EWFA_PAIR_UNDEFINED && stop == nsteps) {
        fin = R[run_index(dist, tk, band)];
        big = big || fin >= 255;
    }
    int run0 = 0, run1 = 0;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int st = 64 * r + lane;
        const int op = r ? opr1 : opr0;
        const int x = st < start ? op - 2 : 0;
        const int incl = r ? 0 : wave_incl_scan_i32(x);  // the whole wave takes part
        const int k = r ? tk + (tot1 - x1_incl + x1) : tk + tot1 + ((int)__builtin_amdgcn_readlane(incl, 63) - incl + x);
        if (st < nsteps && st <= stop) {
            // the path's diagonal at level st lies in [-st, st] (and in the
            // band); the clamp only keeps the read inside the scratch row
            const int run = R[run_index(st, k, band)];
            big |= run >= 255;
            if (r) run1 = run; else run0 = run;
        }
    }
    if (__ballot(big)) {
        if constexpr (BAND) {
            // the emit kernel reads the ops from the scratch
            uint32_t* ow32 = reinterpret_cast<uint32_t*>(a.ops + (size_t)pidx * a.ops_stride);
            if (dist > 0 && lane <= 2 * ((dist - 1) >> 5) + 1) ow32[lane] = (uint32_t)opsv;
        }
        if (lane == 0) {
            *res = ewfa_result_t{dist, 0, status, 0};
            a.fallback_list[atomicAdd(a.fallback_count, 1u)] = pidx;
        }
        return;
    }
    int8_t* out = a.cigars + pr.cigar_off;
    const uint32_t cap = pr.cigar_cap;
    if constexpr (BAND) {
        // at most 64 steps of at most 3 bytes (a run under 255 bases takes at
        // most 2) and a final run: under 256 bytes
        int8_t* S = reinterpret_cast<int8_t*>(R);
        reinterpret_cast<uint32_t*>(S)[lane] = 0u;
        group_sync<1>();
        const bool act = 0 < nsteps && lane < nsteps && lane <= stop;
        const int nb = act ? short_run_bytes(run0) + (opr0 != 0 ? 1 : 0) : 0;
        const int incl = wave_incl_scan_i32(nb);
        if (nb) put_run_op(S, 256u, (uint32_t)(incl - nb), run0, opr0);
        const uint32_t body = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
        const int fin_bytes = (status != EWFA_PAIR_UNDEFINED && stop == nsteps) ? short_run_bytes(fin) : 0;
        const uint32_t len = body + (uint32_t)fin_bytes;
        if (lane == 0 && fin_bytes) put_run_op(S, 256u, body, fin, 0);
        group_sync<1>();
        // whole dwords up to the terminator's, as the emit kernel writes them
        const uint32_t wbytes = min((len & ~3u) + 4u, cap);
        if ((uint32_t)lane * 4u < wbytes)
            reinterpret_cast<uint32_t*>(out)[lane] = reinterpret_cast<const uint32_t*>(S)[lane];
        if (lane == 0) *res = ewfa_result_t{dist, (int32_t)len, len > cap ? EWFA_PAIR_OVERFLOW : status, 0};
        return;
    }
    uint32_t base = 0;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int st = 64 * r + lane;
        const bool act = st < nsteps && st <= stop;
        const int op = r ? opr1 : opr0;
        const int run = r ? run1 : run0;
        const int nb = act ? short_run_bytes(run) + (op != 0 ? 1 : 0) : 0;
        const int incl = wave_incl_scan_i32(nb);
        if (nb) put_run_op(out, cap, base + (uint32_t)(incl - nb), run, op);
        base += (uint32_t)__builtin_amdgcn_readlane(incl, 63);
    }
    const int fin_bytes = (status != EWFA_PAIR_UNDEFINED && stop == nsteps) ? short_run_bytes(fin) : 0;
    const uint32_t len = base + (uint32_t)fin_bytes;
    if (lane == 0) {
        if (fin_bytes) put_run_op(out, cap, base, fin, 0);
        // terminator: zeros up to the next dword boundary inside the region
        const uint32_t w = (len & ~3u) + 4u < cap ? (len & ~3u) + 4u : cap;
        for (uint32_t q = len; q < w; ++q) out[q] = 0;
        *res = ewfa_result_t{dist, (int32_t)len, len > cap ? EWFA_PAIR_OVERFLOW : status, 0};
    }
}

template <bool FUSED>
__device__ __forceinline__ void finish_pair(const AlignArgs& a, uint32_t pidx, const uint64_t* BT, int gtid,
                                            bool done, int d, int limit, int tk, const ewfa_pair_t* prp = nullptr,
                                            uint8_t* R = nullptr, int band = -1, bool longrun = false) {
    ewfa_result_t* res = a.results + pidx;
    const int tka = tk < 0 ? -tk : tk;
    const int cap = a.max_distance;
    int status = EWFA_PAIR_OK;
    int dist = d, start = d;
    if (!done) {
        if (cap > 0 && limit == cap - 1) {
            // the reference's cap: distance = cap, ops of level cap-1 at the
            // target diagonal (src/kernels.cu:259, :288-303)
            status = tka > cap ? EWFA_PAIR_UNDEFINED : EWFA_PAIR_CAPPED;
            dist = cap;
            start = cap - 1;
        } else {
            if (gtid == 0) {
                const uint32_t at = atomicAdd(a.defer_count, 1u);
                a.defer_list[at] = pidx;
                res->status = PAIR_DEFERRED;
            }
            return;
        }
    }
    if (gtid >= 64) return;
    const int lane = gtid;
    uint64_t* ow = a.ops + (size_t)pidx * a.ops_stride;
    if (dist > 0) {
        int wi = (dist - 1) >> 5;
        uint64_t word = 0;
        int k = tk;
        const int last = tka <= start ? 1 : start + 1;  // no walk when |t| > start
        for (int top = start; top >= last; top -= 64) {
            const int lev_l = top - lane;
            const int c0 = max(0, (k + top - 2 * lane) >> 6);
            uint64_t r[3][2];
#pragma unroll
            for (int u = 0; u < 3; ++u) r[u][0] = r[u][1] = 0;
            if (lev_l >= 1) {
                const uint64_t* row = BT + 2 * (size_t)bt_level_base((uint32_t)lev_l);
                const int nch = (lev_l >> 5) + 1;
#pragma unroll
                for (int u = 0; u < 3; ++u)
                    if (c0 + u < nch) {
                        r[u][0] = row[2 * (c0 + u)];
                        r[u][1] = row[2 * (c0 + u) + 1];
                    }
            }
            const int steps = min(64, top - last + 1);
            for (int j = 0; j < steps; ++j) {
                const int lev = top - j;
                const int s = lev - 1;
                while (wi > (s >> 5)) {
                    if (lane == 0) ow[wi] = word;
                    --wi;
                    word = 0;
                }
                const int idx = k + lev;
                const int rel = (idx >> 6) - __builtin_amdgcn_readlane(c0, j);
                const int bit = idx & 63;
                uint32_t wl = 0, wh = 0;
#pragma unroll
                for (int u = 0; u < 3; ++u)
                    if (rel == u) {
                        if (bit >= 32) {
                            wl = __builtin_amdgcn_readlane((int)(uint32_t)(r[u][0] >> 32), j);
                            wh = __builtin_amdgcn_readlane((int)(uint32_t)(r[u][1] >> 32), j);
                        } else {
                            wl = __builtin_amdgcn_readlane((int)(uint32_t)r[u][0], j);
                            wh = __builtin_amdgcn_readlane((int)(uint32_t)r[u][1], j);
                        }
                    }
                const int b = bit & 31;
                // DEL when its DEL-won bit is set, else SUB when its SUB bit
                // is (SUB won, or SUB at least INS), else INS
                const int op = ((wl >> b) & 1u) ? OP_DEL : ((wh >> b) & 1u) ? OP_SUB : OP_INS;
                word |= (uint64_t)op << ((s & 31) << 1);
                k += op - 2;
            }
        }
        while (wi >= 0) {
            if (lane == 0) ow[wi] = word;
            --wi;
            word = 0;
        }
    }
    if constexpr (FUSED) {
        fused_emit<2>(a, pidx, *prp, R, dist, start, status, tk, lane, band, longrun);
        return;
    }
    if (lane == 0) *res = ewfa_result_t{dist, 0, status, 0};
}

// Little-endian 32-bit copies for the block kernel: word j holds bases
// 16j..16j+15, base 16j+r in bits 2r+1..2r with its two bits swapped (the
// bit reversal of the packed big-endian 64-bit word is two such words, low
// first).  Only base equality is asked, which the swap keeps.  Two zero
// 64-bit words follow the data.

// Equal leading bases (0..32) of the 32-base windows at pattern position v
// and text position h: three words and two funnel shifts per sequence.
__device__ __forceinline__ int eq_le3(const uint32_t* P, const uint32_t* T, int v, int h) {
    const uint32_t* p = P + (v >> 4);
    const uint32_t* t = T + (h >> 4);
    const uint32_t sv = (uint32_t)v << 1, sh = (uint32_t)h << 1;
    const uint32_t xl = __builtin_amdgcn_alignbit(p[1], p[0], sv) ^ __builtin_amdgcn_alignbit(t[1], t[0], sh);
    const uint32_t xh = __builtin_amdgcn_alignbit(p[2], p[1], sv) ^ __builtin_amdgcn_alignbit(t[2], t[1], sh);
    const uint64_t x = ((uint64_t)xh << 32) | xl;
    const uint32_t z = x ? (uint32_t)__builtin_ctzll(x) : 64u;
    return (int)(z >> 1);
}

// The same in doubled units: 2v and 2h in, twice the equal bases out.  Twice
// a position is both its words' funnel shift (low five bits) and, shifted
// right by five, its first word.
__device__ __forceinline__ int eq2_le3(const uint32_t* P, const uint32_t* T, int v2, int h2) {
    const uint32_t* p = P + (v2 >> 5);
    const uint32_t* t = T + (h2 >> 5);
    const uint32_t sv = (uint32_t)v2, sh = (uint32_t)h2;
    const uint32_t xl = __builtin_amdgcn_alignbit(p[1], p[0], sv) ^ __builtin_amdgcn_alignbit(t[1], t[0], sh);
    const uint32_t xh = __builtin_amdgcn_alignbit(p[2], p[1], sv) ^ __builtin_amdgcn_alignbit(t[2], t[1], sh);
    const uint64_t x = ((uint64_t)xh << 32) | xl;
    const uint32_t z = x ? (uint32_t)__builtin_ctzll(x) : 64u;
    return (int)(z & ~1u);  // base b of the window sits in bits 2b, 2b+1
}

// The same over a 16-base window (two words per sequence): twice the equal
// leading bases (0..30) when the windows differ, and 62 when they match (the
// first set bit of zero is all ones); callers clamp it.
__device__ __forceinline__ int eq2_lez(const uint32_t* P, const uint32_t* T, int v2, int h2) {
    const uint32_t* p = P + (v2 >> 5);
    const uint32_t* t = T + (h2 >> 5);
    const uint32_t x = __builtin_amdgcn_alignbit(p[1], p[0], (uint32_t)v2) ^
                       __builtin_amdgcn_alignbit(t[1], t[0], (uint32_t)h2);
    uint32_t z;  // v_ffbl_b32 of zero is 0xffffffff; ctz would let the compiler assume z < 32
    asm("v_ffbl_b32 %0, %1" : "=v"(z) : "v"(x));
    return (int)(z & 62u);
}

// Equal bases from (v, h) up to the first mismatch or either end.  Positions
// past an end extend by nothing; the reads are clamped into the copies.
__device__ __forceinline__ int extend_le3(const uint32_t* P, const uint32_t* T, int plen, int tlen, int v, int h) {
    int acc = max(min(eq_le3(P, T, min(v, plen), min(h, tlen)), min(plen - v, tlen - h)), 0);
    if (acc == 32) {
        int e;
        do {
            e = min(eq_le3(P, T, v + acc, h + acc), min(plen - v - acc, tlen - h - acc));
            acc += e;
        } while (e == 32);
    }
    return acc;
}

// One chunk of one level of the block kernel (chunk c of level d: diagonals
// 64c - d .. 64c + 63 - d), for the lanes inside the indices ilo..ihi.  FULL:
// every lane of the chunk is inside (no range test, lane 0 stores the
// ballots).
// BT_PTR: the op ballots are stored through BT (LDS, or HBM with 64-bit
// addresses), else through the buffer resource BTs with a 32-bit offset.
// kc = 128c - 2d: twice the diagonal 64c + lane - d of the chunk's lanes,
// less 2*lane (uniform; its uses below add it to uniform bases, so that each
// is one per-lane add), and bi = base + c, the chunk's backtrace index; both
// kept by the caller, a scalar add a chunk.
template <bool FULL, bool BT_PTR, bool BOUNDED, typename Off>
__device__ __forceinline__ void block_chunk(int kc, uint32_t bi, int d, int ilo, int ihi, int lane, const Off* curm1,
                                            Off* nxt, const uint32_t* P32, const uint32_t* T32, int plen, int tlen,
                                            uint64_t* BT, __amdgpu_buffer_rsrc_t BTs) {
    bool in;
    if constexpr (BOUNDED) {
        // flattened index 64c + lane (diagonal + d) inside ilo..ihi
        const int t = (kc >> 1) + d + lane - ilo;
        in = FULL || (unsigned)t <= (unsigned)(ihi - ilo);
    } else {
        in = FULL || 2 * lane + kc <= 2 * d;
    }
    if (in) {
        // Candidates from level d-1 (src/kernels.cu:131-138): DEL o(k+1),
        // SUB o(k)+1, INS o(k-1)+1, doubled; the larger wins, DEL on ties,
        // then SUB (src/kernels.cu:141-152)
        const Off* q = (curm1 + (kc >> 1)) + lane;  // diagonals k-1, k, k+1
        const int ins = q[0], sub = q[1], del = q[2];
        const int si = max(sub, ins) + 2;
        const int val = max(del, si);
        // op ballots: DEL won; SUB at least INS (the op is DEL when its
        // DEL bit is set, else SUB when this bit is, else INS)
        const uint64_t lo = __builtin_amdgcn_ballot_w64(del >= si);
        const uint64_t hi = __builtin_amdgcn_ballot_w64(sub >= ins);
        const int v = val - 2 * lane - kc;
        // matched bases, capped at the bases left (none past an end); a
        // valid offset is never left of either start, and a read past an
        // end stays inside the slot.  The first step compares 16 bases
        // (most cells' runs end there): eq2_lez gives 32 for a full match
        // or more (an unmatched window), and a result of 32 or more is
        // taken as exactly 32 matched, continued 32 at a time.
        const int mc = min(2 * plen + 2 * lane + kc, 2 * tlen);
        int acc = med3_i32(eq2_lez(P32, T32, v, val), mc - val, 0);
        // (a uniform test and loop, the continuing lanes as the exec mask:
        // most chunks have no lane to continue)
        uint64_t cm = __builtin_amdgcn_ballot_w64(acc >= 32);
        if (cm) {
            acc = min(acc, 32);
            int e = 64;
            do {
                if (__builtin_amdgcn_inverse_ballot_w64(cm)) {
                    e = min(eq2_le3(P32, T32, v + acc, val + acc), mc - val - acc);
                    acc += e;
                }
                cm &= __builtin_amdgcn_ballot_w64(e == 64);
            } while (cm);
        }
        nxt[(kc >> 1) + lane] = (Off)(val + acc);
        if constexpr (FULL && !BT_PTR) {
            // a whole chunk: lanes 0..3 store the four ballot dwords, and the
            // other lanes' stores fall past the buffer's record count and are
            // dropped, so no branch around a one-lane store is needed
            int w = acc;  // (any register: lanes 4..63 are not stored)
            w = writelane_at<0>(w, (int)(uint32_t)lo);
            w = writelane_at<1>(w, (int)(uint32_t)(lo >> 32));
            w = writelane_at<2>(w, (int)(uint32_t)hi);
            w = writelane_at<3>(w, (int)(uint32_t)(hi >> 32));
            const int vo = lane < 4 ? 4 * lane : 0x7FFFFFF0;
            __builtin_amdgcn_raw_buffer_store_b32(w, BTs, vo, (int)(16u * bi), 0);
            return;
        }
        // the chunk's first lane in the range stores its ballots: lane 0,
        // unless a bounded range starts inside the chunk
        bool first;
        if constexpr (BOUNDED && !FULL) {
            const uint64_t ex = __builtin_amdgcn_read_exec();
            first = __builtin_amdgcn_inverse_ballot_w64(ex & (0ull - ex));
        } else {
            first = lane == 0;
        }
        if (first) {
            if constexpr (BT_PTR) {
                uint64_t* bq = BT + 2 * (size_t)bi;
                bq[0] = lo;
                bq[1] = hi;
            } else {
                const u32x4 w = {(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)};
                __builtin_amdgcn_raw_buffer_store_b128(w, BTs, 0, (int)(16u * bi), 0);
            }
        }
    }
}

// One level of the block kernel: the group's waves take chunks of 64
// flattened diagonals in turn over the indices ilo..ihi.  The chunks wholly
// inside the range go without a range test; only the first (with a bound)
// and the last can be partial.  Without a bound (BOUNDED false) the range is
// the whole level 0..2d, whose every chunk starts with a lane inside it.
template <int G, bool BT_PTR, bool BOUNDED, typename Off>
__device__ __forceinline__ void block_level(int d, int ilo, int ihi, int wave, int lane, const Off* cur,
                                            Off* nxt, const uint32_t* P32, const uint32_t* T32, int plen,
                                            int tlen, uint32_t base, uint64_t* BT, __amdgpu_buffer_rsrc_t BTs) {
    const Off* curm1 = cur - 1;
    const int c0 = BOUNDED ? ilo >> 6 : 0;
    const int cend = ihi >> 6;  // the last chunk, partial unless ihi ends one
    // a chunk c with c0 < c < cend (or c0 itself when ilo starts it) is full
    const int cfull = (BOUNDED && (ilo & 63)) ? c0 + 1 : c0;
    int c = c0 + wave;
    int kc = (c << 7) - 2 * d;
    uint32_t bi = base + (uint32_t)c;
    if (BOUNDED && c < cfull) {
        block_chunk<false, BT_PTR, BOUNDED, Off>(kc, bi, d, ilo, ihi, lane, curm1, nxt, P32, T32, plen, tlen, BT, BTs);
        c += G;
        kc += 128 * G;
        bi += G;
    }
    for (; c < cend; c += G, kc += 128 * G, bi += G)
        block_chunk<true, BT_PTR, BOUNDED, Off>(kc, bi, d, ilo, ihi, lane, curm1, nxt, P32, T32, plen, tlen, BT, BTs);
    if (c == cend)
        block_chunk<false, BT_PTR, BOUNDED, Off>(kc, bi, d, ilo, ihi, lane, curm1, nxt, P32, T32, plen, tlen, BT, BTs);
}

// BOUNDED: the pass reads a distance bound per pair (a.bound) and leaves out
// the cells beyond it; a separate instantiation, so that passes without
// bounds keep the registers (and occupancy) of the plain level loop.
// Off: the type of the LDS offsets, int32_t, or int16_t when twice the
// largest offset a pass can reach (max_len + dcap) fits (half the LDS, more
// pairs per CU).
// GSLOT: `slot` is the block's slot in HBM (sequence copies and offsets),
// the backtrace in HBM stored through 64-bit addresses.
template <int G, bool BT_LDS, bool BOUNDED, typename Off = int32_t, bool GSLOT = false>
__device__ void align_pair(const AlignArgs& a, uint32_t pidx, uint64_t* slot, uint64_t* bt_g,
                           int gtid) {
    constexpr bool BT_PTR = BT_LDS || GSLOT;
    const int lane = gtid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(gtid >> 6);  // uniform in the wave
    const ewfa_pair_t pr = a.pairs[pidx];
    const int plen = (int)pr.plen, tlen = (int)pr.tlen;
    ewfa_result_t* res = a.results + pidx;
    const int sw = a.seq_words;
    const int nwp = (plen + 31) >> 5, nwt = (tlen + 31) >> 5;
    if (pr.plen > a.max_len || pr.tlen > a.max_len || nwp + 2 > sw || nwt + 2 > sw) {
        if (gtid == 0) *res = ewfa_result_t{0, 0, EWFA_PAIR_TOO_LONG, 0};
        return;
    }
    const int dcap = a.dcap;
    const int tk = tlen - plen;  // target diagonal, target offset tlen
    const int tka = tk < 0 ? -tk : tk;
    const int cap = a.max_distance;
    int limit = dcap;
    if (cap > 0 && cap - 1 < limit) limit = cap - 1;
    // With an upper bound U on the distance, cell (d, k) is computed only when
    // d + |k - tk| <= U: a path through any other cell costs more than U to
    // reach the target diagonal, and the cells kept read only cells kept, so
    // the values the search and its traceback use are unchanged.
    const uint32_t ub = BOUNDED ? a.bound[pidx] : BOUND_NONE;
    // Narrow slots (a.band_cap > 0): the search keeps the cells within
    // U = min(bound, band_cap, limit), which is exact for every pair whose
    // distance is at most U (the argument above holds for any U of at least
    // the distance); one whose distance is larger ends the search undone at
    // level U and is deferred.  The cells kept span the diagonals
    // kmin = -floor((U - tk)/2) .. kmax = floor((U + tk)/2), at most U + 1,
    // and an offset generation holds those with two cells either side,
    // diagonal k at index k - kmin + 2 of band_cap + 6.  A pair with
    // |tk| > U is deferred before it is staged.
    const bool narrow = BOUNDED && a.band_cap > 0;
    const int U = narrow ? (int)min(min(ub, (uint32_t)a.band_cap), (uint32_t)limit) : (int)ub;
    const bool pruned = narrow || (BOUNDED && ub <= (uint32_t)limit);
    if (narrow && tka > U) {
        if (gtid == 0) {
            const uint32_t at = atomicAdd(a.defer_count, 1u);
            a.defer_list[at] = pidx;
            res->status = PAIR_DEFERRED;
        }
        return;
    }
    const int kmin = narrow ? -((U - tk) >> 1) : 0;
    uint64_t* P = slot;
    uint64_t* T = slot + sw;
    const uint32_t* P32 = reinterpret_cast<const uint32_t*>(P);
    const uint32_t* T32 = reinterpret_cast<const uint32_t*>(T);
    Off* X = reinterpret_cast<Off*>(slot + 2 * sw) + (narrow ? 2 - kmin : dcap + 2);
    Off* Y = X + (narrow ? a.band_cap + 6 : 2 * dcap + 5);
    uint64_t* BT = BT_LDS ? slot + slot_bt_word(sw, dcap) : bt_g;
    // the HBM backtrace slot as a buffer: stores take their offset in a scalar
    // (slots of passes with 32-bit byte offsets; the HBM-slot passes store
    // through BT)
    const __amdgpu_buffer_rsrc_t BTs = __builtin_amdgcn_make_buffer_rsrc(
        bt_g, (short)0, BT_PTR ? 0 : (int)(a.bt_slot_words * 8u), BUF_DWORD3);

    // stage the two sequences little-endian (coalesced 8-byte loads) plus
    // two zero pad words
    const uint64_t* gp = a.words + pr.pattern_off;
    const uint64_t* gt = a.words + pr.text_off;
    for (int i = gtid; i < nwp + 2; i += 64 * G) P[i] = i < nwp ? __builtin_bitreverse64(gp[i]) : 0ull;
    for (int i = gtid; i < nwt + 2; i += 64 * G) T[i] = i < nwt ? __builtin_bitreverse64(gt[i]) : 0ull;
    // offsets are kept doubled (none: -2), positions too, as eq2_le3 takes
    // them.  An LDS slot's two generations start as none, in 8-byte stores
    // (the slot's LDS is rounded up to 16 bytes): a cell is written only at
    // levels at least |k|, so the cells a level reads past those computed
    // before it (|k| = d+1, d+2 at level d+1) still hold none.  An HBM slot
    // writes those sentinels a level at a time instead.
    if constexpr (!GSLOT) {
        const int cells = narrow ? 2 * (a.band_cap + 6) : 2 * (2 * dcap + 5);
        const int words = (cells * (int)sizeof(Off) + 7) >> 3;
        uint64_t* z = slot + 2 * sw;
        const uint64_t none = sizeof(Off) == 2 ? 0xFFFEFFFEFFFEFFFEull : 0xFFFFFFFEFFFFFFFEull;
        for (int i = gtid; i < words; i += 64 * G) z[i] = none;
    }
    group_sync<G>();

    // (diagonals -2..2 lie inside a narrow slot: kmin <= 0 <= kmax, as
    // U >= |tk|)
    if (gtid == 0) {
        X[0] = 2 * extend_le3(P32, T32, plen, tlen, 0, 0);
        if constexpr (GSLOT) {
            X[-1] = -2;
            X[1] = -2;
            X[-2] = -2;
            X[2] = -2;
        }
    }
    group_sync<G>();

    bool done = (tk == 0) && (X[0] == 2 * tlen);
    int d = 0;
    Off* cur = X;
    Off* nxt = Y;
    uint32_t base = 0;  // bt_level_base(d): level d's first chunk
    // a search with a bound ends by level U (at most limit); the ranges below
    // are not empty up to there
    const int dstop = pruned ? U : limit;
    while (!done && d < dstop) {
        ++d;
        if (BOUNDED && pruned) {
            // flattened diagonal indices (diagonal k at k + d) of the level's
            // cells within the bound
            const int ilo = max(0, tk + d - (U - d));
            const int ihi = min(2 * d, tk + d + (U - d));
            block_level<G, BT_PTR, true, Off>(d, ilo, ihi, wave, lane, cur, nxt, P32, T32, plen, tlen, base, BT, BTs);
        } else {
            block_level<G, BT_PTR, false, Off>(d, 0, 2 * d, wave, lane, cur, nxt, P32, T32, plen, tlen, base, BT, BTs);
        }
        // an HBM slot's sentinels for the next level's reads at |k| = d+1,
        // d+2 (an LDS slot holds none there from the start)
        if constexpr (GSLOT) {
            if (gtid < 4) {
                const int e = gtid < 2 ? -(d + 1 + gtid) : (d + 1 + (gtid - 2));
                nxt[e] = -2;
            }
        }
        group_sync<G>();
        base += (uint32_t)((d >> 5) + 1);
        done = (tka <= d) && (nxt[tk] == 2 * tlen);
        Off* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }

    // (a narrow search that stopped undone below the level limit is deferred:
    // the reference's cap applies only at that limit)
    finish_pair<false>(a, pidx, BT, gtid, done, d, narrow ? U : limit, tk);
}

// ---- register wavefront (one wave per pair) --------------------------------

// Its LDS copies are little-endian: 32-bit word j holds bases 16j..16j+15,
// base 16j+r in bits 2r+1..2r with its two bits swapped (the bit reversal of
// the packed big-endian word).  Only base equality is ever asked of a copy,
// which the swap keeps.  64-bit word q holds words q (low half) and q+1
// (high half); a len-base copy takes 2*ceil(len/32) + 2 words, the last two
// zero.

// Global words g (bases 32m..32m+31) and gn (the next, or 0) -> copy words
// 2m and 2m+1.
__device__ __forceinline__ void le_words(uint64_t g, uint64_t gn, uint64_t& w0, uint64_t& w1) {
    const uint64_t a = __builtin_bitreverse32((uint32_t)(g >> 32));
    const uint64_t b = __builtin_bitreverse32((uint32_t)g);
    const uint64_t c = __builtin_bitreverse32((uint32_t)(gn >> 32));
    w0 = (b << 32) | a;
    w1 = (c << 32) | b;
}

// Equal leading bases (0..32) of the 32-base windows at pattern position v
// and text position h.  A window is one ds_read2_b64 and two funnel shifts,
// which use the low five bits of 2i themselves.
__device__ __forceinline__ int eq_le(const uint64_t* P, const uint64_t* T, int v, int h) {
    const uint64_t* p = P + (v >> 4);
    const uint64_t* t = T + (h >> 4);
    const uint64_t p0 = p[0], p1 = p[1], t0 = t[0], t1 = t[1];
    const uint32_t sv = (uint32_t)v << 1, sh = (uint32_t)h << 1;
    const uint32_t xl = __builtin_amdgcn_alignbit((uint32_t)(p0 >> 32), (uint32_t)p0, sv) ^
                        __builtin_amdgcn_alignbit((uint32_t)(t0 >> 32), (uint32_t)t0, sh);
    const uint32_t xh = __builtin_amdgcn_alignbit((uint32_t)(p1 >> 32), (uint32_t)p1, sv) ^
                        __builtin_amdgcn_alignbit((uint32_t)(t1 >> 32), (uint32_t)t1, sh);
    const uint64_t x = ((uint64_t)xh << 32) | xl;
    const uint32_t z = x ? (uint32_t)__builtin_ctzll(x) : 64u;
    return (int)(z >> 1);
}

// Little-endian copy of a len-base sequence from global words g, cooperatively.
__device__ __forceinline__ void stage_le(uint64_t* S, const uint64_t* g, int len, int lane) {
    const int nw = (len + 31) >> 5;
    for (int m = lane; m <= nw; m += 64) {
        uint64_t w0 = 0, w1 = 0;
        if (m < nw) le_words(g[m], m + 1 < nw ? g[m + 1] : 0ull, w0, w1);
        S[2 * m] = w0;
        S[2 * m + 1] = w1;
    }
}

__device__ __forceinline__ int extend_le(const uint64_t* P, const uint64_t* T, int plen, int tlen, int v, int h) {
    int acc = 0;
    for (;;) {
        const int e = min(eq_le(P, T, v + acc, h + acc), min(plen - v - acc, tlen - h - acc));
        acc += max(e, 0);
        if (e < 32) break;
    }
    return acc;
}

// One level of the register wavefront with NA active chunks.  o holds
// offset + 1 per diagonal (0: none), so the lanes a whole-wave DPP shift
// leaves without a source read 0, i.e. no offset.  Candidates for every
// chunk, the first 32-base compare of every chunk back to back, then a
// continuation loop only while some lane matched all 32 bases.  Writes the
// run of each lane to byte rrow + 64c + lane of the run scratch Rs (lanes
// past the level's 2d+1 diagonals write 0 into rows of later levels, which
// those levels overwrite; the buffer's size drops writes past the wave's
// slot) and chunk c's two op ballots to BTl[2c], BTl[2c+1].  vtop:
// the valid lanes of the top chunk (those below it are full).  nidx[c]:
// -(64c + lane).
template <int NA, int NC>
__device__ __forceinline__ void reg_level(int (&o)[NC], const int (&nidx)[NC], int d, uint64_t vtop, int lane,
                                          const uint64_t* P, const uint64_t* T, int plen, int tlen,
                                          uint64_t* BTl, __amdgpu_buffer_rsrc_t Rs, int rrow) {
    int p62[NA], p63[NA];
#pragma unroll
    for (int c = 0; c + 1 < NA; ++c) {
        p62[c] = __builtin_amdgcn_readlane(o[c], 62);
        p63[c] = __builtin_amdgcn_readlane(o[c], 63);
    }
    int val[NA], acc[NA];
    bool valid[NA];
    uint64_t bl[NA], bh[NA];  // op ballots: DEL won, SUB won (else INS)
#pragma unroll
    for (int c = 0; c < NA; ++c) {
        int s1, s2;
        if (c == 0) {
            s1 = __builtin_amdgcn_update_dpp(0, o[0], 0x138, 0xF, 0xF, true);
            s2 = __builtin_amdgcn_update_dpp(0, s1, 0x138, 0xF, 0xF, true);
        } else {
            s1 = __builtin_amdgcn_update_dpp(p63[c - 1], o[c], 0x138, 0xF, 0xF, false);
            s2 = __builtin_amdgcn_update_dpp(p62[c - 1], s1, 0x138, 0xF, 0xF, false);
        }
        // candidates (src/kernels.cu:131-138) in stored units: DEL o (this
        // lane held k+1), SUB s1+1, INS s2+1; the larger wins, DEL on ties,
        // then SUB (src/kernels.cu:141-152)
        const int si = max(s1, s2) + 1;
        const int st = max(o[c], si);
        // lanes 0..2d-64c hold the chunk's diagonals: all of a chunk below
        // the top one, vtop of the top one
        const uint64_t vm = c + 1 < NA ? ~0ull : vtop;
        bl[c] = __builtin_amdgcn_ballot_w64(o[c] >= si) & vm;
        bh[c] = __builtin_amdgcn_ballot_w64(s1 >= s2) & vm & ~bl[c];
        valid[c] = c + 1 < NA ? true : __builtin_amdgcn_inverse_ballot_w64(vtop);
        val[c] = st - 1;
        acc[c] = 0;
        // a valid offset is never left of either sequence start, and at most
        // d past an end: its reads stay inside the slot, and a position past
        // an end extends by nothing
        if (valid[c]) {
            const int h = val[c], v = h + nidx[c] + d;
            acc[c] = med3_i32(eq_le(P, T, v, h), min(plen - v, tlen - h), 0);
        }
    }
    // lanes whose first step matched all 32 bases continue; their positions
    // are inside both sequences
#pragma unroll
    for (int c = 0; c < NA; ++c) {
        if (__builtin_amdgcn_ballot_w64(acc[c] == 32)) {
            int last = acc[c];
            int h = val[c] + acc[c];
            int v = h + nidx[c] + d;
            do {
                if (last == 32) {
                    last = min(eq_le(P, T, v, h), min(plen - v, tlen - h));
                    h += last;
                    v += last;
                }
            } while (__builtin_amdgcn_ballot_w64(last == 32));
            acc[c] = h - val[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NA; ++c) {
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)min(acc[c], 255), Rs, lane, rrow + (c << 6), 0);
        o[c] = valid[c] ? val[c] + 1 + acc[c] : 0;
    }
#pragma unroll
    for (int c = 0; c < NA; ++c) {
        if (lane == 0) {
            BTl[2 * c] = bl[c];
            BTl[2 * c + 1] = bh[c];
        }
    }
}

// Levels d+1..dend with NA chunks, until the target diagonal holds tlen.
// Returns the last level done, plus REG_HIT when the target was reached
// there.  Level d's ballots start at word 2*bt_level_base(d) of BT (NA
// chunks per level in this range), its runs at byte d*d of R.
constexpr int REG_HIT = 1 << 30;

template <int NA, int NC>
__device__ __forceinline__ int reg_levels(int (&o)[NC], const int (&nidx)[NC], int d, int dend, int lane,
                                          const uint64_t* P, const uint64_t* T, int plen, int tlen, int tk,
                                          uint64_t* BT, __amdgpu_buffer_rsrc_t Rs) {
    uint64_t* BTl = BT + 2 * (size_t)bt_level_base((uint32_t)d + 1);
    // valid lanes of the top chunk at level d+1: 2(d+1)+1-64(NA-1), 1..63
    uint64_t vtop = (1ull << (2 * d + 3 - 64 * (NA - 1))) - 1ull;
    const int tka = tk < 0 ? -tk : tk;
    // one exit: stop drops to -1 when the target is reached
    int stop = dend;
    do {
        ++d;
        reg_level<NA, NC>(o, nidx, d, vtop, lane, P, T, plen, tlen, BTl, Rs, d * d);
        BTl += 2 * NA;
        vtop = (vtop << 2) | 3ull;
        // the target diagonal's offset (flattened index tk + d); before level
        // |tk| the diagonal is outside the wavefront, and no offset equals -7
        const int ti = tk + d;
        int tv = __builtin_amdgcn_readlane(o[0], ti & 63);
#pragma unroll
        for (int c = 1; c < NA; ++c) {
            const int tc = __builtin_amdgcn_readlane(o[c], ti & 63);
            tv = (ti >> 6) == c ? tc : tv;
        }
        const int want = d >= tka ? tlen + 1 : -7;
        stop = tv == want ? -1 : stop;
    } while (d < stop);
    return stop < 0 ? d | REG_HIT : d;
}

// A pair's descriptor and, when both sequences total at most 64 words, its
// words in registers (lane j: pattern word j, then text words), loaded ahead
// of use so the global latency overlaps the previous pair's work.
struct PairPrefetch {
    ewfa_pair_t pr;
    uint64_t word;
    int in_regs;
};

// NT: non-temporal loads, for the multi-chunk kernel, so that the streamed
// words do not displace its waves' run scratch in L2.  The banded kernel keeps
// its run rows in LDS and loads normally: the 128-byte line a pair's words
// share with the next pair's (words are laid out pair after pair, and a wave
// takes consecutive pairs) then stays in L2 for it.
template <bool NT>
__device__ __forceinline__ PairPrefetch prefetch_pair(const AlignArgs& a, uint32_t pidx, int lane) {
    PairPrefetch f;
    f.pr = a.pairs[pidx];
    const int nwp = ((int)f.pr.plen + 31) >> 5, nwt = ((int)f.pr.tlen + 31) >> 5;
    f.in_regs = (nwp + nwt <= 64) ? 1 : 0;
    f.word = 0;
    if (f.in_regs) {
        const uint64_t* g = lane < nwp ? a.words + f.pr.pattern_off + lane : a.words + f.pr.text_off + (lane - nwp);
        if (lane < nwp + nwt) f.word = NT ? __builtin_nontemporal_load(g) : *g;
    }
    return f;
}

// Little-endian copies of both sequences from prefetched registers: lane j's
// next word arrives from lane j+1 by a DPP whole-wave left shift.
__device__ __forceinline__ void stage_from_regs(uint64_t* P, uint64_t* T, const PairPrefetch& f, int lane) {
    const int nwp = ((int)f.pr.plen + 31) >> 5, nwt = ((int)f.pr.tlen + 31) >> 5;
    const uint32_t lo = (uint32_t)f.word, hi = (uint32_t)(f.word >> 32);
    const uint32_t nhi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x130, 0xF, 0xF, false);
    const bool is_p = lane < nwp;
    const int m = is_p ? lane : lane - nwp;
    const int nw = is_p ? nwp : nwt;
    const uint64_t gn = (m + 1 < nw) ? ((uint64_t)nhi << 32) : 0ull;
    uint64_t* S = is_p ? P : T;
    if (lane < nwp + nwt) {
        uint64_t w0, w1;
        le_words(((uint64_t)hi << 32) | lo, gn, w0, w1);
        S[2 * m] = w0;
        S[2 * m + 1] = w1;
    }
    if (lane < 2) {
        P[2 * nwp + lane] = 0;
        T[2 * nwt + lane] = 0;
    }
}

// Plain little-endian copies (64-bit word m = the bit reversal of global word
// m: bases 32m..32m+15 in its low 32-bit half, then 32m+16..32m+31) of both
// sequences from prefetched registers, each followed by two zero words.
__device__ __forceinline__ void stage_plain_from_regs(uint64_t* P, uint64_t* T, const PairPrefetch& f, int lane) {
    const int nwp = ((int)f.pr.plen + 31) >> 5, nwt = ((int)f.pr.tlen + 31) >> 5;
    const uint64_t w = __builtin_bitreverse64(f.word);
    if (lane < nwp) P[lane] = w;
    else if (lane < nwp + nwt) T[lane - nwp] = w;
    if (lane < 2) P[nwp + lane] = 0ull;
    else if (lane < 4) T[nwt + lane - 2] = 0ull;
}

// The same from global words, cooperatively.
__device__ __forceinline__ void stage_plain(uint64_t* S, const uint64_t* g, int len, int lane) {
    const int nw = (len + 31) >> 5;
    for (int i = lane; i < nw + 2; i += 64) S[i] = i < nw ? __builtin_bitreverse64(g[i]) : 0ull;
}

// Op of the diagonal at ballot bit `bit` (0..63, uniform): DEL (3) when its
// DEL-won bit in L is set, else SUB (2) when its SUB-over-INS bit in S is,
// else INS (1).  Two scalar bit tests and selects on the 32-bit half of the
// ballots that holds the bit (the test takes the bit index modulo 32).
__device__ __forceinline__ int ballot_op32(uint32_t L, uint32_t S, int bit) {
    int op;
    asm("s_bitcmp1_b32 %2, %3\n\t"
        "s_cselect_b32 %0, 2, 1\n\t"
        "s_bitcmp1_b32 %1, %3\n\t"
        "s_cselect_b32 %0, 3, %0"
        : "=&s"(op)
        : "s"(L), "s"(S), "s"(bit)
        : "scc");
    return op;
}

// Backtrace and fused emit of a banded pass.  Per lane (band lane j holds
// diagonal j - band), two bit planes per 32 levels: dlo/slo for levels
// 1..m, m = min(d, 32), level l at bit m - l, and dhi/shi for levels 33..d,
// level l at bit d - l.  A D bit says DEL won, an S bit SUB at least INS;
// an op is DEL if its D bit is set, else SUB if its S bit is, else INS.
// Status, distance and deferral as finish_pair.
__device__ __forceinline__ void finish_pair_band(const AlignArgs& a, uint32_t pidx, int lane, bool done, int d,
                                                 int limit, int tk, const ewfa_pair_t& pr, uint8_t* R,
                                                 int band, bool longrun, uint32_t dlo, uint32_t slo,
                                                 uint32_t dhi, uint32_t shi) {
    ewfa_result_t* res = a.results + pidx;
    const int tka = tk < 0 ? -tk : tk;
    const int cap = a.max_distance;
    int status = EWFA_PAIR_OK;
    int dist = d, start = d;
    if (!done) {
        if (cap > 0 && limit == cap - 1) {
            // the reference's cap: distance = cap, ops of level cap-1 at the
            // target diagonal (src/kernels.cu:259, :288-303)
            status = tka > cap ? EWFA_PAIR_UNDEFINED : EWFA_PAIR_CAPPED;
            dist = cap;
            start = cap - 1;
        } else {
            if (lane == 0) {
                const uint32_t at = atomicAdd(a.defer_count, 1u);
                a.defer_list[at] = pidx;
                res->status = PAIR_DEFERRED;
            }
            return;
        }
    }
    // the traced ops: 32-bit half word q (steps 16q..16q+15, step s at bits
    // 2(s%16)) in lane q; half words above the walked ones hold zero ops
    int opsv = 0;
    if (dist > 0 && tka <= start && start >= 1) {
        // steps start-1 down to 0 (op s from level s+1) when the target
        // diagonal is inside the last level, else none; each half word fills
        // from its top step down, shifting in two bits a step
        // the path's lane (kept in a scalar register)
        int pl = __builtin_amdgcn_readfirstlane(tk + band);
        const int m = min(d, 32);
        for (int s = start - 1; s >= 0;) {
            const int lo = s & ~15;
            // steps 32 and up (levels 33 and up) are in the upper planes
            const bool upper = lo >= 32;
            const uint32_t Dv = upper ? dhi : dlo, Sv = upper ? shi : slo;
            int pos = (upper ? d : m) - 1 - s;  // level s+1's bit
            uint32_t hw = 0;
            for (; s >= lo; --s, ++pos) {
                const uint32_t L = (uint32_t)__builtin_amdgcn_readlane((int)Dv, pl);
                const uint32_t S = (uint32_t)__builtin_amdgcn_readlane((int)Sv, pl);
                const int op = ballot_op32(L, S, pos);
                hw = (hw << 2) + (uint32_t)op;
                pl += op - 2;
            }
            opsv = writelane_i32(opsv, (int)hw, lo >> 4);
        }
    }
    fused_emit<1, true>(a, pidx, pr, R, dist, start, status, tk, lane, band, longrun, opsv);  // dist <= 64
}

// Banded register wavefront: one VGPR of 64 diagonals for passes that search
// at most L <= 63 levels.  Only the band of cells that can still reach the
// target diagonal tk within L levels matters: (d, k) with |k - tk| <= L - d.
// A cell (d', k') that feeds a cell (d, k) of the band has |k' - k| <= d - d',
// so |k' - tk| <= L - d': the band is closed under taking sources, and the
// traced path, its runs and the target diagonal's offsets all lie in it.
// Over all levels the band spans the L+1 diagonals -floor((L-tk)/2) ..
// ceil((L+tk)/2), so lane j holds diagonal j - A with A = floor((L - tk)/2)
// throughout: the previous level's k+1, k and k-1 sit in lanes j+1, j and
// j-1.  The lanes of diagonals -d..d are computed as in the full wavefront,
// but for the source outside the wave, read as none; where that diagonal is
// in the wavefront the error moves inward one lane per level and never
// reaches the band (its front at level d is below lane d-A-1 <= lowest band
// lane, and likewise at the top edge), so for a pair that reaches its target
// within L levels the distance, ops and runs along the path are those of the
// full wavefront.  A pair that does not is deferred (or capped) exactly as
// before.  Level d's op bits are shifted into two bit planes of each lane
// (four registers over 63 levels), its runs go to row d of the run rows in
// LDS, lane by lane.
__device__ __forceinline__ void align_pair_band(const AlignArgs& a, uint32_t pidx, const PairPrefetch& f,
                                                uint64_t* slot, int lane) {
    const ewfa_pair_t& pr = f.pr;
    const int plen = (int)pr.plen, tlen = (int)pr.tlen;
    const int sw = a.seq_words;
    const int nwp = (plen + 31) >> 5, nwt = (tlen + 31) >> 5;
    if (pr.plen > a.max_len || pr.tlen > a.max_len || nwp + 2 > sw || nwt + 2 > sw) {
        if (lane == 0) a.results[pidx] = ewfa_result_t{0, 0, EWFA_PAIR_TOO_LONG, 0};
        return;
    }
    // slot: head guard, pattern copy, text copy (plain little-endian words),
    // tail guard, then a 64-byte run row per level
    uint64_t* P = slot + BAND_HEAD_WORDS;
    uint64_t* T = P + sw;
    const uint32_t* P32 = reinterpret_cast<const uint32_t*>(P);
    const uint32_t* T32 = reinterpret_cast<const uint32_t*>(T);
    uint8_t* R = reinterpret_cast<uint8_t*>(T + sw + BAND_TAIL_WORDS);
    if (f.in_regs) {
        stage_plain_from_regs(P, T, f, lane);
    } else {
        stage_plain(P, a.words + pr.pattern_off, plen, lane);
        stage_plain(T, a.words + pr.text_off, tlen, lane);
    }
    group_sync<1>();

    const int tk = tlen - plen;
    const int tka = tk < 0 ? -tk : tk;
    int limit = min(a.dcap, 63);
    const int cap = a.max_distance;
    if (cap > 0 && cap - 1 < limit) limit = cap - 1;
    const int A = (limit - tk) >> 1;  // lane of diagonal 0 (when |tk| <= limit)
    const int o00 = extend_le3(P32, T32, plen, tlen, 0, 0);
    if (lane == 0) R[min(max(A, 0), 63)] = (uint8_t)min(o00, 255);
    bool done = (tk == 0) && (o00 == tlen);
    int d = 0;
    // some level ran a continuation of 7+ steps: a run of 255 bases or more
    // takes 32 + 7 * 32 + more
    bool longrun = false;
    int maxit = 0;  // the most continuation steps of a level (uniform)
    // op bit planes, per lane: levels 1..32 in dlo/slo, 33..63 in dhi/shi
    uint32_t dlo = 0, slo = 0, dhi = 0, shi = 0;
    if (!done && tka <= limit && limit >= 1) {
        // lanes hold twice the offset (-2: none) and positions are doubled
        // throughout, as eq2_le3 takes them
        int o = lane == A ? 2 * o00 : -2;
        const int ca = 2 * (A - lane);                  // 2v = 2h + ca on lane's diagonal
        const int mc = 2 * min(plen - (A - lane), tlen);  // twice the bases left from h: mc - 2h
        const int tl = A + tk;                          // lane of the target diagonal
        const int want = 2 * tlen;  // a lane outside the wavefront holds -2
        // DEL and INS sources: lane j+1's and lane j-1's o.  The shifts leave
        // the lane whose source is outside the wave unwritten, so lane 63's
        // DEL source and lane 0's INS source stay none (-2) throughout
        int sdel = -2, sins = -2;
        uint64_t vmask = 1ull << A;  // lanes of diagonals -d..d, clipped to the window
        int lv = 0;  // this level minus one
        int stop = 0;  // 64 once the target diagonal holds tlen, else lv
        // one level, its op bits shifted into the planes dp and sp
        // wr: this lane's byte in the run row of the next level
        uint8_t* wr = R + 64 + lane;
        const auto level = [&](uint32_t& dp, uint32_t& sp) {
            vmask |= (vmask << 1) | (vmask >> 1);
            const bool valid = __builtin_amdgcn_inverse_ballot_w64(vmask);
            // candidates (src/kernels.cu:131-138): DEL from k+1 (lane j+1),
            // SUB from k (lane j), INS from k-1 (lane j-1): doubled, DEL o(k+1),
            // SUB o(k)+2, INS o(k-1)+2; the larger wins, DEL on ties, then SUB
            // (src/kernels.cu:141-152).  A none (-2) gives at most 0, below
            // every candidate from an offset but a DEL of offset 0, which wins
            // that tie as it would against a none.
            sdel = __builtin_amdgcn_update_dpp(sdel, o, 0x130, 0xF, 0xF, false);
            sins = __builtin_amdgcn_update_dpp(sins, o, 0x138, 0xF, 0xF, false);
            const int si = max(o, sins) + 2;
            const int h = max(sdel, si);  // twice the new offset
            dp = shift_in_ge(dp, sdel, si);  // DEL won
            sp = shift_in_ge(sp, o, sins);   // SUB at least INS
            // twice the matched bases, capped at the bases left (none past an
            // end).  Every lane compares: one outside the wavefront holds
            // offset 0, its reads stay inside the slot's guards, and its
            // result is dropped
            int acc = med3_i32(eq2_le3(P32, T32, h + ca, h), mc - h, 0);
            uint64_t cm = __builtin_amdgcn_ballot_w64(acc == 64) & vmask;
            if (cm) {
                int last = acc;
                int hh = h + acc, vv = hh + ca;
                int it = 0;
                do {
                    if (__builtin_amdgcn_inverse_ballot_w64(cm)) {
                        last = min(eq2_le3(P32, T32, vv, hh), mc - hh);
                        hh += last;
                        vv += last;
                    }
                    ++it;
                    cm &= __builtin_amdgcn_ballot_w64(last == 64);
                } while (cm);
                acc = hh - h;
                maxit = max(maxit, it);
            }
            // the run's low byte in row lv+1 (a pair with a run of 255+ is
            // emitted by the emit kernel, which recomputes the runs)
            *wr = (uint8_t)(acc >> 1);
            wr += 64;
            o = valid ? h + acc : -2;
            ++lv;
            stop = __builtin_amdgcn_readlane(o, tl) == want ? 64 : lv;
        };
        const int lim1 = min(limit, 32);
        do level(dlo, slo);
        while (stop < lim1);
        if (stop < limit) {  // limit <= 63
            do level(dhi, shi);
            while (stop < limit);
        }
        d = lv;
        done = stop == 64;
        longrun = maxit >= 7;
    }
    group_sync<1>();
    // (when |tk| > limit nothing is walked, and row 0 is read where it was
    // written)
    finish_pair_band(a, pidx, lane, done, d, limit, tk, pr, R, min(max(A, 0), 63), longrun, dlo, slo, dhi, shi);
}

// Register-resident wavefront for one wave per pair.  The offsets of a level
// live in NC VGPRs: chunk c, lane j holds diagonal k = 64c + j - d.  The next
// level reads the same lane (k+1), the lane before (k) and the lane two
// before (k-1): two DPP whole-wave right shifts, with the chunk below's
// lanes 62/63 entering lane 0/1 through readlane.  Levels 32(n-1)..32n-1
// take n chunks, each range in a loop of its own.  Levels up to 32*NC-1 fit.
template <int NC>
__device__ __forceinline__ void align_pair_reg(const AlignArgs& a, uint32_t pidx, const PairPrefetch& f, uint64_t* slot, int lane) {
    if constexpr (NC == 1) {
        align_pair_band(a, pidx, f, slot, lane);
        return;
    }
    const ewfa_pair_t& pr = f.pr;
    const int plen = (int)pr.plen, tlen = (int)pr.tlen;
    const int sw = a.seq_words;
    const int nwp = (plen + 31) >> 5, nwt = (tlen + 31) >> 5;
    // sw: words of one sequence copy
    if (pr.plen > a.max_len || pr.tlen > a.max_len || 2 * nwp + 2 > sw || 2 * nwt + 2 > sw) {
        if (lane == 0) a.results[pidx] = ewfa_result_t{0, 0, EWFA_PAIR_TOO_LONG, 0};
        return;
    }
    uint64_t* P = slot;
    uint64_t* T = slot + sw;
    uint64_t* BT = slot + 2 * sw;
    if (f.in_regs) {
        stage_from_regs(P, T, f, lane);
    } else {
        stage_le(P, a.words + pr.pattern_off, plen, lane);
        stage_le(T, a.words + pr.text_off, tlen, lane);
    }
    group_sync<1>();

    const int tk = tlen - plen;
    int o[NC], nidx[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        o[c] = 0;
        nidx[c] = -((c << 6) + lane);
    }
    const int o00 = extend_le(P, T, plen, tlen, 0, 0);
    if (lane == 0) o[0] = o00 + 1;
    // run scratch of this wave (the emit is fused), also as a buffer whose
    // size bounds the level loops' stores to the slot
    uint8_t* R = a.runs + (size_t)blockIdx.x * a.runs_slot;
    const __amdgpu_buffer_rsrc_t Rs = __builtin_amdgcn_make_buffer_rsrc(R, (short)0, (int)a.runs_slot, BUF_DWORD3);
    if (lane == 0) R[0] = (uint8_t)min(o00, 255);

    int limit = min(a.dcap, 32 * NC - 1);
    const int cap = a.max_distance;
    if (cap > 0 && cap - 1 < limit) limit = cap - 1;
    bool done = (tk == 0) && (o00 == tlen);
    int d = 0;
    // one loop per chunk count; each starts where the last stopped
#pragma unroll
    for (int na = 1; na <= NC; ++na) {
        const int dend = min(limit, 32 * na - 1);
        if (done || d >= dend) continue;
        int r;
        if (na == 1) r = reg_levels<1, NC>(o, nidx, d, dend, lane, P, T, plen, tlen, tk, BT, Rs);
        else if (na == 2) r = reg_levels<(NC >= 2 ? 2 : 1), NC>(o, nidx, d, dend, lane, P, T, plen, tlen, tk, BT, Rs);
        else if (na == 3) r = reg_levels<(NC >= 3 ? 3 : 1), NC>(o, nidx, d, dend, lane, P, T, plen, tlen, tk, BT, Rs);
        else r = reg_levels<(NC >= 4 ? 4 : 1), NC>(o, nidx, d, dend, lane, P, T, plen, tlen, tk, BT, Rs);
        d = r & (REG_HIT - 1);
        done = (r & REG_HIT) != 0;
    }
    group_sync<1>();
    finish_pair<true>(a, pidx, BT, lane, done, d, limit, tk, &pr, R);
}

template <int NC>
__global__ __launch_bounds__(64, NC <= 3 ? 8 : 6) void align_reg_kernel(AlignArgs a, int persistent) {
    extern __shared__ uint64_t lds[];
    const int lane = threadIdx.x;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    if (!persistent) {
        const uint32_t item = blockIdx.x;
        if (item < count) {
            const uint32_t pidx = a.list ? a.list[item] : item;
            const PairPrefetch f = prefetch_pair<(NC > 1)>(a, pidx, lane);
            align_pair_reg<NC>(a, pidx, f, lds, lane);
        }
        return;
    }
    // an empty list (a deferral pass with nothing deferred) ends here, not
    // on a queue atomic that every block of the grid would wait for in turn
    if (count == 0) return;
    // items [base, end), each pair's descriptor and words prefetched while
    // the pair before aligns
    const auto run = [&](uint32_t base, uint32_t end) {
        uint32_t pidx = a.list ? a.list[base] : base;
        PairPrefetch cur = prefetch_pair<(NC > 1)>(a, pidx, lane);
        for (uint32_t it = base; it < end; ++it) {
            const bool more = it + 1 < end;
            const uint32_t nidx = more ? (a.list ? a.list[it + 1] : it + 1) : 0u;
            PairPrefetch nxt = cur;
            if (more) nxt = prefetch_pair<(NC > 1)>(a, nidx, lane);  // in flight during this pair
            align_pair_reg<NC>(a, pidx, cur, lds, lane);
            group_sync<1>();
            cur = nxt;
            pidx = nidx;
        }
    };
    if (a.xqueue) {
        // the list in XCDS contiguous shares, share q handed out by counter
        // q: a block takes its own XCD's share first, then helps the others
        // in turn, so neighbouring pairs (whose descriptors and packed words
        // share cache lines) mostly go through one XCD's L2
        const uint32_t home = blockIdx.x % (uint32_t)XCDS;
        for (uint32_t r = 0; r < (uint32_t)XCDS; ++r) {
            const uint32_t q = (home + r) % (uint32_t)XCDS;
            const uint32_t lo = (uint32_t)(((uint64_t)count * q) / XCDS);
            const uint32_t hi = (uint32_t)(((uint64_t)count * (q + 1)) / XCDS);
            for (;;) {
                const uint32_t got = wave_take(a.xqueue + q * XQ_STRIDE, a.grab, lane);
                if (got >= hi - lo) break;
                run(lo + got, min(lo + got + a.grab, hi));
            }
        }
        return;
    }
    for (;;) {
        const uint32_t base = wave_take(a.queue, a.grab, lane);
        if (base >= count) break;
        run(base, min(base + a.grab, count));
    }
}

// ---------------------------------------------------------------------------
// align kernel: G waves per pair, one pair slot per block

// GSLOT: each block's slot (sequence copies, offsets) is its gslot_words of
// a.gslot in HBM instead of LDS: the passes whose wavefront LDS cannot hold.
template <int G, bool BT_LDS, bool BOUNDED = false, bool OFF16 = false, bool GSLOT = false>
__global__ __launch_bounds__(64 * G) void align_kernel(AlignArgs a, int persistent) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t s_base;
    const int tid = threadIdx.x;
    uint64_t* btg = BT_LDS ? nullptr : a.bt_global + (size_t)blockIdx.x * a.bt_slot_words;
    uint64_t* slot = GSLOT ? a.gslot + (size_t)blockIdx.x * a.gslot_words : lds;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    using Off = typename std::conditional<OFF16, int16_t, int32_t>::type;
    if (!persistent) {
        const uint32_t item = blockIdx.x;
        if (item < count) align_pair<G, BT_LDS, BOUNDED, Off, GSLOT>(a, a.list ? a.list[item] : item, slot, btg, tid);
        return;
    }
    if (count == 0) return;  // as in align_reg_kernel
    // `grab` items from a work counter, the value seen by the whole group
    const auto take = [&](uint32_t* counter) -> uint32_t {
        if constexpr (G == 1) {
            return wave_take(counter, a.grab, tid);
        } else {
            if (tid == 0) s_base = atomicAdd(counter, a.grab);
            __syncthreads();
            const uint32_t b = s_base;
            __syncthreads();
            return b;
        }
    };
    const auto run = [&](uint32_t base, uint32_t end) {
        for (uint32_t it = base; it < end; ++it) {
            align_pair<G, BT_LDS, BOUNDED, Off, GSLOT>(a, a.list ? a.list[it] : it, slot, btg, tid);
            group_sync<G>();  // the slot is reused by the next pair
        }
    };
    if (a.xqueue) {  // per-XCD shares, as in align_reg_kernel
        const uint32_t home = blockIdx.x % (uint32_t)XCDS;
        for (uint32_t r = 0; r < (uint32_t)XCDS; ++r) {
            const uint32_t q = (home + r) % (uint32_t)XCDS;
            const uint32_t lo = (uint32_t)(((uint64_t)count * q) / XCDS);
            const uint32_t hi = (uint32_t)(((uint64_t)count * (q + 1)) / XCDS);
            for (;;) {
                const uint32_t got = take(a.xqueue + q * XQ_STRIDE);
                if (got >= hi - lo) break;
                run(lo + got, min(lo + got + a.grab, hi));
            }
        }
        return;
    }
    for (;;) {
        const uint32_t base = take(a.queue);
        if (base >= count) break;
        run(base, min(base + a.grab, count));
    }
}

// ---------------------------------------------------------------------------
// distance bound: the wavefront (src/kernels.cu:111-180) restricted to a band
// of 64 diagonals that follows the pair's furthest-reaching cell.  Every
// offset the band keeps is reached by a path inside the band, so the first
// level at which it holds (plen, tlen) is the cost of a real alignment: an
// upper bound on the distance, which the block kernel uses to leave out
// cells.  The band moves by at most one diagonal a level, towards the cell
// with the most bases consumed (counted up to the ends), which it looks for
// every BAND_LOOK levels and otherwise follows by its own shifts.

constexpr int BAND_NONE_OFF = -(1 << 28);  // doubled offset of a diagonal with none
constexpr int BAND_LOOK = 8;

// A packed sequence read in place, for the bound search: the 32 bases from
// a doubled position x2 (its first bit; first base in the high bits), the
// position clamped into the sequence.  Each sequence of n words is followed
// by a readable pad word (index n), so both words read stay within n + 1,
// also for an empty sequence.  Past either end the caller's count of bases
// left drops what such a window matches.  No branch, so that the loads of
// the pattern's and the text's windows are in flight together.
struct PackedSeq {
    const uint64_t* g;
    int len2;  // twice the length
    int nw;    // words of bases
    __device__ __forceinline__ uint64_t win(int x2) const {
        x2 = min(max(x2, 0), max(len2 - 2, 0));
        const int wi = x2 >> 6, sh = x2 & 63;
        const uint64_t lo = g[wi], hi = g[min(wi + 1, nw)];
        return (lo << sh) | ((hi >> 1) >> (63 - sh));  // a funnel shift, also for sh = 0
    }
};

__device__ __forceinline__ uint32_t band_bound(const AlignArgs& a, uint32_t pidx, int lane) {
    const ewfa_pair_t pr = a.pairs[pidx];
    const int plen = (int)pr.plen, tlen = (int)pr.tlen;
    if (pr.plen > a.max_len || pr.tlen > a.max_len) return BOUND_NONE;
    // the sequences are read in place (through the caches), not staged in
    // LDS, so that the searches of many pairs run on a CU at once
    const PackedSeq gP{a.words + pr.pattern_off, 2 * plen, (plen + 31) >> 5};
    const PackedSeq gT{a.words + pr.text_off, 2 * tlen, (tlen + 31) >> 5};
    // twice the equal leading bases of the 16- and 32-base windows at
    // doubled positions (0..32, 0..64)
    const auto eq16 = [&](int v2, int h2) -> int { return __clz((int)((gP.win(v2) ^ gT.win(h2)) >> 32)) & ~1; };
    const auto eq32 = [&](int v2, int h2) -> int {
        return __clzll((long long)(gP.win(v2) ^ gT.win(h2))) & ~1;
    };
    const int tk = tlen - plen;
    const int tka = tk < 0 ? -tk : tk;
    int o00;
    {
        const int most = 2 * min(plen, tlen);
        int acc2 = 0, e;
        do {
            e = min(eq32(acc2, acc2), most - acc2);
            acc2 += e;
        } while (e == 64);
        o00 = acc2 >> 1;
    }
    if (tk == 0 && o00 == tlen) return 0u;
    int limit = a.dcap;
    const int cap = a.max_distance;
    if (cap > 0 && cap - 1 < limit) limit = cap - 1;
    const int pl2 = 2 * plen, tl2 = 2 * tlen;
    // lane j holds diagonal base + j, its offset doubled
    int base = -32;
    int o = lane == 32 ? 2 * o00 : BAND_NONE_OFF;
    int best = 32;
    for (int d = 1; d <= limit; ++d) {
        const int s = best > 33 ? 1 : best < 31 ? -1 : 0;
        base += s;
        best -= s;
        // diagonals k-1, k, k+1 of level d-1 sit in lanes j+s-1, j+s, j+s+1;
        // lanes shifted in from outside the band read none
        int ins, sub, del;
        if (s == 0) {
            ins = __builtin_amdgcn_update_dpp(BAND_NONE_OFF, o, 0x138, 0xF, 0xF, false);  // wave_shr:1
            sub = o;
            del = __builtin_amdgcn_update_dpp(BAND_NONE_OFF, o, 0x130, 0xF, 0xF, false);  // wave_shl:1
        } else if (s > 0) {
            ins = o;
            sub = __builtin_amdgcn_update_dpp(BAND_NONE_OFF, o, 0x130, 0xF, 0xF, false);
            del = __builtin_amdgcn_update_dpp(BAND_NONE_OFF, sub, 0x130, 0xF, 0xF, false);
        } else {
            del = o;
            sub = __builtin_amdgcn_update_dpp(BAND_NONE_OFF, o, 0x138, 0xF, 0xF, false);
            ins = __builtin_amdgcn_update_dpp(BAND_NONE_OFF, sub, 0x138, 0xF, 0xF, false);
        }
        const int val = max(del, max(sub, ins) + 2);
        const int k2 = 2 * (base + lane);
        int acc = 0;
        if (val >= 0) {
            // a real path's offset is never left of either start; the first
            // read (16 bases, then 32 at a time) is clamped to the ends, past
            // which nothing is matched
            const int v2 = val - k2;
            const int left = min(pl2 - v2, tl2 - val);
            acc = med3_i32(eq16(min(v2, pl2), min(val, tl2)), left, 0);
            if (acc == 32) {
                int e;
                do {
                    e = min(eq32(v2 + acc, val + acc), left - acc);
                    acc += e;
                } while (e == 64);
            }
        }
        o = val >= 0 ? val + acc : BAND_NONE_OFF;
        const int tj = tk - base;
        if (tka <= d && tj >= 0 && tj < 64 && __builtin_amdgcn_readlane(o, tj) == tl2) return (uint32_t)d;
        if (d % BAND_LOOK == 0) {
            const int key = o >= 0 ? ((min(o - k2, pl2) + min(o, tl2)) << 6) | lane : -1;
            best = wave_max_i32(key) & 63;
        }
    }
    return BOUND_NONE;
}

// One wave per pair from a work counter (a.queue) over the list; writes
// a.bound[pair] (BOUND_NONE when the band does not reach the end within the
// pass's level limit).
__global__ __launch_bounds__(64) void bound_kernel(AlignArgs a) {
    const int lane = threadIdx.x;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    if (count == 0) return;  // as in align_reg_kernel
    for (;;) {
        const uint32_t item = wave_take(a.queue, 1u, lane);
        if (item >= count) break;
        const uint32_t pidx = a.list ? a.list[item] : item;
        const uint32_t u = band_bound(a, pidx, lane);
        if (lane == 0) a.bound[pidx] = u;
    }
}

// Block q reorders XCD share q of a list (items [count*q/8, count*(q+1)/8),
// the share the pass hands out from counter q) by decreasing distance bound:
// a counting sort over 64 bins of bound >> shift, no bound first.  A share's
// longest searches then start first and its last ones are short, so the
// pass's tail (the last searches running on few CUs) is short too.

__global__ __launch_bounds__(1024) void order_by_bound_kernel(const uint32_t* list, const uint32_t* count_ptr,
                                                              uint32_t count, const uint32_t* bound, uint32_t* out,
                                                              int shift) {
    __shared__ uint32_t at[ORDER_BINS];
    const uint32_t n = count_ptr ? *count_ptr : count;
    const uint32_t q = blockIdx.x;
    const uint32_t lo = (uint32_t)(((uint64_t)n * q) / XCDS), hi = (uint32_t)(((uint64_t)n * (q + 1)) / XCDS);
    const auto bin_of = [&](uint32_t pidx) -> uint32_t {
        const uint32_t u = bound[pidx];
        return u == BOUND_NONE ? 0u : (uint32_t)(ORDER_BINS - 1) - min(u >> shift, (uint32_t)(ORDER_BINS - 1));
    };
    if (threadIdx.x < ORDER_BINS) at[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&at[bin_of(list ? list[i] : i)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = lo;  // bin b's items start at lo + (items of bins before b)
        for (int b = 0; b < ORDER_BINS; ++b) {
            const uint32_t c = at[b];
            at[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint32_t pidx = list ? list[i] : i;
        out[atomicAdd(&at[bin_of(pidx)], 1u)] = pidx;
    }
}

// ---------------------------------------------------------------------------
// emit kernel: one lane per pair

struct StreamWriter {
    int8_t* out;
    uint32_t cap;
    uint32_t len;
    uint32_t acc;

    __device__ __forceinline__ void put(int b) {
        acc |= (uint32_t)(uint8_t)(int8_t)b << ((len & 3u) << 3);
        ++len;
        if ((len & 3u) == 0) {
            if (len <= cap) *reinterpret_cast<uint32_t*>(out + len - 4) = acc;
            acc = 0;
        }
    }
    // match run: -127 chunks then the remainder (src/kernels.cu:436-446)
    __device__ __forceinline__ void run(int r) {
        while (r >= 127) {
            put(-127);
            r -= 127;
        }
        if (r) put(-r);
    }
    __device__ __forceinline__ void finish() {
        if (len & 3u) {
            const uint32_t at = len & ~3u;
            if (at + 4 <= cap) *reinterpret_cast<uint32_t*>(out + at) = acc;  // holds the terminator
        } else if (len + 4 <= cap) {
            *reinterpret_cast<uint32_t*>(out + len) = 0u;
        }
    }
};

__device__ __forceinline__ void emit_one(const EmitArgs& a, uint32_t i);

// Grid-stride over the items, so a launch sized for the largest count can
// run over a count that is only known on the device.
// One lane per pair, a.per_wave pairs per wave (lanes 0..per_wave-1),
// grid-stride over the items.
__global__ __launch_bounds__(256) void emit_kernel(EmitArgs a) {
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t pw = a.per_wave;
    if (lane >= pw) return;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t stride = gridDim.x * 4u * pw;
    for (uint32_t item = wave * pw + lane; item < count; item += stride)
        emit_one(a, a.list ? a.list[item] : item);
}

// A packed sequence in global memory read through a two-word window: the
// emit walk only moves forward, so each word is loaded about once.  window(i)
// is called for 0 <= i < len only; the word after base i's is read only
// while it holds data.
struct GlobalSeq {
    const uint64_t* g;
    int len, w;
    uint64_t a, b;
    __device__ __forceinline__ GlobalSeq(const uint64_t* g_, int len_) : g(g_), len(len_), w(-2), a(0), b(0) {}
    // 32 bases from base i (first base in the high bits)
    __device__ __forceinline__ uint64_t window(int i) {
        const int wi = i >> 5;
        if (wi != w) {
            a = wi == w + 1 ? b : g[wi];
            b = ((wi + 1) << 5) < len ? g[wi + 1] : 0ull;
            w = wi;
        }
        const int s = (i & 31) << 1;
        return s ? ((a << s) | (b >> (64 - s))) : a;
    }
};

// Equal bases from pattern position v, text position h, stopping at either
// end (the extend of src/kernels.cu:27-109, 32 bases per XOR+clz).
__device__ __forceinline__ int extend_seq(GlobalSeq& P, GlobalSeq& T, int v, int h) {
    int acc = 0;
    for (;;) {
        const int rem = min(P.len - v, T.len - h);
        if (rem <= 0) break;
        uint64_t x = P.window(v) ^ T.window(h);
        if (rem < 32) x |= (~0ull) >> (rem << 1);  // bases past an end never match
        if (x == 0) {
            acc += 32;
            v += 32;
            h += 32;
            continue;
        }
        acc += __clzll(x) >> 1;
        break;
    }
    return acc;
}

__device__ __forceinline__ void emit_one(const EmitArgs& a, uint32_t i) {
    ewfa_result_t r = a.results[i];
    if (r.status != EWFA_PAIR_OK && r.status != EWFA_PAIR_CAPPED && r.status != EWFA_PAIR_UNDEFINED)
        return;
    const ewfa_pair_t pr = a.pairs[i];
    if ((pr.cigar_off & 3u) || (pr.cigar_cap & 3u)) {
        a.results[i].status = EWFA_PAIR_BADARG;
        return;
    }
    StreamWriter w{a.cigars + pr.cigar_off, pr.cigar_cap, 0u, 0u};
    if (r.status != EWFA_PAIR_UNDEFINED) {
        GlobalSeq P(a.words + pr.pattern_off, (int)pr.plen);
        GlobalSeq T(a.words + pr.text_off, (int)pr.tlen);
        const uint64_t* ow = a.ops + (size_t)i * a.ops_stride;
        int k = 0, off = 0;
        uint64_t word = 0;
        bool stopped = false;
        for (int s = 0; s < r.distance; ++s) {
            const int e = extend_seq(P, T, off - k, off);
            w.run(e);
            off += e;
            if ((s & 31) == 0) word = ow[s >> 5];
            const int op = (int)((word >> ((s & 31) << 1)) & 3ull);
            if (op == 0) {
                stopped = true;  // op 0 ends the walk (src/kernels.cu:471-478)
                break;
            }
            // one byte and two selects for the three ops, so that the lanes
            // of a wave, each at its own op, do not split three ways
            w.put(op == OP_DEL ? 'D' : op == OP_SUB ? 'X' : 'I');
            k += (op == OP_INS) - (op == OP_DEL);
            off += op != OP_DEL;
        }
        if (!stopped) w.run(extend_seq(P, T, off - k, off));
    }
    w.finish();
    a.results[i].cigar_len = (int32_t)w.len;
    if (w.len > pr.cigar_cap) a.results[i].status = EWFA_PAIR_OVERFLOW;
}

// Same walk with the sequences of the block's 64 pairs staged in LDS first:
// the wave loads each pair's words with coalesced reads, then every lane
// walks its own pair out of LDS.  sw words per sequence slot.
__global__ __launch_bounds__(64) void emit_lds_kernel(EmitArgs a, int sw) {
    extern __shared__ uint64_t lds[];
    const int lane = threadIdx.x;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    const uint32_t item0 = blockIdx.x * 64u;
    if (item0 >= count) return;
    const uint32_t item = item0 + (uint32_t)lane;
    const bool mine = item < count;
    const uint32_t i = mine ? (a.list ? a.list[item] : item) : 0u;
    ewfa_pair_t pr{};
    ewfa_result_t r{0, 0, -1, 0};
    if (mine) {
        pr = a.pairs[i];
        r = a.results[i];
    }
    const bool want = mine && (r.status == EWFA_PAIR_OK || r.status == EWFA_PAIR_CAPPED ||
                               r.status == EWFA_PAIR_UNDEFINED);
    // stage: lane j loads word j of pair p (pattern words then text words),
    // 16 pairs per round so that 16 loads per lane are in flight
    const uint32_t nvalid = min(64u, count - item0);
    const int nw_lane = ((int)pr.plen + 31) / 32 + ((int)pr.tlen + 31) / 32;
    for (uint32_t p0 = 0; p0 < nvalid; p0 += 16) {
        uint64_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint32_t p = p0 + u;
            v[u] = 0;
            if (p < nvalid) {
                const int pl = __builtin_amdgcn_readlane((int)pr.plen, (int)p);
                const int nwp = (pl + 31) >> 5;
                const int nw = __builtin_amdgcn_readlane(nw_lane, (int)p);
                const uint64_t po = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(pr.pattern_off >> 32), (int)p) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pr.pattern_off, (int)p);
                const uint64_t to = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(pr.text_off >> 32), (int)p) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pr.text_off, (int)p);
                if (lane < nw) v[u] = to_lds_order(a.words[lane < nwp ? po + lane : to + (lane - nwp)]);
            }
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint32_t p = p0 + u;
            if (p < nvalid) {
                const int pl = __builtin_amdgcn_readlane((int)pr.plen, (int)p);
                const int tl = __builtin_amdgcn_readlane((int)pr.tlen, (int)p);
                const int nwp = (pl + 31) >> 5, nwt = (tl + 31) >> 5;
                uint64_t* P = lds + (size_t)p * 2 * sw;
                uint64_t* T = P + sw;
                if (nwp + 2 <= sw && nwt + 2 <= sw) {
                    if (lane < nwp) P[lane] = v[u];
                    else if (lane < nwp + nwt) T[lane - nwp] = v[u];
                    if (lane < 2) {
                        P[nwp + lane] = 0;
                        T[nwt + lane] = 0;
                    }
                    // words past the first 64 of a pair (sequences over 1 kbp)
                    if (nwp + nwt > 64) {
                        const uint64_t po = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(pr.pattern_off >> 32), (int)p) << 32) |
                                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pr.pattern_off, (int)p);
                        const uint64_t to = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(pr.text_off >> 32), (int)p) << 32) |
                                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pr.text_off, (int)p);
                        for (int w = lane + 64; w < nwp + nwt; w += 64) {
                            if (w < nwp) P[w] = to_lds_order(a.words[po + w]);
                            else T[w - nwp] = to_lds_order(a.words[to + (w - nwp)]);
                        }
                    }
                }
            }
        }
    }
    group_sync<1>();
    if (!want) return;
    if ((pr.cigar_off & 3u) || (pr.cigar_cap & 3u)) {
        a.results[i].status = EWFA_PAIR_BADARG;
        return;
    }
    const int plen = (int)pr.plen, tlen = (int)pr.tlen;
    if (((plen + 31) >> 5) + 2 > sw || ((tlen + 31) >> 5) + 2 > sw) {
        a.results[i].status = EWFA_PAIR_TOO_LONG;
        return;
    }
    StreamWriter w{a.cigars + pr.cigar_off, pr.cigar_cap, 0u, 0u};
    if (r.status != EWFA_PAIR_UNDEFINED) {
        const uint32_t* P = reinterpret_cast<const uint32_t*>(lds + (size_t)lane * 2 * sw);
        const uint32_t* T = P + 2 * sw;
        const uint64_t* ow = a.ops + (size_t)i * a.ops_stride;
        int k = 0, off = 0;
        uint64_t word = 0;
        bool stopped = false;
        for (int s = 0; s < r.distance; ++s) {
            const int e = extend_w(P, T, plen, tlen, off - k, off);
            w.run(e);
            off += e;
            if ((s & 31) == 0) word = ow[s >> 5];
            const int op = (int)((word >> ((s & 31) << 1)) & 3ull);
            if (op == 0) {
                stopped = true;  // op 0 ends the walk (src/kernels.cu:471-478)
                break;
            }
            w.put(op == OP_DEL ? 'D' : op == OP_SUB ? 'X' : 'I');  // as the global walk
            k += (op == OP_INS) - (op == OP_DEL);
            off += op != OP_DEL;
        }
        if (!stopped) w.run(extend_w(P, T, plen, tlen, off - k, off));
    }
    w.finish();
    a.results[i].cigar_len = (int32_t)w.len;
    if (w.len > pr.cigar_cap) a.results[i].status = EWFA_PAIR_OVERFLOW;
}

// ---------------------------------------------------------------------------
// length buckets

// Sorts pair indices into per-bucket lists by max(plen, tlen): bucket b takes
// lengths up to hi[b] (the last bucket takes the rest).  A block of 1,024
// pairs counts its members per bucket in LDS (one LDS atomic per bucket per
// wave), then takes its ranges of the lists with one global atomic per bucket
// per block; the order inside a bucket does not matter for the results.
__global__ __launch_bounds__(1024) void bucketize_kernel(const ewfa_pair_t* pairs, uint32_t n, BucketTable t,
                                                         uint32_t* lists, uint32_t* counts) {
    __shared__ uint32_t cnt[MAX_BUCKETS], base[MAX_BUCKETS];
    const uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < (uint32_t)t.nb) cnt[threadIdx.x] = 0u;
    __syncthreads();
    int b = -1;
    if (i < n) {
        const uint32_t L = max(pairs[i].plen, pairs[i].tlen);
        b = t.nb - 1;
        for (int q = t.nb - 2; q >= 0; --q)
            if (L <= t.hi[q]) b = q;
    }
    uint32_t local = 0;  // this pair's place among the block's members of bucket b
    for (int q = 0; q < t.nb; ++q) {
        const uint64_t m = __ballot(b == q);
        if (m == 0) continue;
        const int leader = __ffsll((long long)m) - 1;
        uint32_t at = 0;
        if (lane == leader) at = atomicAdd(cnt + q, (uint32_t)__popcll(m));
        at = (uint32_t)__shfl((int)at, leader);
        if (b == q) local = at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)t.nb) {
        const uint32_t c = cnt[threadIdx.x];
        base[threadIdx.x] = c ? atomicAdd(counts + threadIdx.x, c) : 0u;
    }
    __syncthreads();
    if (b >= 0) lists[(size_t)b * n + base[b] + local] = i;
}

// Match run from pattern position v, text position h with the whole wave:
// lane j compares the 32 bases at 32j of each 2,048-base round; the first lane
// that stops (mismatch or end) gives the run.  Positions past either end give 0.
__device__ __forceinline__ int wave_lcp(const uint64_t* P, const uint64_t* T, int plen, int tlen, int v, int h,
                                        int lane) {
    const int rem = min(plen - v, tlen - h);
    for (int base = 0; base < rem; base += 2048) {
        const int off = base + 32 * lane;
        int e = 32;
        if (off < rem) e = min(lead_eq(win32o(P, v + off) ^ win32o(T, h + off)), rem - off);
        const uint64_t stop = __ballot(off < rem && e < 32);
        if (stop) {
            const int j = __ffsll((long long)stop) - 1;
            return base + 32 * j + __builtin_amdgcn_readlane(e, j);
        }
    }
    return rem > 0 ? rem : 0;
}

// Emit with one wave per pair, for sequences too long to stage 64 pairs in
// LDS: the wave stages its pair's sequences (overlapped copies, sw words
// each), walks the ops with wave_lcp for every match run, and lane 0 writes
// the stream.  Grid-stride over the items.
__global__ __launch_bounds__(64) void emit_wave_kernel(EmitArgs a, int sw) {
    extern __shared__ uint64_t lds[];
    uint64_t* P = lds;
    uint64_t* T = lds + sw;
    const int lane = threadIdx.x;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
        const uint32_t i = a.list ? a.list[item] : item;
        const ewfa_result_t r = a.results[i];
        if (r.status != EWFA_PAIR_OK && r.status != EWFA_PAIR_CAPPED && r.status != EWFA_PAIR_UNDEFINED) continue;
        const ewfa_pair_t pr = a.pairs[i];
        if ((pr.cigar_off & 3u) || (pr.cigar_cap & 3u)) {
            if (lane == 0) a.results[i].status = EWFA_PAIR_BADARG;
            continue;
        }
        const int plen = (int)pr.plen, tlen = (int)pr.tlen;
        if (2 * ((plen + 31) >> 5) + 2 > sw || 2 * ((tlen + 31) >> 5) + 2 > sw) {
            if (lane == 0) a.results[i].status = EWFA_PAIR_TOO_LONG;
            continue;
        }
        group_sync<1>();  // the previous pair's walk is done with the copies
        stage_overlapped(P, a.words + pr.pattern_off, plen, lane);
        stage_overlapped(T, a.words + pr.text_off, tlen, lane);
        group_sync<1>();
        StreamWriter w{a.cigars + pr.cigar_off, pr.cigar_cap, 0u, 0u};  // lane 0's
        if (r.status != EWFA_PAIR_UNDEFINED) {
            const uint64_t* ow = a.ops + (size_t)i * a.ops_stride;
            int k = 0, off = 0;
            uint64_t word = 0;
            bool stopped = false;
            for (int st = 0; st < r.distance; ++st) {
                const int e = wave_lcp(P, T, plen, tlen, off - k, off, lane);
                if (lane == 0) w.run(e);
                off += e;
                if ((st & 31) == 0) word = ow[st >> 5];
                const int op = (int)((word >> ((st & 31) << 1)) & 3ull);
                if (op == OP_DEL) {
                    if (lane == 0) w.put('D');
                    --k;
                } else if (op == OP_SUB) {
                    if (lane == 0) w.put('X');
                    ++off;
                } else if (op == OP_INS) {
                    if (lane == 0) w.put('I');
                    ++k;
                    ++off;
                } else {
                    stopped = true;  // op 0 ends the walk (src/kernels.cu:471-478)
                    break;
                }
            }
            if (!stopped) {
                const int e = wave_lcp(P, T, plen, tlen, off - k, off, lane);
                if (lane == 0) w.run(e);
            }
        }
        if (lane == 0) {
            w.finish();
            a.results[i].cigar_len = (int32_t)w.len;
            if (w.len > pr.cigar_cap) a.results[i].status = EWFA_PAIR_OVERFLOW;
        }
    }
}

// Pairs still deferred after the last pass the device can stage: their
// distance is beyond every wavefront budget that fits in LDS.
__global__ __launch_bounds__(256) void mark_too_distant_kernel(const uint32_t* list, const uint32_t* count_ptr,
                                                               ewfa_result_t* results) {
    const uint32_t count = *count_ptr;
    for (uint32_t item = blockIdx.x * 256u + threadIdx.x; item < count; item += gridDim.x * 256u)
        results[list[item]] = ewfa_result_t{0, 0, EWFA_PAIR_TOO_DISTANT, 0};
}

// ---------------------------------------------------------------------------
// CIGAR compaction: the written bytes of every pair's region, packed densely
// in pair order, so that only they cross PCIe.  Three launches: per-block
// byte sums, one block scanning those sums, and a copy in which each block
// rescans its own pairs and its four waves copy one pair's dwords at a time.

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = __shfl_up(v, (unsigned)d);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan over a 256-thread block; *total gets the block's sum.
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* ws, uint64_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t inc = wave_incl_scan_u64(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint64_t pre = 0;
    for (int q = 0; q < w; ++q) pre += ws[q];
    *total = ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
    return pre + inc - v;
}

__global__ __launch_bounds__(256) void compact_sums_kernel(const ewfa_pair_t* pairs, const ewfa_result_t* results,
                                                           size_t n, uint64_t* bsum) {
    __shared__ uint64_t ws[4];
    const size_t base = (size_t)blockIdx.x * COMPACT_BLOCK + threadIdx.x * 4u;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < n) s += written_cigar_bytes(results[base + k], pairs[base + k]);
    uint64_t total;
    (void)block_excl_scan_u64(s, ws, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// One block: exclusive offsets of nb block sums, 1024 at a time.
__global__ __launch_bounds__(256) void compact_scan_kernel(const uint64_t* bsum, uint64_t* boff, uint32_t nb) {
    __shared__ uint64_t ws[4];
    uint64_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += 1024u) {
        const uint32_t b0 = c0 + threadIdx.x * 4u;
        uint64_t v[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = (b0 + k < nb) ? bsum[b0 + k] : 0u;
            s += v[k];
        }
        uint64_t total;
        uint64_t pre = carry + block_excl_scan_u64(s, ws, &total);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (b0 + k < nb) boff[b0 + k] = pre;
            pre += v[k];
        }
        carry += total;
    }
}

__global__ __launch_bounds__(256) void compact_copy_kernel(const ewfa_pair_t* pairs, const ewfa_result_t* results,
                                                           const int8_t* cigars, size_t n, const uint64_t* boff,
                                                           int8_t* dense) {
    __shared__ uint64_t ws[4];
    __shared__ uint64_t off[COMPACT_BLOCK];
    __shared__ uint32_t len[COMPACT_BLOCK];
    const size_t first = (size_t)blockIdx.x * COMPACT_BLOCK;
    const uint32_t t4 = threadIdx.x * 4u;
    uint32_t l[4];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = first + t4 + k;
        l[k] = i < n ? written_cigar_bytes(results[i], pairs[i]) : 0u;
        s += l[k];
    }
    uint64_t total;
    uint64_t pre = boff[blockIdx.x] + block_excl_scan_u64(s, ws, &total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        off[t4 + k] = pre;
        len[t4 + k] = l[k];
        pre += l[k];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t cnt = (uint32_t)min((size_t)COMPACT_BLOCK, n - first);
    for (uint32_t j = (uint32_t)w; j < cnt; j += 4u) {
        const uint32_t nd = len[j] >> 2;
        if (nd == 0) continue;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(cigars + pairs[first + j].cigar_off);
        uint32_t* dst = reinterpret_cast<uint32_t*>(dense + off[j]);
        for (uint32_t q = (uint32_t)lane; q < nd; q += 64u) dst[q] = src[q];
    }
}

// ---------------------------------------------------------------------------
// pack and fill

__global__ __launch_bounds__(256) void pack_kernel(const char* ascii, size_t ascii_stride,
                                                   const uint32_t* len, size_t nseq, uint64_t* words,
                                                   size_t words_stride) {
    const size_t gid = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (gid >= nseq * words_stride) return;
    const size_t j = gid / words_stride;
    const size_t w = gid - j * words_stride;
    size_t L = len[j];
    if (L > ascii_stride) L = ascii_stride;
    const unsigned char* s = reinterpret_cast<const unsigned char*>(ascii) + j * ascii_stride;
    const size_t b0 = w * 32;
    uint64_t out = 0;
    if (b0 < L) {
        const int n = (int)min((size_t)32, L - b0);
        for (int i = 0; i < n; ++i) out |= (uint64_t)((s[b0 + i] & 6u) >> 1) << (62 - 2 * i);
    }
    words[gid] = out;
}

// Zeroes the words i < 64 of p whose bit i of mask is set: the control words
// a deferral pass starts from, in one launch instead of one per word.
__global__ __launch_bounds__(64) void clear_words_kernel(uint32_t* p, uint64_t mask) {
    const uint32_t i = threadIdx.x;
    if ((mask >> i) & 1u) p[i] = 0u;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t value, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = value;
}

// ---------------------------------------------------------------------------
// launchers

template <int G, bool BT_LDS>
static hipError_t launch_align_t(const AlignArgs& a, bool persistent, uint32_t grid, size_t lds,
                                 hipStream_t stream) {
    if (a.bound && a.off16) return hipErrorInvalidValue;  // the bounded kernel keeps 32-bit offsets
    if (a.band_cap > 0 && (!a.bound || a.off16 || a.gslot || BT_LDS)) return hipErrorInvalidValue;  // narrow slots need bounds
    if (a.gslot) {
        // slots in HBM: GSLOT_WAVES waves, 32-bit offsets, no dynamic LDS
        if (G != GSLOT_WAVES || BT_LDS || a.off16 || lds) return hipErrorInvalidValue;
        if (a.bound)
            hipLaunchKernelGGL((align_kernel<GSLOT_WAVES, false, true, false, true>), dim3(grid),
                               dim3(64 * GSLOT_WAVES), 0, stream, a, persistent ? 1 : 0);
        else
            hipLaunchKernelGGL((align_kernel<GSLOT_WAVES, false, false, false, true>), dim3(grid),
                               dim3(64 * GSLOT_WAVES), 0, stream, a, persistent ? 1 : 0);
        return hipGetLastError();
    }
    if (a.bound && !BT_LDS)
        hipLaunchKernelGGL((align_kernel<G, false, true>), dim3(grid), dim3(64 * G), lds, stream, a,
                           persistent ? 1 : 0);
    else if (a.off16 && !BT_LDS)
        hipLaunchKernelGGL((align_kernel<G, false, false, true>), dim3(grid), dim3(64 * G), lds, stream, a,
                           persistent ? 1 : 0);
    else
        hipLaunchKernelGGL((align_kernel<G, BT_LDS>), dim3(grid), dim3(64 * G), lds, stream, a,
                           persistent ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_align(const AlignArgs& a, int waves, bool persistent, uint32_t grid, size_t lds,
                        hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    const bool b = a.bt_lds != 0;
    switch (waves) {
        case 1: return b ? launch_align_t<1, true>(a, persistent, grid, lds, stream)
                         : launch_align_t<1, false>(a, persistent, grid, lds, stream);
        case 2: return b ? launch_align_t<2, true>(a, persistent, grid, lds, stream)
                         : launch_align_t<2, false>(a, persistent, grid, lds, stream);
        case 4: return b ? launch_align_t<4, true>(a, persistent, grid, lds, stream)
                         : launch_align_t<4, false>(a, persistent, grid, lds, stream);
        case 8: return b ? launch_align_t<8, true>(a, persistent, grid, lds, stream)
                         : launch_align_t<8, false>(a, persistent, grid, lds, stream);
        case 16: return b ? launch_align_t<16, true>(a, persistent, grid, lds, stream)
                          : launch_align_t<16, false>(a, persistent, grid, lds, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_order_by_bound(const uint32_t* list, const uint32_t* count_ptr, uint32_t count,
                                 const uint32_t* bound, uint32_t* out, int shift, hipStream_t stream) {
    hipLaunchKernelGGL(order_by_bound_kernel, dim3(XCDS), dim3(1024), 0, stream, list, count_ptr, count, bound, out,
                       shift);
    return hipGetLastError();
}

hipError_t launch_bound(const AlignArgs& a, uint32_t grid, hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(bound_kernel, dim3(grid), dim3(64), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_align_reg(const AlignArgs& a, int nc, bool persistent, uint32_t grid, size_t lds,
                            hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    const int p = persistent ? 1 : 0;
    switch (nc) {
        case 1: hipLaunchKernelGGL(align_reg_kernel<1>, dim3(grid), dim3(64), lds, stream, a, p); break;
        case 2: hipLaunchKernelGGL(align_reg_kernel<2>, dim3(grid), dim3(64), lds, stream, a, p); break;
        case 3: hipLaunchKernelGGL(align_reg_kernel<3>, dim3(grid), dim3(64), lds, stream, a, p); break;
        case 4: hipLaunchKernelGGL(align_reg_kernel<4>, dim3(grid), dim3(64), lds, stream, a, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int align_reg_blocks_per_cu(int nc, size_t lds) {
    const void* f = nullptr;
    switch (nc) {
        case 1: f = reinterpret_cast<const void*>(&align_reg_kernel<1>); break;
        case 2: f = reinterpret_cast<const void*>(&align_reg_kernel<2>); break;
        case 3: f = reinterpret_cast<const void*>(&align_reg_kernel<3>); break;
        case 4: f = reinterpret_cast<const void*>(&align_reg_kernel<4>); break;
        default: return 0;
    }
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, 64, lds) != hipSuccess) return 0;
    return blocks;
}

template <int G>
static const void* align_fn_variant(int variant) {
    switch (variant) {
        case 0: return reinterpret_cast<const void*>(&align_kernel<G, false>);
        case 1: return reinterpret_cast<const void*>(&align_kernel<G, true>);
        case 2: return reinterpret_cast<const void*>(&align_kernel<G, false, true>);
        case 3: return reinterpret_cast<const void*>(&align_kernel<G, false, false, true>);
        case 4:
            if constexpr (G == GSLOT_WAVES) return reinterpret_cast<const void*>(&align_kernel<G, false, false, false, true>);
            return nullptr;
        case 5:
            if constexpr (G == GSLOT_WAVES) return reinterpret_cast<const void*>(&align_kernel<G, false, true, false, true>);
            return nullptr;
        default: return nullptr;
    }
}

// variant 0: backtrace in HBM, 1: in LDS, 2: in HBM with distance bounds,
// 3: in HBM with 16-bit offsets
static const void* align_fn_for(int waves, int variant) {
    switch (waves) {
        case 1: return align_fn_variant<1>(variant);
        case 2: return align_fn_variant<2>(variant);
        case 4: return align_fn_variant<4>(variant);
        case 8: return align_fn_variant<8>(variant);
        case 16: return align_fn_variant<16>(variant);
        default: return nullptr;
    }
}

// The attribute is per kernel and process-wide, so it is only ever raised
// (under a lock): no thread can lower it below a pass another thread is
// about to launch.
hipError_t align_set_max_lds(size_t lds) {
    static std::mutex m;
    static size_t done = 0;
    std::lock_guard<std::mutex> lk(m);
    if (lds <= done) return hipSuccess;
    for (int waves : {1, 2, 4, 8, 16})
        for (int b = 0; b < 4; ++b) {
            const void* f = align_fn_for(waves, b);
            if (!f) return hipErrorInvalidValue;
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
    done = lds;
    return hipSuccess;
}

int align_blocks_per_cu(int waves, size_t lds, int variant) {
    const void* f = align_fn_for(waves, variant);
    if (!f) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, 64 * waves, lds) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_emit(const EmitArgs& a, uint32_t grid_items, uint32_t per_wave, hipStream_t stream) {
    if (grid_items == 0) return hipSuccess;
    if (per_wave < 1 || per_wave > 64) return hipErrorInvalidValue;
    EmitArgs b = a;
    b.per_wave = per_wave;
    const uint32_t per_block = 4u * per_wave;
    const uint32_t blocks = (grid_items + per_block - 1) / per_block;
    hipLaunchKernelGGL(emit_kernel, dim3(blocks < 16384u ? blocks : 16384u), dim3(256), 0, stream, b);
    return hipGetLastError();
}

hipError_t launch_emit_wave(const EmitArgs& a, uint32_t grid_items, int sw, uint32_t max_grid, hipStream_t stream) {
    if (grid_items == 0) return hipSuccess;
    const size_t lds = (size_t)2 * (size_t)sw * 8;
    const uint32_t blocks = grid_items < max_grid ? grid_items : max_grid;
    hipLaunchKernelGGL(emit_wave_kernel, dim3(blocks), dim3(64), lds, stream, a, sw);
    return hipGetLastError();
}

hipError_t launch_mark_too_distant(const uint32_t* list, const uint32_t* count_ptr, ewfa_result_t* results,
                                   uint32_t max_items, hipStream_t stream) {
    if (max_items == 0) return hipSuccess;
    const uint32_t blocks = (max_items + 255) / 256 < 1024u ? (max_items + 255) / 256 : 1024u;
    hipLaunchKernelGGL(mark_too_distant_kernel, dim3(blocks), dim3(256), 0, stream, list, count_ptr, results);
    return hipGetLastError();
}

hipError_t launch_emit_lds(const EmitArgs& a, uint32_t grid_items, int sw, hipStream_t stream) {
    if (grid_items == 0) return hipSuccess;
    const size_t lds = (size_t)64 * 2 * (size_t)sw * 8;
    hipLaunchKernelGGL(emit_lds_kernel, dim3((grid_items + 63) / 64), dim3(64), lds, stream, a, sw);
    return hipGetLastError();
}

hipError_t launch_pack(const char* ascii, size_t ascii_stride, const uint32_t* len, size_t nseq,
                       uint64_t* words, size_t words_stride, hipStream_t stream) {
    const size_t total = nseq * words_stride;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       ascii, ascii_stride, len, nseq, words, words_stride);
    return hipGetLastError();
}

hipError_t launch_bucketize(const ewfa_pair_t* pairs, uint32_t n, const BucketTable& t, uint32_t* lists,
                            uint32_t* counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(bucketize_kernel, dim3((n + 1023) / 1024), dim3(1024), 0, stream, pairs, n, t, lists, counts);
    return hipGetLastError();
}

hipError_t launch_compact(const ewfa_pair_t* pairs, const ewfa_result_t* results, const int8_t* cigars, size_t n,
                          uint64_t* scan, int8_t* dense, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const size_t nb = (n + COMPACT_BLOCK - 1) / COMPACT_BLOCK;
    if (nb > 0xFFFFFFFFu) return hipErrorInvalidValue;
    uint64_t* bsum = scan;
    uint64_t* boff = scan + nb;
    hipLaunchKernelGGL(compact_sums_kernel, dim3((unsigned)nb), dim3(256), 0, stream, pairs, results, n, bsum);
    hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(256), 0, stream, bsum, boff, (uint32_t)nb);
    hipLaunchKernelGGL(compact_copy_kernel, dim3((unsigned)nb), dim3(256), 0, stream, pairs, results, cigars, n,
                       boff, dense);
    return hipGetLastError();
}

hipError_t launch_clear_words(uint32_t* p, uint64_t mask, hipStream_t stream) {
    if (mask == 0) return hipSuccess;
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, stream, p, mask);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p,
                       value, n);
    return hipGetLastError();
}

}  // namespace ewfa
