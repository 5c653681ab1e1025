// kernels.hip -- gfx950 kernels of the WFA edit aligner.
//
//   align_kernel<G, BT_LDS>  wavefront search for one pair per group of G
//                            wave64s: lanes = diagonals, offsets staged in LDS,
//                            2-bit ops per cell recorded as two 64-bit ballots
//                            per (level, 64 diagonals), traceback at the end.
//                            Replaces WF_edit_distance + WF_compute_kernel +
//                            WF_extend_kernel (src/kernels.cu:27-310).
//   emit_kernel              one lane per pair: forward walk over the ops with
//                            re-extension, writes the compressed CIGAR stream.
//                            Replaces generate_cigars[_sh_mem]
//                            (src/kernels.cu:398-625).
//   pack_kernel              ASCII -> 2-bit words.  Replaces compact_sequences
//                            (src/kernels.cu:313-394).
//
// Integer work only; nothing here touches MFMA.
#include "ewfa_internal.h"

namespace ewfa {

// ---------------------------------------------------------------------------
// sequence access

// 32 bases starting at base i of a packed sequence (2 bits per base, first
// base in the high bits).  LDS copies carry a zero pad word, so the second
// word is always readable there; in global memory the read of the next word
// is skipped when it lies past the sequence's data words.
template <bool GUARD, typename Ptr>
__device__ __forceinline__ uint64_t fetch32(Ptr seq, int i, int len) {
    const int w = i >> 5;
    const int s = (i & 31) << 1;
    const uint64_t a = seq[w];
    uint64_t b = 0;
    if (!GUARD || (s != 0 && ((w + 1) << 5) < len)) b = seq[w + 1];
    return s ? ((a << s) | (b >> (64 - s))) : a;
}

// Number of equal bases from pattern position v and text position h, stopping
// at either end (the extend of src/kernels.cu:27-109, 32 bases per XOR+clz).
template <bool GUARD, typename Ptr>
__device__ __forceinline__ int extend(Ptr pat, Ptr txt, int plen, int tlen, int v, int h) {
    int acc = 0;
    for (;;) {
        const int rem = min(plen - v, tlen - h);
        if (rem <= 0) break;
        uint64_t x = fetch32<GUARD>(pat, v, plen) ^ fetch32<GUARD>(txt, h, tlen);
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

template <int G, bool BT_LDS>
__device__ void align_pair(const AlignArgs& a, uint32_t pidx, uint64_t* slot, uint64_t* bt_g,
                           int gtid) {
    const int lane = gtid & 63;
    const int wave = gtid >> 6;
    const ewfa_pair_t pr = a.pairs[pidx];
    const int plen = (int)pr.plen, tlen = (int)pr.tlen;
    ewfa_result_t* res = a.results + pidx;
    const int sw = a.seq_words;
    const int nwp = (plen + 31) >> 5, nwt = (tlen + 31) >> 5;
    if (pr.plen > a.max_len || pr.tlen > a.max_len || nwp + 1 > sw || nwt + 1 > sw) {
        if (gtid == 0) *res = ewfa_result_t{0, 0, EWFA_PAIR_TOO_LONG, 0};
        return;
    }
    const int dcap = a.dcap;
    uint64_t* P = slot;
    uint64_t* T = slot + sw;
    int32_t* X = reinterpret_cast<int32_t*>(slot + 2 * sw) + (dcap + 2);
    int32_t* Y = X + (2 * dcap + 5);
    uint64_t* BT = BT_LDS ? slot + slot_bt_word(sw, dcap) : bt_g;

    // stage the two sequences (coalesced 8-byte loads) plus a zero pad word
    const uint64_t* gp = a.words + pr.pattern_off;
    const uint64_t* gt = a.words + pr.text_off;
    for (int i = gtid; i < nwp; i += 64 * G) P[i] = gp[i];
    for (int i = gtid; i < nwt; i += 64 * G) T[i] = gt[i];
    if (gtid == 0) {
        P[nwp] = 0;
        T[nwt] = 0;
    }
    group_sync<G>();

    const int tk = tlen - plen;  // target diagonal, target offset tlen
    const int tka = tk < 0 ? -tk : tk;
    if (gtid == 0) {
        X[0] = extend<false>(P, T, plen, tlen, 0, 0);
        X[-1] = -1;
        X[1] = -1;
        X[-2] = -1;
        X[2] = -1;
    }
    group_sync<G>();

    const int cap = a.max_distance;
    int limit = dcap;
    if (cap > 0 && cap - 1 < limit) limit = cap - 1;
    bool done = (tk == 0) && (X[0] == tlen);
    int d = 0;
    int32_t* cur = X;
    int32_t* nxt = Y;
    while (!done && d < limit) {
        ++d;
        const int nch = (d >> 5) + 1;
        const uint32_t base = bt_level_base((uint32_t)d);
        for (int c = wave; c < nch; c += G) {
            const int idx = (c << 6) + lane;  // diagonal index 0..2d
            const int k = idx - d;
            int op = 0;
            if (idx <= 2 * d) {
                // candidates from level d-1 (src/kernels.cu:131-138); the max
                // of (value, op) with DEL(3) > SUB(2) > INS(1) on ties
                // (src/kernels.cu:141-152)
                const int del = cur[k + 1];
                const int sub = cur[k] + 1;
                const int ins = cur[k - 1] + 1;
                int val = del;
                op = OP_DEL;
                if (sub > val) {
                    val = sub;
                    op = OP_SUB;
                }
                if (ins > val) {
                    val = ins;
                    op = OP_INS;
                }
                nxt[k] = val + extend<false>(P, T, plen, tlen, val - k, val);
            }
            const uint64_t lo = __ballot(op & 1);
            const uint64_t hi = __ballot(op & 2);
            if (lane == 0) {
                uint64_t* q = BT + 2 * (size_t)(base + (uint32_t)c);
                q[0] = lo;
                q[1] = hi;
            }
        }
        // sentinels for the next level's reads at |k| = d+1, d+2
        if (gtid < 4) {
            const int e = gtid < 2 ? -(d + 1 + gtid) : (d + 1 + (gtid - 2));
            nxt[e] = -1;
        }
        group_sync<G>();
        done = (tka <= d) && (nxt[tk] == tlen);
        int32_t* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }

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

    if (gtid == 0) {
        // traceback: op of (lev, k) selects the predecessor diagonal k+op-2
        uint64_t* ow = a.ops + (size_t)pidx * a.ops_stride;
        if (dist > 0) {
            int wi = (dist - 1) >> 5;
            uint64_t word = 0;
            int k = tk;
            const bool trace = tka <= start;
            for (int lev = start; trace && lev >= 1; --lev) {
                const int s = lev - 1;
                while (wi > (s >> 5)) {
                    ow[wi--] = word;
                    word = 0;
                }
                const int idx = k + lev;
                if (idx < 0 || idx > 2 * lev) break;  // cannot happen for a valid table
                const uint64_t* q = BT + 2 * (size_t)(bt_level_base((uint32_t)lev) + (uint32_t)(idx >> 6));
                const int bit = idx & 63;
                const int op = (int)((q[0] >> bit) & 1ull) | ((int)((q[1] >> bit) & 1ull) << 1);
                word |= (uint64_t)op << ((s & 31) << 1);
                k += op - 2;
            }
            while (wi >= 0) {
                ow[wi--] = word;
                word = 0;
            }
        }
        *res = ewfa_result_t{dist, 0, status, 0};
    }
}

// ---------------------------------------------------------------------------
// align kernel: G waves per pair, one pair slot per block

template <int G, bool BT_LDS>
__global__ __launch_bounds__(64 * G) void align_kernel(AlignArgs a, int persistent) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t s_base;
    const int tid = threadIdx.x;
    uint64_t* btg = BT_LDS ? nullptr : a.bt_global + (size_t)blockIdx.x * a.bt_slot_words;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    if (!persistent) {
        const uint32_t item = blockIdx.x;
        if (item < count) align_pair<G, BT_LDS>(a, a.list ? a.list[item] : item, lds, btg, tid);
        return;
    }
    for (;;) {
        uint32_t base;
        if constexpr (G == 1) {
            uint32_t b = 0;
            if (tid == 0) b = atomicAdd(a.queue, a.grab);
            base = __shfl(b, 0);
        } else {
            if (tid == 0) s_base = atomicAdd(a.queue, a.grab);
            __syncthreads();
            base = s_base;
            __syncthreads();
        }
        if (base >= count) break;
        const uint32_t end = min(base + a.grab, count);
        for (uint32_t it = base; it < end; ++it) {
            align_pair<G, BT_LDS>(a, a.list ? a.list[it] : it, lds, btg, tid);
            group_sync<G>();  // the slot is reused by the next pair
        }
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

__global__ __launch_bounds__(256) void emit_kernel(EmitArgs a) {
    const uint32_t item = blockIdx.x * 256u + threadIdx.x;
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    if (item >= count) return;
    const uint32_t i = a.list ? a.list[item] : item;
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
        const int plen = (int)pr.plen, tlen = (int)pr.tlen;
        const uint64_t* P = a.words + pr.pattern_off;
        const uint64_t* T = a.words + pr.text_off;
        const uint64_t* ow = a.ops + (size_t)i * a.ops_stride;
        int k = 0, off = 0;
        uint64_t word = 0;
        bool stopped = false;
        for (int s = 0; s < r.distance; ++s) {
            const int e = extend<true>(P, T, plen, tlen, off - k, off);
            w.run(e);
            off += e;
            if ((s & 31) == 0) word = ow[s >> 5];
            const int op = (int)((word >> ((s & 31) << 1)) & 3ull);
            if (op == OP_DEL) {
                w.put('D');
                --k;
            } else if (op == OP_SUB) {
                w.put('X');
                ++off;
            } else if (op == OP_INS) {
                w.put('I');
                ++k;
                ++off;
            } else {
                stopped = true;  // op 0 ends the walk (src/kernels.cu:471-478)
                break;
            }
        }
        if (!stopped) w.run(extend<true>(P, T, plen, tlen, off - k, off));
    }
    w.finish();
    a.results[i].cigar_len = (int32_t)w.len;
    if (w.len > pr.cigar_cap) a.results[i].status = EWFA_PAIR_OVERFLOW;
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

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t value, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = value;
}

// ---------------------------------------------------------------------------
// launchers

template <int G, bool BT_LDS>
static hipError_t launch_align_t(const AlignArgs& a, bool persistent, uint32_t grid, size_t lds,
                                 hipStream_t stream) {
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

template <int G, bool BT_LDS>
static const void* align_fn() {
    return reinterpret_cast<const void*>(&align_kernel<G, BT_LDS>);
}

static const void* align_fn_for(int waves, bool bt_lds) {
    switch (waves) {
        case 1: return bt_lds ? align_fn<1, true>() : align_fn<1, false>();
        case 2: return bt_lds ? align_fn<2, true>() : align_fn<2, false>();
        case 4: return bt_lds ? align_fn<4, true>() : align_fn<4, false>();
        case 8: return bt_lds ? align_fn<8, true>() : align_fn<8, false>();
        case 16: return bt_lds ? align_fn<16, true>() : align_fn<16, false>();
        default: return nullptr;
    }
}

hipError_t align_set_max_lds(int waves, size_t lds) {
    for (int b = 0; b < 2; ++b) {
        const void* f = align_fn_for(waves, b != 0);
        if (!f) return hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int align_blocks_per_cu(int waves, size_t lds) {
    const void* f = align_fn_for(waves, true);
    if (!f) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, 64 * waves, lds) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_emit(const EmitArgs& a, uint32_t grid_items, hipStream_t stream) {
    if (grid_items == 0) return hipSuccess;
    hipLaunchKernelGGL(emit_kernel, dim3((grid_items + 255) / 256), dim3(256), 0, stream, a);
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

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p,
                       value, n);
    return hipGetLastError();
}

}  // namespace ewfa
