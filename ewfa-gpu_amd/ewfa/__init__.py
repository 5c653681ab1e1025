"""Python view of libewfa, the MI355X WFA edit aligner (C ABI in include/ewfa.h).

The library is loaded from ../build/libewfa.so next to this package; there is
no fallback: if the shared library is missing, importing this module raises.
Device buffers are raw device addresses, e.g. DeviceBuffer.ptr (allocated
through the library's own HIP runtime: a second HIP runtime in the process,
such as the one bundled with PyTorch, cannot share the GPU with it).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
BUILD_DIR = os.path.join(PKG_ROOT, "build")
LIB_PATH = os.path.join(BUILD_DIR, "libewfa.so")
ALIGNER_PATH = os.path.join(BUILD_DIR, "wfe.aligner")

OK = 0
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NODEV, ERR_IO, ERR_FORMAT, ERR_RANGE, ERR_STATE = range(-1, -9, -1)
PAIR_OK, PAIR_CAPPED, PAIR_OVERFLOW, PAIR_UNDEFINED, PAIR_BADARG, PAIR_TOO_LONG, PAIR_TOO_DISTANT = range(7)
MODE_PLAIN, MODE_PERSISTENT = 0, 1
FLAG_TIMING, FLAG_NO_GRAPH, FLAG_SERIAL = 1, 2, 4

PAIR_DTYPE = np.dtype([
    ("pattern_off", "<u8"), ("text_off", "<u8"), ("cigar_off", "<u8"),
    ("plen", "<u4"), ("tlen", "<u4"), ("cigar_cap", "<u4"), ("flags", "<u4"),
])
RESULT_DTYPE = np.dtype([("distance", "<i4"), ("cigar_len", "<i4"), ("status", "<i4"), ("reserved", "<i4")])
assert PAIR_DTYPE.itemsize == 40 and RESULT_DTYPE.itemsize == 16


class Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("mode", ctypes.c_int), ("max_distance", ctypes.c_int),
                ("batch_size", ctypes.c_int), ("policy_path", ctypes.c_char_p)]


class Synth(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("seed", ctypes.c_uint64), ("len", ctypes.c_uint32),
                ("len_lo", ctypes.c_uint32), ("len_hi", ctypes.c_uint32), ("err", ctypes.c_double),
                ("err_lo", ctypes.c_double), ("err_hi", ctypes.c_double)]


class CliArgs(ctypes.Structure):
    _fields_ = [("file", ctypes.c_char_p), ("num_alignments", ctypes.c_int64), ("seq_len", ctypes.c_int64),
                ("batch_size", ctypes.c_int64), ("cpu_threads", ctypes.c_int64),
                ("batch_set", ctypes.c_int), ("threads_set", ctypes.c_int), ("print_cigars", ctypes.c_int),
                ("max_distance", ctypes.c_int64), ("check", ctypes.c_int), ("standard_cigar", ctypes.c_int),
                ("plain", ctypes.c_int), ("policy", ctypes.c_char_p)]


class BucketPolicy(ctypes.Structure):
    _fields_ = [("len_hi", ctypes.c_uint32), ("waves", ctypes.c_int32), ("dcap", ctypes.c_int32),
                ("bt_lds", ctypes.c_int32), ("grab", ctypes.c_int32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("arena_bytes", ctypes.c_size_t),
                ("max_distance", ctypes.c_int), ("policy_path", ctypes.c_char_p),
                ("max_batch_pairs", ctypes.c_size_t), ("max_batch_cigar_bytes", ctypes.c_size_t),
                ("flags", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("arena_bytes", ctypes.c_size_t), ("arena_peak", ctypes.c_size_t), ("slots", ctypes.c_int),
                ("peak_running", ctypes.c_int), ("batches", ctypes.c_uint64), ("peak_batch_jobs", ctypes.c_int),
                ("upload_s", ctypes.c_double), ("align_s", ctypes.c_double), ("handout_s", ctypes.c_double)]


class PairsFile(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("n_words", ctypes.c_size_t), ("words", ctypes.POINTER(ctypes.c_uint64)),
                ("pairs", ctypes.c_void_p), ("ascii", ctypes.POINTER(ctypes.c_char)),
                ("ascii_off", ctypes.POINTER(ctypes.c_uint64)), ("max_len", ctypes.c_uint32),
                ("map", ctypes.c_void_p), ("map_len", ctypes.c_size_t), ("scan_s", ctypes.c_double),
                ("pack_s", ctypes.c_double)]


# every symbol include/ewfa.h declares
EXPORTS = [
    "ewfa_default_opts", "ewfa_create", "ewfa_destroy", "ewfa_last_error", "ewfa_device_info",
    "ewfa_device_count", "ewfa_pack_ascii", "ewfa_align_device", "ewfa_align_host", "ewfa_upload_words",
    "ewfa_align_resident", "ewfa_upload_words_async", "ewfa_set_flags", "ewfa_last_timing", "ewfa_max_distance", "ewfa_pack_host",
    "ewfa_packed_words", "ewfa_layout_pairs", "ewfa_pack_pairs_host", "ewfa_cigar_text",
    "ewfa_cigar_text_standard", "ewfa_check_cigar", "ewfa_read_pairs", "ewfa_free_pairs",
    "ewfa_synth_max_len", "ewfa_synth_pairs", "ewfa_cli_parse", "ewfa_cli_usage",
    "ewfa_device_alloc", "ewfa_device_free", "ewfa_memcpy", "ewfa_device_synchronize",
    "ewfa_server_create", "ewfa_server_destroy", "ewfa_server_submit", "ewfa_server_poll", "ewfa_server_wait",
    "ewfa_server_job_latency", "ewfa_server_stats", "ewfa_autotune", "ewfa_host_alloc", "ewfa_host_free",
    "ewfa_context_memory", "ewfa_policy_read", "ewfa_align_device_noalloc",
]

# device scratch the allocation-free call holds per GPU (EWFA_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20

_lib = None
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


def lib():
    """The loaded libewfa (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libewfa is not built ({LIB_PATH} missing); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.ewfa_last_error.restype = ctypes.c_char_p
    L.ewfa_create.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(Opts)]
    L.ewfa_destroy.argtypes = [_vp]
    L.ewfa_default_opts.argtypes = [ctypes.POINTER(Opts)]
    L.ewfa_device_info.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.ewfa_pack_ascii.argtypes = [_vp, _sz, _vp, _sz, _vp, _sz, _vp]
    L.ewfa_align_device.argtypes = [_vp, _vp, _vp, _sz, ctypes.c_uint32, _vp, _vp, _vp]
    L.ewfa_align_device_noalloc.argtypes = [_vp, _vp, _vp, _sz, ctypes.c_uint32, _vp, _vp, _vp]
    L.ewfa_align_host.argtypes = [_vp, _vp, _sz, _vp, _sz, _vp, _vp, _sz]
    L.ewfa_upload_words.argtypes = [_vp, _vp, _sz]
    L.ewfa_upload_words_async.argtypes = [_vp, _vp, _sz]
    L.ewfa_align_resident.argtypes = [_vp, _vp, _sz, _vp, _vp, _sz]
    L.ewfa_set_flags.argtypes = [_vp, ctypes.c_int]
    L.ewfa_last_timing.argtypes = [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.ewfa_max_distance.argtypes = [_vp, ctypes.c_uint32]
    L.ewfa_pack_host.argtypes = [ctypes.c_char_p, _sz, _vp]
    L.ewfa_packed_words.argtypes = [_sz]
    L.ewfa_packed_words.restype = _sz
    L.ewfa_layout_pairs.argtypes = [_vp, _vp, _sz, ctypes.c_uint32, _vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]
    L.ewfa_pack_pairs_host.argtypes = [_vp, _vp, _sz, _vp, _sz, _vp, ctypes.c_int]
    L.ewfa_cigar_text.argtypes = [_vp, _sz, ctypes.c_char_p, _sz]
    L.ewfa_cigar_text_standard.argtypes = [_vp, _sz, ctypes.c_char_p, _sz]
    L.ewfa_check_cigar.argtypes = [ctypes.c_char_p, _sz, ctypes.c_char_p, _sz, _vp, _sz]
    L.ewfa_read_pairs.argtypes = [ctypes.c_char_p, _sz, _sz, ctypes.c_int, ctypes.POINTER(PairsFile)]
    L.ewfa_free_pairs.argtypes = [ctypes.POINTER(PairsFile)]
    L.ewfa_synth_max_len.argtypes = [ctypes.POINTER(Synth)]
    L.ewfa_synth_max_len.restype = ctypes.c_uint32
    L.ewfa_synth_pairs.argtypes = [ctypes.POINTER(Synth), _sz, _sz, _vp, _vp, _sz, _vp, _vp, ctypes.c_int]
    L.ewfa_cli_parse.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(CliArgs)]
    L.ewfa_device_alloc.argtypes = [ctypes.c_int, _sz, ctypes.POINTER(_vp)]
    L.ewfa_device_free.argtypes = [_vp]
    L.ewfa_memcpy.argtypes = [_vp, _vp, _sz, ctypes.c_int]
    L.ewfa_device_synchronize.argtypes = [ctypes.c_int]
    L.ewfa_host_alloc.argtypes = [_sz, ctypes.POINTER(_vp)]
    L.ewfa_context_memory.argtypes = [_vp, ctypes.POINTER(_sz)]
    L.ewfa_host_free.argtypes = [_vp]
    L.ewfa_server_create.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(ServerOpts)]
    L.ewfa_server_destroy.argtypes = [_vp]
    L.ewfa_server_submit.argtypes = [_vp, _vp, _sz, _vp, _sz, _vp, _vp, _sz, ctypes.POINTER(ctypes.c_uint64)]
    L.ewfa_server_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    L.ewfa_server_wait.argtypes = [_vp, ctypes.c_uint64]
    L.ewfa_server_job_latency.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.ewfa_server_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    L.ewfa_autotune.argtypes = [ctypes.c_int, _vp, _sz, _vp, _sz, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                ctypes.c_char_p, _sz]
    L.ewfa_policy_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(BucketPolicy), ctypes.c_int]
    _lib = L
    return L


class EwfaError(RuntimeError):
    pass


def _check(rc):
    if rc != OK:
        raise EwfaError(f"libewfa error {rc}: {lib().ewfa_last_error().decode(errors='replace')}")
    return rc


def read_policy(path):
    """The buckets of a policy file as ewfa_create reads it (None: the
    built-in table), as (len_hi, waves, dcap, bt_lds, grab) tuples."""
    buf = (BucketPolicy * 16)()
    n = lib().ewfa_policy_read(path.encode() if path else None, buf, 16)
    if n < 0:
        raise EwfaError(f"libewfa error {n}: policy file {path!r} rejected")
    return [(b.len_hi, b.waves, b.dcap, b.bt_lds, b.grab) for b in buf[:n]]


def device_count():
    return lib().ewfa_device_count()


def device_info(device=0):
    name = ctypes.create_string_buffer(256)
    major, minor = ctypes.c_int(), ctypes.c_int()
    _check(lib().ewfa_device_info(device, name, 256, ctypes.byref(major), ctypes.byref(minor)))
    return name.value.decode(), major.value, minor.value


# ---------------------------------------------------------------------------
# workloads

CONFIGS = {
    # 10,000 pairs, 150 bp, 2% edits split equally between sub/ins/del
    "c1": dict(kind=0, seed=1, len=150, err=0.02, n=10_000),
    # 1,000,000 pairs, 1000 bp, 5% edits
    "c2": dict(kind=0, seed=2, len=1000, err=0.05, n=1_000_000),
    # 500,000 pairs, log-uniform 100..10,000 bp, error U[1%, 10%]
    "c3": dict(kind=1, seed=3, len_lo=100, len_hi=10_000, err_lo=0.01, err_hi=0.10, n=500_000),
    # 20,000 pairs, U[10k, 50k] bp, error U[10%, 15%], 30% of edits as indel runs
    "c4": dict(kind=2, seed=4, len_lo=10_000, len_hi=50_000, err_lo=0.10, err_hi=0.15, n=20_000),
}


def synth_spec(kind=0, seed=1, len=150, err=0.02, len_lo=0, len_hi=0, err_lo=0.0, err_hi=0.0, **_):
    return Synth(kind=kind, seed=seed, len=len, len_lo=len_lo, len_hi=len_hi, err=err, err_lo=err_lo,
                 err_hi=err_hi)


class Workload:
    """n synthetic pairs: ASCII (stride bytes per sequence), lengths, packed
    words and pair descriptors laid out for upload."""

    def __init__(self, spec, n, first=0, cigar_cap=0, threads=8):
        L = lib()
        self.spec = spec
        self.n = n
        self.stride = int(L.ewfa_synth_max_len(ctypes.byref(spec)))
        self.pat = np.zeros(n * self.stride + 1, dtype=np.uint8)
        self.txt = np.zeros(n * self.stride + 1, dtype=np.uint8)
        self.plen = np.zeros(max(n, 1), dtype=np.uint32)
        self.tlen = np.zeros(max(n, 1), dtype=np.uint32)
        _check(L.ewfa_synth_pairs(ctypes.byref(spec), first, n, self.pat.ctypes.data, self.txt.ctypes.data,
                                  self.stride, self.plen.ctypes.data, self.tlen.ctypes.data, threads))
        self.plen = self.plen[:n]
        self.tlen = self.tlen[:n]
        self.layout(cigar_cap, threads)

    def layout(self, cigar_cap, threads=8):
        L = lib()
        self.pairs = np.zeros(max(self.n, 1), dtype=PAIR_DTYPE)
        nw, cb = _sz(), _sz()
        _check(L.ewfa_layout_pairs(self.plen.ctypes.data, self.tlen.ctypes.data, self.n, cigar_cap,
                                   self.pairs.ctypes.data, ctypes.byref(nw), ctypes.byref(cb)))
        self.n_words = nw.value
        self.cigar_bytes = max(cb.value, 4)
        self.words = np.zeros(max(self.n_words, 1), dtype=np.uint64)
        _check(L.ewfa_pack_pairs_host(self.pat.ctypes.data, self.txt.ctypes.data, self.stride,
                                      self.pairs.ctypes.data, self.n, self.words.ctypes.data, threads))
        self.max_len = int(max(self.plen.max(initial=0), self.tlen.max(initial=0)))

    def pattern(self, i):
        return self.pat[i * self.stride: i * self.stride + int(self.plen[i])].tobytes()

    def text(self, i):
        return self.txt[i * self.stride: i * self.stride + int(self.tlen[i])].tobytes()


def slice_pairs(words, pairs, a, b):
    """Pairs [a, b) of a sequential layout as a self-contained job:
    (words view, rebased descriptors, cigar_bytes, first word, first CIGAR byte).
    The words view shares memory with `words`."""
    sub = np.array(pairs[a:b], dtype=PAIR_DTYPE, copy=True)
    if b <= a:
        return words[:1], np.zeros(1, dtype=PAIR_DTYPE), 4, 0, 0
    w0 = int(sub[0]["pattern_off"])
    last = sub[-1]
    w1 = int(last["text_off"]) + int(lib().ewfa_packed_words(int(last["tlen"])))
    c0 = int(sub[0]["cigar_off"])
    c1 = int(last["cigar_off"]) + int(last["cigar_cap"])
    sub["pattern_off"] -= w0
    sub["text_off"] -= w0
    sub["cigar_off"] -= c0
    return words[w0:w1], sub, c1 - c0, w0, c0


def from_sequences(patterns, texts, cigar_cap=0):
    """Lays out explicit (pattern, text) byte strings: (words, pairs, cigar_bytes, max_len)."""
    L = lib()
    n = len(patterns)
    plen = np.array([len(p) for p in patterns] or [0], dtype=np.uint32)
    tlen = np.array([len(t) for t in texts] or [0], dtype=np.uint32)
    pairs = np.zeros(max(n, 1), dtype=PAIR_DTYPE)
    nw, cb = _sz(), _sz()
    _check(L.ewfa_layout_pairs(plen.ctypes.data, tlen.ctypes.data, n, cigar_cap, pairs.ctypes.data,
                               ctypes.byref(nw), ctypes.byref(cb)))
    words = np.zeros(max(nw.value, 1), dtype=np.uint64)
    for i in range(n):
        L.ewfa_pack_host(bytes(patterns[i]), len(patterns[i]), words.ctypes.data + 8 * int(pairs[i]["pattern_off"]))
        L.ewfa_pack_host(bytes(texts[i]), len(texts[i]), words.ctypes.data + 8 * int(pairs[i]["text_off"]))
    max_len = int(max(plen.max(initial=0), tlen.max(initial=0)))
    return words, pairs[:max(n, 1)], max(cb.value, 4), max_len


# ---------------------------------------------------------------------------
# host helpers

def cigar_text(stream):
    """The reference's printed text for a stream (list of ints or int8 array)."""
    arr = np.asarray(stream, dtype=np.int8)
    n = arr.size
    if n == 0:
        arr = np.zeros(1, dtype=np.int8)
    cap = 8 * n + 16
    out = ctypes.create_string_buffer(cap)
    r = lib().ewfa_cigar_text(arr.ctypes.data, n, out, cap)
    if r < 0:
        raise EwfaError("text buffer too small")
    return out.value.decode()


def cigar_text_standard(stream):
    arr = np.asarray(stream, dtype=np.int8)
    n = arr.size
    if n == 0:
        arr = np.zeros(1, dtype=np.int8)
    cap = 16 * n + 16
    out = ctypes.create_string_buffer(cap)
    r = lib().ewfa_cigar_text_standard(arr.ctypes.data, n, out, cap)
    if r < 0:
        raise EwfaError("text buffer too small")
    return out.value.decode()


def check_cigar(pattern, text, stream):
    arr = np.asarray(stream, dtype=np.int8)
    n = arr.size
    if n == 0:
        arr = np.zeros(1, dtype=np.int8)
    return lib().ewfa_check_cigar(bytes(pattern), len(pattern), bytes(text), len(text), arr.ctypes.data, n)


def streams_from(results, pairs, cigars):
    """List of per-pair streams (int8 arrays) from host copies of the outputs."""
    out = []
    for i in range(len(results)):
        ln = int(results[i]["cigar_len"])
        cap = int(pairs[i]["cigar_cap"])
        off = int(pairs[i]["cigar_off"])
        out.append(cigars[off: off + min(max(ln, 0), cap)])
    return out


# ---------------------------------------------------------------------------
# device memory

class DeviceBuffer:
    """A device allocation owned from Python (hipMalloc through libewfa)."""

    def __init__(self, nbytes, device=0):
        self.nbytes = int(nbytes)
        p = _vp()
        _check(lib().ewfa_device_alloc(device, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr, device=0):
        arr = np.ascontiguousarray(arr)
        b = cls(max(arr.nbytes, 4), device)
        if arr.nbytes:
            _check(lib().ewfa_memcpy(b.ptr, arr.ctypes.data, arr.nbytes, 0))
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().ewfa_memcpy(self.ptr, arr.ctypes.data, arr.nbytes, 0))

    def fill_zero(self):
        self.upload(np.zeros(self.nbytes, dtype=np.uint8))

    def download(self, dtype=np.uint8, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(lib().ewfa_memcpy(out.ctypes.data, self.ptr, out.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            lib().ewfa_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A numpy view of pinned host memory (freed with the object)."""

    def __init__(self, count, dtype):
        dtype = np.dtype(dtype)
        p = _vp()
        _check(lib().ewfa_host_alloc(max(count * dtype.itemsize, 4), ctypes.byref(p)))
        self.ptr = p.value
        buf = (ctypes.c_uint8 * max(count * dtype.itemsize, 4)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint8)[: count * dtype.itemsize].view(dtype)

    def __del__(self):
        try:
            if self.ptr:
                lib().ewfa_host_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def synchronize(device=0):
    _check(lib().ewfa_device_synchronize(device))


# ---------------------------------------------------------------------------
# device context

class Context:
    """One libewfa context (device buffers, streams, policy)."""

    def __init__(self, device=0, mode=MODE_PERSISTENT, max_distance=0, batch_size=0, policy_path=None, flags=0):
        L = lib()
        self._policy = policy_path.encode() if policy_path else None
        o = Opts(device=device, mode=mode, max_distance=max_distance, batch_size=batch_size,
                 policy_path=self._policy)
        h = _vp()
        _check(L.ewfa_create(ctypes.byref(h), ctypes.byref(o)))
        self.handle = h
        if flags:
            self.set_flags(flags)

    def close(self):
        if self.handle:
            lib().ewfa_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_flags(self, flags):
        _check(lib().ewfa_set_flags(self.handle, flags))

    def align_device(self, d_words, d_pairs, n, max_len, d_results, d_cigars, stream=None):
        _check(lib().ewfa_align_device(self.handle, d_words, d_pairs, n, max_len, d_results, d_cigars, stream))

    def align_device_noalloc(self, d_words, d_pairs, n, max_len, d_results, d_cigars, stream=None):
        """ewfa_align_device_noalloc: the same results with the library's
        static device scratch and no allocation."""
        _check(lib().ewfa_align_device_noalloc(self.handle, d_words, d_pairs, n, max_len, d_results, d_cigars,
                                               stream))

    def align_host(self, words, pairs, n, cigar_bytes, results=None, cigars=None):
        """Host arrays in, (results, cigars) host arrays out (fresh arrays, or
        the caller's, e.g. views of pinned memory)."""
        if results is None:
            results = np.zeros(max(n, 1), dtype=RESULT_DTYPE)
        if cigars is None:
            cigars = np.zeros(max(cigar_bytes, 4), dtype=np.int8)
        assert results.size >= n and cigars.nbytes >= cigar_bytes
        _check(lib().ewfa_align_host(self.handle, words.ctypes.data, words.size, pairs.ctypes.data, n,
                                     results.ctypes.data, cigars.ctypes.data, min(cigars.nbytes, max(cigar_bytes, 4))))
        return results[:n], cigars

    def last_timing(self):
        a, e = ctypes.c_float(), ctypes.c_float()
        _check(lib().ewfa_last_timing(self.handle, ctypes.byref(a), ctypes.byref(e)))
        return a.value, e.value

    def memory_bytes(self):
        """Device bytes the context holds (scratch, backtrace spill, lists)."""
        b = _sz()
        _check(lib().ewfa_context_memory(self.handle, ctypes.byref(b)))
        return b.value

    def max_distance(self, max_len):
        r = lib().ewfa_max_distance(self.handle, max_len)
        _check(r if r < 0 else OK)
        return r


# ---------------------------------------------------------------------------
# serving

class ServerJob:
    """A submitted job: owns its host buffers until it completes."""

    def __init__(self, server, words, pairs, n, cigar_bytes, results=None, cigars=None):
        self.server = server
        self.words = np.ascontiguousarray(words)
        self.pairs = np.ascontiguousarray(pairs[:max(n, 1)])
        self.n = n
        # caller-provided outputs (e.g. views of pinned memory) or fresh arrays
        self.results = results if results is not None else np.zeros(max(n, 1), dtype=RESULT_DTYPE)
        self.cigars = cigars if cigars is not None else np.zeros(max(cigar_bytes, 4), dtype=np.int8)
        assert self.results.flags.c_contiguous and self.results.size >= n
        assert self.cigars.flags.c_contiguous and self.cigars.nbytes >= cigar_bytes
        jid = ctypes.c_uint64()
        _check(lib().ewfa_server_submit(server.handle, self.words.ctypes.data, self.words.size,
                                        self.pairs.ctypes.data, n, self.results.ctypes.data,
                                        self.cigars.ctypes.data, min(self.cigars.nbytes, max(cigar_bytes, 4)),
                                        ctypes.byref(jid)))
        self.id = jid.value

    def done(self):
        d = ctypes.c_int()
        _check(lib().ewfa_server_poll(self.server.handle, self.id, ctypes.byref(d)))
        return bool(d.value)

    def wait(self):
        _check(lib().ewfa_server_wait(self.server.handle, self.id))
        return self.results[: self.n], self.cigars

    def latency(self):
        s = ctypes.c_double()
        _check(lib().ewfa_server_job_latency(self.server.handle, self.id, ctypes.byref(s)))
        return s.value


class Server:
    """The job server: continuous batching of queued jobs into persistent
    launches, admitted first-come first-served by arena footprint."""

    def __init__(self, device=0, slots=0, arena_bytes=0, max_distance=0, policy_path=None, max_batch_pairs=0,
                 max_batch_cigar_bytes=0, flags=0):
        self._policy = policy_path.encode() if policy_path else None
        o = ServerOpts(device=device, slots=slots, arena_bytes=arena_bytes, max_distance=max_distance,
                       policy_path=self._policy, max_batch_pairs=max_batch_pairs,
                       max_batch_cigar_bytes=max_batch_cigar_bytes, flags=flags)
        h = _vp()
        _check(lib().ewfa_server_create(ctypes.byref(h), ctypes.byref(o)))
        self.handle = h

    def submit(self, words, pairs, n, cigar_bytes, results=None, cigars=None):
        return ServerJob(self, words, pairs, n, cigar_bytes, results, cigars)

    def stats(self):
        st = ServerStats()
        _check(lib().ewfa_server_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().ewfa_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# autotuner

def autotune(words, pairs, n, bucket_hi, out_path, reps=2, device=0):
    """Times candidate configurations per length bucket on a sample and
    writes the fastest to out_path; returns the text report."""
    hi = np.ascontiguousarray(bucket_hi, dtype=np.uint32)
    rep = ctypes.create_string_buffer(1 << 16)
    words = np.ascontiguousarray(words)
    pairs = np.ascontiguousarray(pairs[:max(n, 1)])
    _check(lib().ewfa_autotune(device, words.ctypes.data, words.size, pairs.ctypes.data, n, hi.ctypes.data,
                               len(hi), reps, out_path.encode(), rep, len(rep)))
    return rep.value.decode()
