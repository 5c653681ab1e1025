"""ctypes view of the oracle (oracle/build/liboracle.so) for the tests.

The oracle is test infrastructure: a CPU restatement of the reference's
alignment path.  Only tests, the smoke check and bench.py's CPU baseline load
it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "build", "liboracle.so")
REF_ARG_SO = os.path.join(ORACLE_DIR, "_ref", "libarg_handler.so")


class _Alignment(ctypes.Structure):
    _fields_ = [
        ("distance", ctypes.c_int),
        ("capped", ctypes.c_int),
        ("undefined", ctypes.c_int),
        ("nwords", ctypes.c_int),
        ("ops", ctypes.POINTER(ctypes.c_uint64)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = ctypes.CDLL(ORACLE_SO)
        c_u8p = ctypes.POINTER(ctypes.c_uint8)
        L.or_code.argtypes = [ctypes.c_int]
        L.or_extend.argtypes = [c_u8p, ctypes.c_int, c_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_align.argtypes = [c_u8p, ctypes.c_int, c_u8p, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(_Alignment)]
        L.or_alignment_free.argtypes = [ctypes.POINTER(_Alignment)]
        L.or_op_at.argtypes = [ctypes.POINTER(_Alignment), ctypes.c_int]
        L.or_cigar_stream.argtypes = [c_u8p, ctypes.c_int, c_u8p, ctypes.c_int,
                                      ctypes.POINTER(_Alignment), ctypes.c_void_p, ctypes.c_int]
        L.or_cigar_text.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.or_check_cigar.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_int]
        L.or_levenshtein.argtypes = [c_u8p, ctypes.c_int, c_u8p, ctypes.c_int]
        L.or_pack_u64.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p]
        L.or_pack_ref32.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p]
        L.or_unpack_ref32.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p]
        L.or_read_pairs.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t,
                                    ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        L.or_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_void_p]
        L.or_batch.restype = ctypes.c_int64
        L.or_batch_text.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                             ctypes.c_int64]
        L.or_batch_text.restype = ctypes.c_int64
        L.or_batch_streams.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.or_batch_streams.restype = ctypes.c_int64
        _lib = L
    return _lib


def _b(s):
    return s.encode() if isinstance(s, str) else bytes(s)


def codes(seq):
    b = np.frombuffer(_b(seq), dtype=np.uint8)
    return ((b & 6) >> 1).astype(np.uint8)


def _u8p(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def extend(p, t, k, off):
    pa, pp = _u8p(codes(p))
    ta, tp = _u8p(codes(t))
    return lib().or_extend(pp, len(p), tp, len(t), k, off)


def align(p, t, max_distance=0):
    """Returns (distance, ops list, capped, undefined)."""
    pa, pp = _u8p(codes(p))
    ta, tp = _u8p(codes(t))
    a = _Alignment()
    rc = lib().or_align(pp, len(p), tp, len(t), max_distance, ctypes.byref(a))
    assert rc == 0
    n = a.distance if not a.capped else max_distance
    ops = [lib().or_op_at(ctypes.byref(a), s) for s in range(n)]
    res = (a.distance, ops, bool(a.capped), bool(a.undefined))
    lib().or_alignment_free(ctypes.byref(a))
    return res


def stream(p, t, max_distance=0):
    """Compressed CIGAR stream as a list of signed ints, plus the distance."""
    pa, pp = _u8p(codes(p))
    ta, tp = _u8p(codes(t))
    a = _Alignment()
    assert lib().or_align(pp, len(p), tp, len(t), max_distance, ctypes.byref(a)) == 0
    cap = len(p) + len(t) + 8
    buf = (ctypes.c_int8 * cap)()
    n = lib().or_cigar_stream(pp, len(p), tp, len(t), ctypes.byref(a), buf, cap)
    d = a.distance
    lib().or_alignment_free(ctypes.byref(a))
    assert n <= cap
    return d, [buf[i] for i in range(n)]


def text(stream_bytes):
    arr = np.asarray(stream_bytes, dtype=np.int8)
    if arr.size == 0:
        arr = np.zeros(1, dtype=np.int8)
        n = 0
    else:
        n = len(stream_bytes)
    cap = 8 * n + 16
    out = ctypes.create_string_buffer(cap)
    r = lib().or_cigar_text(arr.ctypes.data, n, out, cap)
    assert r >= 0
    return out.value.decode()


def check(p, t, stream_bytes):
    arr = np.asarray(stream_bytes, dtype=np.int8)
    n = len(stream_bytes)
    if arr.size == 0:
        arr = np.zeros(1, dtype=np.int8)
    return lib().or_check_cigar(_b(p), len(p), _b(t), len(t), arr.ctypes.data, n)


def levenshtein(p, t):
    pa, pp = _u8p(codes(p))
    ta, tp = _u8p(codes(t))
    return lib().or_levenshtein(pp, len(p), tp, len(t))


def pack_u64(seq):
    n = len(seq) // 32 + 2
    out = np.zeros(n, dtype=np.uint64)
    lib().or_pack_u64(_b(seq), len(seq), out.ctypes.data)
    return out[: (len(seq) + 31) // 32 + 1]


def pack_ref32(seq):
    out = np.zeros(max(1, (len(seq) + 15) // 16), dtype=np.uint32)
    lib().or_pack_ref32(_b(seq), len(seq), out.ctypes.data)
    return out


def unpack_ref32(words, n):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    out = ctypes.create_string_buffer(n + 1)
    lib().or_unpack_ref32(words.ctypes.data, n, out)
    return out.raw[:n].decode()


def read_pairs(path, seq_len, n, stride=None):
    """Returns (rc, [(pattern, text)], plen, tlen) with the reference's quirks."""
    if stride is None:
        stride = max(4 * seq_len + 64, 64)
    slab = np.zeros(2 * n * stride + 1, dtype=np.uint8)
    plen = np.zeros(max(n, 1), dtype=np.int64)
    tlen = np.zeros(max(n, 1), dtype=np.int64)
    rc = lib().or_read_pairs(_b(path), seq_len, n, slab.ctypes.data, stride,
                             plen.ctypes.data, tlen.ctypes.data)
    pairs = []
    for i in range(n):
        pl = int(min(plen[i], stride))
        tl = int(min(tlen[i], stride))
        p = slab[2 * i * stride: 2 * i * stride + pl].tobytes()
        t = slab[(2 * i + 1) * stride: (2 * i + 1) * stride + tl].tobytes()
        pairs.append((p, t))
    return rc, pairs, plen[:n], tlen[:n]


def batch(patterns, texts, max_distance=0, threads=1):
    """Runs or_batch over lists of byte strings; returns (ok, dist, slen)."""
    n = len(patterns)
    pat = b"".join(patterns) + b"\0"
    txt = b"".join(texts) + b"\0"
    plen = np.array([len(x) for x in patterns], dtype=np.int32)
    tlen = np.array([len(x) for x in texts], dtype=np.int32)
    po = np.zeros(n, dtype=np.int64)
    to = np.zeros(n, dtype=np.int64)
    if n:
        po[1:] = np.cumsum(plen)[:-1]
        to[1:] = np.cumsum(tlen)[:-1]
    dist = np.zeros(max(n, 1), dtype=np.int32)
    slen = np.zeros(max(n, 1), dtype=np.int32)
    pb = ctypes.create_string_buffer(pat, len(pat))
    tb = ctypes.create_string_buffer(txt, len(txt))
    ok = lib().or_batch(pb, po.ctypes.data, plen.ctypes.data, tb, to.ctypes.data, tlen.ctypes.data,
                        n, max_distance, threads, dist.ctypes.data, slen.ctypes.data)
    return ok, dist[:n], slen[:n]


def batch_streams(patterns, texts, max_distance=0, threads=1):
    """Oracle over lists of byte strings; returns (ok, dist, slen, streams)
    with streams a list of int8 arrays."""
    n = len(patterns)
    pat = np.frombuffer(b"".join(patterns) + b"\0", dtype=np.uint8)
    txt = np.frombuffer(b"".join(texts) + b"\0", dtype=np.uint8)
    plen = np.array([len(x) for x in patterns] or [0], dtype=np.int32)
    tlen = np.array([len(x) for x in texts] or [0], dtype=np.int32)
    po = np.zeros(max(n, 1), dtype=np.int64)
    to = np.zeros(max(n, 1), dtype=np.int64)
    if n:
        po[1:n] = np.cumsum(plen[:n])[:-1]
        to[1:n] = np.cumsum(tlen[:n])[:-1]
    sstride = int(max(8, (plen.astype(np.int64) + tlen).max(initial=0) + 8))
    streams = np.zeros(max(n, 1) * sstride, dtype=np.int8)
    dist = np.zeros(max(n, 1), dtype=np.int32)
    slen = np.zeros(max(n, 1), dtype=np.int32)
    ok = lib().or_batch_streams(pat.ctypes.data, po.ctypes.data, plen.ctypes.data, txt.ctypes.data,
                                to.ctypes.data, tlen.ctypes.data, n, max_distance, threads,
                                dist.ctypes.data, slen.ctypes.data, streams.ctypes.data, sstride)
    out = [streams[i * sstride: i * sstride + min(int(slen[i]), sstride)] for i in range(n)]
    return ok, dist[:n], slen[:n], out


def batch_workload(w, max_distance=0, threads=8, idx=None, text=True):
    """Oracle over pairs of an ewfa.Workload (all, or the indices idx), with
    or without the CIGAR text step.  Returns (ok, dist, slen, streams) with
    streams a list of int8 arrays."""
    if idx is None:
        idx = np.arange(w.n)
    idx = np.asarray(idx, dtype=np.int64)
    n = len(idx)
    stride = w.stride
    pat_off = (idx * stride).astype(np.int64)
    txt_off = (idx * stride).astype(np.int64)
    plen = w.plen[idx].astype(np.int32)
    tlen = w.tlen[idx].astype(np.int32)
    sstride = int(max(8, (plen.astype(np.int64) + tlen).max(initial=0) + 8))
    streams = np.zeros(max(n, 1) * sstride, dtype=np.int8)
    dist = np.zeros(max(n, 1), dtype=np.int32)
    slen = np.zeros(max(n, 1), dtype=np.int32)
    ok = lib().or_batch_text(w.pat.ctypes.data, pat_off.ctypes.data, plen.ctypes.data,
                             w.txt.ctypes.data, txt_off.ctypes.data, tlen.ctypes.data,
                             n, max_distance, threads, 1 if text else 0, dist.ctypes.data, slen.ctypes.data,
                             streams.ctypes.data, sstride)
    out = [streams[i * sstride: i * sstride + min(int(slen[i]), sstride)] for i in range(n)]
    return ok, dist[:n], slen[:n], out
