"""Pins of the oracle (oracle/ewfa_oracle.c) against checks that are not a
retyping of its own code: hand-derived known answers, brute force, the
Levenshtein recurrence written recursively, and an independent WFA
formulation built from the full DP matrix (furthest-reaching offsets, a per
cell op table with the DEL > SUB > INS rule, explicit traceback).
"""
import functools
import itertools
import os
import random
import struct

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle_py as orc

ALPHA = "ACGT"


# ---------------------------------------------------------------------------
# independent formulations used as checks

def py_codes(s):
    """Reference enum A=0, C=1, T=2, G=3 (src/wavefront_structures.h:78-83);
    lowercase, U and N fold onto it through the ASCII bits the reference uses."""
    table = {"A": 0, "a": 0, "C": 1, "c": 1, "T": 2, "t": 2, "U": 2, "u": 2, "G": 3, "g": 3, "N": 3, "n": 3}
    return [table[ch] for ch in s]


def dp_matrix(p, t):
    n, m = len(p), len(t)
    D = [[0] * (m + 1) for _ in range(n + 1)]
    for i in range(n + 1):
        D[i][0] = i
    for j in range(m + 1):
        D[0][j] = j
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            D[i][j] = min(D[i - 1][j] + 1, D[i][j - 1] + 1, D[i - 1][j - 1] + (p[i - 1] != t[j - 1]))
    return D


def lev_recursive(p, t):
    @functools.lru_cache(maxsize=None)
    def go(i, j):
        if i == 0:
            return j
        if j == 0:
            return i
        return min(go(i - 1, j) + 1, go(i, j - 1) + 1, go(i - 1, j - 1) + (p[i - 1] != t[j - 1]))
    return go(len(p), len(t))


def b3_ops(p, t, cap=0):
    """Ops (1-based step order) of the WFA path from the DP matrix: offsets
    F_d[k] = max{h : DP[h-k][h] <= d} (-1 if none), op of (d, k) = lexicographic
    max of (F_{d-1}[k+1], DEL), (F_{d-1}[k]+1, SUB), (F_{d-1}[k-1]+1, INS),
    traced back from (D, tlen-plen); with a cap, from (cap-1, tlen-plen)."""
    pc, tc = py_codes(p), py_codes(t)
    n, m = len(pc), len(tc)
    D = dp_matrix(pc, tc)
    dist = D[n][m]

    def F(d, k):
        best = -1
        for h in range(max(0, k), min(m, n + k) + 1):
            if D[h - k][h] <= d:
                best = h
        return best

    tk = m - n
    if cap and dist >= cap:
        start = cap - 1
        if abs(tk) > start:
            return dist, None
    else:
        start = dist
    k = tk
    ops = []
    for lev in range(start, 0, -1):
        cands = [(F(lev - 1, k + 1), 3), (F(lev - 1, k) + 1, 2), (F(lev - 1, k - 1) + 1, 1)]
        val, op = max(cands)  # tuple order: value, then op code 3 > 2 > 1
        ops.append(op)
        k += op - 2
    ops.reverse()
    return dist, ops


def py_stream(p, t, ops, stop_at_noop=False):
    """Forward walk of A.7 on codes: runs before each op and after the last."""
    pc, tc = py_codes(p), py_codes(t)
    out = []
    k = off = 0

    def ext(k, off):
        v, h = off - k, off
        while v < len(pc) and h < len(tc) and pc[v] == tc[h]:
            v += 1
            h += 1
        return h

    def run(r):
        out.extend([-127] * (r // 127))
        if r % 127:
            out.append(-(r % 127))

    for op in ops:
        e = ext(k, off)
        run(e - off)
        off = e
        if op == 3:
            out.append(ord("D"))
            k -= 1
        elif op == 2:
            out.append(ord("X"))
            off += 1
        else:
            out.append(ord("I"))
            k += 1
            off += 1
    e = ext(k, off)
    run(e - off)
    return out


def rand_seq(rng, n):
    return "".join(rng.choice(ALPHA) for _ in range(n))


def mutate(rng, s, edits):
    s = list(s)
    for _ in range(edits):
        kind = rng.randrange(3)
        pos = rng.randrange(len(s) + 1)
        if kind == 0 and pos < len(s):
            s[pos] = rng.choice([c for c in ALPHA if c != s[pos]])
        elif kind == 1:
            s.insert(pos, rng.choice(ALPHA))
        elif pos < len(s):
            del s[pos]
    return "".join(s)


# ---------------------------------------------------------------------------
# code map and packing

def test_code_map_matches_reference_enum():
    for ch, want in zip("AaCcTtUuGgNn", [0, 0, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3]):
        assert orc.lib().or_code(ord(ch)) == want, ch
    # the reference's LUT {A,C,T,G} inverts the map for the four bases
    rng = random.Random(1)
    for n in [1, 15, 16, 17, 31, 32, 33, 100]:
        s = rand_seq(rng, n)
        assert orc.unpack_ref32(orc.pack_ref32(s), n) == s


def test_ref32_byte_layout():
    # ACGT = codes 0,1,3,2 = 0b00011110 = 0x1E per group of four bases
    w = orc.pack_ref32("ACGT" * 4)
    assert int(w[0]) == 0x1E1E1E1E
    # first base lands in bits 31:30; the reference stores the first packed
    # byte (bases 0..3) at byte 3 of the little-endian word (src/kernels.cu:372-392)
    w = orc.pack_ref32("C" + "A" * 15)
    assert int(w[0]) == 0x40000000
    assert struct.pack("<I", int(w[0]))[3] == 0x40
    w = orc.pack_ref32("AAAAG" + "A" * 11)
    assert struct.pack("<I", int(w[0]))[2] == 0b11000000


def test_u64_layout_is_two_ref32_words():
    rng = random.Random(2)
    for n in [1, 31, 32, 33, 64, 200]:
        s = rand_seq(rng, n)
        w64 = orc.pack_u64(s)
        w32 = list(orc.pack_ref32(s)) + [0, 0]
        for m in range((n + 31) // 32):
            assert int(w64[m]) == (int(w32[2 * m]) << 32) | int(w32[2 * m + 1])
        assert int(w64[-1]) == 0  # pad word


# ---------------------------------------------------------------------------
# extend and distance

def test_extend_brute_force():
    rng = random.Random(3)
    for _ in range(400):
        p = rand_seq(rng, rng.randrange(0, 90))
        t = mutate(rng, p, rng.randrange(0, 4)) if rng.random() < 0.7 else rand_seq(rng, rng.randrange(0, 90))
        pc, tc = py_codes(p), py_codes(t)
        for _ in range(5):
            k = rng.randrange(-len(p) - 2, len(t) + 3)
            off = rng.randrange(max(0, k), max(0, k) + 40)
            v, h = off - k, off
            r = 0
            while v + r < len(pc) and h + r < len(tc) and pc[v + r] == tc[h + r]:
                r += 1
            assert orc.extend(p, t, k, off) == off + r


def test_levenshtein_textbook_and_recursive():
    assert orc.levenshtein("", "") == 0
    assert orc.levenshtein("ACGT", "") == 4
    assert orc.levenshtein("", "GG") == 2
    assert orc.levenshtein("ACGT", "AGT") == 1
    assert orc.levenshtein("ACGT", "ACCGT") == 1
    assert orc.levenshtein("ACGT", "TGCA") == 4
    assert orc.levenshtein("AAAA", "NNNN") == 4
    assert orc.levenshtein("acgt", "ACGT") == 0  # codes ignore case
    rng = random.Random(4)
    for _ in range(300):
        p = rand_seq(rng, rng.randrange(0, 9))
        t = rand_seq(rng, rng.randrange(0, 9))
        assert orc.levenshtein(p, t) == lev_recursive(tuple(py_codes(p)), tuple(py_codes(t)))


def test_align_distance_equals_dp():
    rng = random.Random(5)
    for _ in range(300):
        n = rng.randrange(0, 300)
        p = rand_seq(rng, n)
        t = mutate(rng, p, rng.randrange(0, max(1, n // 8) + 1))
        d, ops, capped, undef = orc.align(p, t)
        assert d == orc.levenshtein(p, t)
        assert len(ops) == d and all(o in (1, 2, 3) for o in ops)


# ---------------------------------------------------------------------------
# ops and path: the independent DP-derived formulation

def test_ops_exhaustive_tiny():
    seqs = [""] + ["".join(x) for n in (1, 2, 3) for x in itertools.product(ALPHA, repeat=n)]
    for p in seqs:
        for t in seqs:
            d, ops, _, _ = orc.align(p, t)
            bd, bops = b3_ops(p, t)
            assert (d, ops) == (bd, bops), (p, t, ops, bops)


@settings(max_examples=300, deadline=None)
@given(st.text(alphabet="ACGT", max_size=14), st.text(alphabet="ACGT", max_size=14))
def test_ops_random_small(p, t):
    d, ops, _, _ = orc.align(p, t)
    assert (d, ops) == b3_ops(p, t)


def test_ops_mutated_medium():
    rng = random.Random(6)
    for _ in range(60):
        p = rand_seq(rng, rng.randrange(20, 60))
        t = mutate(rng, p, rng.randrange(1, 8))
        d, ops, _, _ = orc.align(p, t)
        assert (d, ops) == b3_ops(p, t)
        d2, s = orc.stream(p, t)
        assert s == py_stream(p, t, ops)


def test_cap_matches_dp_formulation():
    rng = random.Random(7)
    seen = 0
    for _ in range(80):
        p = rand_seq(rng, rng.randrange(10, 40))
        t = rand_seq(rng, rng.randrange(10, 40))
        for cap in (3, 5, 8):
            d, ops, capped, undef = orc.align(p, t, max_distance=cap)
            bd, bops = b3_ops(p, t, cap=cap)
            if bd >= cap:
                seen += 1
                assert capped and d == cap
                if bops is None:
                    assert abs(len(t) - len(p)) >= cap
                else:
                    assert ops[: cap - 1] == bops and ops[cap - 1] == 0
            else:
                assert not capped and (d, ops) == (bd, bops)
    assert seen > 50


def test_capped_stream_stops_after_run():
    # D >= cap: the walk emits cap-1 ops, the run after the last of them,
    # and stops at the NOOP (src/kernels.cu:471-478)
    p, t = "A" * 20 + "C" * 10, "G" * 10 + "A" * 20
    d, s = orc.stream(p, t, max_distance=4)
    assert d == 4
    assert sum(1 for b in s if b > 0) == 3


# ---------------------------------------------------------------------------
# known answers (derived by hand from src/kernels.cu and src/wavefront.cuh)

KATS = [
    # pattern, text, distance, stream, printed text
    ("ACGT", "ACGT", 0, [-4], ""),
    ("A", "", 1, [ord("D")], "D"),
    ("AA", "A", 1, [-1, ord("D")], "68MD"),
    ("ACGT", "AGGT", 1, [-1, ord("X"), -2], "88MX"),
    ("", "A", 1, [ord("I")], "I"),
    ("A", "C", 1, [ord("X")], "X"),
    ("AC", "A", 1, [-1, ord("D")], "68MD"),
    ("CA", "A", 1, [ord("D"), -1], "D"),
    # d=1: k=-1 DEL (tie) off 1, k=0 SUB off 1, k=1 INS off 2; d=2, k=0:
    # del=2, sub=2, ins=2 -> DEL from k=1, which came by INS
    ("AC", "CA", 2, [ord("I"), -1, ord("D")], "I68MD"),
    ("", "", 0, [], ""),
    ("T" * 300, "T" * 300, 0, [-127, -127, -46], ""),
]


@pytest.mark.parametrize("p,t,d,stream,text", KATS)
def test_known_answers(p, t, d, stream, text):
    od, os_ = orc.stream(p, t)
    assert od == d
    assert os_ == stream
    assert orc.text(os_) == text
    assert orc.check(p, t, os_) == 1


def test_text_quirks():
    # the number before M is the op byte's value, not the run; a trailing
    # run is never printed
    assert orc.text([-5, ord("X"), -3]) == "88MX"
    assert orc.text([-5, ord("I")]) == "73MI"
    assert orc.text([-127, -127, ord("D"), ord("D"), -3]) == "68MDD"
    assert orc.text([ord("X"), ord("X")]) == "XX"
    assert orc.text([-3, 0, ord("X")]) == ""  # stops at the terminator


# ---------------------------------------------------------------------------
# stream invariants and the validity check

def test_stream_invariants():
    rng = random.Random(8)
    for _ in range(200):
        n = rng.randrange(0, 600)
        p = rand_seq(rng, n)
        t = mutate(rng, p, rng.randrange(0, 12))
        d, s = orc.stream(p, t)
        assert orc.check(p, t, s) == 1
        assert sum(1 for b in s if b > 0) == d
        assert all(-127 <= b <= -1 or b in (ord("X"), ord("I"), ord("D")) for b in s)
        for a, b in zip(s, s[1:]):
            if a < 0 and b < 0:
                assert a == -127  # only full chunks are followed by more run bytes


def test_check_rejects_planted_mistakes():
    rng = random.Random(9)
    for _ in range(100):
        p = rand_seq(rng, 80)
        t = mutate(rng, p, 4)
        d, s = orc.stream(p, t)
        if d == 0:
            continue
        i = next(j for j, b in enumerate(s) if b > 0)
        bad = list(s)
        bad[i] = {ord("X"): ord("I"), ord("I"): ord("D"), ord("D"): ord("X")}[bad[i]]
        assert orc.check(p, t, bad) != 1
        assert orc.check(p, t, s[:-1]) != 1
        assert orc.check(p, t, s + [ord("I")]) != 1
        assert orc.check(p, t, s[:i] + [ord("Q")] + s[i + 1:]) == -1


def test_check_compares_raw_ascii():
    # equal codes but different letters pass the aligner, not the check
    d, s = orc.stream("NNNN", "GGGG")
    assert d == 0 and s == [-4]
    assert orc.check("NNNN", "GGGG", s) == 0


# ---------------------------------------------------------------------------
# reader quirks (utils/sequence_reader.cu:108-201)

def _write(tmp_path, name, data):
    path = os.path.join(tmp_path, name)
    with open(path, "wb") as f:
        f.write(data)
    return path


def test_reader_quirks(tmp_path):
    path = _write(tmp_path, "ok.txt", b">ACGT\n<AGGT\n>AA\n<A\n")
    rc, pairs, pl, tl = orc.read_pairs(path, 10, 2)
    assert rc == 1 and pairs == [(b"ACGT", b"AGGT"), (b"AA", b"A")]
    # last line without newline loses its last base
    path = _write(tmp_path, "nonl.txt", b">ACGT\n<AGGT")
    rc, pairs, pl, tl = orc.read_pairs(path, 10, 1)
    assert pairs == [(b"ACGT", b"AGG")]
    # CRLF: '\r' stays as a base
    path = _write(tmp_path, "crlf.txt", b">ACGT\r\n<ACGT\r\n")
    rc, pairs, pl, tl = orc.read_pairs(path, 10, 1)
    assert pairs == [(b"ACGT\r", b"ACGT\r")]
    # wrong marker
    path = _write(tmp_path, "bad.txt", b"<ACGT\n>ACGT\n")
    rc, pairs, pl, tl = orc.read_pairs(path, 10, 1)
    assert rc == 0
    # short file: missing pairs are empty
    path = _write(tmp_path, "short.txt", b">ACGT\n<ACGT\n")
    rc, pairs, pl, tl = orc.read_pairs(path, 10, 3)
    assert rc == 1 and pairs[1:] == [(b"", b""), (b"", b"")]
    # -l 0 reads nothing
    rc, pairs, pl, tl = orc.read_pairs(path, 0, 1)
    assert rc == 1 and pairs == [(b"", b"")]
    # missing file
    rc, pairs, pl, tl = orc.read_pairs(os.path.join(tmp_path, "none.txt"), 10, 1)
    assert rc == -1
