"""The banded one-chunk register wavefront (passes of at most 63 levels).

Level d keeps only the diagonals that can still reach the target diagonal
within the pass's levels; the results must be those of the full wavefront
on every pair: the oracle's distance and stream for pairs inside the budget,
the deferral passes' for pairs beyond it, and the reference's capped output
under a cap.  The cases put the target diagonal at and around the budget
(pure insertions or deletions at either end, so |tk| = D), mix them with
random pairs, and compare against the oracle and against a policy that
runs the unbanded multi-chunk kernel.
"""
import os
import random
import sys

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def policy_file(tmp_path, name, text):
    path = os.path.join(tmp_path, name + ".policy")
    with open(path, "w") as f:
        f.write(text)
    return path


def rand_seq(rng, n):
    return bytes(rng.choice(b"ACGT") for _ in range(n))


def hand_pairs(budgets, seed=5):
    """Pairs whose distance and target diagonal sit at and around each budget:
    a tail or head added or removed (|tk| = D), the same with substitutions
    inside (|tk| < D), and equal-length pairs with D substitutions (tk = 0)."""
    rng = random.Random(seed)
    P, T = [], []
    for b in budgets:
        for t in (max(b - 1, 1), b, b + 1, b + 2):
            p = rand_seq(rng, 200 + rng.randrange(300))
            tail = rand_seq(rng, t)
            P += [p, p + tail, p, tail + p]
            T += [p + tail, p, tail + p, p]
            q = bytearray(p)
            for pos in rng.sample(range(len(q)), min(t // 2, len(q))):
                q[pos] = ord("A") if q[pos] != ord("A") else ord("C")
            P += [bytes(q)]
            T += [p + rand_seq(rng, max(t - t // 2, 0))]
            q = bytearray(p)
            for pos in rng.sample(range(len(q)), min(t, len(q))):
                q[pos] = ord("G") if q[pos] != ord("G") else ord("T")
            P += [bytes(q)]
            T += [p]
    return P, T


def check(res, pairs, cig, P, T, label, max_distance=0):
    streams = ewfa.streams_from(res, pairs, cig)
    for i, (p, t) in enumerate(zip(P, T)):
        d, s = orc.stream(p, t, max_distance=max_distance) if max_distance else orc.stream(p, t)
        ok = (int(res[i]["distance"]) == d and streams[i].tolist() == s)
        if max_distance == 0:
            ok = ok and int(res[i]["status"]) == ewfa.PAIR_OK
        assert ok, (label, i, len(p), len(t), int(res[i]["status"]), int(res[i]["distance"]), d,
                    streams[i][:24].tolist(), s[:24])


@pytest.mark.parametrize("dcap", [5, 15, 31, 47, 63])
def test_band_budgets_hand_pairs(tmp_path, dcap):
    P, T = hand_pairs([dcap, dcap // 2 + 1, 1])
    say(f"case band dcap={dcap} hand pairs {len(P)}")
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    path = policy_file(tmp_path, f"band{dcap}", f"4000000000 1 {dcap} 1 4\n")
    with ewfa.Context(policy_path=path) as ctx:
        res, cig = ctx.align_host(words, pairs, len(P), cb)
    check(res, pairs, cig, P, T, f"band{dcap}")


def test_band_matches_unbanded_and_oracle(tmp_path):
    """Random mixed pairs: the banded budgets and the three-chunk kernel give
    the same bytes, equal to the oracle's (pairs beyond a budget go through
    the deferral passes)."""
    w = ewfa.Workload(ewfa.synth_spec(kind=1, seed=21, len_lo=100, len_hi=1500, err_lo=0.01, err_hi=0.10), 600)
    ok, d, sl, st = orc.batch_workload(w)
    ref = None
    for name, text in (("band63", "4000000000 1 63 1 4\n"), ("band31", "4000000000 1 31 1 16\n"),
                       ("band9", "4000000000 1 9 1 1\n"), ("chunks3", "4000000000 1 95 1 4\n")):
        say(f"case band mixed 600 pairs policy {name}")
        path = policy_file(tmp_path, name, text)
        with ewfa.Context(policy_path=path) as ctx:
            res, cig = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
        streams = ewfa.streams_from(res, w.pairs, cig)
        assert (res["status"] == ewfa.PAIR_OK).all(), name
        assert np.array_equal(res["distance"], d), name
        assert all(np.array_equal(a, b) for a, b in zip(streams, st)), name
        if ref is None:
            ref = (res, streams)
        assert np.array_equal(ref[0], res), name


@pytest.mark.parametrize("cap", [8, 33, 64])
def test_band_under_reference_cap(tmp_path, cap):
    """With a cap the band spans cap-1 levels; capped and undefined pairs
    come out as the reference's."""
    P, T = hand_pairs([cap - 1, cap, cap + 3], seed=cap)
    say(f"case band cap={cap} hand pairs {len(P)}")
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    path = policy_file(tmp_path, "band63", "4000000000 1 63 1 4\n")
    with ewfa.Context(policy_path=path, max_distance=cap) as ctx:
        res, cig = ctx.align_host(words, pairs, len(P), cb)
    check(res, pairs, cig, P, T, f"cap{cap}", max_distance=cap)
    st = res["status"]
    assert ((st == ewfa.PAIR_CAPPED) | (st == ewfa.PAIR_UNDEFINED)).any()


def test_band_long_runs(tmp_path):
    """Runs at the byte limits: the band kernel keeps only a run's low byte in
    its scratch and sends a pair with a run of 255 or more to the emit
    kernel; streams must equal the oracle's on both sides of the limit."""
    rng = random.Random(77)
    P, T = [], []
    for run in (200, 254, 255, 256, 257, 300, 510, 511, 512, 513, 767, 768, 1000):
        for edits in (1, 3):
            head = rand_seq(rng, 20)
            body = rand_seq(rng, run)
            p = head + body
            t = bytearray(p)
            for pos in rng.sample(range(len(head)), edits):
                t[pos] = ord("A") if t[pos] != ord("A") else ord("C")
            P.append(p)
            T.append(bytes(t))
            P.append(body + head)
            T.append(body + rand_seq(rng, 5) + head)
    say(f"case band long runs {len(P)} pairs")
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    path = policy_file(tmp_path, "band63", "4000000000 1 63 1 4\n")
    with ewfa.Context(policy_path=path) as ctx:
        res, cig = ctx.align_host(words, pairs, len(P), cb)
    check(res, pairs, cig, P, T, "longruns")
