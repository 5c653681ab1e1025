"""Block passes that leave out cells beyond a distance bound.

A block pass with a budget of 2,048 levels or more first bounds each pair's
distance with a 64-diagonal band that follows the furthest-reaching cell, and
then computes only the cells (d, k) with d + |k - tk| <= bound.  The results
must be those of the full wavefront (src/kernels.cu:111-180, the tie-break at
:141-152), so every case here is checked against the oracle bit for bit:
long pairs with indel runs (the band must follow the path's jumps), pairs
whose distance exceeds the reference cap (the band fails and the pass runs
unbounded), low-complexity pairs full of equal-cost paths, each wave count
of the block kernel (a level's first computed chunk can start inside a
chunk, where lane 0 is outside the range), and distances above a bounded
pass's budget (those pairs finish in the deferral passes, bounded again).
Bucket passes keep only the diagonals within the bound in narrow slots
(more blocks per CU); a pair whose bound exceeds the slot searches up to the
slot's size and is deferred when its distance is larger, and a pair whose
length difference exceeds it is deferred before it is staged.
"""
import numpy as np
import pytest

import ewfa
import oracle_py as orc
from test_gpu_stress import oracle_streams, stress_pairs

pytestmark = pytest.mark.gpu

BOUNDED = {
    "block1": "4000000000 1 2048 0 1\n",
    "block4": "4000000000 4 2048 0 1\n",
    "block16": "4000000000 16 2048 0 1\n",
}


def check(res, pairs, cig, d, st, label, capped_ok=False):
    streams = ewfa.streams_from(res, pairs, cig)
    ok_status = (ewfa.PAIR_OK, ewfa.PAIR_CAPPED, ewfa.PAIR_UNDEFINED) if capped_ok else (ewfa.PAIR_OK,)
    for i in range(len(res)):
        good = (int(res[i]["status"]) in ok_status and int(res[i]["distance"]) == int(d[i])
                and np.array_equal(streams[i], st[i]))
        assert good, (label, i, int(res[i]["status"]), int(res[i]["distance"]), int(d[i]),
                      streams[i][:32].tolist(), st[i][:32].tolist())


def policy_file(tmp_path, name, text):
    path = str(tmp_path / f"{name}.policy")
    with open(path, "w") as f:
        f.write(text)
    return path


@pytest.mark.parametrize("name", sorted(BOUNDED))
def test_long_pairs_with_indel_runs(tmp_path, name):
    # 2,000..5,000 bp at 10-15% edits, 30% of them as indel runs: distances
    # of about 200..700 levels
    print(f"case bounded {name}: 120 pairs 2000..5000 bp 10-15% with indel runs", flush=True)
    w = ewfa.Workload(ewfa.synth_spec(kind=2, seed=41, len_lo=2000, len_hi=5000, err_lo=0.10, err_hi=0.15), 120)
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == w.n
    with ewfa.Context(policy_path=policy_file(tmp_path, name, BOUNDED[name])) as ctx:
        res, cig = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    check(res, w.pairs, cig, d, st, name)


def test_low_complexity_pairs(tmp_path):
    P, T = stress_pairs(303, 600)
    print(f"case bounded low-complexity {len(P)} pairs", flush=True)
    d, st = oracle_streams(P, T)
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    for name in ("block4", "block16"):
        with ewfa.Context(policy_path=policy_file(tmp_path, name, BOUNDED[name])) as ctx:
            res, cig = ctx.align_host(words, pairs, len(P), cb)
        check(res, pairs, cig, d, st, name)


def test_reference_cap(tmp_path):
    # with the cap of 64 the band searches 63 levels: pairs beyond it get no
    # bound and keep the capped walk from level 63
    print("case bounded cap 64: 300 pairs 300..3000 bp 2-10%", flush=True)
    w = ewfa.Workload(ewfa.synth_spec(kind=1, seed=42, len_lo=300, len_hi=3000, err_lo=0.02, err_hi=0.10), 300)
    ok, d, sl, st = orc.batch_workload(w, max_distance=64)  # ok counts the uncapped pairs
    assert 0 < ok < w.n
    with ewfa.Context(max_distance=64, policy_path=policy_file(tmp_path, "block4", BOUNDED["block4"])) as ctx:
        res, cig = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    check(res, w.pairs, cig, d, st, "cap64", capped_ok=True)
    assert any(int(r["status"]) == ewfa.PAIR_CAPPED for r in res)


def test_distances_above_the_budget(tmp_path):
    # budget 2,048 on 16 waves; pairs of 8-10 kbp at 25-35% edits have
    # distances of about 2,000-2,700, so most are deferred and finish in
    # the deferral passes
    print("case bounded beyond budget: 16 pairs 8000..10000 bp 25-35%", flush=True)
    w = ewfa.Workload(ewfa.synth_spec(kind=1, seed=43, len_lo=8000, len_hi=10000, err_lo=0.25, err_hi=0.35), 16)
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == w.n
    assert (d > 2048).any() and (d <= 2048).any()
    with ewfa.Context(policy_path=policy_file(tmp_path, "block16", BOUNDED["block16"])) as ctx:
        res, cig = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    check(res, w.pairs, cig, d, st, "beyond-budget")


def test_length_difference_beyond_a_narrow_slot(tmp_path):
    # a budget of 2,048 levels gives the bucket pass narrow slots of 2,048
    # diagonals; pairs whose lengths differ by more (the target diagonal lies
    # outside every slot) are deferred unstaged and finish in the deferral
    # passes, beside pairs that fit
    rng = np.random.default_rng(44)
    P, T = [], []
    for i in range(12):
        n = int(rng.integers(2500, 3500))
        p = rng.integers(0, 4, n)
        extra = int(rng.integers(2100, 2600)) if i % 2 == 0 else int(rng.integers(0, 300))
        t = np.concatenate([p[: n // 2], rng.integers(0, 4, extra), p[n // 2:]])
        if i % 4 in (1, 2):
            p, t = t, p  # the pattern the longer one
        P.append("".join("ACGT"[x] for x in p).encode())
        T.append("".join("ACGT"[x] for x in t).encode())
    print(f"case narrow slot, length differences to 2,600: {len(P)} pairs", flush=True)
    d, st = oracle_streams(P, T)
    assert (d > 2048).any() and (d <= 2048).any()
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    with ewfa.Context(policy_path=policy_file(tmp_path, "block8", "4000000000 8 2048 0 1\n")) as ctx:
        res, cig = ctx.align_host(words, pairs, len(P), cb)
    check(res, pairs, cig, d, st, "narrow-tk")
