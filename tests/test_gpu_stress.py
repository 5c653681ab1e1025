"""Low-complexity and repeat-rich pairs, where many paths tie and the
reference's tie-break (DEL over SUB over INS at equal offsets,
src/kernels.cu:141-152) decides every op: homopolymer blocks, tandem
repeats of unit 1..9 with whole-unit indels, random sequence with embedded
repeats, and lengths from 0 to 2,600 bases so that every kernel family
(banded and multi-chunk register wavefronts, block kernels with the
backtrace in LDS or HBM, the deferral passes) takes some of them.  Every
pair's distance and stream must equal the oracle's under each policy.
"""
import random
import types

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu

POLICIES = {
    "builtin": None,
    "band63": "4000000000 1 63 1 4\n",
    "wave_chunks_dcap127": "4000000000 1 127 1 4\n",
    "block4_lds": "1200 1 63 1 4\n4000000000 4 300 1 1\n",
    "block2_hbm": "4000000000 2 512 0 1\n",
}

BASES = b"ACGT"


def low_complexity(rng, n):
    kind = rng.randrange(4)
    out = bytearray()
    if kind == 0:  # homopolymer blocks
        while len(out) < n:
            out += bytes([rng.choice(BASES)]) * rng.randrange(1, 300)
    elif kind == 1:  # one tandem repeat
        unit = bytes(rng.choice(BASES) for _ in range(rng.randrange(1, 10)))
        out = bytearray(unit * (n // len(unit) + 1))
    elif kind == 2:  # random with embedded repeats
        while len(out) < n:
            if rng.random() < 0.5:
                out += bytes(rng.choice(BASES) for _ in range(rng.randrange(1, 80)))
            else:
                unit = bytes(rng.choice(BASES) for _ in range(rng.randrange(1, 7)))
                out += unit * rng.randrange(2, 60)
    else:  # two-base alphabet
        a, b = rng.sample(list(BASES), 2)
        out = bytearray(rng.choice((a, b)) for _ in range(n))
    return bytes(out[:n])


def mutate(rng, s, rate):
    s = bytearray(s)
    for _ in range(int(round(rate * len(s)))):
        op = rng.randrange(4)
        pos = rng.randrange(len(s) + 1)
        if op == 0 and pos < len(s):
            s[pos] = rng.choice([b for b in BASES if b != s[pos]])
        elif op == 1:
            s[pos:pos] = bytes([rng.choice(BASES)])
        elif op == 2 and pos < len(s):
            del s[pos]
        elif len(s) > 0:  # a whole repeat unit shifted: a run of equal-cost paths
            u = rng.randrange(1, 10)
            if rng.random() < 0.5:
                s[pos:pos] = s[max(pos - u, 0):pos]
            else:
                del s[pos:pos + u]
    return bytes(s)


def stress_pairs(seed, n):
    rng = random.Random(seed)
    P, T = [], []
    for _ in range(n):
        r = rng.random()
        length = rng.randrange(0, 40) if r < 0.1 else rng.randrange(40, 1300) if r < 0.8 else \
            rng.randrange(1300, 2600)
        p = low_complexity(rng, length)
        t = mutate(rng, p, rng.choice((0.0, 0.01, 0.03, 0.08, 0.15)))
        if rng.random() < 0.5:
            p, t = t, p
        P.append(p)
        T.append(t)
    return P, T


def oracle_streams(P, T):
    """The oracle's distances and streams for lists of byte strings."""
    n = len(P)
    stride = max(max((len(x) for x in P + T), default=0), 1)
    pat = np.zeros(n * stride + 1, dtype=np.uint8)
    txt = np.zeros(n * stride + 1, dtype=np.uint8)
    for i, (p, t) in enumerate(zip(P, T)):
        pat[i * stride:i * stride + len(p)] = np.frombuffer(p, dtype=np.uint8)
        txt[i * stride:i * stride + len(t)] = np.frombuffer(t, dtype=np.uint8)
    w = types.SimpleNamespace(n=n, stride=stride, pat=pat, txt=txt,
                              plen=np.array([len(x) for x in P], dtype=np.uint32),
                              tlen=np.array([len(x) for x in T], dtype=np.uint32))
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == n
    return d, st


@pytest.mark.parametrize("seed", [101, 202])
def test_low_complexity_pairs_all_policies(tmp_path, seed):
    P, T = stress_pairs(seed, 1500)
    d, st = oracle_streams(P, T)
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    ref = None
    for name, text in POLICIES.items():
        print(f"case stress seed={seed} {len(P)} pairs policy {name}", flush=True)
        path = None
        if text is not None:
            path = str(tmp_path / f"{name}.policy")
            with open(path, "w") as f:
                f.write(text)
        with ewfa.Context(policy_path=path) as ctx:
            res, cig = ctx.align_host(words, pairs, len(P), cb)
        streams = ewfa.streams_from(res, pairs, cig)
        for i in range(len(P)):
            ok = (int(res[i]["status"]) == ewfa.PAIR_OK and int(res[i]["distance"]) == int(d[i])
                  and np.array_equal(streams[i], st[i]))
            assert ok, (name, i, len(P[i]), len(T[i]), int(res[i]["status"]), int(res[i]["distance"]),
                        int(d[i]), streams[i][:32].tolist(), st[i][:32].tolist())
        if ref is None:
            ref = res
        assert np.array_equal(ref, res), name


def test_sixteen_bit_offsets_near_their_limit(tmp_path):
    """A pass whose max_len + budget is 16,324 keeps its offsets in 16 bits;
    texts of 15,300 bases with a 700-900 base insertion drive doubled
    offsets to within a few hundred of 32,767 (an offset also runs one base a
    level past the text's end).  Results must equal the oracle's."""
    rng = np.random.default_rng(77)
    P, T = [], []
    for i in range(12):
        ins = int(rng.integers(700, 900))
        p = rng.choice(np.frombuffer(BASES, dtype=np.uint8), 15300 - ins)
        t = p.copy()
        sub = rng.random(t.size) < 0.01
        t[sub] = rng.choice(np.frombuffer(BASES, dtype=np.uint8), int(sub.sum()))
        at = int(rng.integers(0, t.size + 1)) if i % 2 else t.size  # inside, or at the end
        t = np.concatenate([t[:at], rng.choice(np.frombuffer(BASES, dtype=np.uint8), ins), t[at:]])
        P.append(p.tobytes())
        T.append(t.tobytes())
    print(f"case 16-bit offsets near the limit: {len(P)} pairs, texts 15300 bp", flush=True)
    d, st = oracle_streams(P, T)
    assert max(d) <= 1024
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    assert ml + 1024 <= 16383
    path = str(tmp_path / "off16.policy")
    with open(path, "w") as f:
        f.write("4000000000 2 1024 0 1\n")
    with ewfa.Context(policy_path=path) as ctx:
        res, cig = ctx.align_host(words, pairs, len(P), cb)
    streams = ewfa.streams_from(res, pairs, cig)
    for i in range(len(P)):
        assert int(res[i]["status"]) == ewfa.PAIR_OK and int(res[i]["distance"]) == int(d[i]), (i, int(res[i]["status"]))
        assert np.array_equal(streams[i], st[i]), i
