"""Committed golden fixtures (tests/golden/*.json, written by make_golden.py):
the oracle and the product's host text reproduce them on the CPU; the HIP
path reproduces them on the GPU."""
import glob
import json
import os

import numpy as np
import pytest

import ewfa
import oracle_py as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.json")))


def load(path):
    with open(path) as f:
        return json.load(f)


def test_fixtures_present():
    names = {os.path.basename(f) for f in FILES}
    assert {"readme_pairs.json", "kats.json", "c1_sample.json", "c2_sample.json", "edge.json",
            "cap64.json"} <= names


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_fixture(path):
    doc = load(path)
    md = doc["max_distance"]
    for c in doc["cases"]:
        p, t = c["pattern"].encode("latin-1"), c["text"].encode("latin-1")
        d, s = orc.stream(p, t, md)
        assert d == c["distance"] and s == c["stream"], c
        assert orc.text(s) == c["printed"]
        assert ewfa.cigar_text(s) == c["printed"]


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_gpu_reproduces_fixture(path):
    doc = load(path)
    md = doc["max_distance"]
    cases = doc["cases"]
    print(f"case golden {os.path.basename(path)} ({len(cases)} pairs, cap {md})", flush=True)
    P = [c["pattern"].encode("latin-1") for c in cases]
    T = [c["text"].encode("latin-1") for c in cases]
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    for mode in (ewfa.MODE_PERSISTENT, ewfa.MODE_PLAIN):
        with ewfa.Context(mode=mode, max_distance=md) as ctx:
            res, cig = ctx.align_host(words, pairs, len(P), cb)
        got = ewfa.streams_from(res, pairs, cig)
        for i, c in enumerate(cases):
            want = ewfa.PAIR_CAPPED if c["capped"] else ewfa.PAIR_OK
            if c["undefined"]:
                want = ewfa.PAIR_UNDEFINED
            assert int(res[i]["status"]) == want, (mode, i)
            assert int(res[i]["distance"]) == c["distance"], (mode, i)
            if not c["undefined"]:
                assert got[i].tolist() == c["stream"], (mode, i)
