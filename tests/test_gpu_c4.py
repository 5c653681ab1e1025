"""Long reads (10-50 kbp at 10-15% edits, 30% of the edit budget as indel
runs) through the default policy, bit for bit against the oracle.

Pairs of this length take the built-in table's two long-read buckets (up to
25 kbp and beyond, run side by side): an 8-wave block per pair with the
backtrace in HBM, and per-pair distance bounds from the banded bound kernel
that limit the cells the block computes and size its narrow slot.  Every
distance and stream must equal the oracle's (the reference's tie-break,
src/kernels.cu:141-152, on every cell of the path), and distances the
Levenshtein DP.
"""
import functools
import os
import sys

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 8))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def c4_pairs(firsts, per):
    """`per` consecutive pairs of the 20,000-pair long-read workload from each
    index in `firsts`, laid out as one batch."""
    spec = ewfa.synth_spec(**ewfa.CONFIGS["c4"])
    pats, txts = [], []
    for f in firsts:
        w = ewfa.Workload(spec, per, first=f)
        pats += [w.pattern(i) for i in range(per)]
        txts += [w.text(i) for i in range(per)]
    return pats, txts


FIRSTS = (0, 6_000, 13_000, 19_992)


@functools.lru_cache(maxsize=1)
def oracle_answers():
    pats, txts = c4_pairs(FIRSTS, 8)
    return orc.batch_streams(pats, txts, threads=THREADS)


@pytest.mark.parametrize("path", ["device", "host"])
def test_c4_long_reads_default_policy(path):
    say(f"case c4 {4 * 8} pairs of 10-50 kbp (from {FIRSTS}) default policy, {path} path")
    pats, txts = c4_pairs(FIRSTS, 8)
    words, pairs, cigar_bytes, max_len = ewfa.from_sequences(pats, txts)
    n = len(pats)
    assert max_len > 10_500  # the long-read bucket runs
    if path == "device":
        dw = ewfa.DeviceBuffer.from_array(words)
        dp = ewfa.DeviceBuffer.from_array(pairs[:n])
        dr = ewfa.DeviceBuffer(n * 16)
        dc = ewfa.DeviceBuffer(cigar_bytes)
        with ewfa.Context() as ctx:
            ctx.align_device(dw.ptr, dp.ptr, n, max_len, dr.ptr, dc.ptr, None)
        ewfa.synchronize()
        res = dr.download(ewfa.RESULT_DTYPE, n)
        cig = dc.download(np.int8)
    else:
        with ewfa.Context() as ctx:
            res, cig = ctx.align_host(words, pairs, n, cigar_bytes)
    say(f"  gpu done, distances {int(res['distance'].min())}..{int(res['distance'].max())}; oracle on {THREADS} threads")
    ok, od, osl, ost = oracle_answers()
    assert ok == n  # every oracle stream passes the reference's validity check
    streams = ewfa.streams_from(res, pairs, cig)
    bad = [i for i in range(n) if int(res[i]["status"]) != ewfa.PAIR_OK or int(res[i]["distance"]) != int(od[i])
           or not np.array_equal(streams[i], ost[i])]
    assert not bad, (f"{len(bad)} of {n} long pairs differ; first {bad[0]}: status {int(res[bad[0]]['status'])} "
                     f"dist {int(res[bad[0]]['distance'])} vs {int(od[bad[0]])}")
    # the distance is also the Levenshtein distance (shortest pairs: O(nm))
    for i in np.argsort([len(p) + len(t) for p, t in zip(pats, txts)])[:2]:
        assert int(res[i]["distance"]) == orc.levenshtein(pats[i], txts[i])
