"""A launch policy changes speed only: the same pairs under very different
policies (wave or block per pair, wavefront in registers or LDS, backtrace
in LDS or HBM, budgets small enough to force the deferral passes) give the
same bytes, equal to the oracle's.  The autotuner runs and writes a policy."""
import os

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu

POLICIES = {
    "builtin": None,
    "wave_regs_dcap31": "1200 1 31 1 4\n4000000000 1 31 1 4\n",
    "wave_lds_dcap200": "4000000000 1 200 1 2\n",
    "block4_lds": "4000000000 4 300 1 1\n",
    "block8_hbm_small": "600 1 63 1 8\n4000000000 8 96 0 1\n",
    "block16_hbm_large": "4000000000 16 1048576 0 1\n",
    "block2_hbm": "2000 2 128 0 2\n4000000000 4 700 0 1\n",
}


def run_policy(w, text, tmp_path, name):
    path = None
    if text is not None:
        path = os.path.join(tmp_path, name + ".policy")
        with open(path, "w") as f:
            f.write(text)
    with ewfa.Context(policy_path=path) as ctx:
        return ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)


def test_policies_give_identical_bytes(tmp_path):
    w = ewfa.Workload(ewfa.synth_spec(kind=1, seed=11, len_lo=100, len_hi=4000, err_lo=0.01, err_hi=0.10), 400)
    ok, d, sl, st = orc.batch_workload(w)
    ref = None
    for name, text in POLICIES.items():
        print(f"case policy {name}", flush=True)
        res, cig = run_policy(w, text, tmp_path, name)
        streams = ewfa.streams_from(res, w.pairs, cig)
        assert (res["status"] == ewfa.PAIR_OK).all(), name
        assert np.array_equal(res["distance"], d), name
        assert all(np.array_equal(a, b) for a, b in zip(streams, st)), name
        if ref is None:
            ref = res
        assert np.array_equal(ref, res), name


def test_autotune_writes_policy(tmp_path):
    w = ewfa.Workload(ewfa.synth_spec(kind=1, seed=12, len_lo=100, len_hi=3000, err_lo=0.01, err_hi=0.08), 300)
    out = os.path.join(tmp_path, "tuned.policy")
    print("case autotune 300 pairs, 3 buckets", flush=True)
    report = ewfa.autotune(w.words, w.pairs, w.n, [500, 1500, 3000], out, reps=1)
    assert "chosen" in report
    lines = [l for l in open(out) if l.strip() and not l.startswith("#")]
    assert len(lines) == 3
    with ewfa.Context(policy_path=out) as ctx:
        res, cig = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    ok, d, sl, st = orc.batch_workload(w)
    assert np.array_equal(res["distance"], d)
