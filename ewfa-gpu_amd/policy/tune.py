#!/usr/bin/env python3
"""Regenerates the tuned launch policies in this folder with the library's
autotuner (ewfa_autotune), on the GPU this runs on:

  c2_1000bp_5pct.policy        300,000 pairs of the 1,000 bp / 5% workload
                               (seed 2), one bucket
  c3_mixed_100_10000bp.policy  the 500,000-pair mixed-length workload
                               (log-uniform 100..10,000 bp, 1..10% edits,
                               seed 3), the built-in table's six buckets

Every candidate's results and streams must be byte-identical, so a policy
changes speed only.  The tuner's report (each candidate's time) is printed.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ewfa  # noqa: E402

LAST = 0xFFFFFFFF
JOBS = {
    "c2": ("c2_1000bp_5pct.policy", "c2", 300_000, [LAST]),
    "c3": ("c3_mixed_100_10000bp.policy", "c3", 500_000, [300, 700, 1200, 2500, 5000, LAST]),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--which", default="c2,c3")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out-dir", default=HERE)
    args = ap.parse_args()
    for key in args.which.split(","):
        fname, cfg, n, buckets = JOBS[key]
        w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS[cfg]), n, threads=16)
        out = os.path.join(args.out_dir, fname)
        print(f"== {fname}: {n} pairs, buckets {buckets}", flush=True)
        print(ewfa.autotune(w.words, w.pairs, w.n, buckets, out, reps=args.reps), flush=True)
        print(open(out).read(), flush=True)


if __name__ == "__main__":
    main()
