#!/usr/bin/env python3
"""Writes the golden fixtures in this folder from the oracle.

Inputs: the three sample pairs of the reference's README (README.md:62-67,
read from the reference tree when present, else kept from the existing
fixture), the hand-derived known answers, samples of the synthetic workloads
and edge cases.  Expected outputs: distance, compressed stream and the
reference's printed text, from the oracle (oracle/ewfa_oracle.c).

    python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "ewfa-gpu_amd"))

import ewfa  # noqa: E402
import oracle_py as orc  # noqa: E402

REFERENCE_README = "/root/reference/README.md"


def readme_pairs():
    out = os.path.join(HERE, "readme_pairs.json")
    if os.path.exists(REFERENCE_README):
        lines = [l.strip() for l in open(REFERENCE_README) if l.startswith((">", "<"))]
        pats = [l[1:] for l in lines if l.startswith(">")]
        txts = [l[1:] for l in lines if l.startswith("<")]
        return list(zip(pats, txts))
    with open(out) as f:
        return [(c["pattern"], c["text"]) for c in json.load(f)["cases"]]


def entry(p, t, max_distance=0):
    pb = p.encode() if isinstance(p, str) else p
    tb = t.encode() if isinstance(t, str) else t
    d, s = orc.stream(pb, tb, max_distance)
    _, _, capped, undefined = orc.align(pb, tb, max_distance)
    return {"pattern": pb.decode("latin-1"), "text": tb.decode("latin-1"), "distance": d, "stream": s,
            "printed": orc.text(s), "capped": capped, "undefined": undefined}


def write(name, cases, max_distance=0, note=""):
    doc = {"max_distance": max_distance, "note": note, "cases": [entry(p, t, max_distance) for p, t in cases]}
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(doc, f, indent=None, separators=(",", ":"))
        f.write("\n")
    print(name, len(cases))


def main():
    write("readme_pairs.json", readme_pairs(), note="the three sample pairs of the reference README (README.md:62-67)")
    kats = [("ACGT", "ACGT"), ("A", ""), ("AA", "A"), ("ACGT", "AGGT"), ("", "A"), ("A", "C"), ("AC", "A"),
            ("CA", "A"), ("AC", "CA"), ("", ""), ("T" * 300, "T" * 300)]
    write("kats.json", kats, note="hand-derived known answers")
    w1 = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c1"]), 200, threads=1)
    write("c1_sample.json", [(w1.pattern(i), w1.text(i)) for i in range(w1.n)],
          note="first 200 pairs of the 150 bp / 2% workload (seed 1)")
    w2 = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c2"]), 24, first=123456, threads=1)
    write("c2_sample.json", [(w2.pattern(i), w2.text(i)) for i in range(w2.n)],
          note="pairs 123456.. of the 1000 bp / 5% workload (seed 2)")
    edge = [("NNNNACGT", "GGGGACGT"), ("acgtacgt", "ACGTACGT"), ("ACGT\r", "ACGT\r"), ("U" * 40, "T" * 40),
            ("A" * 200, "A" * 450), ("CA" * 70, "AC" * 70), ("A" * 10 + "C" * 10, "C" * 10 + "A" * 10),
            ("ACGT" * 16, "ACGT" * 8), ("G" * 33, ""), ("", "T" * 65)]
    write("edge.json", edge, note="case, N, U, CR bytes, long runs, empty sides")
    wc = ewfa.Workload(ewfa.synth_spec(kind=1, seed=9, len_lo=300, len_hi=3000, err_lo=0.02, err_hi=0.10), 40,
                       threads=1)
    write("cap64.json", [(wc.pattern(i), wc.text(i)) for i in range(wc.n)], max_distance=64,
          note="the reference's distance cap of 64 on pairs up to 10% error")


if __name__ == "__main__":
    main()
