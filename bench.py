#!/usr/bin/env python3
"""Benchmark of the WFA edit aligner on one MI355X.

Workload: 1,000,000 pairs of 1,000 bp, uniform ACGT, 5% edits split equally
between substitution, insertion and deletion (seed 2), 2-bit packed uint64
input resident in HBM, the reference's per-pair CIGAR buffer of 1,004 bytes.
One step = distances and full CIGAR streams for all 1,000,000 pairs.

value      : alignments/s of the persistent path (one hipGraph-launched
             persistent kernel over a device work queue, then the emit kernel)
plain      : alignments/s of the reference's launch model (batches of 100,000,
             one workgroup per pair, synchronised per batch) on the same input
roofline   : algorithmic bytes of the path per launch / average duration of
             the align kernel (HIP events on its launch stream)
cpu_baseline: the oracle (a C restatement of the reference's algorithm with
             its CPU validity check) on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ewfa-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "alignments/s with CIGAR @1000bp/5% err (batched vs plain); served p50/p99"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=100_000)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=3000, help="pairs checked against the oracle")
    ap.add_argument("--profile-run", action="store_true",
                    help="one eager persistent pass only (for rocprofv3)")
    ap.add_argument("--profile-copy", action="store_true",
                    help="one 1 GiB device copy only (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(pairs are independent; see DESIGN.md for how the work shards)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import ewfa

    assert ewfa.device_count() > 0, "bench needs a GPU"

    if args.profile_copy:
        x = ewfa.DeviceBuffer.from_array(np.ones(1 << 27, dtype=np.int64))
        y = ewfa.DeviceBuffer(x.nbytes)
        ewfa.synchronize()
        ewfa.lib().ewfa_memcpy(y.ptr, x.ptr, x.nbytes, 2)
        ewfa.synchronize()
        log("copied 1 GiB")
        return

    threads = min(16, os.cpu_count() or 8)
    spec_d = dict(ewfa.CONFIGS["c2"])
    spec = ewfa.synth_spec(**spec_d)
    n = args.pairs
    t0 = time.time()
    w = ewfa.Workload(spec, n, cigar_cap=1004, threads=threads)
    log(f"generated {n} pairs in {time.time() - t0:.1f}s, max_len {w.max_len}")

    dw = ewfa.DeviceBuffer.from_array(w.words)
    dp = ewfa.DeviceBuffer.from_array(w.pairs[:n])
    dr = ewfa.DeviceBuffer(n * 16)
    dc = ewfa.DeviceBuffer(w.cigar_bytes)
    ewfa.synchronize()

    def run(ctx):
        ctx.align_device(dw.ptr, dp.ptr, n, w.max_len, dr.ptr, dc.ptr, None)

    if args.profile_run:
        with ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=ewfa.FLAG_NO_GRAPH) as ctx:
            run(ctx)
            ewfa.synchronize()
            run(ctx)
            ewfa.synchronize()
        log("profile run done")
        return

    # persistent path (the value)
    ctx = ewfa.Context(mode=ewfa.MODE_PERSISTENT)
    for i in range(args.warmup):
        run(ctx)
        log(f"warmup {i + 1}/{args.warmup}")
    ewfa.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(ctx)
    ewfa.synchronize()
    dt = time.perf_counter() - t0
    value = args.steps * n / dt
    log(f"persistent: {args.steps} steps in {dt:.3f}s -> {value / 1e6:.2f} M alignments/s")

    # per-kernel durations from HIP events on the launch stream
    ctx_t = ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=ewfa.FLAG_TIMING)
    a_ms, e_ms = [], []
    for _ in range(3):
        run(ctx_t)
        a, e = ctx_t.last_timing()
        a_ms.append(a)
        e_ms.append(e)
    ctx_t.close()
    align_ms = float(np.mean(a_ms))
    emit_ms = float(np.mean(e_ms))
    log(f"align kernel {align_ms:.3f} ms, emit kernel {emit_ms:.3f} ms (events)")

    # plain path: the reference's per-batch launch model
    ctx_plain = ewfa.Context(mode=ewfa.MODE_PLAIN, batch_size=args.batch)
    run(ctx_plain)
    ewfa.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        run(ctx_plain)
    ewfa.synchronize()
    pdt = time.perf_counter() - t0
    plain_value = args.plain_steps * n / pdt
    log(f"plain: {plain_value / 1e6:.2f} M alignments/s")
    ctx_plain.close()

    # the persistent result again (the plain run overwrote the buffers)
    run(ctx)
    ewfa.synchronize()
    res = dr.download(ewfa.RESULT_DTYPE, n)
    cig = dc.download(np.int8)
    ctx.close()
    ok_status = int((res["status"] == ewfa.PAIR_OK).sum())
    dist = res["distance"].astype(np.int64)
    clen = res["cigar_len"].astype(np.int64)
    plen = w.plen.astype(np.int64)
    tlen = w.tlen.astype(np.int64)
    algo_bytes = int(((plen + 3) // 4 + (tlen + 3) // 4 + 8 + clen).sum())
    cells = int(((dist + 1) ** 2).sum())

    # parity sample against the oracle
    checked = 0
    mismatches = 0
    if args.check > 0:
        import oracle_py as orc
        idx = np.linspace(0, n - 1, num=min(args.check, n)).astype(np.int64)
        ok, od, osl, ost = orc.batch_workload(w, idx=idx, threads=threads)
        got = ewfa.streams_from(res[idx], w.pairs[idx], cig)
        for j, i in enumerate(idx):
            if int(res[i]["distance"]) != int(od[j]) or not np.array_equal(got[j], ost[j]):
                mismatches += 1
        checked = len(idx)
        log(f"oracle check: {checked} pairs, {mismatches} mismatches")

    # CPU baseline: the oracle on a bounded sample of the same workload
    cpu = None
    if not args.no_cpu:
        import oracle_py as orc
        cores = threads
        done = 0
        t0 = time.perf_counter()
        chunk = 100_000
        while done < args.cpu_sample and time.perf_counter() - t0 < args.cpu_seconds:
            idx = np.arange(done, min(done + chunk, args.cpu_sample, n))
            orc.batch_workload(w, idx=idx, threads=cores)
            done += len(idx)
        cdt = time.perf_counter() - t0
        cpu = {"value": round(done / cdt, 1), "unit": "alignments/s", "cores": cores, "kind": "port",
               "sample": f"first {done} pairs of the same workload (1000 bp, 5% edits), oracle align + "
                         f"stream + text + validity check on {cores} host threads, {cdt:.1f}s"}
        log(f"cpu baseline: {cpu['value']:.0f} alignments/s on {cores} threads")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_align_c2.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    achieved = algo_bytes / (align_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "alignments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "c2: 1,000,000 pairs x 1,000 bp, 5% edits (sub/ins/del equal), seed 2, "
                               "2-bit packed u64 input resident in HBM, CIGAR stride 1004 B",
                   "pairs": n, "seq_len": 1000, "error": 0.05, "plain_batch": args.batch,
                   "parallelism": "single GPU, persistent kernel"},
        "plain": {"value": round(plain_value, 1), "unit": "alignments/s",
                  "ms_per_step": round(pdt / args.plain_steps * 1e3, 3)},
        "ratio_vs_plain": round(value / plain_value, 3),
        "kernels_ms": {"align": round(align_ms, 4), "emit": round(emit_ms, 4)},
        "cells_per_s": round(cells / (align_ms / 1e3), 1),
        "pairs_ok": ok_status,
        "oracle_checked": checked,
        "oracle_mismatches": mismatches,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
