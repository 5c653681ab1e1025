#!/usr/bin/env python3
"""Benchmark of the WFA edit aligner on one MI355X.

Headline workload (the JSON `value`): 1,000,000 pairs of 1,000 bp, uniform
ACGT, 5% edits split equally between substitution, insertion and deletion
(seed 2), 2-bit packed uint64 input resident in HBM, the reference's
per-pair CIGAR buffer of 1,004 bytes.  One step = distances and full CIGAR
streams for all 1,000,000 pairs on the persistent path (one hipGraph-launched
persistent kernel over a device work queue, then the emit kernel), under the
launch policy the autotuner wrote for this workload.

Beside it, on the same box and run:
  plain      the reference's launch model (batches of 100,000, one workgroup
             per pair, synchronised per batch): kernels over resident input,
             and end to end (per batch: packed words and descriptors host to
             device, align, emit, results and CIGAR buffers back)
  c3         500,000 pairs, log-uniform 100..10,000 bp, 1..10% edits, one
             launch under the autotuned per-bucket policy
  c4         20,000 pairs of 10..50 kbp at 10..15% edits with indel runs
             (backtrace spilled to HBM), with the device footprint
  served     Poisson arrivals of jobs of 1..4,096 pairs (lengths and errors as
             c3) into the job server; p50/p99 latency from arrival to results
             in host memory, and alignments/s against the c3 batched rate
  roofline   algorithmic bytes per align launch / its average duration from
             HIP events recorded (inside the graph) on its launch stream
  cpu        the oracle (a C restatement of the reference's algorithm with its
             CPU validity check) on a bounded sample, and on c1
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
POLICY_C2 = os.path.join(ROOT, "ewfa-gpu_amd", "policy", "c2_1000bp_5pct.policy")
POLICY_C3 = os.path.join(ROOT, "ewfa-gpu_amd", "policy", "c3_mixed_100_10000bp.policy")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algo_bytes(w, res):
    """Algorithmic bytes: 2-bit inputs, two u32 lengths, the compressed
    CIGAR stream per pair."""
    import numpy as np
    plen = w.plen.astype(np.int64)
    tlen = w.tlen.astype(np.int64)
    clen = res["cigar_len"].astype(np.int64)
    return int(((plen + 3) // 4 + (tlen + 3) // 4 + 8 + clen).sum())


def cells(res):
    import numpy as np
    d = res["distance"].astype(np.int64)
    return int(((d + 1) ** 2).sum())


class Resident:
    """A workload's inputs and outputs in device memory."""

    def __init__(self, w):
        import ewfa
        self.w = w
        self.dw = ewfa.DeviceBuffer.from_array(w.words)
        self.dp = ewfa.DeviceBuffer.from_array(w.pairs[:w.n])
        self.dr = ewfa.DeviceBuffer(w.n * 16)
        self.dc = ewfa.DeviceBuffer(w.cigar_bytes)

    def run(self, ctx):
        ctx.align_device(self.dw.ptr, self.dp.ptr, self.w.n, self.w.max_len, self.dr.ptr, self.dc.ptr, None)

    def results(self):
        import ewfa
        return self.dr.download(ewfa.RESULT_DTYPE, self.w.n)

    def cigars(self):
        import numpy as np
        return self.dc.download(np.int8)

    def nbytes(self):
        return self.dw.nbytes + self.dp.nbytes + self.dr.nbytes + self.dc.nbytes


def timed(run, ctx, steps, warmup):
    """Warm-up, then `steps` timed calls bracketed by device synchronisation;
    returns (seconds, [align ms], [emit ms]) with the per-call kernel times."""
    import ewfa
    for _ in range(warmup):
        run(ctx)
    ewfa.synchronize()
    a_ms, e_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        run(ctx)
        a, e = ctx.last_timing()
        a_ms.append(a)
        e_ms.append(e)
    ewfa.synchronize()
    return time.perf_counter() - t0, a_ms, e_ms


def leg_c2(args, threads):
    import numpy as np
    import ewfa
    n = args.pairs
    t0 = time.time()
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c2"]), n, cigar_cap=1004, threads=threads)
    log(f"c2: generated {n} pairs in {time.time() - t0:.1f}s, max_len {w.max_len}")
    r = Resident(w)
    out = {}

    # the value: persistent path, hipGraph replay, timing events in the graph
    ctx = ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=ewfa.FLAG_TIMING, policy_path=POLICY_C2)
    dt, a_ms, e_ms = timed(r.run, ctx, args.steps, args.warmup)
    out["value"] = args.steps * n / dt
    out["dt"] = dt
    out["align_ms"] = float(np.mean(a_ms))
    out["emit_ms"] = float(np.mean(e_ms))
    log(f"c2 persistent: {args.steps} steps in {dt:.3f}s -> {out['value'] / 1e6:.2f} M aln/s; "
        f"align {out['align_ms']:.3f} ms, emit {out['emit_ms']:.3f} ms per step")
    res = r.results()
    cig = r.cigars()
    out["res"], out["cig"], out["w"] = res, cig, w
    out["algo_bytes"] = algo_bytes(w, res)
    out["cells"] = cells(res)

    # plain: the reference's launch model over the resident input
    ctx_p = ewfa.Context(mode=ewfa.MODE_PLAIN, batch_size=args.batch, flags=ewfa.FLAG_TIMING)
    pdt, pa, pe = timed(r.run, ctx_p, args.plain_steps, 1)
    out["plain_kernels"] = args.plain_steps * n / pdt
    out["plain_align_ms"] = float(np.mean(pa))
    out["plain_emit_ms"] = float(np.mean(pe))
    log(f"c2 plain (resident): {out['plain_kernels'] / 1e6:.2f} M aln/s")
    pres = r.results()
    out["plain_equal"] = bool(np.array_equal(pres, res))

    # plain end to end: per batch, host -> device -> host, synchronous
    words_pin = ewfa.PinnedArray(w.words.size, np.uint64)
    words_pin.array[:] = w.words
    res_pin = ewfa.PinnedArray(args.batch, ewfa.RESULT_DTYPE)
    cig_pin = ewfa.PinnedArray(args.batch * 1004, np.int8)
    batches = []
    for a in range(0, n, args.batch):
        b = min(n, a + args.batch)
        wv, sub, cb, _, _ = ewfa.slice_pairs(words_pin.array, w.pairs, a, b)
        pp = ewfa.PinnedArray(len(sub), ewfa.PAIR_DTYPE)
        pp.array[:] = sub
        batches.append((wv, pp, b - a, cb))

    def e2e(ctx_):
        for wv, pp, nb, cb in batches:
            ctx_.align_host(wv, pp.array, nb, cb, results=res_pin.array, cigars=cig_pin.array)

    ctx_pe = ewfa.Context(mode=ewfa.MODE_PLAIN, batch_size=args.batch)
    e2e(ctx_pe)
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        e2e(ctx_pe)
    out["plain_e2e"] = args.plain_steps * n / (time.perf_counter() - t0)
    log(f"c2 plain end to end: {out['plain_e2e'] / 1e6:.2f} M aln/s")
    ctx_pe.close()

    # persistent end to end: the whole job host -> device -> host in one call
    res_all = ewfa.PinnedArray(n, ewfa.RESULT_DTYPE)
    cig_all = ewfa.PinnedArray(w.cigar_bytes, np.int8)
    pairs_pin = ewfa.PinnedArray(n, ewfa.PAIR_DTYPE)
    pairs_pin.array[:] = w.pairs[:n]
    ctx_e = ewfa.Context(mode=ewfa.MODE_PERSISTENT, policy_path=POLICY_C2)
    run_e = lambda c: c.align_host(words_pin.array, pairs_pin.array, n, w.cigar_bytes,
                                   results=res_all.array, cigars=cig_all.array)
    run_e(ctx_e)
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        run_e(ctx_e)
    out["persistent_e2e"] = args.plain_steps * n / (time.perf_counter() - t0)
    log(f"c2 persistent end to end: {out['persistent_e2e'] / 1e6:.2f} M aln/s")
    out["e2e_equal"] = bool(np.array_equal(res_all.array, res))
    for c in (ctx, ctx_p, ctx_e):
        c.close()
    return out


def leg_c3(args, threads):
    import numpy as np
    import ewfa
    n = args.c3_pairs
    t0 = time.time()
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c3"]), n, threads=threads)
    log(f"c3: generated {n} pairs in {time.time() - t0:.1f}s, max_len {w.max_len}")
    r = Resident(w)
    ctx = ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=ewfa.FLAG_TIMING, policy_path=POLICY_C3)
    dt, a_ms, e_ms = timed(r.run, ctx, args.c3_steps, 1)
    ctx.close()
    res = r.results()
    out = {"value": args.c3_steps * n / dt, "align_ms": float(np.mean(a_ms)), "emit_ms": float(np.mean(e_ms)),
           "ms_per_launch": dt / args.c3_steps * 1e3, "res": res, "cig": r.cigars(), "w": w}
    out["cells"] = cells(res)
    out["algo_bytes"] = algo_bytes(w, res)
    log(f"c3: {out['value'] / 1e6:.2f} M aln/s, align {out['align_ms']:.2f} ms, emit {out['emit_ms']:.2f} ms")
    return out


def leg_c4(args, threads):
    import numpy as np
    import ewfa
    n = args.c4_pairs
    t0 = time.time()
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c4"]), n, threads=threads)
    log(f"c4: generated {n} pairs in {time.time() - t0:.1f}s, max_len {w.max_len}")
    r = Resident(w)
    ctx = ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=ewfa.FLAG_TIMING)
    dt, a_ms, e_ms = timed(r.run, ctx, 1, 1)
    footprint = r.nbytes() + ctx.memory_bytes()
    ctx.close()
    res = r.results()
    out = {"value": n / dt, "align_ms": float(np.mean(a_ms)), "emit_ms": float(np.mean(e_ms)),
           "seconds": dt, "device_bytes": footprint, "res": res, "cig": r.cigars(), "w": w}
    out["cells"] = cells(res)
    log(f"c4: {out['value']:.0f} aln/s ({dt:.3f}s), device footprint {footprint / 2**30:.2f} GiB")
    return out


def serve_layout(w):
    """Per-pair CIGAR regions of max(plen, tlen)/4 + 16 bytes, dword aligned:
    enough for every stream of a pair with at most 10% edits.  A stream holds
    one byte per edit op plus ceil(run/127) per match run, so at most
    2D + 1 + min(plen, tlen)/127 bytes, and D is at most the edits applied."""
    import numpy as np
    cap = ((np.maximum(w.plen, w.tlen).astype(np.int64) // 4 + 16 + 3) // 4) * 4
    off = np.zeros(w.n, dtype=np.int64)
    off[1:] = np.cumsum(cap)[:-1]
    w.pairs["cigar_cap"][:w.n] = cap
    w.pairs["cigar_off"][:w.n] = off
    w.cigar_bytes = int(cap.sum())


def leg_served(args, threads, batched_rate):
    """Poisson arrivals of jobs into the job server.  A pool of distinct jobs
    (inputs in pinned memory, shared by their submissions) is submitted in a
    random order; each submission writes into one of `serve_ring` output sets
    of its job, and waits only if that set is still in flight.  One untimed
    pass of the pool warms the server first."""
    import numpy as np
    import ewfa
    rng = np.random.default_rng(5)
    sizes = rng.integers(1, 4097, args.serve_pool)
    total = int(sizes.sum())
    spec = ewfa.synth_spec(kind=1, seed=5, len_lo=100, len_hi=10_000, err_lo=0.01, err_hi=0.10)
    t0 = time.time()
    w = ewfa.Workload(spec, total, threads=threads)
    serve_layout(w)
    words_pin = ewfa.PinnedArray(w.words.size, np.uint64)
    words_pin.array[:] = w.words
    jobs = []
    a = 0
    for s in sizes:
        b = a + int(s)
        wv, sub, cb, _, _ = ewfa.slice_pairs(words_pin.array, w.pairs, a, b)
        outs = [(np.zeros(b - a, dtype=ewfa.RESULT_DTYPE), np.zeros(cb, dtype=np.int8))
                for _ in range(args.serve_ring)]
        jobs.append((wv, sub, b - a, cb, outs, a))
        a = b
    log(f"served: pool of {len(jobs)} jobs, {total} pairs, {args.serve_ring} output sets each, "
        f"ready in {time.time() - t0:.1f}s")
    # the same pairs as one batch from host memory (words pinned, descriptors
    # and outputs in ordinary host arrays, as the served jobs have them):
    # the end-to-end batched rate, beside the resident one
    res_b = np.zeros(total, dtype=ewfa.RESULT_DTYPE)
    cig_b = np.zeros(max(w.cigar_bytes, 4), dtype=np.int8)
    with ewfa.Context(mode=ewfa.MODE_PERSISTENT) as ctx:
        ctx.align_host(words_pin.array, w.pairs, total, w.cigar_bytes, results=res_b, cigars=cig_b)
        t1 = time.perf_counter()
        for _ in range(3):
            ctx.align_host(words_pin.array, w.pairs, total, w.cigar_bytes, results=res_b, cigars=cig_b)
        batched_e2e = 3 * total / (time.perf_counter() - t1)
    log(f"served: the pool as one batch from host memory: {batched_e2e / 1e6:.2f} M aln/s")
    mean_pairs = total / len(jobs)
    lam = args.serve_load * batched_rate / mean_pairs
    k = max(len(jobs), int(args.serve_seconds * lam))
    order = np.concatenate([rng.permutation(len(jobs)) for _ in range((k + len(jobs) - 1) // len(jobs))])[:k]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, k))
    log(f"served: lambda {lam:.0f} jobs/s ({args.serve_load:.2f} x batched {batched_rate / 1e6:.2f} M aln/s), "
        f"{k} submissions")

    srv = ewfa.Server(slots=args.serve_slots, arena_bytes=args.serve_arena << 30,
                      max_batch_pairs=args.serve_batch, max_batch_cigar_bytes=args.serve_batch_bytes << 20,
                      flags=ewfa.FLAG_SERIAL if args.serve_serial else 0)
    # warm-up: every job once, into its last output set
    warm = [srv.submit(wv, sub, nb, cb, results=outs[-1][0], cigars=outs[-1][1])
            for wv, sub, nb, cb, outs, _ in jobs]
    for h in warm:
        h.wait()
    st0 = srv.stats()
    inflight = [[None] * args.serve_ring for _ in jobs]
    uses = [0] * len(jobs)
    last_set = [args.serve_ring - 1] * len(jobs)
    sub_t = np.zeros(k)
    handles = []
    throttled = 0
    t_start = time.perf_counter()
    for i in range(k):
        target = t_start + arrivals[i]
        while True:
            now = time.perf_counter()
            if now >= target:
                break
            time.sleep(min(target - now, 0.0005))
        j = int(order[i])
        r = uses[j] % args.serve_ring
        uses[j] += 1
        prev = inflight[j][r]
        if prev is not None and not prev.done():
            throttled += 1
            prev.wait()
        wv, sub, nb, cb, outs, _ = jobs[j]
        sub_t[i] = time.perf_counter() - t_start
        h = srv.submit(wv, sub, nb, cb, results=outs[r][0], cigars=outs[r][1])
        inflight[j][r] = h
        last_set[j] = r
        handles.append((h, i, nb))
    for h, _, _ in handles:
        h.wait()
    st = srv.stats()
    lat = np.array([(sub_t[i] - arrivals[i]) + h.latency() for h, i, _ in handles])
    finish = np.array([arrivals[i] + l for (h, i, _), l in zip(handles, lat)])
    pairs_done = int(sum(nb for _, _, nb in handles))
    span = float(finish.max() - arrivals[0])
    srv.close()
    # each job's last outputs, gathered into the pool layout for the checks
    res = np.zeros(total, dtype=ewfa.RESULT_DTYPE)
    cig = np.zeros(w.cigar_bytes, dtype=np.int8)
    for (wv, sub, nb, cb, outs, first), r in zip(jobs, last_set):
        res[first:first + nb] = outs[r][0]
        c0 = int(w.pairs[first]["cigar_off"])
        cig[c0:c0 + cb] = outs[r][1]
    d = lambda key: st[key] - st0[key]
    out = {"lambda_jobs_per_s": round(lam, 1), "offered_frac_of_batched": args.serve_load,
           "offered": round(lam * mean_pairs, 1), "jobs": k, "pairs": pairs_done,
           "mean_job_pairs": round(mean_pairs, 1),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "value": round(pairs_done / span, 1), "unit": "alignments/s",
           "frac_of_batched": round(pairs_done / span / batched_rate, 3),
           "batched_resident": round(batched_rate, 1), "batched_end_to_end": round(batched_e2e, 1),
           "frac_of_batched_end_to_end": round(pairs_done / span / batched_e2e, 3),
           "launches": int(d("batches")), "peak_batch_jobs": int(st["peak_batch_jobs"]),
           "slots": args.serve_slots, "max_batch_pairs": args.serve_batch, "serial": bool(args.serve_serial),
           "throttled_resubmits": throttled,
           "arena_peak_bytes": int(st["arena_peak"]), "failed": int(d("failed")),
           "pairs_not_ok": int((res["status"] != ewfa.PAIR_OK).sum()),
           "server_s": {"enqueue": round(d("upload_s"), 3), "device_wait": round(d("align_s"), 3),
                        "handout": round(d("handout_s"), 3)}}
    out["w"], out["res"], out["cig"] = w, res, cig
    log(f"served: {out['value'] / 1e6:.2f} M aln/s ({out['frac_of_batched']:.2f} of batched, "
        f"{out['frac_of_batched_end_to_end']:.2f} of batched from host memory), "
        f"p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, {out['launches']} launches, "
        f"{throttled} throttled")
    return out


def write_pair_file(w, path):
    """The workload as a '>'/'<' pair file (README.md:58-69)."""
    import numpy as np
    with open(path, "wb") as f:
        buf = []
        for i in range(w.n):
            buf.append(b">" + w.pattern(i) + b"\n<" + w.text(i) + b"\n")
            if len(buf) == 4096:
                f.write(b"".join(buf))
                buf = []
        f.write(b"".join(buf))
    return os.path.getsize(path)


def leg_cli(args, w, threads):
    """wfe.aligner on the 1,000,000-pair file: the reference's own figure
    ("Aligned executed. Wall time", src/aligner.cu:233) with and without -p,
    a sample of the printed lines against the oracle's text, and the pair-file
    ingestion (file mapped, lines found, sequences packed) timed apart."""
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp(prefix="ewfa_cli_")
    try:
        return cli_runs(args, w, threads, tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cli_runs(args, w, threads, tmp):
    import ctypes
    import re
    import subprocess
    import numpy as np
    import ewfa
    import oracle_py as orc
    path = os.path.join(tmp, "c2_pairs.txt")
    t0 = time.time()
    size = write_pair_file(w, path)
    log(f"cli: wrote {w.n} pairs ({size / 1e9:.2f} GB) in {time.time() - t0:.1f}s")
    out = {"pairs": w.n, "file_bytes": size, "batch": args.batch}
    f = ewfa.PairsFile()
    t0 = time.perf_counter()
    rc = ewfa.lib().ewfa_read_pairs(path.encode(), 1000, w.n, threads, ctypes.byref(f))
    out["ingest_s"] = round(time.perf_counter() - t0, 3)
    out["ingest_scan_s"] = round(f.scan_s, 3)
    out["ingest_pack_s"] = round(f.pack_s, 3)
    ewfa.lib().ewfa_free_pairs(ctypes.byref(f))
    assert rc == 0
    # the autotuned c2 policy, as the c2 leg uses; the built-in table (tuned
    # for the mixed-length workload) once more without -p for comparison
    cmd = [ewfa.ALIGNER_PATH, "-f", path, "-n", str(w.n), "-l", "1000", "-b", str(args.batch)]
    out["policy"] = os.path.relpath(POLICY_C2, ROOT)
    wall = lambda text: float(re.search(r"Aligned executed\. Wall time ([0-9.]+)s", text).group(1))
    tuned = ["--policy", POLICY_C2]
    for key, extra in (("no_print", tuned), ("print", tuned + ["-p"]), ("no_print_builtin_policy", [])):
        sink = os.path.join(tmp, f"out_{key}.txt")
        t0 = time.perf_counter()
        with open(sink, "wb") as so:
            r = subprocess.run(cmd + extra, stdout=so, stderr=subprocess.PIPE, timeout=300)
        proc = time.perf_counter() - t0
        assert r.returncode == 0, r.stderr.decode()[-500:]
        with open(sink, "rb") as so:
            so.seek(max(0, os.path.getsize(sink) - 200))
            tail = so.read().decode(errors="replace")
        s = wall(tail)
        out[key] = {"wall_s": s, "value": round(w.n / s, 1), "process_s": round(proc, 3)}
        log(f"cli {key}: wall {s:.3f}s -> {w.n / s / 1e6:.2f} M aln/s (process {proc:.2f}s)")
        if key == "print":
            # every 1000th line against the oracle's text
            want = {}
            idx = np.arange(0, w.n, 1000)
            ok, od, osl, ost = orc.batch_workload(w, idx=idx, threads=threads)
            for j, i in enumerate(idx):
                want[int(i)] = f"{int(i)}: {orc.text(ost[j])}"
            bad = 0
            with open(sink, "rb") as so:
                for line in so:
                    m = re.match(rb"^(\d+): ", line)
                    if m and int(m.group(1)) in want:
                        bad += line.decode().rstrip("\n") != want.pop(int(m.group(1)))
            out[key]["lines_checked"] = len(idx)
            out[key]["mismatches"] = bad + len(want)
        os.unlink(sink)
    return out


def oracle_check(w, res, cig, count, threads):
    """Distances and streams of `count` evenly spaced pairs against the oracle."""
    import numpy as np
    import ewfa
    import oracle_py as orc
    n = w.n
    idx = np.linspace(0, n - 1, num=min(count, n)).astype(np.int64)
    ok, od, osl, ost = orc.batch_workload(w, idx=idx, threads=threads)
    got = ewfa.streams_from(res[idx], w.pairs[idx], cig)
    bad = 0
    for j, i in enumerate(idx):
        if int(res[i]["distance"]) != int(od[j]) or not np.array_equal(got[j], ost[j]):
            bad += 1
    return len(idx), bad


def host_info():
    """The box's CPU as the OS reports it: model name, logical CPUs, and the
    CPUs this process may run on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return model, os.cpu_count() or 1, affinity


def cpu_baseline(w, args, threads):
    """The oracle on the host cores over a bounded prefix of the workload:
    align, CIGAR stream, the reference's validity check, with and without the
    CIGAR text step (the reference's -p path), on `threads` threads; and on
    one thread over a smaller prefix."""
    import numpy as np
    import oracle_py as orc
    n = min(args.cpu_sample, w.n)
    out = {}
    for text in (True, False):
        idx = np.arange(n)
        t0 = time.perf_counter()
        orc.batch_workload(w, idx=idx, threads=threads, text=text)
        out["text" if text else "no_text"] = (n, time.perf_counter() - t0)
    n1 = min(max(n // 20, 1), w.n)
    t0 = time.perf_counter()
    orc.batch_workload(w, idx=np.arange(n1), threads=1, text=True)
    out["one_thread"] = (n1, time.perf_counter() - t0)
    return out


def cpu_c1(threads):
    """c1 (10,000 x 150 bp, 2%) on the oracle at 1 thread and at `threads`."""
    import ewfa
    import oracle_py as orc
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c1"]), ewfa.CONFIGS["c1"]["n"])
    out = {}
    for t in (1, threads):
        t0 = time.perf_counter()
        orc.batch_workload(w, threads=t)
        out[f"threads_{t}"] = round(w.n / (time.perf_counter() - t0), 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=100_000)
    ap.add_argument("--legs", default="c2,cli,c3,c4,served", help="comma list of c2,cli,c3,c4,served")
    ap.add_argument("--c3-pairs", type=int, default=500_000)
    ap.add_argument("--c3-steps", type=int, default=3)
    ap.add_argument("--c4-pairs", type=int, default=20_000)
    ap.add_argument("--serve-load", type=float, default=0.85, help="offered load / batched c3 rate")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-pool", type=int, default=384)
    ap.add_argument("--serve-ring", type=int, default=6, help="output sets per pool job")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-batch", type=int, default=393216, help="max pairs per launch")
    ap.add_argument("--serve-serial", action="store_true", help="each batch's bucket passes on one stream")
    ap.add_argument("--serve-batch-bytes", type=int, default=2048, help="max CIGAR MiB per launch")
    ap.add_argument("--serve-arena", type=int, default=24, help="GiB")
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--c4-check", type=int, default=32, help="long-read pairs checked against the oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=100000, help="c2 pairs checked against the oracle "
                    "(c3 half as many, the served pool a fifth; a few seconds of host work)")
    ap.add_argument("--profile-run", choices=["c2", "c3", "c4", "served"],
                    help="the bench's timed loop over one workload only (for rocprofv3)")
    ap.add_argument("--eager", action="store_true",
                    help="with --profile-run: launch each kernel eagerly instead of replaying a graph")
    ap.add_argument("--profile-copy", action="store_true",
                    help="one 1 GiB device copy only (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(pairs are independent and would shard by index range across GPUs)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import ewfa

    assert ewfa.device_count() > 0, "bench needs a GPU"
    threads = min(16, os.cpu_count() or 8)

    if args.profile_copy:
        x = ewfa.DeviceBuffer.from_array(np.ones(1 << 27, dtype=np.int64))
        y = ewfa.DeviceBuffer(x.nbytes)
        ewfa.synchronize()
        ewfa.lib().ewfa_memcpy(y.ptr, x.ptr, x.nbytes, 2)
        ewfa.synchronize()
        log("copied 1 GiB")
        return

    if args.profile_run == "served":
        # the served leg alone (after the c3 leg it is offered against)
        c3 = leg_c3(args, threads)
        served = leg_served(args, threads, c3["value"])
        log(json.dumps({k: v for k, v in served.items() if k not in ("w", "res", "cig")}))
        return

    if args.profile_run:
        # the bench's own timed loop for one workload (graph replay with the
        # timing events, --warmup then --steps calls), so that a profiler's
        # per-kernel averages compare with the bench's kernels_ms
        name = args.profile_run
        n = {"c2": args.pairs, "c3": args.c3_pairs, "c4": args.c4_pairs}[name]
        cap = 1004 if name == "c2" else 0
        w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS[name]), n, cigar_cap=cap, threads=threads)
        r = Resident(w)
        pol = {"c2": POLICY_C2, "c3": POLICY_C3, "c4": None}[name]
        flags = ewfa.FLAG_TIMING | (ewfa.FLAG_NO_GRAPH if args.eager else 0)
        with ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=flags, policy_path=pol) as ctx:
            dt, a_ms, e_ms = timed(r.run, ctx, args.steps, args.warmup)
        log(f"profile run {name}: {n} pairs, {args.steps} steps in {dt:.3f}s -> {args.steps * n / dt / 1e6:.3f} M aln/s; "
            f"align {np.mean(a_ms):.3f} ms, emit {np.mean(e_ms):.3f} ms per step (events)")
        return

    legs = set(args.legs.split(","))
    c2 = leg_c2(args, threads)
    cli = None
    if "cli" in legs:
        # the pair file takes about 2 GB of temporary space; a box without
        # it (or any failure of this leg) is reported, not fatal
        import shutil
        import tempfile
        free = shutil.disk_usage(tempfile.gettempdir()).free
        if free < 6 << 30:
            cli = {"skipped": f"{free / 2**30:.1f} GiB free in {tempfile.gettempdir()}, 6 needed"}
        else:
            try:
                cli = leg_cli(args, c2["w"], threads)
            except Exception as e:  # reported in the JSON line
                cli = {"error": f"{type(e).__name__}: {e}"[:300]}
        log(f"cli: {cli if 'skipped' in cli or 'error' in cli else 'done'}")
    c3 = leg_c3(args, threads) if "c3" in legs else None
    c4 = leg_c4(args, threads) if "c4" in legs else None
    served = leg_served(args, threads, c3["value"]) if ("served" in legs and c3) else None

    # parity samples against the oracle (and validity on the long reads)
    checks = {}
    if args.check > 0:
        checks["c2"] = oracle_check(c2["w"], c2["res"], c2["cig"], args.check, threads)
        if c3:
            checks["c3"] = oracle_check(c3["w"], c3["res"], c3["cig"], max(args.check // 2, 1), threads)
        if served:
            checks["served"] = oracle_check(served["w"], served["res"], served["cig"], max(args.check // 5, 1),
                                            threads)
        if c4:
            # long reads: evenly spaced pairs, bit for bit (about 10-20 s of
            # oracle work on the host threads)
            checks["c4"] = oracle_check(c4["w"], c4["res"], c4["cig"], args.c4_check, threads)
        log("checks: " + ", ".join(f"{k} {v[0]} pairs {v[1]} bad" for k, v in checks.items()))

    cpu = None
    c1_cpu = None
    if not args.no_cpu:
        model, cores_box, affinity = host_info()
        cb = cpu_baseline(c2["w"], args, threads)
        rate = lambda k: round(cb[k][0] / cb[k][1], 1)
        cpu = {"value": rate("text"), "unit": "alignments/s", "cores": threads, "kind": "port",
               "sample": f"first {cb['text'][0]} pairs of the same 1000 bp / 5% workload: oracle align + CIGAR "
                         f"stream + text + the reference's validity check on {threads} host threads "
                         f"({cb['text'][1]:.1f}s)",
               "without_text": rate("no_text"), "one_thread": rate("one_thread"),
               "cores_box": cores_box, "cpus_allowed": affinity, "cpu_model": model}
        c1_cpu = cpu_c1(threads)
        log(f"cpu baseline: {cpu['value']:.0f} aln/s on {threads} threads ({cpu['without_text']:.0f} without "
            f"text, {cpu['one_thread']:.0f} on one); box: {model}, {cores_box} CPUs, {affinity} allowed; "
            f"c1 {c1_cpu}")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_align_c2.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    n = args.pairs
    achieved = c2["algo_bytes"] / (c2["align_ms"] / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(c2["value"], 1),
        "unit": "alignments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c2["dt"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "1,000,000 pairs x 1,000 bp, uniform ACGT, 5% edits (sub/ins/del equal), seed 2; "
                               "2-bit packed u64 input resident in HBM, CIGAR stride 1004 B; persistent kernel, "
                               "hipGraph replay, autotuned policy",
                   "pairs": n, "seq_len": 1000, "error": 0.05, "plain_batch": args.batch,
                   "parallelism": "single GPU"},
        "kernels_ms": {"align": round(c2["align_ms"], 4), "emit": round(c2["emit_ms"], 4)},
        "cells_per_s": round(c2["cells"] / (c2["align_ms"] / 1e3), 1),
        "plain": {"kernels": round(c2["plain_kernels"], 1), "end_to_end": round(c2["plain_e2e"], 1),
                  "unit": "alignments/s", "align_ms": round(c2["plain_align_ms"], 3),
                  "emit_ms": round(c2["plain_emit_ms"], 3), "identical_outputs": c2["plain_equal"]},
        "persistent_end_to_end": round(c2["persistent_e2e"], 1),
        # like for like: both paths with inputs resident in HBM, and both
        # with host buffers in and out (PCIe included)
        "ratio_vs_plain": {"resident": round(c2["value"] / c2["plain_kernels"], 3),
                           "end_to_end": round(c2["persistent_e2e"] / c2["plain_e2e"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if c3:
        out["c3"] = {"value": round(c3["value"], 1), "unit": "alignments/s", "pairs": c3["w"].n,
                     "ms_per_launch": round(c3["ms_per_launch"], 3), "align_ms": round(c3["align_ms"], 3),
                     "emit_ms": round(c3["emit_ms"], 3),
                     "cells_per_s": round(c3["cells"] / (c3["align_ms"] / 1e3), 1)}
    if c4:
        out["c4"] = {"value": round(c4["value"], 1), "unit": "alignments/s", "pairs": c4["w"].n,
                     "seconds": round(c4["seconds"], 4), "align_ms": round(c4["align_ms"], 3),
                     "emit_ms": round(c4["emit_ms"], 3), "device_bytes": int(c4["device_bytes"]),
                     "cells_per_s": round(c4["cells"] / c4["seconds"], 1)}
    if served:
        out["served"] = {k: v for k, v in served.items() if k not in ("w", "res", "cig")}
    if cli:
        out["cli"] = cli
    if c1_cpu:
        out["cpu_c1"] = dict(c1_cpu, unit="alignments/s", sample="c1: 10,000 x 150 bp, 2% edits, whole")
    out["oracle_checks"] = {k: {"pairs": v[0], "mismatches": v[1]} for k, v in checks.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
