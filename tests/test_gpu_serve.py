"""The job server: every job's results are identical alone and in the mix
(tolerance 0: the work is integer), queued jobs are merged into shared
launches and overlap on the device, and an arena smaller than the offered load
still admits and completes every job."""
import os
import time

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_jobs(count, seed=5, max_pairs=300):
    rng = np.random.default_rng(seed)
    spec = ewfa.synth_spec(kind=1, seed=seed, len_lo=100, len_hi=3000, err_lo=0.01, err_hi=0.10)
    jobs = []
    first = 0
    for _ in range(count):
        n = int(rng.integers(1, max_pairs + 1))
        w = ewfa.Workload(spec, n, first=first, threads=4)
        first += n
        jobs.append(w)
    return jobs


def alone(w):
    with ewfa.Context() as ctx:
        return ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)


def test_each_job_same_alone_and_in_the_mix():
    jobs = make_jobs(24)
    say(f"case serve mix: {len(jobs)} jobs, {sum(w.n for w in jobs)} pairs, 4 slots")
    ref = [alone(w) for w in jobs]
    with ewfa.Server(slots=4) as srv:
        handles = [srv.submit(w.words, w.pairs, w.n, w.cigar_bytes) for w in jobs]
        outs = [h.wait() for h in handles]
        lat = [h.latency() for h in handles]
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["peak_running"] > 1  # admitted jobs ran side by side
    assert st["batches"] < len(jobs) and st["peak_batch_jobs"] > 1  # merged launches
    for w, (r0, c0), (r1, c1) in zip(jobs, ref, outs):
        assert np.array_equal(r0, r1)
        s0 = ewfa.streams_from(r0, w.pairs, c0)
        s1 = ewfa.streams_from(r1, w.pairs, c1)
        assert all(np.array_equal(a, b) for a, b in zip(s0, s1))
    assert all(x > 0 for x in lat)
    # and the mix is right, not just consistent: spot-check against the oracle
    w = jobs[3]
    ok, d, sl, streams = orc.batch_workload(w)
    got = ewfa.streams_from(outs[3][0], w.pairs, outs[3][1])
    assert all(np.array_equal(a, b) for a, b in zip(got, streams))


def test_served_mix_at_full_ranges_every_job_against_the_oracle():
    """The served workload's ranges: jobs of 1..4,096 pairs, lengths
    log-uniform 100..10,000 bp, 1..10% edits, arriving as a Poisson stream
    while earlier jobs run.  Every job equals the oracle bit for bit and its
    own run alone."""
    rng = np.random.default_rng(55)
    spec = ewfa.synth_spec(kind=1, seed=55, len_lo=100, len_hi=10_000, err_lo=0.01, err_hi=0.10)
    sizes = [int(x) for x in rng.integers(1, 4097, 10)] + [4096, 1]
    jobs, first = [], 0
    for n in sizes:
        jobs.append(ewfa.Workload(spec, n, first=first, threads=8))
        first += n
    gaps = rng.exponential(0.004, len(jobs))  # about 250 jobs/s
    say(f"case serve c5 ranges: {len(jobs)} jobs, {first} pairs, Poisson arrivals, 2 slots")
    with ewfa.Server(slots=2, max_batch_pairs=8192) as srv:
        handles = []
        for w, g in zip(jobs, gaps):
            time.sleep(float(g))
            handles.append(srv.submit(w.words, w.pairs, w.n, w.cigar_bytes))
        outs = [h.wait() for h in handles]
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    say("  served; checking every job against the oracle and alone")
    threads = max(1, min(16, os.cpu_count() or 8))
    for k, (w, (r, c)) in enumerate(zip(jobs, outs)):
        ok, d, sl, streams = orc.batch_workload(w, threads=threads)
        assert ok == w.n
        got = ewfa.streams_from(r, w.pairs, c)
        bad = [i for i in range(w.n) if int(r[i]["status"]) != ewfa.PAIR_OK or int(r[i]["distance"]) != int(d[i])
               or not np.array_equal(got[i], streams[i])]
        assert not bad, f"job {k}: {len(bad)} of {w.n} pairs differ from the oracle (first {bad[0]})"
        r0, c0 = alone(w)
        assert np.array_equal(r0, r), f"job {k}: results differ from the job run alone"


def footprint(w):
    """Arena bytes of a job alone: resident words and CIGAR regions, plus
    descriptors, results, op scratch and compaction block sums (256-byte
    blocks; words, descriptors and results of at least 64 KiB, the size the
    server pads their copies to)."""
    r = lambda x: (x + 255) // 256 * 256
    dma = 64 << 10
    ops = w.n * ((w.max_len + 1) // 32 + 1) * 8
    scan = 16 * ((w.n + 1023) // 1024)
    return (r(max(w.n_words * 8, dma)) + r(w.cigar_bytes) + r(max(w.n * 40, dma)) + r(max(w.n * 16, dma)) + r(ops)
            + r(scan))


def test_small_arena_admits_as_memory_frees():
    jobs = make_jobs(12, seed=6, max_pairs=200)
    need = max(footprint(w) for w in jobs)
    arena = int(need * 2.5)  # room for about two jobs at a time
    say(f"case serve arena {arena} bytes, {len(jobs)} jobs")
    with ewfa.Server(slots=4, arena_bytes=arena) as srv:
        handles = [srv.submit(w.words, w.pairs, w.n, w.cigar_bytes) for w in jobs]
        outs = [h.wait() for h in handles]
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["arena_peak"] <= arena
    for w, (r, c) in zip(jobs, outs):
        r0, c0 = alone(w)
        assert np.array_equal(r0, r)


def test_jobs_arriving_during_a_launch_share_the_next():
    big = ewfa.Workload(ewfa.synth_spec(kind=1, seed=8, len_lo=2000, len_hi=4000, err_lo=0.05, err_hi=0.10),
                        3000, threads=4)
    small = make_jobs(10, seed=9, max_pairs=100)
    say("case serve one slot: a long job, then 10 small ones queued behind it")
    with ewfa.Server(slots=1) as srv:
        handles = [srv.submit(big.words, big.pairs, big.n, big.cigar_bytes)]
        handles += [srv.submit(w.words, w.pairs, w.n, w.cigar_bytes) for w in small]
        outs = [h.wait() for h in handles]
        st = srv.stats()
    assert st["completed"] == 11 and st["failed"] == 0
    assert st["batches"] <= 3 and st["peak_batch_jobs"] >= 5
    for w, (r, c) in zip([big] + small, outs):
        r0, c0 = alone(w)
        assert np.array_equal(r0, r)
        s0 = ewfa.streams_from(r0, w.pairs, c0)
        s1 = ewfa.streams_from(r, w.pairs, c)
        assert all(np.array_equal(a, b) for a, b in zip(s0, s1))


def test_job_larger_than_arena_is_refused():
    w = make_jobs(1, seed=7, max_pairs=50)[0]
    with ewfa.Server(slots=1, arena_bytes=1 << 12) as srv:
        with pytest.raises(ewfa.EwfaError):
            srv.submit(w.words, w.pairs, w.n, w.cigar_bytes)
