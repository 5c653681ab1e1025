"""Parity of the HIP path (through the C ABI) with the oracle, bit for bit.

Every case prints its name and sizes before touching the GPU.
"""
import sys

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_host(words, pairs, n, cigar_bytes, mode=ewfa.MODE_PERSISTENT, max_distance=0, flags=0, batch_size=0):
    with ewfa.Context(mode=mode, max_distance=max_distance, flags=flags, batch_size=batch_size) as ctx:
        res, cig = ctx.align_host(words, pairs, n, cigar_bytes)
    return res, cig


def assert_matches_oracle(res, pairs, cig, o_dist, o_streams, label, max_distance=0):
    streams = ewfa.streams_from(res, pairs, cig)
    bad = []
    for i in range(len(res)):
        st = int(res[i]["status"])
        want_status = (ewfa.PAIR_OK,) if max_distance == 0 else (ewfa.PAIR_OK, ewfa.PAIR_CAPPED,
                                                                 ewfa.PAIR_UNDEFINED)
        if st not in want_status or int(res[i]["distance"]) != int(o_dist[i]) or \
                not np.array_equal(streams[i], o_streams[i]):
            bad.append(i)
            if len(bad) > 5:
                break
    if bad:
        i = bad[0]
        raise AssertionError(
            f"{label}: {len(bad)}+ mismatching pairs; first {i}: status {int(res[i]['status'])} "
            f"dist {int(res[i]['distance'])} vs {int(o_dist[i])}; "
            f"gpu {streams[i][:40].tolist()} vs oracle {o_streams[i][:40].tolist()}")


def workload_case(name, n, first=0, **over):
    spec = dict(ewfa.CONFIGS[name])
    spec.update(over)
    return ewfa.Workload(ewfa.synth_spec(**spec), n, first=first)


@pytest.mark.parametrize("mode", [ewfa.MODE_PERSISTENT, ewfa.MODE_PLAIN])
def test_c1_all_pairs(mode):
    say(f"case c1 10000x150bp mode={mode}")
    w = workload_case("c1", 10_000)
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == w.n
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes, mode=mode, batch_size=3000)
    assert_matches_oracle(res, w.pairs, cig, d, st, "c1")


@pytest.mark.parametrize("mode", [ewfa.MODE_PERSISTENT, ewfa.MODE_PLAIN])
def test_c2_subset(mode):
    say(f"case c2 3000x1000bp (pairs 500000..) mode={mode}")
    w = workload_case("c2", 3000, first=500_000)
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == w.n
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes, mode=mode, batch_size=1000)
    assert_matches_oracle(res, w.pairs, cig, d, st, "c2")
    # distance is also the Levenshtein distance
    for i in range(0, w.n, 300):
        assert int(res[i]["distance"]) == orc.levenshtein(w.pattern(i), w.text(i))


def test_c2_reference_stride():
    """The reference's fixed per-pair CIGAR buffer (1004 bytes at -l 1000)."""
    say("case c2 reference stride 1004, 2000 pairs")
    spec = ewfa.synth_spec(**ewfa.CONFIGS["c2"])
    w = ewfa.Workload(spec, 2000, first=7, cigar_cap=1004)
    ok, d, sl, st = orc.batch_workload(w)
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes)
    assert_matches_oracle(res, w.pairs, cig, d, st, "c2-stride")


def test_c3_subset_mixed_lengths():
    say("case c3 1500 pairs log-uniform 100..10000 bp")
    w = workload_case("c3", 1500, first=12345)
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == w.n
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes)
    assert_matches_oracle(res, w.pairs, cig, d, st, "c3")
    assert int(d.max()) > 64  # the deferral passes were exercised


def test_c4_kind_short():
    say("case c4-kind 6 pairs 6000..8000 bp 10-15% with indel runs")
    w = ewfa.Workload(ewfa.synth_spec(kind=2, seed=4, len_lo=6000, len_hi=8000, err_lo=0.10, err_hi=0.15), 6)
    ok, d, sl, st = orc.batch_workload(w)
    assert ok == w.n
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes)
    assert_matches_oracle(res, w.pairs, cig, d, st, "c4-kind")


EDGE = [
    (b"", b""), (b"A", b""), (b"", b"A"), (b"ACGT", b"ACGT"), (b"AA", b"A"), (b"ACGT", b"AGGT"),
    (b"AC", b"CA"), (b"A" * 300, b"A" * 300), (b"A" * 300, b"C" * 300), (b"A" * 200, b"A" * 450),
    (b"N" * 40, b"G" * 40), (b"acgtu" * 30, b"ACGTT" * 30), (b"ACGT" * 16, b"ACGT" * 8),
    (b"A" * 32, b"A" * 33), (b"A" * 64, b"T" + b"A" * 64), (b"G" * 1, b"C" * 1),
    (b"ACGTACGTAC" * 13, b"ACGTACGTAC" * 13 + b"G"), (b"T" * 127, b"T" * 127), (b"T" * 254, b"T" * 254),
    (b"CA" * 70, b"AC" * 70), (b"A" * 10 + b"C" * 10, b"C" * 10 + b"A" * 10),
]


def test_edge_cases():
    say(f"case edge {len(EDGE)} hand pairs")
    P = [p for p, _ in EDGE]
    T = [t for _, t in EDGE]
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    res, cig = run_host(words, pairs, len(P), cb)
    streams = ewfa.streams_from(res, pairs, cig)
    for i, (p, t) in enumerate(EDGE):
        d, s = orc.stream(p, t)
        assert int(res[i]["status"]) == ewfa.PAIR_OK, (p, t)
        assert int(res[i]["distance"]) == d, (p, t)
        assert streams[i].tolist() == s, (p, t, streams[i].tolist(), s)
        assert ewfa.cigar_text(streams[i]) == orc.text(s)


def test_reference_cap_64():
    """max_distance=64 reproduces the reference's capped output."""
    say("case cap64 c3-kind high error 400 pairs")
    w = ewfa.Workload(ewfa.synth_spec(kind=1, seed=9, len_lo=300, len_hi=3000, err_lo=0.02, err_hi=0.10), 400)
    ok, d, sl, st = orc.batch_workload(w, max_distance=64)
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes, max_distance=64)
    assert_matches_oracle(res, w.pairs, cig, d, st, "cap64", max_distance=64)
    capped = res["status"] == ewfa.PAIR_CAPPED
    assert capped.sum() > 10
    assert (res["distance"][capped] == 64).all()


def test_plain_equals_persistent_and_graph_replays():
    say("case replay c2 2000 pairs: plain, eager persistent, graph x3")
    w = workload_case("c2", 2000, first=99)
    ref, refc = run_host(w.words, w.pairs, w.n, w.cigar_bytes, mode=ewfa.MODE_PLAIN, batch_size=700)
    with ewfa.Context(mode=ewfa.MODE_PERSISTENT, flags=ewfa.FLAG_NO_GRAPH) as ctx:
        r1, c1 = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    assert np.array_equal(ref, r1) and np.array_equal(refc, c1)
    with ewfa.Context(mode=ewfa.MODE_PERSISTENT) as ctx:
        for rep in range(3):
            r2, c2 = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
            assert np.array_equal(ref, r2) and np.array_equal(refc, c2), rep


def test_device_api_with_library_buffers():
    say("case device api c2 1000 pairs, device buffers, timing flag")
    w = workload_case("c2", 1000, first=3)
    ok, d, sl, st = orc.batch_workload(w)
    dw = ewfa.DeviceBuffer.from_array(w.words)
    dp = ewfa.DeviceBuffer.from_array(w.pairs[:w.n])
    dr = ewfa.DeviceBuffer(w.n * 16)
    dc = ewfa.DeviceBuffer(w.cigar_bytes)
    # timing events inside the captured graph, then on eager launches
    for flags in (ewfa.FLAG_TIMING, ewfa.FLAG_TIMING | ewfa.FLAG_NO_GRAPH):
        dr.fill_zero()
        with ewfa.Context(flags=flags) as ctx:
            for _ in range(2):
                ctx.align_device(dw.ptr, dp.ptr, w.n, w.max_len, dr.ptr, dc.ptr, None)
                a_ms, e_ms = ctx.last_timing()
                assert a_ms > 0 and e_ms > 0, flags
            assert ctx.memory_bytes() > 0
        ewfa.synchronize()
        res = dr.download(ewfa.RESULT_DTYPE, w.n)
        cig = dc.download(np.int8)
        assert_matches_oracle(res, w.pairs, cig, d, st, f"device-api flags {flags}")


def test_graph_replay_after_scratch_grows():
    """A replayed plan graph must not use scratch buffers that a call in between
    replaced: a small device call (graph captured), a host call on a larger
    c3-like input (the op, list, bucket and backtrace scratch grow), then the
    small call again with the same buffers, sizes and policy."""
    say("case graph replay after scratch growth: c2 600 pairs, then c3-like 6000 pairs")
    w = workload_case("c2", 600, first=11)
    ok, d, sl, st = orc.batch_workload(w)
    big = ewfa.Workload(ewfa.synth_spec(kind=1, seed=12, len_lo=200, len_hi=6000, err_lo=0.02,
                                        err_hi=0.10), 6000)
    dw = ewfa.DeviceBuffer.from_array(w.words)
    dp = ewfa.DeviceBuffer.from_array(w.pairs[:w.n])
    dr = ewfa.DeviceBuffer(w.n * 16)
    dc = ewfa.DeviceBuffer(w.cigar_bytes)
    with ewfa.Context() as ctx:
        ctx.align_device(dw.ptr, dp.ptr, w.n, w.max_len, dr.ptr, dc.ptr, None)
        before = ctx.memory_bytes()
        ctx.align_host(big.words, big.pairs, big.n, big.cigar_bytes)
        assert ctx.memory_bytes() > before  # the scratch was replaced by larger buffers
        dr.fill_zero()
        dc.fill_zero()
        ctx.align_device(dw.ptr, dp.ptr, w.n, w.max_len, dr.ptr, dc.ptr, None)
        ewfa.synchronize()
        res = dr.download(ewfa.RESULT_DTYPE, w.n)
        cig = dc.download(np.int8)
    assert_matches_oracle(res, w.pairs, cig, d, st, "replay after growth")


def test_pack_kernel_matches_host():
    say("case pack kernel 500 sequences, stride 304")
    rng = np.random.default_rng(5)
    stride = 304
    n = 500
    lens = rng.integers(0, stride + 1, n).astype(np.uint32)
    ascii = rng.choice(np.frombuffer(b"ACGTNacgtU", dtype=np.uint8), size=n * stride).astype(np.uint8)
    wstride = int(ewfa.lib().ewfa_packed_words(stride))
    expect = np.zeros(n * wstride, dtype=np.uint64)
    for j in range(n):
        seq = ascii[j * stride: j * stride + int(lens[j])].tobytes()
        ew = orc.pack_u64(seq)
        expect[j * wstride: j * wstride + len(ew)] = ew
    da = ewfa.DeviceBuffer.from_array(ascii)
    dl = ewfa.DeviceBuffer.from_array(lens)
    dw = ewfa.DeviceBuffer.from_array(np.full(n * wstride, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64))
    rc = ewfa.lib().ewfa_pack_ascii(da.ptr, stride, dl.ptr, n, dw.ptr, wstride, None)
    assert rc == 0
    ewfa.synchronize()
    got = dw.download(np.uint64)
    assert np.array_equal(got, expect)


def test_cigar_overflow_status():
    say("case overflow cap 8 bytes, 50 c1 pairs")
    w = workload_case("c1", 50)
    w.layout(cigar_cap=8)
    ok, d, sl, st = orc.batch_workload(w)
    res, cig = run_host(w.words, w.pairs, w.n, w.cigar_bytes)
    for i in range(w.n):
        full = len(st[i])
        if full > 8:
            assert int(res[i]["status"]) == ewfa.PAIR_OVERFLOW
            assert int(res[i]["cigar_len"]) == full
            assert cig[8 * i: 8 * i + 8].tolist() == st[i][:8].tolist()
        else:
            assert int(res[i]["status"]) == ewfa.PAIR_OK
            assert cig[8 * i: 8 * i + full].tolist() == st[i].tolist()


def test_distances_beyond_what_lds_holds():
    """Pairs whose wavefront no longer fits LDS go to deferral passes whose
    slots (sequence copies and offsets) live in HBM: a 100 kbp pair at 10%
    edits (distance about 9,600, above the about 7,100 levels LDS holds for
    that length) equals the oracle bit for bit, and a random 40 kbp pair
    (distance about 21,000) has the Levenshtein distance and a valid stream
    with as many edits; a close pair in the same call is aligned as usual."""
    rng = np.random.default_rng(21)
    L = 40000
    a = bytes(rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), L))
    b = bytes(rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), L))
    c = a[:2000]
    d = c[:1000] + b"A" + c[1000:]
    w = ewfa.Workload(ewfa.synth_spec(kind=0, seed=77, len=100_000, err=0.10), 1)
    p100, t100 = w.pattern(0), w.text(0)
    say("case beyond LDS: 100 kbp pair at 10%, random 40 kbp pair, a one-insertion pair")
    words, pairs, cb, ml = ewfa.from_sequences([p100, a, c], [t100, b, d])
    with ewfa.Context() as ctx:
        assert ctx.max_distance(ml) == ml + 1
        res, cig = ctx.align_host(words, pairs, 3, cb)
    streams = ewfa.streams_from(res, pairs, cig)
    assert all(int(r["status"]) == ewfa.PAIR_OK for r in res)
    say(f"  distances {res['distance'].tolist()}; oracle and DP on the host")
    # the 100 kbp pair: the oracle's distance and stream
    ok, od, osl, ost = orc.batch_streams([p100], [t100], threads=1)
    assert int(res[0]["distance"]) == int(od[0]) > 8000
    assert np.array_equal(streams[0], ost[0])
    # the random pair: exact distance, a valid alignment with that many edits
    assert int(res[1]["distance"]) == orc.levenshtein(a, b)
    assert ewfa.check_cigar(a, b, streams[1]) == 1
    assert int(np.isin(streams[1], [ord("X"), ord("I"), ord("D")]).sum()) == int(res[1]["distance"])
    assert int(res[2]["distance"]) == 1
    assert streams[2].tolist() == orc.stream(c, d)[1]


def test_host_pipeline_several_chunks():
    """ewfa_align_host splits large calls into chunks whose uploads overlap
    the compute of the chunk before; the results equal the device-resident
    call on every pair, and a sample equals the oracle."""
    w = workload_case("c1", 300_000, first=20_000)
    say("case host pipeline: 300,000 pairs of 150 bp (3 chunks)")
    with ewfa.Context() as ctx:
        res, cig = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    dw = ewfa.DeviceBuffer.from_array(w.words)
    dp = ewfa.DeviceBuffer.from_array(w.pairs[:w.n])
    dr = ewfa.DeviceBuffer(w.n * 16)
    dc = ewfa.DeviceBuffer(w.cigar_bytes)
    with ewfa.Context() as ctx:
        ctx.align_device(dw.ptr, dp.ptr, w.n, w.max_len, dr.ptr, dc.ptr, None)
    ewfa.synchronize()
    res_d = dr.download(ewfa.RESULT_DTYPE, w.n)
    cig_d = dc.download(np.int8)
    assert np.array_equal(res, res_d)
    idx = np.arange(0, w.n, 97)
    sa = ewfa.streams_from(res[idx], w.pairs[idx], cig)
    sb = ewfa.streams_from(res_d[idx], w.pairs[idx], cig_d)
    assert all(np.array_equal(a, b) for a, b in zip(sa, sb))
    ok, d, sl, st = orc.batch_workload(w, idx=idx[:3000])
    assert np.array_equal(res["distance"][idx[:3000]], d)
    assert all(np.array_equal(a, b) for a, b in zip(sa[:3000], st))


def test_resident_words_uploaded_in_pieces():
    """ewfa_upload_words_async sends the words in 16 MiB pieces and each
    ewfa_align_resident call waits only for the pieces its pairs read: the
    batches, run while later pieces are in flight, equal the host pipeline on
    every pair; a synchronous upload and a host call afterwards still see
    their own words."""
    import ctypes
    w = workload_case("c2", 40_000, first=3_000)  # about 2.6 M words: two pieces
    say(f"case resident pieces: {w.n} pairs, {w.n_words} words")
    with ewfa.Context() as ctx:
        ref, ref_c = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
    pin = ewfa.PinnedArray(w.n_words, np.uint64)
    pin.array[:] = w.words[:w.n_words]
    L = ewfa.lib()
    res = np.zeros(w.n, dtype=ewfa.RESULT_DTYPE)
    cig = np.zeros(w.cigar_bytes, dtype=np.int8)
    with ewfa.Context() as ctx:
        assert L.ewfa_upload_words_async(ctx.handle, pin.ptr, w.n_words) == ewfa.OK
        for lo, hi in ((0, 7_000), (7_000, 21_000), (21_000, w.n)):
            sub = np.array(w.pairs[lo:hi])
            rb = np.zeros(hi - lo, dtype=ewfa.RESULT_DTYPE)
            assert L.ewfa_align_resident(ctx.handle, sub.ctypes.data, hi - lo, rb.ctypes.data,
                                         cig.ctypes.data, w.cigar_bytes) == ewfa.OK
            res[lo:hi] = rb
        assert np.array_equal(res, ref)
        sa = ewfa.streams_from(res, w.pairs, cig)
        sb = ewfa.streams_from(ref, w.pairs, ref_c)
        assert all(np.array_equal(a, b) for a, b in zip(sa, sb))
        # a second asynchronous upload, then a host call over other words
        assert L.ewfa_upload_words_async(ctx.handle, pin.ptr, w.n_words) == ewfa.OK
        r2, c2 = ctx.align_host(w.words, w.pairs, w.n, w.cigar_bytes)
        assert np.array_equal(r2, ref)
    ok, d, sl, st = orc.batch_workload(w, idx=np.arange(0, w.n, 40))
    assert np.array_equal(res["distance"][::40], d)


def _subst(seq, positions):
    out = bytearray(seq)
    for i in positions:
        out[i] = b"C"[0] if out[i] != b"C"[0] else b"G"[0]
    return bytes(out)


def test_match_runs_across_byte_and_chunk_limits():
    """Match runs of 126, 127, 128, 253, 254, 255 and more bases: the
    register kernel writes streams itself for runs up to 254 and leaves
    longer ones to the emit kernel; both agree with the oracle, in one call."""
    rng = np.random.default_rng(33)
    base = lambda n: bytes(rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), n))
    cases = []
    for runs in ([126, 127, 128], [253, 254, 40], [255, 20], [254, 254, 254], [1000], [127, 127, 127, 127],
                 [300, 299], [10, 500, 10]):
        n = sum(runs) + len(runs) - 1
        p = base(n)
        cuts = list(np.cumsum(runs[:-1]) + np.arange(len(runs) - 1))
        cases.append((p, _subst(p, cuts)))
    cases.append((base(700), base(700)))  # unrelated: short runs, large distance
    say(f"case match runs: {len(cases)} pairs")
    P = [p for p, _ in cases]
    T = [t for _, t in cases]
    words, pairs, cb, ml = ewfa.from_sequences(P, T)
    with ewfa.Context() as ctx:
        res, cig = ctx.align_host(words, pairs, len(P), cb)
    streams = ewfa.streams_from(res, pairs, cig)
    for i, (p, t) in enumerate(cases):
        d, s = orc.stream(p, t)
        assert int(res[i]["status"]) in (ewfa.PAIR_OK,), i
        assert int(res[i]["distance"]) == d, i
        assert streams[i].tolist() == s, (i, streams[i].tolist()[:20], s[:20])
