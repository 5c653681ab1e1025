"""The allocation-free entry point (ewfa_align_device_noalloc) against
ewfa_align_device, bit for bit, on two calls in a row, the second with a
smaller input; the context grows no buffer; a batch whose scratch does not fit
the static array is refused before anything is launched."""
import sys

import numpy as np
import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class OnDevice:
    def __init__(self, w):
        self.w = w
        self.dw = ewfa.DeviceBuffer.from_array(w.words)
        self.dp = ewfa.DeviceBuffer.from_array(w.pairs[:w.n])
        self.dr = ewfa.DeviceBuffer(w.n * 16)
        self.dc = ewfa.DeviceBuffer(w.cigar_bytes)

    def run(self, fn):
        self.dr.fill_zero()
        self.dc.fill_zero()
        ewfa.synchronize()
        fn(self.dw.ptr, self.dp.ptr, self.w.n, self.w.max_len, self.dr.ptr, self.dc.ptr, None)
        ewfa.synchronize()
        return self.dr.download(ewfa.RESULT_DTYPE, self.w.n), self.dc.download(np.int8)


def same(w, a, b, label):
    ra, ca = a
    rb, cb = b
    assert np.array_equal(ra, rb), f"{label}: results differ"
    sa = ewfa.streams_from(ra, w.pairs, ca)
    sb = ewfa.streams_from(rb, w.pairs, cb)
    bad = [i for i in range(w.n) if not np.array_equal(sa[i], sb[i])]
    assert not bad, f"{label}: {len(bad)} streams differ (first {bad[0]})"


def test_noalloc_equals_align_device_on_two_calls():
    big = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c3"]), 3000, first=777)
    small = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c2"]), 500, first=4242)
    say(f"case noalloc: call 1 {big.n} mixed pairs (max_len {big.max_len}), call 2 {small.n} pairs of 1000 bp")
    with ewfa.Context() as ref, ewfa.Context() as ctx:
        held = ctx.memory_bytes()
        for k, w in enumerate((big, small)):
            d = OnDevice(w)
            want = d.run(ref.align_device)
            got = d.run(ctx.align_device_noalloc)
            same(w, want, got, f"call {k + 1}")
            assert int((got[0]["status"] == ewfa.PAIR_OK).sum()) == w.n
        assert ctx.memory_bytes() == held  # no context buffer was grown
    # and the second call's outputs are the oracle's
    ok, dist, slen, streams = orc.batch_workload(small)
    got = ewfa.streams_from(want[0], small.pairs, want[1])
    assert all(int(want[0][i]["distance"]) == int(dist[i]) and np.array_equal(got[i], streams[i])
               for i in range(small.n))


def test_noalloc_refuses_a_batch_beyond_its_scratch():
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c1"]), 3000)
    # op words for 100,001 levels a pair: 3,000 pairs need about 75 MB,
    # more than the static array holds; refused on the host
    say("case noalloc refusal: 3000 pairs declared up to 100,000 bases")
    d = OnDevice(w)
    with ewfa.Context() as ctx:
        with pytest.raises(ewfa.EwfaError, match="allocation-free"):
            ctx.align_device_noalloc(d.dw.ptr, d.dp.ptr, w.n, 100_000, d.dr.ptr, d.dc.ptr, None)
        # the context still works afterwards
        r, c = d.run(ctx.align_device_noalloc)
    assert int((r["status"] == ewfa.PAIR_OK).sum()) == w.n
