"""CPU tests of libewfa's host side: C-ABI exports, text, validity check,
packing, pair-file reader, generator, and the wfe.aligner argument contract
(checked against the reference's own parser, compiled from its source into
oracle/_ref when the reference tree is present)."""
import contextlib
import ctypes
import os
import random
import re
import subprocess
import tempfile

import numpy as np
import pytest

import ewfa
import oracle_py as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ewfa.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ewfa_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = header_functions()
    assert len(names) >= 25
    L = ewfa.lib()
    for n in names:
        assert hasattr(L, n), n
    assert sorted(ewfa.EXPORTS) == names


def test_aligner_binary_built():
    assert os.access(ewfa.ALIGNER_PATH, os.X_OK)


# ---------------------------------------------------------------------------
# text and check against the oracle

def rand_stream(rng):
    out = []
    for _ in range(rng.randrange(0, 40)):
        if rng.random() < 0.5:
            out.append(-rng.randrange(1, 128))
        else:
            out.append(ord(rng.choice("XID")))
    return out


def test_text_matches_oracle():
    rng = random.Random(1)
    for _ in range(500):
        s = rand_stream(rng)
        assert ewfa.cigar_text(s) == orc.text(s)


def test_standard_text():
    assert ewfa.cigar_text_standard([-5, ord("X"), -127, -3, ord("I"), ord("I")]) == "5=1X130=2I"
    assert ewfa.cigar_text_standard([]) == ""


def test_check_matches_oracle():
    rng = random.Random(2)
    for _ in range(300):
        p = "".join(rng.choice("ACGT") for _ in range(rng.randrange(0, 60)))
        t = "".join(rng.choice("ACGT") for _ in range(rng.randrange(0, 60)))
        d, s = orc.stream(p, t)
        assert ewfa.check_cigar(p.encode(), t.encode(), s) == orc.check(p, t, s) == 1
        bad = rand_stream(rng)
        assert ewfa.check_cigar(p.encode(), t.encode(), bad) == orc.check(p, t, bad)


def test_pack_matches_oracle():
    rng = random.Random(3)
    L = ewfa.lib()
    for n in [0, 1, 31, 32, 33, 63, 64, 65, 1000]:
        s = bytes(rng.choice(b"ACGTNacgtU\r") for _ in range(n))
        w = np.full(int(L.ewfa_packed_words(n)), 7, dtype=np.uint64)
        L.ewfa_pack_host(s, n, w.ctypes.data)
        assert np.array_equal(w, orc.pack_u64(s))


# ---------------------------------------------------------------------------
# reader

def read_product(path, seq_len, n):
    f = ewfa.PairsFile()
    rc = ewfa.lib().ewfa_read_pairs(path.encode(), seq_len, n, 2, ctypes.byref(f))
    out = []
    if f.pairs:
        pairs = np.ctypeslib.as_array((ctypes.c_uint8 * (40 * max(n, 1))).from_address(f.pairs))
        pairs = pairs.view(ewfa.PAIR_DTYPE)
        for i in range(n):
            po, to = f.ascii_off[2 * i], f.ascii_off[2 * i + 1]
            pl, tl = int(pairs[i]["plen"]), int(pairs[i]["tlen"])
            out.append((ctypes.string_at(ctypes.addressof(f.ascii.contents) + po, pl),
                        ctypes.string_at(ctypes.addressof(f.ascii.contents) + to, tl)))
    ewfa.lib().ewfa_free_pairs(ctypes.byref(f))
    return rc, out


@pytest.mark.parametrize("data,n,seq_len", [
    (b">ACGT\n<AGGT\n>AA\n<A\n", 2, 10),
    (b">ACGT\n<AGGT", 1, 10),
    (b">ACGT\r\n<ACGT\r\n", 1, 10),
    (b">ACGT\n<ACGT\n", 3, 10),
    (b">ACGT\n<ACGT\n", 1, 0),
    (b">\n<\n>A\n<\n", 2, 5),
    (b">" + b"ACGT" * 300 + b"\n<" + b"ACGA" * 300 + b"\n", 1, 1200),
    (b">ACGT\n<ACGT\n>TT\n<GG\n>CC\n<AA\n", 2, 10),
])
def test_reader_matches_oracle(tmp_path, data, n, seq_len):
    path = os.path.join(tmp_path, "pairs.txt")
    with open(path, "wb") as f:
        f.write(data)
    rc, pairs = read_product(path, seq_len, n)
    orc_rc, opairs, _, _ = orc.read_pairs(path, seq_len, n)
    assert rc == ewfa.OK and orc_rc == 1
    assert pairs == opairs


def test_reader_errors(tmp_path):
    path = os.path.join(tmp_path, "bad.txt")
    with open(path, "wb") as f:
        f.write(b"<ACGT\n>ACGT\n")
    rc, _ = read_product(path, 10, 1)
    assert rc == ewfa.ERR_FORMAT
    rc, _ = read_product(os.path.join(tmp_path, "missing.txt"), 10, 1)
    assert rc == ewfa.ERR_IO


def test_reader_error_message_is_the_references(tmp_path):
    # utils/sequence_reader.cu:157-158: "Invalid sequence format on line %d.\n    %s\n"
    path = os.path.join(tmp_path, "bad.txt")
    with open(path, "wb") as f:
        f.write(b">ACGT\n<ACGT\n>AC\nTT\n")
    rc, _ = read_product(path, 10, 2)
    assert rc == ewfa.ERR_FORMAT
    assert ewfa.lib().ewfa_last_error() == b"Invalid sequence format on line 3.\n    TT\n\n"


@pytest.mark.parametrize("threads,n", [(1, 9000), (3, 9000), (8, 9000), (8, 4000), (5, 20000)])
def test_reader_across_scan_blocks(tmp_path, threads, n):
    # about 20 MB: line starts are found in 8 MiB blocks, one thread a block;
    # lines straddle block ends, the file ends without a newline, and n may
    # stop before the file does or ask for more pairs than it holds
    rng = np.random.default_rng(7)
    lines = []
    for i in range(2 * 9000):
        L = int(rng.integers(0, 2400))
        seq = bytes(rng.choice(list(b"ACGTN"), L).astype(np.uint8))
        lines.append((b">" if i % 2 == 0 else b"<") + seq)
    data = b"\n".join(lines)  # no final newline
    path = os.path.join(tmp_path, "big.txt")
    with open(path, "wb") as f:
        f.write(data)
    f_ = ewfa.PairsFile()
    rc = ewfa.lib().ewfa_read_pairs(path.encode(), 2500, n, threads, ctypes.byref(f_))
    assert rc == ewfa.OK
    got_words = np.ctypeslib.as_array(f_.words, shape=(f_.n_words,)).copy()
    pairs = np.ctypeslib.as_array((ctypes.c_uint8 * (40 * n)).from_address(f_.pairs)).view(ewfa.PAIR_DTYPE).copy()
    ewfa.lib().ewfa_free_pairs(ctypes.byref(f_))
    orc_rc, opairs, _, _ = orc.read_pairs(path, 2500, n)
    assert orc_rc == 1
    for i in range(0, n, 97):
        p, t = opairs[i]
        assert int(pairs[i]["plen"]) == len(p) and int(pairs[i]["tlen"]) == len(t), i
        for seq, off in ((p, int(pairs[i]["pattern_off"])), (t, int(pairs[i]["text_off"]))):
            want = orc.pack_u64(seq)
            assert np.array_equal(got_words[off:off + len(want)], want), i


# ---------------------------------------------------------------------------
# launch policy files

def test_policy_builtin_and_files(tmp_path):
    builtin = ewfa.read_policy(None)
    assert 1 <= len(builtin) <= 8 and builtin[-1][0] == 0xFFFFFFFF
    good = os.path.join(tmp_path, "good.policy")
    with open(good, "w") as f:
        f.write("# len_hi waves dcap bt_lds grab\n300 1 63 1 16\n5000 2 512 0 1\n9000 4 700 0 1\n")
    assert ewfa.read_policy(good) == [(300, 1, 63, 1, 16), (5000, 2, 512, 0, 1), (0xFFFFFFFF, 4, 700, 0, 1)]
    # a plan holds 8 passes: a ninth bucket would never be planned, so the
    # file is refused rather than its longest pairs failing
    nine = os.path.join(tmp_path, "nine.policy")
    with open(nine, "w") as f:
        for i in range(9):
            f.write(f"{(i + 1) * 100} 1 63 1 4\n")
    with pytest.raises(ewfa.EwfaError):
        ewfa.read_policy(nine)
    eight = os.path.join(tmp_path, "eight.policy")
    with open(eight, "w") as f:
        for i in range(8):
            f.write(f"{(i + 1) * 100} 1 63 1 4\n")
    assert len(ewfa.read_policy(eight)) == 8
    with pytest.raises(ewfa.EwfaError):
        ewfa.read_policy(os.path.join(tmp_path, "missing.policy"))


# ---------------------------------------------------------------------------
# generator

def test_synth_deterministic_and_within_budget():
    spec = ewfa.synth_spec(**ewfa.CONFIGS["c2"])
    a = ewfa.Workload(spec, 50, first=100, threads=1)
    b = ewfa.Workload(spec, 50, first=100, threads=4)
    assert np.array_equal(a.words, b.words)
    c = ewfa.Workload(spec, 60, first=90, threads=3)  # same pairs at another offset
    assert c.pattern(10) == a.pattern(0) and c.text(10) == a.text(0)
    for i in range(a.n):
        assert len(a.pattern(i)) == 1000
        assert orc.levenshtein(a.pattern(i), a.text(i)) <= 50
    spec3 = ewfa.synth_spec(**ewfa.CONFIGS["c3"])
    w = ewfa.Workload(spec3, 300)
    assert 100 <= w.plen.min() and w.plen.max() <= 10000
    spec4 = ewfa.synth_spec(**ewfa.CONFIGS["c4"])
    w4 = ewfa.Workload(spec4, 3)
    assert 10000 <= w4.plen.min() and w4.plen.max() <= 50000


def test_layout_regions():
    spec = ewfa.synth_spec(**ewfa.CONFIGS["c1"])
    w = ewfa.Workload(spec, 20)
    p = w.pairs[: w.n]
    assert (p["cigar_off"] % 4 == 0).all() and (p["cigar_cap"] % 4 == 0).all()
    assert (p["cigar_cap"] >= p["plen"] + p["tlen"] + 1).all()
    w.layout(cigar_cap=152)
    assert (w.pairs[: w.n]["cigar_cap"] == 152).all()


# ---------------------------------------------------------------------------
# command-line contract

class ArgValue(ctypes.Union):
    _fields_ = [("int_val", ctypes.c_int64), ("str_val", ctypes.c_char_p), ("char_val", ctypes.c_char),
                ("float_val", ctypes.c_double), ("bool_val", ctypes.c_bool), ("set", ctypes.c_bool)]


class Option(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("description", ctypes.c_char_p), ("short_arg", ctypes.c_char),
                ("long_arg", ctypes.c_char_p), ("required", ctypes.c_bool), ("parsed", ctypes.c_bool),
                ("type", ctypes.c_int), ("value", ArgValue)]


class Options(ctypes.Structure):
    _fields_ = [("options", ctypes.POINTER(Option)), ("len", ctypes.c_int)]


ARG_INT, ARG_STR, ARG_NO_VALUE = 0, 1, 5
# the wfe.aligner table (src/aligner.cu:75-139)
TABLE = [
    (b"Sequences file", b"File containing the sequences to align.", b"f", b"file", True, ARG_STR),
    (b"Number of alignments", b"Number of alignments to read from the file (default=all alignments)", b"n",
     b"num-alignments", True, ARG_INT),
    (b"Sequence length", b"Maximum sequence length.", b"l", b"seq-len", True, ARG_INT),
    (b"Batch size", b"Number of alignments per batch (default=num-alignments).", b"b", b"batch-size", False,
     ARG_INT),
    (b"Number of CPU threads", b"Number of CPU threads, each CPU thread creates two streams to overlap compute "
     b"and memory transfers. (default=1)", b"t", b"cpu-threads", False, ARG_INT),
    (b"Print CIGARS", b"Print CIGARS to stdout", b"p", b"print-cigars", False, ARG_NO_VALUE),
]


def ref_parser():
    if not os.path.exists(orc.REF_ARG_SO):
        pytest.skip("reference argument parser not built (the reference tree is absent)")
    L = ctypes.CDLL(orc.REF_ARG_SO)
    L.parse_args.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), Options]
    L.parse_args.restype = ctypes.c_bool
    L.print_usage.argtypes = [Options]
    return L


def ref_parse(L, argv):
    arr = (Option * len(TABLE))()
    for i, (name, desc, sa, la, req, ty) in enumerate(TABLE):
        arr[i] = Option(name, desc, sa, la, req, False, ty, ArgValue())
    cargv = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    ok = L.parse_args(len(argv), cargv, Options(arr, len(TABLE)))
    got = {}
    for i, (name, desc, sa, la, req, ty) in enumerate(TABLE):
        if arr[i].parsed:
            got[la.decode()] = (arr[i].value.str_val if ty == ARG_STR else
                                arr[i].value.int_val if ty == ARG_INT else True)
    return bool(ok), got


def our_parse(argv):
    a = ewfa.CliArgs()
    cargv = (ctypes.c_char_p * len(argv))(*[x.encode() for x in argv])
    ok = ewfa.lib().ewfa_cli_parse(len(argv), cargv, ctypes.byref(a))
    return bool(ok), a


ARGVS = [
    ["wfe.aligner"],
    ["wfe.aligner", "-f", "x.txt", "-n", "10", "-l", "150"],
    ["wfe.aligner", "-f", "x.txt", "-n", "10", "-l", "150", "-b", "5", "-t", "2", "-p"],
    ["wfe.aligner", "--file", "y", "--num-alignments", "7", "--seq-len", "1000", "--print-cigars"],
    ["wfe.aligner", "-f", "x", "-n", "abc", "-l", "-3"],
    ["wfe.aligner", "-fz", "x", "-nn", "5", "-lq", "9"],
    ["wfe.aligner", "-f", "-n", "-l", "4"],
    ["wfe.aligner", "-f", "x", "-n", "3"],
    ["wfe.aligner", "-f", "x", "-n", "3", "-l"],
    ["wfe.aligner", "-q", "-f", "x", "--bogus", "-n", "3", "-l", "8", "-"],
    ["wfe.aligner", "-f", "a", "-f", "b", "-n", "1", "-n", "2", "-l", "3"],
    ["wfe.aligner", "-p", "5", "-f", "x", "-n", "3", "-l", "8"],
]


@pytest.mark.parametrize("argv", ARGVS)
def test_cli_parse_matches_reference(argv):
    L = ref_parser()
    rok, rvals = ref_parse(L, argv)
    ok, a = our_parse(argv)
    assert ok == rok
    if "file" in rvals:
        assert a.file == rvals["file"]
    if "num-alignments" in rvals:
        assert a.num_alignments == rvals["num-alignments"]
    if "seq-len" in rvals:
        assert a.seq_len == rvals["seq-len"]
    assert bool(a.batch_set) == ("batch-size" in rvals)
    if a.batch_set:
        assert a.batch_size == rvals["batch-size"]
    assert bool(a.threads_set) == ("cpu-threads" in rvals)
    if a.threads_set:
        assert a.cpu_threads == rvals["cpu-threads"]
    assert bool(a.print_cigars) == ("print-cigars" in rvals)


@contextlib.contextmanager
def captured_stderr():
    with tempfile.TemporaryFile() as tmp:
        saved = os.dup(2)
        os.dup2(tmp.fileno(), 2)
        box = {}
        try:
            yield box
        finally:
            libc = ctypes.CDLL(None)
            libc.fflush(None)
            os.dup2(saved, 2)
            os.close(saved)
            tmp.seek(0)
            box["text"] = tmp.read().decode()


def test_cli_usage_matches_reference():
    L = ref_parser()
    arr = (Option * len(TABLE))()
    for i, (name, desc, sa, la, req, ty) in enumerate(TABLE):
        arr[i] = Option(name, desc, sa, la, req, False, ty, ArgValue())
    with captured_stderr() as ref:
        L.print_usage(Options(arr, len(TABLE)))
    with captured_stderr() as ours:
        ewfa.lib().ewfa_cli_usage()
    ref_lines = ref["text"].splitlines()
    our_lines = ours["text"].splitlines()
    assert len(ref_lines) == 7
    assert our_lines[:7] == ref_lines  # the reference's six options, text for text
    assert len(our_lines) > 7          # then the additions


def test_cli_policy_option_parsed():
    for argv in (["wfe.aligner", "-f", "x", "-n", "3", "-l", "8", "--policy", "p.policy"],
                 ["wfe.aligner", "-u", "p.policy", "-f", "x", "-n", "3", "-l", "8"]):
        ok, a = our_parse(argv)
        assert ok and a.policy == b"p.policy"
    ok, a = our_parse(["wfe.aligner", "-f", "x", "-n", "3", "-l", "8"])
    assert ok and a.policy is None


@pytest.mark.skipif(ewfa.device_count() > 0, reason="checks the no-device path")
def test_cli_without_device_fails_cleanly(tmp_path):
    r = subprocess.run([ewfa.ALIGNER_PATH, "-f", "x", "-n", "1", "-l", "1"], capture_output=True, timeout=60)
    assert r.returncode == 1
    assert b"No available HIP devices detected." in r.stderr
