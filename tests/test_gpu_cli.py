"""wfe.aligner end to end on the GPU: the same stdout lines as the
reference's CLI (src/aligner.cu, src/wavefront.cu:293-300) for the same pair
file, with the CIGAR text checked against the oracle's."""
import json
import os
import re
import subprocess

import pytest

import ewfa
import oracle_py as orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


def pair_file(tmp_path, pairs, name="pairs.txt"):
    path = os.path.join(tmp_path, name)
    with open(path, "wb") as f:
        for p, t in pairs:
            f.write(b">" + p + b"\n<" + t + b"\n")
    return path


def run(args, timeout=120):
    r = subprocess.run([ewfa.ALIGNER_PATH] + args, capture_output=True, timeout=timeout)
    return r.returncode, r.stdout.decode(), r.stderr.decode()


def sample_pairs(n=300):
    doc = json.load(open(os.path.join(GOLDEN, "readme_pairs.json")))
    pairs = [(c["pattern"].encode(), c["text"].encode()) for c in doc["cases"]]
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c1"]), n, first=5000)
    pairs += [(w.pattern(i), w.text(i)) for i in range(w.n)]
    return pairs


def expected_lines(pairs, max_distance=0):
    out = []
    for i, (p, t) in enumerate(pairs):
        d, s = orc.stream(p, t, max_distance)
        out.append(f"{i}: {orc.text(s)}")
    return out


def cigar_lines(stdout):
    return [l for l in stdout.splitlines() if re.match(r"^\d+: ", l) or re.match(r"^\d+:$", l)]


def test_print_cigars_matches_oracle(tmp_path):
    pairs = sample_pairs()
    path = pair_file(tmp_path, pairs)
    say(f"case cli -p {len(pairs)} pairs")
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "150", "-p"])
    assert rc == 0, err
    lines = out.splitlines()
    assert re.match(r'^Using device ".*", with compute capability \d+\.\d+$', lines[0])
    assert re.match(r"^Aligned executed\. Wall time \d+\.\d+s$", lines[-1])
    assert lines[1:-1] == expected_lines(pairs)


def test_threads_batches_and_plain(tmp_path):
    pairs = sample_pairs(200)
    path = pair_file(tmp_path, pairs)
    want = expected_lines(pairs)
    say("case cli -t 3 -b 17 / --plain / -t 4 default batch")
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "150", "-p", "-t", "3", "-b", "17"])
    assert rc == 0, err
    got = cigar_lines(out)
    assert sorted(got, key=lambda l: int(l.split(":")[0])) == want
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "150", "-p", "--plain", "-b", "50"])
    assert rc == 0, err
    assert cigar_lines(out) == want
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "150", "-p", "-t", "4"])
    assert rc == 0, err
    assert "[Thread 0] Changing batch size to 50" in err
    assert sorted(cigar_lines(out), key=lambda l: int(l.split(":")[0])) == want


def test_policy_option_and_pieced_upload(tmp_path):
    """--policy with the c2 policy file (one bucket, the banded kernel) and
    with the mixed-length one gives the same lines as the oracle, over
    batches whose words arrive in 16 MiB pieces (about 150,000 pairs of
    1,000 bp read 9.6 M words, five pieces); an unreadable policy file ends
    the run with status 1."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pol_dir = os.path.join(root, "ewfa-gpu_amd", "policy")
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c2"]), 150_000, first=11)
    path = os.path.join(tmp_path, "c2.txt")
    with open(path, "wb") as f:
        for i in range(w.n):
            f.write(b">" + w.pattern(i) + b"\n<" + w.text(i) + b"\n")
    idx = list(range(0, w.n, 997)) + [w.n - 1]
    ok, od, osl, ost = orc.batch_workload(w, idx=idx, threads=8)
    want = {i: f"{i}: {orc.text(ost[j])}" for j, i in enumerate(idx)}
    for pol in ("c2_1000bp_5pct.policy", "c3_mixed_100_10000bp.policy"):
        say(f"case cli --policy {pol}: {w.n} pairs, -b 40000")
        rc, out, err = run(["-f", path, "-n", str(w.n), "-l", "1000", "-b", "40000", "-p",
                            "--policy", os.path.join(pol_dir, pol)], timeout=300)
        assert rc == 0, err
        got = {}
        for line in out.splitlines():
            m = re.match(r"^(\d+): ", line)
            if m and int(m.group(1)) in want:
                got[int(m.group(1))] = line
        assert got == want
    rc, out, err = run(["-f", path, "-n", "10", "-l", "1000", "--policy", os.path.join(tmp_path, "none.policy")])
    assert rc == 1 and "launch policy" in err


def test_print_to_a_file_matches_the_pipe(tmp_path):
    """With stdout a regular file the batch text goes out by positional
    writes from several threads; the file holds the same lines, in the same
    order, as the piped output (one worker and three), and the wall-time line
    comes last."""
    w = ewfa.Workload(ewfa.synth_spec(**ewfa.CONFIGS["c2"]), 30_000, first=77)
    path = os.path.join(tmp_path, "c2.txt")
    with open(path, "wb") as f:
        for i in range(w.n):
            f.write(b">" + w.pattern(i) + b"\n<" + w.text(i) + b"\n")
    for extra in ([], ["-t", "3"]):
        args = ["-f", path, "-n", str(w.n), "-l", "1000", "-b", "7000", "-p"] + extra
        say(f"case cli -p to a file {extra}")
        rc, piped, err = run(args, timeout=300)
        assert rc == 0, err
        sink = os.path.join(tmp_path, "out.txt")
        with open(sink, "wb") as so:
            r = subprocess.run([ewfa.ALIGNER_PATH] + args, stdout=so, stderr=subprocess.PIPE, timeout=300)
        assert r.returncode == 0, r.stderr.decode()
        lines = open(sink, "rb").read().decode().splitlines()
        plines = piped.splitlines()
        assert lines[0] == plines[0] and lines[-1].startswith("Aligned executed. Wall time")
        if extra:  # workers interleave whole batches: compare as sets of lines
            assert sorted(lines[1:-1]) == sorted(plines[1:-1])
        else:
            assert lines[1:-1] == plines[1:-1]
        assert len(lines) == w.n + 2


def test_without_print_only_two_lines(tmp_path):
    pairs = sample_pairs(50)
    path = pair_file(tmp_path, pairs)
    say("case cli no -p")
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "150"])
    assert rc == 0, err
    lines = out.splitlines()
    assert len(lines) == 2 and lines[0].startswith("Using device")


def test_short_file_pads_with_empty_pairs(tmp_path):
    pairs = sample_pairs(5)
    path = pair_file(tmp_path, pairs)
    say("case cli -n beyond the file")
    rc, out, err = run(["-f", path, "-n", "12", "-l", "150", "-p"])
    assert rc == 0, err
    got = cigar_lines(out)
    assert got[: len(pairs)] == expected_lines(pairs)
    assert got[len(pairs):] == [f"{i}: " for i in range(len(pairs), 12)]


def test_reference_cap_and_check(tmp_path):
    doc = json.load(open(os.path.join(GOLDEN, "cap64.json")))
    pairs = [(c["pattern"].encode(), c["text"].encode()) for c in doc["cases"] if not c["undefined"]]
    path = pair_file(tmp_path, pairs)
    say(f"case cli --max-distance 64 ({len(pairs)} pairs) and --check")
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "3000", "-p", "--max-distance", "64"])
    assert rc == 0, err
    assert cigar_lines(out) == expected_lines(pairs, 64)
    capped = [i for i, (p, t) in enumerate(pairs) if orc.align(p, t, 64)[2]]
    assert capped
    for i in capped:
        assert f"Max distance reached at alignment {i}" in out
    rc, out, err = run(["-f", path, "-n", str(len(pairs)), "-l", "3000", "--check"])
    assert rc == 0, err
    assert f"Correct alignments: {len(pairs)}/{len(pairs)}" in err


def test_error_paths(tmp_path):
    pairs = sample_pairs(4)
    path = pair_file(tmp_path, pairs)
    say("case cli error paths")
    rc, out, err = run([])
    assert rc == 1 and out.startswith("Using device") and err.startswith("Options:\n")
    rc, out, err = run(["-f", path, "-n", "4", "-l", "150", "-b", "5"])
    assert rc == 1 and "batch_size must be <= than the number of alignments." in err
    rc, out, err = run(["-f", path, "-n", "4", "-l", "150", "-t", "0"])
    assert rc == 1 and "Minimum needed is 1 thread" in err
    bad = os.path.join(tmp_path, "bad.txt")
    with open(bad, "wb") as f:
        f.write(b"<ACGT\n>ACGT\n")
    rc, out, err = run(["-f", bad, "-n", "1", "-l", "10"])
    assert rc == 1 and "Could not read the sequences from file" in err
