"""The oracle and the host classes under AddressSanitizer and
UndefinedBehaviorSanitizer (host code only; CPU).  The oracle's entry points
run through oracle/san_driver.c; the host classes through a sanitized
host_driver over every mode that needs no GPU (newick and NEXUS reading and
writing, NNI lists, the three data readers, rate-model arguments,
time-modified lengths, the reference-compatible P(t), the model-fit vector),
whose output must equal the normal build's."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
DRIVER = os.path.join(ROOT, "phylommand_amd", "host", "host_driver")


def _clean(err):
    return "AddressSanitizer" not in err and "runtime error" not in err and "LeakSanitizer" not in err


def test_oracle_under_sanitizers():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    out = subprocess.run([os.path.join(ROOT, "oracle", "_san", "san_driver"),
                          os.path.join(G, "example_files", "tree_file.tree"),
                          os.path.join(G, "example_files", "alignment_file.fst")],
                         capture_output=True, text=True, timeout=300, env=ENV)
    assert out.returncode == 0 and _clean(out.stderr), out.stderr[-3000:]
    assert out.stdout.strip().endswith("ok")


CASES = [
    ("parse", f"{G}/parse/input.nwk"),
    ("parsex", f"{G}/nexus/input.nwk"),
    ("nexus", f"{G}/example_files/tree_file.tree"),
    ("nni", f"{G}/nni/tree16.nwk"),
    ("matrix", "nexus", f"{G}/formats/data.nex"),
    ("matrix", "phylip", f"{G}/formats/data.phy"),
    ("matrix", "fasta", f"{G}/example_files/alignment_file.fst"),
    ("ntrees", f"{G}/formats/data.nex"),
    ("ntrees", f"{G}/formats/quirks.nex"),
    ("ratearg", "clade", "2:a,b;0.5:c;;d"),
    ("ratearg", "time", "2:0.1,0.3,1:0.2,"),
    ("tmbl", "2:0.05,0.5:0.2", "0.1", "0.0", "0.3", "0.02", "0.01", "0.5"),
    ("pmatrix", "4", "1", "0,1,2,2,1,0", "1.0,2.5,0.7", "0.3,0.2,0.15,0.35", "0.01", "0.5", "3.0"),
    ("objective", f"{G}/dna3_24x300/tree.nwk", f"{G}/dna3_24x300/aln.fst", "dna", "1", "0,1,2,2,1,0",
     "1.0,2.5,0.7", "0.3,0.2,0.15,0.35", "0", "0", "-", "2:0.05,0.5:0.2", "-", "t"),
]


def test_host_classes_under_sanitizers(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("host_driver not built (python -m phylommand_amd.build)")
    from phylommand_amd import build as B
    san = B.build_host_sanitized(str(tmp_path))
    for args in CASES:
        want = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
        got = subprocess.run([san, *args], capture_output=True, text=True, timeout=300, env=ENV)
        assert got.returncode == want.returncode and _clean(got.stderr), (args, got.stderr[-3000:])
        assert got.stdout == want.stdout, args
