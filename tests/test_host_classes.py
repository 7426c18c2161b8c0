"""The C++ host classes (phylommand_amd/host: tree, sub_model, rate_model,
io) against the golden outputs of the compiled reference.  CPU-only modes of
host_driver; the GPU-backed modes are in test_gpu_host.py."""
import json
import os
import subprocess

import numpy as np
import pytest

from golden_cases import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phylommand_amd", "host", "host_driver")


def drive(*args):
    if not os.path.exists(DRIVER):
        pytest.skip("host_driver not built (python -m phylommand_amd.build)")
    out = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_newick_parse_print_matches_reference(tmp_path):
    d = os.path.join(GOLDEN, "parse")
    inputs = open(os.path.join(d, "input.nwk")).read().strip().split("\n")
    want = open(os.path.join(d, "expected.txt")).read().strip().split("\n")
    for s, w in zip(inputs, want):
        p = tmp_path / "t.nwk"
        p.write_text(s + "\n")
        assert drive("parse", p).strip() == w, s


@pytest.mark.parametrize("n", [5, 8, 16])
def test_nni_lists_match_treebender(n):
    d = os.path.join(GOLDEN, "nni")
    got = drive("nni", os.path.join(d, f"tree{n}.nwk")).strip().split("\n")
    want = open(os.path.join(d, f"expected{n}.txt")).read().strip().split("\n")
    assert got == want


def test_time_mod_length_bit_exact():
    for case in json.load(open(os.path.join(GOLDEN, "time_mod_length.json"))):
        flat = [x for p in case["pairs"] for x in p]
        got = [float(x) for x in drive("tmbl", case["arg"], *flat).split()]
        assert got == case["out"], case["arg"]


def test_pmatrix_bit_exact():
    for key, c in json.load(open(os.path.join(GOLDEN, "pmatrix_refcompat.json"))).items():
        out = drive("pmatrix", c["S"], c["tr"], c["spec"], c["params"], c["freqs"], *c["t"])
        vals = np.array([[float(x) for x in line.split()] for line in out.strip().split("\n")])
        want = np.array(c["P"]).reshape(-1, c["S"])
        np.testing.assert_array_equal(vals, want, err_msg=key)


def test_rate_arguments_build_the_reference_models():
    """-V clade-rate and time-rate arguments, edge cases of their grammar
    included (trailing separators, empty items, items before any rate),
    against the models the compiled reference builds from them."""
    cases = json.load(open(os.path.join(GOLDEN, "rate_args.json")))
    assert len(cases) >= 30
    for c in cases:
        assert drive("ratearg", c["kind"], c["arg"]) == c["out"], (c["kind"], c["arg"])


@pytest.mark.parametrize("mode,src,want", [
    ("nexus", "input.nwk", "expected_nexus.txt"),
    ("parsex", "input.nwk", "expected_parsex.txt"),
    ("nexus", "../example_files/tree_file.tree", "expected_example_nexus.txt"),
])
def test_nexus_output_and_annotations_match_reference(mode, src, want):
    """print_nexus_tree_intro / print_tree_to_nexus / print_nexus_tree_end
    with an empty and then a filled Translate table, and trees read with
    "[&...]" annotations kept, against the compiled reference's output."""
    d = os.path.join(GOLDEN, "nexus")
    assert drive(mode, os.path.join(d, src)) == open(os.path.join(d, want)).read()


def test_data_and_tree_readers_match_reference():
    """fasta, relaxed phylip and NEXUS DATA rows, and NEXUS TREES (Translate
    table, comments, several trees), against what the reference's
    file_parser + matrix_parser read from the same files."""
    d = os.path.join(GOLDEN, "formats")
    cases = json.load(open(os.path.join(d, "expected.json")))
    assert len(cases) == 6
    for c in cases:
        *mode, name = c["args"]
        assert drive(*mode, os.path.join(d, name)) == c["out"], c["args"]
