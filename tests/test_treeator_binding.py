"""The drop-in at the class level: phylommand's own treeator.cpp, unmodified,
built against the host classes (phylommand_amd/host/treeator.mk) instead of
src/simpleML.cpp, src/sub_model.cpp and src/rate_model.cpp.

CPU: the recipe builds and links from the phylommand sources (skipped where
they are absent), the likelihood calls resolve to this project's classes,
and the program runs.  GPU: the built treeator runs the commands whose
output the reference's own treeator wrote (tests/golden/treeator_cli): with
PHYLO_AMD_P_BUILDER=refcompat and PHYLO_AMD_PRECISION=fp64 its stdout is the
reference's, byte for byte; with the defaults (eigen P, FP32 partials) every
number agrees to 1e-5 relative."""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("PHYLOMMAND_SRC", "/root/reference/src")
GOLD = os.path.join(ROOT, "tests", "golden", "treeator_cli")
EXAMPLES = os.path.join(ROOT, "tests", "golden", "example_files")
BUILT = os.path.join(ROOT, "phylommand_amd", "host", "_treeator", "treeator")
HOST_LIB = os.path.join(ROOT, "phylommand_amd", "libphylommand_host.so")


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "treeator.cpp")), reason="phylommand sources absent")
def test_treeator_builds_against_the_host_classes(tmp_path):
    if not os.path.exists(HOST_LIB):
        pytest.skip("host library not built (python -m phylommand_amd.build)")
    out = subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "phylommand_amd", "host", "treeator.mk"),
                          f"REF={REF}", f"OUT={tmp_path}"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    exe = tmp_path / "treeator"
    assert exe.exists()
    # the likelihood path is this project's: treeator.o calls phylo_amd's
    # classes, and nothing in the binary defines phylommand's own simpleML
    obj = subprocess.run(["nm", "-C", str(tmp_path / "treeator.o")], capture_output=True, text=True).stdout
    for sym in ("phylo_amd::simpleML::calculate_log_likelihood", "phylo_amd::simpleML::set_char",
                "phylo_amd::simpleML::simulate_chars", "phylo_amd::simpleML::draw_normalized_likelihood_on_nodes",
                "phylo_amd::tree::tree_file_parser", "phylo_amd::tree::print_tree_to_nexus",
                "phylo_amd::sub_model::sub_model", "phylo_amd::rate_model::pars_clade_rates_arg"):
        assert sym in obj, sym
    binsyms = subprocess.run(["nm", "-C", str(exe)], capture_output=True, text=True).stdout
    assert not re.search(r" [TW] simpleML::", binsyms)
    assert not re.search(r" [TW] sub_model::", binsyms)
    h = subprocess.run([str(exe), "-h"], capture_output=True, text=True, timeout=60)
    assert "Treeator" in h.stdout


def _numbers(text):
    # every number in the text, integers included (a probability the
    # reference prints as "1" may print as "0.999999" here)
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?", text)]


@pytest.mark.gpu
@pytest.mark.parametrize("exact", [True, False])
def test_treeator_cli_matches_the_reference(exact):
    if not os.path.exists(BUILT):
        pytest.skip("treeator binding not built here (needs the phylommand sources at build time)")
    cmds = json.load(open(os.path.join(GOLD, "commands.json")))
    env = dict(os.environ)
    if exact:
        env.update(PHYLO_AMD_P_BUILDER="refcompat", PHYLO_AMD_PRECISION="fp64")
    for name, args in cmds.items():
        print(f"treeator {name} exact={exact}", flush=True)
        out = subprocess.run([BUILT, *args], capture_output=True, text=True, timeout=120, cwd=EXAMPLES, env=env)
        assert out.returncode == 0, (name, out.stderr[-500:])
        want = open(os.path.join(GOLD, name + ".out")).read()
        if exact:
            assert out.stdout == want, name
        else:
            got_n, want_n = _numbers(out.stdout), _numbers(want)
            assert len(got_n) == len(want_n), name
            for a, b in zip(got_n, want_n):
                assert abs(a - b) <= 1e-5 * max(1.0, abs(b)), (name, a, b)
