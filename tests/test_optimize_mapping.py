"""treeator's model fit (treeator.cpp:738-911) through optimize_parameters /
model_objective: the parameter vector (free rates, then n-1 frequencies of a
reversible model unless --fix_freq, then the rate-model parameters not
fixed: cut-offs first), its bounds, its objective (change_non_fixed,
treeator.cpp:1152-1182) and the fit.

The vector layout and bounds need no GPU; the objective's values are checked
on the GPU against direct evaluations of the model the vector describes
(same engine, FP64, the reference's P builder), invalid points must score
-infinity (the reference returns 0.0, treeator.cpp:1123), and fits must reach
at least the lnL of the parameters the data were drawn from."""
import math
import os
import subprocess

import pytest

from golden_cases import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phylommand_amd", "host", "host_driver")
CASE = os.path.join(GOLDEN, "dna3_24x300")          # 3 rates, spec 0,1,2,2,1,0, gaps and ambiguity codes
TREE, ALN = os.path.join(CASE, "tree.nwk"), os.path.join(CASE, "aln.fst")
SPEC = "0,1,2,2,1,0"
ENV = dict(os.environ, PHYLO_AMD_P_BUILDER="refcompat")


def drive(*args, env=ENV):
    if not os.path.exists(DRIVER):
        pytest.skip("host_driver not built (python -m phylommand_amd.build)")
    out = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.strip().split("\n"):
        k, *v = line.split()
        res.setdefault(k, []).append([float(x) for x in v])
    return res


def objective(params, freqs, fixed="-", fix_freq=0, fixed_x="-", targ="-", carg="-", method="o", points=()):
    return drive("objective", TREE, ALN, "dna", 1, SPEC, params, freqs, fixed, fix_freq, fixed_x, targ, carg, method,
                 *[",".join(repr(float(x)) for x in p) for p in points])


def total(params, freqs):
    out = drive("persite", TREE, ALN, "dna", 1, SPEC, params, freqs, "fp64", "refcompat")
    return out["total"][0][0]


def test_parameter_vector_layout_and_bounds():
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fixed="0")
    assert r["x0"][0] == pytest.approx([2.5, 0.7, 0.3, 0.2, 0.15], abs=1e-15)
    assert r["lo"][0] == [0.0] * 5
    assert r["hi"][0][:2] == [1e6, 1e6] and r["hi"][0][2:] == [1.0, 1.0, 1.0]
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fix_freq=1)       # --fix_freq: rates only
    assert r["x0"][0] == pytest.approx([1.0, 2.5, 0.7])
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fixed="0,2", fix_freq=1)
    assert r["x0"][0] == pytest.approx([2.5])
    # rates through time: cut-offs, then time rates; every cut-off starts at the
    # first one's value (get_cut_off(0), treeator.cpp:774) and is bounded by
    # the longest root-to-tip distance
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fixed="0", fix_freq=1, targ="2:0.05,0.5:0.2", method="t")
    x0, hi = r["x0"][0], r["hi"][0]
    assert x0 == pytest.approx([2.5, 0.7, 0.05, 0.05, 2.0, 0.5], rel=1e-6)
    assert hi[2] == hi[3] and 0.2 < hi[2] < 10 and hi[4:] == [1e6, 1e6]
    # a fixed rate-model parameter drops out (the first cut-off here)
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fixed="0", fix_freq=1, fixed_x="0", targ="2:0.05,0.5:0.2",
                  method="t")
    assert r["x0"][0] == pytest.approx([2.5, 0.7, 0.05, 2.0, 0.5], rel=1e-6)
    # clade rates: one per rate class after the model's parameters
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fix_freq=1, carg="3.0:t1,t2", method="A")
    assert r["x0"][0] == pytest.approx([1.0, 2.5, 0.7, 3.0], rel=1e-6)


@pytest.mark.gpu
def test_objective_values_and_invalid_points():
    print("optimize mapping: objective values", flush=True)
    pts = [[2.5, 0.7, 0.3, 0.2, 0.15], [1.7, 0.4, 0.25, 0.25, 0.25], [-1.0, 0.7, 0.3, 0.2, 0.15],
           [2.5, 0.7, 0.6, 0.6, 0.1]]
    f = [v[0] for v in objective("1.0,1.0,1.0", "-", fixed="0", points=pts)["f"]]
    assert f[0] == pytest.approx(total("1.0,2.5,0.7", "0.3,0.2,0.15,0.35"), rel=1e-12, abs=1e-9)
    assert f[1] == pytest.approx(total("1.0,1.7,0.4", "0.25,0.25,0.25,0.25"), rel=1e-12, abs=1e-9)
    assert f[2] == -math.inf and f[3] == -math.inf
    # with --fix_freq the frequencies are the model's own
    f = [v[0] for v in objective("1.0,1.0,1.0", "0.3,0.2,0.15,0.35", fixed="0", fix_freq=1,
                                 points=[[2.5, 0.7]])["f"]]
    assert f[0] == pytest.approx(total("1.0,2.5,0.7", "0.3,0.2,0.15,0.35"), rel=1e-12, abs=1e-9)
    # a rate-model parameter set: the same as the rate model built directly
    r = objective("1.0,2.5,0.7", "0.3,0.2,0.15,0.35", fixed="0,1,2", fix_freq=1, targ="2:0.05", method="t",
                  points=[[0.03, 1.5]])
    want = drive("ratelik", TREE, ALN, "dna", 1, SPEC, "1.0,2.5,0.7", "0.3,0.2,0.15,0.35", "1.5:0.03", "-")
    assert r["f"][0][0] == pytest.approx(want["total"][0][0], rel=1e-12, abs=1e-9)


@pytest.mark.gpu
def test_fits_reach_the_generating_parameters():
    print("optimize mapping: fits", flush=True)
    truth = total("1.0,2.5,0.7", "0.3,0.2,0.15,0.35")
    r = drive("optimize", TREE, ALN, "dna", 1, SPEC, "1.0,1.0,1.0", "-", "0", 0, "-", "-", "-", "o")
    assert r["lnl"][0][0] >= truth - 1e-3
    assert r["rates"][0][0] == 1.0                      # fixed
    assert sum(r["freqs"][0]) == pytest.approx(1.0)
    r = drive("optimize", TREE, ALN, "dna", 1, SPEC, "1.0,1.0,1.0", "0.3,0.2,0.15,0.35", "0", 1, "-", "-", "-", "o")
    assert r["lnl"][0][0] >= truth - 1e-3
    assert r["freqs"][0] == pytest.approx([0.3, 0.2, 0.15, 0.35], abs=1e-15)   # --fix_freq
    # rates through time: the fit never does worse than its start
    start = drive("ratelik", TREE, ALN, "dna", 1, SPEC, "1.0,2.5,0.7", "0.3,0.2,0.15,0.35", "2:0.05", "-")
    r = drive("optimize", TREE, ALN, "dna", 1, SPEC, "1.0,2.5,0.7", "0.3,0.2,0.15,0.35", "0,1,2", 1, "-", "2:0.05",
              "-", "t")
    assert r["lnl"][0][0] >= start["total"][0][0] - 1e-9
    assert 0.0 <= r["cut_offs"][0][0] and r["time_rates"][0][0] >= 0.0
