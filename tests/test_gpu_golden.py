"""GPU parity against the golden vectors the compiled reference wrote: the
library is driven through its C ABI on the same trees, alignments and models."""
import json
import os

import numpy as np
import pytest

import phylommand_amd as P
from golden_cases import GOLDEN, PERSITE_CASES, Case

pytestmark = pytest.mark.gpu


def run_case(c, builder, prec):
    masks, left, right, root, eff = c.library_arrays()
    with P.Context(0, prec) as ctx:
        ctx.set_tips(masks, c.S)
        ctx.set_topology(left, right, root)
        ctx.set_edge_lengths(eff)
        ctx.set_model(c.Q, c.root_w, c.pi if c.tr else None, builder)
        ctx.set_categories([1.0], [1.0])
        return ctx.loglik(per_site=True)


@pytest.mark.parametrize("name", PERSITE_CASES)
def test_golden_refcompat_fp64(oracle, name):
    """The reference's own P builder and FP64 partials: agreement to rounding."""
    c = Case(oracle, name)
    print(f"golden {name} S={c.S} m={c.m}", flush=True)
    tot, site = run_case(c, P.P_REFCOMPAT, P.PREC_FP64)
    np.testing.assert_allclose(site, c.expected_site, rtol=1e-11, atol=1e-13)
    assert abs(tot - c.expected_total) <= 1e-11 * abs(c.expected_total)


@pytest.mark.parametrize("name", PERSITE_CASES)
def test_golden_exact_builders(oracle, name):
    """Exact P(t) (eigen for reversible, Pade otherwise): within the reference's
    own P(t) error (~1e-7 relative), FP32 and FP64 partials."""
    c = Case(oracle, name)
    b = P.P_EIGEN if c.tr else P.P_PADE
    for prec, tol in ((P.PREC_FP64, 2e-6), (P.PREC_FP32, 1e-5)):
        tot, site = run_case(c, b, prec)
        rel = np.abs(site - c.expected_site) / np.abs(c.expected_site)
        assert rel.max() <= tol, (prec, rel.max())


@pytest.mark.parametrize("model", ["jc", "k2"])
def test_golden_example_files(oracle, model):
    meta = json.load(open(os.path.join(GOLDEN, "example_files", "case.json")))
    c = Case(oracle, "example_files", tree_file="tree_file.tree", aln_file="alignment_file.fst",
             expected=f"expected_{model}.txt", model=meta["models"][model])
    tot, site = run_case(c, P.P_REFCOMPAT, P.PREC_FP64)
    np.testing.assert_allclose(site, c.expected_site, rtol=1e-11)
    tot32, site32 = run_case(c, P.P_EIGEN, P.PREC_FP32)
    assert np.max(np.abs(site32 - c.expected_site) / np.abs(c.expected_site)) <= 1e-5
    assert abs(tot32 - c.expected_total) <= 1e-3


def test_golden_pmatrix_refcompat(oracle):
    """phylo_get_pmatrix with the REFCOMPAT builder equals the P matrices
    sub_model::set_P_matrix produced, bit for bit."""
    pm = json.load(open(os.path.join(GOLDEN, "pmatrix_refcompat.json")))
    from golden_cases import dlist
    for key, cfg in pm.items():
        S, tr = cfg["S"], cfg["tr"]
        spec = [int(x) for x in dlist(cfg["spec"])]
        fr = dlist(cfg["freqs"])
        Q = oracle.build_Q(S, tr, spec, dlist(cfg["params"]), fr if (tr and fr) else None)
        ts = cfg["t"]
        n = len(ts)
        # a caterpillar whose tips carry the requested branch lengths
        n_nodes = 2 * n + 1 if n > 1 else 3
        ntaxa = n + 1
        left = -np.ones(2 * ntaxa - 1, np.int32)
        right = -np.ones(2 * ntaxa - 1, np.int32)
        prev = 0
        for k in range(1, ntaxa):
            v = ntaxa + k - 1
            left[v], right[v] = prev, k
            prev = v
        eff = np.zeros(2 * ntaxa - 1)
        eff[1:ntaxa] = ts
        masks = np.full((ntaxa, 3), (1 << S) - 1, np.uint32)
        with P.Context(0, P.PREC_FP64) as ctx:
            ctx.set_tips(masks, S)
            ctx.set_topology(left, right, prev)
            ctx.set_edge_lengths(eff)
            ctx.set_model(Q, np.ones(S), None, P.P_REFCOMPAT)
            for k, want in enumerate(cfg["P"]):
                got = ctx.pmatrix(k + 1, 1, S)[0]
                np.testing.assert_array_equal(got, np.array(want), err_msg=f"{key} t={ts[k]}")
