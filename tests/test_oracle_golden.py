"""Pins the FP64 oracle against golden vectors written by the compiled
reference (tests/golden/make_golden.py).  CPU only."""
import json
import os
import re

import numpy as np
import pytest

from golden_cases import GOLDEN, PERSITE_CASES, Case, dlist


@pytest.mark.parametrize("name", PERSITE_CASES)
def test_persite_lnl_bit_exact_refcompat(oracle, name):
    """With the reference's own P builder the oracle repeats simpleML's
    arithmetic operation for operation, so per-site values are bit-identical."""
    c = Case(oracle, name)
    tot, site = c.problem(oracle).loglik()
    assert site.shape == c.expected_site.shape
    np.testing.assert_array_equal(site, c.expected_site)
    assert tot == c.expected_total or abs(tot - c.expected_total) <= 1e-12 * abs(c.expected_total)


@pytest.mark.parametrize("name", PERSITE_CASES)
def test_persite_lnl_exact_builders_close(oracle, name):
    """EIG / PADE P(t) differ from the reference's squaring by its own ~1e-7
    relative P error (survey probe B4), and nothing more."""
    c = Case(oracle, name)
    for b in ([oracle.EIG, oracle.PADE] if c.tr else [oracle.PADE]):
        _, site = c.problem(oracle, builder=b).loglik()
        rel = np.abs(site - c.expected_site) / np.maximum(np.abs(c.expected_site), 1e-300)
        assert rel.max() < 2e-6, (b, rel.max())


@pytest.mark.parametrize("model", ["jc", "k2"])
def test_example_files_all_sites(oracle, model):
    """The reference's own example_files (103 taxa), every site, root edges zeroed."""
    meta = json.load(open(os.path.join(GOLDEN, "example_files", "case.json")))
    c = Case(oracle, "example_files", tree_file="tree_file.tree", aln_file="alignment_file.fst",
             expected=f"expected_{model}.txt", model=meta["models"][model])
    tot, site = c.problem(oracle).loglik()
    np.testing.assert_array_equal(site, c.expected_site)
    assert abs(tot - c.expected_total) <= 1e-12 * abs(c.expected_total)
    # the CLI scores only character 0 (treeator.cpp:723-724)
    assert np.isfinite(site[0])


def test_pmatrix_refcompat_bit_exact(oracle):
    pm = json.load(open(os.path.join(GOLDEN, "pmatrix_refcompat.json")))
    for key, c in pm.items():
        S, tr = c["S"], c["tr"]
        spec = [int(x) for x in dlist(c["spec"])]
        fr = dlist(c["freqs"])
        Q = oracle.build_Q(S, tr, spec, dlist(c["params"]), fr if (tr and fr) else None)
        for t, P_ref in zip(c["t"], c["P"]):
            P = oracle.pmatrix(Q, t, oracle.REFCOMPAT)
            np.testing.assert_array_equal(P, np.array(P_ref), err_msg=f"{key} t={t}")


def _cutoffs(arg):
    """rate_model::pars_rates_and_times_arg (rate_model.cpp:153-176)."""
    if arg == "-":
        return [], []
    rates, cuts, cls = [], [], []
    for grp in arg.split(","):
        if ":" in grp:
            r, tm = grp.split(":")
            rates.append(np.float32(float(r)))
        else:
            tm = grp
        cuts.append((np.float32(float(tm)), len(rates) - 1))
    cuts.sort(key=lambda x: x[0])
    return [x[0] for x in cuts], [rates[x[1]] for x in cuts]


def test_time_mod_length_bit_exact(oracle):
    for case in json.load(open(os.path.join(GOLDEN, "time_mod_length.json"))):
        cuts, crates = _cutoffs(case["arg"])
        for (b, d), want in zip(case["pairs"], case["out"]):
            got = oracle.time_mod_length(b, d, cuts, crates)
            assert got == want, (case["arg"], b, d, got, want)


def _mrca(tree, taxa):
    anc = None
    for name in taxa:
        v = tree.labels.index(name)
        path = []
        while v >= 0:
            path.append(v)
            v = tree.parent[v]
        anc = path if anc is None else [x for x in anc if x in set(path)]
    return anc[0]


def test_rate_model_likelihoods(oracle):
    d = os.path.join(GOLDEN, "ratemodel_8x50")
    tree = oracle.OTree(open(os.path.join(d, "tree.nwk")).read())
    names, masks = oracle.parse_fasta(open(os.path.join(d, "aln.fst")).read())
    mb = np.zeros((tree.n_nodes, masks.shape[1]), np.uint32)
    for v in tree.tips():
        mb[v] = masks[names.index(tree.labels[v])]
    pi = oracle.freqs(4, [0.25] * 3)
    Q = oracle.build_Q(4, 1, [0] * 6, [4 / 3], [0.25] * 3)
    for case in json.load(open(os.path.join(d, "expected.json"))):
        cuts, crates = _cutoffs(case["time_arg"])
        clade_of = None
        crate = []
        if case["clade_arg"] != "-":
            r, taxa = case["clade_arg"].split(":")
            clade_of = -np.ones(tree.n_nodes, np.int32)
            clade_of[_mrca(tree, taxa.split(","))] = 0
            crate = [np.float32(float(r))]
        eff = oracle.effective_lengths(tree.left, tree.right, tree.blen, tree.root, cuts, crates, clade_of, crate)
        p = oracle.Problem(tree.left, tree.right, tree.root, eff, mb, 4, Q, pi, pi=pi, builder=oracle.REFCOMPAT)
        tot, _ = p.loglik()
        assert abs(tot - case["total"]) <= 1e-12 * abs(case["total"]), (case, tot)


def test_node_partial_labels(oracle):
    """draw_normalized_likelihood_on_nodes, including its running sum across
    sites (simpleML.cpp:204,211), at the 6 significant digits it prints."""
    d = os.path.join(GOLDEN, "nodes_6x3")
    tree = oracle.OTree(open(os.path.join(d, "tree.nwk")).read())
    names, masks = oracle.parse_fasta(open(os.path.join(d, "aln.fst")).read())
    mb = np.zeros((tree.n_nodes, masks.shape[1]), np.uint32)
    for v in tree.tips():
        mb[v] = masks[names.index(tree.labels[v])]
    fr = [0.1, 0.2, 0.3]
    pi = oracle.freqs(4, fr)
    Q = oracle.build_Q(4, 1, [0, 1, 0, 0, 1, 0], [1.0, 0.25], fr)
    eff = oracle.effective_lengths(tree.left, tree.right, tree.blen, tree.root)
    p = oracle.Problem(tree.left, tree.right, tree.root, eff, mb, 4, Q, pi, pi=pi, builder=oracle.REFCOMPAT)
    out = p.node_partials()
    text = open(os.path.join(d, "expected.txt")).read().split("\n", 1)[1]
    labels = re.findall(r"\)\[&(.*?)\]", text)
    # internal nodes in the order their ')' is printed = postorder, left first
    order = []

    def post(v):
        if tree.left[v] >= 0:
            post(tree.left[v])
            post(tree.right[v])
            order.append(v)
    post(tree.root)
    assert len(labels) == len(order)
    for v, lab in zip(order, labels):
        vals = [[float(x) for x in blk.split(",")] for blk in re.findall(r"P=\{(.*?)\}", lab)]
        np.testing.assert_allclose(out[v], np.array(vals), rtol=6e-6, atol=1e-12)


def test_newick_parse_and_print(oracle):
    d = os.path.join(GOLDEN, "parse")
    inputs = open(os.path.join(d, "input.nwk")).read().strip().split("\n")
    expected = open(os.path.join(d, "expected.txt")).read().strip().split("\n")
    for s, want in zip(inputs, expected):
        assert oracle.OTree(s).printed == want, s


@pytest.mark.parametrize("n", [5, 8, 16])
def test_nni_reference_enumeration(oracle, n):
    d = os.path.join(GOLDEN, "nni")
    got = oracle.nni_reference(open(os.path.join(d, f"tree{n}.nwk")).read())
    want = open(os.path.join(d, f"expected{n}.txt")).read().strip().split("\n")
    assert len(want) == 1 + 2 * (n - 4)
    assert got == want
