"""Edge cases of the likelihood path on the GPU against the FP64 oracle: the
full 128-taxon x 100,000-site configuration, deep trees that need the
scalers, the smallest trees, single sites, all-gap sites and taxa, zero and
very long branches, unary nodes, site weights, wide alphabets at their
limits, and inputs the library must reject."""
import numpy as np
import pytest

import phylommand_amd as P
from phylommand_amd import synth
from helpers import gpu_context, oracle_problem

pytestmark = pytest.mark.gpu


def _tree(left, right, root, blen, n_tips):
    left = np.asarray(left, np.int32)
    right = np.asarray(right, np.int32)
    parent = -np.ones(len(left), np.int32)
    for v in range(len(left)):
        for ch in (left[v], right[v]):
            if ch >= 0:
                parent[ch] = v
    return synth.Tree(n_tips, left, right, parent, np.asarray(blen, float), root,
                      [f"t{i}" for i in range(n_tips)])


def _workload(t, m, C=4, S=4, gap=0.05, seed=1, alpha=0.7):
    rng = np.random.default_rng(seed)
    nr = S * (S - 1) // 2
    pi = rng.dirichlet(np.full(S, 5.0))
    Q = synth.normalise_Q(synth.gtr_Q(rng.gamma(1.0, 1.0, nr), pi), pi)
    cr = synth.gamma_rates(alpha, C) if C > 1 else np.ones(1)
    states = synth.simulate(t, Q, pi, m, rng, cr)
    masks = synth.masks_from_states(states[: t.n_tips], gap, rng, S)
    return synth.Workload(t, masks, Q, pi, None, cr, np.full(C, 1.0 / C), alpha)


def _check(oracle, wl, prec, site_tol=None, tot_tol=None):
    ref_tot, ref_site = oracle_problem(oracle, wl).loglik()
    with gpu_context(wl, precision=prec) as ctx:
        tot, site = ctx.loglik(per_site=True)
    assert np.all(np.isfinite(site))
    # relative to |lnL|, absolute near lnL = 0 (all-gap sites)
    rel = np.abs(site - ref_site) / np.maximum(np.abs(ref_site), 1.0)
    if prec == 1:
        assert rel.max() <= (site_tol or 1e-9), rel.max()
        assert abs(tot - ref_tot) <= 1e-9 * max(1.0, abs(ref_tot))
    else:
        assert rel.max() <= (site_tol or 1e-5), rel.max()
        assert abs(tot - ref_tot) <= (tot_tol or 1e-3), abs(tot - ref_tot)
    return tot, site, ref_site


def caterpillar(n, blen):
    """((((t0,t1),t2),t3)...): internal node n + k joins node (n + k - 1) or
    t0, with tip t(k+1)."""
    nn = 2 * n - 1
    left = -np.ones(nn, np.int32)
    right = -np.ones(nn, np.int32)
    left[n], right[n] = 0, 1
    for k in range(1, n - 1):
        left[n + k], right[n + k] = n + k - 1, k + 1
    b = np.full(nn, blen)
    b[nn - 1] = 0.0
    return _tree(left, right, nn - 1, b, n)


def test_full_size_dna_configuration(oracle):
    """GTR+G4, 128-taxon Yule tree, 100,000 sites, 5 % gaps, FP32 partials."""
    print("full size 128 x 100000", flush=True)
    wl = synth.gtr_gamma_workload(n_tips=128, m=100_000, mean_blen=0.05, alpha=0.5, C=4, gap_frac=0.05, seed=2)
    ref_tot, ref_site = oracle_problem(oracle, wl).loglik()
    with gpu_context(wl, precision=0) as ctx:
        tot, site = ctx.loglik(per_site=True)
    d = site - ref_site
    print(f"full size: total diff {tot - ref_tot:+.3e}, per-site max rel {np.max(np.abs(d) / np.abs(ref_site)):.3e}, "
          f"mean diff {d.mean():+.3e}", flush=True)
    assert np.max(np.abs(d) / np.abs(ref_site)) <= 1e-5
    assert abs(tot - ref_tot) <= 1e-3


@pytest.mark.parametrize("prec", [0, 1])
def test_deep_caterpillar_needs_scaling(oracle, prec):
    """300 taxa in a caterpillar with long branches: site likelihoods near
    1e-200, far below the FP32 range, so every level leans on the scalers."""
    print(f"caterpillar 300 prec={prec}", flush=True)
    t = caterpillar(300, 0.9)
    wl = _workload(t, 700, C=4, gap=0.02, seed=3)
    _, _, ref_site = _check(oracle, wl, prec)
    assert ref_site.min() < -300


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("prec", [0, 1])
def test_smallest_trees(oracle, n, prec):
    if n == 2:
        t = _tree([-1, -1, 0], [-1, -1, 1], 2, [0.3, 0.7, 0.0], 2)
    else:
        t = _tree([-1, -1, -1, 0, 3], [-1, -1, -1, 1, 2], 4, [0.1, 0.2, 0.3, 0.05, 0.0], 3)
    wl = _workload(t, 333, C=2, seed=n)
    _check(oracle, wl, prec)


@pytest.mark.parametrize("m", [1, 63, 64, 65])
def test_site_counts_around_the_block(oracle, m):
    wl = synth.gtr_gamma_workload(n_tips=20, m=m, C=4, alpha=0.6, gap_frac=0.1, seed=m)
    for prec in (0, 1):
        _check(oracle, wl, prec)


def test_all_gap_sites_and_taxa(oracle):
    wl = synth.gtr_gamma_workload(n_tips=24, m=500, C=4, alpha=0.6, gap_frac=0.05, seed=9)
    wl.masks[:, 10] = 15                 # an all-gap site has lnL 0
    wl.masks[5, :] = 15                  # a taxon with no data
    for prec in (0, 1):
        _, site, ref = _check(oracle, wl, prec)
        assert abs(site[10]) < 1e-12 and abs(ref[10]) < 1e-12


@pytest.mark.parametrize("blen", [0.0, 2000.0])
def test_zero_and_saturated_branches(oracle, blen):
    """Every branch 0 (P = I: only constant sites are possible, each with
    probability pi_state) or 2000 (P rows = pi even for the slowest gamma
    category)."""
    wl = synth.gtr_gamma_workload(n_tips=16, m=400, C=4, alpha=0.5, gap_frac=0.0, seed=11)
    t = wl.tree
    t.blen[:] = blen
    t.blen[t.root] = 0.0
    if blen == 0.0:
        wl.masks[:] = np.uint32(1) << np.random.default_rng(1).integers(0, 4, 400).astype(np.uint32)
    for prec in (0, 1):
        _, site, _ = _check(oracle, wl, prec)
    if blen == 0.0:
        state = np.log2(wl.masks[0]).astype(int)
        np.testing.assert_allclose(site, np.log(wl.pi[state]), rtol=1e-6)
    else:
        code = np.log2(np.where(wl.masks == 15, 1, wl.masks)).astype(int)
        want = np.where(wl.masks == 15, 0.0, np.log(wl.pi[code])).sum(axis=0)
        np.testing.assert_allclose(site, want, rtol=1e-6)


def test_unary_node(oracle):
    """A unary internal node (one child) on a path: its partial is P . child."""
    # tips 0..3; 4 = (0,1); 5 = unary above 4; 6 = (2,3); 7 = root (5, 6)
    t = _tree([-1, -1, -1, -1, 0, 4, 2, 5], [-1, -1, -1, -1, 1, -1, 3, 6], 7,
              [0.1, 0.2, 0.15, 0.25, 0.05, 0.3, 0.12, 0.0], 4)
    wl = _workload(t, 400, C=2, seed=5)
    for prec in (0, 1):
        _check(oracle, wl, prec)


def test_site_weights_give_the_weighted_total(oracle):
    wl = synth.gtr_gamma_workload(n_tips=20, m=300, C=4, alpha=0.6, gap_frac=0.05, seed=12)
    w = np.random.default_rng(3).integers(1, 5, 300).astype(float)
    _, ref_site = oracle_problem(oracle, wl).loglik()
    with gpu_context(wl) as ctx:
        ctx.set_site_weights(w)
        tot, site = ctx.loglik(per_site=True)
    assert abs(tot - float(np.dot(w, ref_site))) <= 1e-3
    np.testing.assert_allclose(site, ref_site, rtol=1e-5)


def test_site_weights_reach_the_nni_base(oracle):
    """Weights set after an evaluation that kept every partial: the NNI
    base, the no-op variants and every scored variant use the weights."""
    print("site weights then nni_score: 14 taxa x 400 sites", flush=True)
    wl = synth.gtr_gamma_workload(n_tips=14, m=400, C=4, alpha=0.6, gap_frac=0.05, seed=13)
    t = wl.tree
    w = np.random.default_rng(4).integers(1, 6, 400).astype(float)
    mb = np.zeros((t.n_nodes, 400), np.uint32)
    mb[:14] = wl.masks
    with gpu_context(wl) as ctx:
        ctx.set_keep_partials(True)
        unweighted = ctx.loglik()
        ctx.set_site_weights(w)
        base, variants = ctx.nni_score(1)
        weighted = ctx.loglik()
    _, ref_site = oracle_problem(oracle, wl).loglik()
    assert abs(weighted - float(np.dot(w, ref_site))) <= 1e-3
    assert abs(base - weighted) <= 1e-9 * abs(weighted)
    assert abs(base - unweighted) > 1.0
    for v, side, lnl in variants[:6]:
        l2, r2 = oracle.nni_apply(t.left, t.right, t.root, v, side)
        _, site2 = oracle.Problem(l2, r2, t.root, t.blen, mb, 4, wl.Q, wl.pi, pi=wl.pi, cat_rates=wl.cat_rates,
                                  cat_w=wl.cat_weights).loglik()
        assert abs(lnl - float(np.dot(w, site2))) <= 1e-3, (v, side)


@pytest.mark.parametrize("S,C", [(4, 4), (20, 2)])
def test_eigen_pmatrix_follows_new_lengths(oracle, S, C):
    """P built on the device by the eigen builder, read back after the edge
    lengths change (the build runs on the context's own stream)."""
    print(f"eigen pmatrix after new lengths S={S} C={C}", flush=True)
    kw = dict(rate_dist="gamma", freq_conc=2.0) if S > 4 else {}
    wl = synth.gtr_gamma_workload(n_tips=12, m=100, S=S, C=C, alpha=0.8, gap_frac=0.0, seed=21, **kw)
    t = wl.tree
    with gpu_context(wl) as ctx:
        ctx.loglik()
        for scale in (1.0, 2.5, 0.3):
            eff = t.blen * scale
            ctx.set_edge_lengths(eff)
            for v in (0, 5, t.n_tips + 1):
                got = ctx.pmatrix(v, C, S)
                for k in range(C):
                    want = oracle.pmatrix(wl.Q, eff[v] * wl.cat_rates[k], oracle.EIG, wl.pi)
                    np.testing.assert_allclose(got[k], want, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("S,C", [(5, 1), (17, 2), (19, 4), (20, 8), (21, 2), (32, 4)])
def test_wide_alphabet_limits(oracle, S, C):
    wl = synth.gtr_gamma_workload(n_tips=10, m=130, S=S, C=C, alpha=0.8, gap_frac=0.05, seed=S,
                                  rate_dist="gamma", freq_conc=2.0)
    for prec in (0, 1):
        _check(oracle, wl, prec)


def test_invalid_inputs_are_rejected():
    wl = synth.gtr_gamma_workload(n_tips=6, m=50, C=4, seed=1)
    t = wl.tree
    with P.Context(0, P.PREC_FP32) as ctx:
        ctx.set_tips(wl.masks, 4)
        bad_left = t.left.copy()
        bad_left[t.root] = t.root                  # a node its own child
        with pytest.raises(P.PhyloError):
            ctx.set_topology(bad_left, t.right, t.root)
        ctx.set_topology(t.left, t.right, t.root)
        with pytest.raises(P.PhyloError):
            ctx.set_edge_lengths(np.full(t.n_nodes, np.nan))
        with pytest.raises(P.PhyloError):
            ctx.set_model(np.full((4, 4), np.nan), wl.pi, wl.pi)
        with pytest.raises(P.PhyloError):
            ctx.loglik()                           # no model or lengths yet
        ctx.set_edge_lengths(t.blen)
        ctx.set_model(wl.Q, wl.pi, wl.pi)
        ctx.set_categories(np.ones(3))             # 3 categories: no walker for it
        with pytest.raises(P.PhyloError):
            ctx.loglik()
        ctx.set_categories(wl.cat_rates, wl.cat_weights)
        assert np.isfinite(ctx.loglik())           # the context recovers
        zero = wl.masks.copy()
        zero[0, 0] = 0
        with pytest.raises(P.PhyloError):
            ctx.set_tips(zero, 4)                  # a tip with no state
        ctx.set_tips(wl.masks[:5], 4)              # fewer taxa: the 6-taxon topology no longer fits
        with pytest.raises(P.PhyloError):
            ctx.loglik()


def balanced(depth, blen):
    """A perfectly balanced tree of 2^depth tips: every internal node above
    the cherries has two internal children, so half the results of a walk
    wait for a sibling subtree (carry B or memory)."""
    n = 1 << depth
    nn = 2 * n - 1
    left = -np.ones(nn, np.int32)
    right = -np.ones(nn, np.int32)
    level = list(range(n))
    nxt = n
    while len(level) > 1:
        up = []
        for a, b in zip(level[0::2], level[1::2]):
            left[nxt], right[nxt] = a, b
            up.append(nxt)
            nxt += 1
        level = up
    b = np.full(nn, blen)
    b[nn - 1] = 0.0
    return _tree(left, right, nn - 1, b, n)


def two_caterpillars(k, blen):
    """Two caterpillars of k tips under one root: long chains that walk with
    carry A only, then one wait at the root."""
    n = 2 * k
    nn = 2 * n - 1
    left = -np.ones(nn, np.int32)
    right = -np.ones(nn, np.int32)
    v = n
    tops = []
    for side in range(2):
        t0 = side * k
        left[v], right[v] = t0, t0 + 1
        prev = v
        v += 1
        for j in range(2, k):
            left[v], right[v] = t0 + j, prev     # the chain on the right
            prev = v
            v += 1
        tops.append(prev)
    left[v], right[v] = tops
    b = np.full(nn, blen)
    b[v] = 0.0
    return _tree(left, right, v, b, n)


@pytest.mark.parametrize("shape", ["balanced", "two_caterpillars"])
@pytest.mark.parametrize("keep", [False, True])
def test_tree_shapes_against_oracle(oracle, shape, keep):
    """Shapes at the extremes of the walk's register use, both storing modes,
    both precisions; NNI scores of the balanced tree against the oracle."""
    print(f"shape {shape} keep={keep}", flush=True)
    t = balanced(6, 0.08) if shape == "balanced" else two_caterpillars(20, 0.06)
    wl = _workload(t, 517, C=4, seed=21)
    for prec in (0, 1):
        ref_tot, ref_site = oracle_problem(oracle, wl).loglik()
        with gpu_context(wl, precision=prec) as ctx:
            ctx.set_keep_partials(keep)
            tot, site = ctx.loglik(per_site=True)
            rel = np.abs(site - ref_site) / np.maximum(np.abs(ref_site), 1.0)
            assert rel.max() <= (1e-9 if prec else 1e-5), rel.max()
            if shape == "balanced" and prec == 1:
                base, var = ctx.nni_score(1)
                assert abs(base - ref_tot) <= 1e-9 * abs(ref_tot)
                mb = np.zeros((t.n_nodes, wl.masks.shape[1]), np.uint32)
                mb[: t.n_tips] = wl.masks
                for v, side, lnl in var[::7]:
                    l2, r2 = oracle.nni_apply(t.left, t.right, t.root, v, side)
                    want, _ = oracle.Problem(l2, r2, t.root, t.blen, mb, 4, wl.Q, wl.pi, pi=wl.pi,
                                             cat_rates=wl.cat_rates, cat_w=wl.cat_weights).loglik()
                    assert abs(lnl - want) <= 1e-9 * abs(want), (v, side, lnl, want)


def test_contexts_on_concurrent_threads():
    """Fresh contexts evaluated from several host threads at once (as the
    learned-proposal fits do): every first and later total equals the
    single-threaded one.  The folded total of a split launch needs the
    context's counter cleared before its first launch, on its own stream."""
    import concurrent.futures as cf
    print("8 threads x 12 fresh contexts, 32 taxa x 2000 sites", flush=True)
    wl = synth.gtr_gamma_workload(n_tips=32, m=2000, C=4, alpha=0.6, gap_frac=0.05, seed=31)
    with gpu_context(wl, precision=1) as ctx:
        want = ctx.loglik()

    def job(k):
        out = []
        with gpu_context(wl, precision=1) as c:
            for _ in range(3):
                out.append(c.loglik())
        return out

    with cf.ThreadPoolExecutor(8) as ex:
        got = list(ex.map(job, range(96)))
    assert all(v == want for run in got for v in run), {v for run in got for v in run}
