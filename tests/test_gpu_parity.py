"""GPU parity: libphylo_hip (through its C ABI) against the FP64 oracle on the
same seeded inputs."""
import numpy as np
import pytest

from phylommand_amd import synth
from helpers import gpu_context, oracle_problem

pytestmark = pytest.mark.gpu

CASES = [
    # n_tips, m, C, alpha, gap, seed
    (16, 1000, 1, 1.0, 0.0, 1),
    (32, 3000, 4, 0.5, 0.05, 2),
    (64, 5000, 4, 0.5, 0.05, 3),
    (128, 20000, 4, 0.5, 0.05, 4),
    (9, 777, 2, 1.5, 0.1, 5),
    (50, 4099, 8, 0.3, 0.02, 6),
]


@pytest.mark.parametrize("n,m,C,alpha,gap,seed", CASES)
@pytest.mark.parametrize("prec", [0, 1])
def test_dna_loglik_matches_oracle(oracle, n, m, C, alpha, gap, seed, prec):
    print(f"case n={n} m={m} C={C} prec={prec}", flush=True)
    wl = synth.gtr_gamma_workload(n_tips=n, m=m, C=C, alpha=alpha, gap_frac=gap, seed=seed)
    ref_tot, ref_site = oracle_problem(oracle, wl).loglik()
    with gpu_context(wl, precision=prec) as ctx:
        tot, site = ctx.loglik(per_site=True)
    rel = np.abs(site - ref_site) / np.abs(ref_site)
    if prec == 1:
        assert rel.max() <= 1e-9, rel.max()
        assert abs(tot - ref_tot) <= 1e-9 * abs(ref_tot)
    else:
        # per-site relative error <= 1e-5, total |dlnL| <= 1e-3
        assert rel.max() <= 1e-5, rel.max()
        assert abs(tot - ref_tot) <= 1e-3, abs(tot - ref_tot)
    # the total is the compensated sum of the per-site values
    assert abs(tot - np.sum(site)) <= 1e-9 * abs(tot)


WIDE_CASES = [
    # n_tips, m, S, C, alpha, gap, seed
    (8, 500, 20, 1, 1.0, 0.0, 11),
    (64, 20000, 20, 4, 1.0, 0.03, 12),     # the 20-state protein configuration
    (16, 1500, 6, 2, 0.8, 0.05, 13),
    (12, 900, 23, 4, 0.5, 0.05, 14),
    (10, 700, 32, 8, 1.2, 0.02, 15),
    (20, 1200, 13, 4, 0.6, 0.05, 16),
    # more taxa than the matrix-core walker stages in LDS (128): tips read
    # from global memory on the whole-tree path
    (160, 1500, 20, 4, 1.0, 0.03, 17),
    (200, 700, 23, 2, 0.7, 0.05, 18),
]


@pytest.mark.parametrize("n,m,S,C,alpha,gap,seed", WIDE_CASES)
@pytest.mark.parametrize("prec", [0, 1])
def test_wide_state_loglik_matches_oracle(oracle, n, m, S, C, alpha, gap, seed, prec):
    print(f"wide case n={n} m={m} S={S} C={C} prec={prec}", flush=True)
    wl = synth.gtr_gamma_workload(n_tips=n, m=m, S=S, C=C, alpha=alpha, gap_frac=gap, seed=seed,
                                  rate_dist="gamma", freq_conc=2.0)
    ref_tot, ref_site = oracle_problem(oracle, wl).loglik()
    with gpu_context(wl, precision=prec) as ctx:
        tot, site = ctx.loglik(per_site=True)
    rel = np.abs(site - ref_site) / np.abs(ref_site)
    print(f"wide case: total diff {tot - ref_tot:+.3e}, per-site max rel {rel.max():.3e}", flush=True)
    if prec == 1:
        assert rel.max() <= 1e-9, rel.max()
    else:
        assert rel.max() <= 1e-5, rel.max()
        assert abs(tot - ref_tot) <= 1e-3, abs(tot - ref_tot)


@pytest.mark.parametrize("n,m,S,C,prec", [(12, 300, 4, 4, 0), (12, 300, 4, 2, 1), (10, 333, 20, 4, 0),
                                          (9, 200, 20, 4, 1), (8, 150, 7, 2, 0), (8, 150, 29, 1, 0)])
def test_node_partials_match_oracle(oracle, n, m, S, C, prec):
    """phylo_node_partials (every storage layout: DNA float4 rows, wide FP64
    rows, the MFMA walker's operand order) against the oracle's partials,
    compared as value x 2^scaler."""
    print(f"node partials n={n} m={m} S={S} C={C} prec={prec}", flush=True)
    kw = dict(rate_dist="gamma", freq_conc=2.0) if S > 4 else {}
    wl = synth.gtr_gamma_workload(n_tips=n, m=m, S=S, C=C, alpha=0.8, gap_frac=0.05, seed=S + n, **kw)
    clv_o, sc_o, _, _ = oracle_problem(oracle, wl).node_clvs()
    t = wl.tree
    with gpu_context(wl, precision=prec) as ctx:
        ctx.loglik()
        for v in range(t.n_tips, t.n_nodes):
            clv, sc = ctx.node_partials(v, C, S)
            got = clv * np.ldexp(1.0, sc)[:, None, None]
            want = clv_o[v] * np.ldexp(1.0, sc_o[v])[:, None, None]
            scale = np.abs(want).max(axis=(1, 2), keepdims=True)
            tol = 1e-10 if prec == 1 else 2e-5
            np.testing.assert_allclose(got / scale, want / scale, atol=tol, err_msg=f"node {v}")


@pytest.mark.parametrize("S,prec", [(4, 0), (4, 1), (20, 0)])
def test_partial_storing_does_not_change_results(S, prec):
    """loglik stores only the partials its walk reads back unless asked to
    keep them all: totals, per-site values, node partials (completed by one
    storing evaluation) and NNI scores agree bit for bit with the
    keep-everything evaluation."""
    print(f"keep partials S={S} prec={prec}", flush=True)
    kw = dict(rate_dist="gamma", freq_conc=2.0) if S > 4 else {}
    wl = synth.gtr_gamma_workload(n_tips=24, m=517, S=S, C=4, alpha=0.7, gap_frac=0.05, seed=40 + S, **kw)
    t = wl.tree
    out = []
    for keep in (False, True):
        with gpu_context(wl, precision=prec) as ctx:
            ctx.set_keep_partials(keep)
            tot, site = ctx.loglik(per_site=True)
            parts = [ctx.node_partials(v, 4, S) for v in range(t.n_tips, t.n_nodes)]
            base, var = ctx.nni_score(1)
            tot2 = ctx.loglik()
            out.append((tot, site, parts, base, var, tot2))
    (a, b) = out
    assert a[0] == b[0] and a[5] == b[5] == a[0]
    np.testing.assert_array_equal(a[1], b[1])
    for (ca, sa), (cb, sb) in zip(a[2], b[2]):
        np.testing.assert_array_equal(ca, cb)
        np.testing.assert_array_equal(sa, sb)
    assert a[3] == b[3] == a[0]
    np.testing.assert_array_equal(a[4], b[4])


def test_node_partials_follow_changed_inputs(oracle):
    """node_partials after the edge lengths change evaluates the new inputs."""
    wl = synth.gtr_gamma_workload(n_tips=10, m=200, C=4, alpha=0.8, gap_frac=0.05, seed=77)
    t = wl.tree
    eff2 = t.blen * 1.7
    clv_o, sc_o, _, _ = oracle_problem(oracle, wl, eff=eff2).node_clvs()
    with gpu_context(wl, precision=1) as ctx:
        ctx.loglik()
        ctx.set_edge_lengths(eff2)
        for v in range(t.n_tips, t.n_nodes):
            clv, sc = ctx.node_partials(v, 4, 4)
            got = clv * np.ldexp(1.0, sc)[:, None, None]
            want = clv_o[v] * np.ldexp(1.0, sc_o[v])[:, None, None]
            scale = np.abs(want).max(axis=(1, 2), keepdims=True)
            np.testing.assert_allclose(got / scale, want / scale, atol=1e-10, err_msg=f"node {v}")


@pytest.mark.parametrize("S,n,m,C", [(4, 128, 20000, 4), (20, 64, 10000, 4), (4, 64, 20000, 8), (4, 64, 20000, 1)])
@pytest.mark.parametrize("prec", [0, 1])
def test_tail_split_matches_whole_tree_walk(oracle, S, n, m, C, prec):
    """20,000 DNA sites are 313 blocks of 64 (10,000 20-state sites 313
    blocks of 32): past a full round over the GPU's CUs the tail blocks walk
    the tree as subtree pieces plus a short top program.  Per-site values
    against the oracle; lean and storing evaluations agree bit for bit;
    against the whole-tree walk of the tail (phylo_set_tail_split(0)) FP64
    storage is bit-identical and FP32 storage differs only by the rounding of
    the extra stored partials.  Node partials and NNI scores, which read the
    stored partials, follow the same bars.  With 4 rate categories the top
    launch also forms the total (fold_total); otherwise the final-sum kernel
    does."""
    print(f"tail split S={S} n={n} m={m} C={C} prec={prec}", flush=True)
    kw = dict(rate_dist="gamma", freq_conc=2.0) if S > 4 else {}
    wl = synth.gtr_gamma_workload(n_tips=n, m=m, S=S, C=C, alpha=0.5, gap_frac=0.05, seed=21 + S + C, **kw)
    ref_tot, ref_site = oracle_problem(oracle, wl).loglik()
    t = wl.tree
    nodes = [v for v in (t.root, t.left[t.root], t.right[t.root], t.n_tips + 3, t.n_nodes - 2) if v >= t.n_tips]
    res = {}
    for split in (True, False):
        for keep in (False, True):
            with gpu_context(wl, precision=prec) as ctx:
                ctx.set_tail_split(split)
                ctx.set_keep_partials(keep)
                tot, site = ctx.loglik(per_site=True)
                parts = [ctx.node_partials(v, C, S) for v in nodes]
                _, var = ctx.nni_score(1)
                res[(split, keep)] = (tot, site, parts, var)
    for (split, keep), (tot, site, _, _) in res.items():
        rel = np.abs(site - ref_site) / np.abs(ref_site)
        assert rel.max() <= (1e-9 if prec == 1 else 1e-5), (split, keep, rel.max())
        assert abs(tot - ref_tot) <= (1e-9 * abs(ref_tot) if prec == 1 else 1e-3), (split, keep, tot - ref_tot)
    for split in (True, False):
        a, b = res[(split, False)], res[(split, True)]
        assert a[0] == b[0]
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[3], b[3])
    s_, w_ = res[(True, True)], res[(False, True)]
    if prec == 1:
        assert s_[0] == w_[0]
        np.testing.assert_array_equal(s_[1], w_[1])
        np.testing.assert_array_equal(s_[3], w_[3])
        for (ca, sa), (cb, sb) in zip(s_[2], w_[2]):
            np.testing.assert_array_equal(ca, cb)
            np.testing.assert_array_equal(sa, sb)
    else:
        np.testing.assert_allclose(s_[1], w_[1], rtol=1e-6)
        assert abs(s_[0] - w_[0]) <= 1e-3
        np.testing.assert_allclose(s_[3], w_[3], rtol=0, atol=1e-3)
        for (ca, sa), (cb, sb) in zip(s_[2], w_[2]):
            ga = ca * np.ldexp(1.0, sa)[:, None, None]
            gb = cb * np.ldexp(1.0, sb)[:, None, None]
            scale = np.abs(gb).max(axis=(1, 2), keepdims=True)
            np.testing.assert_allclose(ga / scale, gb / scale, atol=1e-6)
