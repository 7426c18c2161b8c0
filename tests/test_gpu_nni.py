"""NNI neighbourhood scoring with path-to-root partial recompute, against
the oracle evaluating every variant tree from scratch."""
import numpy as np
import pytest

from phylommand_amd import synth
from helpers import gpu_context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m,C,S,mode,seed", [
    (8, 300, 1, 4, 0, 1), (16, 500, 4, 4, 1, 2), (40, 2000, 4, 4, 1, 3), (24, 800, 2, 4, 0, 4),
    (12, 400, 4, 20, 1, 5), (5, 100, 1, 4, 0, 6), (64, 4000, 4, 4, 0, 7),
])
def test_nni_scores_match_full_reevaluation(oracle, n, m, C, S, mode, seed):
    print(f"nni n={n} m={m} C={C} S={S} mode={mode}", flush=True)
    kw = dict(rate_dist="gamma", freq_conc=2.0) if S > 4 else {}
    wl = synth.gtr_gamma_workload(n_tips=n, m=m, C=C, S=S, alpha=0.7, gap_frac=0.05, seed=seed, **kw)
    t = wl.tree
    with gpu_context(wl) as ctx:
        base, variants = ctx.nni_score(mode)
    want = oracle.nni_list(t.left, t.right, t.root, mode)
    assert [(v, s) for v, s, _ in variants] == want
    assert len(want) == (2 * (n - 4) if mode == 0 else 2 * (n - 3))
    mb = np.zeros((t.n_nodes, m), np.uint32)
    mb[:n] = wl.masks
    ref_base, _ = oracle.Problem(t.left, t.right, t.root, t.blen, mb, S, wl.Q, wl.pi, pi=wl.pi,
                                 cat_rates=wl.cat_rates, cat_w=wl.cat_weights).loglik()
    assert abs(base - ref_base) <= 1e-3
    for v, side, lnl in variants:
        l2, r2 = oracle.nni_apply(t.left, t.right, t.root, v, side)
        ref, _ = oracle.Problem(l2, r2, t.root, t.blen, mb, S, wl.Q, wl.pi, pi=wl.pi, cat_rates=wl.cat_rates,
                                cat_w=wl.cat_weights).loglik()
        assert abs(lnl - ref) <= 1e-3 + 1e-9 * abs(ref), (v, side, lnl, ref)


@pytest.mark.parametrize("n,m,C,S,n_check,seed", [(40, 7000, 4, 4, None, 8), (64, 2100, 2, 20, 10, 9)])
def test_nni_several_variants_a_workgroup_match_oracle(oracle, n, m, C, S, n_check, seed):
    """Program-major neighbourhood launches with a site-block count that is
    not a multiple of 8: the DNA case fills the GPU with four variants a
    workgroup (at 256 CUs) and ends with a group of two; the 20-state case
    (one variant a workgroup) checks a seeded sample of variants that
    always includes the last two."""
    print(f"nni several a workgroup n={n} m={m} C={C} S={S}", flush=True)
    kw = dict(rate_dist="gamma", freq_conc=2.0) if S > 4 else {}
    wl = synth.gtr_gamma_workload(n_tips=n, m=m, C=C, S=S, alpha=0.7, gap_frac=0.05, seed=seed, **kw)
    t = wl.tree
    with gpu_context(wl) as ctx:
        base, variants = ctx.nni_score(1)
        base2, variants2 = ctx.nni_score(1)
    assert len(variants) == 2 * (n - 3) and len(variants) % 4 == 2
    assert base == base2 and [x[2] for x in variants] == [x[2] for x in variants2]
    mb = np.zeros((t.n_nodes, m), np.uint32)
    mb[:n] = wl.masks
    idx = list(range(len(variants)))
    if n_check is not None:
        rng = np.random.default_rng(seed)
        idx = sorted(set(rng.choice(len(variants) - 2, size=n_check - 2, replace=False).tolist())
                     | {len(variants) - 2, len(variants) - 1})
    for i in idx:
        v, side, lnl = variants[i]
        l2, r2 = oracle.nni_apply(t.left, t.right, t.root, v, side)
        ref, _ = oracle.Problem(l2, r2, t.root, t.blen, mb, S, wl.Q, wl.pi, pi=wl.pi, cat_rates=wl.cat_rates,
                                cat_w=wl.cat_weights).loglik()
        assert abs(lnl - ref) <= 1e-3 + 1e-9 * abs(ref), (i, v, side, lnl, ref)
