"""phylo_loglik_noalloc (the allocation-free evaluation) against phylo_loglik
through the C ABI: bit-identical to the context's whole-tree walk (tail split
off), within the FP32 bounds of the default walk and of the FP64 oracle, on
consecutive calls of shrinking size, and a clear error for problems larger
than its scratch array."""
import numpy as np
import pytest

import phylommand_amd as P
from phylommand_amd import synth
from helpers import oracle_problem

pytestmark = pytest.mark.gpu


def _reference(wl, tail_split):
    t = wl.tree
    with P.Context(0, P.PREC_FP32) as ctx:
        ctx.set_tail_split(tail_split)
        ctx.set_tips(wl.masks, 4)
        ctx.set_topology(t.left, t.right, t.root)
        ctx.set_edge_lengths(t.blen)
        ctx.set_model(wl.Q, wl.pi, wl.pi, P.P_EIGEN)
        ctx.set_categories(wl.cat_rates, wl.cat_weights)
        return ctx.loglik(per_site=True)


def _noalloc(ctx, wl, per_site=True):
    t = wl.tree
    return ctx.loglik_noalloc(wl.masks, t.left, t.right, t.root, t.blen, wl.Q, wl.pi, wl.pi, wl.cat_rates,
                              wl.cat_weights, per_site=per_site)


CASES = [  # (taxa, sites, categories, seed): each call smaller than the one before
    (128, 30_000, 4, 1),
    (24, 3_000, 2, 2),
    (300, 1_500, 1, 3),     # more taxa than the walker stages in LDS
    (6, 70, 8, 4),
]


def test_noalloc_matches_loglik_on_consecutive_calls(oracle):
    wls = [synth.gtr_gamma_workload(n_tips=n, m=m, C=C, alpha=0.5, gap_frac=0.05, seed=seed)
           for n, m, C, seed in CASES]
    with P.Context(0, P.PREC_FP32) as ctx:
        got = []
        for (n, m, C, _), wl in zip(CASES, wls):
            print(f"noalloc taxa={n} sites={m} categories={C}", flush=True)
            got.append(_noalloc(ctx, wl))
        # once more without the per-site output, after the smaller calls
        again = _noalloc(ctx, wls[0], per_site=False)
    for (n, m, C, _), wl, (tot, site) in zip(CASES, wls, got):
        ref_tot, ref_site = _reference(wl, tail_split=False)
        assert tot == ref_tot, (n, m, C, tot, ref_tot)
        np.testing.assert_array_equal(site, ref_site)
        def_tot, def_site = _reference(wl, tail_split=True)
        assert abs(tot - def_tot) <= 1e-3
        np.testing.assert_allclose(site, def_site, rtol=1e-5)
        o_tot, o_site = oracle_problem(oracle, wl).loglik()
        assert abs(tot - o_tot) <= 1e-3, (n, m, C, tot, o_tot)
        np.testing.assert_allclose(site, o_site, rtol=1e-5)
    assert again == got[0][0]


def test_noalloc_rejects_what_does_not_fit():
    wl = synth.gtr_gamma_workload(n_tips=128, m=100_000, C=4, alpha=0.5, gap_frac=0.05, seed=5)
    with P.Context(0, P.PREC_FP32) as ctx:
        with pytest.raises(P.PhyloError, match="scratch"):
            _noalloc(ctx, wl)
        small = synth.gtr_gamma_workload(n_tips=8, m=100, C=4, seed=6)
        with pytest.raises(P.PhyloError):          # 20 states: not the DNA walker
            t = small.tree
            ctx.loglik_noalloc(np.full((8, 100), 0xFFFFF, np.uint32), t.left, t.right, t.root, t.blen,
                               np.eye(20) - 1.0 / 20, np.full(20, 0.05), np.full(20, 0.05), small.cat_rates)
        # the context still works after the rejections
        tot = _noalloc(ctx, small, per_site=False)
        assert tot == _reference(small, tail_split=False)[0]
