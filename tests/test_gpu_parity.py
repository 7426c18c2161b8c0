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
