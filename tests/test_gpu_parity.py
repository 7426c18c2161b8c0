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
