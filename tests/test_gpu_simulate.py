"""phylo_simulate (the GPU forward simulator behind simpleML::simulate_chars)
against exact joint state distributions from the oracle's P(t)."""
import numpy as np
import pytest
from scipy.stats import chi2

import phylommand_amd as P
from phylommand_amd import synth

pytestmark = pytest.mark.gpu


def _ctx(n_tips, S, Q, pi, left, right, root, blen, cats, builder):
    ctx = P.Context(0, P.PREC_FP64)
    ctx.set_tips(np.full((n_tips, 1), (1 << S) - 1 if S < 32 else 0xFFFFFFFF, np.uint32), S)
    ctx.set_topology(left, right, root)
    ctx.set_edge_lengths(blen)
    ctx.set_model(Q, pi, pi, builder)
    ctx.set_categories(cats)
    return ctx


@pytest.mark.parametrize("S,C,builder", [(4, 1, P.P_EIGEN), (4, 4, P.P_EIGEN), (20, 4, P.P_EIGEN), (4, 2, P.P_REFCOMPAT)])
def test_two_tip_joint_distribution(oracle, S, C, builder):
    """Root with two tips: the joint (tip a, tip b) frequency is
    sum_c w_c diag(pi) P(r_c (t_a + t_b)) for a reversible model."""
    print(f"simulate two tips S={S} C={C} builder={builder}", flush=True)
    rng = np.random.default_rng(S * 10 + C)
    pi = rng.dirichlet(np.full(S, 4.0))
    r = rng.gamma(1.0, 1.0, S * (S - 1) // 2)
    Q = synth.normalise_Q(synth.gtr_Q(r, pi), pi)
    cats = synth.gamma_rates(0.7, C) if C > 1 else np.ones(1)
    left = np.array([-1, -1, 0], np.int32)
    right = np.array([-1, -1, 1], np.int32)
    blen = np.array([0.2, 0.45, 0.0])
    ctx = _ctx(2, S, Q, pi, left, right, 2, blen, cats, builder)
    m = 1_000_000
    st = ctx.simulate(m, pi, seed=5)
    assert st.shape == (3, m) and st.max() < S
    counts = np.bincount(st[0].astype(np.int64) * S + st[1], minlength=S * S).reshape(S, S)
    joint = sum(np.diag(pi) @ oracle.pmatrix(Q, c * 0.65, oracle.EIG, pi) for c in cats) / C
    exp = joint * m
    keep = exp >= 5
    stat = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    assert stat < chi2.ppf(1 - 1e-6, keep.sum() - 1), stat
    # root states follow pi
    root = np.bincount(st[2], minlength=S) / m
    np.testing.assert_allclose(root, pi, atol=5 * np.sqrt(pi.max() / m) + 1e-4)
    ctx.close()


def test_simulate_is_seeded_and_rejects_bad_input():
    pi = np.full(4, 0.25)
    Q = synth.normalise_Q(synth.gtr_Q(np.ones(6), pi), pi)
    left = np.array([-1, -1, -1, 0, 3], np.int32)
    right = np.array([-1, -1, -1, 1, 2], np.int32)
    ctx = _ctx(3, 4, Q, pi, left, right, 4, np.array([0.1, 0.2, 0.3, 0.05, 0.0]), np.ones(1), P.P_EIGEN)
    a = ctx.simulate(5000, pi, seed=3)
    b = ctx.simulate(5000, pi, seed=3)
    c = ctx.simulate(5000, pi, seed=4)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    with pytest.raises(P.PhyloError):
        ctx.simulate(0, pi)
    with pytest.raises(P.PhyloError):
        ctx.simulate(10, np.array([np.nan, 0.5, 0.25, 0.25]))
    ctx.close()
