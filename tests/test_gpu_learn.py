"""The learned-proposal kernels on the GPU against the oracle: simulator +
feature kernel (features recomputed by the oracle from the simulated tip
states; site-pattern frequencies against the exact pattern probabilities of
the pruning likelihood), MLP forward / backward / Adam against FP64
backpropagation, and the GTR+Gamma fit."""
import numpy as np
import pytest
from scipy.stats import chi2

import phylommand_amd as P
from phylommand_amd import learn as L

from oracle import learn_oracle as LO

pytestmark = pytest.mark.gpu


def test_sim_features_match_oracle_features():
    print("sim features: 6 alignments x 12 taxa x 700 sites, gaps 0.05", flush=True)
    d = L.draw_models(6, 12, np.random.default_rng(11))
    d.gap[:] = 0.05
    f, masks = L.simulate_features(d, 700, seed=5, want_masks=True)
    assert masks.shape == (6, 12, 700)
    gaps = np.mean(masks == 15)
    assert 0.03 < gaps < 0.07
    for k in range(6):
        ref = LO.features_from_states(LO.states_from_masks(masks[k]))
        np.testing.assert_allclose(f[k], ref, atol=2e-6, err_msg=f"alignment {k}")


def test_sim_features_are_seeded():
    d = L.draw_models(3, 8, np.random.default_rng(1))
    a, ma = L.simulate_features(d, 300, seed=9, want_masks=True)
    b, mb = L.simulate_features(d, 300, seed=9, want_masks=True)
    c, mc = L.simulate_features(d, 300, seed=10, want_masks=True)
    np.testing.assert_array_equal(ma, mb)
    assert not np.array_equal(ma, mc)


def test_simulator_pattern_frequencies(oracle):
    """4 taxa, GTR+Gamma4: counts of the 256 site patterns against the
    probabilities exp(site lnL) of the pruning likelihood (chi-square)."""
    print("simulator pattern frequencies: 4 taxa x 400000 sites", flush=True)
    rng = np.random.default_rng(21)
    d = L.draw_models(1, 4, rng)
    d.rates[0] = [2.5, 0.6, 1.1, 0.8, 0.4, 3.0]
    d.pi[0] = [0.35, 0.15, 0.2, 0.3]
    d.alpha[0] = 0.6
    d.blen[0] = np.where(d.blen[0] > 0, 0.15 + 0.1 * np.arange(7), 0.0)
    d.blen[0][d.roots[0]] = 0.0
    d.gap[:] = 0.0
    m = 400_000
    _, masks = L.simulate_features(d, m, seed=3, want_masks=True)
    st = LO.states_from_masks(masks[0])
    code = st[0] * 64 + st[1] * 16 + st[2] * 4 + st[3]
    counts = np.bincount(code, minlength=256)
    # every pattern as one site of an oracle problem
    pats = np.array([[(p >> (2 * (3 - t))) & 3 for p in range(256)] for t in range(4)])
    nn = 7
    mb = np.zeros((nn, 256), np.uint32)
    mb[:4] = 1 << pats
    Q = L.normalised_Q(d.rates[0], d.pi[0])
    cats = P.gamma_rates(d.alpha[0], 4)
    prob = oracle.Problem(d.left[0], d.right[0], d.roots[0], d.blen[0], mb, 4, Q, d.pi[0], pi=d.pi[0],
                          cat_rates=cats, builder=oracle.EIG)
    _, site = prob.loglik()
    p = np.exp(site)
    assert p.sum() == pytest.approx(1.0, abs=1e-10)
    exp = p * m
    keep = exp >= 5
    stat = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    df = keep.sum() - 1
    assert stat < chi2.ppf(1 - 1e-6, df), (stat, df)
    # and a planted mistake in the oracle side (a swapped pair of branch
    # lengths) is detected by the same statistic
    b2 = d.blen[0].copy()
    kids = [v for v in range(nn) if v != d.roots[0] and v < 4][:2]
    b2[kids[0]], b2[kids[1]] = b2[kids[1]] * 3, b2[kids[0]]
    _, site2 = oracle.Problem(d.left[0], d.right[0], d.roots[0], b2, mb, 4, Q, d.pi[0], pi=d.pi[0],
                              cat_rates=cats, builder=oracle.EIG).loglik()
    exp2 = np.exp(site2) * m
    k2 = exp2 >= 5
    assert np.sum((counts[k2] - exp2[k2]) ** 2 / exp2[k2]) > chi2.ppf(1 - 1e-6, k2.sum() - 1)


def _mlp_case(n_in=256, n_hid=512, n_out=12, B=64, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    t = rng.normal(0, 1, (B, n_out)).astype(np.float32)
    return x, t


def test_mlp_forward_matches_oracle():
    print("mlp forward 256-512-512-12, batch 300 (two device batches)", flush=True)
    net = P.MLP(256, 512, 12, max_batch=256, seed=4)
    w = net.params()
    x, _ = _mlp_case(B=300)
    y = net.forward(x)
    ref, _ = LO.mlp_forward(w, x, 256, 512, 12)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())
    net.close()


def test_mlp_gradient_matches_oracle():
    print("mlp gradient 256-512-512-12, batch 64", flush=True)
    net = P.MLP(256, 512, 12, max_batch=256, seed=5)
    w = net.params()
    x, t = _mlp_case(B=64, seed=1)
    loss, g = net.grad(x, t)
    ref_loss, ref_g = LO.mlp_grad(w, x, t, 256, 512, 12)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    scale = np.abs(ref_g).max()
    np.testing.assert_allclose(g, ref_g, atol=1e-5 * scale)
    assert np.linalg.norm(g - ref_g) <= 1e-5 * np.linalg.norm(ref_g)
    net.close()


def test_mlp_adam_step_matches_oracle():
    net = P.MLP(32, 48, 5, max_batch=64, seed=6)
    w = net.params().astype(np.float64)
    x, t = _mlp_case(32, 48, 5, B=64, seed=2)
    net.train(x, t, epochs=1, batch=64, lr=1e-3)        # one full-batch step
    _, g = LO.mlp_grad(w, x, t, 32, 48, 5)
    w1, _, _ = LO.adam_step(w, g, np.zeros_like(w), np.zeros_like(w), step=1, lr=1e-3)
    got = net.params()
    # elements whose gradient is far from Adam's epsilon move by exactly lr
    big = np.abs(g) > 1e-6
    np.testing.assert_allclose(got[big], w1[big], atol=2e-6)
    np.testing.assert_allclose(got, w1, atol=1.1e-3)
    net.close()


def test_training_learns_the_model_parameters():
    print("train the proposal on 4096 simulated alignments (32 taxa x 1000 sites)", flush=True)
    rng = np.random.default_rng(31)
    d = L.draw_models(4096 + 256, 32, rng)
    f = L.simulate_features(d, 1000, seed=2)
    t = L.targets(d)
    prop = L.Proposal(seed=3)
    losses = prop.fit(f[:4096], t[:4096], epochs=25, lr=1e-3)
    assert losses[-1] < 0.5 * losses[0]
    pred = prop.predict(f[4096:])
    r2 = 1 - ((pred - t[4096:]) ** 2).mean(axis=0) / t[4096:].var(axis=0)
    assert np.all(r2 > 0.5), r2
    assert r2[6:10].min() > 0.9          # base frequencies
    assert r2[11] > 0.9                  # tree length


def _fit_problem(seed=41, n_taxa=16, m=1500):
    d = L.draw_models(1, n_taxa, np.random.default_rng(seed))
    d.gap[:] = 0.02
    _, masks = L.simulate_features(d, m, seed=seed, want_masks=True)
    rel = d.blen[0] / (d.blen[0].sum() / (len(d.blen[0]) - 1))
    ctx = P.Context(0, P.PREC_FP64)
    ctx.set_tips(masks[0], 4)
    ctx.set_topology(d.left[0], d.right[0], int(d.roots[0]))
    return d, rel, ctx, masks[0]


def test_gtr_gamma_lnl_matches_oracle(oracle):
    d, rel, ctx, tip_masks = _fit_problem()
    x = L.fit_start(L.targets(d))[0]
    got = ctx.gtr_gamma_lnl(rel, 4, x)
    Q = L.normalised_Q(d.rates[0], d.pi[0])
    eff = rel * np.exp(x[9])
    masks = np.zeros((len(rel), ctx.n_sites), np.uint32)
    masks[: tip_masks.shape[0]] = tip_masks
    want, _ = oracle.Problem(d.left[0], d.right[0], d.roots[0], eff, masks, 4, Q, d.pi[0], pi=d.pi[0],
                             cat_rates=P.gamma_rates(d.alpha[0], 4), builder=oracle.EIG).loglik()
    assert got == pytest.approx(want, rel=1e-9)
    bad = x.copy()
    bad[0] = np.nan
    assert ctx.gtr_gamma_lnl(rel, 4, bad) == -np.inf
    ctx.close()


def test_fit_reaches_at_least_the_true_parameters():
    d, rel, ctx, _ = _fit_problem()
    x_true = L.fit_start(L.targets(d))[0]
    lnl_true = ctx.gtr_gamma_lnl(rel, 4, x_true)
    opts = P.nm_options(max_evals=6000)
    x, lnl, evals, secs = ctx.fit_gtr_gamma(rel, 4, L.default_start(), opts)
    assert lnl >= lnl_true - 1e-6
    assert ctx.gtr_gamma_lnl(rel, 4, x) == lnl
    assert 0 < evals <= 6000 and secs > 0
    # a second start converges to the same optimum
    _, lnl2, _, _ = ctx.fit_gtr_gamma(rel, 4, x_true, opts)
    assert abs(lnl2 - lnl) < 0.01
    ctx.close()


def test_learned_seed_reduces_fit_cost():
    """What the trained network gains over the method without it: the same
    Nelder-Mead fit from the MLP's prediction (simplex scaled by its held-out
    RMSE) reaches the plain fit's lnL (within 0.01) with fewer likelihood
    evaluations."""
    print("learned seed vs plain: train on 16384, fit 8 held-out alignments", flush=True)
    rng = np.random.default_rng(5)
    d = L.draw_models(16384, 32, rng)
    f = L.simulate_features(d, 2000, seed=11)
    prop = L.Proposal(seed=2)
    prop.fit(f, L.targets(d), schedule=[(15, 1e-3), (10, 3e-4)])
    assert prop.x_rmse.shape == (10,) and np.all(prop.x_rmse > 0) and np.all(prop.x_rmse < 0.5)
    held = L.draw_models(8, 32, np.random.default_rng(99))
    hf, hm = L.simulate_features(held, 2000, seed=77, want_masks=True)
    common = dict(ftol_abs=1e-3, xtol_rel=1e-4, max_evals=8000, restart_step=0.01)
    c = L.compare_fits(prop, held, hf, hm, opts=P.nm_options(**common),
                       seeded_opts=P.nm_options(**common, steps=2 * prop.x_rmse))
    assert np.all(np.abs(c.lnl_plain - c.lnl_seeded) <= 0.01), c.lnl_plain - c.lnl_seeded
    assert c.evals_seeded.sum() * 1.15 <= c.evals_plain.sum(), (c.evals_plain, c.evals_seeded)
