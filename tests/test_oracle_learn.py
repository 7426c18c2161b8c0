"""CPU checks of the learned-proposal oracle (features, MLP backpropagation,
Adam) and of the host-side helpers of the model fit (discrete-gamma rates,
parameter encodings), each against a closed form, an independent brute force
or a library routine."""
import numpy as np
import pytest
from scipy.special import gammainc, gammaincinv

import phylommand_amd as P
from phylommand_amd import learn as L

from oracle import learn_oracle as LO

GAP = LO.GAP


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n_in, n_hid, n_out, B = 5, 7, 3, 4
    n_par = n_in * n_hid + n_hid + n_hid * n_hid + n_hid + n_hid * n_out + n_out
    w = rng.normal(0, 0.7, n_par)
    x = rng.normal(0, 1, (B, n_in))
    t = rng.normal(0, 1, (B, n_out))
    loss, g = LO.mlp_grad(w, x, t, n_in, n_hid, n_out)
    assert loss == pytest.approx(LO.mlp_loss(w, x, t, n_in, n_hid, n_out), rel=1e-14)
    h = 1e-6
    fd = np.zeros(n_par)
    for i in range(n_par):
        e = np.zeros(n_par)
        e[i] = h
        fd[i] = (LO.mlp_loss(w + e, x, t, n_in, n_hid, n_out) - LO.mlp_loss(w - e, x, t, n_in, n_hid, n_out)) / (2 * h)
    np.testing.assert_allclose(g, fd, atol=1e-7)


def test_mlp_oracle_layout_planted_transpose_fails():
    """A weight matrix read transposed changes the output: the layout is pinned."""
    rng = np.random.default_rng(4)
    n_in, n_hid, n_out = 3, 3, 2
    W1 = rng.normal(size=(n_in, n_hid))
    flat = np.concatenate([W1.ravel(), np.zeros(n_hid), np.eye(n_hid).ravel(), np.zeros(n_hid),
                           np.ones((n_hid, n_out)).ravel(), np.zeros(n_out)])
    x = np.abs(rng.normal(size=(1, n_in)))
    y, (h1, _) = LO.mlp_forward(flat, x, n_in, n_hid, n_out)
    np.testing.assert_allclose(h1, np.maximum(x @ W1, 0))
    assert not np.allclose(h1, np.maximum(x @ W1.T, 0))
    np.testing.assert_allclose(y, np.maximum(x @ W1, 0).sum(axis=1, keepdims=True) * np.ones((1, n_out)))


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([0.3, -2.0, 1e-3])
    w, m, v = LO.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), step=1, lr=0.01)
    np.testing.assert_allclose(w, -0.01 * np.sign(g), rtol=1e-4)


def brute_features(st):
    """Independent restatement by explicit loops over pairs and sites."""
    n_taxa, m = st.shape
    pairs = [(i, j) for i in range(n_taxa) for j in range(i + 1, n_taxa)]
    pd, ts, tv = [], [], []
    for i, j in pairs:
        cols = [s for s in range(m) if st[i, s] != GAP and st[j, s] != GAP]
        if not cols:
            continue
        diff = sum(st[i, s] != st[j, s] for s in cols)
        trans = sum({st[i, s], st[j, s]} in ({0, 1}, {2, 3}) for s in cols)
        pd.append(diff / len(cols))
        ts.append(trans / len(cols))
        tv.append((diff - trans) / len(cols))
    distinct = [len({st[t, s] for t in range(n_taxa) if st[t, s] != GAP}) for s in range(m)]
    return pairs, pd, ts, tv, distinct


def test_features_against_brute_force():
    rng = np.random.default_rng(7)
    st = rng.integers(0, 4, (6, 37))
    st[rng.random(st.shape) < 0.1] = GAP
    st[:, 3] = 2                         # an invariant site
    st[:, 5] = GAP                       # an all-gap site
    st[0, :] = GAP                       # a taxon with no data: its pairs drop out
    f = LO.features_from_states(st)
    pairs, pd, ts, tv, distinct = brute_features(st)
    n_pairs = len(pairs)
    ok = st != GAP
    np.testing.assert_allclose(f[0:4], [np.sum(st == j) / ok.sum() for j in range(4)])
    assert f[184] == pytest.approx(sum(pd) / n_pairs)
    assert f[186] == pytest.approx(sum(ts) / n_pairs)
    assert f[187] == pytest.approx(sum(tv) / n_pairs)
    assert f[188] == pytest.approx((~ok).sum() / st.size)
    for d in range(1, 5):
        assert f[115 + d] == pytest.approx(sum(x == d for x in distinct) / st.shape[1])
    assert f[191] == pytest.approx(sum(x == 1 for x in distinct) / st.shape[1])
    # histograms hold mass equal to the pairs / sites present
    assert f[20:52].sum() == pytest.approx(len(pd) / n_pairs)
    assert f[120:152].sum() == pytest.approx(1.0)
    assert f[152:184].sum() == pytest.approx(1.0)
    assert f[4:20].sum() == pytest.approx(len(pd) / n_pairs)
    assert np.all(f[192:] == 0)


def test_feature_bins_known_answers():
    # two taxa, 40 sites: 10 transitions, 6 transversions, 24 identical
    a = np.zeros(40, int)
    b = np.zeros(40, int)
    b[:10] = 1          # A -> G transition
    b[10:16] = 2        # A -> C transversion
    f = LO.features_from_states(np.stack([a, b]))
    assert f[184] == pytest.approx(16 / 40)
    assert f[20 + 16] == 1.0              # p-distance 0.4 = 16/40 -> bin 40*16//40 = 16
    assert f[52 + 20] == 1.0              # ts 0.25 -> bin 80*10//40 = 20
    assert f[84 + 12] == 1.0              # tv 0.15 -> bin 80*6//40 = 12
    # 2 of 2 taxa differ at 16 sites: minor fraction 1/2 -> bin 31 (clamped 32),
    # Gini-Simpson 1/2 -> bin 128*2//12 = 21
    assert f[120 + 31] == pytest.approx(16 / 40)
    assert f[152 + 21] == pytest.approx(16 / 40)
    assert f[189] == pytest.approx(-0.75 * np.log(1 - 0.4 / 0.75))


def test_states_from_masks_roundtrip():
    st = np.array([[0, 1, 2, 3, GAP]])
    masks = np.where(st == GAP, 15, 1 << np.maximum(st, 0))
    np.testing.assert_array_equal(LO.states_from_masks(masks), st)


@pytest.mark.parametrize("alpha", [0.1, 0.5, 1.0, 2.7, 20.0])
@pytest.mark.parametrize("C", [2, 4, 8])
def test_host_gamma_rates_match_scipy(alpha, C):
    got = P.gamma_rates(alpha, C)
    q = gammaincinv(alpha, np.arange(1, C) / C)
    I = np.concatenate([[0.0], gammainc(alpha + 1, q), [1.0]])
    np.testing.assert_allclose(got, np.diff(I) * C, rtol=1e-9, atol=1e-12)
    assert got.mean() == pytest.approx(1.0, rel=1e-12)
    assert np.all(np.diff(got) > 0)


def test_host_gamma_rates_reject_bad_arguments():
    with pytest.raises(P.PhyloError):
        P.gamma_rates(-1.0, 4)


def test_eigen_system_reconstructs_Q():
    rng = np.random.default_rng(2)
    r = np.exp(rng.normal(size=6))
    pi = rng.dirichlet(np.ones(4) * 3)
    e = L.eigen_system(r, pi)
    V, Vi, lam = e[:16].reshape(4, 4), e[16:32].reshape(4, 4), e[32:]
    Q = L.normalised_Q(r, pi)
    np.testing.assert_allclose(V @ np.diag(lam) @ Vi, Q, atol=1e-12)
    np.testing.assert_allclose(V @ Vi, np.eye(4), atol=1e-12)
    assert -np.dot(pi, np.diag(Q)) == pytest.approx(1.0)
    np.testing.assert_allclose(pi @ Q, 0, atol=1e-12)


def test_fit_start_encodes_the_true_model():
    """targets -> fit_start gives the x whose Q (rates relative to C-T,
    frequency logits relative to T) is the drawn model's normalised Q."""
    rng = np.random.default_rng(8)
    d = L.draw_models(5, 8, rng)
    x = L.fit_start(L.targets(d))
    for k in range(5):
        r = np.exp(np.append(x[k, :5], 0.0))
        z = np.exp(np.append(x[k, 5:8], 0.0))
        pi = z / z.sum()
        np.testing.assert_allclose(pi, d.pi[k], rtol=1e-12)
        np.testing.assert_allclose(L.normalised_Q(r, pi), L.normalised_Q(d.rates[k], d.pi[k]), rtol=1e-10)
        assert np.exp(x[k, 8]) == pytest.approx(d.alpha[k])
        assert np.exp(x[k, 9]) == pytest.approx(d.blen[k].sum() / (len(d.blen[k]) - 1))
