"""CPU restatement of the learned-proposal kernels: the alignment features the
GPU simulator computes from its tip states, and the FP64 forward / backward
pass of the ReLU MLP.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product path.  The reference has no learned proposal (treeator.cpp:738-911
starts every fit from fixed values); its simulator, simpleML::simulate_chars
(simpleML.cpp:233-292), is the model of the GPU simulator, and the oracle
checks that simulator by the exact pattern probabilities of the pruning
likelihood (tests/test_oracle_learn.py, tests/test_gpu_learn.py).
"""
from __future__ import annotations

import numpy as np

N_FEATURES = 256
GAP = -1


def features_from_states(states: np.ndarray) -> np.ndarray:
    """states [n_taxa, n_sites] in 0..3 (A, G, C, T) or GAP -> the 256
    features of one alignment, in the GPU kernel's layout:
      0-3    base composition over non-gap tip states
      4-19   symmetrised pairwise joint state frequencies, averaged over pairs
      20-51  histogram of pairwise p-distance, bins of 1/40
      52-83  histogram of pairwise transition distance, bins of 1/80
      84-115 histogram of pairwise transversion distance, bins of 1/80
      116-119 fraction of sites with 1..4 distinct states
      120-151 histogram of the minor-allele fraction, bins of 1/64
      152-183 histogram of Gini-Simpson diversity, bins of 3/128
      184 mean p-distance, 185 its standard deviation over pairs,
      186 mean transition distance, 187 mean transversion distance,
      188 gap fraction, 189 mean Jukes-Cantor distance (capped at 3),
      190 transition / transversion ratio (capped at 20),
      191 fraction of invariant sites; 192-255 zero.
    Pairs with no site where both are non-gap are left out of the sums but
    not of the pair count the averages divide by."""
    st = np.asarray(states)
    n_taxa, n_sites = st.shape
    f = np.zeros(N_FEATURES)
    ok = st != GAP
    base = np.array([(st == j).sum() for j in range(4)], dtype=np.int64)
    nb = base.sum()
    if nb:
        f[0:4] = base / nb
    n_pairs = n_taxa * (n_taxa - 1) // 2
    pair_joint = np.zeros(16)
    h_pd = np.zeros(32)
    h_ts = np.zeros(32)
    h_tv = np.zeros(32)
    s_pd = s_pd2 = s_ts = s_tv = s_jc = 0.0
    for i in range(n_taxa):
        for j in range(i + 1, n_taxa):
            both = ok[i] & ok[j]
            n = int(both.sum())
            if n == 0:
                continue
            J = np.zeros((4, 4), dtype=np.int64)
            np.add.at(J, (st[i][both], st[j][both]), 1)
            same = int(np.trace(J))
            ts = int(J[0, 1] + J[1, 0] + J[2, 3] + J[3, 2])
            tv = n - same - ts
            pair_joint += (0.5 * (J + J.T) / n).ravel() / n_pairs
            h_pd[min(31, 40 * (n - same) // n)] += 1.0 / n_pairs
            h_ts[min(31, 80 * ts // n)] += 1.0 / n_pairs
            h_tv[min(31, 80 * tv // n)] += 1.0 / n_pairs
            pd = (n - same) / n
            s_pd += pd / n_pairs
            s_pd2 += pd * pd / n_pairs
            s_ts += ts / n / n_pairs
            s_tv += tv / n / n_pairs
            s_jc += min(3.0, -0.75 * np.log(max(1e-6, 1.0 - pd / 0.75))) / n_pairs
    f[4:20] = pair_joint
    f[20:52] = h_pd
    f[52:84] = h_ts
    f[84:116] = h_tv
    cnt = np.stack([(st == j).sum(axis=0) for j in range(4)])       # [4, n_sites]
    n_ok = cnt.sum(axis=0)
    distinct = (cnt > 0).sum(axis=0)
    for d in range(1, 5):
        f[115 + d] = (distinct == d).sum() / n_sites
    mx = cnt.max(axis=0)
    sq = (cnt * cnt).sum(axis=0)
    safe = np.maximum(n_ok, 1)
    minor_bin = np.where(n_ok > 0, np.minimum(31, 64 * (n_ok - mx) // safe), 0)
    div_bin = np.where(n_ok > 0, np.minimum(31, 128 * (n_ok * n_ok - sq) // (3 * safe * safe)), 0)
    f[120:152] = np.bincount(minor_bin, minlength=32) / n_sites
    f[152:184] = np.bincount(div_bin, minlength=32) / n_sites
    f[184] = s_pd
    f[185] = np.sqrt(max(0.0, s_pd2 - s_pd * s_pd))
    f[186] = s_ts
    f[187] = s_tv
    f[188] = (~ok).sum() / (n_sites * n_taxa)
    f[189] = s_jc
    f[190] = min(20.0, s_ts / max(1e-6, s_tv))
    f[191] = (distinct == 1).sum() / n_sites
    return f


def states_from_masks(masks: np.ndarray) -> np.ndarray:
    """Tip bit masks (one bit per state, gap = all four bits) -> states."""
    m = np.asarray(masks, dtype=np.int64)
    out = np.full(m.shape, GAP, dtype=np.int64)
    for j in range(4):
        out[m == (1 << j)] = j
    return out


def mlp_unpack(w, n_in, n_hid, n_out):
    """Flat parameters [W1 | b1 | W2 | b2 | W3 | b3], weights row-major with
    the layer's inputs as rows."""
    w = np.asarray(w, dtype=np.float64)
    sizes = [n_in * n_hid, n_hid, n_hid * n_hid, n_hid, n_hid * n_out, n_out]
    parts = np.split(w, np.cumsum(sizes)[:-1])
    return (parts[0].reshape(n_in, n_hid), parts[1], parts[2].reshape(n_hid, n_hid), parts[3],
            parts[4].reshape(n_hid, n_out), parts[5])


def mlp_forward(w, x, n_in, n_hid, n_out):
    W1, b1, W2, b2, W3, b3 = mlp_unpack(w, n_in, n_hid, n_out)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ W3 + b3, (h1, h2)


def mlp_loss(w, x, t, n_in, n_hid, n_out):
    """Mean squared error over all batch x output entries."""
    y, _ = mlp_forward(w, x, n_in, n_hid, n_out)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def mlp_grad(w, x, t, n_in, n_hid, n_out):
    """(loss, flat gradient) by backpropagation in FP64."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(w, n_in, n_hid, n_out)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    y, (h1, h2) = mlp_forward(w, x, n_in, n_hid, n_out)
    B = x.shape[0]
    dy = 2.0 * (y - t) / (B * n_out)
    gW3 = h2.T @ dy
    gb3 = dy.sum(axis=0)
    dh2 = (dy @ W3.T) * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(np.mean((y - t) ** 2)), g


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) with bias correction; returns new
    (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
