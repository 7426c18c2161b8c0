"""Learned starting values for the GTR+Gamma fit.

A GPU forward simulator draws alignments from random GTR+Gamma models on
random 32-taxon Yule trees and reduces each one to 256 site-pattern and
pairwise-distance features in the same kernel (phylo_sim_features); a ReLU
MLP 256 -> 512 -> 512 -> 12 trained on the GPU (phylo_mlp_*) maps features to
the model parameters, and its prediction seeds the Nelder-Mead fit
(phylo_fit_gtr_gamma) in place of treeator's fixed starting values
(treeator.cpp:738-911).

Parameter vector of the fit (include/phylo_host.h): x = [log r_AG, log r_AC,
log r_AT, log r_GC, log r_GT relative to r_CT, logits of pi_A, pi_G, pi_C
relative to pi_T, log alpha, log branch-length scale].  The MLP predicts the
twelve centred quantities [log r - mean log r (6), log pi - mean log pi (4),
log alpha, log mean branch length], standardised with the training set's
mean and deviation.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import MLP, PREC_FP64, Context, gamma_rates, nm_options, sim_features
from .synth import yule_tree

N_OUT = 12
# exchangeability slots in the engine's state order A, G, C, T
PAIRS = ("AG", "AC", "AT", "GC", "GT", "CT")


@dataclass
class ModelDraws:
    """n random GTR+Gamma problems: trees as stacked arrays (tips 0..n_taxa-1)."""
    left: np.ndarray       # [n, n_nodes]
    right: np.ndarray
    roots: np.ndarray      # [n]
    blen: np.ndarray       # [n, n_nodes], root entry 0
    rates: np.ndarray      # [n, 6] exchangeabilities (PAIRS order)
    pi: np.ndarray         # [n, 4]
    alpha: np.ndarray      # [n]
    mean_blen: np.ndarray  # [n] mean non-root branch length
    gap: np.ndarray        # [n] gap fraction


def draw_models(n: int, n_taxa: int, rng: np.random.Generator) -> ModelDraws:
    """Random problems: log exchangeabilities N(0, 0.5) with transitions
    (A-G, C-T) raised by N(1, 0.4); pi ~ Dirichlet(6, 6, 6, 6); alpha
    log-uniform on [0.15, 4]; mean branch length log-uniform on [0.01, 0.25]
    with exponential branch lengths on a Yule tree; gap fraction uniform on
    [0, 0.08]."""
    nn = 2 * n_taxa - 1
    left = np.zeros((n, nn), np.int32)
    right = np.zeros((n, nn), np.int32)
    roots = np.zeros(n, np.int32)
    blen = np.zeros((n, nn))
    mean_blen = np.exp(rng.uniform(np.log(0.01), np.log(0.25), n))
    for k in range(n):
        t = yule_tree(n_taxa, rng, mean_blen=1.0)
        b = t.blen.copy()
        b[t.root] = 0.0
        b *= mean_blen[k] / (b.sum() / (nn - 1))
        left[k], right[k], roots[k], blen[k] = t.left, t.right, t.root, b
    lr = rng.normal(0.0, 0.5, (n, 6))
    lr[:, 0] += rng.normal(1.0, 0.4, n)
    lr[:, 5] += rng.normal(1.0, 0.4, n)
    rates = np.exp(lr)
    pi = rng.dirichlet(np.full(4, 6.0), n)
    alpha = np.exp(rng.uniform(np.log(0.15), np.log(4.0), n))
    gap = rng.uniform(0.0, 0.08, n)
    return ModelDraws(left, right, roots, blen, rates, pi, alpha, mean_blen, gap)


def normalised_Q(rates6, pi) -> np.ndarray:
    Q = np.zeros((4, 4))
    k = 0
    for i in range(4):
        for j in range(i + 1, 4):
            Q[i, j] = rates6[k] * pi[j]
            Q[j, i] = rates6[k] * pi[i]
            k += 1
    np.fill_diagonal(Q, -Q.sum(axis=1))
    return Q / -np.dot(pi, np.diag(Q))


def eigen_system(rates6, pi) -> np.ndarray:
    """[V (16), V^-1 (16), eigenvalues (4)] with Q = V diag(lam) V^-1, from
    the symmetric form D^1/2 Q D^-1/2 (D = diag pi)."""
    Q = normalised_Q(rates6, pi)
    d = np.sqrt(pi)
    B = d[:, None] * Q / d[None, :]
    lam, U = np.linalg.eigh(0.5 * (B + B.T))
    V = U / d[:, None]
    Vi = U.T * d[None, :]
    return np.concatenate([V.ravel(), Vi.ravel(), lam])


def simulate_features(d: ModelDraws, n_sites: int, n_cat: int = 4, seed: int = 1, want_masks: bool = False,
                      chunk: int = 4096):
    """GPU simulation + features for every draw (chunks of `chunk` alignments;
    one gap fraction per chunk: the mean of its draws)."""
    n, nn = d.left.shape
    n_taxa = (nn + 1) // 2
    eig = np.stack([eigen_system(d.rates[k], d.pi[k]) for k in range(n)])
    cats = np.stack([gamma_rates(a, n_cat) for a in d.alpha])
    feats, masks = [], []
    for c0 in range(0, n, chunk):
        s = slice(c0, min(n, c0 + chunk))
        out = sim_features(n_taxa, n_sites, d.left[s], d.right[s], d.roots[s], d.blen[s], eig[s], d.pi[s], cats[s],
                           gap_frac=float(d.gap[s].mean()), seed=seed + c0, want_masks=want_masks)
        if want_masks:
            feats.append(out[0])
            masks.append(out[1])
        else:
            feats.append(out)
    f = np.concatenate(feats)
    return (f, np.concatenate(masks)) if want_masks else f


def targets(d: ModelDraws) -> np.ndarray:
    lr = np.log(d.rates)
    lp = np.log(d.pi)
    return np.concatenate([lr - lr.mean(axis=1, keepdims=True), lp - lp.mean(axis=1, keepdims=True),
                           np.log(d.alpha)[:, None], np.log(d.mean_blen)[:, None]], axis=1)


def fit_start(t: np.ndarray) -> np.ndarray:
    """Twelve centred quantities -> the fit's ten-parameter x (relative
    lengths normalised to mean 1, so the scale is the mean length)."""
    t = np.atleast_2d(t)
    return np.concatenate([t[:, 0:5] - t[:, 5:6], t[:, 6:9] - t[:, 9:10], t[:, 10:12]], axis=1)


def default_start(mean_blen: float = 0.1) -> np.ndarray:
    """The unlearned start: equal exchangeabilities and frequencies, alpha 1."""
    x = np.zeros(10)
    x[9] = np.log(mean_blen)
    return x


class Proposal:
    """Feature / target standardisation plus the GPU MLP.  After fit(),
    x_rmse holds the prediction's root-mean-square error per fit parameter on
    a held-out part of the training draws: the proposal's own scale, used as
    the seeded simplex's edges."""

    def __init__(self, n_hidden: int = 512, seed: int = 1, max_batch: int = 256):
        self.mlp = MLP(256, n_hidden, N_OUT, max_batch=max_batch, seed=seed)
        self.fmu = self.fsd = self.tmu = self.tsd = None
        self.x_rmse = None

    def fit(self, feats, targ, epochs=30, batch=256, lr=1e-3, seed=1, holdout=0.05, schedule=None):
        """Trains on all but the last `holdout` fraction; schedule, a list of
        (epochs, learning rate) stages, replaces (epochs, lr)."""
        n_val = max(1, int(len(feats) * holdout))
        tr_f, tr_t = feats[:-n_val], targ[:-n_val]
        self.fmu = tr_f.mean(axis=0)
        self.fsd = np.maximum(tr_f.std(axis=0), 1e-6)
        self.tmu = tr_t.mean(axis=0)
        self.tsd = np.maximum(tr_t.std(axis=0), 1e-6)
        xs, ts = (tr_f - self.fmu) / self.fsd, (tr_t - self.tmu) / self.tsd
        losses = []
        for k, (ep, rate) in enumerate(schedule or [(epochs, lr)]):
            losses.extend(self.mlp.train(xs, ts, epochs=ep, batch=batch, lr=rate, seed=seed + k))
        losses = np.array(losses)
        err = fit_start(self.predict(feats[-n_val:])) - fit_start(targ[-n_val:])
        self.x_rmse = np.sqrt(np.mean(err ** 2, axis=0))
        return losses

    def predict(self, feats) -> np.ndarray:
        y = self.mlp.forward((np.atleast_2d(feats) - self.fmu) / self.fsd)
        return y * self.tsd + self.tmu


@dataclass
class FitComparison:
    lnl_plain: np.ndarray
    lnl_seeded: np.ndarray
    evals_plain: np.ndarray
    evals_seeded: np.ndarray
    sec_plain: np.ndarray
    sec_seeded: np.ndarray
    wall_plain: float = 0.0      # wall time of all plain fits (one arm after the other)
    wall_seeded: float = 0.0
    workers: int = 1


def compare_fits(prop: Proposal, d: ModelDraws, feats, masks, n_cat=4, opts=None, seeded_opts=None,
                 precision=PREC_FP64, workers: int = 1) -> FitComparison:
    """Fits every held-out alignment twice on its true topology (relative
    branch lengths normalised to mean 1): from the default start and from the
    MLP's prediction.  Returns per-alignment lnL, evaluations and seconds.
    With workers > 1 each arm runs its fits on that many host threads (one
    context, so one HIP stream, per thread): all plain fits, then all seeded
    ones, and wall_plain / wall_seeded time each arm as a whole."""
    if workers > 1:
        return _compare_fits_threaded(prop, d, feats, masks, n_cat, opts, seeded_opts, precision, workers)
    n = len(d.roots)
    out = {k: np.zeros(n) for k in ("lp", "ls", "ep", "es", "tp", "ts")}
    pred = fit_start(prop.predict(feats))
    for k in range(n):
        rel = d.blen[k] / (d.blen[k].sum() / (len(d.blen[k]) - 1))
        with Context(0, precision) as ctx:
            ctx.set_tips(masks[k], 4)
            ctx.set_topology(d.left[k], d.right[k], int(d.roots[k]))
            t0 = time.perf_counter()
            _, lp, ep, _ = ctx.fit_gtr_gamma(rel, n_cat, default_start(), opts)
            tp = time.perf_counter() - t0
            t0 = time.perf_counter()
            _, ls, es, _ = ctx.fit_gtr_gamma(rel, n_cat, pred[k], seeded_opts or opts)
            ts = time.perf_counter() - t0
        out["lp"][k], out["ls"][k], out["ep"][k], out["es"][k], out["tp"][k], out["ts"][k] = lp, ls, ep, es, tp, ts
    return FitComparison(out["lp"], out["ls"], out["ep"], out["es"], out["tp"], out["ts"],
                         float(out["tp"].sum()), float(out["ts"].sum()), 1)


def _compare_fits_threaded(prop, d, feats, masks, n_cat, opts, seeded_opts, precision, workers):
    import threading
    from concurrent.futures import ThreadPoolExecutor
    n = len(d.roots)
    pred = fit_start(prop.predict(feats))
    rel = [d.blen[k] / (d.blen[k].sum() / (len(d.blen[k]) - 1)) for k in range(n)]
    local, made = threading.local(), []

    def fit(k, x0, o):
        # one context per worker thread, reused by its fits (a context is its
        # own HIP stream); the library calls release the GIL, so the threads'
        # fits overlap on the host and their launches on the GPU
        ctx = getattr(local, "ctx", None)
        if ctx is None:
            ctx = local.ctx = Context(0, precision)
            made.append(ctx)
        ctx.set_tips(masks[k], 4)
        ctx.set_topology(d.left[k], d.right[k], int(d.roots[k]))
        t0 = time.perf_counter()
        _, lnl, ev, _ = ctx.fit_gtr_gamma(rel[k], n_cat, x0, o)
        return lnl, ev, time.perf_counter() - t0

    try:
        with ThreadPoolExecutor(workers) as ex:
            t0 = time.perf_counter()
            plain = list(ex.map(lambda k: fit(k, default_start(), opts), range(n)))
            wall_plain = time.perf_counter() - t0
            t0 = time.perf_counter()
            seeded = list(ex.map(lambda k: fit(k, pred[k], seeded_opts or opts), range(n)))
            wall_seeded = time.perf_counter() - t0
    finally:
        for c in made:
            c.close()
    P_, S_ = np.array(plain), np.array(seeded)
    return FitComparison(P_[:, 0], S_[:, 0], P_[:, 1], S_[:, 1], P_[:, 2], S_[:, 2], wall_plain, wall_seeded, workers)
