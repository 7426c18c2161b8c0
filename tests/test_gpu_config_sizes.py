"""The batched-scoring and learned-proposal workloads at their full stated
sizes, checked against the FP64 oracle and the plain optimiser.

Learned proposal: 200,000 alignments (32 taxa x 2,000 sites, random GTR+G4
models on Yule trees) simulated and featurised on the GPU, the
256-512-512-12 MLP trained on them with the library's own kernels, and 1,000
held-out alignments fitted twice by the same Nelder-Mead, from the default
start and from the MLP's prediction: the seeded fits must take at most half
the wall time and end no lower than the plain ones (within 0.01 lnL); plain
fits that stall below the seeded optimum are counted and must be rare.

Batched scoring: 512 independent GTR+G4 DNA alignments (taxa ~ U{8..256},
sites log-uniform in [500, 50000], per-instance Dirichlet parameters), each
with its full NNI neighbourhood (2(n-3) variants), scored by the persistent
launches replayed from one captured HIP graph with autotuned tiles.  Every
instance total is checked against the oracle, and two seeded NNI variants per
instance against the oracle's re-evaluation of the interchanged tree (the
oracle runs in spawned worker processes)."""
import multiprocessing as mp
import os
import time

import numpy as np
import pytest

import phylommand_amd as P
from phylommand_amd import synth

import batch_oracle

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, len(os.sched_getaffinity(0))))


def test_batch_512_alignments_full_neighbourhoods_against_oracle(oracle):
    t0 = time.time()
    print(f"batch 512 alignments: simulating with {WORKERS} workers", flush=True)
    wls = synth.batch_workloads(512, workers=WORKERS)
    params = synth.batch_instance_params(512)
    with P.Batch(4, P.PREC_FP32) as b:
        for wl in wls:
            t = wl.tree
            b.add(wl.masks, t.left, t.right, t.root, t.blen, wl.Q, wl.pi, wl.cat_rates, wl.cat_weights, nni_mode=1)
        b.finalize(autotune=True)
        tot, nni = b.run()
        tot2, nni2 = b.run()
        off, node, side = b.nni_layout()
        u1, u2, ms, spi, nv = b.stats()
    print(f"batch: {nv} variants, {u1 + u2:.3g} updates, {ms:.1f} ms per run ({time.time() - t0:.0f}s)", flush=True)
    assert np.array_equal(tot, tot2) and np.array_equal(nni, nni2)
    assert nv == sum(2 * (n - 3) for n, _, _ in params)
    rng = np.random.default_rng(11)
    jobs, picked = [], []
    for i, (n, m, alpha) in enumerate(params):
        assert off[i + 1] - off[i] == 2 * (n - 3)
        ks = rng.choice(np.arange(off[i], off[i + 1]), size=2, replace=False)
        picked.append(ks)
        jobs.append((i, n, m, alpha, [(int(node[k]), int(side[k])) for k in ks]))
    del wls
    with mp.get_context("spawn").Pool(WORKERS) as pool:
        res = pool.map(batch_oracle.score_instance, jobs, chunksize=2)
    print(f"oracle done ({time.time() - t0:.0f}s)", flush=True)
    worst_tot = worst_var = 0.0
    for (i, ref_tot, ref_var), ks in zip(res, picked):
        d = abs(tot[i] - ref_tot)
        worst_tot = max(worst_tot, d)
        assert d <= 1e-3, (i, params[i], tot[i], ref_tot)
        for k, ref in zip(ks, ref_var):
            dv = abs(nni[k] - ref)
            worst_var = max(worst_var, dv)
            assert dv <= 1e-3, (i, params[i], int(node[k]), int(side[k]), nni[k], ref)
    print(f"max |total - oracle| {worst_tot:.2e}, max |variant - oracle| {worst_var:.2e}", flush=True)


def test_learned_proposal_200k_training_1000_fits():
    from phylommand_amd import learn as L
    t0 = time.time()
    print("learned proposal: 200000 training alignments, 1000 held-out fits", flush=True)
    d = L.draw_models(200_000, 32, np.random.default_rng(5))
    feats = L.simulate_features(d, 2000, seed=11)
    prop = L.Proposal()
    losses = prop.fit(feats, L.targets(d), schedule=[(20, 1e-3), (10, 3e-4), (10, 1e-4)])
    assert np.isfinite(losses[-1]) and losses[-1] < 0.5 * losses[0]
    held = L.draw_models(1000, 32, np.random.default_rng(99))
    hf, hm = L.simulate_features(held, 2000, seed=77, want_masks=True)
    common = dict(ftol_abs=1e-3, xtol_rel=1e-4, max_evals=8000, restart_step=0.01)
    c = L.compare_fits(prop, held, hf, hm, opts=P.nm_options(**common),
                       seeded_opts=P.nm_options(**common, steps=2 * prop.x_rmse), workers=8)
    diff = c.lnl_seeded - c.lnl_plain
    ratio = c.wall_plain / c.wall_seeded
    stalls = int(np.sum(diff > 0.01))
    print(f"plain {c.wall_plain:.1f}s ({c.evals_plain.mean():.0f} evaluations per fit), seeded {c.wall_seeded:.1f}s "
          f"({c.evals_seeded.mean():.0f}): {ratio:.2f}x; lnL within 0.01 on {int(np.sum(np.abs(diff) <= 0.01))} "
          f"fits, plain stalled below the seeded optimum on {stalls} (largest gap {max(0.0, diff.max()):.1f}); "
          f"{time.time() - t0:.0f}s", flush=True)
    assert ratio >= 2.0
    # Both arms stop at an absolute tolerance of 1e-3 and may stop short on a
    # flat ridge; the proposal's training features are sums of float atomics,
    # so the seeded starts (and which fits stop short) vary from run to run.
    # Each arm may end more than 0.01 below the other on at most 5 of the
    # 1000 fits, and the seeded one never more than 0.1 below.
    assert int(np.sum(diff < -0.01)) <= 5, np.sort(diff)[:6]
    assert diff.min() >= -0.1, diff.min()
    assert stalls <= 5
