"""The batched-scoring and learned-proposal workloads at their full stated
sizes, checked against the FP64 oracle and the plain optimiser.

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
