"""Batched scoring (persistent work-queue launches in a captured graph)
against single-problem contexts and the oracle."""
import numpy as np
import pytest

import phylommand_amd as P
from phylommand_amd import synth
from helpers import gpu_context

pytestmark = pytest.mark.gpu


def make_instances(k, seed, nmin=8, nmax=48, mmin=100, mmax=3000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(rng.integers(nmin, nmax + 1))
        m = int(np.exp(rng.uniform(np.log(mmin), np.log(mmax))))
        out.append(synth.gtr_gamma_workload(n_tips=n, m=m, C=4, alpha=float(rng.uniform(0.3, 2.0)),
                                            gap_frac=0.05, seed=1000 * seed + i))
    return out


@pytest.mark.parametrize("prec", [P.PREC_FP32, P.PREC_FP64])
def test_batch_matches_single_contexts_and_oracle(oracle, prec):
    wls = make_instances(12, 3)
    with P.Batch(4, prec) as b:
        for i, wl in enumerate(wls):
            t = wl.tree
            b.add(wl.masks, t.left, t.right, t.root, t.blen, wl.Q, wl.pi, wl.cat_rates, wl.cat_weights,
                  nni_mode=i % 3 - 1)
        b.finalize(autotune=True)
        tot, nni = b.run()
        tot2, nni2 = b.run()                    # graph replays give the same numbers
        off, node, side = b.nni_layout()
        u1, u2, ms, spi, nv = b.stats()
    assert np.array_equal(tot, tot2) and np.array_equal(nni, nni2)
    assert spi in (64, 128, 256, 512) and nv == len(nni)
    for i, wl in enumerate(wls):
        t = wl.tree
        mb = np.zeros((t.n_nodes, wl.masks.shape[1]), np.uint32)
        mb[: t.n_tips] = wl.masks
        ref, _ = oracle.Problem(t.left, t.right, t.root, t.blen, mb, 4, wl.Q, wl.pi, pi=wl.pi,
                                cat_rates=wl.cat_rates, cat_w=wl.cat_weights).loglik()
        assert abs(tot[i] - ref) <= 1e-3, (i, tot[i], ref)
        mode = i % 3 - 1
        if mode < 0:
            assert off[i + 1] == off[i]
            continue
        with gpu_context(wl, precision=prec) as ctx:
            _, single = ctx.nni_score(mode)
        got = [(int(node[k]), int(side[k]), nni[k]) for k in range(off[i], off[i + 1])]
        assert [(a, s) for a, s, _ in got] == [(a, s) for a, s, _ in single]
        for (_, _, x), (_, _, y) in zip(got, single):
            assert abs(x - y) <= 1e-9 * abs(y) + 1e-9


def _run_batch(wls, prec, tiling=None):
    with P.Batch(4, prec) as b:
        for i, wl in enumerate(wls):
            t = wl.tree
            b.add(wl.masks, t.left, t.right, t.root, t.blen, wl.Q, wl.pi, wl.cat_rates, wl.cat_weights,
                  nni_mode=i % 3 - 1)
        if tiling:
            b.set_tiling(*tiling)
        b.finalize(autotune=tiling is None)
        tot, nni = b.run()
        return tot, nni, b.tiling()


@pytest.mark.parametrize("prec", [P.PREC_FP32, P.PREC_FP64])
def test_batch_tilings_agree(prec):
    """The category tiling (one category per wave, or one wave walking all
    four) does not change a bit; work-item sizes change only how per-item sums
    are grouped (relative 1e-14); the autotuner reports the tiling it chose."""
    wls = make_instances(7, 5)
    for spi in (64, 256, 4096):
        tot4, nni4, got4 = _run_batch(wls, prec, (spi, 4))
        tot1, nni1, got1 = _run_batch(wls, prec, (spi, 1))
        assert got4 == (spi, 4) and got1 == (spi, 1)
        np.testing.assert_array_equal(tot1, tot4)
        np.testing.assert_array_equal(nni1, nni4)
        if spi == 64:
            ref_tot, ref_nni = tot4, nni4
        np.testing.assert_allclose(tot4, ref_tot, rtol=1e-14)
        np.testing.assert_allclose(nni4, ref_nni, rtol=1e-14)
    tot, nni, got = _run_batch(wls, prec)
    assert got[0] in (64, 128, 256, 512, 1024) and got[1] in (1, 4)
    np.testing.assert_allclose(tot, ref_tot, rtol=1e-14)
    np.testing.assert_allclose(nni, ref_nni, rtol=1e-14)


def test_batch_rejects_bad_tilings():
    with P.Batch(4, P.PREC_FP32) as b:
        for bad in [(0, 4), (100, 4), (8192, 4), (128, 2), (128, 3)]:
            with pytest.raises(P.PhyloError):
                b.set_tiling(*bad)
