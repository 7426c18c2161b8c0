"""Writes tests/golden/nm_plateau_32x2000.npz: held-out alignment 330 of the
learned-proposal comparison (1,000 GTR+G4 models on 32-taxon Yule trees drawn
with numpy seed 99, 2,000 sites simulated on the GPU with seed 77), the one
fit from the default start that stalled 188.7 lnL units below the fit from
the proposal's start.  Needs a GPU (the simulator is a HIP kernel).

stalled_x is where the plain Nelder-Mead run without the coordinate-reset
probe stopped (lnL -54737.31); the test reads it to check the plateau on the
CPU.  It comes from the fit's Nelder-Mead restated over the oracle's lnL, so
pass it with --stalled-x FILE.npy when regenerating.

    python tests/golden/make_nm_plateau.py --stalled-x x.npy
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stalled-x", required=True)
    ap.add_argument("--seeded-lnl", type=float, default=-54548.58320311)
    ap.add_argument("--stalled-lnl", type=float, default=-54737.31254465)
    a = ap.parse_args()
    from phylommand_amd import learn as L
    k = 330
    held = L.draw_models(1000, 32, np.random.default_rng(99))
    _, masks = L.simulate_features(held, 2000, seed=77, want_masks=True)
    np.savez_compressed(os.path.join(HERE, "nm_plateau_32x2000.npz"), masks=masks[k].astype(np.uint8),
                        left=held.left[k], right=held.right[k], root=held.roots[k], blen=held.blen[k],
                        rates=held.rates[k], pi=held.pi[k], alpha=held.alpha[k], seeded_lnl=a.seeded_lnl,
                        stalled_lnl=a.stalled_lnl, stalled_x=np.load(a.stalled_x))


if __name__ == "__main__":
    main()
