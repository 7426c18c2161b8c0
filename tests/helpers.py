"""Shared test helpers: run one workload through the oracle and the GPU."""
import numpy as np

from phylommand_amd import synth


def oracle_problem(O, wl, eff=None, builder=None, cat=True):
    t = wl.tree
    S = len(wl.pi)
    mb = np.zeros((t.n_nodes, wl.masks.shape[1]), np.uint32)
    mb[: t.n_tips] = wl.masks
    eff = t.blen if eff is None else eff
    return O.Problem(t.left, t.right, t.root, eff, mb, S, wl.Q, wl.pi, pi=wl.pi,
                     cat_rates=wl.cat_rates if cat else (1.0,), cat_w=wl.cat_weights if cat else None,
                     builder=O.EIG if builder is None else builder)


def gpu_context(wl, precision=0, builder=1, eff=None):
    import phylommand_amd as P
    t = wl.tree
    ctx = P.Context(0, precision)
    ctx.set_tips(wl.masks, len(wl.pi))
    ctx.set_topology(t.left, t.right, t.root)
    ctx.set_edge_lengths(t.blen if eff is None else eff)
    ctx.set_model(wl.Q, wl.pi, wl.pi, builder)
    ctx.set_categories(wl.cat_rates, wl.cat_weights)
    return ctx
