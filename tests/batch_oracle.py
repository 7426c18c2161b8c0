"""Oracle side of the full-size batched-scoring test: each worker process
regenerates one instance of the seeded 512-alignment workload and scores
its total and a few NNI variants with the FP64 oracle.  Test
infrastructure only (imported by spawned pool workers)."""
import numpy as np

from phylommand_amd import synth


def score_instance(args):
    """args: (i, n, m, alpha, variants [(node, side)]) -> (i, total, [lnl])."""
    from oracle import oracle as O
    i, n, m, alpha, variants = args
    wl = synth.batch_instance(i, n, m, alpha)
    t = wl.tree
    mb = np.zeros((t.n_nodes, m), np.uint32)
    mb[:n] = wl.masks
    prob = dict(S=4, Q=wl.Q, root_w=wl.pi, pi=wl.pi, cat_rates=wl.cat_rates, cat_w=wl.cat_weights)
    total, _ = O.Problem(t.left, t.right, t.root, t.blen, mb, **prob).loglik()
    out = []
    for v, side in variants:
        l2, r2 = O.nni_apply(t.left, t.right, t.root, v, side)
        out.append(O.Problem(l2, r2, t.root, t.blen, mb, **prob).loglik()[0])
    return i, total, out
