"""Synthetic workloads: seeded Yule trees, substitution models and alignments.

Used to build the benchmark inputs (GTR+G4 DNA at 128 taxa x 100k sites,
20-state protein, batches of uneven instances) and property-test data.  The
generator is plain numpy and never part of the timed path.

Conventions follow the reference's encoding: DNA state order A,G,C,T = 0..3
(matrix_parser.cpp:282-291); tips are bit masks with one bit per state and
gaps / 'N' as the all-states mask (treeator.cpp:463-466).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DNA_LETTERS = "AGCT"          # state order of alphabet::set_alphabet_dna
AA20_LETTERS = "ARNDCQEGHILKMFPSTWYV"


@dataclass
class Tree:
    """Rooted binary tree as arrays.  Tips are nodes 0..n_tips-1 (tip i is
    taxon i); internal nodes n_tips..2*n_tips-2; blen[v] is the edge above v."""
    n_tips: int
    left: np.ndarray
    right: np.ndarray
    parent: np.ndarray
    blen: np.ndarray
    root: int
    names: list = field(default_factory=list)

    @property
    def n_nodes(self) -> int:
        return len(self.left)

    def newick(self, digits: int = 17) -> str:
        fmt = "%%.%dg" % digits

        def rec(v):
            if self.left[v] < 0:
                s = self.names[v]
            else:
                s = "(" + rec(self.left[v]) + "," + rec(self.right[v]) + ")"
            if v != self.root:
                s += ":" + (fmt % self.blen[v])
            return s

        import sys
        old = sys.getrecursionlimit()
        sys.setrecursionlimit(max(old, 10 * self.n_nodes + 100))
        try:
            return rec(self.root) + ";"
        finally:
            sys.setrecursionlimit(old)


def yule_tree(n_tips: int, rng: np.random.Generator, mean_blen: float = 0.1,
              prefix: str = "t") -> Tree:
    """Yule topology by repeatedly splitting a uniformly chosen leaf; iid
    Exp(mean_blen) branch lengths."""
    if n_tips < 2:
        raise ValueError("need at least two tips")
    # build with temporary ids, then renumber tips 0..n-1 and internals after
    children = {0: None}
    leaves = [0]
    nxt = 1
    while len(leaves) < n_tips:
        k = int(rng.integers(len(leaves)))
        v = leaves.pop(k)
        a, b = nxt, nxt + 1
        nxt += 2
        children[v] = (a, b)
        children[a] = None
        children[b] = None
        leaves += [a, b]
    tips = [v for v in range(nxt) if children[v] is None]
    internals = [v for v in range(nxt) if children[v] is not None]
    ren = {}
    for i, v in enumerate(tips):
        ren[v] = i
    for i, v in enumerate(internals):
        ren[v] = n_tips + i
    nn = 2 * n_tips - 1
    left = -np.ones(nn, np.int32)
    right = -np.ones(nn, np.int32)
    parent = -np.ones(nn, np.int32)
    for v, ch in children.items():
        if ch is None:
            continue
        a, b = ren[ch[0]], ren[ch[1]]
        left[ren[v]] = a
        right[ren[v]] = b
        parent[a] = ren[v]
        parent[b] = ren[v]
    blen = rng.exponential(mean_blen, nn)
    root = ren[0]
    blen[root] = 0.0
    names = [f"{prefix}{i}" for i in range(n_tips)]
    return Tree(n_tips, left, right, parent, blen, root, names)


def wrap_with_dummy(t: Tree, stem: float = 1.0, name: str = "DUMMY") -> Tree:
    """Returns (T:stem, DUMMY:stem).  Under the reference's default rate model
    the root's child edges get length 0 (rate_model.cpp:106,122); with the
    dummy taxon all missing, the wrapped tree's likelihood equals the textbook
    likelihood of T."""
    n = t.n_tips + 1
    nn = 2 * n - 1
    # tips: old tips 0..n-2, dummy n-1; internals: old internals shifted by 1, new root last
    left = -np.ones(nn, np.int32)
    right = -np.ones(nn, np.int32)
    parent = -np.ones(nn, np.int32)
    blen = np.zeros(nn)

    def m(v):
        return v if v < t.n_tips else v + 1

    for v in range(t.n_nodes):
        if t.left[v] >= 0:
            left[m(v)] = m(t.left[v])
            right[m(v)] = m(t.right[v])
        if t.parent[v] >= 0:
            parent[m(v)] = m(t.parent[v])
        blen[m(v)] = t.blen[v]
    root = nn - 1
    old_root = m(t.root)
    dummy = n - 1
    left[root], right[root] = old_root, dummy
    parent[old_root] = root
    parent[dummy] = root
    blen[old_root] = stem
    blen[dummy] = stem
    return Tree(n, left, right, parent, blen, root, list(t.names) + [name])


def gtr_Q(rates6, pi) -> np.ndarray:
    """Q for a reversible model in the reference's convention
    (sub_model.cpp:99-114): Q_ij = r_ij * pi_j with rate slots in row-major
    upper-triangle order (sub_model.cpp:3-29), diagonal = -row sum."""
    pi = np.asarray(pi, float)
    S = len(pi)
    r = np.asarray(rates6, float)
    Q = np.zeros((S, S))
    k = 0
    for i in range(S):
        for j in range(i + 1, S):
            Q[i, j] = r[k] * pi[j]
            Q[j, i] = r[k] * pi[i]
            k += 1
    np.fill_diagonal(Q, -Q.sum(axis=1))
    return Q


def normalise_Q(Q, pi) -> np.ndarray:
    """Scales Q to mean rate 1 (-sum_i pi_i Q_ii = 1)."""
    mu = -float(np.dot(pi, np.diag(Q)))
    return Q / mu


def expm_reversible(Q, pi, t) -> np.ndarray:
    """P(t) for a reversible Q via the symmetric eigen-decomposition (numpy)."""
    pi = np.asarray(pi, float)
    d = np.sqrt(pi)
    B = (d[:, None] * Q) / d[None, :]
    B = 0.5 * (B + B.T)
    w, U = np.linalg.eigh(B)
    t = np.atleast_1d(np.asarray(t, float))
    E = np.exp(np.multiply.outer(t, w))               # [T, S]
    P = np.einsum("ik,tk,jk->tij", U, E, U)
    P = P * (d[None, None, :] / d[None, :, None])
    return np.clip(P, 0.0, None)


def gamma_rates(alpha: float, C: int) -> np.ndarray:
    """Mean rates of C equal-probability discrete-gamma categories (Yang 1994)."""
    if C == 1:
        return np.ones(1)
    from scipy.special import gammainc, gammaincinv
    q = gammaincinv(alpha, np.arange(1, C) / C) / alpha
    I = np.concatenate([[0.0], gammainc(alpha + 1, q * alpha), [1.0]])
    return np.diff(I) * C


def simulate(t: Tree, Q, pi, m: int, rng: np.random.Generator, cat_rates=(1.0,)) -> np.ndarray:
    """Simulates m sites down the tree; returns states [n_nodes, m] (int8)."""
    S = len(pi)
    cat_rates = np.asarray(cat_rates, float)
    C = len(cat_rates)
    states = np.zeros((t.n_nodes, m), np.int8)
    cats = rng.integers(C, size=m) if C > 1 else np.zeros(m, np.int64)
    states[t.root] = rng.choice(S, size=m, p=np.asarray(pi) / np.sum(pi))
    order = [t.root]
    i = 0
    while i < len(order):
        v = order[i]
        i += 1
        order += [int(c) for c in (t.left[v], t.right[v]) if c >= 0]
    for v in order[1:]:
        p = t.parent[v]
        P = expm_reversible(Q, pi, t.blen[v] * cat_rates)            # [C, S, S]
        cum = np.cumsum(P, axis=2)
        cum[:, :, -1] = 1.0
        rows = cum[cats, states[p]]                                   # [m, S]
        u = rng.random(m)
        states[v] = (u[:, None] > rows).sum(axis=1).astype(np.int8)
    return states


def masks_from_states(states_tips: np.ndarray, gap_frac: float, rng: np.random.Generator,
                      S: int) -> np.ndarray:
    """One-hot tip masks with a fraction of cells replaced by the all-states mask."""
    masks = (np.uint32(1) << states_tips.astype(np.uint32)).astype(np.uint32)
    if gap_frac > 0:
        g = rng.random(masks.shape) < gap_frac
        masks[g] = np.uint32((1 << S) - 1)
    return masks


def fasta_text(names, masks, letters: str, S: int) -> str:
    """Fasta with one letter per state; the all-states mask is written as '-'
    (DNA 'N' also maps to all states)."""
    full = (1 << S) - 1
    lut = {1 << i: letters[i] for i in range(S)}
    if letters == DNA_LETTERS:
        # IUPAC codes in the A,G,C,T bit order of alphabet::set_alphabet_dna
        lut.update({3: "R", 12: "Y", 6: "S", 9: "W", 10: "K", 5: "M", 14: "B", 11: "D", 13: "H", 7: "V"})
    lut[full] = "-"
    out = []
    for name, row in zip(names, masks):
        out.append(">" + name)
        out.append("".join(lut[int(x)] for x in row))
    return "\n".join(out) + "\n"


@dataclass
class Workload:
    tree: Tree
    masks: np.ndarray        # [n_tips, m] uint32
    Q: np.ndarray
    pi: np.ndarray
    rates: np.ndarray        # exchangeabilities (S*(S-1)/2)
    cat_rates: np.ndarray
    cat_weights: np.ndarray
    alpha: float


def gtr_gamma_workload(n_tips=128, m=100_000, mean_blen=0.05, alpha=0.5, C=4, gap_frac=0.05,
                       seed=2, S=4, rate_conc=1.0, freq_conc=5.0, rate_dist="dirichlet") -> Workload:
    """GTR+G DNA (S=4) or a general reversible S-state model: Yule tree,
    exchangeabilities ~ Dirichlet(rate_conc) (or Gamma(1,1) when
    rate_dist == 'gamma'), frequencies ~ Dirichlet(freq_conc), Q scaled to mean
    rate 1, discrete gamma with C categories, a fraction of cells set to gap."""
    rng = np.random.default_rng(seed)
    t = yule_tree(n_tips, rng, mean_blen)
    nr = S * (S - 1) // 2
    if rate_dist == "gamma":
        rates = rng.gamma(1.0, 1.0, nr)
    else:
        rates = rng.dirichlet(np.full(nr, rate_conc))
    pi = rng.dirichlet(np.full(S, freq_conc))
    Q = gtr_Q(rates, pi)
    mu = -float(np.dot(pi, np.diag(Q)))
    rates = rates / mu
    Q = gtr_Q(rates, pi)
    cr = gamma_rates(alpha, C) if C > 1 else np.ones(1)
    states = simulate(t, Q, pi, m, rng, cr)
    masks = masks_from_states(states[: t.n_tips], gap_frac, rng, S)
    return Workload(t, masks, Q, pi, rates, cr, np.full(C, 1.0 / C), alpha)


# ---- the batched-scoring workload: many alignments of uneven shape ---------

def batch_instance_params(n_inst: int = 512, seed: int = 3):
    """Per-instance (taxa, sites, gamma shape) of the batched workload: taxa ~
    U{8..256}, sites log-uniform in [500, 50000], shape ~ U(0.3, 2); instance
    i is gtr_gamma_workload(seed=10000 + i) with these (mean length 0.05,
    4 categories, 5% gaps)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        n = int(rng.integers(8, 257))
        m = int(np.exp(rng.uniform(np.log(500), np.log(50_000))))
        out.append((n, m, float(rng.uniform(0.3, 2.0))))
    return out


def batch_instance(i: int, n: int, m: int, alpha: float) -> Workload:
    return gtr_gamma_workload(n_tips=n, m=m, C=4, alpha=alpha, gap_frac=0.05, seed=10_000 + i, mean_blen=0.05)


def _batch_instance_packed(args):
    i, n, m, alpha = args
    wl = batch_instance(i, n, m, alpha)
    wl.masks = wl.masks.astype(np.uint8)          # DNA state sets fit a byte: a quarter of the transfer
    return wl


def batch_workloads(n_inst: int = 512, seed: int = 3, workers: int = 16):
    """The batched workload's instances, simulated by `workers` spawned
    processes (masks as uint8).  Spawned rather than forked, so a caller that
    has already initialised the GPU can use it."""
    import multiprocessing as mp
    params = [(i, *p) for i, p in enumerate(batch_instance_params(n_inst, seed))]
    if workers <= 1:
        return [_batch_instance_packed(a) for a in params]
    with mp.get_context("spawn").Pool(workers) as pool:
        return pool.map(_batch_instance_packed, params, chunksize=4)
