"""Pins the two oracle functions that no reference fixture can reach.

orc_simulate: the reference draws from libc rand() seeded by the clock
(simpleML.cpp:233-292, tree.h:61-64), so no draw can be replayed.  Its
distribution can: on a 4-taxon tree the 256 tip patterns have the exact
probabilities exp(site lnL) of the pruning likelihood, which is itself pinned
by the golden per-site fixtures.  A chi-square test against them passes for
the right simulator and fails for each plausible mistake, planted on the
oracle's own inputs: P = I, swapped branch lengths, rate categories ignored,
the root drawn from the wrong frequencies.

orc_nni_list mode 1 (every internal edge of the unrooted tree, the
interchange of tree.cpp:2480-2518 applied at each): checked by brute force
on split sets.  Two binary unrooted trees are NNI neighbours exactly when
their split sets differ in one split each way (Robinson-Foulds distance 2),
so the neighbourhood is the set of all unrooted topologies at that distance,
enumerated here by stepwise addition for n = 5..8.  The variants must be
2(n-3) distinct trees forming exactly that set.  Wrong root-edge rules
(the root edge counted twice or not at all) and identity variants fail.
CPU only.
"""
import itertools

import numpy as np
import pytest
from scipy.stats import chi2

from phylommand_amd import synth


# ---------------------------------------------------------------- simulate

def _pattern_masks():
    pats = np.array([[(p >> (2 * (3 - t))) & 3 for p in range(256)] for t in range(4)])
    return (np.uint32(1) << pats.astype(np.uint32)).astype(np.uint32)


def _chi2(counts, p, m):
    exp = p * m
    keep = exp >= 5
    stat = float(np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep]))
    return stat, chi2.ppf(1 - 1e-6, keep.sum() - 1)


@pytest.fixture(scope="module")
def sim_case(oracle):
    rng = np.random.default_rng(31)
    t = synth.yule_tree(4, rng, 0.2)
    # distinct, well separated lengths so a swap is visible
    t.blen[:] = [0.05, 0.45, 0.15, 0.8, 0.3, 0.6, 0.0][: t.n_nodes]
    t.blen[t.root] = 0.0
    pi = np.array([0.4, 0.1, 0.2, 0.3])
    Q = synth.normalise_Q(synth.gtr_Q([2.5, 0.6, 1.1, 0.8, 0.4, 3.0], pi), pi)
    cats = oracle.gamma_rates(0.4, 4)
    mb = np.zeros((t.n_nodes, 256), np.uint32)
    mb[:4] = _pattern_masks()
    _, site = oracle.Problem(t.left, t.right, t.root, t.blen, mb, 4, Q, pi, pi=pi, cat_rates=cats,
                             builder=oracle.EIG).loglik()
    p = np.exp(site)
    assert p.sum() == pytest.approx(1.0, abs=1e-10)
    return t, Q, pi, cats, p


def _counts(oracle, t, Q, pi, root_freq, cats, blen, m, seed):
    st = oracle.simulate(t.left, t.right, t.root, blen, m, Q, pi, root_freq, cat_rates=cats, seed=seed)
    code = st[0] * 64 + st[1] * 16 + st[2] * 4 + st[3]
    return np.bincount(code, minlength=256)


M_SIM = 300_000


def test_simulate_pattern_frequencies(oracle, sim_case):
    t, Q, pi, cats, p = sim_case
    counts = _counts(oracle, t, Q, pi, pi, cats, t.blen, M_SIM, seed=7)
    stat, crit = _chi2(counts, p, M_SIM)
    assert stat < crit, (stat, crit)


@pytest.mark.parametrize("mistake", ["p_identity", "swapped_lengths", "no_categories", "root_freq"])
def test_simulate_planted_mistakes_fail(oracle, sim_case, mistake):
    t, Q, pi, cats, p = sim_case
    blen, c, rf = t.blen.copy(), cats, pi
    if mistake == "p_identity":
        blen[:] = 0.0
    elif mistake == "swapped_lengths":
        blen[0], blen[1] = blen[1], blen[0]
    elif mistake == "no_categories":
        c = np.ones(1)
    elif mistake == "root_freq":
        rf = np.full(4, 0.25)
    counts = _counts(oracle, t, Q, pi, rf, c, blen, M_SIM, seed=7)
    stat, crit = _chi2(counts, p, M_SIM)
    assert stat > crit, (mistake, stat, crit)


# ---------------------------------------------------------------- NNI mode 1

def _splits(left, right, root, n_tips):
    """Non-trivial splits of the unrooted tree, each as the frozenset of taxa
    on the side without taxon 0."""
    below = {}

    def rec(v):
        if left[v] < 0:
            below[v] = frozenset([v])
        else:
            below[v] = rec(left[v]) | rec(right[v])
        return below[v]

    rec(root)
    full = frozenset(range(n_tips))
    out = set()
    for v, s in below.items():
        if v == root:
            continue
        side = full - s if 0 in s else s
        if 2 <= len(side) <= n_tips - 2:
            out.add(side)
    return frozenset(out)


def _all_unrooted_split_sets(n):
    """Split sets of every binary unrooted tree on taxa 0..n-1, by stepwise
    addition of taxa onto every edge of the trees on fewer taxa."""
    trees = [[(0, -1), (1, -1), (2, -1)]]      # edges of the 3-taxon star; -1 = internal node
    nxt = -2
    for k in range(3, n):
        new = []
        for edges in trees:
            for i in range(len(edges)):
                a, b = edges[i]
                u = nxt
                e2 = edges[:i] + edges[i + 1:] + [(a, u), (u, b), (k, u)]
                new.append(e2)
        nxt -= 1
        trees = new
    full = frozenset(range(n))
    result = set()
    for edges in trees:
        adj = {}
        for a, b in edges:
            adj.setdefault(a, []).append(b)
            adj.setdefault(b, []).append(a)
        splits = set()
        for a, b in edges:
            if a >= 0 or b >= 0:
                continue                    # terminal edge
            # taxa reachable from b without crossing (a, b)
            seen, stack = {a, b}, [b]
            taxa = set()
            while stack:
                x = stack.pop()
                for y in adj[x]:
                    if y not in seen:
                        seen.add(y)
                        stack.append(y)
                        if y >= 0:
                            taxa.add(y)
            side = frozenset(full - taxa) if 0 in taxa else frozenset(taxa)
            splits.add(side)
        result.add(frozenset(splits))
    return result


def _check_neighbourhood(oracle, t, variants, universe):
    """Raises AssertionError unless `variants` ((node, side) pairs) is the
    complete NNI neighbourhood of t, each variant once."""
    n = t.n_tips
    base = _splits(t.left, t.right, t.root, n)
    assert len(base) == n - 3
    got = []
    for v, side in variants:
        l2, r2 = oracle.nni_apply(t.left, t.right, t.root, v, side)
        s2 = _splits(l2, r2, t.root, n)
        assert len(s2) == n - 3, "variant is not a binary tree on the same taxa"
        assert len(base - s2) == 1 and len(s2 - base) == 1, f"variant ({v},{side}) is not one NNI away"
        got.append(s2)
    assert len(set(got)) == len(got), "duplicate variants"
    want = {s for s in universe if len(s - base) == 1 and len(base - s) == 1}
    assert len(want) == 2 * (n - 3)
    assert set(got) == want, "variants are not the whole neighbourhood"


def _trees(n):
    rng = np.random.default_rng(100 + n)
    out = [synth.yule_tree(n, rng) for _ in range(6)]
    # caterpillar rooted at its end: one root child is a tip
    left = -np.ones(2 * n - 1, np.int32)
    right = -np.ones(2 * n - 1, np.int32)
    prev = 0
    for i in range(1, n):
        v = n + i - 1
        left[v], right[v] = prev, i
        prev = v
    par = -np.ones(2 * n - 1, np.int32)
    for v in range(2 * n - 1):
        for c in (left[v], right[v]):
            if c >= 0:
                par[c] = v
    out.append(synth.Tree(n, left, right, par, np.full(2 * n - 1, 0.1), 2 * n - 2, [f"t{i}" for i in range(n)]))
    return out


@pytest.mark.parametrize("n", [5, 6, 7, 8])
def test_nni_mode1_is_the_full_neighbourhood(oracle, n):
    universe = _all_unrooted_split_sets(n)
    double_factorial = int(np.prod(np.arange(2 * n - 5, 0, -2)))
    assert len(universe) == double_factorial          # (2n-5)!! unrooted topologies
    for t in _trees(n):
        variants = oracle.nni_list(t.left, t.right, t.root, mode=1)
        assert len(variants) == 2 * (n - 3)
        _check_neighbourhood(oracle, t, variants, universe)


def test_nni_mode1_planted_mistakes_fail(oracle):
    """The same check rejects the plausible ways a mode-1 enumeration goes
    wrong around the root."""
    n = 7
    universe = _all_unrooted_split_sets(n)
    rng = np.random.default_rng(5)
    t = None
    for _ in range(200):                  # a tree whose root children are both internal
        c = synth.yule_tree(n, rng)
        if c.left[c.left[c.root]] >= 0 and c.left[c.right[c.root]] >= 0:
            t = c
            break
    assert t is not None
    good = oracle.nni_list(t.left, t.right, t.root, mode=1)
    _check_neighbourhood(oracle, t, good, universe)
    a, b = t.left[t.root], t.right[t.root]
    root_twice = good + [(b, 0), (b, 1)]               # the root edge counted from both sides
    root_missing = [x for x in good if x[0] != a]      # the root edge left out
    identity = good[:-1] + [(t.root, 0)]               # an interchange that changes nothing
    for bad in (root_twice, root_missing, identity):
        with pytest.raises(AssertionError):
            _check_neighbourhood(oracle, t, bad, universe)
