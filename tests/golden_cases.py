"""Loads the committed golden fixtures (tests/golden/) into arrays the oracle
and the GPU library both take."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

PERSITE_CASES = ["c1_jc16", "t4_jc_rootzero", "t4_jc_wrapped", "t4_nontr", "t4_2param", "t4_unrooted",
                 "dna3_24x300", "nontr_dna_10x200", "binary_8x40", "protein20_8x60"]


def dlist(s):
    return [] if s == "-" else [float(x) for x in s.split(",")]


def read_expected(path):
    site, total = [], None
    for line in open(path):
        line = line.strip()
        if not line:
            continue
        if line.startswith("total"):
            total = float(line.split()[1])
        else:
            site.append(float(line))
    return np.array(site), total


class Case:
    """One golden likelihood case: tree, alignment, model, expected per-site lnL."""

    def __init__(self, O, name, tree_file="tree.nwk", aln_file="aln.fst", expected="expected.txt", model=None):
        d = os.path.join(GOLDEN, name)
        self.name = name
        meta = json.load(open(os.path.join(d, "case.json")))
        if model is not None:
            spec, params, freqs = model
            meta = dict(meta, spec=spec, params=params, freqs=freqs, tr=1, alphabet="dna")
        self.meta = meta
        self.tree = O.OTree(open(os.path.join(d, tree_file)).read())
        alpha = meta["alphabet"]
        if alpha == "dna":
            a = 0
        elif alpha == "protein":
            a = 1
        elif alpha == "binary":
            a = 2
        else:
            a = O.alphabet_table(open(os.path.join(d, alpha)).read())
        names, masks = O.parse_fasta(open(os.path.join(d, aln_file)).read(), a)
        self.names = names
        self.S = O.lib().orc_n_states(masks.ctypes.data_as(O._up), masks.size)
        self.m = masks.shape[1]
        # tip rows in tree node numbering (set_char looks a taxon up by label, simpleML.h:45)
        self.masks_by_node = np.zeros((self.tree.n_nodes, self.m), np.uint32)
        for v in self.tree.tips():
            self.masks_by_node[v] = masks[names.index(self.tree.labels[v])]
        self.tr = bool(meta["tr"])
        spec = [int(x) for x in dlist(meta["spec"])]
        rates = dlist(meta["params"])
        fr = dlist(meta["freqs"])
        self.spec, self.rates = spec, rates
        self.freqs_in = fr if (self.tr and fr) else None
        self.Q = O.build_Q(self.S, self.tr, spec, rates, self.freqs_in)
        self.pi = O.freqs(self.S, self.freqs_in)
        self.root_w = self.pi if self.tr else np.ones(self.S)
        self.eff = O.effective_lengths(self.tree.left, self.tree.right, self.tree.blen, self.tree.root)
        self.expected_site, self.expected_total = read_expected(os.path.join(d, expected))

    def problem(self, O, builder=None):
        return O.Problem(self.tree.left, self.tree.right, self.tree.root, self.eff, self.masks_by_node, self.S,
                         self.Q, self.root_w, pi=self.pi, builder=O.REFCOMPAT if builder is None else builder)

    def library_arrays(self):
        """Renumbers nodes for libphylo_hip: tips 0..n_tips-1 first, internal after."""
        t = self.tree
        tips = t.tips()
        internal = [v for v in range(t.n_nodes) if v not in set(tips)]
        order = tips + internal
        new = {v: i for i, v in enumerate(order)}
        left = np.array([new[t.left[v]] if t.left[v] >= 0 else -1 for v in order], np.int32)
        right = np.array([new[t.right[v]] if t.right[v] >= 0 else -1 for v in order], np.int32)
        eff = np.array([self.eff[v] for v in order])
        masks = np.array([self.masks_by_node[v] for v in tips], np.uint32)
        return masks, left, right, new[t.root], eff
