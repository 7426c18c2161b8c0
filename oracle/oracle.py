"""ctypes binding of the FP64 CPU oracle (oracle/phylo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker -- never by the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "libphylo_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

REFCOMPAT, EIG, PADE = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint32)
_fp = ctypes.POINTER(ctypes.c_float)
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = ctypes.CDLL(SO)
        L.orc_tree_parse.restype = ctypes.c_void_p
        L.orc_tree_parse.argtypes = [ctypes.c_char_p]
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_n_nodes.argtypes = [ctypes.c_void_p]
        L.orc_tree_root.argtypes = [ctypes.c_void_p]
        L.orc_tree_arrays.argtypes = [ctypes.c_void_p, _ip, _ip, _ip, _dp]
        L.orc_tree_label.restype = ctypes.c_char_p
        L.orc_tree_label.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_tree_newick.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_parse_fasta.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                      _up, ctypes.c_int, _ip, _ip]
        L.orc_set_custom_alphabet.argtypes = [_up]
        L.orc_alphabet_mask.restype = ctypes.c_uint32
        L.orc_alphabet_mask.argtypes = [ctypes.c_int, ctypes.c_char]
        L.orc_n_states.argtypes = [_up, ctypes.c_long]
        L.orc_rate_pos.argtypes = [ctypes.c_int] * 4
        L.orc_freqs.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_build_Q.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp]
        L.orc_pmatrix.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_time_mod_length.restype = ctypes.c_double
        L.orc_time_mod_length.argtypes = [ctypes.c_double, ctypes.c_float, ctypes.c_int, _fp, _fp]
        L.orc_effective_lengths.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, _fp, _fp,
                                            _ip, ctypes.c_int, _fp, ctypes.c_int, _dp]
        L.orc_gammainc.restype = ctypes.c_double
        L.orc_gammainc.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_gammaincinv.restype = ctypes.c_double
        L.orc_gammaincinv.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_gamma_rates.argtypes = [ctypes.c_double, ctypes.c_int, _dp]
        common = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _up, ctypes.c_int, _dp, _dp,
                  _dp, _dp, ctypes.c_int]
        L.orc_loglik.argtypes = common + [ctypes.c_int, _dp, _dp, _dp]
        L.orc_node_clvs.argtypes = common + [_dp, _dp, _ip, _dp, _dp]
        L.orc_node_partials.argtypes = common + [_dp, ctypes.c_int, _dp]
        L.orc_nni_reference.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_nni_list.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip]
        L.orc_nni_apply.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_simulate.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                   _dp, ctypes.c_int, _dp, ctypes.c_uint64, _ip]
        L.orc_time_reference_cost.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int,
                                              _up, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OTree:
    """A tree parsed by the oracle's newick restatement (node 0 = root)."""

    def __init__(self, newick: str):
        L = lib()
        h = L.orc_tree_parse(newick.encode())
        n = L.orc_tree_n_nodes(h)
        self.n_nodes = n
        self.root = L.orc_tree_root(h)
        self.left = np.zeros(n, np.int32)
        self.right = np.zeros(n, np.int32)
        self.parent = np.zeros(n, np.int32)
        self.blen = np.zeros(n)
        if n:
            L.orc_tree_arrays(h, self.left.ctypes.data_as(_ip), self.right.ctypes.data_as(_ip),
                              self.parent.ctypes.data_as(_ip), self.blen.ctypes.data_as(_dp))
        self.labels = [L.orc_tree_label(h, i).decode() for i in range(n)]
        need = L.orc_tree_newick(h, None, 0)
        buf = ctypes.create_string_buffer(need + 1)
        L.orc_tree_newick(h, buf, need + 1)
        self.printed = buf.value.decode()
        L.orc_tree_free(h)

    def tips(self):
        return [v for v in range(self.n_nodes) if self.left[v] < 0 and self.right[v] < 0]


def alphabet_table(text: str) -> np.ndarray:
    """Character -> state-mask table of a whitespace alphabet file: each line is
    a character followed by the state numbers it stands for
    (alphabet_parser::pars_whitespace, matrix_parser.cpp:227-272)."""
    table = np.zeros(256, np.uint32)
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        ch = line.lstrip()[0]
        m = 0
        for tok in parts[1:]:
            if tok.isdigit() and int(tok) < 24:
                m |= 1 << int(tok)
        table[ord(ch)] = m
    return table


def parse_fasta(text: str, alphabet=0, set_empty: bool = True):
    """Returns (names, masks[n_taxa, n_sites]).  alphabet: 0 dna, 1 protein,
    2 binary, or a 256-entry table from alphabet_table()."""
    L = lib()
    if not isinstance(alphabet, int):
        tab = np.ascontiguousarray(alphabet, dtype=np.uint32)
        L.orc_set_custom_alphabet(tab.ctypes.data_as(_up))
        alphabet = 3
    nsites = ctypes.c_int(0)
    nt = L.orc_parse_fasta(text.encode(), alphabet, int(set_empty), None, 0, None, 0, None, ctypes.byref(nsites))
    names = ctypes.create_string_buffer(len(text) + 16)
    masks = np.zeros((nt, max(nsites.value, 1)), np.uint32)
    rl = np.zeros(max(nt, 1), np.int32)
    L.orc_parse_fasta(text.encode(), alphabet, int(set_empty), names, len(names), masks.ctypes.data_as(_up),
                      masks.shape[1], rl.ctypes.data_as(_ip), ctypes.byref(nsites))
    raw = names.raw.split(b"\0")[:nt]
    return [x.decode() for x in raw], masks[:, : nsites.value]


def build_Q(S, tr, spec, rates, freqs=None):
    Q = np.zeros((S, S))
    f = None if freqs is None else _d(list(freqs) + [0.0] * (S - len(freqs)))
    lib().orc_build_Q(S, int(tr), _i(spec).ctypes.data_as(_ip), _d(rates).ctypes.data_as(_dp),
                      None if f is None else f.ctypes.data_as(_dp), Q.ctypes.data_as(_dp))
    return Q


def freqs(S, freqs_in=None):
    out = np.zeros(S)
    f = None if freqs_in is None else _d(list(freqs_in) + [0.0] * (S - len(freqs_in)))
    lib().orc_freqs(S, None if f is None else f.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def pmatrix(Q, t, builder=EIG, pi=None):
    Q = _d(Q)
    S = Q.shape[0]
    P = np.zeros((S, S))
    p = _d(pi) if pi is not None else np.full(S, 1.0 / S)
    rc = lib().orc_pmatrix(S, Q.ctypes.data_as(_dp), p.ctypes.data_as(_dp), float(t), builder, P.ctypes.data_as(_dp))
    assert rc == 0
    return P


def gamma_rates(alpha, C):
    r = np.zeros(C)
    lib().orc_gamma_rates(float(alpha), C, r.ctypes.data_as(_dp))
    return r


def time_mod_length(b, d, cuts=(), cut_rates=()):
    c = np.ascontiguousarray(cuts, dtype=np.float32)
    r = np.ascontiguousarray(cut_rates, dtype=np.float32)
    return lib().orc_time_mod_length(float(b), ctypes.c_float(d), len(c), c.ctypes.data_as(_fp),
                                     r.ctypes.data_as(_fp))


def effective_lengths(left, right, blen, root, cuts=(), cut_rates=(), clade_of_node=None, clade_rates=(),
                      root_quirk=True):
    n = len(left)
    eff = np.zeros(n)
    c = np.ascontiguousarray(cuts, dtype=np.float32)
    r = np.ascontiguousarray(cut_rates, dtype=np.float32)
    cr = np.ascontiguousarray(clade_rates, dtype=np.float32)
    con = None if clade_of_node is None else _i(clade_of_node)
    lib().orc_effective_lengths(n, int(root), _i(left).ctypes.data_as(_ip), _i(right).ctypes.data_as(_ip),
                                _d(blen).ctypes.data_as(_dp), len(c), c.ctypes.data_as(_fp), r.ctypes.data_as(_fp),
                                None if con is None else con.ctypes.data_as(_ip), len(cr), cr.ctypes.data_as(_fp),
                                int(root_quirk), eff.ctypes.data_as(_dp))
    return eff


class Problem:
    """Inputs of one likelihood evaluation in the oracle's node numbering:
    masks_by_node [n_nodes, m] (tips only are read)."""

    def __init__(self, left, right, root, eff, masks_by_node, S, Q, root_w, pi=None, cat_rates=(1.0,),
                 cat_w=None, builder=EIG):
        self.left, self.right, self.root = _i(left), _i(right), int(root)
        self.eff = _d(eff)
        self.masks = np.ascontiguousarray(masks_by_node, dtype=np.uint32)
        self.m = self.masks.shape[1]
        self.S = S
        self.Q = _d(Q)
        self.root_w = _d(root_w)
        self.pi = _d(pi if pi is not None else root_w)
        self.cat_rates = _d(cat_rates)
        self.C = len(self.cat_rates)
        self.cat_w = _d(cat_w if cat_w is not None else np.full(self.C, 1.0 / self.C))
        self.builder = builder

    def _args(self):
        return [len(self.left), self.root, self.left.ctypes.data_as(_ip), self.right.ctypes.data_as(_ip),
                self.eff.ctypes.data_as(_dp), self.m, self.S, self.masks.ctypes.data_as(_up), self.C,
                self.cat_rates.ctypes.data_as(_dp), self.cat_w.ctypes.data_as(_dp), self.Q.ctypes.data_as(_dp),
                self.pi.ctypes.data_as(_dp), self.builder]

    def loglik(self, per_site_p=False):
        site = np.zeros(self.m)
        tot = np.zeros(1)
        rc = lib().orc_loglik(*self._args(), int(per_site_p), self.root_w.ctypes.data_as(_dp),
                              site.ctypes.data_as(_dp), tot.ctypes.data_as(_dp))
        assert rc == 0
        return float(tot[0]), site

    def node_clvs(self):
        n = len(self.left)
        clv = np.zeros((n, self.m, self.C, self.S))
        sc = np.zeros((n, self.m), np.int32)
        site = np.zeros(self.m)
        tot = np.zeros(1)
        rc = lib().orc_node_clvs(*self._args(), self.root_w.ctypes.data_as(_dp), clv.ctypes.data_as(_dp),
                                 sc.ctypes.data_as(_ip), site.ctypes.data_as(_dp), tot.ctypes.data_as(_dp))
        assert rc == 0
        return clv, sc, site, float(tot[0])

    def node_partials(self, reset_per_site=False):
        n = len(self.left)
        out = np.zeros((n, self.m, self.S))
        rc = lib().orc_node_partials(*self._args(), self.root_w.ctypes.data_as(_dp), int(reset_per_site),
                                     out.ctypes.data_as(_dp))
        assert rc == 0
        return out

    def time_reference_cost(self):
        tot = np.zeros(1)
        sec = np.zeros(1)
        rc = lib().orc_time_reference_cost(len(self.left), self.root, self.left.ctypes.data_as(_ip),
                                           self.right.ctypes.data_as(_ip), self.eff.ctypes.data_as(_dp), self.m,
                                           self.S, self.masks.ctypes.data_as(_up), self.C,
                                           self.cat_rates.ctypes.data_as(_dp), self.cat_w.ctypes.data_as(_dp),
                                           self.Q.ctypes.data_as(_dp), self.root_w.ctypes.data_as(_dp),
                                           tot.ctypes.data_as(_dp), sec.ctypes.data_as(_dp))
        assert rc == 0
        return float(tot[0]), float(sec[0])


def nni_reference(newick: str):
    cap = 64 * (len(newick) + 64) * 1024
    buf = ctypes.create_string_buffer(cap)
    n = lib().orc_nni_reference(newick.encode(), buf, cap)
    return buf.value.decode().strip().split("\n")[:n]


def simulate(left, right, root, blen, m, Q, pi, freqs_root, cat_rates=(1.0,), seed=1):
    n = len(left)
    S = len(pi)
    states = np.zeros((n, m), np.int32)
    cr = _d(cat_rates)
    lib().orc_simulate(n, int(root), _i(left).ctypes.data_as(_ip), _i(right).ctypes.data_as(_ip),
                       _d(blen).ctypes.data_as(_dp), m, S, _d(Q).ctypes.data_as(_dp), _d(pi).ctypes.data_as(_dp),
                       _d(freqs_root).ctypes.data_as(_dp), len(cr), cr.ctypes.data_as(_dp), seed,
                       states.ctypes.data_as(_ip))
    return states


def nni_list(left, right, root, mode=0):
    """[(node, side)] of the NNI variants (mode 0: treebender --nni all; 1: full
    unrooted neighbourhood)."""
    n = len(left)
    nodes = np.zeros(2 * n, np.int32)
    sides = np.zeros(2 * n, np.int32)
    k = lib().orc_nni_list(n, int(root), _i(left).ctypes.data_as(_ip), _i(right).ctypes.data_as(_ip), mode,
                           nodes.ctypes.data_as(_ip), sides.ctypes.data_as(_ip))
    return list(zip(nodes[:k].tolist(), sides[:k].tolist()))


def nni_apply(left, right, root, node, side):
    n = len(left)
    ol = np.zeros(n, np.int32)
    orr = np.zeros(n, np.int32)
    lib().orc_nni_apply(n, int(root), _i(left).ctypes.data_as(_ip), _i(right).ctypes.data_as(_ip), int(node),
                        int(side), ol.ctypes.data_as(_ip), orr.ctypes.data_as(_ip))
    return ol, orr
