"""phylommand_amd: MI355X-native Felsenstein-pruning likelihood for phylommand's
treeator path.

The compute lives in libphylo_hip.so (hand-written gfx950 HIP kernels behind
the C ABI declared in include/phylo_hip.h).  This module is a thin ctypes
binding over that ABI, used by the tests and the benchmark; the
simpleML/sub_model/rate_model-compatible C++ host classes are in
phylommand_amd/host/.  There is no CPU fallback: if the library is missing or
no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphylo_hip.so")

PHYLO_OK = 0
PREC_FP32, PREC_FP64 = 0, 1
P_REFCOMPAT, P_EIGEN, P_PADE = 0, 1, 2

_lib = None

# (name, restype, argtypes) of every entry point in include/phylo_hip.h
_c_int, _c_dbl_p, _c_int_p, _c_void_p = ctypes.c_int, ctypes.POINTER(ctypes.c_double), \
    ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p
_c_flt_p = ctypes.POINTER(ctypes.c_float)
SIGNATURES = [
    ("phylo_version", _c_int, []),
    ("phylo_create", _c_int, [_c_int, ctypes.POINTER(_c_void_p)]),
    ("phylo_destroy", None, [_c_void_p]),
    ("phylo_last_error", ctypes.c_char_p, [_c_void_p]),
    ("phylo_set_precision", _c_int, [_c_void_p, _c_int]),
    ("phylo_set_keep_partials", _c_int, [_c_void_p, _c_int]),
    ("phylo_set_tail_split", _c_int, [_c_void_p, _c_int]),
    ("phylo_set_tips", _c_int, [_c_void_p, _c_int, _c_int, _c_int, ctypes.POINTER(ctypes.c_uint32)]),
    ("phylo_set_site_weights", _c_int, [_c_void_p, _c_dbl_p]),
    ("phylo_set_topology", _c_int, [_c_void_p, _c_int, _c_int, _c_int_p, _c_int_p]),
    ("phylo_set_edge_lengths", _c_int, [_c_void_p, _c_dbl_p]),
    ("phylo_set_model", _c_int, [_c_void_p, _c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_int]),
    ("phylo_set_categories", _c_int, [_c_void_p, _c_int, _c_dbl_p, _c_dbl_p]),
    ("phylo_loglik", _c_int, [_c_void_p, _c_dbl_p, _c_dbl_p]),
    ("phylo_node_partials", _c_int, [_c_void_p, _c_int, _c_dbl_p, _c_int_p]),
    ("phylo_get_pmatrix", _c_int, [_c_void_p, _c_int, _c_dbl_p]),
    ("phylo_nni_score", _c_int, [_c_void_p, _c_int, _c_int, _c_dbl_p, _c_int_p, _c_int_p,
                                 ctypes.POINTER(ctypes.c_int), _c_dbl_p]),
    ("phylo_nni_ops", _c_int, [_c_void_p, ctypes.POINTER(ctypes.c_long)]),
    ("phylo_synchronize", _c_int, [_c_void_p]),
    ("phylo_batch_create", _c_int, [_c_int, _c_int, _c_int, ctypes.POINTER(_c_void_p)]),
    ("phylo_batch_last_error", ctypes.c_char_p, [_c_void_p]),
    ("phylo_batch_add", _c_int, [_c_void_p, _c_int, _c_int, _c_int, ctypes.POINTER(ctypes.c_uint32), _c_int, _c_int,
                                 _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                 _c_int, ctypes.POINTER(ctypes.c_int)]),
    ("phylo_batch_finalize", _c_int, [_c_void_p, _c_int]),
    ("phylo_batch_run", _c_int, [_c_void_p, _c_dbl_p, _c_dbl_p]),
    ("phylo_batch_nni_layout", _c_int, [_c_void_p, _c_int_p, _c_int_p, _c_int_p]),
    ("phylo_batch_stats", _c_int, [_c_void_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]),
    ("phylo_batch_set_tiling", _c_int, [_c_void_p, _c_int, _c_int]),
    ("phylo_batch_tiling", _c_int, [_c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)]),
    ("phylo_batch_destroy", None, [_c_void_p]),
    ("phylo_last_kernel_ms", _c_int, [_c_void_p, _c_dbl_p]),
    ("phylo_simulate", _c_int, [_c_void_p, _c_int, _c_dbl_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint8)]),
    ("phylo_loglik_noalloc", _c_int, [_c_void_p, _c_int, _c_int, _c_int, ctypes.POINTER(ctypes.c_uint32), _c_int,
                                      _c_int, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_int,
                                      _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    # include/phylo_learn.h
    ("phylo_sim_features", _c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p,
                                    _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.c_double, ctypes.c_uint64, _c_flt_p,
                                    ctypes.POINTER(ctypes.c_uint32)]),
    ("phylo_mlp_create", _c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_uint64,
                                  ctypes.POINTER(_c_void_p)]),
    ("phylo_mlp_destroy", None, [_c_void_p]),
    ("phylo_mlp_last_error", ctypes.c_char_p, [_c_void_p]),
    ("phylo_mlp_n_params", ctypes.c_long, [_c_void_p]),
    ("phylo_mlp_params", _c_int, [_c_void_p, _c_flt_p, _c_flt_p]),
    ("phylo_mlp_forward", _c_int, [_c_void_p, _c_int, _c_flt_p, _c_flt_p]),
    ("phylo_mlp_grad", _c_int, [_c_void_p, _c_int, _c_flt_p, _c_flt_p, _c_flt_p, _c_flt_p]),
    ("phylo_mlp_set_data", _c_int, [_c_void_p, _c_int, _c_flt_p, _c_flt_p]),
    ("phylo_mlp_train", _c_int, [_c_void_p, _c_int, _c_int, ctypes.c_float, ctypes.c_uint64, _c_flt_p]),
]

# (name, restype, argtypes) of every entry point in include/phylo_host.h,
# exported by libphylommand_host.so
HOST_LIB_PATH = os.path.join(HERE, "libphylommand_host.so")
HOST_SIGNATURES = [
    ("phylo_gamma_rates", _c_int, [ctypes.c_double, _c_int, _c_dbl_p]),
    ("phylo_gtr_gamma_lnl", ctypes.c_double, [_c_void_p, _c_int, _c_dbl_p, _c_int, _c_dbl_p]),
    ("phylo_nm_defaults", _c_int, [ctypes.c_void_p]),
    ("phylo_fit_gtr_gamma", _c_int, [_c_void_p, _c_int, _c_dbl_p, _c_int, _c_dbl_p, ctypes.c_void_p, _c_dbl_p,
                                     _c_dbl_p, ctypes.POINTER(ctypes.c_int), _c_dbl_p]),
]
# which library exports the entry points of each header
HEADER_LIBRARY = {"phylo_hip.h": LIB_PATH, "phylo_learn.h": LIB_PATH, "phylo_host.h": HOST_LIB_PATH}


class PhyloError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libphylo_hip error {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH):
    """Loads libphylo_hip.so and declares every exported symbol.  Raises if the
    library has not been built (python -m phylommand_amd.build)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m phylommand_amd.build`")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


_host_lib = None


class NMOptions(ctypes.Structure):
    """phylo_nm_options (include/phylo_host.h)."""
    _fields_ = [("max_evals", ctypes.c_int), ("ftol_abs", ctypes.c_double), ("xtol_rel", ctypes.c_double),
                ("initial_step", ctypes.c_double), ("adaptive", ctypes.c_int), ("max_restarts", ctypes.c_int),
                ("steps", ctypes.POINTER(ctypes.c_double)), ("restart_step", ctypes.c_double)]


def nm_options(**kw) -> NMOptions:
    """The library's Nelder-Mead defaults with the given fields replaced
    (steps: one simplex edge per parameter)."""
    o = NMOptions()
    load_host_library().phylo_nm_defaults(ctypes.byref(o))
    for k, v in kw.items():
        if v is None:
            continue
        if k == "steps":
            arr = np.ascontiguousarray(v, dtype=np.float64)
            o._steps_keep = arr
            o.steps = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        else:
            setattr(o, k, v)
    return o


def load_host_library():
    """Loads libphylommand_host.so (the C++ host layer; it links
    libphylo_hip.so, so contexts created through either are shared)."""
    global _host_lib
    if _host_lib is not None:
        return _host_lib
    load_library()
    if not os.path.exists(HOST_LIB_PATH):
        raise FileNotFoundError(f"{HOST_LIB_PATH} is missing: build it with `python -m phylommand_amd.build`")
    lib = ctypes.CDLL(HOST_LIB_PATH)
    for name, res, args in HOST_SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _host_lib = lib
    return lib


def gamma_rates(alpha: float, n_cat: int) -> np.ndarray:
    """Mean rates of n_cat equal-probability discrete-gamma categories."""
    r = np.zeros(n_cat)
    rc = load_host_library().phylo_gamma_rates(float(alpha), int(n_cat), _dp(r))
    if rc != PHYLO_OK:
        raise PhyloError(rc, "phylo_gamma_rates: invalid arguments")
    return r


def _dp(a):
    return a.ctypes.data_as(_c_dbl_p)


def _ip(a):
    return a.ctypes.data_as(_c_int_p)


class Context:
    """One libphylo_hip context (its own HIP stream and device buffers)."""

    def __init__(self, device: int = 0, precision: int = PREC_FP32):
        self.lib = load_library()
        h = _c_void_p()
        rc = self.lib.phylo_create(device, ctypes.byref(h))
        if rc != PHYLO_OK:
            raise PhyloError(rc, "phylo_create failed (no GPU visible?)")
        self.h = h
        # output cells of the per-evaluation calls, passed by reference (a
        # numpy buffer's ctypes pointer costs a few microseconds a call)
        self._tot = ctypes.c_double(0.0)
        self._ms = ctypes.c_double(0.0)
        self._check(self.lib.phylo_set_precision(self.h, precision))
        self.n_sites = 0
        self.n_nodes = 0

    def _check(self, rc):
        if rc != PHYLO_OK:
            raise PhyloError(rc, self.lib.phylo_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.phylo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_precision(self, precision):
        self._check(self.lib.phylo_set_precision(self.h, precision))

    def set_keep_partials(self, keep: bool):
        """Store every internal node's partial in loglik (default: only those
        the walk reads back; node_partials / nni_score evaluate again when
        they need the rest)."""
        self._check(self.lib.phylo_set_keep_partials(self.h, int(bool(keep))))

    def set_tail_split(self, enable: bool):
        """Walk the tail of a whole-tree DNA evaluation (site blocks past the
        last full round over the CUs) as subtree pieces plus a short second
        launch (default), or as whole-tree workgroups."""
        self._check(self.lib.phylo_set_tail_split(self.h, int(bool(enable))))

    def set_tips(self, masks: np.ndarray, n_states: int):
        masks = np.ascontiguousarray(masks, dtype=np.uint32)
        n_taxa, m = masks.shape
        self.n_sites = m
        self._check(self.lib.phylo_set_tips(self.h, n_taxa, m, n_states,
                                            masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))

    def set_site_weights(self, w):
        if w is None:
            self._check(self.lib.phylo_set_site_weights(self.h, None))
        else:
            w = np.ascontiguousarray(w, dtype=np.float64)
            self._check(self.lib.phylo_set_site_weights(self.h, _dp(w)))

    def set_topology(self, left, right, root: int):
        left = np.ascontiguousarray(left, dtype=np.int32)
        right = np.ascontiguousarray(right, dtype=np.int32)
        self.n_nodes = len(left)
        self._check(self.lib.phylo_set_topology(self.h, len(left), int(root), _ip(left), _ip(right)))

    def set_edge_lengths(self, eff):
        eff = np.ascontiguousarray(eff, dtype=np.float64)
        self._check(self.lib.phylo_set_edge_lengths(self.h, _dp(eff)))

    def set_model(self, Q, root_weights, pi=None, builder: int = P_EIGEN):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        rw = np.ascontiguousarray(root_weights, dtype=np.float64)
        pp = None if pi is None else _dp(np.ascontiguousarray(pi, dtype=np.float64))
        self._pi_keep = None if pi is None else np.ascontiguousarray(pi, dtype=np.float64)
        if pi is not None:
            pp = _dp(self._pi_keep)
        self._check(self.lib.phylo_set_model(self.h, Q.shape[0], _dp(Q), _dp(rw), pp, builder))

    def set_categories(self, rates, weights=None):
        rates = np.ascontiguousarray(rates, dtype=np.float64)
        if weights is None:
            weights = np.full(len(rates), 1.0 / len(rates))
        weights = np.ascontiguousarray(weights, dtype=np.float64)
        self._check(self.lib.phylo_set_categories(self.h, len(rates), _dp(rates), _dp(weights)))

    def loglik(self, per_site: bool = False):
        site = np.zeros(self.n_sites) if per_site else None
        self._check(self.lib.phylo_loglik(self.h, _dp(site) if per_site else None, ctypes.byref(self._tot)))
        return (self._tot.value, site) if per_site else self._tot.value

    def node_partials(self, node: int, n_cat: int, n_states: int):
        clv = np.zeros((self.n_sites, n_cat, n_states))
        sc = np.zeros(self.n_sites, np.int32)
        self._check(self.lib.phylo_node_partials(self.h, int(node), _dp(clv), _ip(sc)))
        return clv, sc

    def pmatrix(self, node: int, n_cat: int, n_states: int):
        P = np.zeros((n_cat, n_states, n_states))
        self._check(self.lib.phylo_get_pmatrix(self.h, int(node), _dp(P)))
        return P

    def nni_score(self, mode: int = 1):
        """Returns (base_lnl, [(node, side, lnl)]) for every NNI variant."""
        nv = ctypes.c_int(0)
        base = np.zeros(1)
        self._check(self.lib.phylo_nni_score(self.h, mode, 0, None, None, None, ctypes.byref(nv), _dp(base)))
        n = nv.value
        lnl = np.zeros(max(n, 1))
        nodes = np.zeros(max(n, 1), np.int32)
        sides = np.zeros(max(n, 1), np.int32)
        self._check(self.lib.phylo_nni_score(self.h, mode, n, _dp(lnl), _ip(nodes), _ip(sides), ctypes.byref(nv),
                                             _dp(base)))
        return float(base[0]), list(zip(nodes[:n].tolist(), sides[:n].tolist(), lnl[:n].tolist()))

    def loglik_noalloc(self, masks, left, right, root, eff, Q, root_weights, pi, cat_rates, cat_weights=None,
                       per_site: bool = False):
        """One evaluation through phylo_loglik_noalloc: no device allocation,
        the workspace is the library's fixed scratch array; the context gives
        only its device and stream."""
        masks = np.ascontiguousarray(masks, dtype=np.uint32)
        n_taxa, m = masks.shape
        left = np.ascontiguousarray(left, dtype=np.int32)
        right = np.ascontiguousarray(right, dtype=np.int32)
        eff = np.ascontiguousarray(eff, dtype=np.float64)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        rw = np.ascontiguousarray(root_weights, dtype=np.float64)
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        cr = np.ascontiguousarray(cat_rates, dtype=np.float64)
        cw = np.full(len(cr), 1.0 / len(cr)) if cat_weights is None else np.ascontiguousarray(cat_weights, np.float64)
        tot = np.zeros(1)
        site = np.zeros(m) if per_site else None
        self._check(self.lib.phylo_loglik_noalloc(
            self.h, n_taxa, m, Q.shape[0], masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(left),
            int(root), _ip(left), _ip(right), _dp(eff), _dp(Q), _dp(rw), _dp(pi), len(cr), _dp(cr), _dp(cw),
            _dp(site) if per_site else None, _dp(tot)))
        return (float(tot[0]), site) if per_site else float(tot[0])

    def nni_ops(self) -> int:
        n = ctypes.c_long(0)
        self._check(self.lib.phylo_nni_ops(self.h, ctypes.byref(n)))
        return n.value

    def synchronize(self):
        self._check(self.lib.phylo_synchronize(self.h))

    def last_kernel_ms(self) -> float:
        self._check(self.lib.phylo_last_kernel_ms(self.h, ctypes.byref(self._ms)))
        return self._ms.value

    def simulate(self, n_sites: int, root_freq, seed: int = 1) -> np.ndarray:
        """States [n_nodes, n_sites] drawn down the context's tree and model."""
        rf = np.ascontiguousarray(root_freq, dtype=np.float64)
        out = np.zeros((self.n_nodes, n_sites), np.uint8)
        self._check(self.lib.phylo_simulate(self.h, int(n_sites), _dp(rf), int(seed),
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def gtr_gamma_lnl(self, rel_len, n_cat: int, x) -> float:
        """lnL of the GTR+Gamma point x (see include/phylo_host.h)."""
        rel = np.ascontiguousarray(rel_len, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        return float(load_host_library().phylo_gtr_gamma_lnl(self.h, len(rel), _dp(rel), int(n_cat), _dp(x)))

    def fit_gtr_gamma(self, rel_len, n_cat: int, x0, opts: "NMOptions | None" = None):
        """Nelder-Mead GTR+Gamma fit from x0 (opts: nm_options(...), None for
        the defaults); returns (x, lnl, evals, seconds)."""
        rel = np.ascontiguousarray(rel_len, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x = np.zeros(10)
        lnl, secs = np.zeros(1), np.zeros(1)
        ev = ctypes.c_int(0)
        self._check(load_host_library().phylo_fit_gtr_gamma(self.h, len(rel), _dp(rel), int(n_cat), _dp(x0),
                                                            None if opts is None else ctypes.byref(opts), _dp(x),
                                                            _dp(lnl), ctypes.byref(ev), _dp(secs)))
        return x, float(lnl[0]), ev.value, float(secs[0])


class Batch:
    """Many independent DNA problems (and their NNI neighbourhoods) scored by
    persistent work-queue launches replayed from one captured HIP graph."""

    def __init__(self, n_cat: int = 4, precision: int = PREC_FP32, device: int = 0):
        self.lib = load_library()
        h = _c_void_p()
        rc = self.lib.phylo_batch_create(device, n_cat, precision, ctypes.byref(h))
        if rc != PHYLO_OK:
            raise PhyloError(rc, "phylo_batch_create failed")
        self.h = h
        self.n_cat = n_cat
        self.n = 0
        self.n_var = 0

    def _check(self, rc):
        if rc != PHYLO_OK:
            raise PhyloError(rc, self.lib.phylo_batch_last_error(self.h).decode())

    def add(self, masks, left, right, root, eff, Q, pi, cat_rates, cat_w=None, nni_mode=-1, root_w=None):
        masks = np.ascontiguousarray(masks, dtype=np.uint32)
        left = np.ascontiguousarray(left, dtype=np.int32)
        right = np.ascontiguousarray(right, dtype=np.int32)
        eff = np.ascontiguousarray(eff, dtype=np.float64)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        pi = np.ascontiguousarray(pi, dtype=np.float64)
        rw = pi if root_w is None else np.ascontiguousarray(root_w, dtype=np.float64)
        cr = np.ascontiguousarray(cat_rates, dtype=np.float64)
        cw = np.full(len(cr), 1.0 / len(cr)) if cat_w is None else np.ascontiguousarray(cat_w, dtype=np.float64)
        idx = ctypes.c_int(-1)
        self._check(self.lib.phylo_batch_add(self.h, masks.shape[0], masks.shape[1], Q.shape[0],
                                             masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(left),
                                             int(root), _ip(left), _ip(right), _dp(eff), _dp(Q), _dp(rw), _dp(pi),
                                             _dp(cr), _dp(cw), nni_mode, ctypes.byref(idx)))
        self.n += 1
        return idx.value

    def finalize(self, autotune: bool = True):
        self._check(self.lib.phylo_batch_finalize(self.h, int(autotune)))
        _, _, _, _, self.n_var = self.stats()

    def run(self):
        tot = np.zeros(max(self.n, 1))
        nni = np.zeros(max(self.n_var, 1))
        self._check(self.lib.phylo_batch_run(self.h, _dp(tot), _dp(nni)))
        return tot[: self.n], nni[: self.n_var]

    def nni_layout(self):
        off = np.zeros(self.n + 1, np.int32)
        node = np.zeros(max(self.n_var, 1), np.int32)
        side = np.zeros(max(self.n_var, 1), np.int32)
        self._check(self.lib.phylo_batch_nni_layout(self.h, _ip(off), _ip(node), _ip(side)))
        return off, node[: self.n_var], side[: self.n_var]

    def stats(self):
        u1, u2, ms = np.zeros(1), np.zeros(1), np.zeros(1)
        spi, nv = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.phylo_batch_stats(self.h, _dp(u1), _dp(u2), _dp(ms), ctypes.byref(spi), ctypes.byref(nv)))
        return float(u1[0]), float(u2[0]), float(ms[0]), spi.value, nv.value

    def set_tiling(self, sites_per_item: int, category_waves: int):
        """Tiling used by finalize(autotune=False)."""
        self._check(self.lib.phylo_batch_set_tiling(self.h, int(sites_per_item), int(category_waves)))

    def tiling(self):
        """(sites per work item, category waves per workgroup) chosen by finalize."""
        return self.launch_shape()[:2]

    def launch_shape(self):
        """(sites per work item, category waves per workgroup, resident workgroups per CU)."""
        spi, cw, wpc = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.phylo_batch_tiling(self.h, ctypes.byref(spi), ctypes.byref(cw), ctypes.byref(wpc)))
        return spi.value, cw.value, wpc.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.phylo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _fp(a):
    return a.ctypes.data_as(_c_flt_p)


def sim_features(n_taxa, n_sites, left, right, roots, blen, eig, pi, cat_rates, gap_frac=0.0, seed=1,
                 want_masks=False, device=0):
    """Simulates len(roots) alignments on the GPU and returns their features
    [n_align, 256] (and the tip masks [n_align, n_taxa, n_sites])."""
    lib = load_library()
    left = np.ascontiguousarray(left, dtype=np.int32)
    right = np.ascontiguousarray(right, dtype=np.int32)
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    blen = np.ascontiguousarray(blen, dtype=np.float64)
    eig = np.ascontiguousarray(eig, dtype=np.float64)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    cat_rates = np.ascontiguousarray(cat_rates, dtype=np.float64)
    n_align = len(roots)
    n_cat = cat_rates.shape[1]
    feat = np.zeros((n_align, 256), np.float32)
    masks = np.zeros((n_align, n_taxa, n_sites), np.uint32) if want_masks else None
    rc = lib.phylo_sim_features(device, n_align, n_taxa, n_sites, n_cat, _ip(left), _ip(right), _ip(roots), _dp(blen),
                                _dp(eig), _dp(pi), _dp(cat_rates), float(gap_frac), int(seed), _fp(feat),
                                masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if want_masks else None)
    if rc != PHYLO_OK:
        raise PhyloError(rc, "phylo_sim_features failed")
    return (feat, masks) if want_masks else feat


class MLP:
    """n_in -> n_hidden -> n_hidden -> n_out ReLU MLP trained on the GPU."""

    def __init__(self, n_in=256, n_hidden=512, n_out=12, max_batch=256, seed=1, device=0):
        self.lib = load_library()
        h = _c_void_p()
        rc = self.lib.phylo_mlp_create(device, n_in, n_hidden, n_out, max_batch, seed, ctypes.byref(h))
        if rc != PHYLO_OK:
            raise PhyloError(rc, "phylo_mlp_create failed")
        self.h = h
        self.n_in, self.n_hidden, self.n_out, self.max_batch = n_in, n_hidden, n_out, max_batch
        self.n_params = self.lib.phylo_mlp_n_params(self.h)

    def _check(self, rc):
        if rc != PHYLO_OK:
            raise PhyloError(rc, self.lib.phylo_mlp_last_error(self.h).decode())

    def params(self):
        w = np.zeros(self.n_params, np.float32)
        self._check(self.lib.phylo_mlp_params(self.h, _fp(w), None))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.n_params
        self._check(self.lib.phylo_mlp_params(self.h, None, _fp(w)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        self._check(self.lib.phylo_mlp_forward(self.h, x.shape[0], _fp(x), _fp(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self._check(self.lib.phylo_mlp_grad(self.h, x.shape[0], _fp(x), _fp(t), _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x, t, epochs=20, batch=256, lr=1e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._check(self.lib.phylo_mlp_set_data(self.h, x.shape[0], _fp(x), _fp(t)))
        losses = np.zeros(epochs, np.float32)
        self._check(self.lib.phylo_mlp_train(self.h, epochs, batch, float(lr), int(seed), _fp(losses)))
        return losses

    def close(self):
        if getattr(self, "h", None):
            self.lib.phylo_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
