"""The C ABI library loads and exports every symbol include/*.h declares
(no compute calls: CPU only)."""
import ctypes
import glob
import os
import re

import phylommand_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(phylo_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"phylo_create", "phylo_loglik", "phylo_set_model", "phylo_node_partials"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = {n for n, _, _ in P.SIGNATURES}
    assert declared_functions() <= bound, declared_functions() - bound


def test_version_without_gpu():
    assert P.load_library().phylo_version() >= 1
