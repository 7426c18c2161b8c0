"""The C ABI library loads and exports every symbol include/*.h declares
(no compute calls: CPU only)."""
import ctypes
import glob
import os
import re

import phylommand_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header=None):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", header or "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(phylo_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"phylo_create", "phylo_loglik", "phylo_set_model", "phylo_node_partials"} <= names


def test_every_header_has_a_library():
    headers = {os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))}
    assert headers == set(P.HEADER_LIBRARY)


def test_library_exports_every_declared_symbol():
    for header, path in P.HEADER_LIBRARY.items():
        lib = ctypes.CDLL(path)
        missing = [n for n in sorted(declared_functions(header)) if not hasattr(lib, n)]
        assert not missing, (header, missing)


def test_binding_covers_every_declared_symbol():
    bound = {n for n, _, _ in P.SIGNATURES} | {n for n, _, _ in P.HOST_SIGNATURES}
    assert declared_functions() <= bound, declared_functions() - bound
    P.load_host_library()


def test_version_without_gpu():
    assert P.load_library().phylo_version() >= 1
