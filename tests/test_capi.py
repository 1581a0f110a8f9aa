"""The C-ABI library loads and exports every function include/fem25hip.h
declares (no device calls)."""
import ctypes as C
import os
import re

from fem25_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "fem25hip.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(f25_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("f25_mesh_create", "f25_asmble", "f25_apply_batch_dev", "f25_gradient",
              "f25_cfmat_dist", "f25_adjflam_dist", "f25_asmb_peff", "f25_graph_build"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_types_every_export():
    assert set(declared()) <= set(_lib.EXPORTS), set(declared()) - set(_lib.EXPORTS)
    _lib.load()
