"""The C-ABI library loads and exports every function include/muscade_hip.h declares (no GPU needed)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "muscade_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(os.path.join(ROOT, "muscade.jl_amd", "muscade_amd", "libmuscade_hip.so"))
    names = declared()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from muscade_amd import _lib
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_strerror_without_gpu():
    from muscade_amd import _lib
    assert _lib.lib.mh_strerror(_lib.MH_ERR_NAN).decode().startswith("element returned NaN")
    assert _lib.lib.mh_strerror(0).decode() == "ok"
