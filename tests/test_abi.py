"""The C-ABI library loads and exports every symbol its headers declare.

CPU-only: no compute call is made (there is no GPU here); the library is
opened with dlopen and each declared `aeth_*` / `aether_*` function of
include/*.h must resolve. Also checks the Python binding's table matches.
"""
import ctypes
import glob
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:aeth|aether)_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "aeth_eval_cfg0" in names and "aeth_rng_fill" in names


def test_library_exports_every_declared_symbol():
    from aether_amd import _lib

    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from aether_amd import _lib

    declared = declared_symbols()
    bound = set(_lib.SIGNATURES)
    assert bound <= declared, bound - declared


def test_abi_version_and_error_string():
    from aether_amd import _lib

    L = _lib.lib()
    assert L.aeth_abi_version() == 1
    assert isinstance(L.aeth_last_error(), bytes)
