"""The C-ABI library loads (no GPU needed to load it) and exports every entry
point include/moons_amd.h declares; the Python binding covers all of them."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "moons_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(moons_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "moons_pcg_sf_solve" in syms and "moons_clean_div" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    from moons_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmoons_amd.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (moons_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_python_binding_covers_header():
    from moons_amd import _lib

    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_library_loads_through_binding():
    import moons_amd

    lib = moons_amd.load()
    assert lib.moons_last_error() is not None


def test_product_does_not_reference_the_oracle():
    """The product package never imports or links the checker."""
    pkg = os.path.join(ROOT, "moons_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".f90")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower(), f"{f} mentions the oracle"
