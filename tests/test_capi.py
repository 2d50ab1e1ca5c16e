"""The C-ABI library loads and exports every entry point include/dftbaby_amd.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dftbaby_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dftb_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "dftb_trans_charges" in names and "dftb_casida_matvec" in names and "dftb_gemm" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from dftbaby_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers every declaration with a signature
    assert set(declared_functions()) <= set(_lib.EXPORTED), set(declared_functions()) - set(_lib.EXPORTED)


def test_product_path_does_not_import_oracle():
    """Only tests, smoke() and bench.py's CPU baseline may touch oracle/."""
    pkg = os.path.join(ROOT, "dftbaby_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "from oracle" not in src, f
                assert "tddftb_oracle" not in src, f
