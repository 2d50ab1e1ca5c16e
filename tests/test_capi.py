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


def test_gemm_rejects_invalid_transposed_copy_and_hadamard_arguments():
    """Argument checks of dftb_gemm that return before any device work: a
    Hadamard factor needs hmod > 0; the transposed copy needs the BF16x3 input,
    the bf16-pair output and M a multiple of 8."""
    from dftbaby_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    fake = 1 << 20                                   # never dereferenced: every case fails its checks first

    def args(**kw):
        g = _lib.GemmArgs()
        g.A = g.A2 = g.B = g.B2 = g.C = g.C2 = fake
        g.lda = g.ldb = g.K = 64
        g.ldc, g.M, g.N, g.batch, g.ksplit, g.kchunk = 64, 64, 64, 1, 1, 64
        g.b_trans, g.alpha, g.hmod = 1, 1.0, 1
        g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_BF16PAIR
        for k, v in kw.items():
            setattr(g, k, v)
        return g

    cases = [args(had=fake, hmod=0, ldh=8),                                    # Hadamard modulus 0
             args(Ct=fake, Ct2=fake, ldct=64, sCt=4096, out_type=_lib.GEMM_OUT_F64),   # copy needs pair output
             args(Ct=fake, Ct2=fake, ldct=64, sCt=4096, M=60),                 # M not a multiple of 8
             args(Ct=fake, Ct2=fake, ldct=64, sCt=4096, in_type=_lib.GEMM_IN_F64, b_trans=1)]  # FP64 kernel
    for g in cases:
        assert lib.dftb_gemm(ctypes.byref(g), None) == 1


def test_noalloc_trans_charges_rejects_sizes_beyond_its_scratch():
    """dftb_trans_charges_noalloc checks its arguments and the fixed scratch size
    before any device work: a norb * (nocc + nvirt) beyond the 64 MiB array and
    invalid sizes return a status, not a launch."""
    from dftbaby_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    lib = _lib.load()
    fake = ctypes.c_void_p(1 << 20)             # never dereferenced
    opa = (ctypes.c_int * 1)(4)
    # norb = 3000 with the full space (nocc + nvirt = 3000): 72 MB of S C products
    assert lib.dftb_trans_charges_noalloc(fake, fake, fake, opa, 750, 3000, 1500, 1500, fake, fake, fake,
                                          None) == _lib.DFTB_ERR_WORKSPACE
    assert lib.dftb_trans_charges_noalloc(fake, fake, fake, opa, 0, 4, 2, 2, fake, fake, fake, None) == \
        _lib.DFTB_ERR_ARG
    assert lib.dftb_trans_charges_noalloc(fake, fake, fake, opa, 1, 4, -1, 2, fake, fake, fake, None) == \
        _lib.DFTB_ERR_ARG
