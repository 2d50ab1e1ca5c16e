"""ctypes binding of the C ABI declared in include/dftbaby_amd.h.

The shared library is built in-tree (``python -m dftbaby_amd.build``).  There
is no fallback: if it is missing every product entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdftbaby_amd.so")

GEMM_IN_F64, GEMM_IN_F32, GEMM_IN_BF16X3 = 0, 1, 2
GEMM_OUT_F64, GEMM_OUT_F32, GEMM_OUT_BF16PAIR = 0, 1, 2

DFTB_OK = 0
DFTB_ERR_ARG = 1
DFTB_ERR_LAUNCH = 2
DFTB_ERR_NOT_SPD = 3
DFTB_ERR_WORKSPACE = 4

_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_long
_d = ctypes.c_double
_sz = ctypes.c_size_t


class GemmArgs(ctypes.Structure):
    _fields_ = [("A", _vp), ("A2", _vp), ("B", _vp), ("B2", _vp), ("C", _vp), ("C2", _vp),
                ("lda", _l), ("ldb", _l), ("ldc", _l),
                ("sA", _l), ("sB", _l), ("sC", _l),
                ("cdiv", _l), ("cstr", _l), ("cs2", _l),
                ("rdiv", _l), ("rstr", _l), ("rs2", _l),
                ("had", _vp), ("hmod", _l), ("ldh", _l), ("hmap", _vp),
                ("M", _i), ("N", _i), ("K", _i),
                ("batch", _i),
                ("ksplit", _i), ("kchunk", _i),
                ("sCk", _l),
                ("alpha", _d), ("beta", _d),
                ("a_trans", _i), ("b_trans", _i),
                ("in_type", _i), ("out_type", _i),
                ("tile", _i),
                ("Ct", _vp), ("Ct2", _vp), ("ldct", _l), ("sCt", _l)]


class CasidaOperator(ctypes.Structure):
    _fields_ = [("nat", _i), ("nocc", _i), ("nvirt", _i), ("opad", _i), ("vpad", _i),
                ("lc", _i), ("precision", _i), ("atom_chunk", _i), ("tile_expand", _i), ("tile_contract", _i),
                ("atom_limit", _i),
                ("omega", _vp), ("gamma", _vp), ("q_c", _vp),
                ("q_oo", _vp * 2), ("q_ov", _vp * 2), ("qf_ov", _vp * 2), ("qf_vv", _vp * 2),
                ("exchange", _i), ("npad", _i),
                ("ao_v", _vp * 2), ("ao_o", _vp * 2), ("ao_cyvt", _vp * 2), ("ao_w", _vp * 2), ("gam_ao", _vp),
                ("amap", _vp)]


class EnsMolecule(ctypes.Structure):
    _fields_ = [("nat", _i), ("nocc", _i), ("nvirt", _i), ("opad", _i), ("vpad", _i), ("k", _i), ("cap", _i),
                ("nguess", _i), ("lc", _i), ("precision", _i),
                ("omega", _vp), ("gamma", _vp), ("q_c", _vp), ("q_oo", _vp), ("q_ov", _vp), ("qf_ov", _vp),
                ("qf_vv", _vp), ("q32", _vp * 4), ("guess", _vp),
                ("bs", _vp), ("bp", _vp), ("bm", _vp), ("Mp", _vp), ("Mm", _vp), ("work", _vp),
                ("w", _vp), ("C", _vp), ("XmY", _vp), ("XpY", _vp), ("resid", _vp), ("status", _vp)]


_SIGS = {
    "dftb_gemm": (_i, [ctypes.POINTER(GemmArgs), _vp]),
    "dftb_gemm_timing_enable": (_i, [_i]),
    "dftb_gemm_timing_read": (_i, [_i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long),
                                   ctypes.POINTER(ctypes.c_double)]),
    "dftb_trans_charges_workspace": (_sz, [_i, _i, _i, _i]),
    "dftb_trans_charges": (_i, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int), _i, _i, _i, _i,
                                _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_trans_charges_noalloc": (_i, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int), _i, _i, _i, _i,
                                        _vp, _vp, _vp, _vp]),
    "dftb_grad_pairs": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp, _i, _i,
                             _vp, _vp]),
    "dftb_grad_vrep": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp]),
    "dftb_slako_overlap12": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _i, _i, _vp, _vp]),
    "dftb_slako_matrices": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _i, _i, _d, _vp, _vp, _vp, _vp]),
    "dftb_trans_charges_engine": (_i, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int), _i, _i, _i, _i,
                                       _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_casida_matvec_workspace": (_sz, [ctypes.POINTER(CasidaOperator), _i]),
    "dftb_casida_matvec": (_i, [ctypes.POINTER(CasidaOperator), _vp, _i, _vp, _vp, _vp, _sz, _vp]),
    "dftb_hermitian_matvec_workspace": (_sz, [ctypes.POINTER(CasidaOperator), _i]),
    "dftb_hermitian_matvec": (_i, [ctypes.POINTER(CasidaOperator), _vp, _vp, _vp, _i, _vp, _vp, _sz, _vp]),
    "dftb_fold_gamma": (_i, [_vp, _i, _vp, _l, _l, _vp, _vp]),
    "dftb_convert_f32": (_i, [_vp, _l, _vp, _vp]),
    "dftb_convert_bf16pair": (_i, [_vp, _l, _vp, _vp, _vp]),
    "dftb_convert_bf16pair_atom_major": (_i, [_vp, _l, _l, _i, _l, _vp, _vp, _vp]),
    "dftb_norms_workspace": (_sz, [_i]),
    "dftb_residual": (_i, [_vp, _vp, _vp, _i, _l, _vp, _vp, _vp, _sz, _vp]),
    "dftb_gram_workspace": (_sz, [_i, _i, _l]),
    "dftb_gram": (_i, [_vp, _l, _i, _vp, _l, _i, _l, _vp, _l, _vp, _sz, _vp]),
    "dftb_precondition": (_i, [_vp, _vp, _vp, _i, _vp, _l, _vp, _vp]),
    "dftb_row_norms": (_i, [_vp, _i, _l, _vp, _vp, _sz, _vp]),
    "dftb_scale_rows": (_i, [_vp, _vp, _i, _l, _vp]),
    "dftb_scale2": (_i, [_vp, _vp, _vp, _i, _l, _vp, _vp]),
    "dftb_weighted_sumsq": (_i, [_vp, _vp, _i, _l, _vp, _vp]),
    "dftb_oia_normalize": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "dftb_orthonormalize_workspace": (_sz, [_i, _i, _l]),
    "dftb_orthonormalize": (_i, [_vp, _i, _vp, _i, _l, _d, _vp, _vp, _sz, _vp]),
    "dftb_subspace_workspace": (_sz, [_i, _i]),
    "dftb_subspace_nonhermitian": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_subspace_nonhermitian_rl": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_subspace_tb": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_sqrtm_workspace": (_sz, [_i]),
    "dftb_sqrtm": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_subspace_hermitian": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dftb_syevx_workspace": (_sz, [_i, _i]),
    "dftb_syevx": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "dftb_unit_vectors": (_i, [_vp, _i, _l, _vp, _vp]),
    "dftb_symmetrize": (_i, [_vp, _i, _i, _vp, _vp]),
    "dftb_coldot": (_i, [_vp, _l, _vp, _l, _i, _l, _vp, _vp]),
    "dftb_ensemble_workspace": (_sz, [ctypes.POINTER(EnsMolecule)]),
    "dftb_ensemble_solve": (_i, [_vp, _i, _d, _i, _vp]),
    "dftb_pair_features": (_i, [_vp, _vp, _vp, _vp, _vp, _l, _i, _vp, _vp]),
    "dftb_restricted_ab": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "dftb_mlp_bias_act": (_i, [_vp, _vp, _l, _i, _i, _vp, _vp]),
    "dftb_mlp_act_bwd": (_i, [_vp, _vp, _l, _vp]),
    "dftb_mlp_colsum": (_i, [_vp, _l, _i, _vp, _vp]),
    "dftb_mlp_mse": (_i, [_vp, _vp, _vp, _l, _i, _vp, _vp, _i, _vp]),
    "dftb_adam": (_i, [_vp, _vp, _vp, _vp, _l, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                       _i, ctypes.c_float, _vp]),
    "dftb_mlp_normalize": (_i, [_vp, _l, _i, _vp, _vp, _vp]),
}

EXPORTED = sorted(_SIGS)

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("dftbaby_amd HIP library not built (%s); run `python -m dftbaby_amd.build`"
                               % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class HipError(RuntimeError):
    pass


def check(rc, what):
    if rc == DFTB_OK:
        return
    names = {DFTB_ERR_ARG: "invalid argument", DFTB_ERR_LAUNCH: "HIP launch error",
             DFTB_ERR_NOT_SPD: "subspace matrix not positive definite",
             DFTB_ERR_WORKSPACE: "workspace too small"}
    raise HipError("%s failed: %s (status %d)" % (what, names.get(rc, "unknown"), rc))


def call(name, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc
