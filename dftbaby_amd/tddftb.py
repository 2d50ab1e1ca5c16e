"""GPU twin of DFTBaby's f2py extension module ``DFTB.extensions.tddftb``.

``tddftb.tddftb.trans_charges``, ``tddftb.tddftb.apbv`` and
``tddftb.tddftb.ambv`` take and return numpy arrays with the signatures and
layouts of the Fortran routines they replace (``DFTB/extensions/tddftb.f90``):

* ``trans_charges(orbs_occ, orbs_virt, S, orbsPerAtom) -> (q_oo, q_ov, q_vv)``
  with q_oo (Nat, nocc, nocc), q_ov (Nat, nocc, nvirt), q_vv (Nat, nvirt, nvirt)
  (tddftb.f90:12-112, called at LR_TDDFTB.py:303-306), through
  ``dftb_trans_charges``;
* ``apbv(g, g_lr, q_oo, q_ov, q_vv, omega, vs) -> us`` = (A+B) vs and
  ``ambv(...)`` = (A-B) vs for vs (nocc, nvirt, nvec) (tddftb.f90:114-287,
  called at Solver.py:678, 689), through ``dftb_casida_matvec`` on the
  transition-charge exchange in FP64.

A caller switches by importing this module instead of the f2py one, e.g. for
the ``SOLVER_IMPLEMENTATION`` branch of Solver.py:22, 673-692.  The operator
built from the charges (gamma_lr folded into q_ov, q_vv once) is kept for the
next call with the same argument objects, so a Davidson loop that passes the
same arrays every iteration builds it once.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .engine import CasidaOperatorGPU, _device, _p, _stream


def _trans_charges(orbs_occ, orbs_virt, S, orbsPerAtom):
    dev = _device()
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    Co, Cv, St = t(orbs_occ), t(orbs_virt), t(S)
    norb, no = Co.shape
    nv = Cv.shape[1]
    opa = np.ascontiguousarray(orbsPerAtom, dtype=np.int32)
    nat = len(opa)
    if St.shape != (norb, norb) or Cv.shape[0] != norb or int(opa.sum()) != norb:
        raise ValueError("trans_charges: inconsistent shapes (orbs_occ %s, orbs_virt %s, S %s, sum(orbsPerAtom) %d)"
                         % (tuple(Co.shape), tuple(Cv.shape), tuple(St.shape), int(opa.sum())))
    q = [torch.empty(s, dtype=torch.float64, device=dev) for s in ((nat, no, no), (nat, no, nv), (nat, nv, nv))]
    nb = _lib.load().dftb_trans_charges_workspace(norb, no, nv, nat)
    ws = torch.empty(max(int(nb), 1), dtype=torch.uint8, device=dev)
    _lib.call("dftb_trans_charges", _p(Co), _p(Cv), _p(St), opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
              nat, norb, no, nv, _p(q[0]), _p(q[1]), _p(q[2]), _p(ws), int(nb), _stream())
    return tuple(x.cpu().numpy() for x in q)


_CACHE = {"args": None, "op": None}


def _operator(g, g_lr, q_oo, q_ov, q_vv, omega):
    args = (g, g_lr, q_oo, q_ov, q_vv, omega)
    cached = _CACHE["args"]
    if cached is not None and all(a is b for a, b in zip(args, cached)):
        return _CACHE["op"]
    op = CasidaOperatorGPU.from_charges(q_oo, q_vv, q_ov, omega, g, g_lr, lc=1)
    _CACHE["args"], _CACHE["op"] = args, op
    return op


def _product(which, g, g_lr, q_oo, q_ov, q_vv, omega, vs):
    vs = np.asarray(vs, dtype=np.float64)
    nat, no, nv = np.shape(q_ov)
    if vs.ndim != 3 or vs.shape[:2] != (no, nv):
        raise ValueError("vs must have shape (nocc, nvirt, nvec) = (%d, %d, n), got %s" % (no, nv, vs.shape))
    op = _operator(g, g_lr, q_oo, q_ov, q_vv, omega)
    V = op.to_padded(np.moveaxis(vs, 2, 0))
    P, M = op.matvec(V, 0)
    out = op.from_padded(P if which == "p" else M).cpu().numpy()
    return np.asfortranarray(np.moveaxis(out, 0, 2))


class _Module:
    """Namespace with the routine names of the f2py module (tddftb.tddftb.*)."""

    @staticmethod
    def trans_charges(orbs_occ, orbs_virt, S, orbsPerAtom):
        return _trans_charges(orbs_occ, orbs_virt, S, orbsPerAtom)

    @staticmethod
    def apbv(g, g_lr, q_oo, q_ov, q_vv, omega, vs):
        return _product("p", g, g_lr, q_oo, q_ov, q_vv, omega, vs)

    @staticmethod
    def ambv(g, g_lr, q_oo, q_ov, q_vv, omega, vs):
        return _product("m", g, g_lr, q_oo, q_ov, q_vv, omega, vs)


tddftb = _Module()
