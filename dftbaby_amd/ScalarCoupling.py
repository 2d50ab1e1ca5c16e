"""Overlaps of TD-DFTB wavefunctions at two geometries, on the GPU.

Drop-in for ``DFTB.ScalarCoupling.ScalarCoupling`` (``DFTB/ScalarCoupling.py:138-265``):
``ci_overlap(atomlist1, orbs1, C1, atomlist2, orbs2, C2, Nst, threshold)``
returns Sci[I][J] = <Psi_I(1)|Psi_J(2)> for the ground state (0) and the lowest
Nst - 1 excited states, with the reference's definitions: the AO overlap between
the geometries (``overlap_AB``, here the ``dftb_slako_overlap12`` kernel), the MO
overlap Smo = orbs1^T S_AO orbs2, and for the singlet configuration state
functions |Psi_ia> the determinants of the occupied block of Smo with the row
of occupied orbital i and / or the column of j replaced by virtual ones;
excitations whose largest coefficient over the states is below ``threshold`` are
skipped (masked here).

The reference evaluates three n x n determinants per pair of excitations
(n = Nelec/2).  With A = Smo[occ0, occ0], B = A^{-1}, the matrix determinant
lemma gives every replaced determinant from B:

  row i -> a:            det0 X[a,i],  X = (Smo[virt, occ0] B)[:, occ]
  column j -> b:         det0 Y[j,b],  Y = (B Smo[occ0, virt])[occ, :]
  both (corner s_ab):    det0 (X[a,i] Y[j,b] - B[occ_j, occ_i] K[a,b]),
                         K = Smo[virt, occ0] B Smo[occ0, virt] - Smo[virt, virt]

so that Sci[I][J] = det0^2 (2 x_I y_J - <C2_J, B_oo C1_I K>) with x_I = <C1_I, X^T>,
y_J = <C2_J, Y>; Sci[0][J] = sqrt(2) det0^2 y_J, Sci[I][0] = sqrt(2) det0^2 x_I and
Sci[0][0] = det0 (as the reference sets it).  All products are MFMA FP64 GEMMs;
A^{-1} and det(A) come from the ROCm solver library through torch.
"""
import numpy as np
import torch

from . import _lib
from .engine import _p, _stream
from .ExcGradients import _mm
from .groundstate import slako as SK


def overlap_AB(atomlist1, atomlist2, dev=None):
    """AO overlap between two geometries of one molecule (ScalarCoupling.py:19-131)."""
    dev = dev or torch.device("cuda")
    Z = [int(z) for z, _ in atomlist1]
    if [int(z) for z, _ in atomlist2] != Z:
        raise ValueError("both geometries must hold the same atoms in the same order")
    T = SK.tables()
    atom, ls, ms, _, types = SK.orbital_lists(Z, T)
    ti = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
    tf = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    p1 = tf([x for _, x in atomlist1])
    p2 = tf([x for _, x in atomlist2])
    knots, coefs, _ = T.device_arrays(dev)
    n = len(atom)
    S = torch.empty((n, n), dtype=torch.float64, device=dev)
    # the index arrays stay referenced until the launch is enqueued (a freed
    # temporary's block would be handed to the next one before the kernel reads it)
    oa, ol, om, ty = (ti(a) for a in (atom, ls, ms, types))
    _lib.call("dftb_slako_overlap12", _p(p1), _p(p2), _p(oa), _p(ol), _p(om), _p(ty), len(Z), n,
              _p(knots), _p(coefs), T.npair, T.npts, _p(S), _stream())
    return S


class ScalarCoupling:
    def __init__(self, tddftb):
        self.tddftb = tddftb

    def ci_overlap(self, atomlist1, orbs1, C1, atomlist2, orbs2, C2, Nst, threshold=1.0e-3):
        td = self.tddftb
        dev = td.op.device if getattr(td, "op", None) is not None else torch.device("cuda")
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        Sao = overlap_AB(atomlist1, atomlist2, dev)
        Smo = _mm(_mm(t(orbs1), Sao, ta=True), t(orbs2))
        nelec = td.dftb2.Nelec_val
        assert nelec % 2 == 0
        n = nelec // 2
        occ, virt = td.getActiveOrbitals()
        occ = torch.as_tensor(np.asarray(occ), device=dev)
        virt = torch.as_tensor(np.asarray(virt), device=dev)
        A = Smo[:n, :n].contiguous()
        det0 = float(torch.linalg.det(A).cpu())
        B = torch.linalg.inv(A)
        R = Smo[virt][:, :n].contiguous()            # rows of the virtual orbitals (geometry 1)
        Cc = Smo[:n][:, virt].contiguous()           # columns of the virtual orbitals (geometry 2)
        RB = _mm(R, B)
        X = RB[:, occ].contiguous()                  # [v][o]
        Y = _mm(B, Cc)[occ].contiguous()             # [o][v]
        K = _mm(RB, Cc) - Smo[virt][:, virt]
        Boo = B[occ][:, occ].contiguous()
        C1 = t(C1)[:Nst - 1]
        C2 = t(C2)[:Nst - 1]
        C1 = C1 * (C1.abs().amax(0) >= threshold)
        C2 = C2 * (C2.abs().amax(0) >= threshold)
        x = (C1 * X.T[None]).sum((1, 2))
        y = (C2 * Y[None]).sum((1, 2))
        k = C1.shape[0]
        cross = torch.empty((k, k), dtype=torch.float64, device=dev)
        for I in range(k):
            P = _mm(_mm(Boo, C1[I]), K)
            cross[I] = (C2 * P[None]).sum((1, 2))
        d2 = det0 * det0
        Sci = np.zeros((Nst, Nst))
        Sci[1:, 1:] = (d2 * (2.0 * torch.outer(x, y) - cross)).cpu().numpy()
        Sci[0, 1:] = (np.sqrt(2.0) * d2 * y).cpu().numpy()
        Sci[1:, 0] = (np.sqrt(2.0) * d2 * x).cpu().numpy()
        Sci[0, 0] = det0
        return Sci
