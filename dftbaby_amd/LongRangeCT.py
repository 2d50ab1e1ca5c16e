"""Long-range charge-transfer correction of the TD-DFTB diagonal (DFTB/LongRangeCT.py:13-93).

Neugebauer, Gritsenko and Baerends, J. Chem. Phys. 124, 214102 (2006): the
orbital-energy difference of an excitation i -> a whose orbital densities do not
overlap is raised by

    T_ia = exp(-(S_ia / Sc)^2) exp(-(Rc / R_ia)^2) (D_ia - 1/R_ia + (D_ia - 1/R_ia)^2 / (2 w_ia)),

clipped to [0, 1e10], with the density overlap S_ia = q_ii^T Omega q_aa (Gaussian
overlap Omega of the atomic fluctuation functions), the distance R_ia = |sum_A
(q^ii_A - q^aa_A) R_A| between the orbital centres (Mulliken approximation) and
D_ia = -eps_a.  The shifts enter A and B on the diagonal (Solver.py:46-49): the
Hermitian Davidson sees (w^2 + 2 w T) on its diagonal.

S_ia and the orbital centres are GEMMs of the diagonal transition charges on the
device; the element-wise formula runs once per geometry on the host, as in the
reference.
"""
import numpy as np
import torch

from .engine import gemm

HARTREE_TO_EV = 27.211396132   # DFTB/AtomicData.py


class LongRangeCTCorrection_Neugebauer:
    """LongRangeCT.py:13-58: S_ia, R_ia, D_ia of the active space of `tddftb`
    (its operator must be built: the diagonal charges q^ii_A, q^aa_A live there)."""

    def __init__(self, tddftb):
        self.tddftb = tddftb
        op = tddftb.op
        dev = op.device
        o, v, nat = op.nocc, op.nvirt, op.nat
        qii = op.diag_oo()                                   # [o][nat]
        if qii is None:
            raise RuntimeError("occupied-occupied charges are needed for the CT correction")
        qaa = op.qaa                                         # [v][nat]
        Om = torch.as_tensor(np.asarray(tddftb.dftb2.getGaussianOverlap(), dtype=np.float64), device=dev)
        # S_ia = sum_AB q^ii_A Omega_AB q^aa_B  (LongRangeCT.py:19-35)
        X = torch.empty((o, nat), dtype=torch.float64, device=dev)
        gemm(qii, Om, X, o, nat, nat, nat, nat, nat)
        S = torch.empty((o, v), dtype=torch.float64, device=dev)
        gemm(X, qaa, S, o, v, nat, nat, nat, v, b_trans=1)
        # orbital centres sum_A q^pp_A R_A  (LongRangeCT.py:37-49)
        pos = np.array([p for (_, p) in tddftb.dftb2.getGeometry()], dtype=np.float64)   # bohr
        R = torch.as_tensor(pos, device=dev)
        ci = torch.empty((o, 3), dtype=torch.float64, device=dev)
        ca = torch.empty((v, 3), dtype=torch.float64, device=dev)
        gemm(qii, R, ci, o, 3, nat, nat, 3, 3)
        gemm(qaa, R, ca, v, 3, nat, nat, 3, 3)
        self.Sia = S.cpu().numpy()
        ci, ca = ci.cpu().numpy(), ca.cpu().numpy()
        d = ci[:, None, :] - ca[None, :, :]
        self.Ria = np.sqrt((d ** 2).sum(axis=2))
        # asymptotic limit D_ia ~ -eps_a  (LongRangeCT.py:53-57)
        occ, virt = tddftb.getActiveOrbitals()
        orbe = tddftb.dftb2.getKSEnergies()
        self.Dia = -np.outer(np.ones(o), orbe[virt])

    def diagonal_shifts(self, omega, Sc=0.0001, Rc=0.1, verbose=1):
        """T_ia (LongRangeCT.py:59-93), clipped to [0, 1e10]; verbose > 0 prints the
        excitations with CT character (T_ia > 1e-3) up to 40 eV."""
        assert omega.shape == self.Ria.shape
        Sia, Ria, Dia = self.Sia, self.Ria, self.Dia
        with np.errstate(divide="ignore", invalid="ignore"):
            Tia = np.exp(-(Sia / Sc) ** 2) * np.exp(-(Rc / Ria) ** 2) \
                * (Dia - 1.0 / Ria + (Dia - 1.0 / Ria) ** 2 / (2 * omega))
        Tia = np.clip(Tia, 0.0, 1.0e10)
        if verbose > 0:
            occ, virt = self.tddftb.getActiveOrbitals()
            thresh = 1.0e-3
            indx = omega.argsort(axis=None, kind="mergesort")
            indi, inda = np.unravel_index(indx, omega.shape)
            print("Excitations i->a with Charge Transfer Character (Tia > %e)" % thresh)
            print("=" * 103)
            print("  i  ->   a         Sia              Ria              Tia              w_ia / eV        w_ia+Tia / eV ")
            print("-" * 103)
            for i, a in zip(indi, inda):
                if Tia[i, a] > thresh:
                    print(" %3d -> %3d    %13.8f    %13.8f    %13.8f    %13.8f    %13.8f"
                          % (occ[i], virt[a], Sia[i, a], Ria[i, a], Tia[i, a], omega[i, a] * HARTREE_TO_EV,
                             (omega[i, a] + Tia[i, a]) * HARTREE_TO_EV))
                if omega[i, a] > 40.0 / HARTREE_TO_EV:
                    break
            print("")
        return Tia


LongRangeCTCorrection = LongRangeCTCorrection_Neugebauer
