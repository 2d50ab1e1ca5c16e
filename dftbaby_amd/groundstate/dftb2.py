"""Host-side SCC ground-state producer (the input side of the excited-state path).

The excited-state engine consumes a converged closed-shell ground state through
the getters that ``DFTB/DFTB2.py:2405-2454`` exposes to ``LR_TDDFTB``
(``getKSCoefficients``, ``getKSEnergies``, ``getOccupation``,
``getOverlapMatrix``, ``getGamma``, ``getGamma_lr``, ``getGaussianOverlap``,
``getDipoleMatrix``, ``getGeometry``, ``getValorbs``, ``orbsPerAtom``,
``Nelec_val``).  This module produces such a state for H/C molecules with the
reference's "homegrown" parameters (converted by ``make_slako_data``):

* H0, S and the dipole matrix from the Slater-Koster tables, assembled as
  DFTBaby's default Fortran path does (``DFTB/extensions/slako.f90:77-394`` with
  the spline evaluation of ``splines.f90:25-94``), see ``groundstate/slako.py``;
  on the GPU by the ``dftb_slako_matrices`` kernel.
* gamma and gamma_lr for Gaussian charge fluctuations
  (``DFTB/GammaApproximation.py:8-82,122-151,210-243,388-405``), sigma = 1/(sqrt(pi) U).
* SCC iterations with Mulliken charges and, for ``long_range_correction=1``,
  the long-range exchange operator on the difference density
  (``DFTB/DFTB2.py:1240-1266``), accelerated by Pulay DIIS on the Fock matrix.
* T=0 occupations (``DFTB/DFTB2.py:2594-2610``).
* The Gaussian overlap Omega_AB for the Lambda2 diagnostic
  (``DFTB/DFTB2.py:1287-1317``).

* The repulsive energy from the tabulated pair potentials
  (``DFTB/DFTB2.py:2340-2359``, ``RepulsivePotential.py:80-101``); getEnergy
  returns the total energy E_tot = E_elec + E_rep as the reference's does.
"""
import math
import os

import numpy as np
import torch
from scipy.special import erf

from . import slako

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "slako_hc.npz")


class SCFNotConverged(Exception):
    pass


class _Parameters:
    """Hubbard U, valence orbitals, electron counts and free-atom populations of H and C."""

    def __init__(self):
        d = np.load(_DATA)
        self.U = {1: float(d["h_U"]), 6: float(d["c_U"])}
        # valence orbitals (n, l, m) per element, DFTB/DFTB2.py:240-245
        self.valorbs = {1: [(0, 0, 0)], 6: [(1, 0, 0), (1, 1, -1), (1, 1, 0), (1, 1, 1)]}
        self.ne_val = {1: 1, 6: 4}
        # reference (neutral atom) populations per orbital
        self.occ_ref = {1: [1.0], 6: [2.0, 2.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0]}


_TABLES = None


def _tables():
    global _TABLES
    if _TABLES is None:
        _TABLES = _Parameters()
    return _TABLES


class DFTB2:
    """Minimal DFTB2 / LC-DFTB ground state for H/C molecules (see module doc)."""

    def __init__(self, atomlist, long_range_correction=1, long_range_radius=3.03,
                 distance_cutoff=30.0, verbose=0, parameter_set="homegrown",
                 mulliken_dipoles=0, use_symmetry=0, device=None, **opts):
        self.long_range_correction = int(long_range_correction)
        self.long_range_radius = float(long_range_radius)
        self.distance_cutoff = distance_cutoff
        self.verbose = verbose
        self.parameter_set = parameter_set
        self.mulliken_dipoles = mulliken_dipoles
        self.use_symmetry = use_symmetry
        self.device = device
        self._T = _tables()
        self.hubbard_U = dict(self._T.U)
        self.valorbs = self._T.valorbs
        self.setGeometry(atomlist)

    # -- geometry dependent set-up ------------------------------------------------
    def setGeometry(self, atomlist, keep_charge=False, **opts):
        """New nuclear positions.  keep_charge=True (as PES.py:62,86 passes it)
        keeps the converged density of the previous geometry of the same molecule
        as the SCC's starting point (the reference keeps its partial charges,
        DFTB2.setGeometry); otherwise the SCC starts from the neutral reference
        density."""
        for Z, _ in atomlist:
            if Z not in (1, 6):
                raise ValueError("only H and C are parametrised in this producer")
        same = getattr(self, "Z", None) is not None and [int(Z) for Z, _ in atomlist] == list(self.Z)
        self._P_start = self.P if (keep_charge and same and getattr(self, "P", None) is not None) else None
        self.atomlist = [(int(Z), list(map(float, x))) for Z, x in atomlist]
        self.Z = np.array([Z for Z, _ in self.atomlist])
        self.pos = np.array([x for _, x in self.atomlist], dtype=np.float64)
        self.orbsPerAtom = np.array([len(self.valorbs[Z]) for Z in self.Z], dtype=np.int32)
        self.dim = int(self.orbsPerAtom.sum())
        self.Nelec_val = int(sum(self._T.ne_val[Z] for Z in self.Z))
        self.orb_atom = np.repeat(np.arange(len(self.Z)), self.orbsPerAtom)
        for a in ("orbs", "orbe", "f", "S", "H0", "P", "_D"):
            if hasattr(self, a):
                delattr(self, a)

    def _constructH0andS(self):
        S, H0, _ = slako.matrices_host(self.Z, self.pos, self.distance_cutoff, dipoles=False)
        return S, H0

    def _gamma_matrices(self):
        U = np.array([self.hubbard_U[Z] for Z in self.Z])
        sig = 1.0 / (np.sqrt(np.pi) * U)
        r = np.linalg.norm(self.pos[:, None, :] - self.pos[None, :, :], axis=-1)
        s2 = sig[:, None] ** 2 + sig[None, :] ** 2
        off = ~np.eye(len(U), dtype=bool)
        g = np.zeros_like(r)
        C = 1.0 / np.sqrt(2.0 * s2)
        g[off] = erf(C[off] * r[off]) / r[off]
        g[~off] = 1.0 / (np.sqrt(np.pi) * sig)
        glr = np.zeros_like(r)
        if self.long_range_correction == 1:
            Rlr = self.long_range_radius
            Clr = 1.0 / np.sqrt(2.0 * (s2 + 0.5 * Rlr ** 2))
            glr[off] = erf(Clr[off] * r[off]) / r[off]
            glr[~off] = 1.0 / np.sqrt(np.pi * (sig ** 2 + 0.25 * Rlr ** 2))
        return g, glr

    def _construct_gaussian_overlap(self):
        U = np.array([self.hubbard_U[Z] for Z in self.Z])
        sig = 1.329 / math.sqrt(8.0 * math.log(2.0)) / U
        sAB = sig[:, None] ** 2 + sig[None, :] ** 2
        r2 = np.sum((self.pos[:, None, :] - self.pos[None, :, :]) ** 2, axis=-1)
        return 1.0 / (2.0 * np.pi * sAB) ** 1.5 * np.exp(-0.5 * r2 / sAB)

    # -- SCC --------------------------------------------------------------------
    def getEnergy(self, scf_conv=1e-10, maxiter=300, diis_size=8, mixing=0.3, **opts):
        dev = self.device
        if dev is None:
            dev = "cuda" if (torch.cuda.is_available() and self.dim > 400) else "cpu"
        self._dev = dev
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)
        if str(dev).startswith("cuda"):
            St, H0t, _ = slako.matrices_device(self.Z, self.pos, dev, self.distance_cutoff, dipoles=False)
            self.S, self.H0 = St.cpu().numpy(), H0t.cpu().numpy()
        else:
            self.S, self.H0 = self._constructH0andS()
            St, H0t = t(self.S), t(self.H0)
        self.gamma, self.gamma_lr = self._gamma_matrices()
        self.gaussOmega = self._construct_gaussian_overlap()
        oa = torch.as_tensor(self.orb_atom, device=dev)
        nat = len(self.Z)
        G = t(self.gamma)
        Glr_ao = t(self.gamma_lr)[oa][:, oa]
        q0 = t([self._T.ne_val[Z] for Z in self.Z])
        P0 = torch.diag(t(np.concatenate([self._T.occ_ref[Z] for Z in self.Z])))
        L = torch.linalg.cholesky(St)
        Linv = torch.linalg.solve_triangular(L, torch.eye(self.dim, dtype=torch.float64, device=dev), upper=False)
        nocc = self.Nelec_val // 2
        if self.Nelec_val % 2:
            raise ValueError("closed-shell ground state needs an even electron count")
        P = t(self._P_start) if getattr(self, "_P_start", None) is not None else P0.clone()
        fock_hist, err_hist = [], []
        converged = False
        for it in range(maxiter):
            dq = torch.zeros(nat, dtype=torch.float64, device=dev).index_add_(0, oa, (P * St).sum(1)) - q0
            V = G @ dq
            F = H0t + 0.5 * St * (V[oa][:, None] + V[oa][None, :])
            if self.long_range_correction == 1:
                dP = P - P0
                SdP = St @ dP
                dPS = dP @ St
                Hx = (Glr_ao * SdP) @ St + Glr_ao * (SdP @ St) + St @ (dP * Glr_ao) @ St + St @ (dPS * Glr_ao)
                F = F - Hx / 8.0
            err = F @ P @ St - St @ P @ F
            errn = float(err.abs().max())
            fock_hist.append(F)
            err_hist.append(Linv @ err @ Linv.T)
            if len(fock_hist) > diis_size:
                fock_hist.pop(0)
                err_hist.pop(0)
            if it >= 2 and len(fock_hist) >= 3:
                m = len(fock_hist)
                B = torch.empty((m + 1, m + 1), dtype=torch.float64, device=dev)
                for i in range(m):
                    for j in range(i, m):
                        B[i, j] = B[j, i] = (err_hist[i] * err_hist[j]).sum()
                B[m, :] = -1.0
                B[:, m] = -1.0
                B[m, m] = 0.0
                rhs = torch.zeros(m + 1, dtype=torch.float64, device=dev)
                rhs[m] = -1.0
                try:
                    c = torch.linalg.solve(B, rhs)[:m]
                    Fx = sum(ci * Fi for ci, Fi in zip(c, fock_hist))
                except RuntimeError:
                    Fx = F
            else:
                Fx = F
            Fp = Linv @ Fx @ Linv.T
            e, Cp = torch.linalg.eigh(0.5 * (Fp + Fp.T))
            C = torch.linalg.solve_triangular(L.T, Cp, upper=True)
            Pn = 2.0 * C[:, :nocc] @ C[:, :nocc].T
            if it < 2:
                Pn = mixing * Pn + (1.0 - mixing) * P if it > 0 else Pn
            dPmax = float((Pn - P).abs().max())
            P = Pn
            if self.verbose > 1:
                print("SCC it %3d  commutator %.3e  dP %.3e" % (it, errn, dPmax))
            if errn < scf_conv ** 0.5 * 1e-2 and dPmax < scf_conv ** 0.5 * 1e-2:
                converged = True
                break
        if not converged:
            raise SCFNotConverged("SCC did not converge in %d iterations (commutator %.3e)" % (maxiter, errn))
        self.orbe = e.cpu().numpy()
        self.orbs = C.cpu().numpy()
        self.P = P.cpu().numpy()
        self.dq = dq.cpu().numpy()
        self.scc_iterations = it + 1
        _, self.f = fermi_occupation_T0(self.orbe, self.Nelec_val)
        # electronic energy (DFTB2.py:2241-2254): band structure + second order +
        # long-range exchange of the difference density (DFTB2.py:1267-1285); no
        # repulsive energy
        self.E_bs = float((P * H0t).sum())
        self.E_coulomb = float(0.5 * dq @ G @ dq)
        E = self.E_bs + self.E_coulomb
        self.E_HF_x = 0.0
        if self.long_range_correction == 1:
            dP = P - P0
            self.E_HF_x = float(-((St @ dP @ St) * dP * Glr_ao).sum() - ((St @ dP) * (dP @ St) * Glr_ao).sum()) / 8.0
            E += self.E_HF_x
        self.en_elec = E
        self.E_elec = E
        # nuclear repulsion (DFTB2.py:2340-2359) and the total energy (DFTB2.py:2274-2275)
        self.E_rep = slako.repulsive_energy(self.Z, self.pos)
        self.E_nuc = self.E_rep
        self.E_tot = E + self.E_rep
        return self.E_tot

    # -- getters used by the excited-state path (DFTB/DFTB2.py:2405-2454) --------
    def getKSCoefficients(self):
        return self.orbs

    def getKSEnergies(self):
        return self.orbe

    def getOccupation(self):
        return self.f

    def getOverlapMatrix(self):
        if not hasattr(self, "S"):
            self.S, self.H0 = self._constructH0andS()
        return self.S

    def getGamma(self):
        return self.gamma

    def getGamma_lr(self):
        return self.gamma_lr

    def getGaussianOverlap(self):
        return self.gaussOmega

    def getDipoleMatrix(self):
        """[norb][norb][3] dipole matrix from the Slater-Koster dipole tables
        (DFTB2.py:2446-2452, slako.f90:238-394)."""
        if getattr(self, "_D", None) is None:
            dev = getattr(self, "_dev", "cpu")
            if str(dev).startswith("cuda"):
                self._D = np.ascontiguousarray(np.moveaxis(self.getDipoleMatrixDevice(dev).cpu().numpy(), 0, 2))
            else:
                self._D = slako.matrices_host(self.Z, self.pos, self.distance_cutoff)[2]
        return self._D

    def getDipoleMatrixDevice(self, dev):
        """The dipole matrix as three device planes [3][norb][norb] (x, y, z),
        formed by the dftb_slako_matrices kernel."""
        return slako.matrices_device(self.Z, self.pos, dev, self.distance_cutoff)[2]

    def getGeometry(self):
        return self.atomlist

    def getValorbs(self):
        return self.valorbs

    def getFrontierOrbitals(self):
        HOMO = np.where(self.f > 0.0)[0][-1]
        LUMO = np.where(self.f == 0.0)[0][0]
        return HOMO, LUMO

    def getHOMO_LUMO_gap(self):
        H, L = self.getFrontierOrbitals()
        return self.orbe[L] - self.orbe[H]


def fermi_occupation_T0(orbe, Nelec_paired, Nelec_unpaired=0):
    """T=0 aufbau occupations, DFTB/DFTB2.py:2594-2610."""
    f = np.zeros(len(orbe))
    for a in np.argsort(orbe, kind="stable"):
        if Nelec_paired > 0.0:
            f[a] = min(2.0, Nelec_paired)
            Nelec_paired -= 2.0
        elif Nelec_unpaired > 0.0:
            f[a] = min(1.0, Nelec_unpaired)
            Nelec_unpaired -= 1.0
    return 0.0, f
