"""Potential energy surfaces for dynamics and optimisation (``DFTB/PES.py`` twin).

``PotentialEnergySurfaces(atomlist, Nst=2, **options)`` keeps one LR_TDDFTB
instance and evaluates, for nuclear coordinates x (a flat vector in bohr):

* ``getEnergies(x)``: total energies of S0 .. S_{Nst-1} (``PES.py:56-71``);
* ``getEnergy_S0(x)`` (``:72-80``);
* ``getEnergiesAndGradient(x, gradient_state)`` (``:81-124``): the energies and
  the total gradient gradVrep + gradE0 + gradExc of one state, from
  ``ExcGradients.Gradients`` on the GPU;
* ``getEnergyAndGradient_S0(x, has_scf=False)`` (``:125-167``);
* ``getScalarCouplings(threshold)`` (``:175-253``): the CI-wavefunction overlap
  with the previous step (``ScalarCoupling.ci_overlap`` on the GPU), the phase
  alignment of the states, and the antisymmetrised coupling matrix;
* ``getTransitionDipoles()`` (``:266-296``), ``saveLastCalculation``,
  ``alignCoefficients``, ``resetXY``, ``resetSCF``;
* ``getEnergiesBatch(xs)``: no counterpart; the energies at many geometries with
  all excited-state problems in one launch of the persistent ensemble solver.

Between calls the excited-state vectors of the previous geometry seed the
Davidson (LR_TDDFTB's instance-held warm start, ``LR_TDDFTB.py:29-32``).  The
options split into those of the ground state (DFTB2), those of
``LR_TDDFTB.getEnergies`` and the SCC options, as the reference's option parser
does; ``nstates`` is Nst - 1.  QM/MM, cavities and implicit solvent are not
part of this build.
"""
import inspect

import numpy as np

from .ExcGradients import Gradients
from .LR_TDDFTB import LR_TDDFTB
from .ScalarCoupling import ScalarCoupling

_SCF_OPTIONS = ("scf_conv", "maxiter", "diis_size", "mixing")


def _vector2atomlist(x, atomlist):
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    return [(Z, list(xi)) for (Z, _), xi in zip(atomlist, x)]


class PotentialEnergySurfaces:
    def __init__(self, atomlist, Nst=2, **kwds):
        te = set(inspect.signature(LR_TDDFTB.getEnergies).parameters) - {"self", "scf_opts"}
        self.options = {k: v for k, v in kwds.items() if k in te}
        self.scf_options = {k: v for k, v in kwds.items() if k in _SCF_OPTIONS}
        dftb_opts = {k: v for k, v in kwds.items() if k not in te and k not in _SCF_OPTIONS and k != "charge"}
        self.atomlist = atomlist
        self.tddftb = LR_TDDFTB(atomlist, **dftb_opts)
        self.grads = Gradients(self.tddftb)
        self.Nst = Nst
        self.options.update(self.scf_options)
        self.options["nstates"] = Nst - 1
        self.last_calculation = None
        self.tdip_old = None

    def _set(self, x):
        self.atomlist = _vector2atomlist(x, self.atomlist)
        self.tddftb.setGeometry(self.atomlist, keep_charge=True, update_symmetry=False)

    def _energies(self):
        E0 = self.tddftb.dftb2.E_tot
        Exc = self.tddftb.getExcEnergies()
        return np.hstack(([E0], E0 + Exc))[:self.Nst]

    def getEnergies(self, x):
        self._set(x)
        self.tddftb.getEnergies(**self.options)
        return self._energies()

    def getEnergy_S0(self, x):
        self._set(x)
        return np.array([self.tddftb.dftb2.getEnergy(**self.scf_options)])

    def getEnergiesAndGradient(self, x, gradient_state):
        self._set(x)
        self.tddftb.getEnergies(**self.options)
        energies = self._energies()
        gradVrep, gradE0, gradExc = self.grads.gradient(I=gradient_state, save_intermediates_CPKS=1)
        return energies, gradVrep + gradE0 + gradExc

    def getEnergyAndGradient_S0(self, x, has_scf=False):
        if has_scf:
            E0 = self.tddftb.dftb2.E_tot
        else:
            self._set(x)
            E0 = self.tddftb.dftb2.getEnergy(**self.scf_options)
        gradVrep, gradE0, gradExc = self.grads.gradient(I=0, save_intermediates_CPKS=1)
        return E0, gradVrep + gradE0 + gradExc

    def resetXY(self):
        self.tddftb.XmY = None
        self.tddftb.XpY = None

    def resetSCF(self):
        gs = self.tddftb.dftb2
        for a in ("P", "dq"):
            if hasattr(gs, a):
                delattr(gs, a)

    def getScalarCouplings(self, threshold=0.01):
        """coupl[A,B] ~ <Psi_A|d/dt Psi_B> dt and the overlap matrix (PES.py:175-253)."""
        atomlist2 = self.atomlist
        orbs2 = self.tddftb.dftb2.getKSCoefficients()
        C2 = np.asarray(self.tddftb.Cij)[:self.Nst - 1]
        SC = ScalarCoupling(self.tddftb)
        if self.last_calculation is None:
            atomlist1, orbs1, C1 = atomlist2, orbs2, C2
        else:
            atomlist1, orbs1, C1 = self.last_calculation
        Sci = SC.ci_overlap(atomlist1, orbs1, C1, atomlist2, orbs2, C2, self.Nst, threshold=threshold)
        # align phases: the eigenvectors (and orbitals) carry arbitrary signs
        signs = np.sign(np.diag(Sci))
        self.P = np.diag(signs)
        C2 = np.tensordot(self.P[1:, :][:, 1:], C2, axes=(0, 0))
        Sci = Sci @ self.P
        if hasattr(self, "last_coupl"):
            s = np.sign(self.last_coupl[0, 1:] / Sci[0, 1:])
            w = np.abs(Sci[0, 1:] - self.last_coupl[0, 1:])
            sign0I = np.sign(np.sum(w * s) / np.sum(w))
            Sci[0, 1:] *= sign0I
            Sci[1:, 0] *= sign0I
        olap = Sci.copy()
        coupl = Sci
        coupl[np.diag_indices_from(coupl)] = 0.0
        err = np.sum(np.abs(coupl + coupl.T))
        if err > 1.0e-1:
            print("WARNING: Scalar coupling matrix is not antisymmetric, error = %s" % err)
        coupl = 0.5 * (coupl - coupl.T)
        self.last_coupl = coupl
        self.last_calculation = (atomlist2, orbs2, C2)
        return coupl, olap

    def getEnergiesBatch(self, xs, precision=0, conv=None, maxiter=None):
        """Energies of S0 .. S_{Nst-1} at many geometries of this molecule (a loop
        over geometries as PES.py:63-118 would run it, e.g. for sampling), with
        all excited-state problems solved in one persistent launch of the
        ensemble solver (lc = 1).  Returns [len(xs)][Nst]."""
        from .ensemble import build_operators, solve_operators
        from .groundstate.dftb2 import DFTB2
        gs0 = self.tddftb.dftb2
        opts = dict(long_range_correction=gs0.long_range_correction, long_range_radius=gs0.long_range_radius,
                    distance_cutoff=gs0.distance_cutoff)
        states, E0 = [], []
        for x in xs:
            gs = DFTB2(_vector2atomlist(x, self.atomlist), **opts)
            E0.append(gs.getEnergy(**self.scf_options))
            states.append(gs)
        ops = build_operators(states, self.options.get("nr_active_occ"), self.options.get("nr_active_virt"))
        res = solve_operators(ops, self.Nst - 1, conv=conv or self.options.get("diag_conv", 1e-5),
                              maxiter=maxiter or self.options.get("diag_maxiter", 50), precision=precision)
        return np.array([np.hstack(([e], e + np.asarray(w)))[:self.Nst] for e, w in zip(E0, res.Omega)])

    def saveLastCalculation(self):
        self.last_calculation = (self.atomlist, self.tddftb.dftb2.getKSCoefficients(),
                                 np.asarray(self.tddftb.Cij)[:self.Nst - 1])

    def alignCoefficients(self, c):
        if hasattr(self, "P"):
            c = np.dot(self.P[1:, :][:, 1:], c)
        return c

    def getTransitionDipoles(self):
        """Transition dipoles between all Nst states, signs aligned with the
        previous call (PES.py:266-296)."""
        tdip = self.tddftb.getTransitionDipolesExc(self.Nst)
        tdip_old = self.tdip_old if self.tdip_old is not None else tdip
        for A in range(self.Nst):
            for B in range(self.Nst):
                if np.dot(tdip_old[A, B, :], tdip[A, B, :]) < 0.0:
                    tdip[A, B, :] *= -1.0
        self.tdip_old = tdip
        return tdip
