"""Drop-in LR-TD-DFTB driver (the API of DFTB/LR_TDDFTB.py) on the MI355X engine.

``LR_TDDFTB(atomlist, **dftb2_opts).getEnergies(**opts) -> (Omega, oscillator_strength)``
with the reference's keywords (LR_TDDFTB.py:37-45) and attributes (Omega, Cij,
XmY, XpY, oscillator_strength, tdipX/Y/Z, Lambda2, qtrans_oo/ov/vv, nocc,
nvirt, active_occupied_orbs, active_virtual_orbs, occupied_orbs,
virtual_orbs, dftb2).  The ground state comes from ``self.dftb2``: any object
with DFTBaby's DFTB2 getters (DFTB2.py:2405-2454); by default the host-side
producer in ``dftbaby_amd.groundstate``.

Every excited-state pass runs on the GPU: transition charges (kernel (a)), the
Casida matvecs (b) and the Davidson loop (c), Oia / Lambda2 and the transition
dipoles.  The long-range CT correction (ct_correction=1, lc = 0 only) shifts
the diagonal as in LR_TDDFTB.py:429-437 (``dftbaby_amd.LongRangeCT``);
stability=1 prints the lowest eigenvalues of the stability matrix
(``dftbaby_amd.StabilityAnalysis``).  Out of scope (raise NotImplementedError):
multipole charges (mulliken_dipoles=1), symmetry selection and Tamm-Dancoff
(which the reference itself disables, Solver.py:194).
"""
import time

import numpy as np
import torch

from . import engine
from .engine import (CasidaOperatorGPU, ExcitedStatesError, ExcitedStatesNotConverged, _p, _stream,  # noqa: F401
                     gemm, gram)
from .Solver import precision_code
from . import _lib

HARTREE_TO_EV = 27.211396132
HARTREE_TO_NM = 45.563352527    # lambda (nm) = HARTREE_TO_NM / (energy in Hartree), AtomicData.py:3
EBOHR_TO_DEBYE = 2.54174623


class LR_TDDFTB:
    Omega = None
    # device results of the last solve; the host arrays Cij, XmY, XpY (reference
    # layout [k][nocc][nvirt]) are copied from them on first access
    _C_dev = _XmY_dev = _XpY_dev = None
    _Cij_h = _XmY_h = _XpY_h = None
    # Optional learned start (a dftbaby_amd.learned.PairMLP): when set and no
    # previous solution is available as a warm start, the lc Davidson starts
    # from the model's initial space and preconditioner omega; plain Davidson
    # is the fallback when that run does not converge.
    guess_model = None
    # long-range exchange products: "ao" (AO-basis GEMMs) or "charges" (the
    # reference's gamma_lr-folded transition-charge contractions); same results
    exchange = "ao"
    # True: the launch shapes of the (A+-B) products are picked by the project's
    # autotuner (dftbaby_amd.autotune: measured once per operator size and
    # precision, cached) before the Davidson
    autotune = False

    def __init__(self, atomlist=None, ground_state=None, **opts):
        if ground_state is None:
            from .groundstate.dftb2 import DFTB2
            ground_state = DFTB2(atomlist, **opts)
        self.dftb2 = ground_state
        self.timings = {}

    def _host_vectors(self, name):
        h = getattr(self, "_%s_h" % name)
        if h is None:
            d = getattr(self, "_%s_dev" % ("C" if name == "Cij" else name))
            if d is not None:
                h = self.op.from_padded(d).cpu().numpy()
                setattr(self, "_%s_h" % name, h)
        return h

    def _set_host_vectors(self, name, value):
        setattr(self, "_%s_h" % name, value)
        setattr(self, "_%s_dev" % ("C" if name == "Cij" else name), None)

    Cij = property(lambda self: self._host_vectors("Cij"), lambda self, v: self._set_host_vectors("Cij", v))
    XmY = property(lambda self: self._host_vectors("XmY"), lambda self, v: self._set_host_vectors("XmY", v))
    XpY = property(lambda self: self._host_vectors("XpY"), lambda self, v: self._set_host_vectors("XpY", v))

    def setGeometry(self, atomlist, **opts):
        self.dftb2.setGeometry(atomlist, **opts)

    def _t(self, name, t0):
        self.timings[name] = self.timings.get(name, 0.0) + (time.time() - t0)

    # ------------------------------------------------------------------ API
    def getEnergies(self, nr_active_occ=None, nr_active_virt=None, select_lm=None, oszis="Dipoles",
                    response_method="Casida", multiplicity="S", nstates=None, diag_ifact=1, diag_conv=1.0e-5,
                    diag_maxiter=50, diag_check=0, diag_L2threshold=0.0, diag_selected_irreps=None,
                    ct_correction=0, stability=0, precision="fp64", **scf_opts):
        t0 = time.time()
        self.en0 = self.dftb2.getEnergy(**scf_opts)
        self._t("ground state", t0)
        if select_lm is not None:
            self.setActiveLMOrbitals(nr_active_occ, nr_active_virt, select_lm=select_lm)
        else:
            self.setActiveOrbitals(nr_active_occ, nr_active_virt)
        self.setMultiplicity(multiplicity)
        self.nstates = nstates
        if nstates is not None:
            assert nstates > 0, "nstates==0, no excited states can be calculated!"
        self._solve_lr_equations(response_method=response_method, multiplicity=multiplicity, nstates=nstates,
                                 diag_ifact=diag_ifact, diag_conv=diag_conv, diag_maxiter=diag_maxiter,
                                 diag_check=diag_check, diag_L2threshold=diag_L2threshold,
                                 diag_selected_irreps=diag_selected_irreps, ct_correction=ct_correction,
                                 stability=stability, precision=precision)
        t0 = time.time()
        self._tdipMO_h = None
        if oszis == "Dipoles" and getattr(self.dftb2, "parameter_set", "homegrown") == "homegrown":
            self._TransitionDipoles()
        else:
            self._TransitionDipolesMulliken()
        self.getOscillatorStrengths()
        self._Lambda2_diagnostics()
        self.Irreps = ["" for _ in range(len(self.Omega))]
        self._t("properties", t0)
        if getattr(self.dftb2, "verbose", 0) > 0:
            # the reference computes these inside getEnergies only to print them
            # (LR_TDDFTB.py:89-112): done here when output is requested
            if oszis == "Dipoles" and hasattr(self, "tdipMO_X"):
                self.getTransitionDipolesExc(len(self.Omega))
            if nstates is not None:
                self.determineActiveSpace()
            self._writeExcitations()
        return self.Omega, self.oscillator_strength

    def setActiveOrbitals(self, nr_active_occ=30, nr_active_virt=30):
        """LR_TDDFTB.py:167-188."""
        f = self.dftb2.getOccupation()
        orbe = self.dftb2.getKSEnergies()
        self.occupied_orbs = np.where(f > 0.1)[0]
        self.virtual_orbs = np.where(f <= 0.1)[0]
        self.active_occupied_orbs = self.occupied_orbs[
            np.argsort(orbe[self.occupied_orbs], kind="stable")[::-1][:nr_active_occ][::-1]]
        self.active_virtual_orbs = self.virtual_orbs[np.argsort(orbe[self.virtual_orbs], kind="stable")[:nr_active_virt]]
        self.nocc = len(self.active_occupied_orbs)
        self.nvirt = len(self.active_virtual_orbs)

    def setActiveLMOrbitals(self, nr_active_occ=1000, nr_active_virt=1000, select_lm=((1, 0),)):
        """LR_TDDFTB.py:125-166: keep MOs with weight on AOs of the given (l, m)."""
        atomlist = self.dftb2.getGeometry()
        valorbs = self.dftb2.getValorbs()
        orbs = self.dftb2.getKSCoefficients()
        lm = []
        iao = 0
        for Zi, _ in atomlist:
            for (ni, li, mi) in valorbs[Zi]:
                if (li, mi) in list(select_lm):
                    lm.append(iao)
                iao += 1
        lm = np.array(lm, dtype=int)
        is_lm = np.sum(np.abs(orbs[lm, :]) ** 2, axis=0) > 1.0e-15
        f = self.dftb2.getOccupation()
        orbe = self.dftb2.getKSEnergies()
        self.occupied_orbs = np.where((f > 0.1) & is_lm)[0]
        self.virtual_orbs = np.where((f <= 0.1) & is_lm)[0]
        self.active_occupied_orbs = self.occupied_orbs[
            np.argsort(orbe[self.occupied_orbs], kind="stable")[::-1][:nr_active_occ][::-1]]
        self.active_virtual_orbs = self.virtual_orbs[np.argsort(orbe[self.virtual_orbs], kind="stable")[:nr_active_virt]]
        self.nocc = len(self.active_occupied_orbs)
        self.nvirt = len(self.active_virtual_orbs)

    def getActiveOrbitals(self):
        return self.active_occupied_orbs, self.active_virtual_orbs

    def setMultiplicity(self, multiplicity):
        assert multiplicity in ["S", "T"], "multiplicity has to be either 'S' for Singlet or 'T' for Triplet"
        self.multiplicity = multiplicity

    def getXY(self):
        return self.XmY, self.XpY

    def getExcEnergies(self):
        return self.Omega

    def getTransitionCharges(self):
        """(q_oo, q_vv, q_ov) in the reference layout (LR_TDDFTB.py:376-377)."""
        return self.qtrans_oo, self.qtrans_vv, self.qtrans_ov

    # ------------------------------------------------------------- core solve
    def _build_operator(self, coulomb=True):
        """Operator of the active space; coulomb=False drops the Coulomb kernel (the
        triplet matrices of the dense Casida, Solver.py:71-80: K_singlet only for "S")."""
        t0 = time.time()
        self._qref = None          # reference-layout charges belong to the previous geometry
        orbs = self.dftb2.getKSCoefficients()
        S = self.dftb2.getOverlapMatrix()
        orbe = self.dftb2.getKSEnergies()
        ao, av = self.active_occupied_orbs, self.active_virtual_orbs
        self.omega = orbe[av][None, :] - orbe[ao][:, None]
        gamma = self.dftb2.getGamma()
        if not coulomb:
            gamma = 0.0 * np.asarray(gamma)
        lc = int(getattr(self.dftb2, "long_range_correction", 0))
        gamma_lr = self.dftb2.getGamma_lr() if lc == 1 else 0.0 * gamma
        if getattr(self.dftb2, "mulliken_dipoles", 0) == 1:
            raise NotImplementedError("multipole transition charges (mulliken_dipoles=1) are out of scope")
        self.op = CasidaOperatorGPU.from_orbitals(orbs[:, ao], orbs[:, av], S, self.dftb2.orbsPerAtom,
                                                  self.omega, gamma, gamma_lr, lc, exchange=self.exchange)
        torch.cuda.synchronize()
        self._t("transition charges", t0)
        self._Lambda2_calcOia()

    def _Lambda2_calcOia(self):
        """O_ia = q_ii^T Omega q_aa / sqrt(O_ii O_aa), LR_TDDFTB.py:538-560 (three GEMMs)."""
        op = self.op
        dev = op.device
        Om = torch.as_tensor(np.asarray(self.dftb2.getGaussianOverlap(), dtype=np.float64), device=dev)
        qii = op.diag_oo()
        qaa = op.qaa
        if qii is None or qaa is None:
            raise RuntimeError("Lambda2 needs the diagonal charges q^ii_A and q^aa_A of the operator")
        o, v, nat = op.nocc, op.nvirt, op.nat
        X = torch.empty((o, nat), dtype=torch.float64, device=dev)
        gemm(qii, Om, X, o, nat, nat, nat, nat, nat)
        Y = torch.empty((v, nat), dtype=torch.float64, device=dev)
        gemm(qaa, Om, Y, v, nat, nat, nat, nat, nat)
        Oia = torch.zeros((op.opad, op.vpad), dtype=torch.float64, device=dev)
        gemm(X, qaa, Oia, o, v, nat, nat, nat, op.vpad, b_trans=1)
        Oii = torch.empty(o, dtype=torch.float64, device=dev)
        Oaa = torch.empty(v, dtype=torch.float64, device=dev)
        # row-wise dots X[i].qii[i], Y[a].qaa[a] as weighted sums: sum_A X[i][A] qii[i][A]
        Gi = torch.empty((o, o), dtype=torch.float64, device=dev)
        gemm(X, qii, Gi, o, o, nat, nat, nat, o, b_trans=1)
        Ga = torch.empty((v, v), dtype=torch.float64, device=dev)
        gemm(Y, qaa, Ga, v, v, nat, nat, nat, v, b_trans=1)
        Oii.copy_(torch.diagonal(Gi))
        Oaa.copy_(torch.diagonal(Ga))
        _lib.call("dftb_oia_normalize", _p(Oia), _p(Oii), _p(Oaa), o, v, op.opad, op.vpad, _stream())
        self.Oia_dev = Oia.reshape(-1)
        self.Oia = Oia[:o, :v].cpu().numpy()

    def _solve_lr_equations(self, response_method="Casida", multiplicity="S", nstates=None, diag_ifact=1,
                            diag_conv=1.0e-10, diag_maxiter=10, diag_check=1, diag_L2threshold=0.0,
                            diag_selected_irreps=None, ct_correction=0, stability=0, precision="fp64"):
        """LR_TDDFTB.py:403-527 on the GPU."""
        if response_method == "Tamm-Dancoff":
            raise Exception("Revise the TDA code!!!!")   # Solver.py:194: disabled in the reference
        if response_method != "Casida":
            raise Exception("response_method should be \"Casida\" or \"Tamm-Dancoff\"")
        # the dense Casida honours the multiplicity (triplets: no Coulomb kernel); the
        # Davidsons compute singlets, as in the reference (LR_TDDFTB.py:471-492)
        self._build_operator(coulomb=not (nstates is None and multiplicity == "T"))
        op = self.op
        lc = op.lc
        prec = precision_code(precision)
        self.solve_precision = prec           # the Z-vector solve of the gradients follows it
        # stability of the ground state (LR_TDDFTB.py:423-427)
        if stability == 1:
            from .StabilityAnalysis import check_stability
            g = self.dftb2.getGamma()
            glr = self.dftb2.getGamma_lr() if lc == 1 else 0.0 * g
            qoo, qov, qvv = self._charges_ref()
            self.stability_eigenvalues = check_stability(g, glr, qoo, qvv, qov, self.omega, self.nocc, self.nvirt)
        # long-range CT correction: shifts of the orbital-energy differences (LR_TDDFTB.py:429-437)
        self.ct_correction = ct_correction
        omega_shift = None
        if ct_correction == 1:
            assert lc == 0, ("CT correction --ct_correction=1 cannot be combined with long-range correction "
                             "--long_range_correction=1!")
            from .LongRangeCT import LongRangeCTCorrection
            omega_shift = LongRangeCTCorrection(self).diagonal_shifts(self.omega,
                                                                       verbose=getattr(self.dftb2, "verbose", 0))
        self.omega_shift = omega_shift
        if self.autotune and nstates is not None:
            from .autotune import autotune
            self.tuning = autotune(op, prec)
        verbose = getattr(self.dftb2, "verbose", 0)
        t0 = time.time()
        if nstates is None:
            w, C, XmY, XpY = engine.casida_dense_gpu(op, omega_shift=omega_shift)
        elif lc == 0:
            w, C, XmY, XpY, st = engine.hermitian_davidson_gpu(
                op, nstates, omega_shift=omega_shift, ifact=diag_ifact, conv=diag_conv, maxiter=diag_maxiter,
                XpYguess=self.XpY, Oia=self.Oia, L2_thresh=diag_L2threshold, precision=prec, verbose=verbose)
            self.davidson_stats = st
        else:
            res = None
            if self.guess_model is not None and self.XpY is None:
                from .learned import learned_start
                V, pom, _ = learned_start(op, self.guess_model, nstates)
                try:
                    res = engine.nonhermitian_davidson_gpu(
                        op, nstates, conv=diag_conv, maxiter=diag_maxiter, precision=prec, verbose=verbose,
                        initial_space=V, precond_omega=pom)
                    res[4].learned = True
                except (engine.ExcitedStatesNotConverged, engine.ExcitedStatesError):
                    res = None
            if res is None:
                res = engine.nonhermitian_davidson_gpu(
                    op, nstates, ifact=diag_ifact, conv=diag_conv, maxiter=diag_maxiter, XmYguess=self.XmY,
                    XpYguess=self.XpY, w_guess=self.Omega, precision=prec, verbose=verbose)
            w, C, XmY, XpY, st = res
            self.davidson_stats = st
        torch.cuda.synchronize()
        self._t("Davidson" if nstates is not None else "Casida", t0)
        self._C_dev, self._XmY_dev, self._XpY_dev, self._w_dev = C, XmY, XpY, w
        self._Cij_h = self._XmY_h = self._XpY_h = None
        self.Omega = w.cpu().numpy()
        if diag_check == 1 and nstates is not None:
            Om_full, C_full, XmY_full, XpY_full = engine.casida_dense_gpu(op, omega_shift=omega_shift, nev=nstates)
            Om_full = Om_full.cpu().numpy()
            err = np.sum(np.abs(self.Omega - Om_full[:nstates]))
            assert err < 1.0e-10, "eigenvalues differ, err = %s" % err
            XmY_full = op.from_padded(XmY_full).cpu().numpy()
            XpY_full = op.from_padded(XpY_full).cpu().numpy()
            err = 0.0
            for n in range(nstates):
                err += min(np.abs(self.XmY[n] - XmY_full[n]).sum(), np.abs(self.XmY[n] + XmY_full[n]).sum())
                err += min(np.abs(self.XpY[n] - XpY_full[n]).sum(), np.abs(self.XpY[n] + XpY_full[n]).sum())
            assert err < 1.0e-10, "Davidson-like diagonalization and full diagonalization disagree, error = %s" % err

    # ------------------------------------------------------- reference layouts
    @property
    def qtrans_ov(self):
        op = self.op
        return op.q["c"][:, :op.nocc, :op.nvirt].cpu().numpy()

    @property
    def qtrans_oo(self):
        op = self.op
        if "oo" in op.q:
            return op.q["oo"][:, :, :op.nocc].permute(1, 0, 2).cpu().numpy()
        return self._charges_ref()[0]

    @property
    def qtrans_vv(self):
        return self._charges_ref()[2]

    def _charges_ref(self):
        """(q_oo, q_ov, q_vv) in the reference layout via dftb_trans_charges."""
        if getattr(self, "_qref", None) is None:
            import ctypes
            op = self.op
            dev = op.device
            orbs = self.dftb2.getKSCoefficients()
            S = self.dftb2.getOverlapMatrix()
            ao, av = self.active_occupied_orbs, self.active_virtual_orbs
            t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
            Co, Cv, St = t(orbs[:, ao]), t(orbs[:, av]), t(S)
            nat, o, v = op.nat, op.nocc, op.nvirt
            qoo = torch.empty((nat, o, o), dtype=torch.float64, device=dev)
            qov = torch.empty((nat, o, v), dtype=torch.float64, device=dev)
            qvv = torch.empty((nat, v, v), dtype=torch.float64, device=dev)
            opa = np.ascontiguousarray(self.dftb2.orbsPerAtom, dtype=np.int32)
            w, wn = op.ws.get(_lib.load().dftb_trans_charges_workspace(St.shape[0], o, v, nat))
            _lib.call("dftb_trans_charges", _p(Co), _p(Cv), _p(St), opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                      nat, St.shape[0], o, v, _p(qoo), _p(qov), _p(qvv), _p(w), wn, _stream())
            self._qref = (qoo.cpu().numpy(), qov.cpu().numpy(), qvv.cpu().numpy())
        return self._qref

    # ----------------------------------------------------------- properties
    def _TransitionDipoles(self):
        """d_I = sum_ia C^I_ia <i|r|a> with <p|r_c|q> = (C^T D_c C)_pq
        (LR_TDDFTB.py:574-592) and the ground state's dipole matrix D
        (getDipoleMatrix, DFTB2.py:2446-2452: Slater-Koster dipole tables).  The
        active occupied x virtual block takes two GEMMs per component,
        X = D_c C_v and C_o^T X; the full MO matrices tdipMO_X/Y/Z are formed on
        first access.  A ground state with getDipoleMatrixDevice (the bundled
        producer) hands over device planes [3][norb][norb] from the
        dftb_slako_matrices kernel; any other one its numpy [norb][norb][3]."""
        op = self.op
        dev = op.device
        gs = self.dftb2
        if hasattr(gs, "getDipoleMatrixDevice"):
            Dp = gs.getDipoleMatrixDevice(dev)
        else:
            D = np.asarray(gs.getDipoleMatrix(), dtype=np.float64)
            Dp = torch.as_tensor(np.ascontiguousarray(np.moveaxis(D, 2, 0)), device=dev)
        C = torch.as_tensor(np.ascontiguousarray(gs.getKSCoefficients(), dtype=np.float64), device=dev)
        norb = C.shape[0]
        o, v = op.nocc, op.nvirt
        Co = C[:, torch.as_tensor(self.active_occupied_orbs, device=dev)].contiguous()
        Cv = C[:, torch.as_tensor(self.active_virtual_orbs, device=dev)].contiguous()
        M = torch.zeros((3, op.opad, op.vpad), dtype=torch.float64, device=dev)
        X = torch.empty((norb, v), dtype=torch.float64, device=dev)
        for c in range(3):
            gemm(Dp[c], Cv, X, norb, v, norb, norb, v, v)
            gemm(Co, X, M[c], o, v, norb, o, v, op.vpad, a_trans=1)
        self._tdipMO_src = (Dp, C)
        self._tdipMO_h = [None, None, None]
        self._tdip_from(M.reshape(3, -1))
        return self.tdipX, self.tdipY, self.tdipZ

    def _tdipMO(self, c):
        """<p|r_c|q> over all MOs (host, [norb][norb])."""
        h = getattr(self, "_tdipMO_h", None)
        if h is None:
            raise AttributeError("transition dipoles between MOs are formed by getEnergies(oszis='Dipoles')")
        if h[c] is None:
            Dp, C = self._tdipMO_src
            norb = C.shape[0]
            X = torch.empty_like(C)
            gemm(Dp[c], C, X, norb, norb, norb, norb, norb, norb)
            T = torch.empty_like(C)
            gemm(C, X, T, norb, norb, norb, norb, norb, norb, a_trans=1)
            h[c] = T.cpu().numpy()
        return h[c]

    tdipMO_X = property(lambda self: self._tdipMO(0))
    tdipMO_Y = property(lambda self: self._tdipMO(1))
    tdipMO_Z = property(lambda self: self._tdipMO(2))

    def _tdip_from(self, Mflat):
        op = self.op
        k = self._C_dev.shape[0]
        out = torch.empty((k, 3), dtype=torch.float64, device=op.device)
        gram(self._C_dev, k, Mflat, 3, op.N, out, op.ws)     # split over the pair index
        t = out.cpu().numpy()
        self.tdipX, self.tdipY, self.tdipZ = t[:, 0].copy(), t[:, 1].copy(), t[:, 2].copy()

    def _TransitionDipolesMulliken(self):
        """d_I = sum_A R_A sum_ia q^{ia}_A C^I_ia  (LR_TDDFTB.py:567-573)."""
        op = self.op
        dev = op.device
        pos = torch.as_tensor(np.array([x for _, x in self.dftb2.getGeometry()], dtype=np.float64), device=dev)
        # M[c][x] = sum_A R_A^c q_c[A][x]
        M = torch.empty((3, op.N), dtype=torch.float64, device=dev)
        gemm(pos, op.q["c"], M, 3, op.N, op.nat, 3, op.N, op.N, a_trans=1, tile=64)
        self._tdip_from(M)

    def getTransitionDipolesExc(self, Nst):
        """Dipole matrix elements between the ground and the lowest Nst-1 excited states
        (LR_TDDFTB.py:593-697), in e*bohr with electrons counted positive:

          tdip[0,0] = d0 (electronic + nuclear), tdip[0,I] = sum_ia C^I_ia <i|r|a>,
          tdip[A,B] = <C_B, d0 C_A - D_oo^T C_A + C_A D_vv>   (A, B >= 1),

        the last line being the reference's P0/Pij/Pab contraction rewritten as two
        GEMMs per component.
        """
        op = self.op
        dev = op.device
        o, v = op.nocc, op.nvirt
        occ = np.asarray(self.occupied_orbs)
        ao, av = np.asarray(self.active_occupied_orbs), np.asarray(self.active_virtual_orbs)
        tMO = (self.tdipMO_X, self.tdipMO_Y, self.tdipMO_Z)
        d0_elec = np.array([2.0 * np.trace(t[np.ix_(occ, occ)]) for t in tMO])
        valence = {1: 1, 6: 4, 7: 5, 8: 6}
        d0_nuc = -sum(valence[Z] * np.asarray(pos) for Z, pos in self.dftb2.getGeometry())
        d0 = d0_elec + d0_nuc
        nex = min(Nst - 1, self._C_dev.shape[0])
        tdip = np.zeros((Nst, Nst, 3))
        tdip[0, 0] = d0
        for c in range(3):
            tdip[0, 1:nex + 1, c] = (self.tdipX, self.tdipY, self.tdipZ)[c][:nex]
            tdip[1:nex + 1, 0, c] = tdip[0, 1:nex + 1, c]
        if nex == 0:
            return tdip
        C = self._C_dev[:nex].reshape(nex, op.opad, op.vpad)
        for c in range(3):
            Doo = torch.as_tensor(np.ascontiguousarray(tMO[c][np.ix_(ao, ao)]), device=dev)
            Dvv = torch.as_tensor(np.ascontiguousarray(tMO[c][np.ix_(av, av)]), device=dev)
            Y = torch.zeros((nex, op.opad, op.vpad), dtype=torch.float64, device=dev)
            for a in range(nex):
                Ca = C[a]
                # Y_a = d0 C_a - D_oo^T C_a + C_a D_vv
                gemm(Doo, Ca, Y[a], o, v, o, o, op.vpad, op.vpad, a_trans=1, alpha=-1.0)
                gemm(Ca, Dvv, Y[a], o, v, v, op.vpad, v, op.vpad, beta=1.0)
                Y[a, :o, :v] += d0[c] * Ca[:o, :v]
            G = torch.empty((nex, nex), dtype=torch.float64, device=dev)
            gram(self._C_dev, nex, Y.reshape(nex, -1), nex, op.N, G, op.ws)
            tdip[1:nex + 1, 1:nex + 1, c] = G.cpu().numpy().T
        self.tdip_exc = tdip
        return tdip

    def getOscillatorStrengths(self):
        """LR_TDDFTB.py:698-705."""
        if self.multiplicity == "S":
            self.oscillator_strength = 2.0 / 3.0 * self.Omega * (np.abs(self.tdipX) ** 2 + np.abs(self.tdipY) ** 2
                                                                 + np.abs(self.tdipZ) ** 2)
        else:
            self.oscillator_strength = 0.0 * self.Omega
        return self.oscillator_strength

    def _Lambda2_diagnostics(self):
        op = self.op
        k = self._C_dev.shape[0]
        out = torch.empty(k, dtype=torch.float64, device=op.device)
        _lib.call("dftb_weighted_sumsq", _p(self._C_dev), _p(self.Oia_dev), k, op.N, _p(out), _stream())
        self.Lambda2 = out.cpu().numpy()
        return self.Lambda2

    def determineActiveSpace(self, threshold=1.0e-5):
        """LR_TDDFTB.py:196-243 (host analysis of the k excitation vectors): the
        active space that holds 1 - threshold of every state's probability, with
        the reference's report when the ground state is verbose."""
        Nst = self.Cij.shape[0]
        occ_mins, virt_maxs = [], []
        for n in range(Nst):
            Cn = self.Cij[n]
            Cn2 = np.ravel(Cn * Cn)
            order = np.argsort(-Cn2, kind="stable")
            Sn = np.cumsum(Cn2[order])
            active = order[np.where(Sn <= 1.0 - threshold)]
            if len(active) == 0:
                active = order[:1]
            occ, virt = np.unravel_index(active, Cn.shape)
            occ_mins.append(self.active_occupied_orbs[occ.min()])
            virt_maxs.append(self.active_virtual_orbs[virt.max()])
        HOMO = self.occupied_orbs.max()
        nr_active_occ = HOMO - min(occ_mins) + 1
        nr_active_virt = max(virt_maxs) - HOMO
        if getattr(self.dftb2, "verbose", 0) > 0:
            print(self.active_space_report(occ_mins, virt_maxs, nr_active_occ, nr_active_virt, threshold))
        return nr_active_occ, nr_active_virt

    def active_space_report(self, occ_mins, virt_maxs, nr_active_occ, nr_active_virt, threshold):
        """The text determineActiveSpace prints (LR_TDDFTB.py:228-241)."""
        Nst = len(occ_mins)
        lines = ["", "Active Space", "============"]
        for n in range(Nst):
            lines.append("State %3d: contains excitations from window  occ %3d - virt %3d"
                         % (n + 1, occ_mins[n], virt_maxs[n]))
        lines += ["",
                  "The suggested active space for describing the lowest %d states consists of" % Nst,
                  "the highest %d occupied and the lowest %d virtual orbitals." % (nr_active_occ, nr_active_virt),
                  "This active space accounts for %2.7f of the total probability for each state." % (1.0 - threshold),
                  "You should check that the active space contains all neighbouring (almost) degenerate orbitals.",
                  "Dimension of active space: %d" % (nr_active_occ * nr_active_virt),
                  "Dimension of  full  space: %d" % (len(self.occupied_orbs) * len(self.virtual_orbs)),
                  ""]
        return "\n".join(lines)

    def _dominant_excitations(self, I, dominant_probability=0.99):
        C2 = np.abs(self.Cij[I]) ** 2
        order = np.argsort(-C2, axis=None, kind="stable")
        ex_occ, ex_virt, C = [], [], []
        prob = 0.0
        for idx in order:
            if prob >= dominant_probability:
                break
            i, a = np.unravel_index(idx, C2.shape)
            dp = C2[i, a]
            if dp > 1.0e-10:
                ex_occ.append(self.active_occupied_orbs[i])
                ex_virt.append(self.active_virtual_orbs[a])
                C.append(self.Cij[I, i, a])
            prob += dp
        return ex_occ, ex_virt, C

    def _writeExcitations(self, print_max_states=1000000, nr_dominant_ex=2):
        """Excited-state table of LR_TDDFTB.py:759-806."""
        print(self.excitation_table(print_max_states, nr_dominant_ex))

    def excitation_table(self, print_max_states=1000000, nr_dominant_ex=2):
        txt = " ground state energy: %s %s\n\n" % (("%.7f hartree" % self.en0).rjust(15),
                                                   ("%.7f eV" % (self.en0 * HARTREE_TO_EV)).rjust(15))
        txt += "           Excited States\n           ==============\n"
        symstr = "   "          # no symmetry group: three blanks (LR_TDDFTB.py:768-771)
        txt += ("     Spn %s     exc. en. /hartree      exc. en. / eV   exc. en. / nm    osc. strength"
                "             dominant contrib.               en_KS / eV      transition dipole [TDx TDy TDz]"
                "   /\\ diagn.\n" % symstr)
        orbe = self.dftb2.getKSEnergies()
        for I in range(min(print_max_states, len(self.Omega))):
            domex = self._dominant_excitations(I)
            domex_str = 7 * " "
            domKS = 6 * " "
            for count, (iocc, avirt, Cia) in enumerate(zip(*domex)):
                if count == 0:
                    domKS = "%2.4f" % ((orbe[avirt] - orbe[iocc]) * HARTREE_TO_EV)
                domex_str += "%d->%d(%+.3e)" % (iocc + 1, avirt + 1, Cia)
                if count < min(nr_dominant_ex, len(domex[0])) - 1:
                    domex_str += ","
                if count == nr_dominant_ex - 1:
                    break
            txt += "%s %s %s: %s %s %s  %s %s %s %s %s\n" % (
                str(I + 1).rjust(5), self.multiplicity, self.Irreps[I].rjust(3),
                ("%.7f hartree" % self.Omega[I]).rjust(20), ("%.7f eV" % (self.Omega[I] * HARTREE_TO_EV)).rjust(17),
                ("%.7f nm" % (HARTREE_TO_NM / self.Omega[I])).rjust(17),
                ("%.7f" % self.oscillator_strength[I]).rjust(12), domex_str.rjust(45), domKS.rjust(8),
                ("[%+2.7f %+2.7f %+2.7f]" % (self.tdipX[I], self.tdipY[I], self.tdipZ[I])).rjust(40),
                ("%.4f" % self.Lambda2[I]).rjust(7))
        return txt

    def saveTransitionCharges(self, save_transition_charges=None):
        """Mulliken transition charges between the ground and each excited state,
        q^I_A = sum_ia q^{ia}_A C^I_ia (one GEMM over the pair index), written as
        LR_TDDFTB.py:1008-1026 does: three comment lines, then one row per atom and
        one column per state, in units of -e (np.savetxt of -q)."""
        if save_transition_charges is None:
            return None
        import os
        op = self.op
        k = self._C_dev.shape[0]
        out = torch.empty((op.nat, k), dtype=torch.float64, device=op.device)
        gram(op.q["c"].reshape(op.nat, op.N), op.nat, self._C_dev, k, op.N, out, op.ws)
        tc = out.cpu().numpy()
        with open(os.path.expandvars(os.path.expanduser(save_transition_charges)), "w") as fh:
            fh.write("# Mulliken transition charges (in unit of -e) between ground and excited states\n")
            fh.write("# ROWS: atoms\n")
            fh.write("# COLUMNS: excited states\n")
            np.savetxt(fh, -tc)
        return tc

    def saveAbsorptionSpectrum(self, spectrum_file="absorption_spectrum.dat"):
        """LR_TDDFTB.py:807-832 (no symmetry labels: one file)."""
        with open(spectrum_file, "w") as fh:
            fh.write("# Absorption spectrum\n")
            fh.write("# exc. energy / hartree  osc.strength\n")
            for I in range(len(self.Omega)):
                fh.write("%s   %s\n" % (("%.7f" % self.Omega[I]).rjust(20),
                                        ("%.7f" % self.oscillator_strength[I]).rjust(20)))
