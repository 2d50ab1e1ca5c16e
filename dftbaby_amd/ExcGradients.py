"""Analytic gradients of LC-TD-DFTB excitation energies on the GPU.

Drop-in for ``DFTB.ExcGradients.Gradients`` (``DFTB/ExcGradients.py:213-711``) for
the long-range corrected case: ``Gradients(tddftb).gradient(I)`` returns
``(gradVrep, gradE0, gradExc)``, each of length 3 Nat (derivative of the total
energy of state I = their sum).  The terms are the reference's (``gradient_lc``):
the ground-state gradient (``:477-496``), the relaxed excitation-energy gradient
from U, V, T, the H^+ / H^- contractions, Q, the Z-vector equation (solved by
``engine.zvector_gpu`` on the (A+B) operator, in the precision of the excited-state
solve with FP64 refinement to the reference's residual test), W, and the contractions with the
derivatives of S, H0, gamma and gamma_lr (``:590-711``).

What differs is how they are formed (identical results to rounding):

* No [3 Nat][Norb][Norb] derivative tensor is built.  Every term is linear in
  one of dS/dR, dH0/dR, dgamma_AO/dR, dgamma_lr_AO/dR, so the contractions with
  the operators F and F^lr (``:348-424``) are rewritten by moving the
  derivative to one side of a Frobenius product, e.g. for term 1 of F^lr
  <M, gamma_lr o (dS (S v)^T)> = <(M o gamma_lr) S v, dS>; all of them are
  summed into four Norb x Norb matrices (MFMA FP64 GEMMs) that the kernel
  ``dftb_grad_pairs`` contracts with the derivative blocks it evaluates per atom
  pair from the Slater-Koster splines (the reference's semantics: values from
  splev_uniform, derivatives from splder, the lighter element first).
* The H^+ / H^- contractions of transition charges (``:567-590``) are formed in
  the AO basis with q^{pq}_A = 1/2 sum_{mu in A} (C_mu,p Y_mu,q + Y_mu,p C_mu,q),
  Y = S C: the Coulomb part from orbital-resolved Mulliken populations, the two
  exchange parts as L_p^T [Gamma_2 o (R_r v R_s^T)] L_q (and its transposed
  bracket) with L = [C; Y], R = [Y; C] and Gamma_2 the 2 x 2 block repetition
  of gamma_lr on orbital pairs - no q_oo / q_vv.

Without the long-range correction ``gradient_nolc`` (``:713-981``) is the same
computation with gamma_lr = 0 (no F^lr terms, H^+ reduced to its Coulomb term, no
H^- terms, the lc = 0 (A+B) operator).  gradVrep (``gradient_Vrep``, ``:181-210``) comes from the
``dftb_grad_vrep`` kernel on the repulsive-potential splines (their derivative
is clamped below the first table point, where scipy would extrapolate; atoms
that close do not occur).
A restricted active space is handled as the reference does: the amplitudes are
padded with zeros to the full space (``:236-263``).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .engine import CasidaOperatorGPU, _p, _stream, gemm, zvector_gpu
from .groundstate import slako as SK


def _mm(A, B, ta=False, tb=False, alpha=1.0, out=None, beta=0.0):
    """out = alpha op(A) op(B) + beta out on the MFMA FP64 GEMM (row-major)."""
    A = A.contiguous()
    B = B.contiguous()
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    if out is None:
        out = torch.empty((M, N), dtype=torch.float64, device=A.device)
    gemm(A, B, out, M, N, K, A.shape[1], B.shape[1], out.shape[1], a_trans=int(ta), b_trans=int(tb), alpha=alpha,
         beta=beta)
    return out


class Gradients:
    """Gradients of the excited states of an LR_TDDFTB instance (ExcGradients.py:213)."""

    # products of the Z-vector solve: None = those of the excited-state solve
    # (precision code 0 FP64, 1 FP32, 2 BF16x3)
    zvector_precision = None

    def __init__(self, tddftb):
        self.tddftb = tddftb
        self.dftb = tddftb.dftb2
        self.Zia = None
        self.I = None

    def gradient(self, I=0, check_zvector=0, save_intermediates_CPKS=0):
        """(gradVrep, gradE0, gradExc) of state I (0 = ground state), through
        gradient_lc or, without the long-range correction, gradient_nolc
        (ExcGradients.py:234-285)."""
        if self.dftb.long_range_correction == 0:
            return self.gradient_nolc(I=I, check_zvector=check_zvector, save_intermediates_CPKS=save_intermediates_CPKS)
        return self.gradient_lc(I=I, check_zvector=check_zvector, save_intermediates_CPKS=save_intermediates_CPKS)

    def gradient_nolc(self, I=0, check_zvector=0, save_intermediates_CPKS=0):
        """ExcGradients.py:713-981: the terms of gradient_lc without gamma_lr (no
        F^lr, H^+ reduced to its Coulomb term, no H^- terms, the lc = 0 (A+B))."""
        return self.gradient_lc(I=I, check_zvector=check_zvector, save_intermediates_CPKS=save_intermediates_CPKS)

    # ------------------------------------------------------------------ pieces
    def _setup(self):
        gs = self.dftb
        td = self.tddftb
        dev = td.op.device if getattr(td, "op", None) is not None else torch.device("cuda")
        self.dev = dev
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        geo = gs.getGeometry()
        self.Z = np.array([z for z, _ in geo], dtype=np.int64)
        self.pos = np.array([x for _, x in geo], dtype=np.float64)
        opa = np.asarray(gs.orbsPerAtom, dtype=np.int64)
        self.amap = np.repeat(np.arange(len(self.Z)), opa)
        f = gs.getOccupation()
        orbe = gs.getKSEnergies()
        orbs = gs.getKSCoefficients()
        self.occ = np.where(f > 0.1)[0]
        self.virt = np.where(f <= 0.1)[0]
        self.S = t(gs.getOverlapMatrix())
        C = t(orbs)
        self.Co = C[:, torch.as_tensor(self.occ, device=dev)].contiguous()
        self.Cv = C[:, torch.as_tensor(self.virt, device=dev)].contiguous()
        self.eo = t(orbe[self.occ])
        self.ev = t(orbe[self.virt])
        self.Yo = _mm(self.S, self.Co)
        self.Yv = _mm(self.S, self.Cv)
        am = torch.as_tensor(self.amap, device=dev)
        self.lc = int(gs.long_range_correction) == 1
        self.g = t(gs.getGamma())
        self.glr = t(gs.getGamma_lr()) if self.lc else torch.zeros_like(self.g)
        self.gAO = self.g[am][:, am].contiguous()
        self.GAO = self.glr[am][:, am].contiguous()
        self.G2 = self.GAO.repeat(2, 2).contiguous()
        n = self.S.shape[0]
        self.AS = torch.zeros((n, n), dtype=torch.float64, device=dev)
        self.AH = torch.zeros((n, n), dtype=torch.float64, device=dev)
        self.BG = torch.zeros((n, n), dtype=torch.float64, device=dev)
        self.BL = torch.zeros((n, n), dtype=torch.float64, device=dev)
        # pair-kernel inputs
        T = SK.tables()
        atom, ls, ms, _, types = SK.orbital_lists(self.Z, T)
        ti = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
        self.k_pos = t(self.pos)
        self.k_type = ti(types)
        self.k_off = ti(np.concatenate([[0], np.cumsum(opa)]))
        self.k_l, self.k_m = ti(ls), ti(ms)
        self.k_sigma = t([1.0 / (np.sqrt(np.pi) * gs.hubbard_U[int(z)]) for z in self.Z])
        self.k_knots, self.k_coefs, self.k_dcoefs = T.device_arrays(dev)
        self.k_T = T

    def _clear(self):
        for M in (self.AS, self.AH, self.BG, self.BL):
            M.zero_()

    def _pairs(self):
        """grad[3 Nat] = <AS, dS/dR> + <AH, dH0/dR> + <BG, dgamma/dR> + <BL, dgamma_lr/dR>."""
        gs = self.dftb
        nat, n = len(self.Z), self.S.shape[0]
        out = torch.empty(3 * nat, dtype=torch.float64, device=self.dev)
        cutoff = float(getattr(gs, "distance_cutoff", 30.0))
        _lib.call("dftb_grad_pairs", _p(self.k_pos), _p(self.k_type), _p(self.k_off), _p(self.k_l), _p(self.k_m),
                  _p(self.k_sigma), ctypes.c_double(float(gs.long_range_radius)), _p(self.k_knots),
                  _p(self.k_coefs), _p(self.k_dcoefs), self.k_T.npair, self.k_T.npts, ctypes.c_double(cutoff),
                  _p(self.AS), _p(self.AH), _p(self.BG), _p(self.BL) if self.lc else None, nat, n, _p(out), _stream())
        return out.cpu().numpy()

    def _grad_vrep(self):
        """gradient_Vrep (ExcGradients.py:181-210) by dftb_grad_vrep."""
        nat = len(self.Z)
        T = self.k_T
        knots, dcoefs, dmax = T.device_rep_arrays(self.dev)
        out = torch.empty(3 * nat, dtype=torch.float64, device=self.dev)
        _lib.call("dftb_grad_vrep", _p(self.k_pos), _p(self.k_type), _p(knots), _p(dcoefs), _p(dmax), T.npair,
                  T.rep_npts, nat, _p(out), _stream())
        return out.cpu().numpy()

    def _phi_F(self, v, M, wgt):
        """Adds wgt <F(v), M> (ExcGradients.py:348-366) to the contraction matrices:
        <F(v), M> = 1/4 (<M o (gSv 1^T + 1 gSv^T) + diag(gamma w) (v + v^T), dS/dR>
        + <w Sv^T, dgamma/dR>), Sv = rowsum(S o (v + v^T)), gSv = gamma_AO Sv,
        w = rowsum(S o (M + M^T))."""
        vp = v + v.T
        Sv = (self.S * vp).sum(1, keepdim=True)
        gSv = _mm(self.gAO, Sv)
        w = (self.S * (M + M.T)).sum(1, keepdim=True)
        gw = _mm(self.gAO, w)
        c = 0.25 * wgt
        self.AS.add_(c * (M * (gSv + gSv.T) + gw * vp))
        self.BG.add_(c * (w * Sv.T))

    def _phi_Flr(self, v, M, wgt):
        """Adds wgt <F^lr(v), M> (ExcGradients.py:367-424, its twelve terms) to the
        contraction matrices: eight terms against dS/dR, four against
        dgamma_lr/dR (each term's derivative moved to one side of the product)."""
        S, G = self.S, self.GAO
        Sv = _mm(S, v)
        SvT = _mm(S, v, tb=True)
        gv = G * v
        MG = M * G
        MS = _mm(M, S)
        MtS = _mm(M, S, ta=True)
        A = _mm(MG, Sv)                                   # term 1
        _mm(MS * G, v, out=A, beta=1.0)                   # term 2
        _mm(M, Sv * G, out=A, beta=1.0)                   # term 3
        _mm(MS, gv, out=A, beta=1.0)                      # term 4
        _mm(_mm(v, S), MG, out=A, beta=1.0)               # term 5
        _mm(SvT * G, M, ta=True, out=A, beta=1.0)         # term 6
        _mm(MtS * G, v, tb=True, out=A, beta=1.0)         # term 7
        _mm(MtS, gv, tb=True, out=A, beta=1.0)            # term 8
        B = M * _mm(SvT, S)                               # term 9:  M o (S v^T S)
        B += MS * SvT                                     # term 10
        B += MtS * Sv                                     # term 11
        B += _mm(MS, S, ta=True) * v                      # term 12: (S M^T S) o v
        c = 0.25 * wgt
        self.AS.add_(c * A)
        self.BL.add_(c * B)

    def _H(self, p, q, r, s, v):
        """The three parts (Coulomb, exchange, transposed exchange) of H^+_pq[v_rs]
        (ExcGradients.py:567-590) for orbital spaces p, q, r, s in 'o' / 'v':
        H^+ = t1 - t2 - t3, H^- (q_ps, q_qr, q_pr, q_qs) = t3 - t2."""
        sp = {"o": (self.Co, self.Yo), "v": (self.Cv, self.Yv)}
        (Cp, Yp), (Cq, Yq), (Cr, Yr), (Cs, Ys) = sp[p], sp[q], sp[r], sp[s]
        u = 0.5 * ((_mm(Cr, v) * Ys).sum(1, keepdim=True) + (_mm(Yr, v) * Cs).sum(1, keepdim=True))
        uu = _mm(self.gAO, u)
        t1 = _mm(Cp, uu * Yq, ta=True, alpha=2.0)
        _mm(Yp, uu * Cq, ta=True, alpha=2.0, out=t1, beta=1.0)
        if not self.lc:
            return t1, 0.0, 0.0
        Lp = torch.cat([Cp, Yp], 0)
        Lq = torch.cat([Cq, Yq], 0)
        Rr = torch.cat([Yr, Cr], 0)
        Rs = torch.cat([Ys, Cs], 0)
        Hd = _mm(_mm(Rr, v), Rs, tb=True)
        Hd.mul_(self.G2)
        t2 = _mm(Lp, _mm(Hd, Lq), ta=True, alpha=0.25)
        t3 = _mm(Lp, _mm(Hd, Lq, ta=True), ta=True, alpha=0.25)
        return t1, t2, t3

    def _Hp(self, p, q, r, s, v):
        t1, t2, t3 = self._H(p, q, r, s, v)
        return t1 - t2 - t3

    def _Hm(self, p, q, r, s, v):
        _, t2, t3 = self._H(p, q, r, s, v)
        return t3 - t2

    def _amplitudes(self, I):
        """X - Y, X + Y of state I (1-based) over the full orbital space."""
        td = self.tddftb
        XmY = np.asarray(td.XmY)[I - 1]
        XpY = np.asarray(td.XpY)[I - 1]
        o, v = len(self.occ), len(self.virt)
        ao = np.asarray(td.active_occupied_orbs)
        av = np.asarray(td.active_virtual_orbs)
        if XmY.shape == (o, v) and np.array_equal(ao, self.occ) and np.array_equal(av, self.virt):
            return XmY, XpY
        io = np.searchsorted(self.occ, ao)
        iv = np.searchsorted(self.virt, av)
        full_m = np.zeros((o, v))
        full_p = np.zeros((o, v))
        full_m[np.ix_(io, iv)] = XmY
        full_p[np.ix_(io, iv)] = XpY
        return full_m, full_p

    def _zvector_operator(self):
        td = self.tddftb
        op = getattr(td, "op", None)
        if (op is not None and op.nocc == len(self.occ) and op.nvirt == len(self.virt)
                and np.array_equal(np.asarray(td.active_occupied_orbs), self.occ)):
            return op
        gs = self.dftb
        orbs, orbe = gs.getKSCoefficients(), gs.getKSEnergies()
        omega = orbe[self.virt][None, :] - orbe[self.occ][:, None]
        lc = int(gs.long_range_correction)
        glr = gs.getGamma_lr() if lc else 0.0 * gs.getGamma()
        return CasidaOperatorGPU.from_orbitals(orbs[:, self.occ], orbs[:, self.virt], gs.getOverlapMatrix(),
                                               gs.orbsPerAtom, omega, gs.getGamma(), glr, lc)

    # ------------------------------------------------------------------ gradient
    def gradient_lc(self, I=0, check_zvector=0, save_intermediates_CPKS=0):
        """(gradVrep, gradE0, gradExc) of ExcGradients.py:286-711 (numpy, 3 Nat)."""
        self._setup()
        Co, Cv, S = self.Co, self.Cv, self.S
        nat = len(self.Z)
        n = S.shape[0]
        D = _mm(Co, Co, tb=True, alpha=2.0)
        occ_ref = np.concatenate([[1.0] if z == 1 else [2.0, 2.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0] for z in self.Z])
        dD = D - torch.diag(torch.as_tensor(occ_ref, device=self.dev))
        Den = _mm(Co * self.eo[None, :], Co, tb=True, alpha=2.0)
        # ground-state gradient (ExcGradients.py:477-496)
        self._clear()
        self.AH.copy_(D)
        self.AS.sub_(Den)
        self._phi_F(dD, dD, 0.5)
        if self.lc:
            self._phi_Flr(dD, dD, -0.25)
        gradE0 = self._pairs()
        gradVrep = self._grad_vrep()
        if I == 0:
            return gradVrep, gradE0, np.zeros(3 * nat)
        # excitation-energy gradient (ExcGradients.py:536-711)
        XmY, XpY = self._amplitudes(I)
        if abs(np.sum(XmY * XpY) - 1.0) >= 1e-10:
            raise AssertionError("(X-Y).(X+Y) = %s, not 1" % np.sum(XmY * XpY))
        Omega = float(np.asarray(self.tddftb.Omega)[I - 1])
        Xm = torch.as_tensor(XmY, device=self.dev)
        Xp = torch.as_tensor(XpY, device=self.dev)
        eo, ev = self.eo[:, None], self.ev[None, :]
        Tab = _mm(Xp, Xp, ta=True, alpha=0.5)
        _mm(Xm, Xm, ta=True, alpha=0.5, out=Tab, beta=1.0)
        Tij = _mm(Xp, Xp, tb=True, alpha=0.5)
        _mm(Xm, Xm, tb=True, alpha=0.5, out=Tij, beta=1.0)
        Uab = _mm(Xp, Xm, ta=True)
        _mm(Xm, Xp, ta=True, out=Uab, beta=1.0)
        Uij = _mm(Xp, Xm, tb=True)
        _mm(Xm, Xp, tb=True, out=Uij, beta=1.0)
        Vab = _mm(eo * Xp, Xp, ta=True)
        _mm(eo * Xm, Xm, ta=True, out=Vab, beta=1.0)
        Vij = _mm(Xp * ev, Xp, tb=True)
        _mm(Xm * ev, Xm, tb=True, out=Vij, beta=1.0)
        Gij = self._Hp("o", "o", "v", "v", Tab) - self._Hp("o", "o", "o", "o", Tij)
        Qij = Omega * Uij - Vij + Gij
        Qia = _mm(Xp, self._Hp("v", "v", "o", "v", Xp), tb=True)
        if self.lc:
            _mm(Xm, self._Hm("v", "v", "o", "v", Xm), tb=True, out=Qia, beta=1.0)
        Qia += self._Hp("o", "v", "v", "v", Tab)
        Qia -= self._Hp("o", "v", "o", "o", Tij)
        Qai = _mm(Xp, self._Hp("o", "o", "o", "v", Xp), ta=True)
        if self.lc:
            _mm(Xm, self._Hm("o", "o", "o", "v", Xm), ta=True, out=Qai, beta=1.0)
        Qab = Omega * Uab + Vab
        Ria = Qai.T - Qia
        # Z-vector equation (ExcGradients.py:620-640)
        op = self._zvector_operator()
        guess = self.Zia[None] if (self.I == I and self.Zia is not None) else None
        # the Z-vector in the precision of the excited-state solve (FP32 / BF16x3 products
        # with FP64 refinement to the same residual test), FP64 otherwise
        zprec = self.zvector_precision
        if zprec is None:
            zprec = getattr(self.tddftb, "solve_precision", 0) if op is getattr(self.tddftb, "op", None) else 0
        Zia_h = zvector_gpu(op, Ria[None].cpu().numpy(), X0=guess, precision=zprec)[0]
        self.Zia, self.I = Zia_h, I
        Zia = torch.as_tensor(Zia_h, device=self.dev)
        # W (ExcGradients.py:672-682)
        Wij = Qij + self._Hp("o", "o", "o", "v", Zia)
        Wij.diagonal().mul_(0.5)
        Wia = Qai.T + eo * Zia
        Wab = Qab.clone()
        Wab.diagonal().mul_(0.5)
        o = Co.shape[1]
        W = torch.zeros((n, n), dtype=torch.float64, device=self.dev)
        W[:o, :o] = torch.triu(Wij)
        W[:o, o:] = Wia
        W[o:, o:] = torch.triu(Wab)
        Call = torch.cat([Co, Cv], 1)
        Wao = _mm(_mm(Call, W), Call, tb=True)
        # assembly (ExcGradients.py:684-711)
        M = _mm(_mm(Cv, Tab), Cv, tb=True)
        _mm(_mm(Co, Tij), Co, tb=True, alpha=-1.0, out=M, beta=1.0)
        _mm(_mm(Co, Zia), Cv, tb=True, out=M, beta=1.0)
        XpYao = _mm(_mm(Co, Xp), Cv, tb=True)
        XmYao = _mm(_mm(Co, Xm), Cv, tb=True)
        self._clear()
        self.AH.copy_(M)
        self.AS.sub_(Wao)
        self._phi_F(dD, M, 1.0)
        self._phi_F(XpYao, XpYao, 2.0)
        if self.lc:
            self._phi_Flr(dD, M, -0.5)
            self._phi_Flr(XpYao + XpYao.T, XpYao, -0.5)
            self._phi_Flr(XmYao - XmYao.T, XmYao, 0.5)
        gradExc = self._pairs()
        return gradVrep, gradE0, gradExc
