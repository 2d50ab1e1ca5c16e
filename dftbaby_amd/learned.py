"""Learned Davidson initial space and preconditioner shift (the engine's learned part).

A small per-pair MLP (6 features -> 3 hidden layers of width 128, SiLU -> 2
outputs) is trained on converged eigenvectors that the plain Davidson itself
produces:

* output 0, the *importance score*: log10(sum_I C_{I,ia}^2 + 1e-10) over the
  computed states I;
* output 1, the *preconditioner shift*: delta_ia = sum_I w (Omega_I - omega_ia) /
  (sum_I w + 1e-3) with w = C_{I,ia}^2, i.e. how far the excitation energies a
  pair takes part in sit from its orbital-energy difference.

Features per pair (kernel dftb_pair_features): omega_ia, ||q^{ia}||_2, the
diagonals A_{ia,ia} and B_{ia,ia} of Solver.buildAB, A-B and sqrt((A+B)(A-B)),
normalised with the training-set mean/std.

Inference replaces the reference's initial expansion (unit vectors at the
lowest omega, Solver.py:232-251): the top-M pairs by predicted score span a
restricted Casida problem (A+-B)_{PP} (kernel dftb_restricted_ab), solved on the
GPU; its lowest roots, embedded in the full space, are the initial expansion
vectors, and omega + delta replaces omega in the preconditioner
1/(w - omega).  The Davidson loop, its FP64 residual test and its convergence
threshold are unchanged, so the plain method stays the judge; if the learned
start fails (ExcitedStatesError) the solver falls back to the plain start.

Forward, backward and Adam run on the GPU: the matmuls on the FP32 MFMA GEMM,
bias/activation/loss/optimiser steps in csrc/learned.hip.
"""
import ctypes
import json

import numpy as np
import torch

from . import _lib
from ._lib import call
from .engine import _p, _stream, gemm

F32 = _lib.GEMM_IN_F32
O32 = _lib.GEMM_OUT_F32
NFEAT = 6
NOUT = 2


class PairMLP:
    """FP32 MLP on the GPU with hand-written forward/backward."""

    def __init__(self, nin=NFEAT, width=128, nout=NOUT, seed=0, device=None):
        self.device = device or torch.device("cuda")
        self.shapes = [(nin, width), (width,), (width, width), (width,), (width, width), (width,),
                       (width, nout), (nout,)]
        n = sum(int(np.prod(s)) for s in self.shapes)
        rng = np.random.default_rng(seed)
        flat = np.zeros(n, dtype=np.float32)
        off = 0
        for s in self.shapes:
            k = int(np.prod(s))
            if len(s) == 2:
                flat[off:off + k] = (rng.standard_normal(k) * np.sqrt(2.0 / s[0])).astype(np.float32)
            off += k
        self.params = torch.as_tensor(flat, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0
        self.mean = torch.zeros(nin, dtype=torch.float32, device=self.device)
        self.inv_std = torch.ones(nin, dtype=torch.float32, device=self.device)
        self.nin, self.width, self.nout = nin, width, nout

    def _views(self, flat):
        out, off = [], 0
        for s in self.shapes:
            k = int(np.prod(s))
            out.append(flat[off:off + k].view(*s))
            off += k
        return out

    # ---- forward / backward ----------------------------------------------------
    def forward(self, X, keep=False):
        """X: [rows][nin] float32 (normalised).  Returns Y [rows][nout]."""
        W1, b1, W2, b2, W3, b3, W4, b4 = self._views(self.params)
        rows = X.shape[0]
        st = _stream()
        acts = [X]
        zs = []
        H = X
        for (W, b, act) in ((W1, b1, 1), (W2, b2, 1), (W3, b3, 1), (W4, b4, 0)):
            nin, nout = W.shape
            Z = torch.empty((rows, nout), dtype=torch.float32, device=self.device)
            gemm(H, W, Z, rows, nout, nin, nin, nout, nout, in_type=F32, out_type=O32)
            Hn = torch.empty_like(Z)
            call("dftb_mlp_bias_act", _p(Z), _p(b), rows, nout, act, _p(Hn), st)
            zs.append(Z)
            acts.append(Hn)
            H = Hn
        if keep:
            self._cache = (acts, zs)
        return H

    def backward(self, G):
        """G = dL/dY [rows][nout]; accumulates parameter gradients into self.grads."""
        acts, zs = self._cache
        Ws = self._views(self.params)[0::2]
        gW = self._views(self.grads)
        st = _stream()
        rows = G.shape[0]
        for layer in (3, 2, 1, 0):
            W = Ws[layer]
            nin, nout = W.shape
            Hin = acts[layer]
            # dW = Hin^T G ; db = colsum(G)
            gemm(Hin, G, gW[2 * layer], nin, nout, rows, nin, nout, nout, a_trans=1, in_type=F32, out_type=O32,
                 tile=64)
            call("dftb_mlp_colsum", _p(G), rows, nout, _p(gW[2 * layer + 1]), st)
            if layer > 0:
                Gp = torch.empty((rows, nin), dtype=torch.float32, device=self.device)
                gemm(G, W, Gp, rows, nin, nout, nout, nout, nin, b_trans=1, in_type=F32, out_type=O32)
                call("dftb_mlp_act_bwd", _p(Gp), _p(zs[layer - 1]), Gp.numel(), st)
                G = Gp

    def loss_and_grad(self, X, T, w):
        Y = self.forward(X, keep=True)
        G = torch.empty_like(Y)
        nb = 64
        part = torch.zeros(nb, dtype=torch.float32, device=self.device)
        call("dftb_mlp_mse", _p(Y), _p(T), _p(w), Y.shape[0], Y.shape[1], _p(G), _p(part), nb, _stream())
        self.backward(G)
        return part

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        call("dftb_adam", _p(self.params), _p(self.grads), _p(self.m), _p(self.v), self.params.numel(),
             ctypes.c_float(lr), ctypes.c_float(b1), ctypes.c_float(b2), ctypes.c_float(eps), self.step_count,
             ctypes.c_float(wd), _stream())

    # ---- training ----------------------------------------------------------------
    def fit(self, X, T, w=(1.0, 1.0), steps=2000, batch=4096, lr=2e-3, seed=0, log=None):
        """X [n][nin] raw features, T [n][nout] targets (device float32 tensors)."""
        mean = X.mean(0)
        std = X.std(0).clamp_min(1e-6)
        self.mean.copy_(mean)
        self.inv_std.copy_(1.0 / std)
        Xn = X.clone()
        call("dftb_mlp_normalize", _p(Xn), Xn.shape[0], Xn.shape[1], _p(self.mean), _p(self.inv_std), _stream())
        wt = torch.as_tensor(np.asarray(w, dtype=np.float32), device=self.device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = X.shape[0]
        hist = []
        for s in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), generator=g).to(self.device)
            xb = Xn[idx].contiguous()
            tb = T[idx].contiguous()
            lr_s = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps))
            part = self.loss_and_grad(xb, tb, wt)
            self.adam(lr=lr_s)
            if s % 200 == 0 or s == steps - 1:
                loss = float(part.sum().cpu())
                hist.append((s, loss))
                if log:
                    log("mlp step %d loss %.4e" % (s, loss))
        return hist

    def predict(self, Xraw):
        X = Xraw.clone()
        call("dftb_mlp_normalize", _p(X), X.shape[0], X.shape[1], _p(self.mean), _p(self.inv_std), _stream())
        return self.forward(X)

    # ---- persistence -------------------------------------------------------------
    def state(self):
        return {"params": self.params.cpu().numpy().tolist(), "mean": self.mean.cpu().numpy().tolist(),
                "inv_std": self.inv_std.cpu().numpy().tolist(), "nin": self.nin, "width": self.width,
                "nout": self.nout}

    def save(self, path):
        with open(path, "w") as fh:
            json.dump(self.state(), fh)

    @classmethod
    def load(cls, path, device=None):
        d = json.load(open(path))
        m = cls(d["nin"], d["width"], d["nout"], device=device)
        m.params.copy_(torch.as_tensor(np.asarray(d["params"], dtype=np.float32), device=m.device))
        m.mean.copy_(torch.as_tensor(np.asarray(d["mean"], dtype=np.float32), device=m.device))
        m.inv_std.copy_(torch.as_tensor(np.asarray(d["inv_std"], dtype=np.float32), device=m.device))
        return m


# ---------------------------------------------------------------------------
# features and targets
# ---------------------------------------------------------------------------
def _diagonal_integrals(op, chunk=1 << 17):
    """(ia|ia), (ia|ia)_lr and (ii|aa)_lr on the padded pair grid (FP64), and ||q^ia||^2.
    gamma q^c is formed for `chunk` pairs at a time (nat x chunk FP64 scratch)."""
    dev = op.device
    st = _stream()
    N, nat = op.N, op.nat
    qc = op.q["c"].reshape(nat, N)
    c = min(N, int(chunk))
    P = torch.empty((nat, c), dtype=torch.float64, device=dev)
    cc = torch.empty(N, dtype=torch.float64, device=dev)
    cl = torch.zeros(N, dtype=torch.float64, device=dev)
    qn2 = torch.empty(N, dtype=torch.float64, device=dev)
    for c0 in range(0, N, c):
        n = min(c, N - c0)
        q0 = qc[:, c0:]
        gemm(op.gamma, qc, P, nat, n, nat, nat, N, c, B_off=c0)
        call("dftb_coldot", _p(P), c, _p(q0), N, nat, n, _p(cc[c0:]), st)
        if op.lc:
            gemm(op.gamma_lr, qc, P, nat, n, nat, nat, N, c, B_off=c0)
            call("dftb_coldot", _p(P), c, _p(q0), N, nat, n, _p(cl[c0:]), st)
        call("dftb_coldot", _p(q0), N, _p(q0), N, nat, n, _p(qn2[c0:]), st)
    del P
    xl = torch.zeros((op.opad, op.vpad), dtype=torch.float64, device=dev)
    if op.lc:
        X = torch.empty((op.nocc, nat), dtype=torch.float64, device=dev)
        gemm(op.qii, op.gamma_lr, X, op.nocc, nat, nat, nat, nat, nat)
        gemm(X, op.qaa, xl, op.nocc, op.nvirt, nat, nat, nat, op.vpad, b_trans=1)
    return cc, cl, xl.reshape(-1), qn2


def _casida_diagonal(op, ints):
    cc, cl, xl, _ = ints
    w = op.omega
    apb = w + 4.0 * cc - xl - cl
    amb = w - xl + cl
    d = torch.sqrt(torch.clamp(apb * amb, min=1e-8))
    mask = real_mask(op)
    return torch.where(mask, d, w).contiguous()


def _features(op, ints):
    cc, cl, xl, qn2 = ints
    F = torch.empty((op.N, NFEAT), dtype=torch.float32, device=op.device)
    call("dftb_pair_features", _p(op.omega), _p(qn2), _p(cc), _p(cl), _p(xl), op.N, NFEAT, _p(F), _stream())
    return F


def casida_diagonal_omega(op):
    """sqrt((A+B)_ia,ia (A-B)_ia,ia) on real pairs (omega on the padding): the
    diagonal of the Casida operator, a sharper preconditioner than omega when
    the Coulomb and long-range exchange diagonals are large."""
    return _casida_diagonal(op, _diagonal_integrals(op))


def pair_features(op):
    """[N][6] float32 features on the padded pair grid of a CasidaOperatorGPU:
    omega, ||q^ia||, A_d, B_d, A_d - B_d, sqrt((A+B)_d (A-B)_d)."""
    return _features(op, _diagonal_integrals(op))


def real_mask(op):
    m = torch.zeros((op.opad, op.vpad), dtype=torch.bool, device=op.device)
    m[:op.nocc, :op.nvirt] = True
    return m.reshape(-1)


def pair_targets(op, Omega, C):
    """Targets [N][2] from converged Omega [k] and C [k][N] (device)."""
    w = (C * C)                                     # [k][N]
    ws = w.sum(0)
    score = torch.log10(ws + 1e-10)
    om = op.omega[None, :]
    shift = (w * (Omega[:, None] - om)).sum(0) / (ws + 1e-3)
    return torch.stack([score, shift], 1).to(torch.float32)


def training_rows(op, Omega, C, n_top=2000, n_rand=6000, seed=0):
    """Features and targets of the important pairs plus a random sample."""
    F = pair_features(op)
    T = pair_targets(op, Omega, C)
    mask = real_mask(op)
    idx = torch.nonzero(mask).reshape(-1)
    score = T[idx, 0]
    top = idx[torch.argsort(score, descending=True)[:n_top]]
    g = torch.Generator(device="cpu").manual_seed(seed)
    rnd = idx[torch.randint(0, len(idx), (min(n_rand, len(idx)),), generator=g).to(op.device)]
    sel = torch.cat([top, rnd])
    return F[sel].contiguous(), T[sel].contiguous()


# ---------------------------------------------------------------------------
# inference: learned initial space + preconditioner omega
# ---------------------------------------------------------------------------
def learned_start(op, model, nstates, M=None, use_shift=False, rank="model"):
    """Initial expansion vectors (2k x N, padded) and preconditioner diagonal (N).

    The M highest-scoring pairs (rank="model": the network's score; "omega":
    the smallest orbital-energy gaps, no network) span a restricted Casida
    problem whose lowest k roots give the initial space.  The preconditioner
    diagonal is the Casida diagonal sqrt((A+B)_d (A-B)_d); use_shift adds the
    network's second output to it (measured to slow convergence on the
    graphene flakes, so off by default)."""
    from .engine import Workspace
    dev = op.device
    st = _stream()
    k = int(nstates)
    ints = _diagonal_integrals(op)
    F = _features(op, ints)
    Y = model.predict(F) if model is not None else torch.zeros((op.N, NOUT), dtype=torch.float32, device=dev)
    mask = real_mask(op)
    score = Y[:, 0].clone() if rank == "model" else -op.omega.to(torch.float32)
    score[~mask] = -1e30
    M = int(M or max(8 * k, 128))
    M = min(M, op.nocc * op.nvirt)
    sc = score.cpu().numpy()
    sel = np.argpartition(-sc, M - 1)[:M]
    sel = np.sort(sel)
    ip = torch.as_tensor((sel // op.vpad).astype(np.int32), device=dev)
    ap = torch.as_tensor((sel % op.vpad).astype(np.int32), device=dev)
    # Coulomb block (p|q) = q_p^T gamma q_q
    qsel = op.q["c"].reshape(op.nat, op.N)[:, torch.as_tensor(sel, device=dev)].contiguous()   # [nat][M]
    X = torch.empty((op.nat, M), dtype=torch.float64, device=dev)
    gemm(op.gamma, qsel, X, op.nat, M, op.nat, op.nat, M, M)
    Kc = torch.empty((M, M), dtype=torch.float64, device=dev)
    gemm(qsel, X, Kc, M, M, op.nat, M, M, M, a_trans=1)
    ApB = torch.empty((M, M), dtype=torch.float64, device=dev)
    AmB = torch.empty((M, M), dtype=torch.float64, device=dev)
    if op.lc and op.exchange == "ao":
        # the operator keeps no transition-charge layouts: build them for the
        # occupied / virtual orbitals the selected pairs touch
        from .engine import CasidaOperatorGPU
        ii, aa = sel // op.vpad, sel % op.vpad
        isel, ilocal = np.unique(ii, return_inverse=True)
        asel, alocal = np.unique(aa, return_inverse=True)
        Co, Cv, St, opa = op.orbs
        om = op.omega_real[torch.as_tensor(isel, device=dev)][:, torch.as_tensor(asel, device=dev)]
        sub = CasidaOperatorGPU.from_orbitals(Co[:, torch.as_tensor(isel, device=dev)].contiguous(),
                                              Cv[:, torch.as_tensor(asel, device=dev)].contiguous(), St, opa,
                                              om, op.gamma, op.gamma_lr, 1, dev, exchange="charges")
        ipl = torch.as_tensor(ilocal.astype(np.int32), device=dev)
        apl = torch.as_tensor(alocal.astype(np.int32), device=dev)
        call("dftb_restricted_ab", _p(ipl), _p(apl), M, sub.nat, sub.opad, sub.vpad, _p(sub.omega), _p(Kc),
             _p(sub.q["oo"]), _p(sub.q["ov"]), _p(sub.q["qf_ov"]), _p(sub.q["qf_vv"]), 1, _p(ApB), _p(AmB), st)
        del sub
    elif op.lc:
        call("dftb_restricted_ab", _p(ip), _p(ap), M, op.nat, op.opad, op.vpad, _p(op.omega), _p(Kc),
             _p(op.q["oo"]), _p(op.q["ov"]), _p(op.q["qf_ov"]), _p(op.q["qf_vv"]), 1, _p(ApB), _p(AmB), st)
    else:
        call("dftb_restricted_ab", _p(ip), _p(ap), M, op.nat, op.opad, op.vpad, _p(op.omega), _p(Kc),
             None, None, None, None, 0, _p(ApB), _p(AmB), st)
    # restricted Casida: S = (A-B)^{1/2}, R = S (A+B) S, lowest k roots; X+Y = S F / sqrt(Omega)
    ws = Workspace(dev)
    S = torch.empty_like(AmB)
    Si = torch.empty_like(AmB)
    it = torch.zeros(1, dtype=torch.int32)
    er = torch.zeros(1, dtype=torch.float64)
    w, wn = ws.get(_lib.load().dftb_sqrtm_workspace(M))
    call("dftb_sqrtm", _p(AmB), M, M, _p(S), _p(Si), ctypes.c_void_p(it.data_ptr()), ctypes.c_void_p(er.data_ptr()),
         _p(w), wn, st)
    T1 = torch.empty_like(ApB)
    gemm(ApB, S, T1, M, M, M, M, M, M)
    R = torch.empty_like(ApB)
    gemm(S, T1, R, M, M, M, M, M, M)
    kk = min(k, M)
    w2 = torch.empty(kk, dtype=torch.float64, device=dev)
    Fv = torch.empty((M, kk), dtype=torch.float64, device=dev)
    w, wn = ws.get(_lib.load().dftb_syevx_workspace(M, kk))
    call("dftb_syevx", _p(R), M, M, kk, _p(w2), _p(Fv), _p(w), wn, st)
    # X+Y = S F / sqrt(Omega) and X-Y = S^{-1} F sqrt(Omega) of the restricted roots: both
    # directions seed the non-Hermitian subspace (as the warm start [X-Y; X+Y] does)
    XpY = torch.empty((M, kk), dtype=torch.float64, device=dev)
    gemm(S, Fv, XpY, M, kk, M, M, kk, kk)
    XmY = torch.empty((M, kk), dtype=torch.float64, device=dev)
    gemm(Si, Fv, XmY, M, kk, M, M, kk, kk)
    V = torch.zeros((2 * kk, op.N), dtype=torch.float64, device=dev)
    cols = torch.as_tensor(sel, device=dev)
    V[:kk, cols] = XmY.T
    V[kk:, cols] = XpY.T
    prec_omega = _casida_diagonal(op, ints)
    if use_shift:
        prec_omega = prec_omega + Y[:, 1].to(torch.float64) * mask.to(torch.float64)
    return V, prec_omega.contiguous(), torch.sqrt(w2.clamp_min(0))
