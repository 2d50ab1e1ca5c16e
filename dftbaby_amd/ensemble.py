"""Batched excited states of a molecular ensemble in one persistent launch.

The ensemble caller of the reference is a loop over geometries, each running
``LR_TDDFTB.getEnergies`` (e.g. ``DFTB/PES.py:63-118`` for a swarm of
surface-hopping trajectories).  Here all molecules are solved together: the
transition charges of every molecule are built on the GPU, then one launch of
``dftb_ensemble_solve`` runs, per workgroup, the whole non-Hermitian Davidson
loop (Solver.py:637-917) of one molecule, largest molecules first.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from ._lib import call
from .engine import CasidaOperatorGPU, ExcitedStatesError, ExcitedStatesNotConverged, _p, _stream

EnsMolecule = _lib.EnsMolecule

ENS_MAXL = 1024
ENS_MAXK = 32


class EnsembleResult:
    def __init__(self):
        self.Omega = []
        self.XpY = []
        self.XmY = []
        self.C = []
        self.iterations = []
        self.vectors = []      # expansion vectors (one product pair each) per molecule
        self.resid = []
        self.t_build = 0.0
        self.t_solve = 0.0


def build_operators(ground_states, nr_active_occ=None, nr_active_virt=None):
    """One CasidaOperatorGPU per ground state (transition charges on the GPU)."""
    ops = []
    for gs in ground_states:
        f, orbe = gs.getOccupation(), gs.getKSEnergies()
        occ = np.where(f > 0.1)[0]
        virt = np.where(f <= 0.1)[0]
        ao = occ[np.argsort(orbe[occ], kind="stable")[::-1][:nr_active_occ][::-1]]
        av = virt[np.argsort(orbe[virt], kind="stable")[:nr_active_virt]]
        C = gs.getKSCoefficients()
        omega = orbe[av][None, :] - orbe[ao][:, None]
        lc = int(getattr(gs, "long_range_correction", 1))
        g = gs.getGamma()
        glr = gs.getGamma_lr() if lc else 0.0 * g
        ops.append(CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega,
                                                   g, glr, lc, exchange="charges"))
    return ops


def solve_operators(ops, nstates, conv=1.0e-6, maxiter=50, ifact=1, precision=0):
    """Solve the lowest nstates roots of every operator in one launch.

    precision 0 = FP64, 1 = FP32 exchange contractions (FP64 elsewhere).
    Returns EnsembleResult with per-molecule numpy arrays in input order.
    """
    if not ops:
        return EnsembleResult()
    k = int(nstates)
    if k > ENS_MAXK:
        raise ValueError("at most %d states per molecule in the ensemble solver" % ENS_MAXK)
    dev = ops[0].device
    st = _stream()
    t0 = time.time()
    order = sorted(range(len(ops)), key=lambda i: -ops[i].N * ops[i].nat)
    descs = (EnsMolecule * len(ops))()
    keep = []      # tensors that must outlive the launch
    outs = []
    for slot, i in enumerate(order):
        op = ops[i]
        if not op.lc:
            raise ValueError("the ensemble solver handles long_range_correction=1 (non-Hermitian) molecules")
        kmax = op.nocc * op.nvirt
        kk = min(k, kmax)
        ng = min(ifact * kk, kmax)
        cap = int(min(kmax, ng + 2 * kk * maxiter, ENS_MAXL))
        flat = np.ravel(op.omega_real[:op.nocc, :op.nvirt].cpu().numpy())
        idx = np.argsort(flat, kind="stable")[:ng]
        guess = torch.as_tensor((idx // op.nvirt) * op.vpad + idx % op.nvirt, dtype=torch.int64, device=dev)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        bs, bp, bm = z(cap, op.N), z(cap, op.N), z(cap, op.N)
        Mp, Mm = z(cap, cap), z(cap, cap)
        w, C, XmY, XpY, resid = z(kk), z(kk, op.N), z(kk, op.N), z(kk, op.N), z(kk)
        status = torch.zeros(3, dtype=torch.int32, device=dev)
        d = descs[slot]
        d.nat, d.nocc, d.nvirt, d.opad, d.vpad = op.nat, op.nocc, op.nvirt, op.opad, op.vpad
        d.k, d.cap, d.nguess, d.lc, d.precision = kk, cap, ng, 1, int(precision)
        d.omega, d.gamma, d.q_c = op.omega.data_ptr(), op.gamma.data_ptr(), op.q["c"].data_ptr()
        d.q_oo, d.q_ov = op.q["oo"].data_ptr(), op.q["ov"].data_ptr()
        d.qf_ov, d.qf_vv = op.q["qf_ov"].data_ptr(), op.q["qf_vv"].data_ptr()
        if precision == 1:
            op.prepare(1)
            cp = op.qx[1]
            for j, key in enumerate(("oo", "ov", "qf_ov", "qf_vv")):
                d.q32[j] = cp[key][0].data_ptr()
        d.guess = guess.data_ptr()
        d.bs, d.bp, d.bm, d.Mp, d.Mm = bs.data_ptr(), bp.data_ptr(), bm.data_ptr(), Mp.data_ptr(), Mm.data_ptr()
        nb = _lib.load().dftb_ensemble_workspace(ctypes.byref(d))
        work = torch.empty(int(nb), dtype=torch.uint8, device=dev)
        d.work = work.data_ptr()
        d.w, d.C, d.XmY, d.XpY, d.resid = w.data_ptr(), C.data_ptr(), XmY.data_ptr(), XpY.data_ptr(), resid.data_ptr()
        d.status = status.data_ptr()
        keep.extend([guess, bs, bp, bm, Mp, Mm, work])
        outs.append((i, w, C, XmY, XpY, resid, status))
    raw = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).to(dev)
    torch.cuda.synchronize()
    res = EnsembleResult()
    res.t_build = time.time() - t0
    t1 = time.time()
    call("dftb_ensemble_solve", _p(raw), len(ops), ctypes.c_double(conv), int(maxiter), st)
    torch.cuda.synchronize()
    res.t_solve = time.time() - t1
    per = [None] * len(ops)
    for (i, w, C, XmY, XpY, resid, status) in outs:
        it, code, l = status.cpu().numpy().tolist()
        op = ops[i]
        if code == 0:
            raise ExcitedStatesNotConverged("ensemble member %d did not converge after %d iterations" % (i, it))
        if code != 1:
            raise ExcitedStatesError("ensemble member %d failed (code %d after %d iterations)" % (i, code, it))
        per[i] = (w.cpu().numpy(), op.from_padded(C).cpu().numpy(), op.from_padded(XmY).cpu().numpy(),
                  op.from_padded(XpY).cpu().numpy(), it, resid.cpu().numpy(), l)
    for (w, C, XmY, XpY, it, r, l) in per:
        res.Omega.append(w)
        res.C.append(C)
        res.XmY.append(XmY)
        res.XpY.append(XpY)
        res.iterations.append(it)
        res.vectors.append(l)
        res.resid.append(r)
    del keep
    return res
