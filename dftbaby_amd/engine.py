"""GPU excited-state engine: operator set-up and the Davidson drivers.

The host loop is Python; every numeric pass over device data is a HIP kernel
of libdftbaby_amd (transition charges, MFMA GEMMs, matvecs, residuals,
orthonormalisation, the subspace eigensolver).  PyTorch only allocates device
memory, provides the stream and moves/reshapes data (slicing, padding,
gathers).

Algorithms follow DFTB/Solver.py:
  * HermitianDavidson   (Solver.py:404-578): long_range_correction = 0
  * nonHermitianDavidson (Solver.py:637-917): long_range_correction = 1
with these deliberate differences (same converged results):
  * products of the expansion vectors are computed once, when a vector enters
    the subspace, and reused for the subspace matrices and the residuals (the
    reference recomputes (A+-B) on all l vectors and on R, L every iteration);
  * (A+B)v and (A-B)v are formed together, sharing the exchange contractions;
  * sqrtm(Mm) is a coupled Newton-Schulz iteration on MFMA GEMMs and only
    the lowest nstates eigenpairs of Mh are computed;
  * new vectors are orthonormalised by two passes of block Gram-Schmidt plus
    Cholesky-QR instead of a Householder QR of [bs | Q] (same span); vectors
    that are numerically dependent are dropped rather than kept;
  * the initial guess uses a stable sort of the orbital-energy differences.
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from ._lib import GemmArgs, call


# Optional progress hook: called as PROGRESS(text) once per Davidson iteration
# (bench.py points it at stderr so long solves show they are alive).
PROGRESS = None


class ExcitedStatesNotConverged(Exception):
    pass


class ExcitedStatesError(Exception):
    pass


def pad8(n):
    return (int(n) + 7) // 8 * 8


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("dftbaby_amd needs a ROCm GPU (torch.cuda is not available)")
    return torch.device("cuda", torch.cuda.current_device())


class Workspace:
    """A growable device scratch buffer handed to C-ABI calls."""

    def __init__(self, device):
        self.device = device
        self.buf = torch.empty(0, dtype=torch.uint8, device=device)

    def get(self, nbytes):
        nbytes = (int(nbytes) + 255) // 256 * 256
        if self.buf.numel() < nbytes:
            self.buf = torch.empty((int(nbytes * 1.1) + 4096) // 256 * 256, dtype=torch.uint8, device=self.device)
        return self.buf, self.buf.numel()


def gemm(A, B, C, M, N, K, lda, ldb, ldc, a_trans=0, b_trans=0, alpha=1.0, beta=0.0, sA=0, sB=0, sC=0,
         batch=1, in_type=_lib.GEMM_IN_F64, out_type=_lib.GEMM_OUT_F64, ksplit=1, kchunk=0, sCk=0, tile=0,
         A_off=0, B_off=0, C_off=0):
    """C = alpha op(A) op(B) + beta C on MFMA (all FP64 unless in/out types say otherwise)."""
    g = GemmArgs()
    esz_in = {0: 8, 1: 4, 2: 2}[in_type]
    esz_out = {0: 8, 1: 4, 2: 2}[out_type]
    g.A = A.data_ptr() + A_off * esz_in
    g.B = B.data_ptr() + B_off * esz_in
    g.C = C.data_ptr() + C_off * esz_out
    g.lda, g.ldb, g.ldc = lda, ldb, ldc
    g.sA, g.sB, g.sC = sA, sB, sC
    g.cdiv, g.cstr, g.cs2 = 0, 0, 1
    g.M, g.N, g.K = int(M), int(N), int(K)
    g.batch = batch
    g.ksplit, g.kchunk, g.sCk = ksplit, kchunk, sCk
    g.alpha, g.beta = alpha, beta
    g.a_trans, g.b_trans = int(a_trans), int(b_trans)
    g.in_type, g.out_type = in_type, out_type
    g.tile = tile
    call("dftb_gemm", ctypes.byref(g), _stream())


def gram(X, P, Y, Q, n, out, ws, ldx=None, ldy=None, ldo=None, x_off=0, y_off=0, o_off=0):
    """out[p][q] = sum_e X[p][e] Y[q][e]  (tall-skinny, split-K and an in-order
    on-device reduction: dftb_gram)."""
    ldx = ldx or n
    ldy = ldy or n
    ldo = ldo or Q
    if P == 0 or Q == 0:
        return
    w, wn = ws.get(_lib.load().dftb_gram_workspace(int(P), int(Q), int(n)))
    call("dftb_gram", ctypes.c_void_p(X.data_ptr() + 8 * x_off), int(ldx), int(P),
         ctypes.c_void_p(Y.data_ptr() + 8 * y_off), int(ldy), int(Q), int(n),
         ctypes.c_void_p(out.data_ptr() + 8 * o_off), int(ldo), _p(w), wn, _stream())


def row_norms(X, k, n, norms, ws):
    """norms[j] = ||X[j]||_2 for the k rows of X (length n)."""
    w, wn = ws.get(_lib.load().dftb_norms_workspace(int(k)))
    call("dftb_row_norms", _p(X), int(k), int(n), _p(norms), _p(w), wn, _stream())


def residual(P, X, w, k, n, W, norms, ws):
    """W[j] = P[j] - w[j] X[j] and norms[j] = ||W[j]|| (j < k)."""
    wsb, wn = ws.get(_lib.load().dftb_norms_workspace(int(k)))
    call("dftb_residual", _p(P), _p(X), _p(w), int(k), int(n), _p(W), _p(norms), _p(wsb), wn, _stream())


def gemm_timing(on):
    """Start (and clear) or stop the per-launch event timing of the GEMM kernels."""
    call("dftb_gemm_timing_enable", int(bool(on)))


GEMM_CLASSES = {"f64_128": 0, "f64_64": 1, "f32_128": 2, "f32_64": 3, "bf16x3_pipe": 4, "bf16x3_tile": 5,
                "f64_128_t": 6}


def gemm_timing_read(cls):
    """(device seconds, launches, algorithmic flops) recorded for one kernel class."""
    ms, n, fl = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
    call("dftb_gemm_timing_read", GEMM_CLASSES[cls], ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl))
    return ms.value * 1e-3, int(n.value), fl.value


class CasidaOperatorGPU:
    """Device-resident operator of the LR-TD-DFTB problem for one geometry.

    Holds the transition charges in the engine layouts (see include/dftbaby_amd.h),
    the gamma_lr-folded charges, omega and gamma, plus the FP32 / BF16 operand
    copies for the mixed-precision modes.
    """

    def __init__(self, nat, nocc, nvirt, omega, gamma, gamma_lr, lc, device=None):
        self.device = device or _device()
        self.nat, self.nocc, self.nvirt = int(nat), int(nocc), int(nvirt)
        self.opad, self.vpad = pad8(nocc), pad8(nvirt)
        self.N = self.opad * self.vpad
        self.lc = int(lc)
        dev = self.device
        om = torch.ones((self.opad, self.vpad), dtype=torch.float64, device=dev)
        om[:nocc, :nvirt] = (omega.to(dev, torch.float64) if torch.is_tensor(omega)
                             else torch.as_tensor(np.asarray(omega, dtype=np.float64), device=dev))
        self.omega = om.reshape(-1).contiguous()
        self.omega_real = torch.zeros_like(om)
        self.omega_real[:nocc, :nvirt] = om[:nocc, :nvirt]
        g = lambda a: (a.to(dev, torch.float64).contiguous() if torch.is_tensor(a)
                       else torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev).contiguous())
        self.gamma = g(gamma)
        self.gamma_lr = g(gamma_lr)
        self.ws = Workspace(dev)
        self.q = {}          # FP64 engine-layout charges
        self.qx = {}         # low-precision operand copies keyed by precision
        self.qii = self.qaa = None
        self.exchange = "charges"
        self.ao = None
        self.npad = 0
        self.orbs = None
        self.tuning = {}     # precision -> {"atom_chunk", "tile_expand", "tile_contract"} (autotune.py)

    # -- construction ---------------------------------------------------------
    @classmethod
    def from_orbitals(cls, orbs_occ, orbs_virt, S, orbsPerAtom, omega, gamma, gamma_lr, lc, device=None,
                      exchange="ao"):
        """Transition charges on the GPU (kernel (a)) straight into the engine layouts.

        exchange="ao": the long-range exchange products run in the AO basis
        (no q_oo / q_vv / folded charges are stored); "charges": through the
        gamma_lr-folded transition charges (the reference's contraction)."""
        if exchange not in ("ao", "charges"):
            raise ValueError("exchange must be 'ao' or 'charges'")
        opa = np.ascontiguousarray(np.asarray(orbsPerAtom, dtype=np.int32))
        nat = len(opa)
        self = cls(nat, orbs_occ.shape[1], orbs_virt.shape[1], omega, gamma, gamma_lr, lc, device)
        dev = self.device
        t = lambda a: (a.to(dev, torch.float64).contiguous() if torch.is_tensor(a)
                       else torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev))
        Co, Cv, St = t(orbs_occ), t(orbs_virt), t(S)
        norb = St.shape[0]
        o, v, op, vp = self.nocc, self.nvirt, self.opad, self.vpad
        self.exchange = exchange if self.lc else "charges"
        self.orbs = (Co, Cv, St, opa)
        z = lambda *s: torch.empty(s, dtype=torch.float64, device=dev)
        ao = self.exchange == "ao"
        q_c = z(nat, op, vp)
        q_oo = None if ao else z(o, nat, op)
        q_ov = z(o, nat, vp) if (self.lc and not ao) else None
        q_vv = None if ao else z(vp, nat, vp)
        nb = _lib.load().dftb_trans_charges_workspace(norb, o, v, nat)
        w, wn = self.ws.get(nb)
        call("dftb_trans_charges_engine", _p(Co), _p(Cv), _p(St),
             opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), nat, norb, o, v,
             _p(q_oo), _p(q_ov), _p(q_vv), _p(q_c), _p(w), wn, _stream())
        self.q["c"] = q_c
        if ao:
            self._build_ao(Co, Cv, St, opa)
        else:
            self._diagonals_from_engine(q_c, q_vv)
            self.qii = q_oo[:, :, :o].diagonal(dim1=0, dim2=2).T.contiguous()   # q^{ii}_A as [o][nat]
            if self.lc:
                self.q["oo"], self.q["ov"] = q_oo, q_ov
                self._fold(q_vv)
        del q_vv, q_oo
        return self

    def _build_ao(self, Co, Cv, St, opa):
        """AO-basis exchange operands ([Y_v; C_v], [Y_o; C_o], [C_v; Y_v]^T, [C_o; Y_o]^T with Y = S C,
        gamma_lr expanded to orbital pairs) and the diagonal charges q^{ii}_A, q^{aa}_A."""
        dev = self.device
        norb = St.shape[0]
        o, v, op, vp, nat = self.nocc, self.nvirt, self.opad, self.vpad, self.nat
        npad = pad8(norb)
        Yo = torch.empty((norb, o), dtype=torch.float64, device=dev)
        Yv = torch.empty((norb, v), dtype=torch.float64, device=dev)
        gemm(St, Co, Yo, norb, o, norb, norb, o, o)
        gemm(St, Cv, Yv, norb, v, norb, norb, v, v)
        zz = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        ao_v, ao_o = zz(2 * npad, vp), zz(2 * npad, op)
        ao_v[:norb, :v], ao_v[npad:npad + norb, :v] = Yv, Cv
        ao_o[:norb, :o], ao_o[npad:npad + norb, :o] = Yo, Co
        ao_cyvt, ao_w = zz(vp, 2 * npad), zz(op, 2 * npad)
        ao_cyvt[:v, :norb], ao_cyvt[:v, npad:npad + norb] = Cv.T, Yv.T
        ao_w[:o, :norb], ao_w[:o, npad:npad + norb] = Co.T, Yo.T
        amap = torch.as_tensor(np.repeat(np.arange(nat), opa), device=dev)
        # gamma_lr with a zero row/column for the padded orbitals, indexed through the atom map
        gam = zz(nat + 1, nat + 1)
        if self.lc:
            gam[:nat, :nat] = self.gamma_lr
        amap_pad = torch.full((npad,), nat, dtype=torch.int32, device=dev)
        amap_pad[:norb] = amap.to(torch.int32)
        self.npad = npad
        self.ao = {"v": ao_v, "o": ao_o, "cyvt": ao_cyvt, "w": ao_w, "gam": gam.contiguous(), "amap": amap_pad}
        # q^{ii}_A = sum_{mu in A} C_mu,i Y_mu,i ; q^{aa}_A likewise  ->  [o][nat], [v][nat]
        self.qii = zz(nat, o).index_add_(0, amap, Co * Yo).T.contiguous()
        self.qaa = zz(nat, v).index_add_(0, amap, Cv * Yv).T.contiguous()

    @classmethod
    def from_charges(cls, q_oo, q_vv, q_ov, omega, gamma, gamma_lr, lc, device=None):
        """Operator from charges in the reference layout (Nat, o, o), (Nat, v, v), (Nat, o, v)."""
        dev = device or _device()
        t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev) if not torch.is_tensor(a) \
            else a.to(dev, torch.float64)
        qov = t(q_ov)
        nat, o, v = qov.shape
        self = cls(nat, o, v, omega, gamma, gamma_lr, lc, dev)
        op, vp = self.opad, self.vpad
        q_c = torch.zeros((nat, op, vp), dtype=torch.float64, device=dev)
        q_c[:, :o, :v] = qov
        self.q["c"] = q_c
        qoo = t(q_oo) if q_oo is not None else None
        qvv = t(q_vv) if q_vv is not None else None
        if qoo is not None and qvv is not None:
            self.qii = torch.diagonal(qoo, dim1=1, dim2=2).T.contiguous()
            self.qaa = torch.diagonal(qvv, dim1=1, dim2=2).T.contiguous()
        if self.lc:
            e_oo = torch.zeros((o, nat, op), dtype=torch.float64, device=dev)
            e_oo[:, :, :o] = qoo.permute(1, 0, 2)
            e_ov = torch.zeros((o, nat, vp), dtype=torch.float64, device=dev)
            e_ov[:, :, :v] = qov.permute(1, 0, 2)
            e_vv = torch.zeros((vp, nat, vp), dtype=torch.float64, device=dev)
            e_vv[:v, :, :v] = qvv.permute(1, 0, 2)
            self.q["oo"], self.q["ov"] = e_oo, e_ov
            self._fold(e_vv)
        return self

    def _diagonals_from_engine(self, q_c, q_vv):
        v = self.nvirt
        self.qaa = q_vv[:v, :, :v].diagonal(dim1=0, dim2=2).T.contiguous()   # q^{aa}_A as [v][nat]

    def _fold(self, q_vv_e):
        o, nat, vp = self.nocc, self.nat, self.vpad
        qf_ov = torch.empty((o, nat, vp), dtype=torch.float64, device=self.device)
        call("dftb_fold_gamma", _p(self.gamma_lr), nat, _p(self.q["ov"]), o, vp, _p(qf_ov), _stream())
        qf_vv = torch.empty((vp, nat, vp), dtype=torch.float64, device=self.device)
        call("dftb_fold_gamma", _p(self.gamma_lr), nat, _p(q_vv_e), vp, vp, _p(qf_vv), _stream())
        self.q["qf_ov"], self.q["qf_vv"] = qf_ov, qf_vv

    def diag_oo(self):
        """q^{ii}_A as [o][nat] (for Lambda2)."""
        if self.qii is not None:
            return self.qii
        if "oo" in self.q:
            o = self.nocc
            self.qii = self.q["oo"][:, :, :o].diagonal(dim1=0, dim2=2).T.contiguous()
            return self.qii
        return None

    def prepare(self, precision):
        """Create the operand copies a precision mode needs (0 fp64, 1 fp32, 2 bf16x3)."""
        if precision == 0 or precision in self.qx:
            return
        dev = self.device
        cp = {}
        qc32 = torch.empty(self.q["c"].shape, dtype=torch.float32, device=dev)
        call("dftb_convert_f32", _p(self.q["c"]), self.q["c"].numel(), _p(qc32), _stream())
        cp["c"] = qc32
        if self.lc and self.exchange == "ao":
            for key in ("v", "o", "cyvt", "w"):
                x = self.ao[key]
                if precision == 1:
                    y = torch.empty(x.shape, dtype=torch.float32, device=dev)
                    call("dftb_convert_f32", _p(x), x.numel(), _p(y), _stream())
                    cp["ao_" + key] = (y, None)
                else:
                    hi = torch.empty(x.shape, dtype=torch.int16, device=dev)
                    lo = torch.empty(x.shape, dtype=torch.int16, device=dev)
                    call("dftb_convert_bf16pair", _p(x), x.numel(), _p(hi), _p(lo), _stream())
                    cp["ao_" + key] = (hi, lo)
        elif self.lc:
            if precision == 1:
                for key in ("oo", "ov", "qf_ov", "qf_vv"):
                    x = self.q[key]
                    y = torch.empty(x.shape, dtype=torch.float32, device=dev)
                    call("dftb_convert_f32", _p(x), x.numel(), _p(y), _stream())
                    cp[key] = (y, None)
            else:
                for key in ("oo", "ov"):
                    x = self.q[key]
                    hi = torch.empty(x.shape, dtype=torch.int16, device=dev)
                    lo = torch.empty(x.shape, dtype=torch.int16, device=dev)
                    call("dftb_convert_bf16pair", _p(x), x.numel(), _p(hi), _p(lo), _stream())
                    cp[key] = (hi, lo)
                nat, op, vp, o = self.nat, self.opad, self.vpad, self.nocc
                hi = torch.empty((nat, vp, op), dtype=torch.int16, device=dev)
                lo = torch.empty((nat, vp, op), dtype=torch.int16, device=dev)
                call("dftb_convert_bf16pair_atom_major", _p(self.q["qf_ov"]), o, op, nat, vp, _p(hi), _p(lo),
                     _stream())
                cp["qf_ov"] = (hi, lo)
                hi = torch.empty((nat, vp, vp), dtype=torch.int16, device=dev)
                lo = torch.empty((nat, vp, vp), dtype=torch.int16, device=dev)
                call("dftb_convert_bf16pair_atom_major", _p(self.q["qf_vv"]), vp, vp, nat, vp, _p(hi), _p(lo),
                     _stream())
                cp["qf_vv"] = (hi, lo)
        self.qx[precision] = cp

    def _struct(self, precision, atom_chunk=0, tiles=None, atom_limit=0):
        s = _lib.CasidaOperator()
        s.nat, s.nocc, s.nvirt, s.opad, s.vpad = self.nat, self.nocc, self.nvirt, self.opad, self.vpad
        tuned = self.tuning.get(precision, {})
        s.lc, s.precision = self.lc, precision
        s.atom_chunk = atom_chunk or tuned.get("atom_chunk", 0)
        te, tc = tiles if tiles is not None else (tuned.get("tile_expand", 0), tuned.get("tile_contract", 0))
        s.tile_expand, s.tile_contract, s.atom_limit = te, tc, atom_limit
        s.omega = self.omega.data_ptr()
        s.gamma = self.gamma.data_ptr()
        if self.lc and self.exchange == "ao":
            s.exchange, s.npad = 1, self.npad
            s.gam_ao = self.ao["gam"].data_ptr()
            s.amap = self.ao["amap"].data_ptr()
            if precision == 0:
                s.q_c = self.q["c"].data_ptr()
                for key in ("v", "o", "cyvt", "w"):
                    getattr(s, "ao_" + key)[0] = self.ao[key].data_ptr()
            else:
                self.prepare(precision)
                cp = self.qx[precision]
                s.q_c = cp["c"].data_ptr()
                for key in ("v", "o", "cyvt", "w"):
                    hi, lo = cp["ao_" + key]
                    getattr(s, "ao_" + key)[0] = hi.data_ptr()
                    getattr(s, "ao_" + key)[1] = lo.data_ptr() if lo is not None else 0
            return s
        if precision == 0:
            s.q_c = self.q["c"].data_ptr()
            if self.lc:
                for key, fld in (("oo", "q_oo"), ("ov", "q_ov"), ("qf_ov", "qf_ov"), ("qf_vv", "qf_vv")):
                    getattr(s, fld)[0] = self.q[key].data_ptr()
        else:
            self.prepare(precision)
            cp = self.qx[precision]
            s.q_c = cp["c"].data_ptr()
            if self.lc:
                for key, fld in (("oo", "q_oo"), ("ov", "q_ov"), ("qf_ov", "qf_ov"), ("qf_vv", "qf_vv")):
                    hi, lo = cp[key]
                    getattr(s, fld)[0] = hi.data_ptr()
                    getattr(s, fld)[1] = lo.data_ptr() if lo is not None else 0
        return s

    # -- products ---------------------------------------------------------------
    def matvec(self, V, precision=0, out_p=None, out_m=None, atom_chunk=0, tiles=None, atom_limit=0):
        """((A+B)V, (A-B)V) for V [m][opad*vpad] (FP64, zero padding).  With the AO
        exchange, blocks of at most AO_COULOMB_MAX vectors (the Z-vector solve) take
        the Coulomb term from the AO operands in FP64 (_coulomb_ao) instead of two
        passes over the [nat][N] charges."""
        m = V.shape[0]
        if out_p is None:
            out_p = torch.empty((m, self.N), dtype=torch.float64, device=self.device)
        if out_m is None:
            out_m = torch.empty((m, self.N), dtype=torch.float64, device=self.device)
        if m == 0:
            return out_p, out_m
        s = self._struct(precision, atom_chunk, tiles, atom_limit)
        ao_coulomb = m <= self.AO_COULOMB_MAX and self.lc and self.exchange == "ao" and atom_limit == 0
        if ao_coulomb:
            s.q_c = None                      # the library leaves the Coulomb term out
        nb = _lib.load().dftb_casida_matvec_workspace(ctypes.byref(s), m)
        w, wn = self.ws.get(nb)
        call("dftb_casida_matvec", ctypes.byref(s), _p(V), m, _p(out_p), _p(out_m), _p(w), wn, _stream())
        if ao_coulomb:
            self._coulomb_ao(V, out_p)
        return out_p, out_m

    AO_COULOMB_MAX = 2

    def _coulomb_ao(self, V, out_p):
        """out_p += 4 q^T gamma q V through the AO operands, with the transition charges
        q^{ia}_A = 1/2 sum_{mu in A} (C_o,mu,i Y_v,mu,a + Y_o,mu,i C_v,mu,a), Y = S C:
        u_A = 1/2 sum_{mu in A} rowdot([C_o; Y_o] V, [Y_v; C_v])_mu, g = 4 gamma u,
        W = 1/2 [Y_o; C_o]^T [g C_v; g Y_v] -- two FP64 GEMMs of 2 npad x opad x vpad
        per vector instead of streaming the charges twice."""
        m = V.shape[0]
        npad, nat, op_, vp = self.npad, self.nat, self.opad, self.vpad
        dev = self.device
        if getattr(self, "_ao_cv", None) is None:
            self._ao_cv = self.ao["cyvt"].T.contiguous()            # [C_v; Y_v]
            self._amap_l = self.ao["amap"].to(torch.int64)
        X = torch.empty((m, 2 * npad, vp), dtype=torch.float64, device=dev)
        gemm(self.ao["w"], V, X, 2 * npad, vp, op_, 2 * npad, vp, vp, a_trans=1, batch=m, sB=self.N,
             sC=2 * npad * vp)                                     # [C_o; Y_o] V
        r = (X * self.ao["v"][None]).sum(2)                        # rowdot with [Y_v; C_v]
        r = 0.5 * (r[:, :npad] + r[:, npad:])
        u = torch.zeros((m, nat + 1), dtype=torch.float64, device=dev).index_add_(1, self._amap_l, r)
        g = torch.zeros((m, nat + 1), dtype=torch.float64, device=dev)
        g[:, :nat] = 4.0 * (u[:, :nat] @ self.gamma)
        gmu = g[:, self._amap_l]                                   # per orbital, padding 0
        B = (torch.cat([gmu, gmu], 1)[:, :, None] * self._ao_cv[None]).contiguous()
        gemm(self.ao["o"], B, out_p, op_, vp, 2 * npad, op_, vp, vp, a_trans=1, alpha=0.5, beta=1.0, batch=m,
             sB=2 * npad * vp, sC=self.N)

    def hermitian_matvec(self, V, sqrt_omega, om, precision=0, out=None):
        m = V.shape[0]
        if out is None:
            out = torch.empty((m, self.N), dtype=torch.float64, device=self.device)
        if m == 0:
            return out
        s = self._struct(min(precision, 1))
        nb = _lib.load().dftb_hermitian_matvec_workspace(ctypes.byref(s), m)
        w, wn = self.ws.get(nb)
        call("dftb_hermitian_matvec", ctypes.byref(s), _p(sqrt_omega), _p(om), _p(V), m, _p(out), _p(w), wn,
             _stream())
        return out

    # -- layout helpers -----------------------------------------------------------
    def to_padded(self, X):
        """(k, o, v) -> [k][opad*vpad] FP64 on the device."""
        X = torch.as_tensor(np.asarray(X, dtype=np.float64)) if not torch.is_tensor(X) else X
        k = X.shape[0]
        out = torch.zeros((k, self.opad, self.vpad), dtype=torch.float64, device=self.device)
        out[:, :self.nocc, :self.nvirt] = X.to(self.device, torch.float64)
        return out.reshape(k, self.N)

    def from_padded(self, X):
        k = X.shape[0]
        return X.reshape(k, self.opad, self.vpad)[:, :self.nocc, :self.nvirt]

    def flat_index(self, i, a):
        return i * self.vpad + a


class _Subspace:
    """Expansion vectors and their products, preallocated on the device.  The
    capacity grows on demand (the filled rows and the l x l projected blocks are
    copied) up to `limit`, the memory-derived bound."""

    def __init__(self, N, cap, device, nprod, limit=None):
        self.N, self.device = N, device
        self.cap = cap
        self.limit = max(cap, limit or cap)
        self.bs = torch.zeros((cap, N), dtype=torch.float64, device=device)
        self.prod = [torch.zeros((cap, N), dtype=torch.float64, device=device) for _ in range(nprod)]
        self.M = [torch.zeros((cap, cap), dtype=torch.float64, device=device) for _ in range(nprod)]
        self.l = 0

    def grow(self, l, need):
        """Room for `need` vectors, keeping rows [0, l); False beyond the limit.
        Callers drop their own references to the arrays first, so one array at a
        time is held twice."""
        if need <= self.cap:
            return True
        if need > self.limit:
            return False
        new = min(self.limit, max(_round_cap(need), 2 * self.cap))
        dev, N = self.device, self.N

        def rows(old):
            t = torch.zeros((new, N), dtype=torch.float64, device=dev)
            t[:l].copy_(old[:l])
            return t

        def square(old):
            t = torch.zeros((new, new), dtype=torch.float64, device=dev)
            t[:l, :l].copy_(old[:l, :l])
            return t

        self.bs = rows(self.bs)
        for i in range(len(self.prod)):
            self.prod[i] = rows(self.prod[i])
        for i in range(len(self.M)):
            self.M[i] = square(self.M[i])
        self.cap = new
        return True


def _round_cap(n, q=256):
    return ((int(n) + q - 1) // q) * q


def _subspace_cap(N, l0, k_add, maxiter, nmat, device, frac=0.6):
    """Memory-derived bound on the expansion space (vectors)."""
    want = l0 + k_add * maxiter
    free, total = torch.cuda.mem_get_info(device)
    free += torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)   # torch's cached blocks
    by_mem = int(frac * free / (8.0 * N * nmat))
    return max(l0 + k_add, min(want, by_mem))


def _initial_cap(l0, k_add, limit, iters=16):
    """First allocation: room for `iters` full expansion steps, in multiples of
    256 vectors so that solves of one size reuse torch's cached blocks."""
    return max(min(limit, l0 + k_add), min(limit, _round_cap(l0 + k_add * iters)))


def _initial_indices(values, nocc, nvirt, lmax):
    """Flat (i, a) indices of the lmax smallest values (stable), Solver.py:232-251."""
    flat = np.ravel(np.asarray(values))
    if 0 < lmax < flat.size:
        # partial selection, then a stable sort of the candidates: the same indices
        # and order as a full stable argsort (ties at the threshold by index)
        thr = flat[np.argpartition(flat, lmax - 1)[:lmax]].max()
        less = np.flatnonzero(flat < thr)
        cand = np.concatenate([less, np.flatnonzero(flat == thr)[:lmax - len(less)]])
        idx = cand[np.argsort(flat[cand], kind="stable")]
    else:
        idx = np.argsort(flat, kind="stable")[:lmax]
    return idx // nvirt, idx % nvirt


class DavidsonStats:
    def __init__(self):
        self.iterations = 0
        self.history = []
        self.matvec_vectors = 0
        self.t_total = 0.0
        self.learned = False
        self.refined = False
        self.restarts = 0
        self.l = 0
        self._events = []

    def event_pair(self):
        """Start/stop events on the launch stream around one block matvec; the
        start is recorded now, the caller records the stop."""
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        self._events.append((e0, e1))
        return e0, e1

    @property
    def matvec_time(self):
        """Device time (s) spent in the block matvecs."""
        if self._events:
            self._events[-1][1].synchronize()
        return sum(a.elapsed_time(b) for a, b in self._events) * 1e-3


def _precondition_select(norms_L, norms_R, k, l, kmax, conv):
    """Indices of new expansion vectors, Solver.py:842-884 (non-Hermitian)."""
    dkmax = min(kmax - l, 2 * k)
    eps = 0.01 * conv
    ncL = int(np.sum(norms_L > eps))
    ncR = int(np.sum(norms_R > eps))
    dkR = min(int(np.ceil(dkmax / 2.0)), ncL)
    dkL = min(dkmax - dkR, ncR)
    dk = dkR + dkL
    sel = []
    for n in range(k):
        if len(sel) == dk:
            break
        if norms_L[n] > eps:
            sel.append(n)
    for n in range(k):
        if len(sel) == dk:
            break
        if norms_R[n] > eps:
            sel.append(k + n)
    if len(sel) != dk:
        raise AssertionError("nb = %d vectors were added, but dk = %d vectors should have been added" % (len(sel), dk))
    return sel


def _orthonormalize_block(bs, l, m, N, ws, st):
    """Orthonormalise rows bs[l:l+m] against bs[:l] and among themselves; compact; return new count."""
    keep = torch.empty(m, dtype=torch.int32, device=bs.device)
    nb = _lib.load().dftb_orthonormalize_workspace(l, m, N)
    wsb, wsn = ws.get(nb)
    call("dftb_orthonormalize", _p(bs), l, _p(bs[l:l + m]), m, N, 1.0e-8, _p(keep), _p(wsb), wsn, st)
    kept = np.where(keep.cpu().numpy() != 0)[0]
    if 0 < len(kept) < m:
        Q = bs[l:l + m]
        Q[:len(kept)] = Q[torch.as_tensor(kept, device=bs.device)].clone()
    return l + len(kept)


def _collapse(sub, Rb, Lb, l, k, N, ws, st):
    """Thick restart: replace the basis by the orthonormalised span of the current
    Ritz coordinates [Rb, Lb] (X+Y and X-Y); basis vectors, their stored products
    and the projected matrices follow by the same linear combination (no matvec)."""
    dev = sub.bs.device
    X = torch.cat([Rb[:l, :k].T, Lb[:l, :k].T], 0).contiguous()       # rows: subspace coordinates
    keep = torch.empty(2 * k, dtype=torch.int32, device=dev)
    nb = _lib.load().dftb_orthonormalize_workspace(0, 2 * k, l)
    wsb, wsn = ws.get(nb)
    call("dftb_orthonormalize", None, 0, _p(X), 2 * k, l, 1.0e-8, _p(keep), _p(wsb), wsn, st)
    kept = torch.as_tensor(np.where(keep.cpu().numpy() != 0)[0], device=dev)
    Q = X[kept].contiguous()                                          # [r][l], orthonormal rows
    r = Q.shape[0]
    tmp = torch.empty((r, N), dtype=torch.float64, device=dev)
    for B in [sub.bs] + sub.prod:
        gemm(Q, B, tmp, r, N, l, l, N, N)
        B[:r].copy_(tmp)
    T1 = torch.empty((r, l), dtype=torch.float64, device=dev)
    Mn = torch.empty((r, r), dtype=torch.float64, device=dev)
    for Mx in sub.M:
        gemm(Q, Mx, T1, r, l, l, l, sub.cap, l)
        gemm(T1, Q, Mn, r, r, l, l, l, r, b_trans=1)
        Mx[:r, :r].copy_(Mn)
    return r


def nonhermitian_davidson_gpu(op, nstates, ifact=1, conv=1.0e-14, maxiter=10, XmYguess=None, XpYguess=None,
                              w_guess=None, precision=0, verbose=0, stats=None, initial_space=None,
                              precond_omega=None, refine=True, restart=None):
    """Non-Hermitian Davidson on the GPU (Solver.py:637-917).

    Returns padded device tensors (Omega [k], C, XmY, XpY [k][N]) and stats.
    With precision > 0 and refine, a run whose largest residual stops falling
    (the noise floor of the low-precision products) restarts from its current
    X-Y, X+Y with FP64 products, which converge to conv in a few iterations.
    restart (subspace size; None or 0 = never, as the reference): before the
    basis would exceed it, the basis collapses to span(X+Y, X-Y) of the current
    roots (thick restart), which bounds the subspace eigensolve and the basis
    memory at the price of more iterations (on the graphene flake 19 -> 26 at
    480 vectors), so it is off by default.
    """
    dev = op.device
    stats = stats or DavidsonStats()
    it0 = stats.iterations
    t_start = time.time()
    k = int(nstates)
    o, v, N = op.nocc, op.nvirt, op.N
    kmax = o * v
    st = _stream()
    ws = op.ws
    if initial_space is not None:
        XpYguess = XmYguess = None
        lmax = int(initial_space.shape[0])
    elif XpYguess is None:
        lmax = min(ifact * k, kmax)
    else:
        lmax = k + min(k, kmax - k)
    limit = min(kmax, _subspace_cap(N, lmax, 2 * k, maxiter, 3, dev))
    cap = _initial_cap(lmax, 2 * k, limit)
    sub = _Subspace(N, cap, dev, 2, limit)
    bs, bp, bm = sub.bs, sub.prod[0], sub.prod[1]
    Mp, Mm = sub.M
    if initial_space is not None:
        bs[:lmax] = initial_space
        lmax = _orthonormalize_block(bs, 0, lmax, N, ws, st)
    elif XpYguess is None:
        ii, aa = _initial_indices(op.omega_real[:o, :v].cpu().numpy(), o, v, lmax)
        idx = torch.as_tensor(ii * op.vpad + aa, dtype=torch.int64, device=dev)
        call("dftb_unit_vectors", _p(idx), lmax, N, _p(bs), st)
    else:
        G = torch.cat([op.to_padded(XmYguess), op.to_padded(XpYguess)[:lmax - k]], 0)
        nrm = torch.empty(lmax, dtype=torch.float64, device=dev)
        row_norms(G, lmax, N, nrm, ws)
        inv = 1.0 / nrm
        call("dftb_scale2", _p(G), _p(inv), None, lmax, N, _p(bs), st)
        lmax = _orthonormalize_block(bs, 0, lmax, N, ws, st)
    l = lmax
    new0 = 0
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    diag = torch.zeros(2, dtype=torch.float64, device=dev)
    asym = torch.zeros(3, dtype=torch.float64, device=dev)
    wdev = torch.zeros(k, dtype=torch.float64, device=dev)
    Tb = torch.zeros((cap, k), dtype=torch.float64, device=dev)
    Rb = torch.zeros((cap, k), dtype=torch.float64, device=dev)
    Lb = torch.zeros((cap, k), dtype=torch.float64, device=dev)
    LR = torch.empty((2 * k, N), dtype=torch.float64, device=dev)          # [L; R] = [X-Y; X+Y]
    L, R = LR[:k], LR[k:]
    PM = torch.empty((2 * k, N), dtype=torch.float64, device=dev)          # [(A+B) R; (A-B) L]
    PR, ML = PM[:k], PM[k:]
    T2 = None                                                              # products of a new block
    W = torch.empty((2 * k, N), dtype=torch.float64, device=dev)
    norms = torch.empty(2 * k, dtype=torch.float64, device=dev)
    keep = torch.empty(2 * k, dtype=torch.int32, device=dev)
    w2 = None
    refine_now = False
    best_res, stall = np.inf, 0
    for it in range(maxiter):
        if XpYguess is None or it > 0:
            m = l - new0
            if m > 0:
                if T2 is None or T2.shape[0] < 2 * m:
                    T2 = torch.empty((2 * m, N), dtype=torch.float64, device=dev)
                ev = stats.event_pair()
                op.matvec(bs[new0:l], precision, out_p=T2[:m], out_m=T2[m:2 * m])
                ev[1].record()
                stats.matvec_vectors += m
                # new columns of M+- = bs^T (A+-B) bs from one pass over the basis;
                # the new rows left of them by symmetry of A+-B
                G = torch.empty((l, 2 * m), dtype=torch.float64, device=dev)
                gram(bs, l, T2, 2 * m, N, G, ws)
                Mp[:l, new0:l] = G[:, :m]
                Mm[:l, new0:l] = G[:, m:]
                if new0 > 0:
                    Mp[new0:l, :new0] = Mp[:new0, new0:l].T
                    Mm[new0:l, :new0] = Mm[:new0, new0:l].T
                bp[new0:l].copy_(T2[:m])
                bm[new0:l].copy_(T2[m:2 * m])
            if precision > 0:
                # low-precision products: use the symmetric part of M+-, sanity-check the asymmetry
                for Mx in (Mp, Mm):
                    call("dftb_symmetrize", _p(Mx), l, cap, _p(asym), st)
                    a = float(asym.cpu()[0])
                    if a > 1.0e-3:
                        raise ExcitedStatesError("subspace matrix not symmetric (relative asymmetry %.3e)" % a)
            nb = _lib.load().dftb_subspace_workspace(l, k)
            wsb, wsn = ws.get(nb)
            # roots, Rb and Lb through a Cholesky factor of Mm; Tb (for C) once at the end
            call("dftb_subspace_nonhermitian_rl", _p(Mp), _p(Mm), l, cap, k, _p(wdev), _p(Rb), _p(Lb),
                 _p(info), _p(diag), _p(wsb), wsn, st)
            inf = info.cpu().numpy()
            dg = diag.cpu().numpy()
            if inf[0] != 0:
                raise ExcitedStatesError("subspace matrix of (A-B) is not positive definite (code %d)" % inf[1])
            if dg[0] > 1.0e-10:
                raise ExcitedStatesError("Mh is not hermitian! error = %s" % dg[0])
            if dg[1] > 1.0e-3:
                raise ExcitedStatesError("(X+Y) and (X-Y) vectors not orthonormal! error = %s" % dg[1])
            # R = bs^T Rb, L = bs^T Lb and their products from the stored ones
            LRb = torch.cat([Lb[:l], Rb[:l]], 1)          # one pass over the basis for L and R
            gemm(LRb, bs, LR, 2 * k, N, l, 2 * k, N, N, a_trans=1, tile=64)
            gemm(Rb, bp, PR, k, N, l, k, N, N, a_trans=1, tile=64)
            gemm(Lb, bm, ML, k, N, l, k, N, N, a_trans=1, tile=64)
        else:
            R.copy_(op.to_padded(XpYguess))
            L.copy_(op.to_padded(XmYguess))
            p_, m_ = op.matvec(LR, precision)
            stats.matvec_vectors += 2 * k
            PR.copy_(p_[k:])
            ML.copy_(m_[:k])
            wdev.copy_(torch.as_tensor(np.asarray(w_guess[:k], dtype=np.float64), device=dev))
        # WL = (A+B) R - w L and WR = (A-B) L - w R (Solver.py:809-812) in one pass
        residual(PM, LR, torch.cat([wdev, wdev]), 2 * k, N, W, norms, ws)
        nr = norms.cpu().numpy()
        nL, nR = nr[:k], nr[k:]
        nsum = nL + nR
        w_host = wdev.cpu().numpy()
        stats.history.append((l, nsum.copy(), w_host.copy()))
        if PROGRESS is not None:
            PROGRESS("davidson it %d: l=%d max res %.3e (%.1f s)" % (it, l, float(nsum.max()), time.time() - t_start))
        if verbose > 0:
            print("Iteration %d: ( %d expansion vectors )   time: %2.6f s" % (it, l, time.time() - t_start))
            for n in range(k):
                print("  Eigenvalue %3d      %2.7f    %s   res. norm = %e" %
                      (n + 1, w_host[n], "converged    " if nsum[n] < conv else "not converged", nsum[n]))
        stats.iterations = it0 + it + 1
        if (nsum < conv).all() and it > 0:
            break
        if precision > 0 and refine:
            # the low-precision noise floor only matters close to convergence: stagnation is
            # counted once the largest residual is within 100 x conv (early plateaus are normal)
            r = float(nsum.max())
            if r < 0.7 * best_res or r > max(100.0 * conv, 1e-4):
                best_res, stall = min(best_res, r), 0
            else:
                stall += 1
            if stall >= 3 or it == maxiter - 1:
                refine_now = True
                break
        lrs = int(restart or 0)
        if lrs > 0 and it > 0 and l + 2 * k > lrs and l > 2 * k:
            l = _collapse(sub, Rb, Lb, l, k, N, ws, st)
            new0 = l
            stats.restarts += 1
        sel = _precondition_select(nL, nR, k, l, kmax, conv)
        dk = len(sel)
        if dk == 0:
            continue
        if l + dk > cap:
            bs = bp = bm = Mp = Mm = None
            if not sub.grow(l, l + dk):
                raise ExcitedStatesNotConverged("expansion space exhausted the device memory budget (%d vectors)"
                                                % sub.limit)
            cap = sub.cap
            bs, bp, bm = sub.bs, sub.prod[0], sub.prod[1]
            Mp, Mm = sub.M
            Tb = torch.zeros((cap, k), dtype=torch.float64, device=dev)
            Rb = torch.zeros((cap, k), dtype=torch.float64, device=dev)
            Lb = torch.zeros((cap, k), dtype=torch.float64, device=dev)
        wsel = torch.cat([wdev, wdev])
        seli = torch.as_tensor(sel, dtype=torch.int32, device=dev)
        Q = bs[l:l + dk]
        pom = op.omega if precond_omega is None else precond_omega
        call("dftb_precondition", _p(W), _p(wsel), _p(seli), dk, _p(pom), N, _p(Q), st)
        nb = _lib.load().dftb_orthonormalize_workspace(l, dk, N)
        wsb, wsn = ws.get(nb)
        call("dftb_orthonormalize", _p(bs), l, _p(Q), dk, N, 1.0e-8, _p(keep), _p(wsb), wsn, st)
        kp = keep[:dk].cpu().numpy()
        kept = np.where(kp != 0)[0]
        if len(kept) < dk:
            if len(kept) > 0:
                Q[:len(kept)] = Q[torch.as_tensor(kept, device=dev)].clone()
        # after a warm-started iteration 0 (no products yet) every vector needs its products
        new0 = l if (XpYguess is None or it > 0) else 0
        l = l + len(kept)
    else:
        raise ExcitedStatesNotConverged(
            "Davidson-like diagonalization for non-Hermitian eigenvalue problem did not converge after %s "
            "iterations (threshold: %e) !!!" % (maxiter, conv))
    if refine_now:
        V0 = LR.clone()
        stats.refined = True
        out = nonhermitian_davidson_gpu(op, k, conv=conv, maxiter=max(maxiter - (it + 1), 10), precision=0,
                                        verbose=verbose, stats=stats, initial_space=V0, precond_omega=precond_omega,
                                        refine=False)
        stats.t_total = time.time() - t_start
        return out
    # Tb of the last subspace step (the eigenvectors of Mmsq Mp Mmsq that belong to its
    # Rb), then C = bs Tb (Solver.py:906-910)
    wsb, wsn = ws.get(_lib.load().dftb_subspace_workspace(l, k))
    call("dftb_subspace_tb", _p(Mm), l, cap, k, _p(wdev), _p(Rb), _p(Tb), _p(wsb), wsn, st)
    C = torch.empty((k, N), dtype=torch.float64, device=dev)
    gemm(Tb, bs, C, k, N, l, k, N, N, a_trans=1, tile=64)
    stats.t_total = time.time() - t_start
    stats.l = l
    return wdev[:k].clone(), C, L.clone(), R.clone(), stats


def hermitian_davidson_gpu(op, nstates, omega_shift=None, ifact=1, conv=1.0e-14, maxiter=10, XpYguess=None,
                           Oia=None, L2_thresh=0.0, precision=0, verbose=0, stats=None):
    """Hermitian Davidson on the GPU (Solver.py:404-578), lc = 0."""
    dev = op.device
    stats = stats or DavidsonStats()
    t_start = time.time()
    k = int(nstates)
    o, v, N = op.nocc, op.nvirt, op.N
    kmax = o * v
    st = _stream()
    ws = op.ws
    omega = op.omega
    shift = torch.zeros_like(omega) if omega_shift is None else op.to_padded(np.asarray(omega_shift)[None])[0]
    sq = torch.sqrt(omega)
    sq_inv = 1.0 / sq
    om = omega * omega + 2.0 * omega * shift
    lmax = min(ifact * k, kmax)
    limit = min(kmax, _subspace_cap(N, lmax, k, maxiter, 2, dev))
    cap = _initial_cap(lmax, k, limit)
    sub = _Subspace(N, cap, dev, 1, limit)
    bs, rbs = sub.bs, sub.prod[0]
    H = sub.M[0]
    if XpYguess is None:
        omr = torch.sqrt(om.reshape(op.opad, op.vpad)[:o, :v]).cpu().numpy()
        ii, aa = _initial_indices(omr, o, v, lmax)
        idx = torch.as_tensor(ii * op.vpad + aa, dtype=torch.int64, device=dev)
        call("dftb_unit_vectors", _p(idx), lmax, N, _p(bs), st)
    else:
        G = op.to_padded(np.asarray(XpYguess)[:lmax])
        call("dftb_scale2", _p(G), None, _p(sq_inv), lmax, N, _p(G), st)
        nrm = torch.empty(lmax, dtype=torch.float64, device=dev)
        row_norms(G, lmax, N, nrm, ws)
        call("dftb_scale2", _p(G), _p(1.0 / nrm), None, lmax, N, _p(bs), st)
        lmax = _orthonormalize_block(bs, 0, lmax, N, ws, st)
    l = lmax
    new0 = 0
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    use_l2 = L2_thresh > 0.0 and Oia is not None
    T = W = None
    norms = torch.empty(max(k, 1), dtype=torch.float64, device=dev)
    keep = torch.empty(k, dtype=torch.int32, device=dev)
    for it in range(maxiter):
        m = l - new0
        if m > 0:
            ev = stats.event_pair()
            op.hermitian_matvec(bs[new0:l], sq, om, precision, out=rbs[new0:l])
            ev[1].record()
            stats.matvec_vectors += m
            # new columns of H = bs^T R bs; the new rows left of them by symmetry of R
            gram(bs, l, rbs[new0:l], m, N, H, ws, ldo=cap, o_off=new0)
            if new0 > 0:
                H[new0:l, :new0] = H[:new0, new0:l].T
        nev = l if use_l2 else k
        w2 = torch.empty(nev, dtype=torch.float64, device=dev)
        Tb = torch.empty((l, nev), dtype=torch.float64, device=dev)
        nb = _lib.load().dftb_subspace_workspace(l, nev)
        wsb, wsn = ws.get(nb)
        call("dftb_subspace_hermitian", _p(H), l, cap, nev, _p(w2), _p(Tb), _p(info), _p(wsb), wsn, st)
        T = torch.empty((nev, N), dtype=torch.float64, device=dev)
        gemm(Tb, bs, T, nev, N, l, nev, N, N, a_trans=1, tile=64)
        RT = torch.empty((nev, N), dtype=torch.float64, device=dev)
        gemm(Tb, rbs, RT, nev, N, l, nev, N, N, a_trans=1, tile=64)
        if use_l2:
            # Lambda2 ordering of the Ritz vectors (Solver.py:348-375, 496-501)
            L2 = torch.empty(nev, dtype=torch.float64, device=dev)
            oia = op.to_padded(np.asarray(Oia)[None])[0]
            call("dftb_weighted_sumsq", _p(T), _p(oia), nev, N, _p(L2), st)
            l2 = L2.cpu().numpy()
            order = np.concatenate([np.where(l2 > L2_thresh)[0], np.where(l2 <= L2_thresh)[0]])
            oi = torch.as_tensor(order, device=dev)
            w2 = w2[oi].contiguous()
            T = T[oi].contiguous()
            RT = RT[oi].contiguous()
        W = torch.empty((k, N), dtype=torch.float64, device=dev)
        residual(RT, T, w2, k, N, W, norms, ws)
        nr = norms[:k].cpu().numpy()
        w_host = np.sqrt(w2[:k].cpu().numpy())
        stats.history.append((l, nr.copy(), w_host.copy()))
        stats.iterations = it + 1
        if PROGRESS is not None:
            PROGRESS("davidson it %d: l=%d max res %.3e (%.1f s)" % (it, l, float(nr.max()), time.time() - t_start))
        if verbose > 0:
            print("Iteration %d: ( %d expansion vectors )   time: %2.6f s" % (it, l, time.time() - t_start))
            for n in range(k):
                print("  Eigenvalue %3d      %2.7f    %s   res. norm = %e" %
                      (n + 1, w_host[n], "converged    " if nr[n] < conv else "not converged", nr[n]))
        if (nr < conv).all():
            break
        dkmax = min(kmax - l, k)
        eps = 0.01 * conv
        nc = int(np.sum(nr > eps))
        dk = min(dkmax, nc)
        sel = [n for n in range(dkmax) if nr[n] > eps]
        if len(sel) != dk:
            raise AssertionError("nb == dk violated (%d != %d)" % (len(sel), dk))
        if dk == 0:
            continue
        if l + dk > cap:
            bs = rbs = H = None
            if not sub.grow(l, l + dk):
                raise ExcitedStatesNotConverged("expansion space exhausted the device memory budget (%d vectors)"
                                                % sub.limit)
            cap = sub.cap
            bs, rbs = sub.bs, sub.prod[0]
            H = sub.M[0]
        wv = torch.sqrt(w2[:k]).contiguous()
        seli = torch.as_tensor(sel, dtype=torch.int32, device=dev)
        Q = bs[l:l + dk]
        call("dftb_precondition", _p(W), _p(wv), _p(seli), dk, _p(omega), N, _p(Q), st)
        nb = _lib.load().dftb_orthonormalize_workspace(l, dk, N)
        wsb, wsn = ws.get(nb)
        call("dftb_orthonormalize", _p(bs), l, _p(Q), dk, N, 1.0e-8, _p(keep), _p(wsb), wsn, st)
        kp = keep[:dk].cpu().numpy()
        kept = np.where(kp != 0)[0]
        if 0 < len(kept) < dk:
            Q[:len(kept)] = Q[torch.as_tensor(kept, device=dev)].clone()
        new0 = l
        l = l + len(kept)
    else:
        raise ExcitedStatesNotConverged(
            "Davidson-like diagonalization for Hermitian eigenvalue problem did not converge after %s "
            "iterations (threshold: %e) !!!" % (maxiter, conv))
    w = torch.sqrt(w2[:k]).contiguous()
    C = T[:k].contiguous()
    # X+Y = sqrt(omega)/sqrt(Omega) T ;  X-Y = sqrt(Omega)/sqrt(omega) T   (Solver.py:561-568)
    XpY = torch.empty_like(C)
    XmY = torch.empty_like(C)
    call("dftb_scale2", _p(C), _p(1.0 / torch.sqrt(w)), _p(sq), k, N, _p(XpY), st)
    call("dftb_scale2", _p(C), _p(torch.sqrt(w)), _p(sq_inv), k, N, _p(XmY), st)
    stats.t_total = time.time() - t_start
    stats.l = l
    return w, C, XmY, XpY, stats


def casida_dense_gpu(op, chunk=512, omega_shift=None, nev=None):
    """Full Casida solve (Solver.py:98-181) for small systems: all n_ov roots, or
    the lowest nev of them (the diag_check comparison, LR_TDDFTB.py:493-518, needs
    only nstates).  omega_shift (o x v, the CT correction) is added to the
    diagonals of A and B (Solver.py:46-49), i.e. twice to A+B.

    (A+-B) are assembled column block by column block with the matvec kernel,
    (A-B)^{1/2} is diagonal (lc = 0) or the Newton-Schulz square root, and the
    Hermitian problem (A-B)^{1/2}(A+B)(A-B)^{1/2} F = Omega^2 F is solved by the
    GPU eigensolver.  Returns padded device tensors (Omega, C, XmY, XpY).
    """
    dev = op.device
    st = _stream()
    o, v, N = op.nocc, op.nvirt, op.N
    n = o * v
    if n > 4096:
        raise ValueError("dense Casida is limited to n_ov <= 4096 (got %d); use nstates" % n)
    real = (torch.arange(o, device=dev)[:, None] * op.vpad + torch.arange(v, device=dev)[None, :]).reshape(-1)
    ApB = torch.empty((n, n), dtype=torch.float64, device=dev)
    AmB = torch.empty((n, n), dtype=torch.float64, device=dev)
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        V = torch.empty((m, N), dtype=torch.float64, device=dev)
        call("dftb_unit_vectors", _p(real[c0:c0 + m].contiguous()), m, N, _p(V), st)
        P, M = op.matvec(V, 0)
        ApB[c0:c0 + m] = P[:, real]
        AmB[c0:c0 + m] = M[:, real]
    if omega_shift is not None:
        sh = torch.as_tensor(np.asarray(omega_shift, dtype=np.float64).reshape(-1), device=dev)
        ApB.diagonal().add_(2.0 * sh)
    ws = op.ws
    dAmB = torch.diagonal(AmB).contiguous()
    offdiag = float(torch.linalg.norm(AmB - torch.diag(dAmB)))
    R = torch.empty_like(ApB)
    if offdiag < 1.0e-10:
        sq = torch.sqrt(dAmB)
        call("dftb_scale2", _p(ApB), _p(sq), _p(sq), n, n, _p(R), st)
        Sq = Sinv = None
    else:
        Sq = torch.empty_like(AmB)
        Sinv = torch.empty_like(AmB)
        it = torch.zeros(1, dtype=torch.int32)
        er = torch.zeros(1, dtype=torch.float64)
        w, wn = ws.get(_lib.load().dftb_sqrtm_workspace(n))
        call("dftb_sqrtm", _p(AmB), n, n, _p(Sq), _p(Sinv), ctypes.c_void_p(it.data_ptr()),
             ctypes.c_void_p(er.data_ptr()), _p(w), wn, st)
        X = torch.empty_like(ApB)
        gemm(ApB, Sq, X, n, n, n, n, n, n)
        gemm(Sq, X, R, n, n, n, n, n, n)
    m = n if nev is None else max(1, min(int(nev), n))
    w2 = torch.empty(m, dtype=torch.float64, device=dev)
    F = torch.empty((n, m), dtype=torch.float64, device=dev)
    w, wn = ws.get(_lib.load().dftb_syevx_workspace(n, m))
    call("dftb_syevx", _p(R), n, n, m, _p(w2), _p(F), _p(w), wn, st)
    Om = torch.sqrt(w2)
    # X+Y = (A-B)^{1/2} F / sqrt(Omega); X-Y = (A+B)(X+Y)/Omega; C = (A-B)^{-1/2} (X+Y) sqrt(Omega)
    XpYc = torch.empty_like(F)
    if Sq is None:
        call("dftb_scale2", _p(F), _p(sq), _p(1.0 / torch.sqrt(Om)), n, m, _p(XpYc), st)
    else:
        gemm(Sq, F, XpYc, n, m, n, n, m, m)
        call("dftb_scale2", _p(XpYc), None, _p(1.0 / torch.sqrt(Om)), n, m, _p(XpYc), st)
    XmYc = torch.empty_like(F)
    gemm(ApB, XpYc, XmYc, n, m, n, n, m, m)
    call("dftb_scale2", _p(XmYc), None, _p(1.0 / Om), n, m, _p(XmYc), st)
    Cc = torch.empty_like(F)
    if Sq is None:
        call("dftb_scale2", _p(F), None, None, n, m, _p(Cc), st)      # C = F when A-B is diagonal
    else:
        gemm(Sinv, XpYc, Cc, n, m, n, n, m, m)
        call("dftb_scale2", _p(Cc), None, _p(torch.sqrt(Om)), n, m, _p(Cc), st)

    def padded(Xc):   # columns are states -> rows [state][opad*vpad]
        out = torch.zeros((m, N), dtype=torch.float64, device=dev)
        out[:, real] = Xc.T
        return out
    return Om, padded(Cc), padded(XmYc), padded(XpYc)


def zvector_gpu(op, B, X0=None, maxiter=1000, conv=1.0e-14, precision=0, verbose=0):
    """Solve (A+B) X = B for k right-hand sides on the GPU (Solver.py:921-1012,
    the Z-vector equations of ExcGradients.py:627-635).

    Block Krylov (Galerkin) iteration: the basis starts from diag(omega)^{-1} B (or
    X0), the projected system bs^T (A+B) bs Xb = bs^T B is solved by a Cholesky
    factorisation (it is symmetric positive definite), X = bs^T Xb, the residual
    (A+B) X - B comes from the stored products (no extra matvec), and each
    residual above 0.01 conv, preconditioned by diag(omega)^{-1}, expands the basis.
    The basis is kept orthonormal (same span, hence the same X, as the reference's
    Gram-Schmidt of the whole set).  B and X0 are (k, o, v); returns X (k, o, v).
    Raises Exception (the reference's) when maxiter iterations do not converge.

    precision > 0 (FP32 / BF16x3 products): iterative refinement with the same
    convergence test.  Each outer step forms the residual R = B - (A+B) X with FP64
    products and solves (A+B) D = R with the Krylov iteration above on the
    low-precision products to 1e-5 of |R| (their noise floor is ~1e-7 relative),
    X += D; it stops when every FP64 residual norm is below conv.
    """
    Bd = op.to_padded(B)
    X0d = op.to_padded(X0) if X0 is not None else None
    if precision == 0:
        X, ok, it = _zvector_krylov(op, Bd, X0d, maxiter, conv, 0, verbose)
        if not ok:
            raise Exception("Z-vector equations: Krylov solver for linear system did not converge after %s "
                            "iterations (threshold: %e) !!!" % (it, conv))
        return op.from_padded(X).cpu().numpy().copy()
    ws = op.ws
    k, N = Bd.shape
    X = X0d.clone() if X0d is not None else torch.zeros((k, N), dtype=torch.float64, device=op.device)
    norms = torch.empty(k, dtype=torch.float64, device=op.device)
    used = 0
    for outer in range(40):
        R = Bd - op.matvec(X, 0)[0] if (X0d is not None or outer > 0) else Bd.clone()
        row_norms(R, k, N, norms, ws)
        nr = norms.cpu().numpy()
        if verbose > 0:
            print("Z-vector refinement %d: FP64 residual norms %s" % (outer, nr))
        if (nr < conv).all():
            return op.from_padded(X).cpu().numpy().copy()
        if used >= maxiter:
            break
        D, _, it = _zvector_krylov(op, R, None, min(300, maxiter - used), max(conv, 1.0e-5 * float(nr.max())),
                                   precision, verbose)
        used += it
        X += D
    raise Exception("Z-vector equations: Krylov solver for linear system did not converge after %s "
                    "iterations (threshold: %e) !!!" % (used, conv))


def _zvector_krylov(op, Bd, X0d, maxiter, conv, precision, verbose):
    """The Krylov iteration of zvector_gpu on padded device arrays: returns
    (X [k][N], converged, iterations); a basis that stops growing (no residual
    above 0.01 conv, or no independent new vector) ends it unconverged."""
    dev = op.device
    st = _stream()
    ws = op.ws
    N, o, v = op.N, op.nocc, op.nvirt
    k = Bd.shape[0]
    kmax = o * v
    ainv = torch.where(op.omega_real.reshape(-1) != 0, 1.0 / op.omega, torch.zeros_like(op.omega))
    cap = min(kmax, 64 * k)
    bs = torch.zeros((cap, N), dtype=torch.float64, device=dev)
    P = torch.zeros((cap, N), dtype=torch.float64, device=dev)

    def grow(need):
        nonlocal bs, P, cap
        if need <= cap:
            return
        new = min(kmax, max(need, 2 * cap))
        for name in ("bs", "P"):
            old = bs if name == "bs" else P
            t = torch.zeros((new, N), dtype=torch.float64, device=dev)
            t[:cap].copy_(old)
            if name == "bs":
                bs = t
            else:
                P = t
        cap = new

    if X0d is None:
        call("dftb_scale2", _p(Bd), None, _p(ainv), k, N, _p(bs[:k]), st)
    else:
        bs[:k].copy_(X0d)
    l = _orthonormalize_block(bs, 0, k, N, ws, st)
    op.matvec(bs[:l], precision, out_p=P[:l])
    norms = torch.empty(k, dtype=torch.float64, device=dev)
    X = torch.empty((k, N), dtype=torch.float64, device=dev)
    W = torch.empty((k, N), dtype=torch.float64, device=dev)
    t0 = time.time()
    it = 0
    for it in range(maxiter):
        Ab = torch.empty((l, l), dtype=torch.float64, device=dev)
        gram(bs, l, P, l, N, Ab, ws)
        Bb = torch.empty((l, k), dtype=torch.float64, device=dev)
        gram(bs, l, Bd, k, N, Bb, ws)
        asym = torch.zeros(3, dtype=torch.float64, device=dev)
        call("dftb_symmetrize", _p(Ab), l, l, _p(asym), st)
        # Xb = Ab^{-1} Bb (la.solve, Solver.py:969) by a Cholesky factorisation of the
        # symmetric positive definite projected (A+B) (ROCm solver library)
        Lc, info = torch.linalg.cholesky_ex(Ab)
        if int(info.cpu()) != 0:
            raise ExcitedStatesError("Z-vector equations: the projected (A+B) is not positive definite")
        Xb = torch.cholesky_solve(Bb, Lc).contiguous()
        gemm(Xb, bs, X, k, N, l, k, N, N, a_trans=1)
        gemm(Xb, P, W, k, N, l, k, N, N, a_trans=1)              # (A+B) X from the stored products
        W.sub_(Bd)
        row_norms(W, k, N, norms, ws)
        nr = norms.cpu().numpy()
        if verbose > 0:
            print("Iteration %d: ( %d expansion vectors )   time: %2.6f s" % (it, l, time.time() - t0))
            for n in range(k):
                print("  Solution vector  %3d      %s   res. norm = %e"
                      % (n, "converged    " if nr[n] < conv else "not converged", nr[n]))
        if (nr < conv).all():
            if verbose > 0:
                print("Z-vector equations CONVERGED (residual norms < %s) after %d iterations" % (conv, it + 1))
            return X, True, it + 1
        dkmax = min(kmax - l, k)
        sel = [n for n in range(dkmax) if nr[n] > 0.01 * conv]
        if not sel:
            break
        grow(l + len(sel))
        idx = torch.as_tensor(sel, dtype=torch.int64, device=dev)
        call("dftb_scale2", _p(W[idx].contiguous()), None, _p(ainv), len(sel), N, _p(bs[l:l + len(sel)]), st)
        l_new = _orthonormalize_block(bs, l, len(sel), N, ws, st)
        if l_new == l:
            break
        op.matvec(bs[l:l_new], precision, out_p=P[l:l_new])
        l = l_new
    return X, False, it + 1


def stability_gpu(q_oo, q_vv, q_ov, omega, gamma, gamma_lr, nev=6, chunk=512):
    """Lowest eigenvalues of the restricted -> unrestricted stability matrix
    R = diag(omega) - (ij|ab) - (ia|jb)_lr (StabilityAnalysis.py:24-56).

    R is the A block ((A+B) + (A-B)) / 2 of a Casida operator whose Coulomb
    kernel is -gamma_lr / 2 and whose exchange kernel is gamma: its columns come
    from the matvec kernel on unit vectors, its lowest nev eigenvalues from the
    GPU eigensolver (the reference builds R densely and calls eigsh).  Charges in
    the reference layout; n_ov <= 4096."""
    op = CasidaOperatorGPU.from_charges(q_oo, q_vv, q_ov, omega, -0.5 * np.asarray(gamma_lr, dtype=np.float64),
                                        gamma, lc=1)
    dev = op.device
    st = _stream()
    o, v, N = op.nocc, op.nvirt, op.N
    n = o * v
    if n > 4096:
        raise ValueError("the stability matrix is limited to n_ov <= 4096 (got %d)" % n)
    real = (torch.arange(o, device=dev)[:, None] * op.vpad + torch.arange(v, device=dev)[None, :]).reshape(-1)
    R = torch.empty((n, n), dtype=torch.float64, device=dev)
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        V = torch.empty((m, N), dtype=torch.float64, device=dev)
        call("dftb_unit_vectors", _p(real[c0:c0 + m].contiguous()), m, N, _p(V), st)
        P, M = op.matvec(V, 0)
        R[c0:c0 + m] = 0.5 * (P[:, real] + M[:, real])
    k = min(nev, n)
    w = torch.empty(k, dtype=torch.float64, device=dev)
    F = torch.empty((n, k), dtype=torch.float64, device=dev)
    ws, wn = op.ws.get(_lib.load().dftb_syevx_workspace(n, k))
    call("dftb_syevx", _p(R), n, n, k, _p(w), _p(F), _p(ws), wn, st)
    return w.cpu().numpy()
