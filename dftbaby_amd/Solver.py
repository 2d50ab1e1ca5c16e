"""Drop-in replacements of DFTBaby's iterative TD-DFTB solvers (DFTB/Solver.py).

Same function names, positional/keyword arguments and return values as the
reference (numpy arrays in, ``(Omega[k], C[k,o,v], XmY[k,o,v], XpY[k,o,v])``
out), computed on the GPU by ``dftbaby_amd.engine``.  The extra keyword
``precision`` ("fp64" default, "fp32", "bf16x3") selects the arithmetic of the
(A+-B) products; the Davidson residual, orthonormalisation and subspace
eigensolve are FP64 in every mode.

Reference: Casida Solver.py:98-181, HermitianDavidson Solver.py:404-578,
nonHermitianDavidson Solver.py:637-917, Krylov_solver_Zvector
Solver.py:921-1012, exceptions Solver.py:27-33.
"""
import numpy as np

from .engine import (CasidaOperatorGPU, ExcitedStatesError, ExcitedStatesNotConverged,  # noqa: F401
                     casida_dense_gpu, hermitian_davidson_gpu, nonhermitian_davidson_gpu, zvector_gpu)

_PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, 0: 0, 1: 1, 2: 2}


def precision_code(precision):
    if precision not in _PRECISIONS:
        raise ValueError("precision must be 'fp64', 'fp32' or 'bf16x3', not %r" % (precision,))
    return _PRECISIONS[precision]


def _out(op, tensors):
    return tuple(op.from_padded(t).cpu().numpy().copy() for t in tensors)


def Casida(gamma, gamma_lr, qtrans_oo, qtrans_vv, qtrans_ov, omega, omega_shift, df, nocc, nvirt, multiplicity="S"):
    """Full Casida solve for all nocc*nvirt roots (Solver.py:98-181), n_ov <= 4096:
    (A+-B) assembled from the GPU products, the (A-B)^{1/2}(A+B)(A-B)^{1/2} problem
    solved by the GPU eigensolver.  Triplets drop the Coulomb kernel (71-80); the CT
    shift enters the diagonals of A and B (46-49).  Returns (Omega, C, XmY, XpY)
    with the vectors as (nocc*nvirt, nocc, nvirt).  Closed-shell occupations only
    (df = 2 on every pair)."""
    if not np.allclose(np.asarray(df, dtype=np.float64), 2.0):
        raise NotImplementedError("Casida: closed-shell occupation differences (df = 2) only")
    assert multiplicity in ("S", "T")
    g = np.asarray(gamma, dtype=np.float64)
    glr = np.asarray(gamma_lr, dtype=np.float64)
    lc = int(np.any(glr != 0.0))
    op = CasidaOperatorGPU.from_charges(qtrans_oo, qtrans_vv, qtrans_ov, omega, g if multiplicity == "S" else 0.0 * g,
                                        glr, lc=lc)
    shift = None if omega_shift is None or not np.any(omega_shift) else omega_shift
    w, C, XmY, XpY = casida_dense_gpu(op, omega_shift=shift)
    C, XmY, XpY = _out(op, (C, XmY, XpY))
    return w.cpu().numpy(), C, XmY, XpY


def HermitianDavidson(g0, q_ov, omega, omega_shift, nocc, nvirt, XmYguess, XpYguess, Oia, selector=None,
                      multiplicity="S", nstates=4, ifact=1, conv=1.0e-14, maxiter=10, L2_thresh=0.5, verbose=1,
                      precision="fp64"):
    """(A-B)^{1/2}(A+B)(A-B)^{1/2} T = Omega^2 T for the lowest nstates roots (lc = 0).

    As in the reference, ``multiplicity`` is not used (singlets are computed).
    """
    if selector is not None:
        raise NotImplementedError("symmetry selectors are not supported by the GPU solver")
    op = CasidaOperatorGPU.from_charges(None, None, q_ov, omega, g0, 0.0 * np.asarray(g0), lc=0)
    w, C, XmY, XpY, _ = hermitian_davidson_gpu(op, nstates, omega_shift=omega_shift, ifact=ifact, conv=conv,
                                               maxiter=maxiter, XpYguess=XpYguess, Oia=Oia, L2_thresh=L2_thresh,
                                               precision=precision_code(precision), verbose=verbose)
    C, XmY, XpY = _out(op, (C, XmY, XpY))
    return w.cpu().numpy(), C, XmY, XpY


def nonHermitianDavidson(g0, g0_lr, q_oo, q_vv, q_ov, omega, nocc, nvirt, XmYguess, XpYguess, w_guess,
                         selector=None, multiplicity="S", nstates=4, ifact=1, conv=1.0e-14, maxiter=10, verbose=1,
                         lc=1, precision="fp64"):
    """Stratmann-Scuseria-Frisch solver of the non-Hermitian Casida problem (lc = 1)."""
    assert multiplicity == "S"
    if selector is not None:
        raise NotImplementedError("symmetry selectors are not supported by the GPU solver")
    op = CasidaOperatorGPU.from_charges(q_oo, q_vv, q_ov, omega, g0, g0_lr, lc=lc)
    w, C, XmY, XpY, _ = nonhermitian_davidson_gpu(op, nstates, ifact=ifact, conv=conv, maxiter=maxiter,
                                                  XmYguess=XmYguess, XpYguess=XpYguess, w_guess=w_guess,
                                                  precision=precision_code(precision), verbose=verbose)
    C, XmY, XpY = _out(op, (C, XmY, XpY))
    return w.cpu().numpy(), C, XmY, XpY


def Krylov_solver_Zvector(A, Adiag, B, X0, maxiter=1000, conv=1.0e-14, verbose=1, precision="fp64"):
    """Solve (A+B) X = B iteratively (Solver.py:921-1012), as ExcGradients.py:627-635
    calls it for the Z-vector equations.

    ``A`` is the (A+B) operator: the reference passes a closure over
    ``_ApBv(g0, g0_lr, qoo, qov, qvv, omega, vs, 1)``; here it is the
    ``CasidaOperatorGPU`` built from the same inputs (``CasidaOperatorGPU.from_charges``
    or the operator of an ``LR_TDDFTB`` object), whose products run on the GPU.
    ``Adiag`` is the orbital-energy difference omega (nocc, nvirt) of that operator
    (the preconditioner).  B: (nocc, nvirt, k); X0: (nocc, nvirt, k) or None.
    Returns X (nocc, nvirt, k); raises Exception when it does not converge.
    """
    if not isinstance(A, CasidaOperatorGPU):
        raise TypeError("Krylov_solver_Zvector: A must be the CasidaOperatorGPU of the (A+B) products")
    Ad = np.asarray(Adiag, dtype=np.float64)
    if not np.allclose(Ad, A.omega_real[:A.nocc, :A.nvirt].cpu().numpy(), rtol=0.0, atol=1e-12):
        raise ValueError("Krylov_solver_Zvector: Adiag must be the operator's orbital-energy differences")
    Bk = np.moveaxis(np.asarray(B, dtype=np.float64), 2, 0)
    X0k = None if X0 is None else np.moveaxis(np.asarray(X0, dtype=np.float64), 2, 0)
    X = zvector_gpu(A, Bk, X0=X0k, maxiter=maxiter, conv=conv, precision=precision_code(precision),
                    verbose=verbose)
    return np.moveaxis(X, 0, 2)
