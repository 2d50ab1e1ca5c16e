"""GPU Davidson solvers against the dense Casida oracle (the reference's diag_check criterion)."""
import numpy as np
import pytest

from conftest import ground_state
from oracle import tddftb_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup(name, lc, sigma=0.0, seed=0):
    gs = ground_state(name, lc, sigma, seed)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g = gs.getGamma()
    glr = gs.getGamma_lr() if lc else 0 * g
    return gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr


def _sign_err(a, b):
    return sum(min(np.abs(x - y).sum(), np.abs(x + y).sum()) for x, y in zip(a, b))


@pytest.mark.parametrize("name,sigma", [("C8H10", 0.05), ("C12H14", 0.05)])
def test_nonhermitian_davidson_vs_casida(name, sigma):
    from dftbaby_amd.engine import CasidaOperatorGPU, nonhermitian_davidson_gpu
    gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr = _setup(name, 1, sigma, 11)
    o, v = omega.shape
    full = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    k = 6
    C = gs.getKSCoefficients()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega, g, glr, 1)
    # restart=0: the reference's trajectory (unbounded basis), so iteration counts compare
    w, Cg, XmY, XpY, st = nonhermitian_davidson_gpu(op, k, conv=1e-10, maxiter=100, restart=0)
    w = w.cpu().numpy()
    assert np.abs(w - full[0][:k]).max() < 1e-10
    xpy = op.from_padded(XpY).cpu().numpy()
    xmy = op.from_padded(XmY).cpu().numpy()
    cc = op.from_padded(Cg).cpu().numpy()
    assert _sign_err(xpy, full[3][:k]) < 1e-6
    assert _sign_err(xmy, full[2][:k]) < 1e-6
    # C = bs Tb depends on the expansion space, not only on the roots (the reference's
    # own diag_check leaves C out, LR_TDDFTB.py:513-516): compare with the oracle's
    # Davidson, which builds the same spans iteration by iteration.
    ref = O.nonHermitianDavidson(g, glr, q_oo, q_vv, q_ov, omega, o, v, None, None, None, nstates=k,
                                 conv=1e-10, maxiter=100)
    assert np.abs(ref[0] - w).max() < 1e-10
    assert st.iterations == len(O.nonHermitianDavidson.history)
    # the spans agree only up to rounding amplified in the last, nearly converged
    # expansion vectors: C agrees to ~1e-5 per element, checked per state in 2-norm
    for a, b in zip(cc, ref[1]):
        assert min(np.linalg.norm(a - b), np.linalg.norm(a + b)) < 1e-4


def test_hermitian_davidson_vs_casida_c20h22():
    """Config-1 molecule (planar C20H22, lc=0, 10 singlets)."""
    from dftbaby_amd.engine import CasidaOperatorGPU, hermitian_davidson_gpu
    gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr = _setup("C20H22", 0)
    o, v = omega.shape
    full = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    k = 10
    C = gs.getKSCoefficients()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega, g, glr, 0)
    w, Cg, XmY, XpY, st = hermitian_davidson_gpu(op, k, conv=1e-10, maxiter=100)
    w = w.cpu().numpy()
    assert np.abs(w - full[0][:k]).max() < 1e-10
    assert _sign_err(op.from_padded(XpY).cpu().numpy(), full[3][:k]) < 1e-6
    assert _sign_err(op.from_padded(XmY).cpu().numpy(), full[2][:k]) < 1e-6


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_precision_davidson(precision):
    """FP32 / BF16x3 matvecs with FP64 Davidson: energies within 1e-6 Ha of FP64."""
    from dftbaby_amd.engine import CasidaOperatorGPU, nonhermitian_davidson_gpu
    gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr = _setup("C12H14", 1, 0.05, 5)
    o, v = omega.shape
    C = gs.getKSCoefficients()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega, g, glr, 1)
    w64 = nonhermitian_davidson_gpu(op, 8, conv=1e-6, maxiter=100)[0].cpu().numpy()
    wlo = nonhermitian_davidson_gpu(op, 8, conv=1e-5, maxiter=100, precision=precision)[0].cpu().numpy()
    assert np.abs(w64 - wlo).max() < 1e-6


def test_thick_restart_converges_to_casida_roots():
    """A small restart threshold collapses the basis to span(X+Y, X-Y) every few
    iterations; the converged roots and vectors are unchanged."""
    from dftbaby_amd.engine import CasidaOperatorGPU, nonhermitian_davidson_gpu
    gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr = _setup("C12H14", 1, 0.05, 11)
    o, v = omega.shape
    full = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    k = 6
    C = gs.getKSCoefficients()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega, g, glr, 1)
    w, Cg, XmY, XpY, st = nonhermitian_davidson_gpu(op, k, conv=1e-9, maxiter=200, restart=8 * k)
    assert st.restarts > 0
    assert np.abs(w.cpu().numpy() - full[0][:k]).max() < 1e-10
    assert _sign_err(op.from_padded(XpY).cpu().numpy(), full[3][:k]) < 1e-6
    assert _sign_err(op.from_padded(XmY).cpu().numpy(), full[2][:k]) < 1e-6


def test_mixed_precision_refinement_reaches_fp64_tolerance():
    """BF16x3 products cannot reach a 1e-9 residual; the solver refines with FP64
    products from its own roots and converges to the FP64 answer."""
    from dftbaby_amd.engine import CasidaOperatorGPU, nonhermitian_davidson_gpu
    gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr = _setup("C12H14", 1, 0.05, 5)
    o, v = omega.shape
    full = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    C = gs.getKSCoefficients()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega, g, glr, 1)
    w, Cg, XmY, XpY, st = nonhermitian_davidson_gpu(op, 6, conv=1e-9, maxiter=200, precision=2)
    assert st.refined
    assert np.abs(w.cpu().numpy() - full[0][:6]).max() < 1e-10


def test_expansion_space_grows_on_demand(monkeypatch):
    """The expansion space starts at the smallest capacity and grows (copying the
    filled rows and projected blocks) in both solvers; the roots are unchanged."""
    from dftbaby_amd import engine
    from dftbaby_amd.engine import CasidaOperatorGPU, nonhermitian_davidson_gpu, hermitian_davidson_gpu
    gs, ao, av, q_oo, q_ov, q_vv, omega, df, g, glr = _setup("C12H14", 1, 0.05, 11)
    o, v = omega.shape
    full = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    k = 6
    C = gs.getKSCoefficients()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], gs.getOverlapMatrix(), gs.orbsPerAtom, omega, g, glr, 1)
    grown = []
    orig = engine._Subspace.grow

    def grow(self, l, need):
        grown.append((self.cap, need))
        return orig(self, l, need)

    monkeypatch.setattr(engine, "_initial_cap", lambda l0, k_add, limit, iters=16: min(limit, l0 + k_add))
    monkeypatch.setattr(engine._Subspace, "grow", grow)
    w, Cg, XmY, XpY, st = nonhermitian_davidson_gpu(op, k, conv=1e-9, maxiter=200)
    assert grown
    assert np.abs(w.cpu().numpy() - full[0][:k]).max() < 1e-10
    assert _sign_err(op.from_padded(XpY).cpu().numpy(), full[3][:k]) < 1e-6
    # Hermitian solver (no long-range exchange)
    gs2, ao2, av2, q_oo2, q_ov2, q_vv2, om2, df2, g2, glr2 = _setup("C20H22", 0)
    o2, v2 = om2.shape
    full2 = O.Casida(g2, glr2, q_oo2, q_vv2, q_ov2, om2, 0 * om2, df2, o2, v2)
    C2 = gs2.getKSCoefficients()
    op2 = CasidaOperatorGPU.from_orbitals(C2[:, ao2], C2[:, av2], gs2.getOverlapMatrix(), gs2.orbsPerAtom, om2, g2,
                                          glr2, 0)
    grown.clear()
    w2, C2g, XmY2, XpY2, st2 = hermitian_davidson_gpu(op2, k, conv=1e-9, maxiter=200)
    assert grown
    assert np.abs(w2.cpu().numpy() - full2[0][:k]).max() < 1e-10
