"""The drop-in LR_TDDFTB / Solver API on the GPU against the oracle (run with -m gpu)."""
import os

import numpy as np
import pytest

from conftest import ROOT, ground_state
from oracle import tddftb_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _td(name, lc, sigma=0.05, seed=13):
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state(name, lc, sigma, seed)
    gs.verbose = 0
    return LR_TDDFTB(ground_state=gs), gs


@pytest.mark.parametrize("lc", [0, 1])
def test_dense_casida_path(lc):
    """nstates=None: the full Casida solve (Solver.py:98-181) on the GPU."""
    td, gs = _td("C8H10", lc)
    Omega, osc = td.getEnergies(nstates=None)
    ref = O.solve_from_ground_state(gs, nstates=None)
    np.testing.assert_allclose(Omega, ref["Omega"], atol=1e-10)
    np.testing.assert_allclose(osc[:10], ref["osc"][:10], atol=1e-9)
    for n in range(10):
        a, b = td.XpY[n], ref["XpY"][n]
        assert min(np.abs(a - b).max(), np.abs(a + b).max()) < 1e-8


def test_diag_check_and_attributes(tmp_path):
    td, gs = _td("C8H10", 1)
    Omega, osc = td.getEnergies(nstates=5, diag_conv=1e-11, diag_maxiter=100, diag_check=1)
    ref = O.solve_from_ground_state(gs, nstates=5, diag_conv=1e-11, diag_maxiter=100)
    np.testing.assert_allclose(Omega, ref["Omega"], atol=1e-10)
    q_oo, q_vv, q_ov = td.getTransitionCharges()
    np.testing.assert_allclose(q_ov, ref["q_ov"], atol=1e-13)
    np.testing.assert_allclose(q_oo, ref["q_oo"], atol=1e-13)
    np.testing.assert_allclose(q_vv, ref["q_vv"], atol=1e-13)
    np.testing.assert_allclose(td.Oia, ref["Oia"], atol=1e-13)
    np.testing.assert_allclose(td.Lambda2, ref["Lambda2"], atol=1e-6)
    assert td.XmY.shape == (5, td.nocc, td.nvirt)
    f = str(tmp_path / "spec.dat")
    td.saveAbsorptionSpectrum(f)
    rows = np.loadtxt(f)
    np.testing.assert_allclose(rows[:, 0], np.round(Omega, 7), atol=1e-7)


def test_mulliken_dipoles_and_exc_dipoles():
    td, gs = _td("C8H10", 0)
    td.getEnergies(nstates=4, diag_conv=1e-10, diag_maxiter=100, oszis="Mulliken")
    ref = O.solve_from_ground_state(gs, nstates=4, diag_conv=1e-10, diag_maxiter=100, oszis="Mulliken")
    np.testing.assert_allclose(td.oscillator_strength, ref["osc"], rtol=1e-8, atol=1e-10)
    td.getEnergies(nstates=4, diag_conv=1e-10, diag_maxiter=100, oszis="Dipoles")
    tdip = td.getTransitionDipolesExc(5)
    # reference formula (LR_TDDFTB.py:666-681) with the same C and MO dipoles
    C = td.Cij
    ao, av = td.active_occupied_orbs, td.active_virtual_orbs
    for c, t in enumerate((td.tdipMO_X, td.tdipMO_Y, td.tdipMO_Z)):
        Doo, Dvv = t[np.ix_(ao, ao)], t[np.ix_(av, av)]
        for A in range(1, 5):
            for B in range(A, 5):
                P0 = np.sum(C[A - 1] * C[B - 1])
                Pij = C[A - 1] @ C[B - 1].T
                Pab = C[A - 1].T @ C[B - 1]
                expect = tdip[0, 0, c] * P0 - np.sum(Pij * Doo) + np.sum(Pab * Dvv)
                assert abs(tdip[A, B, c] - expect) < 1e-9
        np.testing.assert_allclose(tdip[0, 1:5, c], (td.tdipX, td.tdipY, td.tdipZ)[c], atol=1e-12)


def test_solver_module_signatures():
    from dftbaby_amd import Solver
    gs = ground_state("C8H10", 1, 0.05, 17)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    got = Solver.nonHermitianDavidson(g, glr, q_oo, q_vv, q_ov, omega, o, v, None, None, None, nstates=4,
                                      conv=1e-9, maxiter=100, verbose=0)
    ref = O.nonHermitianDavidson(g, glr, q_oo, q_vv, q_ov, omega, o, v, None, None, None, nstates=4, conv=1e-9,
                                 maxiter=100)
    np.testing.assert_allclose(got[0], ref[0], atol=1e-10)
    # warm start from the previous solution converges at once to the same roots
    again = Solver.nonHermitianDavidson(g, glr, q_oo, q_vv, q_ov, omega, o, v, got[2], got[3], got[0], nstates=4,
                                        conv=1e-9, maxiter=100, verbose=0)
    np.testing.assert_allclose(again[0], ref[0], atol=1e-10)


def test_error_paths():
    from dftbaby_amd import Solver
    from dftbaby_amd.engine import ExcitedStatesNotConverged, ExcitedStatesError
    td, gs = _td("C8H10", 1)
    with pytest.raises(ExcitedStatesNotConverged):
        td.getEnergies(nstates=4, diag_conv=1e-14, diag_maxiter=1)
    with pytest.raises(Exception, match="Revise the TDA code"):
        td.getEnergies(nstates=4, response_method="Tamm-Dancoff")
    # an asymmetric gamma_lr makes Mh non-Hermitian (Solver.py:766-770)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    glr = gs.getGamma_lr().copy()
    glr[0, 1] += 0.05
    with pytest.raises(ExcitedStatesError):
        Solver.nonHermitianDavidson(gs.getGamma(), glr, q_oo, q_vv, q_ov, omega, len(ao), len(av), None, None,
                                    None, nstates=4, conv=1e-8, maxiter=50, verbose=0)


@pytest.mark.parametrize("ifact", [None, 1, 4])
def test_ct_correction(ifact):
    """ct_correction=1 (LR_TDDFTB.py:429-437) on C8H10 + H2 40 bohr apart: the
    shifts of LongRangeCT.py:13-93 computed on the device equal the oracle's; the
    dense path (ifact None) equals the oracle's dense Casida with the same shifts;
    the Davidson (14 roots) equals the oracle's HermitianDavidson with the shifts
    (ifact 1: the block-diagonal problem of the separated fragments leaves two
    roots out of reach of the 14 starting vectors, as in the reference) and, from
    56 starting vectors, the dense spectrum, which the shift moves by 0.013 Ha."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10-H2", 0)
    gs.verbose = 0
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    pos = np.array([p for _, p in gs.getGeometry()])
    T, _, _ = O.ct_shifts(q_oo, q_vv, gs.getGaussianOverlap(), pos, orbe[av], omega)
    g = gs.getGamma()
    o, v = omega.shape
    td = LR_TDDFTB(ground_state=gs)
    if ifact is None:
        Omega, _ = td.getEnergies(nstates=None, ct_correction=1)
        ref = O.Casida(g, 0 * g, q_oo, q_vv, q_ov, omega, T, df, o, v)[0]
    else:
        k = 14
        Omega, _ = td.getEnergies(nstates=k, ct_correction=1, diag_ifact=ifact, diag_conv=1e-11, diag_maxiter=200)
        Oia = O.Lambda2_calcOia(q_oo, q_vv, gs.getGaussianOverlap())
        ref = O.HermitianDavidson(g, q_ov, omega, T, o, v, None, None, Oia, nstates=k, ifact=ifact, conv=1e-11,
                                  maxiter=200, L2_thresh=0.0)[0]
        if ifact == 4:
            dense = O.Casida(g, 0 * g, q_oo, q_vv, q_ov, omega, T, df, o, v)[0]
            dense0 = O.Casida(g, 0 * g, q_oo, q_vv, q_ov, omega, 0 * T, df, o, v)[0]
            np.testing.assert_allclose(Omega, dense[:k], atol=1e-10)
            assert np.abs(dense[:k] - dense0[:k]).max() > 0.01
    np.testing.assert_allclose(td.omega_shift, T, atol=1e-12)
    assert T.max() > 0.04
    np.testing.assert_allclose(Omega, ref[:len(Omega)], atol=1e-10)
    with pytest.raises(AssertionError):
        LR_TDDFTB(ground_state=ground_state("C8H10", 1)).getEnergies(nstates=2, ct_correction=1)


def test_zvector_solver(rng):
    """Solver.Krylov_solver_Zvector on the GPU (A+B) operator against the dense
    solution of the oracle's buildAB and the oracle's own Krylov iteration; a warm
    start from the solution converges at once."""
    from dftbaby_amd import Solver
    from dftbaby_amd.engine import CasidaOperatorGPU
    gs = ground_state("C8H10", 1, 0.05, 2)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    A, B = O.buildAB(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    rhs = rng.standard_normal((o, v, 3))
    Xd = np.linalg.solve(A + B, rhs.reshape(o * v, 3)).reshape(o, v, 3)
    op = CasidaOperatorGPU.from_charges(q_oo, q_vv, q_ov, omega, g, glr, lc=1)
    X = Solver.Krylov_solver_Zvector(op, omega, rhs, None, conv=1e-10, verbose=0)
    assert X.shape == rhs.shape
    assert np.abs(X - Xd).max() < 1e-9 * np.abs(Xd).max()
    Xo = O.Krylov_solver_Zvector(lambda vs: O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1), omega, rhs, None,
                                 conv=1e-10)
    assert np.abs(X - Xo).max() < 1e-9 * np.abs(Xd).max()
    Xw = Solver.Krylov_solver_Zvector(op, omega, rhs, Xd, maxiter=1, conv=1e-10, verbose=0)
    assert np.abs(Xw - Xd).max() < 1e-9 * np.abs(Xd).max()
    with pytest.raises(TypeError):
        Solver.Krylov_solver_Zvector(lambda x: x, omega, rhs, None)


@pytest.mark.parametrize("lc", [0, 1])
def test_stability_analysis(lc, capsys):
    """stability=1 (LR_TDDFTB.py:423-427, StabilityAnalysis.py:24-56): the lowest
    eigenvalues of R from the GPU (matvec columns, GPU eigensolver) equal the
    oracle's dense ones, and the report is printed."""
    td, gs = _td("C8H10", lc)
    td.getEnergies(nstates=2, stability=1, diag_conv=1e-8, diag_maxiter=100)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g = gs.getGamma()
    glr = gs.getGamma_lr() if lc else 0 * g
    w, _ = O.check_stability(g, glr, q_oo, q_vv, q_ov, omega, *omega.shape)
    np.testing.assert_allclose(td.stability_eigenvalues, w, atol=1e-10)
    assert "lowest eigenvalue of restricted/unrestricted stability matrix R" in capsys.readouterr().out


@pytest.mark.parametrize("precision", [1, 2])
def test_zvector_mixed_precision_refinement(precision, rng):
    """zvector_gpu with FP32 / BF16x3 products and FP64 refinement: the solution
    of the dense (A+B) X = rhs to the FP64 residual test, on the AO operator."""
    from dftbaby_amd.engine import CasidaOperatorGPU, zvector_gpu
    gs = ground_state("C12H14", 1, 0.05, 5)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    A, B = O.buildAB(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    rhs = rng.standard_normal((2, o, v))
    Xd = np.linalg.solve(A + B, rhs.reshape(2, o * v).T).T.reshape(2, o, v)
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], S, gs.orbsPerAtom, omega, g, glr, 1, exchange="ao")
    X = zvector_gpu(op, rhs, conv=1e-11, precision=precision)
    assert np.abs(X - Xd).max() < 1e-9 * np.abs(Xd).max()
    P, _ = op.matvec(op.to_padded(X), 0)
    res = np.linalg.norm((op.from_padded(P).cpu().numpy() - rhs).reshape(2, -1), axis=1)
    assert res.max() < 1e-11


@pytest.mark.parametrize("lc", [0, 1])
def test_dense_casida_triplets(lc):
    """nstates=None with multiplicity="T": the dense Casida of the triplet matrices
    (no Coulomb kernel, Solver.py:71-80) as the oracle's, oscillator strengths 0."""
    td, gs = _td("C8H10", lc)
    Omega, osc = td.getEnergies(nstates=None, multiplicity="T")
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g = gs.getGamma()
    glr = gs.getGamma_lr() if lc else 0 * g
    o, v = omega.shape
    ref = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v, multiplicity="T")[0]
    refS = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v, multiplicity="S")[0]
    np.testing.assert_allclose(Omega, ref, atol=1e-10)
    assert np.abs(ref[:5] - refS[:5]).max() > 1e-3
    assert np.all(osc == 0.0)


def _nondegenerate(Om, tol=1e-6):
    gaps = np.diff(Om)
    ok = np.ones(len(Om), dtype=bool)
    ok[:-1] &= gaps > tol
    ok[1:] &= gaps > tol
    return np.where(ok)[0]


@pytest.mark.parametrize("lc,mult,ct", [(0, "S", False), (0, "T", False), (1, "S", False), (1, "T", False),
                                        (0, "S", True)])
def test_solver_casida(lc, mult, ct):
    """Solver.Casida (Solver.py:98-181) on the GPU against the oracle: all roots,
    X-Y and X+Y of every non-degenerate root up to sign, reference layouts
    (n_ov, nocc, nvirt); lc inferred from gamma_lr as Solver.py:44 does (a zero
    gamma_lr means lc = 0); with the CT shifts of LongRangeCT.py on the C8H10 + H2
    dimer (Solver.py:46-49)."""
    from dftbaby_amd import Solver
    gs = ground_state("C8H10-H2", 0) if ct else ground_state("C8H10", lc, 0.05, 13)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g = gs.getGamma()
    glr = gs.getGamma_lr() if lc else 0.0 * g
    o, v = omega.shape
    shift = 0 * omega
    if ct:
        pos = np.array([p for _, p in gs.getGeometry()])
        shift, _, _ = O.ct_shifts(q_oo, q_vv, gs.getGaussianOverlap(), pos, orbe[av], omega)
        assert shift.max() > 0.04
    Om, Cg, XmY, XpY = Solver.Casida(g, glr, q_oo, q_vv, q_ov, omega, shift, df, o, v, multiplicity=mult)
    ref = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, shift, df, o, v, multiplicity=mult)
    assert XpY.shape == (o * v, o, v)
    np.testing.assert_allclose(Om, ref[0], atol=1e-10)
    idx = _nondegenerate(ref[0])
    assert len(idx) > 0.8 * len(Om)
    for n in idx:
        for a, b in ((XmY[n], ref[2][n]), (XpY[n], ref[3][n]), (Cg[n], ref[1][n])):
            assert min(np.abs(a - b).max(), np.abs(a + b).max()) < 1e-7, n


def test_f2py_module_dropin_against_fortran_golden():
    """dftbaby_amd.tddftb.tddftb.{trans_charges, apbv, ambv} (the f2py module's
    names, arguments and layouts, tddftb.f90:12-287) on the GPU against the
    outputs of DFTBaby's own Fortran committed in tests/golden."""
    import os
    from conftest import ROOT
    from dftbaby_amd import tddftb
    d = np.load(os.path.join(ROOT, "tests", "golden", "tddftb_f90_golden.npz"))
    q_oo, q_ov, q_vv = tddftb.tddftb.trans_charges(d["Co"], d["Cv"], d["S"], d["opa"])
    for a, key in ((q_oo, "q_oo"), (q_ov, "q_ov"), (q_vv, "q_vv")):
        assert a.shape == d[key].shape
        np.testing.assert_allclose(a, d[key], rtol=0, atol=1e-14)
    args = (d["g"], d["glr"], d["q_oo"], d["q_ov"], d["q_vv"], d["omega"])
    us_p = tddftb.tddftb.apbv(*args, d["vs"])
    us_m = tddftb.tddftb.ambv(*args, d["vs"])
    assert us_p.shape == d["ApB"].shape
    np.testing.assert_allclose(us_p, d["ApB"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(us_m, d["AmB"], rtol=0, atol=1e-12)
    # a reference-layout block of expansion vectors from a real molecule against the oracle
    gs = ground_state("C8H10", 1, 0.05, 4)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q = tddftb.tddftb.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    vs = np.random.default_rng(5).standard_normal((len(ao), len(av), 7))
    g, glr = gs.getGamma(), gs.getGamma_lr()
    np.testing.assert_allclose(tddftb.tddftb.apbv(g, glr, q[0], q[1], q[2], omega, vs),
                               O.ApBv(g, glr, q[0], q[1], q[2], omega, vs, 1), rtol=0, atol=1e-12)
    np.testing.assert_allclose(tddftb.tddftb.ambv(g, glr, q[0], q[1], q[2], omega, vs),
                               O.AmBv(g, glr, q[0], q[1], q[2], omega, vs, 1), rtol=0, atol=1e-12)


class _DenseDipoleGroundState:
    """A ground state exposing only DFTBaby's DFTB2 getters (DFTB2.py:2405-2454):
    a numpy getDipoleMatrix [norb][norb][3] and no device planes, as DFTBaby's
    own object."""

    def __init__(self, gs):
        self._gs = gs
        for name in ("orbsPerAtom", "long_range_correction", "Nelec_val", "mulliken_dipoles", "parameter_set",
                     "verbose"):
            setattr(self, name, getattr(gs, name))
        self.en = gs.en_elec
        self.dipole_calls = 0

    def getEnergy(self, **opts):
        return self.en

    def getDipoleMatrix(self):
        self.dipole_calls += 1
        return self._gs.getDipoleMatrix()

    def __getattr__(self, name):
        if name.startswith("get") and name != "getDipoleMatrixDevice":
            return getattr(self._gs, name)
        raise AttributeError(name)


@pytest.mark.parametrize("lc", [0, 1])
def test_dense_dipole_matrix_ground_state(lc):
    """LR_TDDFTB on a ground state that only has the numpy getDipoleMatrix
    (tdip = C^T D C, LR_TDDFTB.py:574-592); its transition dipoles, oscillator
    strengths and MO dipole matrices equal the oracle's with the same dipole
    matrix."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", lc, 0.05, 23)
    wrapped = _DenseDipoleGroundState(gs)
    assert not hasattr(wrapped, "getDipoleMatrixDevice")
    td = LR_TDDFTB(ground_state=wrapped)
    Omega, osc = td.getEnergies(nstates=5, diag_conv=1e-10, diag_maxiter=100)
    assert wrapped.dipole_calls == 1                    # the numpy dipole matrix was used
    ref = O.solve_from_ground_state(gs, nstates=5, diag_conv=1e-10, diag_maxiter=100)
    np.testing.assert_allclose(Omega, ref["Omega"], atol=1e-10)
    # exact: the oracle's dense formula applied to this solve's own C
    own = O.transition_dipoles(td.Cij, gs.getKSCoefficients(), gs.getDipoleMatrix(), td.active_occupied_orbs,
                               td.active_virtual_orbs)
    for c in range(3):
        np.testing.assert_allclose((td.tdipX, td.tdipY, td.tdipZ)[c], own[c], rtol=0, atol=1e-12)
    # against the oracle's solve: C = bs Tb of the non-Hermitian Davidson depends on
    # the expansion space (the projected square root of A - B), beyond the residual;
    # two spaces that differ in rounding give f within ~1e-4 relative for lc = 1
    np.testing.assert_allclose(osc, ref["osc"], rtol=1e-7 if lc == 0 else 1e-3, atol=1e-10)
    D, orbs = gs.getDipoleMatrix(), gs.getKSCoefficients()
    np.testing.assert_allclose(td.tdipMO_Z, orbs.T @ D[:, :, 2] @ orbs, atol=1e-12)


def test_slako_kernel_against_reference_fortran_golden():
    """dftb_slako_matrices (S, H0, dipole planes) against the reference Fortran's
    H0andS / DipoleMatrix outputs (tests/golden/slako_f90_golden.npz): displaced
    C8H10, the same rotated, molecules 10-35 bohr apart (clamped table ends,
    proximity cutoff)."""
    from dftbaby_amd.groundstate import slako as SK
    gd = np.load(os.path.join(ROOT, "tests", "golden", "slako_f90_golden.npz"))
    for name in ("c8h10", "c8h10_rotated", "far_pairs"):
        Z, pos = gd[name + "_Z"], gd[name + "_pos"]
        S, H, D = SK.matrices_device(Z, pos, "cuda")
        np.testing.assert_allclose(S.cpu().numpy(), gd[name + "_S"], rtol=0, atol=1e-15, err_msg=name)
        np.testing.assert_allclose(H.cpu().numpy(), gd[name + "_H0"], rtol=0, atol=1e-15, err_msg=name)
        Dh = np.moveaxis(D.cpu().numpy(), 0, 2)
        np.testing.assert_allclose(Dh, gd[name + "_D"], rtol=4e-15, atol=1e-15, err_msg=name)


def test_slako_kernel_flake_against_oracle_and_host():
    """The kernel on the C604H68 flake (Nat 672, Norb 2484, pairs beyond the
    10 bohr table end and the 30 bohr cutoff) against the producer's host
    assembly (exactly symmetric, equal to a few ulps: FMA contraction), and the
    first 24 atoms against the oracle's scalar restatement."""
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate import slako as SK
    from oracle import slako_oracle as K
    fl = g.flake_c604h68()
    Z = [a for a, _ in fl]
    pos = np.array([x for _, x in fl])
    S, H, D = SK.matrices_device(Z, pos, "cuda")
    S, H, D = S.cpu().numpy(), H.cpu().numpy(), np.moveaxis(D.cpu().numpy(), 0, 2)
    assert np.array_equal(S, S.T) and np.array_equal(H, H.T)
    assert all(np.array_equal(D[:, :, c], D[:, :, c].T) for c in range(3))
    Sh, Hh, Dh = SK.matrices_host(Z, pos)
    np.testing.assert_allclose(S, Sh, rtol=8e-15, atol=1e-15)
    np.testing.assert_allclose(H, Hh, rtol=8e-15, atol=1e-15)
    np.testing.assert_allclose(D, Dh, rtol=8e-15, atol=1e-15)
    # oracle on a corner of the flake: the first 24 atoms (their pairs with every
    # other atom are covered by the host comparison above)
    sub = list(range(24))
    So, Ho = K.h0_and_s([Z[i] for i in sub], pos[sub])
    Do = K.dipole_matrix([Z[i] for i in sub], pos[sub], So)
    n = len(So)
    np.testing.assert_allclose(S[:n, :n], So, rtol=8e-15, atol=1e-15)
    np.testing.assert_allclose(H[:n, :n], Ho, rtol=8e-15, atol=1e-15)
    np.testing.assert_allclose(D[:n, :n], Do, rtol=8e-15, atol=1e-15)


def test_bundled_producer_dipoles_from_device_planes():
    """The bundled producer hands LR_TDDFTB the kernel's dipole planes; the
    transition dipoles equal the oracle's formula with the host dipole matrix."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", 1, 0.05, 23)
    td = LR_TDDFTB(ground_state=gs)
    td.getEnergies(nstates=5, diag_conv=1e-10, diag_maxiter=100)
    own = O.transition_dipoles(td.Cij, gs.getKSCoefficients(), gs.getDipoleMatrix(), td.active_occupied_orbs,
                               td.active_virtual_orbs)
    for c in range(3):
        np.testing.assert_allclose((td.tdipX, td.tdipY, td.tdipZ)[c], own[c], rtol=0, atol=1e-12)


@pytest.mark.parametrize("lc", [0, 1])
def test_davidson_triplet_request_computes_singlets(lc):
    """nstates with multiplicity="T": neither Davidson receives the multiplicity
    (LR_TDDFTB.py:471-492), so both return the singlet roots, with oscillator
    strengths 0 (LR_TDDFTB.py:700-705)."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", lc, 0.05, 29)
    gs.verbose = 0
    wS, fS = LR_TDDFTB(ground_state=gs).getEnergies(nstates=4, diag_conv=1e-10, diag_maxiter=100)
    wT, fT = LR_TDDFTB(ground_state=gs).getEnergies(nstates=4, diag_conv=1e-10, diag_maxiter=100, multiplicity="T")
    np.testing.assert_allclose(wT, wS, atol=1e-12)
    assert np.all(fT == 0.0) and fS.max() > 0.0


def test_select_lm_pi_space_vs_oracle():
    """getEnergies(select_lm=[(1, 0)]) on the planar C8H10 (LR_TDDFTB.py:125-166):
    the pi -> pi* space only, against the oracle's dense Casida on the same space."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", 1)
    gs.verbose = 0
    td = LR_TDDFTB(ground_state=gs)
    Omega, osc = td.getEnergies(select_lm=[(1, 0)], nr_active_occ=1000, nr_active_virt=1000, nstates=5,
                                diag_conv=1e-10, diag_maxiter=100)
    ref = O.solve_from_ground_state(gs, nstates=None, nr_active_occ=1000, nr_active_virt=1000, select_lm=[(1, 0)])
    assert (td.nocc, td.nvirt) == (4, 4)
    np.testing.assert_allclose(Omega, ref["Omega"][:5], atol=1e-10)
    np.testing.assert_allclose(osc, ref["osc"][:5], rtol=1e-7, atol=1e-10)


def test_save_transition_charges(tmp_path):
    """saveTransitionCharges (LR_TDDFTB.py:1008-1026): the three header lines, then
    -sum_ia q^{ia}_A C^I_ia per atom (rows) and state (columns), as np.savetxt
    writes it, from the oracle's charges and the solve's own C."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", 1, 0.05, 31)
    gs.verbose = 0
    td = LR_TDDFTB(ground_state=gs)
    td.getEnergies(nstates=4, diag_conv=1e-9, diag_maxiter=100)
    f = tmp_path / "tq.dat"
    assert td.saveTransitionCharges(None) is None and not f.exists()
    td.saveTransitionCharges(str(f))
    text = f.read_text().split("\n")
    assert text[:3] == ["# Mulliken transition charges (in unit of -e) between ground and excited states",
                        "# ROWS: atoms", "# COLUMNS: excited states"]
    got = np.loadtxt(str(f))
    ref = O.solve_from_ground_state(gs, nstates=4, diag_conv=1e-9, diag_maxiter=100)
    expect = -np.tensordot(ref["q_ov"], td.Cij, axes=([1, 2], [1, 2]))
    assert got.shape == (len(gs.getGeometry()), 4)
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-12)
