"""Pin the CPU oracle (oracle/tddftb_oracle.py) before trusting it.

* against DFTBaby's own Fortran kernels (DFTB/extensions/tddftb.f90
  trans_charges / apbv / ambv) compiled by oracle/Makefile into oracle/_ref,
  when that build exists, and against committed golden vectors it produced
  (tests/golden/tddftb_f90_golden.npz, written by tests/golden/make_golden.py);
* against closed forms / invariants: sum_A q_A^{pq} = (C^T S C)_{pq} = delta_pq,
  q^{pq} = q^{qp}, brute-force loops, dense buildAB products;
* Davidson vs dense Casida (the reference's diag_check criterion).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, ground_state
from oracle import tddftb_oracle as O

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtddftb_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "tddftb_f90_golden.npz")


def _active(gs):
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    return f, orbe, C, S, ao, av


def f90_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference tree and amdflang)")
    return ctypes.CDLL(REF_SO)


def f90_trans_charges(lib, Co, Cv, S, opa):
    norb, no = Co.shape
    nv = Cv.shape[1]
    nat = len(opa)
    F = lambda a: np.asfortranarray(a, dtype=np.float64)
    Co_f, Cv_f, S_f = F(Co), F(Cv), F(S)
    opa_f = np.asfortranarray(opa, dtype=np.int32)
    q_oo = np.zeros((nat, no, no), order="F")
    q_ov = np.zeros((nat, no, nv), order="F")
    q_vv = np.zeros((nat, nv, nv), order="F")
    ip = lambda x: ctypes.byref(ctypes.c_int(x))
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib._QMtddftbPtrans_charges(dp(Co_f), dp(Cv_f), dp(S_f), dp(opa_f), ip(nat), ip(norb), ip(no), ip(nv),
                                dp(q_oo), dp(q_ov), dp(q_vv))
    return q_oo, q_ov, q_vv


def f90_products(lib, g, glr, q_oo, q_ov, q_vv, omega, vs):
    nat, no, nv = q_ov.shape
    nvec = vs.shape[2]
    F = lambda a: np.asfortranarray(a, dtype=np.float64)
    args = [F(g), F(glr), F(q_oo), F(q_ov), F(q_vv), F(omega), F(vs)]
    ip = lambda x: ctypes.byref(ctypes.c_int(x))
    dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    out = []
    for name in ("_QMtddftbPapbv", "_QMtddftbPambv"):
        us = np.zeros((no, nv, nvec), order="F")
        getattr(lib, name)(*[dp(a) for a in args], ip(nat), ip(no), ip(nv), ip(nvec), dp(us))
        out.append(np.array(us))
    return out


# --------------------------------------------------------------------------- a2
@pytest.mark.parametrize("name,lc", [("ethylene", 1), ("benzene", 1), ("C8H10", 0)])
def test_trans_charges_vs_reference_fortran(name, lc):
    lib = f90_lib()
    gs = ground_state(name, lc)
    f, orbe, C, S, ao, av = _active(gs)
    ref = f90_trans_charges(lib, C[:, ao], C[:, av], S, gs.orbsPerAtom)
    got = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_trans_charges_golden_vectors():
    """Committed Fortran outputs (usable where oracle/_ref cannot be built)."""
    if not os.path.exists(GOLDEN):
        pytest.skip("golden vectors not generated")
    d = np.load(GOLDEN)
    got = O.trans_charges(d["Co"], d["Cv"], d["S"], d["opa"])
    for a, key in zip(got, ("q_oo", "q_ov", "q_vv")):
        np.testing.assert_allclose(a, d[key], rtol=0, atol=1e-14)
    ApB = O.ApBv(d["g"], d["glr"], d["q_oo"], d["q_ov"], d["q_vv"], d["omega"], d["vs"], 1)
    AmB = O.AmBv(d["g"], d["glr"], d["q_oo"], d["q_ov"], d["q_vv"], d["omega"], d["vs"], 1)
    np.testing.assert_allclose(ApB, d["ApB"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(AmB, d["AmB"], rtol=0, atol=1e-12)


def test_trans_charges_invariants_and_bruteforce():
    gs = ground_state("C8H10", 1, 0.05, 3)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    # sum over atoms: (C^T S C)_{pq} = delta_pq
    np.testing.assert_allclose(q_oo.sum(0), np.eye(len(ao)), atol=1e-12)
    np.testing.assert_allclose(q_vv.sum(0), np.eye(len(av)), atol=1e-12)
    np.testing.assert_allclose(q_ov.sum(0), 0.0, atol=1e-12)
    np.testing.assert_allclose(q_oo, q_oo.transpose(0, 2, 1), atol=0)
    np.testing.assert_allclose(q_vv, q_vv.transpose(0, 2, 1), atol=0)
    # brute-force loop on a few orbitals
    sub_o, sub_v = ao[-3:], av[:4]
    bf = O.trans_charges_bruteforce(C[:, sub_o], C[:, sub_v], S, gs.orbsPerAtom)
    ve = O.trans_charges(C[:, sub_o], C[:, sub_v], S, gs.orbsPerAtom)
    for a, b in zip(bf, ve):
        np.testing.assert_allclose(a, b, atol=1e-14)
    # planted mistake: per-atom charges without the symmetrising term,
    # q^{ab}_A = sum_{mu in A} C_{mu a} (SC)_{mu b}.  They still sum to C^T S C = 1
    # over the atoms, so only the symmetry and brute-force checks can catch them.
    Cv, Scv = C[:, sub_v], (S @ C)[:, sub_v]
    amap = np.repeat(np.arange(len(gs.orbsPerAtom)), gs.orbsPerAtom)
    wrong = np.zeros_like(bf[2])
    np.add.at(wrong, amap, Cv[:, :, None] * Scv[:, None, :])
    np.testing.assert_allclose(wrong.sum(0), np.eye(len(sub_v)), atol=1e-12)
    assert np.abs(wrong - wrong.transpose(0, 2, 1)).max() > 1e-4
    assert np.abs(wrong - bf[2]).max() > 1e-4


# --------------------------------------------------------------------------- a7
def test_products_vs_reference_fortran(rng):
    lib = f90_lib()
    gs = ground_state("C8H10", 1, 0.05, 4)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    vs = rng.standard_normal((len(ao), len(av), 3))
    refP, refM = f90_products(lib, g, glr, q_oo, q_ov, q_vv, omega, vs)
    np.testing.assert_allclose(O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1), refP, rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.AmBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1), refM, rtol=0, atol=1e-12)


def test_products_vs_dense_buildAB(rng):
    gs = ground_state("C8H10", 1, 0.05, 4)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    A, B = O.buildAB(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    vs = rng.standard_normal((o, v, 2))
    P = O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    M = O.AmBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    for n in range(2):
        np.testing.assert_allclose(P[:, :, n].ravel(), (A + B) @ vs[:, :, n].ravel(), atol=1e-12)
        np.testing.assert_allclose(M[:, :, n].ravel(), (A - B) @ vs[:, :, n].ravel(), atol=1e-12)
    # symmetric A, B; lc=0 makes A-B diagonal
    np.testing.assert_allclose(A, A.T, atol=1e-13)
    np.testing.assert_allclose(B, B.T, atol=1e-13)
    A0, B0 = O.buildAB(g, 0 * glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    np.testing.assert_allclose(A0 - B0, np.diag(omega.ravel()), atol=1e-14)


# --------------------------------------------------------------------------- a6/a7/a8
@pytest.mark.parametrize("name,lc", [("C8H10", 1), ("C12H14", 1)])
def test_ao_exchange_identity(name, lc, rng):
    """The AO-basis exchange sequence the GPU runs (four GEMMs per term over the
    stacked [Y; C] coefficients, gamma_lr expanded to orbital pairs as a Hadamard
    factor) equals the reference's transition-charge products exactly."""
    gs = ground_state(name, lc, 0.05, 3)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    Co, Cv = C[:, ao], C[:, av]
    q_oo, q_ov, q_vv = O.trans_charges(Co, Cv, S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    V = rng.standard_normal((o, v))
    refP = O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, V[:, :, None], 1)[:, :, 0]
    refM = O.AmBv(g, glr, q_oo, q_ov, q_vv, omega, V[:, :, None], 1)[:, :, 0]
    Yo, Yv = S @ Co, S @ Cv
    amap = np.repeat(np.arange(len(gs.orbsPerAtom)), gs.orbsPerAtom)
    G = glr[np.ix_(amap, amap)]
    G2 = np.block([[G, G], [G, G]])
    Av, Ao = np.vstack([Yv, Cv]), np.vstack([Yo, Co])
    CYvT, W = np.vstack([Cv, Yv]).T, np.vstack([Co, Yo]).T
    X1 = 0.25 * W @ (CYvT @ (G2 * (Ao @ (Av @ V.T).T)).T).T
    X2 = 0.25 * W @ (CYvT @ (G2 * (Av @ (Ao @ V).T)).T).T
    coul = np.einsum("Aia,AB,Bjb,jb->ia", q_ov, g, q_ov, V)
    P = omega * V + 4 * coul - X1 - X2
    M = omega * V - X1 + X2
    assert np.abs(P - refP).max() < 1e-13 * np.abs(refP).max()
    assert np.abs(M - refM).max() < 1e-13 * np.abs(refM).max()
    # the six-GEMM form the GPU runs: the K2 bracket is the transpose of the K1
    # bracket (gamma_lr symmetric), so X2 follows from BH in the NN form
    BH = G2 * (Ao @ (Av @ V.T).T)
    assert np.abs(G2 * (Av @ (Ao @ V).T) - BH.T).max() < 1e-13 * np.abs(BH).max()
    X2s = 0.25 * W @ (CYvT @ BH).T
    assert np.abs(X2s - X2).max() < 1e-13 * np.abs(X2).max()


def test_casida_self_consistency():
    """Casida's own asserts (Solver.py:145-173) on an lc geometry."""
    gs = ground_state("C8H10", 1, 0.05, 2)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    Om, Cc, XmY, XpY = O.Casida(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    A, B = O.buildAB(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    assert np.all(np.diff(Om) >= 0) and Om[0] > 0
    for n in range(5):
        xp, xm = XpY[n].ravel(), XmY[n].ravel()
        assert abs(xp @ xm - 1) < 1e-10
        np.testing.assert_allclose((A + B) @ xp, Om[n] * xm, atol=1e-9)
        np.testing.assert_allclose((A - B) @ xm, Om[n] * xp, atol=1e-9)
        assert abs(Cc[n].ravel() @ Cc[n].ravel() - 1) < 1e-10


@pytest.mark.parametrize("lc", [0, 1])
def test_davidson_matches_casida(lc):
    """The reference's diag_check: eigenvalues < 1e-10, X+-Y up to a global sign."""
    gs = ground_state("C8H10", lc, 0.05, 9)
    full = O.solve_from_ground_state(gs, nstates=None)
    dav = O.solve_from_ground_state(gs, nstates=5, diag_conv=1e-11, diag_maxiter=100)
    assert np.abs(dav["Omega"] - full["Omega"][:5]).sum() < 1e-10
    err = 0.0
    for n in range(5):
        for key in ("XmY", "XpY"):
            a, b = dav[key][n], full[key][n]
            err += min(np.abs(a - b).sum(), np.abs(a + b).sum())
    assert err < 1e-6


def test_oscillator_strengths_closed_form():
    gs = ground_state("C8H10", 0, 0.05, 1)
    r = O.solve_from_ground_state(gs, nstates=None)
    C = r["C"][:4]
    orbs, D = gs.getKSCoefficients(), gs.getDipoleMatrix()
    ao, av = r["act_occ"], r["act_virt"]
    for c in range(3):
        dense = np.einsum("nia,mi,mlc,la->n", C, orbs[:, ao], D[..., c:c + 1], orbs[:, av])
        np.testing.assert_allclose(r["tdip"][c][:4], dense, atol=1e-12)
    f = 2.0 / 3.0 * r["Omega"][:4] * sum(r["tdip"][c][:4] ** 2 for c in range(3))
    np.testing.assert_allclose(r["osc"][:4], f, rtol=1e-14)
    assert O.oscillator_strengths(r["Omega"], r["tdip"], "T").max() == 0.0


def test_oia_loop_and_lambda2():
    gs = ground_state("C8H10", 1, 0.05, 6)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    Om = gs.getGaussianOverlap()
    np.testing.assert_allclose(O.Lambda2_calcOia(q_oo, q_vv, Om), O.Lambda2_calcOia_loop(q_oo, q_vv, Om),
                               atol=1e-14)
    Oia = O.Lambda2_calcOia(q_oo, q_vv, Om)
    assert np.all(Oia >= 0) and np.all(Oia <= 1 + 1e-12)


def test_davidson_not_converged_raises():
    gs = ground_state("C8H10", 1, 0.05, 9)
    with pytest.raises(O.ExcitedStatesNotConverged):
        O.solve_from_ground_state(gs, nstates=5, diag_conv=1e-14, diag_maxiter=1)


def test_mlp_oracle_gradient_finite_differences():
    """The CPU MLP reference against central differences (pins the GPU gradient test)."""
    from oracle import mlp_oracle
    rng = np.random.default_rng(0)
    shapes = [(6, 8), (8,), (8, 8), (8,), (8, 8), (8,), (8, 2), (2,)]
    n = sum(int(np.prod(s)) for s in shapes)
    p = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((7, 6))
    T = rng.standard_normal((7, 2))
    w = np.array([1.0, 0.3])
    loss, g = mlp_oracle.loss_and_grad(p, shapes, X, T, w)
    h = 1e-6
    for i in rng.choice(n, 25, replace=False):
        e = np.zeros(n)
        e[i] = h
        fd = (mlp_oracle.loss_and_grad(p + e, shapes, X, T, w)[0] - mlp_oracle.loss_and_grad(p - e, shapes, X, T, w)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd))


def test_ct_shifts_separated_fragments():
    """Long-range CT correction (LongRangeCT.py:13-76) on C8H10 + H2 40 bohr apart:
    every orbital lives on one centrosymmetric fragment, so an excitation between
    the fragments has S_ia -> 0 and R_ia = the distance of the fragment centres
    (40 bohr), and T_ia must equal the asymptotic form exp(-(Rc/R)^2) (D - 1/R +
    (D - 1/R)^2 / (2 w)) with D = -eps_a (clipped at 0); excitations inside a
    fragment with R_ia ~ 0 get no shift.  The centres are not exact: the
    reference's assembly forms each C-C element with the later atom's orbital
    first, and its sp and ps hamiltonian tables are not exact negatives of each
    other (1e-5 relative), so H0 breaks the inversion symmetry of the polyene;
    near-degenerate orbital pairs mix under that perturbation and their centres
    move by up to 0.1 bohr (the cross-fragment distance is kept to 1e-5)."""
    gs = ground_state("C8H10-H2", 0)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    pos = np.array([p for _, p in gs.getGeometry()])
    T, Sia, Ria = O.ct_shifts(q_oo, q_vv, gs.getGaussianOverlap(), pos, orbe[av], omega)
    # fragment of each orbital from its Mulliken weight on the two H2 atoms (the last two)
    w_h2 = np.einsum("mp,mp->p", C[-2:], (S @ C)[-2:])
    assert np.all((w_h2 < 1e-8) | (w_h2 > 1 - 1e-8))
    on_h2 = w_h2 > 0.5
    cross = on_h2[ao][:, None] != on_h2[av][None, :]
    R = np.linalg.norm(pos[-2:].mean(axis=0) - pos[:-2].mean(axis=0))
    assert abs(R - 40.0) < 1e-12
    np.testing.assert_allclose(Ria[cross], R, rtol=1e-5)
    assert np.abs(Sia[cross]).max() < 1e-30
    D = -np.broadcast_to(orbe[av][None, :], omega.shape)[cross]
    x = D - 1.0 / R
    closed = np.clip(np.exp(-(0.1 / R) ** 2) * (x + x ** 2 / (2 * omega[cross])), 0.0, None)
    np.testing.assert_allclose(T[cross], closed, rtol=1e-6, atol=1e-14)
    assert T[cross].max() > 0.04                       # H2 sigma -> C8H10 LUMO is shifted
    intra_T, intra_R = T[~cross], Ria[~cross]
    centred = intra_R < 1e-3
    assert centred.any() and np.all(intra_T[centred] == 0.0)


def test_zvector_krylov_solves_dense_system(rng):
    """Krylov_solver_Zvector (Solver.py:921-1012) with the reference's (A+B)
    products reproduces the dense solution (A+B)^{-1} B of buildAB's matrices."""
    gs = ground_state("C8H10", 1, 0.05, 2)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    A, B = O.buildAB(g, glr, q_oo, q_vv, q_ov, omega, 0 * omega, df, o, v)
    rhs = rng.standard_normal((o, v, 2))
    X = O.Krylov_solver_Zvector(lambda vs: O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1), omega, rhs, None,
                                conv=1e-10)
    Xd = np.linalg.solve(A + B, rhs.reshape(o * v, 2)).reshape(o, v, 2)
    assert np.abs(X - Xd).max() < 1e-9 * np.abs(Xd).max()


def test_stability_matrix_elementwise():
    """StabilityAnalysis.py:24-56: the stability matrix of the oracle equals its
    element-wise definition R_{ia,jb} = w_ia d_ij d_ab - sum_AB q^ij_A g_AB q^ab_B
    - sum_AB q^ia_A glr_AB q^jb_B on ethylene (lc = 1)."""
    gs = ground_state("ethylene", 1, 0.05, 4)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    o, v = omega.shape
    w, R = O.check_stability(g, glr, q_oo, q_vv, q_ov, omega, o, v)
    Rl = np.zeros((o, v, o, v))
    for i in range(o):
        for a in range(v):
            for j in range(o):
                for b in range(v):
                    Rl[i, a, j, b] = ((omega[i, a] if (i == j and a == b) else 0.0)
                                      - q_oo[:, i, j] @ g @ q_vv[:, a, b] - q_ov[:, i, a] @ glr @ q_ov[:, j, b])
    assert np.abs(R - Rl.reshape(o * v, o * v)).max() < 1e-13
    np.testing.assert_allclose(w, np.linalg.eigvalsh(Rl.reshape(o * v, o * v))[:len(w)], atol=1e-13)


def test_hermitian_davidson_with_ct_shift_matches_dense():
    """HermitianDavidson with omega_shift (Solver.py:404-578: diagonal w^2 + 2 w T)
    finds the lowest roots of the dense Casida with the same shift (Solver.py:46-49:
    T on the diagonals of A and B), and the shift moves them."""
    gs = ground_state("C8H10-H2", 0)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, df = O.orbital_differences(orbe, f, ao, av)
    pos = np.array([p for _, p in gs.getGeometry()])
    T, _, _ = O.ct_shifts(q_oo, q_vv, gs.getGaussianOverlap(), pos, orbe[av], omega)
    g = gs.getGamma()
    o, v = omega.shape
    Oia = O.Lambda2_calcOia(q_oo, q_vv, gs.getGaussianOverlap())
    dense = O.Casida(g, 0 * g, q_oo, q_vv, q_ov, omega, T, df, o, v)[0]
    dense0 = O.Casida(g, 0 * g, q_oo, q_vv, q_ov, omega, 0 * T, df, o, v)[0]
    w = O.HermitianDavidson(g, q_ov, omega, T, o, v, None, None, Oia, nstates=14, ifact=4, conv=1e-11,
                            maxiter=200, L2_thresh=0.0)[0]
    np.testing.assert_allclose(w, dense[:14], atol=1e-10)
    assert np.abs(dense[:14] - dense0[:14]).max() > 0.01


# --------------------------------------------------------------------------- a9
def test_dipoles_equal_mulliken_on_point_charge_dipole_matrix(rng):
    """With the dipole matrix D^c = S o (R^c_mu + R^c_nu) / 2 of AO centres,
    <i|r_c|a> = sum_A R^c_A q^{ia}_A exactly (expand q^{ia}_A = 1/2 sum_{mu in A}
    [C_mu,i (SC)_mu,a + C_mu,a (SC)_mu,i] and swap mu, nu in the second term), so
    the "Dipoles" (LR_TDDFTB.py:574-592) and "Mulliken" (567-573) transition dipoles
    of the oracle must agree for any coefficients C^I_ia.  Planted mistake: the
    unsymmetrised D = S o R_mu breaks the equality."""
    gs = ground_state("C8H10", 1, 0.05, 21)
    f, orbe, C, S, ao, av = _active(gs)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    pos = np.array([p for _, p in gs.getGeometry()])
    Rmu = pos[np.repeat(np.arange(len(pos)), gs.orbsPerAtom)]
    D = S[:, :, None] * 0.5 * (Rmu[:, None, :] + Rmu[None, :, :])
    Cij = rng.standard_normal((5, len(ao), len(av)))
    dip = O.transition_dipoles(Cij, C, D, ao, av)
    mul = O.transition_dipoles_mulliken(Cij, q_ov, pos)
    scale = max(np.abs(d).max() for d in dip)
    for c in range(3):
        assert np.abs(dip[c] - mul[c]).max() < 1e-12 * scale
    wrongD = S[:, :, None] * Rmu[:, None, :]
    wrong = O.transition_dipoles(Cij, C, wrongD, ao, av)
    assert max(np.abs(wrong[c] - mul[c]).max() for c in range(3)) > 1e-3 * scale


# --------------------------------------------------------------------------- a1
def _active_bruteforce(f, orbe, nocc_act, nvirt_act):
    """Top-k by pairwise comparison: occupied p (f > 0.1) is active when fewer than
    nocc_act occupied orbitals rank above it in (energy, index); virtual a when
    fewer than nvirt_act rank below it.  Both sets ascending in (energy, index)."""
    n = len(f)
    occ = [p for p in range(n) if f[p] > 0.1]
    virt = [p for p in range(n) if f[p] <= 0.1]
    key = lambda p: (orbe[p], p)
    act_o = [p for p in occ if nocc_act is None or sum(key(q) > key(p) for q in occ) < nocc_act]
    act_v = [p for p in virt if nvirt_act is None or sum(key(q) < key(p) for q in virt) < nvirt_act]
    return np.array(sorted(act_o, key=key), dtype=int), np.array(sorted(act_v, key=key), dtype=int)


@pytest.mark.parametrize("seed", range(6))
def test_active_orbitals_truncation_bruteforce(seed):
    """setActiveOrbitals (LR_TDDFTB.py:167-188): the highest nr_active_occ occupied
    and lowest nr_active_virt virtual orbitals, each ascending in energy, on
    shuffled energies with ties and fractional occupations; the oracle and the
    drop-in's host selection against the pairwise brute force (index work:
    exact)."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    r = np.random.default_rng(seed)
    n = 23
    orbe = np.round(r.standard_normal(n), 1)          # rounding makes ties
    f = r.choice([0.0, 0.05, 0.1, 0.5, 1.0, 2.0], size=n)
    for no, nv in [(None, None), (3, 4), (1, 1), (50, 2), (0, 5)]:
        ref_o, ref_v = _active_bruteforce(f, orbe, no, nv)
        _, _, ao, av = O.active_orbitals(f, orbe, no, nv)
        np.testing.assert_array_equal(ao, ref_o)
        np.testing.assert_array_equal(av, ref_v)

        class GS:
            getOccupation = staticmethod(lambda: f)
            getKSEnergies = staticmethod(lambda: orbe)
        td = LR_TDDFTB(ground_state=GS())
        td.setActiveOrbitals(no, nv)
        np.testing.assert_array_equal(td.active_occupied_orbs, ref_o)
        np.testing.assert_array_equal(td.active_virtual_orbs, ref_v)
        assert td.nocc == len(ref_o) and td.nvirt == len(ref_v)
    # planted mistake: the lowest instead of the highest occupied orbitals
    occ = np.where(f > 0.1)[0]
    wrong = occ[np.argsort(orbe[occ], kind="stable")[:3]]
    assert not np.array_equal(wrong, _active_bruteforce(f, orbe, 3, 4)[0])


def test_active_lm_orbitals_pi_system_of_planar_polyene():
    """setActiveLMOrbitals (LR_TDDFTB.py:125-166) with select_lm = [(1, 0)] (p_z)
    on the planar (z = 0) C8H10: the pi system is exactly the 8 MOs with p_z
    weight (4 occupied, 4 virtual); the oracle and the drop-in's host selection
    agree with a count of p_z weights, and truncation keeps the frontier pi
    orbitals."""
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", 1)
    orbs, f, orbe = gs.getKSCoefficients(), gs.getOccupation(), gs.getKSEnergies()
    pz = [mu for mu, (A, k) in enumerate((A, k) for A, (Z, _) in enumerate(gs.getGeometry())
                                         for k in range(len(gs.getValorbs()[Z])))
          if gs.getValorbs()[gs.getGeometry()[A][0]][k][1:] == (1, 0)]
    assert len(pz) == 8
    w = np.array([sum(orbs[mu, p] ** 2 for mu in pz) for p in range(orbs.shape[1])])
    pi = np.where(w > 1e-15)[0]
    assert len(pi) == 8 and np.all((w < 1e-20) | (w > 1e-3))
    occ, virt, ao, av = O.active_lm_orbitals(orbs, f, orbe, gs.getGeometry(), gs.getValorbs(), 1000, 1000, [(1, 0)])
    np.testing.assert_array_equal(occ, pi[f[pi] > 0.1])
    np.testing.assert_array_equal(virt, pi[f[pi] <= 0.1])
    assert len(ao) == 4 and len(av) == 4
    _, _, ao2, av2 = O.active_lm_orbitals(orbs, f, orbe, gs.getGeometry(), gs.getValorbs(), 2, 3, [(1, 0)])
    np.testing.assert_array_equal(ao2, ao[-2:])
    np.testing.assert_array_equal(av2, av[:3])
    td = LR_TDDFTB(ground_state=gs)
    td.setActiveLMOrbitals(2, 3, select_lm=[(1, 0)])
    np.testing.assert_array_equal(td.active_occupied_orbs, ao2)
    np.testing.assert_array_equal(td.active_virtual_orbs, av2)
    np.testing.assert_array_equal(td.occupied_orbs, occ)
