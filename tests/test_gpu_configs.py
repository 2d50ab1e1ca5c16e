"""The workloads the engine is measured on, at their real sizes (run with -m gpu).

* Flake C604H68 (Nat 672, Norb 2484, lc, R_lr 3.03), reduced active space 60 x 60:
  the GPU Davidson (FP64 and BF16x3) against the oracle's dense Casida
  (Solver.py:98-181), the comparison the reference's diag_check makes
  (LR_TDDFTB.py:493-518): energies, and X+Y up to sign for non-degenerate roots
  to the accuracy the residual tolerance gives them (residual / gap).
* Flake C604H68, full space (o = v = 1242, n_ov 1.54e6), 20 singlets: FP64 and
  BF16x3 AO-exchange solves; per-state energy and oscillator-strength errors of
  the mixed solve; an FP64 residual certificate of the FP64 solve computed with
  the transition-charge products in the reference's contraction order
  (independent of the AO chain the solve used); (X-Y)^T (X+Y) = 1.
* AO vs transition-charge exchange, both FP64, full Davidson on the C180H36 flake.
* The uneven ensemble: 256 polyenes C_nH_{n+2} (n even 8..40, sigma 0.05 A), 8
  singlets, tolerance 1e-6, one persistent launch: every member's roots certified
  by their FP64 residual and biorthonormality, 32 members (C40H42 among them)
  against the single-molecule GPU FP64 Davidson, three small ones against the
  oracle Davidson.
"""
import numpy as np
import pytest

from conftest import progress
from oracle import tddftb_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

_GS = {}


def _flake():
    """SCC ground state of the flake (on the GPU; cached for the module)."""
    if "flake" not in _GS:
        from dftbaby_amd.groundstate import geometry as g
        from dftbaby_amd.groundstate.dftb2 import DFTB2
        gs = DFTB2(g.flake_c604h68(), long_range_correction=1, long_range_radius=3.03)
        gs.getEnergy(scf_conv=1e-8)
        gs.verbose = 0
        _GS["flake"] = gs
    return _GS["flake"]


def _resident(gs):
    """The SCC state is converged once; later getEnergies calls reuse it."""
    gs.getEnergy = lambda **kw: gs.en_elec
    return gs


def _degenerate_groups(Omega, tol=1e-4):
    groups, cur = [], [0]
    for n in range(1, len(Omega)):
        if Omega[n] - Omega[n - 1] < tol:
            cur.append(n)
        else:
            groups.append(cur)
            cur = [n]
    groups.append(cur)
    return groups


@pytest.mark.parametrize("precision,conv,tol", [("fp64", 1e-9, 1e-10), ("bf16x3", 1e-7, 1e-6)])
def test_flake_reduced_active_space_vs_dense_casida(precision, conv, tol):
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = _resident(_flake())
    progress("flake 60x60 active space, %s" % precision)
    td = LR_TDDFTB(ground_state=gs)
    Omega, osc = td.getEnergies(nr_active_occ=60, nr_active_virt=60, nstates=20, diag_conv=conv, diag_maxiter=100,
                                precision=precision)
    progress("GPU solve done; oracle dense Casida (n_ov 3600)")
    if "ref60" not in _GS:
        _GS["ref60"] = O.solve_from_ground_state(gs, nstates=None, nr_active_occ=60, nr_active_virt=60)
    ref = _GS["ref60"]
    assert td.nocc == 60 and td.nvirt == 60
    np.testing.assert_array_equal(td.active_occupied_orbs, ref["act_occ"])
    np.testing.assert_array_equal(td.active_virtual_orbs, ref["act_virt"])
    dE = np.abs(Omega - ref["Omega"][:20])
    progress("max |dE| = %.3e Ha" % dE.max())
    assert dE.max() < tol
    # f comes from C = bs Tb of the non-Hermitian Davidson (Solver.py:906-910), which
    # carries the expansion space's projected (A-B)^{1/2}: it approaches the dense
    # Casida's C with the basis, not with the residual (up to 2e-2 relative here,
    # for the BF16x3 run refined from its own roots).
    # Exact: f from this solve's own C through the oracle's dipole formula.
    tdip = O.transition_dipoles(td.Cij, gs.getKSCoefficients(), gs.getDipoleMatrix(), td.active_occupied_orbs,
                                td.active_virtual_orbs)
    np.testing.assert_allclose(osc, O.oscillator_strengths(Omega, tdip), rtol=1e-10, atol=1e-14)
    df = []
    for grp in _degenerate_groups(ref["Omega"][:20]):
        df.append(abs(osc[grp].sum() - ref["osc"][grp].sum()))
        assert df[-1] < 3e-2 * ref["osc"][grp].sum() + 1e-4
        if len(grp) == 1:
            n = grp[0]
            a, b = td.XpY[n], ref["XpY"][n]
            assert min(np.abs(a - b).max(), np.abs(a + b).max()) < (1e-6 if precision == "fp64" else 1e-3)
    progress("max |d sum f| per group vs the dense Casida: %.2e" % max(df))


def test_flake_full_space_fp64_mixed_and_residual_certificate():
    import ctypes
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    from dftbaby_amd.engine import CasidaOperatorGPU, gram
    gs = _resident(_flake())
    conv = 1e-7
    progress("flake full space, fp64")
    td = LR_TDDFTB(ground_state=gs)
    Om64, f64 = td.getEnergies(nstates=20, diag_conv=conv, diag_maxiter=100, precision="fp64")
    Om64, f64 = Om64.copy(), f64.copy()
    op = td.op
    XpY, XmY, w = td._XpY_dev.clone(), td._XmY_dev.clone(), td._w_dev.clone()
    assert op.nocc == 1242 and op.nvirt == 1242 and op.nat == 672
    # (X-Y)^T (X+Y) = 1
    G = torch.empty((20, 20), dtype=torch.float64, device=op.device)
    gram(XmY, 20, XpY, 20, op.N, G, op.ws)
    assert np.abs(G.cpu().numpy() - np.eye(20)).max() < 1e-8
    # residual certificate with the transition-charge (reference-order) products in FP64
    orbs, S = gs.getKSCoefficients(), gs.getOverlapMatrix()
    ao, av = td.active_occupied_orbs, td.active_virtual_orbs
    ref_op = CasidaOperatorGPU.from_orbitals(orbs[:, ao], orbs[:, av], S, gs.orbsPerAtom, td.omega,
                                             gs.getGamma(), gs.getGamma_lr(), 1, exchange="charges")
    progress("reference-order FP64 products of X+-Y")
    P, _ = ref_op.matvec(XpY, 0)
    _, M = ref_op.matvec(XmY, 0)
    wc = w[:, None]
    rL = torch.linalg.norm(P - wc * XmY, dim=1).cpu().numpy()
    rR = torch.linalg.norm(M - wc * XpY, dim=1).cpu().numpy()
    progress("certificate (reference-order FP64 products): max ||(A+B)(X+Y) - w(X-Y)|| + ||(A-B)(X-Y) - w(X+Y)|| "
          "= %.3e" % (rL + rR).max())
    assert (rL + rR).max() < 10 * conv
    del ref_op, P, M
    torch.cuda.empty_cache()
    # mixed precision with FP64 refinement at the same tolerance: energies
    progress("flake full space, bf16x3 at diag_conv 1e-7 (refined with FP64 products when it stalls)")
    td2 = LR_TDDFTB(ground_state=gs)
    Om, f = td2.getEnergies(nstates=20, diag_conv=conv, diag_maxiter=100, precision="bf16x3")
    progress("per-state |dE| (Ha): " + " ".join("%.1e" % x for x in np.abs(Om - Om64)))
    assert np.abs(Om - Om64).max() <= 1e-6
    # the measured setting (diag_conv 1e-5): per-state energy and oscillator-strength
    # errors of the mixed solve against the FP64 solve.  f comes from C = bs Tb
    # (Solver.py:906-910), which carries the expansion space's projected
    # (A-B)^(1/2) and converges more slowly than the residual: two runs that build
    # different spaces (680 vs 727 vectors here) agree on f to ~1e-3, so the sums
    # over near-degenerate groups are held to 5e-3 (energies to 1e-6 Ha)
    res = {}
    for prec in ("fp64", "bf16x3"):
        t = LR_TDDFTB(ground_state=gs)
        res[prec] = t.getEnergies(nstates=20, diag_conv=1e-5, diag_maxiter=100, precision=prec)
        res[prec] = (res[prec][0].copy(), res[prec][1].copy(), t.davidson_stats.refined)
    (E64, f64, _), (Em, fm, refined) = res["fp64"], res["bf16x3"]
    assert not refined
    dE, df = np.abs(Em - E64), np.abs(fm - f64)
    progress("diag_conv 1e-5, per-state |dE| (Ha): " + " ".join("%.1e" % x for x in dE))
    progress("diag_conv 1e-5, per-state |df|     : " + " ".join("%.1e" % x for x in df))
    assert dE.max() <= 1e-6
    for grp in _degenerate_groups(E64):
        assert abs(fm[grp].sum() - f64[grp].sum()) < 5e-3


def test_ao_and_charges_exchange_agree_fp64():
    """Full FP64 Davidson with both exchange formulations on the C180H36 flake."""
    from conftest import ground_state
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C180H36", 1, 0.05, 11)
    gs.verbose = 0
    out = {}
    for exch in ("ao", "charges"):
        td = LR_TDDFTB(ground_state=gs)
        td.exchange = exch
        Om, f = td.getEnergies(nstates=10, diag_conv=1e-9, diag_maxiter=100, precision="fp64")
        out[exch] = (Om.copy(), f.copy())
    dE = np.abs(out["ao"][0] - out["charges"][0]).max()
    progress("C180H36 AO vs charges FP64: max |dE| = %.2e Ha" % dE)
    assert dE < 1e-10
    # f from C = bs Tb follows the expansion space, which rounding moves slightly
    np.testing.assert_allclose(out["ao"][1], out["charges"][1], rtol=1e-5, atol=1e-8)


def test_polyene_ensemble_256():
    from dftbaby_amd import engine
    from dftbaby_amd.ensemble import build_operators, solve_operators
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate.dftb2 import DFTB2
    members = g.polyene_ensemble(256)
    gss = []
    for j, (n, atoms) in enumerate(members):
        gs = DFTB2(atoms, long_range_correction=1, long_range_radius=3.03, device="cpu")
        gs.getEnergy(scf_conv=1e-8)
        gss.append(gs)
        if j % 64 == 63:
            progress("ensemble ground states: %d" % (j + 1))
    progress("ensemble: 256 ground states done")
    ops = build_operators(gss)
    k, conv = 8, 1e-6
    res = solve_operators(ops, k, conv=conv, maxiter=100, precision=1)
    progress("ensemble launch %.2f s, iterations %d..%d" % (res.t_solve, min(res.iterations), max(res.iterations)))
    # every member: FP64 residual certificate of the launch's X+-Y with the member's
    # own FP64 products, and (X-Y)^T (X+Y) = 1
    worst = 0.0
    for i, op in enumerate(ops):
        XpY, XmY = op.to_padded(res.XpY[i]), op.to_padded(res.XmY[i])
        w = torch.as_tensor(res.Omega[i], device=op.device)[:, None]
        P, _ = op.matvec(XpY, 0)
        _, M = op.matvec(XmY, 0)
        r = (torch.linalg.norm(P - w * XmY, dim=1) + torch.linalg.norm(M - w * XpY, dim=1)).cpu().numpy()
        worst = max(worst, float(r.max()))
        assert r.max() < 5 * conv, (i, members[i][0], r.max())
        G = np.einsum("kia,lia->kl", res.XmY[i], res.XpY[i])
        assert np.abs(G - np.eye(k)).max() < 1e-6, i
    progress("ensemble: max FP64 residual of the launch's roots %.2e" % worst)
    # lowest roots: 32 members (the largest, C40H42, first) against the single-molecule GPU FP64 Davidson
    big = [i for i, (n, _) in enumerate(members) if n == 40][0]
    sample = [big] + [i for i in range(0, 256, 8) if i != big][:31]
    worst = 0.0
    for i in sample:
        w, C, XmY, XpY, st = engine.nonhermitian_davidson_gpu(ops[i], k, conv=1e-8, maxiter=100, precision=0)
        dE = float(np.abs(res.Omega[i] - w.cpu().numpy()).max())
        worst = max(worst, dE)
        assert dE < 1e-6, (i, members[i][0], dE)
    progress("ensemble vs single-molecule FP64 on %d members: max |dE| = %.2e Ha" % (len(sample), worst))
    # three small members against the oracle Davidson (on C40H42 the oracle's
    # per-vector products over the whole basis every iteration take minutes; it is
    # covered by the single-molecule GPU FP64 Davidson above, which the oracle pins
    # on the smaller molecules, and by its residual certificate)
    small = [i for i, (n, _) in enumerate(members) if n <= 16][:3]
    for i in small:
        progress("oracle Davidson on member %d (C%dH%d)" % (i, members[i][0], members[i][0] + 2))
        ref = O.solve_from_ground_state(gss[i], nstates=k, diag_conv=1e-8, diag_maxiter=100)
        assert np.abs(res.Omega[i] - ref["Omega"]).max() < 1e-6, (i, members[i][0])
