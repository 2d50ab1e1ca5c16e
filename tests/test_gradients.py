"""Pin the gradient oracle (oracle/gradients_oracle.py) on the CPU.

* dS/dR, dH0/dR against DFTBaby's own Fortran (slako.f90 gradients_H0andS built
  into oracle/_ref) when that build exists, and against its golden vectors
  (tests/golden/slako_f90_golden.npz).
* The whole TD-DFTB gradient (ground state and excitation energies,
  gradient_lc, ExcGradients.py:286-711, and gradient_nolc, :713-981) against
  central finite differences of the SCC energy
  and of the excitation energies from the dense Casida solve.  Finite differences
  need a smooth energy, so this check runs on a ground state whose S and H0 come
  from the splines themselves (not splev_uniform's shifted interval), with the
  derivative tensors dS/dR, dH0/dR of that same S, H0 taken by finite
  differences: it pins F, F^lr, H^+-, Q, the Z-vector, W and the assembly; the
  Fortran comparison above pins the reference's own dS/dR, dH0/dR.
* A planted mistake in the oracle (one term of F^lr dropped) breaks the
  agreement with the finite differences.
"""
import numpy as np
import pytest

from conftest import ROOT  # noqa: F401
from oracle import gradients_oracle as G
from oracle import slako_oracle as K
from oracle import tddftb_oracle as O
from test_slako import GOLDEN, f90_gradients, f90_lib, gradient_geometries


def test_sk_gradients_against_reference_fortran():
    lib = f90_lib()
    tabs = K.load_tables()
    for name, Z, pos in gradient_geometries():
        gS, gH = G.sk_gradients(Z, pos, tabs)
        fS, fH = f90_gradients(lib, Z, pos, tabs[0])
        np.testing.assert_allclose(gS, fS, rtol=0, atol=1e-15, err_msg=name)
        np.testing.assert_allclose(gH, fH, rtol=0, atol=1e-15, err_msg=name)


def test_sk_gradients_against_golden_vectors():
    gd = np.load(GOLDEN)
    tabs = K.load_tables()
    for name in ("ethylene", "ethylene_pair"):
        gS, gH = G.sk_gradients(gd[name + "_Z"], gd[name + "_pos"], tabs)
        np.testing.assert_allclose(gS, gd[name + "_gradS"], rtol=0, atol=1e-15, err_msg=name)
        np.testing.assert_allclose(gH, gd[name + "_gradH0"], rtol=0, atol=1e-15, err_msg=name)


def _exact_ground_state_class():
    from dftbaby_amd.groundstate.dftb2 import DFTB2

    class ExactSplineDFTB2(DFTB2):
        def _constructH0andS(self):
            return G.h0_and_s_exact(self.Z, self.pos, distance_cutoff=self.distance_cutoff)
    return ExactSplineDFTB2


def _solve(cls, atoms, lc=1):
    gs = cls(atoms, long_range_correction=lc, long_range_radius=3.03, device="cpu")
    gs.getEnergy(scf_conv=1e-14, maxiter=500)
    ref = O.solve_from_ground_state(gs, nstates=None)
    return gs, gs.E_elec, ref


def _fd_sk_tensors(Z, pos, h=1e-5):
    nat = len(Z)
    S0, _ = G.h0_and_s_exact(Z, pos)
    n = len(S0)
    gS = np.zeros((3 * nat, n, n))
    gH = np.zeros((3 * nat, n, n))
    for k in range(3 * nat):
        A, c = divmod(k, 3)
        p1, p2 = pos.copy(), pos.copy()
        p1[A, c] += h
        p2[A, c] -= h
        S1, H1 = G.h0_and_s_exact(Z, p1)
        S2, H2 = G.h0_and_s_exact(Z, p2)
        gS[k] = (S1 - S2) / (2 * h)
        gH[k] = (H1 - H2) / (2 * h)
    return gS, gH


@pytest.fixture(scope="module", params=[1, 0], ids=["lc1", "lc0"])
def fd_case(request):
    from dftbaby_amd.groundstate import geometry as g
    lc = request.param
    cls = _exact_ground_state_class()
    mol = g.displace(g.ethylene(), 0.05, 7)
    Z = [z for z, _ in mol]
    pos = np.array([x for _, x in mol])
    gs, E, ref = _solve(cls, mol, lc)
    gS, gH = _fd_sk_tensors(Z, pos)
    h = 1e-4
    fd = {}
    for k in range(3 * len(mol)):
        A, c = divmod(k, 3)
        vals = []
        for s in (1.0, -1.0):
            p = pos.copy()
            p[A, c] += s * h
            _, E2, r2 = _solve(cls, [(z, list(x)) for z, x in zip(Z, p)], lc)
            vals.append((E2, r2["Omega"][:3]))
        fd[k] = ((vals[0][0] - vals[1][0]) / (2 * h), (vals[0][1] - vals[1][1]) / (2 * h))
    return gs, ref, gS, gH, fd


@pytest.mark.parametrize("state", [0, 1, 2])
def test_gradient_lc_against_finite_differences(fd_case, state):
    gs, ref, gS, gH, fd = fd_case
    gE0, gExc, _ = G.gradient_lc(gs, ref["XmY"][state], ref["XpY"][state], ref["Omega"][state], gradS=gS, gradH0=gH)
    fdE = np.array([fd[k][0] for k in sorted(fd)])
    fdW = np.array([fd[k][1][state] for k in sorted(fd)])
    assert np.abs(gE0 - fdE).max() < 2e-7, np.abs(gE0 - fdE).max()
    assert np.abs(gExc - fdW).max() < 2e-7, np.abs(gExc - fdW).max()


def test_planted_mistake_in_flr_breaks_finite_differences(fd_case, monkeypatch):
    gs, ref, gS, gH, fd = fd_case
    if gs.long_range_correction == 0:
        pytest.skip("F^lr does not enter without the long-range correction")
    full = G.Flr_op

    def without_term12(v, S, gradS, g0lr_AO, g1lr_AO):
        # F^lr minus its 12th term (S (dgamma o v)^T S / 4)
        dgv = g1lr_AO * v[None]
        t12 = np.swapaxes(np.tensordot(S, np.tensordot(dgv, S, axes=(1, 1)), axes=(1, 1)), 0, 1)
        return full(v, S, gradS, g0lr_AO, g1lr_AO) - 0.25 * t12
    monkeypatch.setattr(G, "Flr_op", without_term12)
    _, gExc, _ = G.gradient_lc(gs, ref["XmY"][0], ref["XpY"][0], ref["Omega"][0], gradS=gS, gradH0=gH)
    fdW = np.array([fd[k][1][0] for k in sorted(fd)])
    assert np.abs(gExc - fdW).max() > 1e-4


def test_repulsive_gradient_finite_differences_and_producer():
    """gradient_Vrep against central differences of the repulsive energy (a
    displaced C8H10: C-C, C-H and H-H pairs inside and beyond the table ends);
    the producer's E_rep against the oracle's."""
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate import slako as SK
    mol = g.displace(g.polyene(8), 0.08, 11)
    Z = [z for z, _ in mol]
    pos = np.array([x for _, x in mol])
    grad = G.gradient_vrep(Z, pos)
    h = 1e-5
    fd = np.zeros_like(grad)
    for k in range(len(grad)):
        A, c = divmod(k, 3)
        p1, p2 = pos.copy(), pos.copy()
        p1[A, c] += h
        p2[A, c] -= h
        fd[k] = (G.repulsive_energy(Z, p1) - G.repulsive_energy(Z, p2)) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=0, atol=1e-8)
    assert np.abs(grad).max() > 1e-2
    assert abs(SK.repulsive_energy(Z, pos) - G.repulsive_energy(Z, pos)) < 1e-12
