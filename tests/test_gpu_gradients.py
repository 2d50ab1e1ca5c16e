"""GPU analytic LC-TD-DFTB gradients (dftbaby_amd.ExcGradients) against the oracle
(run with -m gpu).

* C8H10 and C12H14 with displaced atoms (lc = 1, and C8H10 with lc = 0,
  gradient_nolc): repulsive, ground-state and excitation-energy gradients of
  the lowest states against
  oracle/gradients_oracle (gradient_Vrep, gradient_lc:
  (the reference's tensor formulas over [3 Nat][Norb][Norb] derivative arrays,
  pinned on the CPU by the Fortran dS/dR, dH0/dR and by finite differences).
* A restricted active space: the amplitudes are padded to the full space as the
  reference does (ExcGradients.py:236-263).
* The C604H68 flake: the gradient of the rotated molecule is the rotated
  gradient (the oracle's tensors would need 100 GB there).
"""
import numpy as np
import pytest

from conftest import ground_state, progress
from oracle import gradients_oracle as G

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("name,seed,lc", [("C8H10", 23, 1), ("C12H14", 5, 1), ("C8H10", 23, 0)])
def test_gradients_match_oracle(name, seed, lc):
    from dftbaby_amd.ExcGradients import Gradients
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state(name, lc, 0.05, seed)
    td = LR_TDDFTB(ground_state=gs)
    td.getEnergies(nstates=4, diag_conv=1e-10, diag_maxiter=100)
    grads = Gradients(td)
    for I in (0, 1, 2):
        gV, gE0, gExc = grads.gradient(I=I)
        rV = G.gradient_vrep([z for z, _ in gs.getGeometry()], np.array([x for _, x in gs.getGeometry()]))
        np.testing.assert_allclose(gV, rV, rtol=0, atol=1e-12 * max(1.0, np.abs(rV).max()))
        if I == 0:
            rE0, _, _ = G.gradient_lc(gs, td.XmY[0], td.XpY[0], td.Omega[0])
            assert np.all(gExc == 0.0)
        else:
            rE0, rExc, rZ = G.gradient_lc(gs, td.XmY[I - 1], td.XpY[I - 1], td.Omega[I - 1])
            np.testing.assert_allclose(grads.Zia, rZ, rtol=0, atol=1e-10)
            np.testing.assert_allclose(gExc, rExc, rtol=0, atol=1e-9 * max(1.0, np.abs(rExc).max()))
        np.testing.assert_allclose(gE0, rE0, rtol=0, atol=1e-10 * max(1.0, np.abs(rE0).max()))
        progress("%s state %d: max |grad| %.3e" % (name, I, np.abs(gExc if I else gE0).max()))


def test_mixed_precision_zvector_gradient():
    """After a BF16x3 excited-state solve (refined to diag_conv 1e-10) the Z-vector
    equations are solved with BF16x3 products and FP64 refinement to the same
    residual test: Z and the excitation gradient equal the oracle's on the same
    X-Y, X+Y."""
    from dftbaby_amd.ExcGradients import Gradients
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C12H14", 1, 0.05, 5)
    td = LR_TDDFTB(ground_state=gs)
    td.getEnergies(nstates=3, diag_conv=1e-10, diag_maxiter=100, precision="bf16x3")
    assert td.solve_precision == 2
    grads = Gradients(td)
    gV, gE0, gExc = grads.gradient(I=1)
    rE0, rExc, rZ = G.gradient_lc(gs, td.XmY[0], td.XpY[0], td.Omega[0])
    np.testing.assert_allclose(grads.Zia, rZ, rtol=0, atol=1e-10)
    np.testing.assert_allclose(gExc, rExc, rtol=0, atol=1e-9 * max(1.0, np.abs(rExc).max()))


def test_gradient_with_restricted_active_space():
    """nr_active_occ / nr_active_virt: the gradient uses the amplitudes padded
    with zeros over the full space (the reference's own treatment)."""
    from dftbaby_amd.ExcGradients import Gradients
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    gs = ground_state("C8H10", 1, 0.05, 23)
    td = LR_TDDFTB(ground_state=gs)
    td.getEnergies(nstates=3, nr_active_occ=8, nr_active_virt=9, diag_conv=1e-10, diag_maxiter=100)
    _, _, gExc = Gradients(td).gradient(I=1)
    f = gs.getOccupation()
    occ, virt = np.where(f > 0.1)[0], np.where(f <= 0.1)[0]
    XmY = np.zeros((len(occ), len(virt)))
    XpY = np.zeros((len(occ), len(virt)))
    io, iv = np.searchsorted(occ, td.active_occupied_orbs), np.searchsorted(virt, td.active_virtual_orbs)
    XmY[np.ix_(io, iv)] = td.XmY[0]
    XpY[np.ix_(io, iv)] = td.XpY[0]
    _, rExc, _ = G.gradient_lc(gs, XmY, XpY, td.Omega[0])
    np.testing.assert_allclose(gExc, rExc, rtol=0, atol=1e-9 * max(1.0, np.abs(rExc).max()))


def test_flake_gradient_rotation_covariance():
    """C604H68: the excitation-energy gradient of S1 and the ground-state gradient
    of the rotated flake equal the rotated gradients (rotation-invariant energies;
    the reference's spline quirks depend on distances and atom order only)."""
    from dftbaby_amd.ExcGradients import Gradients
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate.dftb2 import DFTB2
    fl = g.flake_c604h68()
    Z = [a for a, _ in fl]
    pos = np.array([x for _, x in fl])
    th = 0.6
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    R = R @ np.array([[1.0, 0.0, 0.0], [0.0, np.cos(0.5), -np.sin(0.5)], [0.0, np.sin(0.5), np.cos(0.5)]])
    out = []
    for P in (pos, pos @ R.T):
        gs = DFTB2([(z, list(x)) for z, x in zip(Z, P)], long_range_correction=1, long_range_radius=3.03)
        gs.getEnergy(scf_conv=1e-14, maxiter=500)
        td = LR_TDDFTB(ground_state=gs)
        td.getEnergies(nstates=4, diag_conv=1e-8, diag_maxiter=100, precision="fp64")
        _, gE0, gExc = Gradients(td).gradient(I=1)
        progress("flake gradient: |gE0| %.3e |gExc| %.3e" % (np.abs(gE0).max(), np.abs(gExc).max()))
        out.append((gE0.reshape(-1, 3), gExc.reshape(-1, 3)))
    (e0, x0), (e1, x1) = out
    # to the SCC convergence (commutator 1e-9) and the Davidson residual (1e-8)
    np.testing.assert_allclose(e1, e0 @ R.T, rtol=0, atol=1e-6 * np.abs(e0).max())
    np.testing.assert_allclose(x1, x0 @ R.T, rtol=0, atol=1e-5 * np.abs(x0).max())
