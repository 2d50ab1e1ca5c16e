"""The dynamics caller on the GPU (run with -m gpu): ScalarCoupling.ci_overlap and
PES.PotentialEnergySurfaces against the oracle and against the single-purpose
entry points they combine.

* dftb_slako_overlap12 against the reference Fortran's golden vectors.
* ci_overlap (GEMM form through the matrix determinant lemma) against the
  oracle's determinant loops (ScalarCoupling.py:141-265) on C8H10 at two
  geometries, two thresholds.
* PotentialEnergySurfaces: energies = E_tot + excitation energies, the gradient
  of getEnergiesAndGradient = the sum of Gradients' three parts, three steps of
  velocity-Verlet dynamics with getScalarCouplings (antisymmetric couplings,
  overlaps of each state with itself near 1 after the phase alignment), and
  getEnergiesBatch (one ensemble launch) against getEnergies per geometry.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, progress
from oracle import coupling_oracle as CO

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_overlap12_kernel_against_reference_fortran_golden():
    from dftbaby_amd.ScalarCoupling import overlap_AB
    gd = np.load(os.path.join(ROOT, "tests", "golden", "slako_f90_golden.npz"))
    Z = gd["c8h10_step_Z"]
    a1 = [(int(z), list(x)) for z, x in zip(Z, gd["c8h10_step_pos1"])]
    a2 = [(int(z), list(x)) for z, x in zip(Z, gd["c8h10_step_pos2"])]
    S = overlap_AB(a1, a2).cpu().numpy()
    np.testing.assert_allclose(S, gd["c8h10_step_S12"], rtol=0, atol=1e-15)


def _two_geometries():
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    from dftbaby_amd.groundstate import geometry as g
    mol = g.displace(g.polyene(8), 0.05, 21)
    Z = [z for z, _ in mol]
    p1 = np.array([x for _, x in mol])
    p2 = p1 + np.random.default_rng(8).normal(0.0, 0.01, p1.shape)
    tds = []
    for p in (p1, p2):
        td = LR_TDDFTB([(z, list(x)) for z, x in zip(Z, p)], long_range_correction=1)
        td.getEnergies(nstates=4, diag_conv=1e-8, diag_maxiter=100)
        tds.append(td)
    return Z, p1, p2, tds


@pytest.mark.parametrize("threshold", [1e-3, 0.01])
def test_ci_overlap_matches_oracle(threshold):
    from dftbaby_amd.ScalarCoupling import ScalarCoupling
    Z, p1, p2, (td1, td2) = _two_geometries()
    a1 = [(z, list(x)) for z, x in zip(Z, p1)]
    a2 = [(z, list(x)) for z, x in zip(Z, p2)]
    o1, o2 = td1.dftb2.getKSCoefficients(), td2.dftb2.getKSCoefficients()
    C1, C2 = td1.Cij, td2.Cij
    Nst = 5
    Sci = ScalarCoupling(td2).ci_overlap(a1, o1, C1, a2, o2, C2, Nst, threshold=threshold)
    occ, virt = td2.getActiveOrbitals()
    ref = CO.ci_overlap(Z, p1, o1, C1, p2, o2, C2, Nst, td2.dftb2.Nelec_val, occ, virt, threshold=threshold)
    progress("ci_overlap threshold %g: max |Sci - ref| = %.2e, diag %s" % (threshold, np.abs(Sci - ref).max(),
                                                                           np.round(np.diag(Sci), 4)))
    np.testing.assert_allclose(Sci, ref, rtol=0, atol=1e-10)
    assert np.abs(np.abs(np.diag(Sci)[1:]) - 1.0).max() < 0.05


def test_pes_energies_gradient_couplings_and_batch():
    from dftbaby_amd.ExcGradients import Gradients
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    from dftbaby_amd.PES import PotentialEnergySurfaces
    from dftbaby_amd.groundstate import geometry as g
    mol = g.displace(g.polyene(8), 0.05, 21)
    x0 = np.array([x for _, x in mol]).ravel()
    Nst = 3
    pes = PotentialEnergySurfaces(mol, Nst=Nst, long_range_correction=1, diag_conv=1e-9, diag_maxiter=100)
    E = pes.getEnergies(x0)
    td = LR_TDDFTB(mol, long_range_correction=1)
    w, _ = td.getEnergies(nstates=Nst - 1, diag_conv=1e-9, diag_maxiter=100)
    assert abs(E[0] - td.dftb2.E_tot) < 1e-9
    np.testing.assert_allclose(E[1:], td.dftb2.E_tot + w, atol=1e-8)
    E1, grad = pes.getEnergiesAndGradient(x0, gradient_state=1)
    np.testing.assert_allclose(E1, E, atol=1e-9)
    gV, gE0, gExc = Gradients(pes.tddftb).gradient(I=1)
    np.testing.assert_allclose(grad, gV + gE0 + gExc, rtol=0, atol=1e-10)
    # three velocity-Verlet steps on S1 (masses in electron masses, dt 10 a.u.)
    mass = np.repeat([1837.15 * (12.0 if z == 6 else 1.008) for z, _ in mol], 3)
    dt = 10.0
    x, v = x0.copy(), np.zeros_like(x0)
    f = -grad
    pes.saveLastCalculation()
    for step in range(3):
        x = x + dt * v + 0.5 * dt * dt * f / mass
        E2, g2 = pes.getEnergiesAndGradient(x, gradient_state=1)
        v = v + 0.5 * dt * (f - g2) / mass
        f = -g2
        coupl, olap = pes.getScalarCouplings(threshold=0.01)
        progress("step %d: E = %s, |coupl| max %.2e, olap diag %s" % (step, np.round(E2, 6), np.abs(coupl).max(),
                                                                     np.round(np.diag(olap), 4)))
        np.testing.assert_allclose(coupl, -coupl.T, atol=1e-14)
        assert np.all(np.diag(olap)[1:] > 0.9)
        assert np.all(np.diag(coupl) == 0.0)
    # many geometries in one ensemble launch
    xs = [x0, x0 + 0.02 * np.random.default_rng(1).standard_normal(x0.shape), x]
    Eb = pes.getEnergiesBatch(xs, precision=0, conv=1e-8, maxiter=100)
    for xi, Ei in zip(xs, Eb):
        np.testing.assert_allclose(Ei, pes.getEnergies(xi), atol=1e-7)
