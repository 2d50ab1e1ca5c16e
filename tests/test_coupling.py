"""Pin the CI-overlap oracle (oracle/coupling_oracle.py) on the CPU.

* The AO overlap between two geometries (slako.f90 overlap12) against the
  reference Fortran built into oracle/_ref, and against its golden vectors.
* ci_overlap (ScalarCoupling.py:141-265) against an independent construction
  from explicit spin-orbital Slater determinants (each singlet CSF the sum of its
  alpha and beta excitations, determinant overlaps = det(alpha) det(beta)), on
  ethylene at two geometries with no threshold; the reference's one departure
  from that construction, Sci[0][0] = det(alpha block) instead of its square,
  is kept.
* With Smo = 1 (one geometry's S against itself), Sci is the identity
  (orthonormal CI vectors); a planted mistake (the geometries' roles swapped in the AO overlap)
  breaks the brute-force match.
"""
import numpy as np
import pytest

from conftest import ground_state  # noqa: F401
from oracle import coupling_oracle as CO
from oracle import slako_oracle as K
from oracle import tddftb_oracle as O
from test_slako import GOLDEN, f90_lib, f90_overlap12, overlap12_geometries


def test_overlap12_against_reference_fortran_and_golden():
    tabs = K.load_tables()
    gd = np.load(GOLDEN)
    for name, Z, p1, p2 in overlap12_geometries():
        S = K.overlap12(Z, p1, p2, tabs)
        np.testing.assert_allclose(S, gd[name + "_S12"], rtol=0, atol=1e-15)
    lib = f90_lib()
    for name, Z, p1, p2 in overlap12_geometries():
        np.testing.assert_allclose(K.overlap12(Z, p1, p2, tabs), f90_overlap12(lib, Z, p1, p2, tabs[0]), rtol=0,
                                   atol=1e-15)


def _ethylene_pair():
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate.dftb2 import DFTB2
    mol = g.displace(g.ethylene(), 0.05, 7)
    Z = [z for z, _ in mol]
    p1 = np.array([x for _, x in mol])
    p2 = p1 + np.random.default_rng(3).normal(0.0, 0.01, p1.shape)     # |step| < 2 grid steps of the tables
    out = []
    for p in (p1, p2):
        gs = DFTB2([(z, list(x)) for z, x in zip(Z, p)], long_range_correction=1, device="cpu")
        gs.getEnergy(scf_conv=1e-12)
        ref = O.solve_from_ground_state(gs, nstates=None)
        out.append((gs, p, ref))
    return Z, out


def test_ci_overlap_against_spin_determinants():
    Z, ((gs1, p1, r1), (gs2, p2, r2)) = _ethylene_pair()
    nelec = gs1.Nelec_val
    occ, virt = r1["act_occ"], r1["act_virt"]
    Nst = 4
    C1, C2 = r1["C"][:Nst - 1], r2["C"][:Nst - 1]
    o1, o2 = gs1.getKSCoefficients(), gs2.getKSCoefficients()
    Sci = CO.ci_overlap(Z, p1, o1, C1, p2, o2, C2, Nst, nelec, occ, virt, threshold=0.0)
    Smo = o1.T @ K.overlap12(Z, p1, p2) @ o2
    ex, g_ex, ex_g, d00 = CO.spin_determinant_overlaps(Smo, nelec, occ, virt)
    want = np.einsum("Iia,iajb,Jjb->IJ", C1, ex, C2)
    np.testing.assert_allclose(Sci[1:, 1:], want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Sci[0, 1:], np.einsum("Jjb,jb->J", C2, g_ex), rtol=0, atol=1e-12)
    np.testing.assert_allclose(Sci[1:, 0], np.einsum("Iia,ia->I", C1, ex_g), rtol=0, atol=1e-12)
    assert abs(Sci[0, 0] ** 2 - d00) < 1e-12
    assert np.abs(Sci[1:, 1:]).max() > 0.5


def test_planted_mistake_swapped_geometries(monkeypatch):
    """The oracle with the two geometries' roles swapped in the AO overlap (a
    plausible slip) no longer matches the spin-determinant construction."""
    Z, ((gs1, p1, r1), (gs2, p2, r2)) = _ethylene_pair()
    nelec, occ, virt = gs1.Nelec_val, r1["act_occ"], r1["act_virt"]
    C1, C2 = r1["C"][:3], r2["C"][:3]
    o1, o2 = gs1.getKSCoefficients(), gs2.getKSCoefficients()
    Smo = o1.T @ K.overlap12(Z, p1, p2) @ o2
    ex, _, _, _ = CO.spin_determinant_overlaps(Smo, nelec, occ, virt)
    want = np.einsum("Iia,iajb,Jjb->IJ", C1, ex, C2)
    full = K.overlap12
    monkeypatch.setattr(CO.K, "overlap12", lambda Z_, a, b: full(Z_, b, a))
    Sci = CO.ci_overlap(Z, p1, o1, C1, p2, o2, C2, 4, nelec, occ, virt, threshold=0.0)
    assert np.abs(Sci[1:, 1:] - want).max() > 1e-4


def test_ci_overlap_identity_at_equal_geometry(monkeypatch):
    """With Smo = C^T S C = 1 (the AO overlap of one geometry with itself) the CI
    overlap is the identity.  overlap12 itself does not give S at equal
    geometries: it evaluates the same-atom blocks from the tables at r = 0 and
    does not mirror the off-site ones (Sci[0][0] = 1.0005 for ethylene), so the
    AO overlap is replaced by S here to test the determinant algebra alone."""
    Z, ((gs1, p1, r1), _) = _ethylene_pair()
    monkeypatch.setattr(CO.K, "overlap12", lambda *a: gs1.getOverlapMatrix())
    o1 = gs1.getKSCoefficients()
    Nst = 4
    C = r1["C"][:Nst - 1]
    Sci = CO.ci_overlap(Z, p1, o1, C, p1, o1, C, Nst, gs1.Nelec_val, r1["act_occ"], r1["act_virt"], threshold=0.0)
    np.testing.assert_allclose(Sci, np.eye(Nst), rtol=0, atol=1e-10)


@pytest.mark.parametrize("threshold", [1e-3, 0.05])
def test_ci_overlap_threshold_masks_excitations(threshold):
    """With a threshold the skipped excitations are exactly those whose largest
    coefficient is below it: the result equals the unthresholded overlap of the
    masked coefficient arrays."""
    Z, ((gs1, p1, r1), (gs2, p2, r2)) = _ethylene_pair()
    Nst = 3
    C1, C2 = r1["C"][:Nst - 1], r2["C"][:Nst - 1]
    o1, o2 = gs1.getKSCoefficients(), gs2.getKSCoefficients()
    args = (gs1.Nelec_val, r1["act_occ"], r1["act_virt"])
    Sci = CO.ci_overlap(Z, p1, o1, C1, p2, o2, C2, Nst, *args, threshold=threshold)
    m1 = np.abs(C1).max(0) >= threshold
    m2 = np.abs(C2).max(0) >= threshold
    Sm = CO.ci_overlap(Z, p1, o1, C1 * m1, p2, o2, C2 * m2, Nst, *args, threshold=0.0)
    np.testing.assert_allclose(Sci, Sm, rtol=0, atol=1e-13)
