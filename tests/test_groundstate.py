"""The host ground-state producer (dftbaby_amd/groundstate) and the gamma matrices it
hands to the excited-state engine, against closed forms (no GPU needed).

gamma, gamma_lr: GammaApproximation.py:122-151, 210-243 (Gaussian fluctuation
functions, sigma = 1 / (sqrt(pi) U)); Hubbard U from (IE - EA) / 27.211396132
(Parameters.py:25-30, AtomicData.py:131-134).
"""
import numpy as np

from conftest import ground_state


def test_hubbard_parameters():
    gs = ground_state("C8H10", 1)
    assert abs(gs.hubbard_U[6] - 0.3674262) < 1e-6
    assert abs(gs.hubbard_U[1] - 0.4720158) < 1e-6


def test_gamma_closed_forms():
    """gamma_AA = U_A; gamma^lr_AA = [pi (sigma_A^2 + R_lr^2 / 4)]^{-1/2}; both
    symmetric; both tend to 1/R between atoms 30 bohr apart (erf -> 1)."""
    gs = ground_state("C8H10-H2", 1)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    Z = np.array([z for z, _ in gs.getGeometry()])
    pos = np.array([p for _, p in gs.getGeometry()])
    U = np.array([gs.hubbard_U[z] for z in Z])
    sig = 1.0 / (np.sqrt(np.pi) * U)
    Rlr = gs.long_range_radius
    assert abs(Rlr - 3.03) < 1e-12
    np.testing.assert_allclose(np.diag(g), U, rtol=1e-14)
    np.testing.assert_allclose(np.diag(glr), 1.0 / np.sqrt(np.pi * (sig ** 2 + 0.25 * Rlr ** 2)), rtol=1e-14)
    assert np.abs(g - g.T).max() == 0.0 and np.abs(glr - glr.T).max() == 0.0
    far = np.linalg.norm(pos[:, None] - pos[None, :], axis=-1)
    h2 = slice(len(Z) - 2, len(Z))
    R = far[:-2, h2]
    np.testing.assert_allclose(g[:-2, h2], 1.0 / R, rtol=1e-12)
    np.testing.assert_allclose(glr[:-2, h2], 1.0 / R, rtol=1e-12)
    # the long-range kernel is the screened one at short range
    near = far[:-2, :-2] > 0
    assert np.all(glr[:-2, :-2][near] < g[:-2, :-2][near])


def test_scc_orthonormal_orbitals_and_neutral_charges():
    """C^T S C = 1 for the SCC orbitals, the Mulliken charges of the neutral
    molecule sum to zero, and the closed shell has 2 electrons per occupied MO."""
    gs = ground_state("C8H10", 1, 0.05, 13)
    C, S = gs.getKSCoefficients(), gs.getOverlapMatrix()
    np.testing.assert_allclose(C.T @ S @ C, np.eye(C.shape[1]), atol=1e-10)
    assert abs(np.sum(gs.dq)) < 1e-10
    f = gs.getOccupation()
    assert set(np.round(f, 12)) <= {0.0, 2.0}
    assert np.sum(f) == sum(4 if z == 6 else 1 for z, _ in gs.getGeometry())
