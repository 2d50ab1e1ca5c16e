"""CPU restatement of DFTBaby's CI-wavefunction overlaps (test infrastructure).

Only tests may import this module; the product never does.  It follows
``ScalarCoupling.ci_overlap`` (``DFTB/ScalarCoupling.py:141-265``) loop by loop:
the overlap of the TD-DFTB "wavefunctions" sum_ia C^I_ia |Psi_ia> (singlet
configuration state functions, closed shell) at two geometries, from the MO
overlap Smo = orbs1^T S_AO(1,2) orbs2 and determinants of its occupied blocks
with one row and/or column replaced (eq. 9.39 of A. Humeniuk's thesis), skipping
excitations whose largest coefficient is below the threshold.  Also
``PES.getScalarCouplings``' phase alignment (``DFTB/PES.py:175-253``).
"""
import numpy as np

from . import slako_oracle as K


def ci_overlap(Z, pos1, orbs1, C1, pos2, orbs2, C2, Nst, nelec, occ, virt, threshold=1.0e-3):
    """Sci[I][J] = <Psi_I(1)|Psi_J(2)>, I, J = 0 (ground state) .. Nst-1."""
    Sao = K.overlap12(Z, pos1, pos2)
    Smo = orbs1.T @ (Sao @ orbs2)
    Sci = np.zeros((Nst, Nst))
    occs_0 = np.arange(nelec // 2)
    nocc, nvirt = len(occ), len(virt)
    det_ij = np.linalg.det(Smo[occs_0, :][:, occs_0])
    for i in range(nocc):
        for a in range(nvirt):
            if abs(C1[:, i, a]).max() < threshold:
                continue
            occs_ia = occs_0.copy()
            occs_ia[occ[i]] = virt[a]
            det_aj = np.linalg.det(Smo[occs_ia, :][:, occs_0])
            for j in range(nocc):
                for b in range(nvirt):
                    if abs(C2[:, j, b]).max() < threshold:
                        continue
                    occs_jb = occs_0.copy()
                    occs_jb[occ[j]] = virt[b]
                    det_ab = np.linalg.det(Smo[occs_ia, :][:, occs_jb])
                    det_ib = np.linalg.det(Smo[occs_0, :][:, occs_jb])
                    for I in range(1, Nst):
                        for J in range(1, Nst):
                            Sci[I, J] += C1[I - 1, i, a] * C2[J - 1, j, b] * (det_ab * det_ij + det_aj * det_ib)
    for i in range(nocc):
        for a in range(nvirt):
            if abs(C2[:, i, a]).max() < threshold:
                continue
            occs_ia = occs_0.copy()
            occs_ia[occ[i]] = virt[a]
            det_ia = np.linalg.det(Smo[occs_0, :][:, occs_ia])
            for J in range(1, Nst):
                Sci[0, J] += C2[J - 1, i, a] * np.sqrt(2) * (det_ia * det_ij)
    for i in range(nocc):
        for a in range(nvirt):
            if abs(C1[:, i, a]).max() < threshold:
                continue
            occs_ia = occs_0.copy()
            occs_ia[occ[i]] = virt[a]
            det_ia = np.linalg.det(Smo[occs_ia, :][:, occs_0])
            for I in range(1, Nst):
                Sci[I, 0] += C1[I - 1, i, a] * np.sqrt(2) * (det_ia * det_ij)
    Sci[0, 0] = det_ij
    return Sci


def spin_determinant_overlaps(Smo, nelec, occ, virt):
    """<Psi_ia|Psi'_jb>, <Psi_0|Psi'_jb>, <Psi_ia|Psi'_0>, <Psi_0|Psi'_0> from
    explicit spin-orbital Slater determinants: |Psi_ia> = (|i->a alpha> + |i->a
    beta>) / sqrt(2), each determinant's overlap the product of its alpha and
    beta determinants of Smo (an independent construction of the CSF overlaps
    that ci_overlap uses)."""
    n = nelec // 2
    occs_0 = np.arange(n)
    D = lambda rows, cols: np.linalg.det(Smo[np.ix_(rows, cols)])

    def csf(i, a):
        o = occs_0.copy()
        o[occ[i]] = virt[a]
        return [(o, occs_0, 1 / np.sqrt(2)), (occs_0, o, 1 / np.sqrt(2))]      # (alpha, beta, coefficient)
    no, nv = len(occ), len(virt)
    ex = np.zeros((no, nv, no, nv))
    g_ex = np.zeros((no, nv))
    ex_g = np.zeros((no, nv))
    for i in range(no):
        for a in range(nv):
            for j in range(no):
                for b in range(nv):
                    ex[i, a, j, b] = sum(c1 * c2 * D(a1, a2) * D(b1, b2)
                                         for a1, b1, c1 in csf(i, a) for a2, b2, c2 in csf(j, b))
            g_ex[i, a] = sum(c2 * D(occs_0, a2) * D(occs_0, b2) for a2, b2, c2 in csf(i, a))
            ex_g[i, a] = sum(c1 * D(a1, occs_0) * D(b1, occs_0) for a1, b1, c1 in csf(i, a))
    return ex, g_ex, ex_g, D(occs_0, occs_0) ** 2
