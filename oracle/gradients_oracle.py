"""CPU restatement of DFTBaby's analytic LC-TD-DFTB gradients (test infrastructure).

Only tests may import this module; the product never does.  It follows, term by
term and in the same tensor layouts (derivative index k = 3 A + c first):

* ``gradients_H0andS`` (``DFTB/extensions/slako.f90:402-569``): dS/dR and dH0/dR
  [3 Nat][Norb][Norb].  Every ordered pair (mu, nu) of orbitals on different atoms
  i = atom(mu), j = atom(nu) within the proximity cutoff gives the derivative with
  respect to R_j of the element, written to (mu, nu) and (nu, mu); the element is
  formed with the orbital on the lighter type first (else swapped, direction
  reversed), from the spline values of ``splev_uniform`` and the derivatives of
  FITPACK's ``splder`` (exact knot interval, argument clamped to the table).  The
  rule gradients (``slako_transformations_grad_lmax2.f90``) are those of the
  tensor rules of ``slako_oracle``: grad_{R_2} V = sum_k dV/dT_k T_k'(r) e +
  (1 - e e^T) grad_e V / r.
* ``GammaMatrix.gamma_AOwise`` (``DFTB/GammaApproximation.py:388-450``) with the
  Gaussian and Gaussian-LC derivatives (``:141-147, 234-238``).
* ``Gradients.gradient_lc`` (``DFTB/ExcGradients.py:286-711``): the operators F
  and F^lr (``:348-424``), the ground-state gradient (``:477-496``), U, V, T, the
  H^+ / H^- contractions, Q, the Z-vector equation (Krylov solver of
  ``Solver.py:921-1012``), W and the assembly of the excitation-energy gradient
  (``:590-711``), all over the full orbital space (lc = 1, the default "new" lc
  implementation: F^lr of D - D0); ``gradient_nolc`` (``:713-981``) is the same
  with gamma_lr = 0, which removes F^lr, the exchange terms of H^+ and H^-.
* the repulsive energy and its gradient (``DFTB2.py:2340-2359``,
  ``ExcGradients.py:181-210``) with the splines of ``RepulsivePotential.py:80-101``
  (``splrep(d, Vrep, s=0)``, zero at and beyond the last table point).
"""

import numpy as np
from scipy.interpolate import splev, splrep
from scipy.special import erf

from . import slako_oracle as K
from . import tddftb_oracle as O


# ----------------------------------------------------------------------------- SK
def _rule_grad(l1, m1, l2, m2, e, r, T, Td):
    """(value, gradient w.r.t. the second atom) of the S/H0 rule of slako_oracle."""
    P = K._PIDX
    dVdT = np.zeros(7)
    ge = np.zeros(3)
    if l1 == 0 and l2 == 0:
        dVdT[0] = 1.0
    elif l1 == 0:
        j = P[m2]
        dVdT[2] = e[j]
        ge[j] = T[2]
    elif l2 == 0:
        i = P[m1]
        dVdT[4] = e[i]
        ge[i] = T[4]
    else:
        i, j = P[m1], P[m2]
        if i == j:
            other = [k for k in range(3) if k != i]
            dVdT[5] = e[other[0]] ** 2 + e[other[1]] ** 2
            dVdT[6] = e[i] ** 2
            for k in other:
                ge[k] = 2.0 * e[k] * T[5]
            ge[i] = 2.0 * e[i] * T[6]
        else:
            dVdT[6] = e[i] * e[j]
            dVdT[5] = -e[i] * e[j]
            ge[i] += e[j] * (T[6] - T[5])
            ge[j] += e[i] * (T[6] - T[5])
    val = K._sk(l1, m1, l2, m2, e, T)
    grad = float(np.dot(dVdT, Td[:7])) * e + (ge - np.dot(e, ge) * e) / r
    return val, grad


def sk_gradients(Z, pos, tables=None, distance_cutoff=30.0, exact=False):
    """gradS, gradH0 [3 Nat][Norb][Norb] as slako.f90:402-569 forms them; with
    exact=True the table values come from the spline itself (correct interval),
    for checks by finite differences."""
    tabs, _ = tables or K.load_tables()
    Z = [int(z) for z in Z]
    pos = np.asarray(pos, dtype=np.float64)
    orbs = K._orbitals(Z)
    n, nat = len(orbs), len(Z)
    gS = np.zeros((3 * nat, n, n))
    gH = np.zeros((3 * nat, n, n))
    val = K.splev_exact if exact else K.splev_uniform
    cache = {}

    def tables_at(za, zb, r):
        key = (min(za, zb), max(za, zb), r)
        if key not in cache:
            tt = tabs[(min(za, zb), max(za, zb))]
            out = []
            for kind in ("S", "H"):
                v, d = np.zeros(7), np.zeros(7)
                for idx, tck in tt[kind].items():
                    v[idx] = val(tck, r)
                    d[idx] = K.splder1(tck, r)
                out += [v, d]
            cache[key] = out
        return cache[key]

    for mu in range(n):
        for nu in range(n):
            if mu == nu:
                continue
            (i, zi, l1, m1), (j, zj, l2, m2) = orbs[mu], orbs[nu]
            if i == j:
                continue
            r, e_ij = K._direction(pos, i, j)
            if not r < distance_cutoff:
                continue
            Sv, Sd, Hv, Hd = tables_at(zi, zj, r)
            if zi <= zj:
                _, gs = _rule_grad(l1, m1, l2, m2, e_ij, r, Sv, Sd)
                _, gh = _rule_grad(l1, m1, l2, m2, e_ij, r, Hv, Hd)
            else:
                _, gs = _rule_grad(l2, m2, l1, m1, -e_ij, r, Sv, Sd)
                _, gh = _rule_grad(l2, m2, l1, m1, -e_ij, r, Hv, Hd)
                gs, gh = -gs, -gh
            gS[3 * j:3 * j + 3, mu, nu] += gs
            gS[3 * j:3 * j + 3, nu, mu] += gs
            gH[3 * j:3 * j + 3, mu, nu] += gh
            gH[3 * j:3 * j + 3, nu, mu] += gh
    return gS, gH


def h0_and_s_exact(Z, pos, tables=None, distance_cutoff=30.0):
    """S, H0 with the spline itself in place of splev_uniform (for finite differences)."""
    saved = K.splev_uniform
    try:
        K.splev_uniform = K.splev_exact
        return K.h0_and_s(Z, pos, tables, distance_cutoff)
    finally:
        K.splev_uniform = saved


# ----------------------------------------------------------------------------- Vrep
def _rep_tables():
    d = np.load(K._DATA)
    out = {}
    for (z1, z2), name in K._PAIRS.items():
        x = d["%s_rep_d" % name]
        out[(z1, z2)] = (splrep(x, d["%s_rep_V" % name], s=0), x[-1])
    return out


def _vrep(tabs, z1, z2, R, deriv=0):
    tck, dmax = tabs[(min(z1, z2), max(z1, z2))]
    return 0.0 if R >= dmax else float(splev(R, tck, der=deriv))


def repulsive_energy(Z, pos):
    """DFTB2.getRepulsiveEnergy (DFTB2.py:2340-2359)."""
    tabs = _rep_tables()
    Z = [int(z) for z in Z]
    pos = np.asarray(pos, dtype=np.float64)
    E = 0.0
    for i in range(1, len(Z)):
        for j in range(i):
            E += _vrep(tabs, Z[i], Z[j], np.linalg.norm(pos[i] - pos[j]))
    return E


def gradient_vrep(Z, pos):
    """gradient_Vrep (ExcGradients.py:181-210)."""
    tabs = _rep_tables()
    Z = [int(z) for z in Z]
    pos = np.asarray(pos, dtype=np.float64)
    grad = np.zeros(3 * len(Z))
    for i in range(len(Z)):
        for j in range(len(Z)):
            if i == j:
                continue
            R = pos[i] - pos[j]
            Rij = np.linalg.norm(R)
            grad[3 * i:3 * i + 3] += _vrep(tabs, Z[i], Z[j], Rij, deriv=1) * R / Rij
    return grad


# ----------------------------------------------------------------------------- gamma
def _gauss_deriv(c, R):
    return 2.0 * c / np.sqrt(np.pi) * np.exp(-(c * R) ** 2) / R - erf(c * R) / R ** 2


def gamma_aowise(Z, pos, U, orbs_per_atom, Rlr=None):
    """g0, g1_AT, g0_AO, g1_AO of GammaApproximation.py:388-450 (Gaussian gamma;
    with Rlr the long-range Gaussian-LC gamma)."""
    Z = [int(z) for z in Z]
    pos = np.asarray(pos, dtype=np.float64)
    nat = len(Z)
    sig = np.array([1.0 / (np.sqrt(np.pi) * U[z]) for z in Z])
    g0 = np.zeros((nat, nat))
    g1 = np.zeros((nat, nat))
    for i in range(nat):
        for j in range(nat):
            if i == j:
                if Rlr is None:
                    g0[i, j] = 1.0 / (np.sqrt(np.pi) * sig[i])
                else:
                    g0[i, j] = 1.0 / np.sqrt(np.pi * (sig[i] ** 2 + 0.25 * Rlr ** 2))
            elif i < j:
                R = np.linalg.norm(pos[i] - pos[j])
                s2 = sig[i] ** 2 + sig[j] ** 2 + (0.0 if Rlr is None else 0.5 * Rlr ** 2)
                c = 1.0 / np.sqrt(2.0 * s2)
                g0[i, j] = erf(c * R) / R
                g1[i, j] = _gauss_deriv(c, R)
            else:
                g0[i, j], g1[i, j] = g0[j, i], g1[j, i]
    amap = np.repeat(np.arange(nat), orbs_per_atom)
    n = len(amap)
    g0_AO = g0[amap][:, amap]
    g1_AO = np.zeros((3 * nat, n, n))
    g1_AT = np.zeros((3 * nat, nat, nat))
    for i in range(nat):
        for j in range(nat):
            if i == j:
                continue
            d = pos[i] - pos[j]
            e = d / np.linalg.norm(d)
            g1_AT[3 * i:3 * i + 3, i, j] = g1[i, j] * e
            mi, mj = np.where(amap == i)[0], np.where(amap == j)[0]
            blk = g1[i, j] * e
            for c in range(3):
                g1_AO[3 * i + c][np.ix_(mi, mj)] += blk[c]
                g1_AO[3 * i + c][np.ix_(mj, mi)] += blk[c]
    return g0, g1_AT, g0_AO, g1_AO


# ----------------------------------------------------------------------------- F, Flr
def F_op(v, S, gradS, g0_AO, g1_AO):
    """F(v) of ExcGradients.py:348-366."""
    nat3, n, _ = gradS.shape
    vp = v + v.T
    Sv = np.sum(S * vp, axis=1)
    gSv = g0_AO @ Sv
    gdSv = np.zeros((nat3, n))
    dgSv = np.zeros((nat3, n))
    for k in range(nat3):
        gdSv[k] = g0_AO @ np.sum(gradS[k] * vp, axis=1)
        dgSv[k] = g1_AO[k] @ Sv
    f = (gradS * (gSv[None, :, None] + gSv[None, None, :])
         + S[None] * (dgSv[:, :, None] + dgSv[:, None, :])
         + S[None] * (gdSv[:, :, None] + gdSv[:, None, :]))
    return 0.25 * f


def Flr_op(v, S, gradS, g0lr_AO, g1lr_AO):
    """F^lr(v) of ExcGradients.py:367-424 (its twelve terms, in order)."""
    nat3 = gradS.shape[0]
    Sv = S @ v
    SvT = S @ v.T
    gv = g0lr_AO * v
    dSv24 = np.tensordot(gradS, v, axes=(2, 1))
    dSv23 = np.tensordot(gradS, v, axes=(2, 0))
    dgv = g1lr_AO * v[None]
    flr = np.zeros_like(gradS)
    tmp1 = np.tensordot(gradS, Sv, axes=(2, 1))
    flr += g0lr_AO[None] * tmp1
    tmp2 = dSv24 * g0lr_AO[None]
    flr += np.tensordot(tmp2, S, axes=(2, 1))
    flr += np.tensordot(gradS, Sv * g0lr_AO, axes=(2, 1))
    flr += np.tensordot(gradS, S @ gv, axes=(2, 1))
    tmp5 = np.swapaxes(np.tensordot(S, dSv23, axes=(1, 2)), 0, 1)
    flr += g0lr_AO[None] * tmp5
    flr += np.swapaxes(np.tensordot(SvT * g0lr_AO, gradS, axes=(1, 2)), 0, 1)
    tmp7 = dSv23 * g0lr_AO[None]
    flr += np.swapaxes(np.tensordot(S, tmp7, axes=(1, 2)), 0, 1)
    tmp8 = np.tensordot(gradS, gv, axes=(2, 0))
    flr += np.swapaxes(np.tensordot(S, tmp8, axes=(1, 2)), 0, 1)
    tmp9 = S @ Sv.T
    flr += g1lr_AO * tmp9[None]
    tmp10 = SvT[None] * g1lr_AO
    flr += np.tensordot(tmp10, S, axes=(2, 1))
    tmp11 = Sv[None] * g1lr_AO
    flr += np.swapaxes(np.tensordot(S, tmp11, axes=(1, 2)), 0, 1)
    tmp12 = np.tensordot(dgv, S, axes=(1, 1))
    flr += np.swapaxes(np.tensordot(S, tmp12, axes=(1, 1)), 0, 1)
    return 0.25 * flr


# ----------------------------------------------------------------------------- H+-
def Hplus(g0, g0lr, q_pq, q_rs, q_pr, q_qs, q_ps, q_qr, v_rs):
    """ExcGradients.py:567-581."""
    tmp = np.tensordot(q_rs, v_rs, axes=([1, 2], [0, 1]))
    h = 4 * np.tensordot(q_pq, g0 @ tmp, axes=(0, 0))
    tmp = np.tensordot(q_qs, v_rs, axes=(2, 1))
    h -= np.tensordot(q_pr, np.tensordot(g0lr, tmp, axes=(1, 0)), axes=([0, 2], [0, 2]))
    tmp = np.tensordot(q_qr, v_rs, axes=(2, 0))
    h -= np.tensordot(q_ps, np.tensordot(g0lr, tmp, axes=(1, 0)), axes=([0, 2], [0, 2]))
    return h


def Hminus(g0lr, q_ps, q_qr, q_pr, q_qs, v_rs):
    """ExcGradients.py:582-590."""
    tmp = np.tensordot(q_qr, v_rs, axes=(2, 0))
    h = np.tensordot(q_ps, np.tensordot(g0lr, tmp, axes=(1, 0)), axes=([0, 2], [0, 2]))
    tmp = np.tensordot(q_qs, v_rs, axes=(2, 1))
    h -= np.tensordot(q_pr, np.tensordot(g0lr, tmp, axes=(1, 0)), axes=([0, 2], [0, 2]))
    return h


# ----------------------------------------------------------------------------- gradient
def gradient_lc(gs, XmY, XpY, Omega, gradS=None, gradH0=None, exact=False, zconv=1e-14):
    """(gradE0, gradExc, Zia) of ExcGradients.py:286-711 for one excited state with
    amplitudes XmY, XpY (nocc x nvirt over the full orbital space) and energy
    Omega, on a ground-state object with the DFTB2 getters; for lc = 0 the same
    formulas with gamma_lr = 0 are gradient_nolc (ExcGradients.py:713-981)."""
    Z = np.array([z for z, _ in gs.getGeometry()])
    pos = np.array([x for _, x in gs.getGeometry()])
    U = gs.hubbard_U
    opa = gs.orbsPerAtom
    g0, _, g0_AO, g1_AO = gamma_aowise(Z, pos, U, opa)
    g0lr, _, g0lr_AO, g1lr_AO = gamma_aowise(Z, pos, U, opa, Rlr=gs.long_range_radius)
    if gs.long_range_correction == 0:
        # gradient_nolc (ExcGradients.py:713-981) is gradient_lc with gamma_lr = 0:
        # no F^lr, H^+ reduced to its Coulomb term, no H^- terms, lc = 0 (A+B)
        g0lr, g0lr_AO, g1lr_AO = 0.0 * g0lr, 0.0 * g0lr_AO, 0.0 * g1lr_AO
    f = gs.getOccupation()
    orbe = gs.getKSEnergies()
    orbs = gs.getKSCoefficients()
    occ = np.where(f > 0.1)[0]
    virt = np.where(f <= 0.1)[0]
    orbs_occ, orbs_virt = orbs[:, occ], orbs[:, virt]
    orbe_occ, orbe_virt = orbe[occ], orbe[virt]
    ei, ea = np.diag(orbe_occ), np.diag(orbe_virt)
    nocc, nvirt = len(occ), len(virt)
    S = gs.getOverlapMatrix()
    if gradS is None:
        gradS, gradH0 = sk_gradients(Z, pos, exact=exact)
    D = 2 * orbs_occ @ orbs_occ.T
    D0 = np.diag(np.concatenate([[1.0] if z == 1 else [2.0, 2.0 / 3, 2.0 / 3, 2.0 / 3] for z in Z]))
    dD = D - D0
    FDmD0 = F_op(dD, S, gradS, g0_AO, g1_AO)
    FlrDmD0 = Flr_op(dD, S, gradS, g0lr_AO, g1lr_AO)
    Den = 2 * orbs_occ @ ei @ orbs_occ.T
    gradE0 = np.tensordot(gradH0, D, axes=([1, 2], [0, 1]))
    gradE0 += 0.5 * np.tensordot(FDmD0, dD, axes=([1, 2], [0, 1]))
    gradE0 -= 0.25 * np.tensordot(FlrDmD0, dD, axes=([1, 2], [0, 1]))
    gradE0 -= np.tensordot(gradS, Den, axes=([1, 2], [0, 1]))
    assert abs(np.sum(XmY * XpY) - 1.0) < 1e-10
    qoo, qov, qvv = O.trans_charges(orbs_occ, orbs_virt, S, opa)
    Uab = np.tensordot(XpY, XmY, axes=(0, 0)) + np.tensordot(XmY, XpY, axes=(0, 0))
    Uij = np.tensordot(XpY, XmY, axes=(1, 1)) + np.tensordot(XmY, XpY, axes=(1, 1))
    Vab = np.tensordot(ei @ XpY, XpY, axes=(0, 0)) + np.tensordot(ei @ XmY, XmY, axes=(0, 0))
    Vij = np.tensordot(XpY @ ea, XpY, axes=(1, 1)) + np.tensordot(XmY @ ea, XmY, axes=(1, 1))
    Tab = 0.5 * (np.tensordot(XpY, XpY, axes=(0, 0)) + np.tensordot(XmY, XmY, axes=(0, 0)))
    Tij = 0.5 * (np.tensordot(XpY, XpY, axes=(1, 1)) + np.tensordot(XmY, XmY, axes=(1, 1)))
    Hp = lambda *a: Hplus(g0, g0lr, *a)
    Hm = lambda *a: Hminus(g0lr, *a)
    Gij = Hp(qoo, qvv, qov, qov, qov, qov, Tab) - Hp(qoo, qoo, qoo, qoo, qoo, qoo, Tij)
    qvo = np.swapaxes(qov, 1, 2)
    Qij = Omega * Uij - Vij + Gij
    Qia = np.tensordot(XpY, Hp(qvv, qov, qvo, qvv, qvv, qvo, XpY), axes=(1, 1))
    Qia += np.tensordot(XmY, Hm(qvv, qvo, qvo, qvv, XmY), axes=(1, 1))
    Qia += Hp(qov, qvv, qov, qvv, qov, qvv, Tab)
    Qia -= Hp(qov, qoo, qoo, qvo, qoo, qvo, Tij)
    Qai = (np.tensordot(XpY, Hp(qoo, qov, qoo, qov, qov, qoo, XpY), axes=(0, 0))
           + np.tensordot(XmY, Hm(qov, qoo, qoo, qov, XmY), axes=(0, 0)))
    Qab = Omega * Uab + Vab
    Ria = Qai.T - Qia
    omega = orbe_virt[None, :] - orbe_occ[:, None]
    ApB = lambda vs: O.ApBv(g0, g0lr, qoo, qov, qvv, omega, vs, 1)
    Zia = O.Krylov_solver_Zvector(ApB, omega, Ria.reshape(nocc, nvirt, 1), None, conv=zconv)[:, :, 0]
    Wij = Qij + Hp(qoo, qov, qoo, qov, qov, qoo, Zia)
    Wij[np.diag_indices_from(Wij)] /= 2.0
    Wia = Qai.T + ei @ Zia
    Wab = Qab.copy()
    Wab[np.diag_indices_from(Wab)] /= 2.0
    W = np.block([[Wij, Wia], [Wia.T, Wab]])
    gradH = gradH0 + FDmD0 - 0.5 * FlrDmD0
    Too = orbs_occ @ Tij @ orbs_occ.T
    Tvv = orbs_virt @ Tab @ orbs_virt.T
    Zao = orbs_occ @ Zia @ orbs_virt.T
    allorbs = np.hstack((orbs_occ, orbs_virt))
    Wao = allorbs @ np.triu(W) @ allorbs.T
    XpYao = orbs_occ @ XpY @ orbs_virt.T
    XmYao = orbs_occ @ XmY @ orbs_virt.T
    fv = F_op(XpYao, S, gradS, g0_AO, g1_AO)
    flr_p = Flr_op(XpYao + XpYao.T, S, gradS, g0lr_AO, g1lr_AO)
    flr_m = -Flr_op(XmYao - XmYao.T, S, gradS, g0lr_AO, g1lr_AO)
    gradExc = np.tensordot(gradH, Tvv - Too + Zao, axes=([1, 2], [0, 1]))
    gradExc -= np.tensordot(gradS, Wao, axes=([1, 2], [0, 1]))
    gradExc += 2 * np.tensordot(XpYao, fv, axes=([0, 1], [1, 2]))
    gradExc -= 0.5 * np.tensordot(XpYao, flr_p, axes=([0, 1], [1, 2]))
    gradExc -= 0.5 * np.tensordot(XmYao, flr_m, axes=([0, 1], [1, 2]))
    return gradE0, gradExc, Zia
