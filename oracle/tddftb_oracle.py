"""CPU FP64 oracle for the LR-TD-DFTB excited-state path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker / the timed CPU baseline.

Each function restates one piece of DFTBaby's algorithm in Python 3 + numpy,
keeping the reference's contraction order so that it can also be timed as "the
reference CPU path" (the reference itself is Python 2 plus f2py Fortran and
cannot run here).  Citations are ``file:line`` in the DFTBaby source tree.

Pinning (see tests/test_oracle.py):
  trans_charges        - the reference's own Fortran (tddftb.f90) built with
                         amdflang into oracle/_ref when present; closed-form
                         invariants sum_A q_A^{pq} = (C^T S C)_{pq} = delta_pq,
                         symmetry q^{pq} = q^{qp}; brute-force triple loop.
  ApBv / AmBv          - the reference's Fortran apbv/ambv (oracle/_ref) and
                         dense buildAB products on small systems.
  buildAB / Casida     - Casida's own self-consistency asserts
                         (Solver.py:145-173), the lc=0 closed form
                         A-B = diag(omega), and TDA/RPA limits.
  HermitianDavidson /
  nonHermitianDavidson - the reference's diag_check criterion: eigenvalues
                         agree with dense Casida to < 1e-10 and X+-Y up to a
                         global sign (LR_TDDFTB.py:493-518).
  oscillator strengths - closed form f = 2/3 Omega |d|^2 checked against an
                         independent dense evaluation; Mulliken variant from q.
  Oia / Lambda2        - brute-force double loop of LR_TDDFTB.py:538-560.
  Krylov_solver_Zvector- the dense solution (A+B)^{-1} B of buildAB's matrices.
  check_stability      - R assembled element by element from its definition
                         (loops over i, a, j, b) on a small molecule.
  ct_shifts            - separated fragments: S_ia -> 0 and R_ia -> the distance
                         between the centres of two centrosymmetric fragments,
                         so T_ia reduces to the closed form of the asymptotic
                         correction; local excitations (large S_ia) get T = 0.
"""
import numpy as np
import scipy.linalg as sla


class ExcitedStatesNotConverged(Exception):
    pass


class ExcitedStatesError(Exception):
    pass


# ---------------------------------------------------------------------------
# a1: active orbitals (LR_TDDFTB.py:167-188)
# ---------------------------------------------------------------------------
def active_orbitals(f, orbe, nr_active_occ=None, nr_active_virt=None):
    occupied = np.where(f > 0.1)[0]
    virtual = np.where(f <= 0.1)[0]
    act_occ = occupied[np.argsort(orbe[occupied], kind="stable")[::-1][:nr_active_occ][::-1]]
    act_virt = virtual[np.argsort(orbe[virtual], kind="stable")[:nr_active_virt]]
    return occupied, virtual, act_occ, act_virt


def active_lm_orbitals(orbs, f, orbe, atomlist, valorbs, nr_active_occ=1000, nr_active_virt=1000,
                       select_lm=((1, 0),)):
    """setActiveLMOrbitals (LR_TDDFTB.py:125-166): only MOs with weight
    sum |C_mu,p|^2 > 1e-15 on AOs whose (l, m) is in select_lm enter the
    occupied/virtual sets; then the same top-k selection as active_orbitals."""
    lm = []
    iao = 0
    for Zi, _ in atomlist:
        for (ni, li, mi) in valorbs[Zi]:
            if (li, mi) in list(select_lm):
                lm.append(iao)
            iao += 1
    lm = np.array(lm, dtype=int)
    is_lm = np.zeros(orbs.shape[1], dtype=bool)
    for imo in range(orbs.shape[1]):
        is_lm[imo] = np.sum(np.abs(orbs[lm, imo]) ** 2) > 1.0e-15
    occupied = np.where((f > 0.1) & is_lm)[0]
    virtual = np.where((f <= 0.1) & is_lm)[0]
    act_occ = occupied[np.argsort(orbe[occupied], kind="stable")[::-1][:nr_active_occ][::-1]]
    act_virt = virtual[np.argsort(orbe[virtual], kind="stable")[:nr_active_virt]]
    return occupied, virtual, act_occ, act_virt


# ---------------------------------------------------------------------------
# a2: Mulliken transition charges (tddftb.f90:12-112, LR_TDDFTB.py:257-319)
# ---------------------------------------------------------------------------
def trans_charges(orbs_occ, orbs_virt, S, orbsPerAtom):
    """q^{pq}_A = 1/2 sum_{mu in A} [C_{mu p}(SC)_{mu q} + C_{mu q}(SC)_{mu p}]."""
    Sc_occ = S @ orbs_occ
    Sc_virt = S @ orbs_virt
    nat = len(orbsPerAtom)
    nocc = orbs_occ.shape[1]
    nvirt = orbs_virt.shape[1]
    q_oo = np.zeros((nat, nocc, nocc))
    q_ov = np.zeros((nat, nocc, nvirt))
    q_vv = np.zeros((nat, nvirt, nvirt))
    mu = 0
    for A in range(nat):
        sl = slice(mu, mu + int(orbsPerAtom[A]))
        co, cv, so, sv = orbs_occ[sl], orbs_virt[sl], Sc_occ[sl], Sc_virt[sl]
        q_oo[A] = 0.5 * (co.T @ so + so.T @ co)
        q_ov[A] = 0.5 * (co.T @ sv + so.T @ cv)
        q_vv[A] = 0.5 * (cv.T @ sv + sv.T @ cv)
        mu += int(orbsPerAtom[A])
    return q_oo, q_ov, q_vv


def trans_charges_bruteforce(orbs_occ, orbs_virt, S, orbsPerAtom):
    """Element-by-element loop of the Python twin at LR_TDDFTB.py:268-299 (tiny inputs only)."""
    Sc_o = S @ orbs_occ
    Sc_v = S @ orbs_virt
    nat = len(orbsPerAtom)
    no, nv = orbs_occ.shape[1], orbs_virt.shape[1]
    q_oo = np.zeros((nat, no, no)); q_ov = np.zeros((nat, no, nv)); q_vv = np.zeros((nat, nv, nv))
    mu = 0
    for A in range(nat):
        for _ in range(int(orbsPerAtom[A])):
            for i in range(no):
                for a in range(nv):
                    q_ov[A, i, a] += 0.5 * (orbs_occ[mu, i] * Sc_v[mu, a] + orbs_virt[mu, a] * Sc_o[mu, i])
                for j in range(no):
                    q_oo[A, i, j] += 0.5 * (orbs_occ[mu, i] * Sc_o[mu, j] + orbs_occ[mu, j] * Sc_o[mu, i])
            for a in range(nv):
                for b in range(nv):
                    q_vv[A, a, b] += 0.5 * (orbs_virt[mu, a] * Sc_v[mu, b] + orbs_virt[mu, b] * Sc_v[mu, a])
            mu += 1
    return q_oo, q_ov, q_vv


# ---------------------------------------------------------------------------
# a3: orbital energy / occupation differences (LR_TDDFTB.py:378-399)
# ---------------------------------------------------------------------------
def orbital_differences(orbe, f, act_occ, act_virt):
    omega = orbe[act_virt][None, :] - orbe[act_occ][:, None]
    df = f[act_occ][:, None] - f[act_virt][None, :]
    return omega, df


# ---------------------------------------------------------------------------
# a7 matvecs: _ApBv / _AmBv (Solver.py:583-635); vs has shape (nocc, nvirt, nvec)
# ---------------------------------------------------------------------------
def ApBv(g0, g0_lr, q_oo, q_ov, q_vv, omega, vs, lc):
    us = np.zeros(vs.shape)
    for l in range(vs.shape[2]):
        v = vs[:, :, l]
        u = omega * v
        tmp = np.tensordot(q_ov, v, axes=([1, 2], [0, 1]))
        tmp2 = np.dot(g0, tmp)
        u += 4 * np.tensordot(q_ov, tmp2, axes=(0, 0))
        if lc == 1:
            tmp = np.tensordot(q_vv, v, axes=(2, 1))
            tmp2 = np.tensordot(g0_lr, tmp, axes=(1, 0))
            u -= np.tensordot(q_oo, tmp2, axes=([0, 2], [0, 2]))
            tmp = np.tensordot(q_ov, v, axes=(1, 0))
            tmp2 = np.tensordot(g0_lr, tmp, axes=(1, 0))
            u -= np.tensordot(q_ov, tmp2, axes=([0, 2], [0, 2]))
        us[:, :, l] = u
    return us


def AmBv(g0, g0_lr, q_oo, q_ov, q_vv, omega, vs, lc):
    us = np.zeros(vs.shape)
    for l in range(vs.shape[2]):
        v = vs[:, :, l]
        u = omega * v
        if lc == 1:
            tmp = np.tensordot(q_ov, v, axes=(1, 0))
            tmp2 = np.tensordot(g0_lr, tmp, axes=(1, 0))
            u += np.tensordot(q_ov, tmp2, axes=([0, 2], [0, 2]))
            tmp = np.tensordot(q_vv, v, axes=(2, 1))
            tmp2 = np.tensordot(g0_lr, tmp, axes=(1, 0))
            u -= np.tensordot(q_oo, tmp2, axes=([0, 2], [0, 2]))
        us[:, :, l] = u
    return us


# ---------------------------------------------------------------------------
# a8: dense A, B and Casida (Solver.py:35-181)
# ---------------------------------------------------------------------------
def buildAB(gamma, gamma_lr, q_oo, q_vv, q_ov, omega, omega_shift, df, nocc, nvirt, multiplicity="S"):
    K_lr_A = np.tensordot(q_oo, np.tensordot(gamma_lr, q_vv, axes=(1, 0)), axes=(0, 0))
    K_lr_A = np.swapaxes(K_lr_A, 1, 2)
    K_lr_B = np.tensordot(q_ov, np.tensordot(gamma_lr, q_ov, axes=(1, 0)), axes=(0, 0))
    K_lr_B = np.swapaxes(K_lr_B, 1, 3)
    K_A = -K_lr_A
    K_B = -K_lr_B
    if multiplicity == "S":
        K_singlet = 2.0 * np.tensordot(q_ov, np.tensordot(gamma, q_ov, axes=(1, 0)), axes=(0, 0))
        K_A = K_A + K_singlet
        K_B = K_B + K_singlet
    n = nocc * nvirt
    KA = np.reshape(K_A, (n, n))
    KB = np.reshape(K_B, (n, n))
    dfhalf = np.reshape(df / 2.0, n)
    om = np.reshape(omega, n)
    A = np.diag(dfhalf * om) + dfhalf[:, None] * KA * dfhalf[None, :]
    B = dfhalf[:, None] * KB * dfhalf[None, :]
    sh = np.reshape(omega_shift, n)
    A[np.diag_indices(n)] += sh
    B[np.diag_indices(n)] += sh
    return A, B


def Casida(gamma, gamma_lr, q_oo, q_vv, q_ov, omega, omega_shift, df, nocc, nvirt, multiplicity="S"):
    A, B = buildAB(gamma, gamma_lr, q_oo, q_vv, q_ov, omega, omega_shift, df, nocc, nvirt, multiplicity)
    AmB = A - B
    offdiag = np.linalg.norm(np.diag(np.diag(AmB)) - AmB)
    if offdiag < 1.0e-10:
        sqAmB = np.diag(np.sqrt(np.diag(AmB)))
    else:
        sqAmB = np.real(sla.sqrtm(AmB))
    R = sqAmB @ (A + B) @ sqAmB
    assert np.linalg.norm(R - R.T) < 1.0e-10, "TD-DFTB matrix not hermitian"
    Omega2, F = sla.eigh(R)
    idx = np.argsort(Omega2, kind="stable")
    Omega2 = Omega2[idx]
    F = F[:, idx]
    Omega = np.sqrt(Omega2)
    XpY = (sqAmB @ F) / np.sqrt(Omega)
    XmY = ((A + B) @ XpY) / Omega
    C = np.linalg.solve(sqAmB, XpY * np.sqrt(Omega))
    shape = (nocc * nvirt, nocc, nvirt)
    return (Omega, np.reshape(C.T, shape), np.reshape(XmY.T, shape), np.reshape(XpY.T, shape))


# ---------------------------------------------------------------------------
# a6/a7: Davidson solvers (Solver.py:232-251, 404-578, 637-917)
# ---------------------------------------------------------------------------
def initial_expansion_vectors(omega, lmax):
    nocc, nvirt = omega.shape
    bs = np.zeros((nocc, nvirt, lmax))
    idx = np.argsort(np.ravel(omega), kind="stable")
    for l in range(lmax):
        i, a = divmod(int(idx[l]), nvirt)
        bs[i, a, l] = 1.0
    return bs


def _norm(v):
    return np.sqrt(np.tensordot(v, v, axes=([0, 1], [0, 1])))


def _reorder_vectors_Lambda2(Oia, w2, T, L2_thresh):
    nst = T.shape[2]
    L2 = np.array([np.tensordot(T[:, :, n] ** 2, Oia, axes=([0, 1], [0, 1])) for n in range(nst)])
    good = np.where(L2 > L2_thresh)[0]
    bad = np.where(L2 <= L2_thresh)[0]
    order = np.concatenate([good, bad])
    return w2[order], T[:, :, order]


def HermitianDavidson(g0, q_ov, omega, omega_shift, nocc, nvirt, XmYguess, XpYguess, Oia,
                      selector=None, multiplicity="S", nstates=4, ifact=1, conv=1.0e-14,
                      maxiter=10, L2_thresh=0.5, verbose=0):
    omega2 = omega ** 2
    omega_sq = np.sqrt(omega)
    omega_sq_inv = 1.0 / omega_sq
    wq_ov = q_ov * omega_sq
    om = omega2 + 2 * omega * omega_shift

    def Rv(vs):
        us = np.zeros(vs.shape)
        for l in range(vs.shape[2]):
            v = vs[:, :, l]
            u = om * v
            tmp = np.tensordot(wq_ov, v, axes=([1, 2], [0, 1]))
            tmp2 = np.dot(g0, tmp)
            u += 4 * np.tensordot(wq_ov, tmp2, axes=(0, 0))
            us[:, :, l] = u
        return us

    kmax = nocc * nvirt
    lmax = min(ifact * nstates, kmax)
    if XpYguess is None:
        bs = initial_expansion_vectors(np.sqrt(om), lmax)
    else:
        bs = np.zeros((nocc, nvirt, lmax))
        for n in range(lmax):
            bs[:, :, n] = omega_sq_inv * XpYguess[n]
            bs[:, :, n] /= _norm(bs[:, :, n])
    l = lmax
    k = nstates
    history = []
    for it in range(maxiter):
        rbs = Rv(bs)
        Hb = np.tensordot(bs, rbs, axes=([0, 1], [0, 1]))
        w2, Tb = np.linalg.eigh(Hb)
        T = np.tensordot(bs, Tb, axes=(2, 0))
        if L2_thresh > 0.0:
            w2, T = _reorder_vectors_Lambda2(Oia, w2, T, L2_thresh)
        w = np.sqrt(w2)
        W = Rv(T) - w2 * T
        norms = np.array([_norm(W[:, :, n]) for n in range(k)])
        history.append((l, norms.copy()))
        if (norms < conv).all():
            break
        dkmax = min(kmax - l, k)
        eps = 0.01 * conv
        nc = int(np.sum(norms > eps))
        dk = min(dkmax, nc)
        Qs = np.zeros((nocc, nvirt, dk))
        nb = 0
        for n in range(dkmax):
            wD = w[n] - omega
            small = abs(wD) < 1.0e-6
            wD[small] = 1.0e-6 * omega[small]
            if norms[n] > eps:
                Qs[:, :, nb] = 1.0 / wD * W[:, :, n]
                nb += 1
        assert nb == dk
        bs_new = np.concatenate([bs, Qs], axis=2)
        Q, _ = sla.qr(np.reshape(bs_new, (nocc * nvirt, l + dk)), mode="economic")
        bs = np.reshape(Q, (nocc, nvirt, l + dk))
        l = bs.shape[-1]
    else:
        raise ExcitedStatesNotConverged("Hermitian Davidson did not converge after %d iterations" % maxiter)
    Omega = w[:k]
    XpY = np.zeros((nocc, nvirt, k)); XmY = np.zeros((nocc, nvirt, k)); C = np.zeros((nocc, nvirt, k))
    for n in range(k):
        XpY[:, :, n] = omega_sq / np.sqrt(Omega[n]) * T[:, :, n]
        XmY[:, :, n] = np.sqrt(Omega[n]) * omega_sq_inv * T[:, :, n]
        C[:, :, n] = T[:, :, n]
    out = (Omega, np.rollaxis(C, 2), np.rollaxis(XmY, 2), np.rollaxis(XpY, 2))
    HermitianDavidson.history = history
    return out


def nonHermitianDavidson(g0, g0_lr, q_oo, q_vv, q_ov, omega, nocc, nvirt, XmYguess, XpYguess, w_guess,
                         selector=None, multiplicity="S", nstates=4, ifact=1, conv=1.0e-14,
                         maxiter=10, verbose=0, lc=1):
    assert multiplicity == "S"

    def ApB(vs):
        return ApBv(g0, g0_lr, q_oo, q_ov, q_vv, omega, vs, lc)

    def AmB(vs):
        return AmBv(g0, g0_lr, q_oo, q_ov, q_vv, omega, vs, lc)

    kmax = nocc * nvirt
    if XpYguess is None:
        lmax = min(ifact * nstates, kmax)
        bs = initial_expansion_vectors(omega, lmax)
    else:
        assert XmYguess.shape[0] == nstates and XpYguess.shape[0] == nstates
        lmaxp = nstates
        lmaxm = min(nstates, kmax - nstates)
        lmax = lmaxp + lmaxm
        bs = np.zeros((nocc, nvirt, lmax))
        bs[:, :, :lmaxp] = np.rollaxis(XmYguess, 0, 3)
        bs[:, :, lmaxp:] = np.rollaxis(XpYguess, 0, 3)[:, :, :lmaxm]
        for n in range(lmax):
            bs[:, :, n] /= _norm(bs[:, :, n])
    l = lmax
    k = nstates
    history = []
    Tb = None
    for it in range(maxiter):
        if (XpYguess is None) or it > 0:
            bp = ApB(bs)
            bm = AmB(bs)
            Mp = np.tensordot(bs, bp, axes=([0, 1], [0, 1]))
            Mm = np.tensordot(bs, bm, axes=([0, 1], [0, 1]))
            Mmsq = np.real(sla.sqrtm(Mm))
            Mh = Mmsq @ Mp @ Mmsq
            err = np.sum(abs(Mh - Mh.T))
            if err > 1.0e-10:
                raise ExcitedStatesError("Mh is not hermitian! error = %s" % err)
            w2, Tb = np.linalg.eigh(Mh)
            w = np.sqrt(w2)
            wsq = np.sqrt(w)
            Rb = (Mmsq @ Tb) / wsq
            Lb = (Mp @ Rb) / w
            err = np.sum(abs(Lb.T @ Rb - np.eye(len(w))))
            if err > 1.0e-3:
                raise ExcitedStatesError("(X+Y) and (X-Y) vectors not orthonormal! error = %s" % err)
            L = np.tensordot(bs, Lb, axes=(2, 0))
            R = np.tensordot(bs, Rb, axes=(2, 0))
        else:
            R = np.rollaxis(XpYguess, 0, 3)
            L = np.rollaxis(XmYguess, 0, 3)
            w = w_guess
        WL = ApB(R) - L * w
        WR = AmB(L) - R * w
        normsL = np.array([_norm(WL[:, :, n]) for n in range(k)])
        normsR = np.array([_norm(WR[:, :, n]) for n in range(k)])
        norms = normsL + normsR
        history.append((l, norms.copy()))
        if (norms < conv).all() and it > 0:
            break
        dkmax = min(kmax - l, 2 * k)
        eps = 0.01 * conv
        ncL = int(np.sum(normsL > eps))
        ncR = int(np.sum(normsR > eps))
        dkR = min(int(np.ceil(dkmax / 2.0)), ncL)
        dkL = min(dkmax - dkR, ncR)
        dk = dkR + dkL
        Qs = np.zeros((nocc, nvirt, dk))
        nb = 0
        for WX, nrm in ((WL, normsL), (WR, normsR)):
            for n in range(k):
                if nb == dk:
                    break
                wD = w[n] - omega
                small = abs(wD) < 1.0e-6
                wD[small] = 1.0e-6 * omega[small]
                if nrm[n] > eps:
                    Qs[:, :, nb] = 1.0 / wD * WX[:, :, n]
                    nb += 1
        assert nb == dk
        bs_new = np.concatenate([bs, Qs], axis=2)
        Q, _ = sla.qr(np.reshape(bs_new, (nocc * nvirt, l + dk)), mode="economic")
        bs = np.reshape(Q, (nocc, nvirt, l + dk))
        l = bs.shape[-1]
    else:
        raise ExcitedStatesNotConverged("non-Hermitian Davidson did not converge after %d iterations" % maxiter)
    Omega = w[:k]
    XpY = R[:, :, :k]
    XmY = L[:, :, :k]
    T = np.tensordot(bs, Tb, axes=(2, 0))
    C = T[:, :, :k]
    nonHermitianDavidson.history = history
    return Omega, np.rollaxis(C, 2), np.rollaxis(XmY, 2), np.rollaxis(XpY, 2)


# ---------------------------------------------------------------------------
# a5, a9, a10: Lambda2, transition dipoles, oscillator strengths
# ---------------------------------------------------------------------------
def Lambda2_calcOia(q_oo, q_vv, gaussOmega):
    """O_ia = q_ii^T Omega q_aa / sqrt(O_ii O_aa)  (LR_TDDFTB.py:538-560)."""
    qii = np.einsum("Aii->Ai", q_oo)
    qaa = np.einsum("Aaa->Aa", q_vv)
    Oii = np.einsum("Ai,AB,Bi->i", qii, gaussOmega, qii)
    Oaa = np.einsum("Aa,AB,Ba->a", qaa, gaussOmega, qaa)
    Oia = qii.T @ gaussOmega @ qaa
    return Oia / np.sqrt(Oii[:, None] * Oaa[None, :])


def Lambda2_calcOia_loop(q_oo, q_vv, gaussOmega):
    no, nv = q_oo.shape[1], q_vv.shape[1]
    Oia = np.zeros((no, nv))
    for i in range(no):
        Oii = q_oo[:, i, i] @ (gaussOmega @ q_oo[:, i, i])
        for a in range(nv):
            Oaa = q_vv[:, a, a] @ (gaussOmega @ q_vv[:, a, a])
            Oia[i, a] = q_oo[:, i, i] @ (gaussOmega @ q_vv[:, a, a]) / np.sqrt(Oii * Oaa)
    return Oia


def check_stability(gamma, gamma_lr, qtrans_oo, qtrans_vv, qtrans_ov, omega, nocc, nvirt, k=6):
    """Lowest eigenvalues of R = diag(omega) - (ij|ab) - (ia|jb)_lr,
    StabilityAnalysis.py:24-56 (dense eigh instead of eigsh(which='SA'))."""
    ijab = np.tensordot(qtrans_oo, np.tensordot(gamma, qtrans_vv, axes=(1, 0)), axes=(0, 0))
    ijab = np.swapaxes(ijab, 1, 2)
    iajb_lr = np.tensordot(qtrans_ov, np.tensordot(gamma_lr, qtrans_ov, axes=(1, 0)), axes=(0, 0))
    n = nocc * nvirt
    R = np.diag(np.reshape(omega, n)) - np.reshape(ijab, (n, n)) - np.reshape(iajb_lr, (n, n))
    return np.linalg.eigvalsh(R)[:min(k, n)], R


def gram_schmidt_2d(W):
    """Modified Gram-Schmidt of the columns of W (nocc, nvirt, n), Solver.py:291-340
    (the linear-dependence pruning of 317-323 is not needed by the callers here)."""
    V = np.zeros(W.shape)
    V[:, :, 0] = W[:, :, 0] / _norm(W[:, :, 0])
    for n in range(1, W.shape[-1]):
        V[:, :, n] = np.copy(W[:, :, n])
        for m in range(n):
            V[:, :, n] -= np.sum(V[:, :, m] * V[:, :, n]) * V[:, :, m]
        V[:, :, n] /= _norm(V[:, :, n])
    return V


def Krylov_solver_Zvector(A, Adiag, B, X0, maxiter=1000, conv=1.0e-14):
    """A.X = B by the Krylov (Galerkin) iteration of Solver.py:921-1012: basis
    Ainv B (or X0), projected solve, residual, Ainv-preconditioned residuals as new
    vectors, Gram-Schmidt of the whole set."""
    nocc, nvirt, k = B.shape
    kmax = nocc * nvirt
    l = k
    Ainv = 1.0 / Adiag
    if X0 is None:
        bs = np.zeros((nocc, nvirt, k))
        for n in range(k):
            bs[:, :, n] = Ainv * B[:, :, n]
    else:
        bs = X0
    for it in range(maxiter):
        Ab = np.tensordot(bs, A(bs), axes=([0, 1], [0, 1]))
        Bb = np.tensordot(bs, B, axes=([0, 1], [0, 1]))
        Xb = np.linalg.solve(Ab, Bb)
        X = np.tensordot(bs, Xb, axes=(2, 0))
        W = A(X) - B
        norms = np.array([_norm(W[:, :, n]) for n in range(k)])
        if (norms < conv).all():
            return X
        dkmax = min(kmax - l, k)
        eps = 0.01 * conv
        dk = min(dkmax, int(np.sum(norms > eps)))
        Qs = np.zeros((nocc, nvirt, dk))
        nb = 0
        for n in range(dkmax):
            if norms[n] > eps:
                Qs[:, :, nb] = Ainv * W[:, :, n]
                nb += 1
        bs = gram_schmidt_2d(np.concatenate([bs, Qs], axis=2))
        l = bs.shape[-1]
    raise Exception("Z-vector equations: Krylov solver for linear system did not converge after %s iterations "
                    "(threshold: %e) !!!" % (it + 1, conv))


def ct_shifts(q_oo, q_vv, gaussOmega, positions, orbe_virt, omega, Sc=0.0001, Rc=0.1):
    """Long-range CT correction T_ia of LongRangeCT.py:13-76 (Neugebauer et al.,
    J. Chem. Phys. 124, 214102 (2006)): S_ia = q_ii^T Omega q_aa (19-35), the
    distance R_ia of the Mulliken orbital centres (37-49), D_ia = -eps_a (53-57),
    T_ia = exp(-(S/Sc)^2) exp(-(Rc/R)^2) (D - 1/R + (D - 1/R)^2 / (2 w)) clipped
    to [0, 1e10] (59-68)."""
    nocc, nvirt = omega.shape
    Nat = q_oo.shape[0]
    qi = np.zeros((Nat, nocc))
    qa = np.zeros((Nat, nvirt))
    for i in range(nocc):
        qi[:, i] = q_oo[:, i, i]
    for a in range(nvirt):
        qa[:, a] = q_vv[:, a, a]
    Sia = np.dot(qi.T, np.dot(gaussOmega, qa))
    Xia = np.zeros((nocc, nvirt))
    Yia = np.zeros((nocc, nvirt))
    Zia = np.zeros((nocc, nvirt))
    for A in range(Nat):
        xA, yA, zA = positions[A]
        dqia = np.outer(qi[A, :], np.ones(nvirt)) - np.outer(np.ones(nocc), qa[A, :])
        Xia += dqia * xA
        Yia += dqia * yA
        Zia += dqia * zA
    Ria = np.sqrt(Xia ** 2 + Yia ** 2 + Zia ** 2)
    Dia = -np.outer(np.ones(nocc), orbe_virt)
    with np.errstate(divide="ignore", invalid="ignore"):
        Tia = np.exp(-(Sia / Sc) ** 2) * np.exp(-(Rc / Ria) ** 2) * (Dia - 1.0 / Ria + (Dia - 1.0 / Ria) ** 2 / (2 * omega))
    return np.clip(Tia, 0.0, 1.0e10), Sia, Ria


def Lambda2(Cij, Oia):
    """LR_TDDFTB.py:561-565."""
    return np.sum(np.sum(Cij ** 2 * Oia, axis=-1), axis=-1)


def transition_dipoles(Cij, orbs, DipMat, act_occ, act_virt):
    """d_I = sum_ia C^I_ia <i|r|a>  (LR_TDDFTB.py:574-592).  The reference forms
    the full MO matrices orbs^T D orbs and selects the active block; the block
    orbs[:, occ]^T D orbs[:, virt] is formed directly (same numbers)."""
    out = []
    Co, Cv = orbs[:, act_occ], orbs[:, act_virt]
    for c in range(3):
        Dc = np.ascontiguousarray(DipMat[:, :, c])
        tact = Co.T @ (Dc @ Cv)
        out.append(np.tensordot(Cij, tact, axes=([1, 2], [0, 1])))
    return tuple(out)


def transition_dipoles_mulliken(Cij, q_ov, positions):
    """d_I = sum_A R_A sum_ia q^{ia}_A C^I_ia  (LR_TDDFTB.py:567-573)."""
    fvec = np.zeros((3, Cij.shape[0]))
    for A in range(q_ov.shape[0]):
        tdip = np.tensordot(q_ov[A], Cij, axes=([0, 1], [1, 2]))
        fvec += np.outer(positions[A], tdip)
    return fvec[0], fvec[1], fvec[2]


def oscillator_strengths(Omega, tdip, multiplicity="S"):
    """f_I = 2/3 Omega_I |d_I|^2 for singlets, 0 for triplets (LR_TDDFTB.py:698-705)."""
    if multiplicity != "S":
        return 0.0 * Omega
    tx, ty, tz = tdip
    return 2.0 / 3.0 * Omega * (abs(tx) ** 2 + abs(ty) ** 2 + abs(tz) ** 2)


# ---------------------------------------------------------------------------
# driver mirroring LR_TDDFTB._solve_lr_equations (LR_TDDFTB.py:403-527)
# ---------------------------------------------------------------------------
def solve_from_ground_state(gs, nstates=None, nr_active_occ=None, nr_active_virt=None,
                            diag_ifact=1, diag_conv=1.0e-5, diag_maxiter=50, oszis="Dipoles", select_lm=None):
    """Full oracle solve on a ground-state object exposing the DFTB2 getters."""
    f = gs.getOccupation()
    orbe = gs.getKSEnergies()
    orbs = gs.getKSCoefficients()
    S = gs.getOverlapMatrix()
    if select_lm is not None:
        occ, virt, ao, av = active_lm_orbitals(orbs, f, orbe, gs.getGeometry(), gs.getValorbs(), nr_active_occ,
                                               nr_active_virt, select_lm)
    else:
        occ, virt, ao, av = active_orbitals(f, orbe, nr_active_occ, nr_active_virt)
    q_oo, q_ov, q_vv = trans_charges(orbs[:, ao], orbs[:, av], S, gs.orbsPerAtom)
    Oia = Lambda2_calcOia(q_oo, q_vv, gs.getGaussianOverlap())
    gamma = gs.getGamma()
    gamma_lr = gs.getGamma_lr() if gs.long_range_correction == 1 else 0.0 * gamma
    omega, df = orbital_differences(orbe, f, ao, av)
    no, nv = len(ao), len(av)
    shift = 0.0 * omega
    if nstates is None:
        Omega, C, XmY, XpY = Casida(gamma, gamma_lr, q_oo, q_vv, q_ov, omega, shift, df, no, nv)
    elif gs.long_range_correction == 0:
        Omega, C, XmY, XpY = HermitianDavidson(gamma, q_ov, omega, shift, no, nv, None, None, Oia,
                                               nstates=nstates, ifact=diag_ifact, conv=diag_conv,
                                               maxiter=diag_maxiter, L2_thresh=0.0)
    else:
        Omega, C, XmY, XpY = nonHermitianDavidson(gamma, gamma_lr, q_oo, q_vv, q_ov, omega, no, nv,
                                                  None, None, None, nstates=nstates, ifact=diag_ifact,
                                                  conv=diag_conv, maxiter=diag_maxiter, lc=1)
    if oszis == "Dipoles":
        tdip = transition_dipoles(C, orbs, gs.getDipoleMatrix(), ao, av)
    else:
        pos = np.array([x for _, x in gs.getGeometry()])
        tdip = transition_dipoles_mulliken(C, q_ov, pos)
    osc = oscillator_strengths(Omega, tdip)
    return dict(Omega=Omega, C=C, XmY=XmY, XpY=XpY, osc=osc, tdip=tdip, Lambda2=Lambda2(C, Oia),
                q_oo=q_oo, q_ov=q_ov, q_vv=q_vv, omega=omega, Oia=Oia, act_occ=ao, act_virt=av)
