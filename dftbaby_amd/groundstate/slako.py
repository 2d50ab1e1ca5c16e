"""Slater-Koster matrices S, H0 and the dipole matrix D for the ground-state producer.

The tables are DFTBaby's "homegrown" H/C tables (``slako_hc.npz``, converted by
``make_slako_data``), splined as ``DFTB/SlaterKoster/SKIntegrals.py:801-827`` does
(``scipy.interpolate.splrep(d, table, s=None)``: the cubic interpolating spline)
and evaluated as the default Fortran path does (``DFTB/extensions/splines.f90:25-94``,
``splev_uniform``): the knot interval of a distance x is taken as
``INT((x - t_1) / dt) + 1`` with dt the spacing of the interior knots, which is two
intervals left of the one that holds x, and distances outside the table get the
end values.  The assembly follows ``DFTB/extensions/slako.f90:77-394``: element
(mu, nu) with mu >= nu is formed with the orbital on the lighter element first
(otherwise the two are swapped and the direction reversed), then mirrored; atom
pairs at or beyond ``distance_cutoff`` (30 bohr by default, ``DFTB2.py:67,601-620``)
are zero; dipole elements are shifted by the position of the second orbital's
atom times S (``slako.f90:365,375``) and same-atom dipoles come from the tables at
r = 0 along (0, 0, 1).

Two implementations of the same assembly: ``matrices_device`` runs the HIP
kernel (``dftb_slako_matrices``) and returns device tensors; ``matrices_host``
is the vectorised numpy form used when the producer runs on the CPU.
"""
import ctypes
import os

import numpy as np
from scipy.interpolate import splrep

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "slako_hc.npz")

VALORBS = {1: [(0, 0)], 6: [(0, 0), (1, -1), (1, 0), (1, 1)]}   # (l, m), DFTB2.py:240-245
ATOM_TYPE = {1: 0, 6: 1}                                         # lightest element first
NTAB = 29                                                        # S 0-6, H0 7-13, dipole 14-28
_PAIRS = {(1, 1): "h_h", (1, 6): "h_c", (6, 6): "c_c"}
_PIDX = {-1: 1, 0: 2, 1: 0}


class Tables:
    """Knots [npair][npts] and coefficients [npair][NTAB][npts] of every table."""

    def __init__(self, path=_DATA):
        d = np.load(path)
        self.npair = 3
        knots, coefs = None, None
        for (z1, z2), name in _PAIRS.items():
            i, j = ATOM_TYPE[z1], ATOM_TYPE[z2]
            ip = j * (j + 1) // 2 + i
            grid = d["%s_d" % name]
            for kind, base in (("S", 0), ("H", 7), ("D", 14)):
                for key, vals in zip(d["%s_%s_keys" % (name, kind)], d["%s_%s_vals" % (name, kind)]):
                    t, c, k = splrep(grid, vals, s=None)
                    if k != 3:
                        raise ValueError("cubic splines expected")
                    if knots is None:
                        knots = np.zeros((self.npair, len(t)))
                        coefs = np.zeros((self.npair, NTAB, len(t)))
                    knots[ip] = t
                    coefs[ip, base + int(key[2])] = c
        self.knots, self.coefs = knots, coefs
        self.npts = knots.shape[1]
        # first-derivative coefficients of the S and H0 tables (FITPACK splder's
        # recurrence, a degree-2 spline on the knots t[1..npts-2])
        n = self.npts
        nk1 = n - 4
        self.dcoefs = np.zeros((self.npair, 14, n))
        for ip in range(self.npair):
            t = knots[ip]
            fac = t[np.arange(nk1 - 1) + 4] - t[np.arange(nk1 - 1) + 1]
            for row in range(14):
                c = coefs[ip, row]
                diff = c[1:nk1] - c[:nk1 - 1]
                self.dcoefs[ip, row, :nk1 - 1] = np.where(fac > 0.0, 3.0 * diff / np.where(fac > 0.0, fac, 1.0), 0.0)
        # repulsive potentials (RepulsivePotential.py:80-101): cubic interpolating
        # spline of the table, zero at and beyond its last point
        self.rep_tck, self.rep_dmax = {}, {}
        rk, rc = [], []
        for (z1, z2), name in _PAIRS.items():
            i, j = ATOM_TYPE[z1], ATOM_TYPE[z2]
            ip = j * (j + 1) // 2 + i
            tck = splrep(d["%s_rep_d" % name], d["%s_rep_V" % name], s=0)
            self.rep_tck[ip], self.rep_dmax[ip] = tck, float(d["%s_rep_d" % name][-1])
        rn = max(len(tck[0]) for tck in self.rep_tck.values())
        self.rep_npts = rn
        self.rep_knots = np.zeros((self.npair, rn))
        self.rep_dcoefs = np.zeros((self.npair, rn))
        self.rep_dmax_arr = np.zeros(self.npair)
        for ip, (t, c, k) in self.rep_tck.items():
            if len(t) != rn:
                raise ValueError("repulsive tables of one length expected")
            nk1 = rn - 4
            fac = t[np.arange(nk1 - 1) + 4] - t[np.arange(nk1 - 1) + 1]
            self.rep_knots[ip] = t
            self.rep_dcoefs[ip, :nk1 - 1] = 3.0 * (c[1:nk1] - c[:nk1 - 1]) / fac
            self.rep_dmax_arr[ip] = self.rep_dmax[ip]
        self.orbital_energies = {1: [float(d["h_energy_1s"])],
                                 6: [float(d["c_energy_2s"])] + 3 * [float(d["c_energy_2p"])]}
        self._dev = {}

    def device_rep_arrays(self, dev):
        import torch
        key = "rep" + str(dev)
        if key not in self._dev:
            self._dev[key] = tuple(torch.as_tensor(a, device=dev).contiguous()
                                   for a in (self.rep_knots, self.rep_dcoefs, self.rep_dmax_arr))
        return self._dev[key]

    def device_arrays(self, dev):
        import torch
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = (torch.as_tensor(self.knots, device=dev).contiguous(),
                              torch.as_tensor(self.coefs, device=dev).contiguous(),
                              torch.as_tensor(self.dcoefs, device=dev).contiguous())
        return self._dev[key]


_TABLES = None


def tables():
    global _TABLES
    if _TABLES is None:
        _TABLES = Tables()
    return _TABLES


def orbital_lists(Z, T=None):
    """Per orbital: atom, l, m, H0 diagonal; per atom: type."""
    T = T or tables()
    atom, ls, ms, en = [], [], [], []
    for A, z in enumerate(Z):
        for o, (l, m) in enumerate(VALORBS[int(z)]):
            atom.append(A)
            ls.append(l)
            ms.append(m)
            en.append(T.orbital_energies[int(z)][o])
    types = [ATOM_TYPE[int(z)] for z in Z]
    return (np.array(atom, dtype=np.int32), np.array(ls, dtype=np.int32), np.array(ms, dtype=np.int32),
            np.array(en, dtype=np.float64), np.array(types, dtype=np.int32))


# ----------------------------------------------------------------------------- device
def matrices_device(Z, pos, dev, distance_cutoff=30.0, dipoles=True):
    """(S, H0, D) as device tensors; D is [3][norb][norb] (x, y, z planes) or None."""
    import torch

    from .. import _lib
    from ..engine import _p, _stream
    T = tables()
    atom, ls, ms, en, types = orbital_lists(Z, T)
    norb, nat = len(atom), len(Z)
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    P = t(np.asarray(pos, dtype=np.float64).reshape(nat, 3), torch.float64)
    oa, ol, om, ty = (t(a, torch.int32) for a in (atom, ls, ms, types))
    e = t(en, torch.float64)
    knots, coefs, _ = T.device_arrays(dev)
    S = torch.empty((norb, norb), dtype=torch.float64, device=dev)
    H = torch.empty((norb, norb), dtype=torch.float64, device=dev)
    D = torch.empty((3, norb, norb), dtype=torch.float64, device=dev) if dipoles else None
    _lib.call("dftb_slako_matrices", _p(P), _p(oa), _p(ol), _p(om), _p(ty), _p(e), nat, norb, _p(knots), _p(coefs),
              T.npair, T.npts, ctypes.c_double(distance_cutoff), _p(S), _p(H), _p(D) if D is not None else None,
              _stream())
    return S, H, D


# ----------------------------------------------------------------------------- host
def _basis(t, x):
    """splev_uniform's interval and fpbspl's cubic basis for an array of x."""
    n = len(t)
    k1, nk1 = 4, n - 4
    tb, te = t[k1 - 1], t[nk1]
    arg = np.array(x, dtype=np.float64)
    l = np.empty(len(arg), dtype=np.int64)
    lo, hi = arg <= tb, arg >= te
    mid = ~lo & ~hi
    l[lo], l[hi] = k1, nk1
    arg[lo], arg[hi] = tb, te
    dt = t[k1 + 1] - t[k1]
    l[mid] = ((arg[mid] - t[0]) / dt).astype(np.int64) + 1
    l = np.clip(l, k1, nk1)
    h = np.zeros((len(arg), 4))
    h[:, 0] = 1.0
    for j in range(1, 4):
        hh = h[:, :j].copy()
        h[:, 0] = 0.0
        for i in range(1, j + 1):
            li = l + i
            lj = li - j
            f = hh[:, i - 1] / (t[li - 1] - t[lj - 1])
            h[:, i - 1] = h[:, i - 1] + f * (t[li - 1] - arg)
            h[:, i] = f * (arg - t[lj - 1])
    return l - k1, h


def _values(T, ip, r):
    """All NTAB table values of pair type ip at distances r: [len(r)][NTAB]."""
    c0, h = _basis(T.knots[ip], r)
    idx = c0[:, None] + np.arange(4)[None, :]
    C = T.coefs[ip][:, idx]                          # [NTAB][P][4]
    return np.einsum("pj,tpj->pt", h, C)


def _other_sq(e, i):
    x2, y2, z2 = e[:, 0] * e[:, 0], e[:, 1] * e[:, 1], e[:, 2] * e[:, 2]
    return (y2 + z2, x2 + z2, x2 + y2)[i]


def _rule(l1, m1, l2, m2, e, V, base):
    T = lambda k: V[:, base + k]
    if l1 == 0 and l2 == 0:
        return T(0)
    if l1 == 0:
        return e[:, _PIDX[m2]] * T(2)
    if l2 == 0:
        return e[:, _PIDX[m1]] * T(4)
    i, j = _PIDX[m1], _PIDX[m2]
    if i == j:
        return _other_sq(e, i) * T(5) + e[:, i] * e[:, i] * T(6)
    return e[:, i] * e[:, j] * (T(6) - T(5))


def _rule_dipole(l1, m1, M, l2, m2, e, V):
    D = lambda k: V[:, 14 + k]
    if l1 == 0 and l2 == 0:
        return e[:, M] * D(3)
    if l1 == 0:
        j = _PIDX[m2]
        if M == j:
            return _other_sq(e, j) * D(1) + e[:, j] * e[:, j] * D(4)
        return e[:, M] * e[:, j] * (D(4) - D(1))
    if l2 == 0:
        i = _PIDX[m1]
        if M == i:
            return _other_sq(e, i) * D(6) + e[:, i] * e[:, i] * D(13)
        return e[:, i] * e[:, M] * (D(13) - D(6))
    i, j = _PIDX[m1], _PIDX[m2]
    v = (D(14) - D(7) - 2.0 * D(8)) * e[:, i] * e[:, M] * e[:, j]
    if i == M:
        v = v + D(7) * e[:, j]
    if M == j:
        v = v + D(8) * e[:, i]
    if i == j:
        v = v + D(8) * e[:, M]
    return v


def matrices_host(Z, pos, distance_cutoff=30.0, dipoles=True):
    """(S, H0, D) as numpy arrays; D is [norb][norb][3] (the reference's layout)."""
    T = tables()
    Z = np.asarray([int(z) for z in Z])
    pos = np.asarray(pos, dtype=np.float64).reshape(len(Z), 3)
    atom, ls, ms, en, types = orbital_lists(Z, T)
    norb = len(atom)
    off = np.concatenate([[0], np.cumsum([len(VALORBS[z]) for z in Z])[:-1]]).astype(np.int64)
    S = np.zeros((norb, norb))
    H = np.zeros((norb, norb))
    D = np.zeros((3, norb, norb)) if dipoles else None
    iA, iB = np.tril_indices(len(Z), -1)              # A > B: A's orbitals come later
    dvec = pos[iB] - pos[iA]
    r = np.sqrt(dvec[:, 0] * dvec[:, 0] + dvec[:, 1] * dvec[:, 1] + dvec[:, 2] * dvec[:, 2])
    keep = r < distance_cutoff
    iA, iB, dvec, r = iA[keep], iB[keep], dvec[keep], r[keep]
    for za in (1, 6):
        for zb in (1, 6):
            g = (Z[iA] == za) & (Z[iB] == zb)
            if not g.any():
                continue
            A, B, rr = iA[g], iB[g], r[g]
            e = dvec[g] / rr[:, None]
            swap = ATOM_TYPE[za] > ATOM_TYPE[zb]
            if swap:
                e = -e
            tmin, tmax = sorted((ATOM_TYPE[za], ATOM_TYPE[zb]))
            V = _values(T, tmax * (tmax + 1) // 2 + tmin, rr)
            for a, (la, ma) in enumerate(VALORBS[za]):
                for b, (lb, mb) in enumerate(VALORBS[zb]):
                    mu, nu = off[A] + a, off[B] + b
                    (l1, m1), (l2, m2) = ((lb, mb), (la, ma)) if swap else ((la, ma), (lb, mb))
                    sv = _rule(l1, m1, l2, m2, e, V, 0)
                    S[mu, nu] = sv
                    H[mu, nu] = _rule(l1, m1, l2, m2, e, V, 7)
                    if dipoles:
                        second = A if swap else B
                        for M in range(3):
                            D[M, mu, nu] = _rule_dipole(l1, m1, M, l2, m2, e, V) + pos[second, M] * sv
    S = S + S.T
    H = H + H.T
    S[np.diag_indices(norb)] = 1.0
    H[np.diag_indices(norb)] = en
    if dipoles:
        # lower triangle of each atom's own block: tables at r = 0, direction z
        ez = np.array([[0.0, 0.0, 1.0]])
        for z in (1, 6):
            atoms = np.where(Z == z)[0]
            if len(atoms) == 0:
                continue
            V = _values(T, ATOM_TYPE[z] * (ATOM_TYPE[z] + 1) // 2 + ATOM_TYPE[z], np.zeros(1))
            orbs = VALORBS[z]
            for a, (la, ma) in enumerate(orbs):
                for b, (lb, mb) in enumerate(orbs[:a + 1]):
                    mu, nu = off[atoms] + a, off[atoms] + b
                    for M in range(3):
                        val = _rule_dipole(la, ma, M, lb, mb, ez, V)[0]
                        D[M, mu, nu] = val + (pos[atoms, M] if a == b else 0.0)
                        if a != b:
                            D[M, nu, mu] = D[M, mu, nu]
        low = np.tril_indices(norb, -1)
        for M in range(3):
            Dm = D[M]
            # off-site lower triangle was filled, the upper one is its mirror
            blk = atom[low[0]] != atom[low[1]]
            Dm[low[1][blk], low[0][blk]] = Dm[low[0][blk], low[1][blk]]
        D = np.ascontiguousarray(np.moveaxis(D, 0, 2))
    return S, H, D


def repulsive_energy(Z, pos):
    """E_rep = sum over atom pairs of Vrep(R) (DFTB2.py:2340-2359)."""
    from scipy.interpolate import splev
    T = tables()
    Z = np.asarray([int(z) for z in Z])
    pos = np.asarray(pos, dtype=np.float64).reshape(len(Z), 3)
    iA, iB = np.tril_indices(len(Z), -1)
    R = np.linalg.norm(pos[iA] - pos[iB], axis=1)
    ta = np.array([ATOM_TYPE[z] for z in Z])
    tmin, tmax = np.minimum(ta[iA], ta[iB]), np.maximum(ta[iA], ta[iB])
    ip = tmax * (tmax + 1) // 2 + tmin
    E = 0.0
    for p, tck in T.rep_tck.items():
        m = (ip == p) & (R < T.rep_dmax[p])
        if m.any():
            E += float(np.sum(splev(R[m], tck)))
    return E
