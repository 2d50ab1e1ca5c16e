"""Synthetic molecules for the excited-state workloads.

Geometries are lists of ``(Z, [x, y, z])`` in bohr, the atom-list convention of
``DFTB/XYZ.py:70`` (read_xyz converts Angstrom to bohr).

* ``polyene(n)``: planar all-trans C_nH_{n+2} with C=C 1.34, C-C 1.46,
  C-H 1.09 Angstrom and 120 degree angles.
* ``graphene_flake(rows, cols)``: rectangular honeycomb cut of ``rows`` zigzag
  chains with ``cols`` carbons each (C-C 1.42 Angstrom); every carbon with
  fewer than three carbon neighbours is capped with hydrogens (C-H 1.09) along
  the missing bond directions.
* ``graphene_flake_armchair(widths)``: elongated-hexagon flake with armchair
  edges only (C604H68 by default), H-terminated.
* ``displace(atoms, sigma_angstrom, seed)``: seeded Gaussian displacement of
  every Cartesian component.
"""
import numpy as np

ANGSTROM = 1.0 / 0.52917721092  # bohr per Angstrom


def _rot2(v, deg):
    t = np.deg2rad(deg)
    c, s = np.cos(t), np.sin(t)
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1]])


def polyene(n, cc_double=1.34, cc_single=1.46, ch=1.09):
    """All-trans C_nH_{n+2} (n even) in the xy plane, coordinates in bohr."""
    assert n >= 2 and n % 2 == 0
    up = np.array([np.cos(np.pi / 6), np.sin(np.pi / 6)])
    dn = np.array([np.cos(np.pi / 6), -np.sin(np.pi / 6)])
    cpos = [np.zeros(2)]
    for k in range(1, n):
        length = cc_double if (k % 2 == 1) else cc_single
        step = up if (k % 2 == 1) else dn
        cpos.append(cpos[-1] + length * step)
    cpos = np.array(cpos)
    hpos = []
    for k in range(n):
        nbrs = [cpos[j] for j in (k - 1, k + 1) if 0 <= j < n]
        units = [(p - cpos[k]) / np.linalg.norm(p - cpos[k]) for p in nbrs]
        if len(units) == 2:
            d = -(units[0] + units[1])
            d /= np.linalg.norm(d)
            hpos.append(cpos[k] + ch * d)
        else:
            u = units[0]
            for ang in (120.0, -120.0):
                hpos.append(cpos[k] + ch * _rot2(u, ang))
    atoms = [(6, [p[0], p[1], 0.0]) for p in cpos] + [(1, [p[0], p[1], 0.0]) for p in hpos]
    return [(Z, list(np.array(x) * ANGSTROM)) for Z, x in atoms]


def graphene_flake(rows=24, cols=25, cc=1.42, ch=1.09, prune=False):
    """Rectangular H-terminated graphene flake (rows zigzag chains x cols atoms).

    prune=True removes carbons left with a single carbon neighbour (the flake
    corners) before capping, so every carbon is sp2 with one H at most.
    """
    a = cc
    pos = {}
    for r in range(rows):
        for j in range(cols):
            x = j * np.sqrt(3.0) / 2.0 * a
            y = r * 1.5 * a + (0.5 * a if (j + r) % 2 == 1 else 0.0)
            pos[(r, j)] = np.array([x, y])
    keys = list(pos.keys())
    cxy = np.array([pos[k] for k in keys])
    while True:
        d = np.linalg.norm(cxy[:, None, :] - cxy[None, :, :], axis=-1)
        bonded = (d < 1.1 * a) & (d > 0.1)
        if not prune:
            break
        deg = bonded.sum(1)
        if (deg > 1).all():
            break
        cxy = cxy[deg > 1]
    nat = len(cxy)
    hxy = []
    for i in range(nat):
        nb = np.where(bonded[i])[0]
        units = [(cxy[j] - cxy[i]) / np.linalg.norm(cxy[j] - cxy[i]) for j in nb]
        if len(units) == 2:
            dvec = -(units[0] + units[1])
            dvec /= np.linalg.norm(dvec)
            hxy.append(cxy[i] + ch * dvec)
        elif len(units) == 1:
            for ang in (120.0, -120.0):
                hxy.append(cxy[i] + ch * _rot2(units[0], ang))
    if len(hxy) % 2 == 1:
        # closed shell needs an even electron count (4 per C, 1 per H)
        hxy = hxy[:-1]
    atoms = [(6, [p[0], p[1], 0.0]) for p in cxy] + [(1, [p[0], p[1], 0.0]) for p in hxy]
    return [(Z, list(np.array(x) * ANGSTROM)) for Z, x in atoms]


def _cap_hydrogens(cxy, a, ch):
    """Bond graph of the carbons and one H per missing bond (two for a carbon
    with a single carbon neighbour), along the bisector of the missing bonds."""
    d = np.linalg.norm(cxy[:, None, :] - cxy[None, :, :], axis=-1)
    bonded = (d < 1.1 * a) & (d > 0.1)
    hxy = []
    for i in range(len(cxy)):
        nb = np.where(bonded[i])[0]
        units = [(cxy[j] - cxy[i]) / np.linalg.norm(cxy[j] - cxy[i]) for j in nb]
        if len(units) == 2:
            dvec = -(units[0] + units[1])
            dvec /= np.linalg.norm(dvec)
            hxy.append(cxy[i] + ch * dvec)
        elif len(units) == 1:
            for ang in (120.0, -120.0):
                hxy.append(cxy[i] + ch * _rot2(units[0], ang))
    return bonded, hxy


def graphene_flake_armchair(widths=(18.5, 22.5, 22.5), cc=1.42, ch=1.09):
    """H-terminated graphene flake with armchair edges only.

    The honeycomb (C-C ``cc`` Angstrom) is cut by three pairs of parallel lines
    along the armchair directions, at distances ``widths`` (Angstrom) from a ring
    centre, i.e. an elongated hexagon; carbons left with fewer than two carbon
    neighbours are removed and every two-coordinated carbon gets one H.  The
    default (18.5, 22.5, 22.5) gives C604H68 (Nat 672, 1242 doubly occupied
    valence orbitals).  Armchair edges carry no edge states, so the long-range
    corrected ground state stays stable (A - B positive definite), which
    zigzag-edged rectangles of this size do not.
    """
    a = cc
    a1 = np.array([np.sqrt(3.0) * a, 0.0])
    a2 = np.array([np.sqrt(3.0) / 2.0 * a, 1.5 * a])
    R = int(2 * max(widths) / a) + 4
    ij = np.array([(i, j) for i in range(-R, R + 1) for j in range(-R, R + 1)], dtype=float)
    cells = ij[:, :1] * a1 + ij[:, 1:] * a2
    pts = np.concatenate([cells, cells + np.array([0.0, a])]) - np.array([np.sqrt(3.0) / 2.0 * a, 0.5 * a])
    keep = np.ones(len(pts), bool)
    for k, w in enumerate(widths):
        t = np.deg2rad(60.0 * k)
        keep &= np.abs(pts @ np.array([np.cos(t), np.sin(t)])) <= w
    cxy = pts[keep]
    while True:
        d = np.linalg.norm(cxy[:, None, :] - cxy[None, :, :], axis=-1)
        deg = ((d < 1.1 * a) & (d > 0.1)).sum(1)
        if (deg > 1).all():
            break
        cxy = cxy[deg > 1]
    order = np.lexsort((cxy[:, 0], np.round(cxy[:, 1], 6)))
    cxy = cxy[order]
    _, hxy = _cap_hydrogens(cxy, a, ch)
    atoms = [(6, [p[0], p[1], 0.0]) for p in cxy] + [(1, [p[0], p[1], 0.0]) for p in hxy]
    return [(Z, list(np.array(x) * ANGSTROM)) for Z, x in atoms]


def displace(atoms, sigma_angstrom, seed):
    rng = np.random.default_rng(seed)
    out = []
    for Z, x in atoms:
        out.append((Z, list(np.array(x) + rng.normal(0.0, sigma_angstrom * ANGSTROM, 3))))
    return out


def ethylene():
    return polyene(2)


def benzene(cc=1.39, ch=1.09):
    atoms = []
    for k in range(6):
        t = np.deg2rad(60.0 * k)
        atoms.append((6, [cc * np.cos(t), cc * np.sin(t), 0.0]))
    for k in range(6):
        t = np.deg2rad(60.0 * k)
        atoms.append((1, [(cc + ch) * np.cos(t), (cc + ch) * np.sin(t), 0.0]))
    return [(Z, list(np.array(x) * ANGSTROM)) for Z, x in atoms]


def polyene_ensemble(count=256, seed=20261202, sigma=0.05):
    """Uneven ensemble of all-trans polyenes C_nH_{n+2}: n drawn uniformly from
    the even values 8..40, each geometry with its own seeded Gaussian
    displacement (sigma in Angstrom).  Returns [(n, atoms)]."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.arange(8, 41, 2), size=count)
    return [(int(n), displace(polyene(int(n)), sigma, seed + 1 + i)) for i, n in enumerate(ns)]


def flake_c604h68(sigma=0.02, seed=20261202):
    """The armchair graphene flake C604H68 (widths 18.5, 22.5, 22.5 A; Nat 672,
    Norb 2484) with seeded Gaussian displacements of sigma Angstrom."""
    return displace(graphene_flake_armchair((18.5, 22.5, 22.5)), sigma, seed)
