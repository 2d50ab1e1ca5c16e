"""Convert DFTBaby's "homegrown" H/C parameter tables into one numpy data file.

This is a build-time data step, not part of the solver.  It reads the
Slater-Koster tables ``DFTB/SlaterKoster/slako_tables/{h_h,h_c,c_c}.py`` (overlap
S, hamiltonian H and dipole integrals, keyed (l1, l2, table index)) and the
repulsive potentials ``DFTB/RepulsivePotential/reppot_tables/{h_h,h_c,c_c}.py`` as
*data* (the files are generated modules holding ``array([...])`` literals; they
are parsed with ``ast`` and never imported or executed) and the free
pseudo-atom orbital energies of ``DFTB/SlaterKoster/free_pseudo_atoms/{h,c}.py``
(``energy_1s``, ``energy_2s``, ``energy_2p``).  The Hubbard parameters follow
``DFTB/Parameters.py:25-30``: U = (IE - EA) / hartree_to_eV with the atomic
ionisation energies and electron affinities of ``DFTB/AtomicData.py:131-134``.

Usage:  python -m dftbaby_amd.groundstate.make_slako_data /root/reference
writes ``slako_hc.npz`` next to this file.
"""
import ast
import os
import sys

import numpy as np


def _literal(node):
    if isinstance(node, ast.Call) and getattr(node.func, "id", None) == "array":
        return np.array(ast.literal_eval(node.args[0]), dtype=np.float64)
    if isinstance(node, ast.Dict):
        return {ast.literal_eval(k): _literal(v) for k, v in zip(node.keys, node.values)}
    return ast.literal_eval(node)


def parse_table(path, names):
    tree = ast.parse(open(path).read())
    out = {}
    for st in tree.body:
        if isinstance(st, ast.Assign) and len(st.targets) == 1 and isinstance(st.targets[0], ast.Name):
            name = st.targets[0].id
            if name in names:
                out[name] = _literal(st.value)
    return out


def main(refroot):
    skdir = os.path.join(refroot, "DFTB", "SlaterKoster", "slako_tables")
    atdir = os.path.join(refroot, "DFTB", "SlaterKoster", "free_pseudo_atoms")
    data = {}
    for pair in ("h_h", "h_c", "c_c"):
        t = parse_table(os.path.join(skdir, pair + ".py"), {"Z1", "Z2", "d", "S", "H", "Dipole"})
        data["%s_Z" % pair] = np.array([t["Z1"], t["Z2"]], dtype=np.int64)
        data["%s_d" % pair] = t["d"]
        t["D"] = t["Dipole"]
        for kind in ("S", "H", "D"):
            keys = sorted(t[kind].keys())
            data["%s_%s_keys" % (pair, kind)] = np.array(keys, dtype=np.int64)
            data["%s_%s_vals" % (pair, kind)] = np.stack([t[kind][k] for k in keys])
    # repulsive potentials DFTB/RepulsivePotential/reppot_tables/{h_h,h_c,c_c}.py
    # (Z1, Z2, d, Vrep), splined by RepulsivePotential.py:80-86
    repdir = os.path.join(refroot, "DFTB", "RepulsivePotential", "reppot_tables")
    for pair in ("h_h", "h_c", "c_c"):
        t = parse_table(os.path.join(repdir, pair + ".py"), {"Z1", "Z2", "d", "Vrep"})
        data["%s_rep_d" % pair] = np.asarray(t["d"], dtype=np.float64)
        data["%s_rep_V" % pair] = np.asarray(t["Vrep"], dtype=np.float64)
    for el in ("h", "c"):
        t = parse_table(os.path.join(atdir, el + ".py"),
                        {"energy_1s", "energy_2s", "energy_2p"})
        for k, v in t.items():
            data["%s_%s" % (el, k)] = np.float64(v)
    # DFTB/AtomicData.py:2 and :131-134 (H and C entries only)
    hartree_to_eV = 27.211396132
    ie = {"h": 13.5984, "c": 11.2603}
    ea = {"h": 0.75419, "c": 1.262119}
    for el in ("h", "c"):
        data["%s_U" % el] = np.float64((ie[el] - ea[el]) / hartree_to_eV)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "slako_hc.npz")
    np.savez_compressed(out, **data)
    print("wrote", out, "keys:", len(data))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
