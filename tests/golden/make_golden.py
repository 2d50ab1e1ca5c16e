"""Write tests/golden/tddftb_f90_golden.npz: inputs and outputs of DFTBaby's own
Fortran kernels (DFTB/extensions/tddftb.f90 trans_charges, apbv, ambv), built
from the reference tree by oracle/Makefile into oracle/_ref.  Run in a
container that has the reference tree:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import ground_state  # noqa: E402
from oracle import tddftb_oracle as O  # noqa: E402
from test_oracle import f90_lib, f90_products, f90_trans_charges  # noqa: E402


def main():
    lib = f90_lib()
    gs = ground_state("C8H10", 1, 0.05, 21)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    ao, av = ao[-8:], av[:9]          # an active space keeps the fixture small
    Co, Cv = C[:, ao], C[:, av]
    opa = np.asarray(gs.orbsPerAtom, dtype=np.int32)
    q_oo, q_ov, q_vv = f90_trans_charges(lib, Co, Cv, S, opa)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    vs = np.random.default_rng(5).standard_normal((len(ao), len(av), 3))
    ApB, AmB = f90_products(lib, g, glr, q_oo, q_ov, q_vv, omega, vs)
    out = os.path.join(HERE, "tddftb_f90_golden.npz")
    np.savez_compressed(out, Co=Co, Cv=Cv, S=S, opa=opa, q_oo=q_oo, q_ov=q_ov, q_vv=q_vv, omega=omega, g=g,
                        glr=glr, vs=vs, ApB=ApB, AmB=AmB)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
