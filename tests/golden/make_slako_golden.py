"""Write tests/golden/slako_f90_golden.npz: S, H0 and the dipole matrix D from
DFTBaby's own Fortran (DFTB/extensions/slako.f90 H0andS and DipoleMatrix, with
splines.f90), built from the reference tree by oracle/Makefile into oracle/_ref,
for the geometries of tests/test_slako.py::golden_geometries, and dS/dR, dH0/dR
(gradients_H0andS) for those of gradient_geometries, and the overlap between two
geometries (overlap12) for those of overlap12_geometries.  Run in a container
that has the reference tree:  python tests/golden/make_slako_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import slako_oracle as K  # noqa: E402
from test_slako import (f90_gradients, f90_lib, f90_matrices, f90_overlap12, golden_geometries,  # noqa: E402
                        gradient_geometries, overlap12_geometries)


def main():
    lib = f90_lib()
    tabs = K.load_tables()
    data = {}
    for name, Z, pos in golden_geometries():
        S, H, D = f90_matrices(lib, Z, pos, *tabs)
        data.update({name + "_Z": np.asarray(Z, dtype=np.int64), name + "_pos": np.asarray(pos), name + "_S": S,
                     name + "_H0": H, name + "_D": D})
    for name, Z, pos in gradient_geometries():
        gS, gH = f90_gradients(lib, Z, pos, tabs[0])
        data.update({name + "_Z": np.asarray(Z, dtype=np.int64), name + "_pos": np.asarray(pos),
                     name + "_gradS": gS, name + "_gradH0": gH})
    for name, Z, p1, p2 in overlap12_geometries():
        data.update({name + "_Z": np.asarray(Z, dtype=np.int64), name + "_pos1": p1, name + "_pos2": p2,
                     name + "_S12": f90_overlap12(lib, Z, p1, p2, tabs[0])})
    out = os.path.join(HERE, "slako_f90_golden.npz")
    np.savez_compressed(out, **data)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
