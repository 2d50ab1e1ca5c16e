import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


_GS_CACHE = {}


def _ct_dimer(separation=40.0):
    """All-trans C8H10 centred at the origin and H2 (0.74 Angstrom) on the z axis
    at `separation` bohr: two centrosymmetric fragments beyond the 30 bohr
    Slater-Koster proximity cutoff, so every orbital lives on one fragment, and a bound
    polyene LUMO (eps < -1/separation) that makes the H2 -> C8H10 excitations
    charge-transfer states the correction shifts."""
    from dftbaby_amd.groundstate import geometry as g
    mol = g.polyene(8)
    c = np.mean([x for _, x in mol], axis=0)
    atoms = [(Z, list(np.array(x) - c)) for Z, x in mol]
    h = 0.5 * 0.74 * g.ANGSTROM
    atoms += [(1, [0.0, 0.0, separation - h]), (1, [0.0, 0.0, separation + h])]
    return atoms


def ground_state(name, lc, sigma=0.0, seed=0):
    """Small ground states for parity tests (cached per session)."""
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate.dftb2 import DFTB2
    key = (name, lc, sigma, seed)
    if key not in _GS_CACHE:
        builders = {"ethylene": g.ethylene, "benzene": g.benzene,
                    "C8H10": lambda: g.polyene(8), "C12H14": lambda: g.polyene(12),
                    "C20H22": lambda: g.polyene(20),
                    "C180H36": lambda: g.graphene_flake_armchair((11.5, 11.5, 11.5)),
                    "C8H10-H2": _ct_dimer}
        atoms = builders[name]()
        if sigma > 0:
            atoms = g.displace(atoms, sigma, seed)
        gs = DFTB2(atoms, long_range_correction=lc, device="cpu")
        gs.getEnergy()
        _GS_CACHE[key] = gs
    return _GS_CACHE[key]


def progress(msg):
    """A progress line on the process's stderr, past pytest's output capture, so a
    long GPU test shows it is working."""
    sys.__stderr__.write("[test] %s\n" % msg)
    sys.__stderr__.flush()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
