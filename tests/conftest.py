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


def ground_state(name, lc, sigma=0.0, seed=0):
    """Small ground states for parity tests (cached per session)."""
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate.dftb2 import DFTB2
    key = (name, lc, sigma, seed)
    if key not in _GS_CACHE:
        builders = {"ethylene": g.ethylene, "benzene": g.benzene,
                    "C8H10": lambda: g.polyene(8), "C12H14": lambda: g.polyene(12),
                    "C20H22": lambda: g.polyene(20),
                    "C180H36": lambda: g.graphene_flake_armchair((11.5, 11.5, 11.5))}
        atoms = builders[name]()
        if sigma > 0:
            atoms = g.displace(atoms, sigma, seed)
        gs = DFTB2(atoms, long_range_correction=lc, device="cpu")
        gs.getEnergy()
        _GS_CACHE[key] = gs
    return _GS_CACHE[key]


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
