"""Host-side logic of the engine (no GPU needed)."""
import numpy as np


def test_initial_indices_match_stable_argsort():
    """The partial selection of the Davidson start (Solver.py:232-251 sorts all
    pairs) returns the same indices in the same order as a full stable argsort,
    including ties at the selection threshold."""
    from dftbaby_amd.engine import _initial_indices
    rng = np.random.default_rng(0)
    for _ in range(200):
        o, v = int(rng.integers(1, 30)), int(rng.integers(1, 30))
        vals = np.round(rng.random((o, v)) * rng.integers(1, 6), 1)
        lmax = int(rng.integers(1, o * v + 1))
        i, a = _initial_indices(vals, o, v, lmax)
        assert np.array_equal(i * v + a, np.argsort(vals.ravel(), kind="stable")[:lmax])
