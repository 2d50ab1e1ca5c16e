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


def test_zvector_solver_needs_the_gpu_operator():
    """Solver.Krylov_solver_Zvector takes the (A+B) operator object, not a Python
    callable (the products run on the GPU); the check precedes any device work."""
    import numpy as np
    import pytest
    from dftbaby_amd import Solver
    with pytest.raises(TypeError):
        Solver.Krylov_solver_Zvector(lambda x: x, np.ones((2, 3)), np.ones((2, 3, 1)), None)
