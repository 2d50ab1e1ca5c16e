"""Stability of the closed-shell ground state (DFTB/StabilityAnalysis.py:1-56).

The restricted closed-shell ground state is stable against restricted ->
unrestricted relaxation when

    R_{ia,jb} = (e_a - e_i) delta_ij delta_ab - (ij|ab) - (ia|jb)_lr

is positive semidefinite (Seeger and Pople, J. Chem. Phys. 66, 3045 (1977)).
The matrix is assembled and its lowest eigenvalues found on the GPU
(``engine.stability_gpu``).
"""
import numpy as np

from .engine import stability_gpu


def check_stability(gamma, gamma_lr, qtrans_oo, qtrans_vv, qtrans_ov, omega, nocc, nvirt):
    """Prints the lowest (up to 6) eigenvalues of R as StabilityAnalysis.py:24-56
    does, and returns them (ascending)."""
    print("Stability of closed-shell DFTB ground state")
    w = stability_gpu(qtrans_oo, qtrans_vv, qtrans_ov, np.reshape(omega, (nocc, nvirt)), gamma, gamma_lr)
    print(w)
    print("  lowest eigenvalue of restricted/unrestricted stability matrix R: %e" % w[0])
    if w[0] < 0.0:
        print("   !!! restricted/unrestricted instability detected !!!")
    return w
