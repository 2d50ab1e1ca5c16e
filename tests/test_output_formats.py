"""Text outputs of the drop-in, byte for byte against the reference's formats
(no GPU: the state of a finished solve is set on the host).

* the excited-state table of _writeExcitations (LR_TDDFTB.py:759-806),
* absorption_spectrum.dat of saveAbsorptionSpectrum (LR_TDDFTB.py:807-832),
* the determineActiveSpace report (LR_TDDFTB.py:196-243).

The expected lines are literals worked out by hand from the reference's format
strings (Python 2 string.rjust(s, n) == s.rjust(n)) for round input values.
"""
import numpy as np

from dftbaby_amd.LR_TDDFTB import LR_TDDFTB


class _GS:
    verbose = 0

    def getKSEnergies(self):
        return np.array([-0.9, -0.8, -0.7, -0.5, -0.25, 0.1])


def _finished_solve():
    td = LR_TDDFTB(ground_state=_GS())
    td.en0 = -1.5
    td.multiplicity = "S"
    td.Omega = np.array([0.2, 0.25])
    td.oscillator_strength = np.array([0.5, 0.0123456789])
    td.tdipX, td.tdipY, td.tdipZ = np.array([0.1, 1.0]), np.array([-0.2, 0.0]), np.array([0.0, -12.5])
    td.Lambda2 = np.array([0.5, 0.98765])
    td.Irreps = ["", ""]
    td.occupied_orbs = np.array([0, 1, 2, 3])
    td.virtual_orbs = np.array([4, 5])
    td.active_occupied_orbs = np.array([2, 3])
    td.active_virtual_orbs = np.array([4, 5])
    C = np.zeros((2, 2, 2))
    C[0, 1, 0] = 1.0                                   # 4 -> 5 only
    C[1, 1, 1] = np.sqrt(0.6)                          # 4 -> 6 (0.6), 3 -> 5 (0.3), 3 -> 6 (0.1)
    C[1, 0, 0] = -np.sqrt(0.3)
    C[1, 0, 1] = np.sqrt(0.1)
    td._Cij_h = C
    return td


HEADER = ("     Spn" + " " * 9 + "exc. en. /hartree      exc. en. / eV   exc. en. / nm    osc. strength"
          "             dominant contrib.               en_KS / eV      transition dipole [TDx TDy TDz]   /\\ diagn.")


def test_excitation_table_bytes():
    txt = _finished_solve().excitation_table()
    lines = txt.split("\n")
    assert lines[0] == " ground state energy: -1.5000000 hartree  -40.8170942 eV"
    assert lines[1] == ""
    assert lines[2] == "           Excited States"
    assert lines[3] == "           ==============" and lines[4] == HEADER
    row1 = ("    1 S    :    0.2000000 hartree      5.4422792 eV    227.8167626 nm     0.5000000"
            + " " * 23 + "       4->5(+1.000e+00)   6.8028       [+0.1000000 -0.2000000 +0.0000000]  0.5000")
    assert lines[5] == row1
    # the two largest contributions: 4->6 (0.6) then 3->5 (0.3); en_KS of the first
    # (eps_6 - eps_4 = 0.6 Ha = 16.3268 eV)
    row2 = ("    2 S    :    0.2500000 hartree      6.8028490 eV    182.2534101 nm     0.0123457"
            + " " * 6 + "       4->6(+7.746e-01),3->5(-5.477e-01)  16.3268      [+1.0000000 +0.0000000 -12.5000000]"
            + "  0.9877")
    assert lines[6] == row2
    assert lines[7] == "" and len(lines) == 8


def test_absorption_spectrum_file_bytes(tmp_path):
    td = _finished_solve()
    f = tmp_path / "absorption_spectrum.dat"
    td.saveAbsorptionSpectrum(str(f))
    assert f.read_text() == ("# Absorption spectrum\n# exc. energy / hartree  osc.strength\n"
                             "           0.2000000              0.5000000\n"
                             "           0.2500000              0.0123457\n")


def test_active_space_report_bytes(capsys):
    """State 1 is a single excitation (3 -> 4 in 0-based MO indices); state 2
    needs 4 -> 6 and 3 -> 5 to reach 1 - threshold (the contribution that crosses
    it, 3 -> 6, is not counted, as in the reference's cumulative test)."""
    td = _finished_solve()
    td.dftb2.verbose = 1
    assert td.determineActiveSpace() == (2, 2)
    out = capsys.readouterr().out
    assert out == ("\nActive Space\n============\n"
                   "State   1: contains excitations from window  occ   3 - virt   4\n"
                   "State   2: contains excitations from window  occ   2 - virt   5\n"
                   "\n"
                   "The suggested active space for describing the lowest 2 states consists of\n"
                   "the highest 2 occupied and the lowest 2 virtual orbitals.\n"
                   "This active space accounts for 0.9999900 of the total probability for each state.\n"
                   "You should check that the active space contains all neighbouring (almost) degenerate orbitals.\n"
                   "Dimension of active space: 4\n"
                   "Dimension of  full  space: 8\n"
                   "\n")
