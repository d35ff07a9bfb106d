"""Units: the CODATA-2014 Bohr and Hartree values ase.units computes, which
the reference uses (constants.py:3,66-79).  Built from the fundamental
constants as ase does, since ase is not a dependency here."""
import math


class Units:
    _c = 299792458.0
    _mu0 = 4.0e-7 * math.pi
    _hplanck = 6.626070040e-34
    _e = 1.6021766208e-19
    _me = 9.10938356e-31
    _eps0 = 1.0 / _mu0 / _c ** 2
    _hbar = _hplanck / (2.0 * math.pi)
    Bohr = 4.0e10 * math.pi * _eps0 * _hbar ** 2 / _me / _e ** 2
    Ha = _me * _e ** 3 / 16.0 / math.pi ** 2 / _eps0 ** 2 / _hbar ** 2
    Hartree = Ha
    Angstrom = 1.0
    eV = 1.0
    alpha = 7.2973525664e-3


ENERGY_CONV = {"Hartree": {"eV": Units.Ha}, "eV": {"Hartree": 1.0 / Units.Ha}}
LEN_CONV = {"Angstrom": {"Bohr": 1.0 / Units.Bohr}, "Bohr": {"Angstrom": Units.Bohr}}
SPEED_OF_LIGHT = 1.0 / Units.alpha
ZERO = 1e-30
