"""Unit constants: ASE's CODATA-2014 set, which the reference takes from
``ase.units`` (``constants.py:3``; used by ``functional/pseudo/recpot.py:14-15``
and ``ions.py:71-77``).  ASE is not installed here, so the two values are
rebuilt from the CODATA-2014 fundamental constants the same way ASE does."""
import math

_c = 299792458.0
_mu0 = 4.0e-7 * math.pi
_hplanck = 6.626070040e-34
_e = 1.6021766208e-19
_me = 9.10938356e-31
_eps0 = 1.0 / _mu0 / _c ** 2
_hbar = _hplanck / (2.0 * math.pi)

# Bohr radius in Angstrom and Hartree in eV
Bohr = 4.0e10 * math.pi * _eps0 * _hbar ** 2 / _me / _e ** 2
Ha = _me * _e ** 3 / 16.0 / math.pi ** 2 / _eps0 ** 2 / _hbar ** 2

ZERO = 1e-30
