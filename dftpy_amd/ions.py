"""Ions in atomic units (ions.py:5-156 of the reference, without ASE)."""
import numpy as np

from .constants import Units


class Ions:
    """Atoms with positions and cell in bohr.  ``units='ase'`` converts from
    Angstrom input like ``Ions(..., units='ase')`` of the reference."""

    def __init__(self, symbols=None, positions=None, cell=None, scaled_positions=None, charges=None, units="au",
                 pbc=None, **kwargs):
        self.symbols = list(symbols)
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3).copy()
        if scaled_positions is not None:
            self.positions = np.asarray(scaled_positions, dtype=np.float64).reshape(-1, 3) @ self.cell
        else:
            self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3).copy()
        if units != "au":
            self.cell /= Units.Bohr
            self.positions /= Units.Bohr
        self._charges = None if charges is None else np.asarray(charges, dtype=np.float64)

    @property
    def nat(self):
        return len(self.symbols)

    def __len__(self):
        return self.nat

    def get_scaled_positions(self):
        """Unwrapped fractional coordinates (the reference's Ions carry no pbc)."""
        return np.linalg.solve(self.cell.T, self.positions.T).T

    def get_positions(self):
        return self.positions.copy()

    @property
    def symbols_uniq(self):
        return np.sort(np.unique(self.symbols))

    def set_charges(self, charges=None):
        if isinstance(charges, dict):
            charges = [charges[s] for s in self.symbols]
        elif isinstance(charges, (int, float)):
            charges = np.ones(self.nat) * charges
        self._charges = np.asarray(charges, dtype=np.float64)

    @property
    def charges(self):
        if self._charges is None:
            raise AttributeError("Please call 'set_charges' before use 'charges'.")
        return self._charges

    def get_ncharges(self):
        return float(self.charges.sum())

    @property
    def zval(self):
        z = {}
        for s, c in zip(self.symbols, self.charges):
            z.setdefault(s, c)
        return z

    @property
    def volume(self):
        return abs(float(np.linalg.det(self.cell)))
