"""Ions in atomic units (ions.py:5-156 of the reference, without ASE)."""
import numpy as np

from .constants import Units

# Element symbols indexed by atomic number (0 is a dummy), for formats that
# store atomic numbers instead of symbols.
chemical_symbols = ["X"] + """
H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn
Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce
Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn
Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl
Mc Lv Ts Og""".split()
atomic_numbers = {s: z for z, s in enumerate(chemical_symbols)}


class Ions:
    """Atoms with positions and cell in bohr.  ``units='ase'`` converts from
    Angstrom input like ``Ions(..., units='ase')`` of the reference."""

    def __init__(self, symbols=None, positions=None, cell=None, scaled_positions=None, charges=None, units="au",
                 pbc=None, numbers=None, **kwargs):
        if symbols is None:
            if numbers is None:
                raise ValueError("Ions need symbols or atomic numbers")
            symbols = [chemical_symbols[int(z)] for z in np.asarray(numbers).ravel()]
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
    def numbers(self):
        """Atomic numbers (0 for a symbol that is not an element)."""
        return np.array([atomic_numbers.get(s, 0) for s in self.symbols], dtype=np.int64)

    @property
    def nat(self):
        return len(self.symbols)

    def __len__(self):
        return self.nat

    def get_scaled_positions(self):
        """Unwrapped fractional coordinates (the reference's Ions carry no pbc):
        f with f @ cell = positions."""
        return self.positions @ np.linalg.inv(self.cell)

    def get_positions(self):
        return self.positions.copy()

    def species_index(self, keys):
        """int32 index of each atom's symbol in the sorted list ``keys`` (kept
        for the current symbols, which are compared on every call)."""
        keys = tuple(keys)
        hit = self.__dict__.get("_species_index")
        if hit is not None and hit[0] == keys and hit[1] == self.symbols:
            return hit[2]
        pos = {k: i for i, k in enumerate(keys)}
        missing = sorted(set(self.symbols) - set(keys))
        if missing:
            raise ValueError("no table for species %s" % missing)
        idx = np.fromiter((pos[s] for s in self.symbols), dtype=np.int32, count=len(self.symbols))
        idx.flags.writeable = False
        self._species_index = (keys, list(self.symbols), idx)
        return idx

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
