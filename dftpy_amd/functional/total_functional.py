"""TotalFunctional (functional/total_functional.py:5-146 of the reference):
sum of member functionals plus the Ewald ion-ion energy of the PSEUDO
member, evaluated in one fused HIP call."""
import numpy as np

from . import AbstractFunctional, FunctionalOutput, evaluate


class TotalFunctional(AbstractFunctional):
    def __init__(self, ewald=None, **kwargs):
        self.funcDict = {k: v for k, v in kwargs.items() if v is not None}
        self._ewald = ewald
        self.UpdateNameType()

    def UpdateNameType(self):
        self.name = ""
        self.type = ""
        self.terms = 0
        for key, f in self.funcDict.items():
            if not isinstance(f, AbstractFunctional):
                raise TypeError("{} must be AbstractFunctional".format(key))
            setattr(self, key, f)
            self.name += f.name + " "
            self.type += f.type + " "
            self.terms |= f.terms

    def UpdateFunctional(self, keysToRemove=[], newFuncDict={}):
        for k in keysToRemove:
            self.funcDict.pop(k, None)
        self.funcDict.update(newFuncDict)
        self.UpdateNameType()

    def Subset(self, keys):
        return TotalFunctional(**{k: self.funcDict[k] for k in keys})

    def __getitem__(self, key):
        return self.funcDict[key]

    def __setitem__(self, key, value):
        self.funcDict[key] = value
        self.UpdateNameType()

    def __repr__(self):
        return repr(self.funcDict)

    @property
    def ewald(self):
        if self._ewald is not None:
            return self._ewald
        for f in self.funcDict.values():
            if f.type == "PSEUDO":
                return f.get_ewald()
        return None

    def eion(self):
        ew = self.ewald
        return ew.energy if ew is not None else 0.0

    def compute(self, rho, calcType={"E", "V"}, split=False, **kwargs):
        if split:
            return self.get_energy_potential(rho, calcType, split=True)
        eion = self.eion() if "E" in set(calcType) else 0.0
        return evaluate(list(self.funcDict.values()), rho, calcType, name="TOTAL", eion=eion)

    def get_energy_potential(self, rho, calcType={"E", "V"}, split=False, **kwargs):
        """Without split: one fused evaluation of all members.  With split
        (total_functional.py:113-146): a dict with 'KEDF-TF', 'KEDF-VW',
        'KEDF-NL', the other members by type, 'II' (Ewald energy) and
        'TOTAL', one evaluation per part."""
        if not split:
            return self.compute(rho, calcType)
        out = {"TOTAL": FunctionalOutput(name="TOTAL", energy=0.0)}
        for func in self.funcDict.values():
            if func.type == "KEDF":
                for k2, r in func.compute(rho, calcType, split=True).items():
                    key = "KEDF-" + k2.split("-")[-1]
                    out[key] = r + out.get(key)
                    out["TOTAL"] += r
            else:
                r = func(rho, calcType)
                out[func.type] = r
                out["TOTAL"] += r
        ew = self.ewald
        if ew is not None:
            ii = FunctionalOutput(name="II", energy=ew.energy)
            out["II"] = ii
            out["TOTAL"] += ii
        return out

    def Energy(self, rho, **kwargs):
        return self.compute(rho, calcType={"E"}).energy

    def get_forces(self, rho, ions=None, split=False, **kwargs):
        """Forces of the members that have them (the local PP) plus the Ewald
        forces (total_functional.py:148-174), nat x 3 in Ha/bohr; split: dict
        by member key with 'II' and 'TOTAL'."""
        ew = self.ewald
        if ions is None:
            if ew is None:
                raise ValueError("Please provide ions for forces")
            ions = ew.ions
        forces = {"TOTAL": np.zeros((ions.nat, 3))}
        for key, func in self.funcDict.items():
            if hasattr(func, "forces"):
                f = func.forces(rho)
                forces[key] = f
                forces["TOTAL"] += f
        if ew is not None:
            forces["II"] = ew.forces
            forces["TOTAL"] += forces["II"]
        return forces if split else forces["TOTAL"]

    def get_stress(self, rho, split=False, **kwargs):
        """Stress tensor (total_functional.py:176-209): KEDF parts as
        'KEDF-TF', 'KEDF-VW', 'KEDF-NL', other members by type, Ewald as 'II'."""
        stress = {"TOTAL": np.zeros((3, 3))}
        for key, func in self.funcDict.items():
            if func.type == "KEDF":
                for k2, v in func.stress(rho, split=True).items():
                    stress["TOTAL"] += v
                    k = "KEDF-" + k2.split("-")[-1]
                    stress[k] = stress.get(k, 0.0) + v
            elif hasattr(func, "stress"):
                v = func.stress(rho)
                stress["TOTAL"] += v
                stress[func.type] = v
        ew = self.ewald
        if ew is not None:
            stress["II"] = ew.stress
            stress["TOTAL"] += stress["II"]
        return stress if split else stress["TOTAL"]

