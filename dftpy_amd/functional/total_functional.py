"""TotalFunctional (functional/total_functional.py:5-146 of the reference):
sum of member functionals plus the Ewald ion-ion energy of the PSEUDO
member, evaluated in one fused HIP call."""
from . import AbstractFunctional, evaluate


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
            raise NotImplementedError("split output is not supported")
        eion = self.eion() if "E" in set(calcType) else 0.0
        return evaluate(list(self.funcDict.values()), rho, calcType, name="TOTAL", eion=eion)

    def get_energy_potential(self, rho, calcType={"E", "V"}, **kwargs):
        return self.compute(rho, calcType, **kwargs)

    def Energy(self, rho, **kwargs):
        return self.compute(rho, calcType={"E"}).energy
