"""Functionals with the reference's interface (functional/__init__.py:1-50,
abstract_functional.py:12-28, functional_output.py:8-131).

Every functional is a callable ``f(rho, calcType={'E','V'}) ->
FunctionalOutput``.  All of them evaluate through one fused HIP pipeline
(``ofdft_eval``) restricted to their own terms; a TotalFunctional runs the
union of its members' terms in a single call.
"""
import copy

import numpy as np
import torch

from .. import _lib
from ..field import DirectField

__all__ = ["FunctionalOutput", "AbstractFunctional", "KEDF", "XC", "Hartree", "LocalPseudo", "TotalFunctional",
           "Functional", "ZeroFunctional"]


class FunctionalOutput:
    """Energy (float, Ha) and potential (DirectField) of one evaluation
    (functional_output.py:8-108)."""

    def __init__(self, name=None, energy=None, potential=None, energydensity=None, **kwargs):
        if name is None:
            raise AttributeError("Functional name must be specified")
        self.name = name
        if energy is not None:
            self.energy = energy
        if potential is not None:
            self.potential = potential
        if energydensity is not None:
            self.energydensity = energydensity
        if kwargs.get("v2rho2") is not None:
            self.v2rho2 = kwargs["v2rho2"]
        self.parts = kwargs.get("parts")

    def _items(self):
        for k in ("energy", "potential", "energydensity", "v2rho2"):
            if hasattr(self, k):
                yield k, getattr(self, k)

    def __add__(self, other):
        if other is None:
            return self.copy()
        res = FunctionalOutput(name=self.name if self.name == other.name else self.name + other.name)
        for k, v in self._items():
            setattr(res, k, v + getattr(other, k) if hasattr(other, k) else v)
        for k, v in other._items():
            if not hasattr(res, k):
                setattr(res, k, v)
        return res

    def __sub__(self, other):
        res = self.copy()
        for k, v in self._items():
            setattr(res, k, v - getattr(other, k) if hasattr(other, k) else v)
        for k, v in other._items():
            if not hasattr(self, k):
                setattr(res, k, -v)
        return res

    def __mul__(self, x):
        res = self.copy()
        for k, v in self._items():
            setattr(res, k, v * x)
        return res

    def __truediv__(self, x):
        if x == 0:
            raise ValueError("Dividing zero")
        res = self.copy()
        for k, v in self._items():
            setattr(res, k, v / x)
        return res

    def copy(self):
        res = FunctionalOutput(name=self.name)
        for k, v in self._items():
            setattr(res, k, v.copy() if hasattr(v, "copy") else copy.copy(v))
        return res


class AbstractFunctional:
    type = ""
    name = ""
    terms = 0

    def __call__(self, rho, *args, **kwargs):
        return self.compute(rho, *args, **kwargs)

    def kedf_options(self):
        return None

    def compute(self, rho, calcType={"E", "V"}, **kwargs):
        return evaluate([self], rho, calcType, name=self.name)

    def __repr__(self):
        return self.name


def _energy_of(terms, e):
    s = 0.0
    for bit, idx in ((_lib.TERM_TF, _lib.E_TF), (_lib.TERM_VW, _lib.E_VW), (_lib.TERM_WT | _lib.TERM_LMGP, _lib.E_NL),
                     (_lib.TERM_XC, _lib.E_XC), (_lib.TERM_HARTREE, _lib.E_H), (_lib.TERM_PP, _lib.E_PP)):
        if terms & bit:
            s += e[idx]
    return s


def configure(plan, funcs, rho_grid, eion=0.0):
    """Push the members' settings into the plan; returns the term mask and
    the keep-alive vloc tensor."""
    terms = 0
    kedf = dict(x=1.0, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0, rho0=None, sigma=None)
    vloc = None
    core = None
    for f in funcs:
        terms |= f.terms
        o = f.kedf_options()
        if o:
            kedf.update({k: v for k, v in o.items() if k in kedf})
        if f.terms & _lib.TERM_PP:
            vloc = f.vreal_for(rho_grid)
        if f.terms & _lib.TERM_XC and getattr(f, "core_density", None) is not None:
            core = f.core_for(rho_grid)
    rho0 = kedf["rho0"]
    plan.kedf_settings = {k: kedf[k] for k in ("x", "y", "alpha", "beta")}
    plan.has_core = core is not None
    # the plan keeps its settings between calls: skip the library calls when
    # nothing changed (vloc and the core density are read through their
    # pointers at evaluation time)
    vptr = vloc.data.data_ptr() if vloc is not None else None
    cptr = core.data_ptr() if core is not None else None
    keep = (vloc, core)
    key = (terms, tuple((k, kedf[k]) for k in sorted(kedf)), vptr, cptr, float(eion))
    if getattr(plan, "config_key", None) == key:
        return terms, keep
    _lib.check(_lib.lib.ofdft_set_kedf(plan.handle, float(kedf["x"]), float(kedf["y"]), float(kedf["alpha"]),
                                       float(kedf["beta"]), float(rho0) if rho0 else -1.0,
                                       float(kedf["sigma"]) if kedf["sigma"] else -1.0))
    if terms & _lib.TERM_LMGP and not getattr(plan, "lmgp_ready", False):
        from .lmgp import lmgp_tables
        eta, kc, dc = lmgp_tables()
        _lib.check(_lib.lib.ofdft_set_lmgp(plan.handle, int(eta.size), float(eta[-1]), _lib.dptr(kc), _lib.dptr(dc),
                                           1.15))
        plan.lmgp_ready = True
    _lib.check(_lib.lib.ofdft_set_terms(plan.handle, terms))
    _lib.check(_lib.lib.ofdft_set_vloc(plan.handle, vptr))
    _lib.check(_lib.lib.ofdft_set_core(plan.handle, cptr))
    _lib.check(_lib.lib.ofdft_set_eion(plan.handle, float(eion)))
    plan.config_key = key
    return terms, keep


def evaluate(funcs, rho, calcType, name="TOTAL", eion=0.0):
    """One fused evaluation of the union of ``funcs`` on ``rho``."""
    if not isinstance(rho, DirectField):
        raise TypeError("rho must be a DirectField")
    if rho.rank > 1:
        raise NotImplementedError("spin-polarised densities are not supported")
    calc = set(calcType)
    if calc - {"E", "V", "D", "V2"}:
        raise NotImplementedError("calcType %s not supported (only 'E', 'V', 'D' and 'V2')" %
                                  sorted(calc - {"E", "V", "D", "V2"}))
    grid = rho.grid
    plan = grid.plan(grid.dtype)
    terms, vloc = configure(plan, funcs, grid, eion)
    flags = (_lib.CALC_E if "E" in calc else 0) | (_lib.CALC_V if "V" in calc else 0)
    src = rho.data.contiguous()
    pot = torch.empty_like(src) if "V" in calc else None
    en = torch.empty(_lib.E_COUNT, dtype=torch.float64, device=src.device) if "E" in calc else None
    # energy density (kedf/tf.py:68-73, vw.py:125-126, wt.py:123-129) and
    # second derivative (tf.py:77-79, semilocal_xc.py:291-303) of the terms
    # that have them
    dens = torch.empty_like(src) if "D" in calc and terms & (_lib.TERM_TF | _lib.TERM_VW | _lib.TERM_WT |
                                                              _lib.TERM_LMGP) else None
    v2 = torch.empty_like(src) if "V2" in calc and terms & (_lib.TERM_TF | _lib.TERM_XC) else None
    if dens is not None or v2 is not None:
        _lib.check(_lib.lib.ofdft_eval_fields(plan.handle, src.data_ptr(), flags,
                                              pot.data_ptr() if pot is not None else None,
                                              en.data_ptr() if en is not None else None,
                                              dens.data_ptr() if dens is not None else None,
                                              v2.data_ptr() if v2 is not None else None, _lib.stream()))
    else:
        _lib.check(_lib.lib.ofdft_eval(plan.handle, src.data_ptr(), flags,
                                       pot.data_ptr() if pot is not None else None,
                                       en.data_ptr() if en is not None else None, _lib.stream()))
    out = FunctionalOutput(name=name)
    if dens is not None:
        out.energydensity = DirectField(grid, data=dens)
    if v2 is not None:
        out.v2rho2 = DirectField(grid, data=v2)
    if en is not None:
        e = en.cpu().numpy()
        out.energy = _energy_of(terms, e) + eion
        out.parts = dict(TF=e[_lib.E_TF], vW=e[_lib.E_VW], NL=e[_lib.E_NL], XC=e[_lib.E_XC], HARTREE=e[_lib.E_H],
                         PSEUDO=e[_lib.E_PP], II=eion, N=e[_lib.E_NELEC])
    if pot is not None:
        out.potential = DirectField(grid, data=pot)
    del vloc
    return out


def _fp64(rho):
    """rho on the FP64 grid of its shape (forces and stress run in FP64)."""
    if rho.grid.dtype == torch.float64:
        return rho
    return DirectField(rho.grid.with_dtype(torch.float64), data=rho.data.double())


def functional_stress(funcs, rho):
    """Stress tensors of the TF, vW, WT-nonlocal, LDA and Hartree terms of
    ``funcs`` at rho (ofdft_functional_stress; kedf/tf.py:52-62,
    kedf/vw.py:97-111, kedf/wt.py:75-105, xc/semilocal_xc.py:308-322,
    hartree.py:67-91).  Returns a dict of 3x3 arrays (zero for absent terms)."""
    r = _fp64(rho)
    grid = r.grid
    plan = grid.plan(torch.float64)
    terms, vloc = configure(plan, funcs, grid)
    if terms & _lib.TERM_LMGP:
        raise NotImplementedError("stress of the LMGP functional is not implemented")
    if plan.has_core:
        raise NotImplementedError("the XC stress with a core density is not implemented")
    src = r.data.contiguous()
    en = torch.empty(_lib.E_COUNT, dtype=torch.float64, device=src.device)
    _lib.check(_lib.lib.ofdft_eval(plan.handle, src.data_ptr(), _lib.CALC_E, None, en.data_ptr(), _lib.stream()))
    e = np.ascontiguousarray(en.cpu().numpy())
    out = np.zeros(45)
    _lib.check(_lib.lib.ofdft_functional_stress(plan.handle, src.data_ptr(), _lib.dptr(e), _lib.dptr(out),
                                                _lib.stream()))
    del vloc
    p = out.reshape(5, 3, 3)
    return dict(TF=p[0].copy(), VW=p[1].copy(), NL=p[2].copy(), XC=p[3].copy(), HARTREE=p[4].copy())


class ZeroFunctional(AbstractFunctional):
    name = "ZERO"

    def __init__(self, *args, **kwargs):
        pass

    def compute(self, rho, calcType={"E", "V"}, **kwargs):
        out = FunctionalOutput(name="ZERO")
        if "E" in calcType:
            out.energy = 0.0
        if "V" in calcType:
            out.potential = DirectField(rho.grid)
        return out


# KEDF name -> term bits (functional/kedf/__init__.py:27-69)
KEDF_TERMS = {
    "TF": _lib.TERM_TF, "VW": _lib.TERM_VW, "WT-NL": _lib.TERM_WT,
    "TFVW": _lib.TERM_TF | _lib.TERM_VW, "X_TF_Y_VW": _lib.TERM_TF | _lib.TERM_VW, "XTFYVW": _lib.TERM_TF | _lib.TERM_VW,
    "WT": _lib.TERM_TF | _lib.TERM_VW | _lib.TERM_WT, "NONE": 0,
    "LMGP-NL": _lib.TERM_LMGP, "LMGP": _lib.TERM_TF | _lib.TERM_VW | _lib.TERM_LMGP,
}


KEDF_PARTS = (("TF", _lib.TERM_TF), ("VW", _lib.TERM_VW), ("WT-NL", _lib.TERM_WT), ("LMGP-NL", _lib.TERM_LMGP))


class _TermSubset(AbstractFunctional):
    """Some of a KEDF's terms with the KEDF's settings (split output)."""

    def __init__(self, terms, options, name):
        self.type = "KEDF"
        self.name = name
        self.terms = terms
        self.options = options

    def kedf_options(self):
        return self.options


class KEDF(AbstractFunctional):
    """``KEDF(name='WT', **options)`` (functional/kedf/__init__.py:90-184).
    Options: x (TF scale), y (vW scale), alpha, beta, rho0, sigma."""

    def __init__(self, name="WT", kedf=None, **kwargs):
        self.type = "KEDF"
        self.name = (kedf or name).upper()
        if self.name not in KEDF_TERMS:
            raise NotImplementedError("KEDF %s is not implemented on the GPU path" % self.name)
        self.terms = KEDF_TERMS[self.name]
        self.options = dict(kwargs)
        self.options["kedf"] = self.name

    def kedf_options(self):
        return self.options

    def compute(self, rho, calcType={"E", "V"}, split=False, **kwargs):
        """split: dict of the parts keyed as kedf/__init__.py:144-172 names
        them ('TF', 'VW', 'WT-NL' or 'LMGP-NL'), one evaluation each."""
        if not split:
            return evaluate([self], rho, calcType, name="KEDF")
        parts = {}
        for key, bit in KEDF_PARTS:
            if self.terms & bit:
                parts[key] = evaluate([_TermSubset(bit, self.options, key)], rho, calcType, name=key)
        return parts

    def stress(self, density, split=False, **kwargs):
        """Sum of the stresses of the KEDF's parts (kedf/__init__.py:186-226);
        split: dict keyed TF / VW / WT-NL."""
        st = functional_stress([self], density)
        parts = {}
        if self.terms & _lib.TERM_TF:
            parts["TF"] = st["TF"]
        if self.terms & _lib.TERM_VW:
            parts["VW"] = st["VW"]
        if self.terms & _lib.TERM_WT:
            parts["WT-NL"] = st["NL"]
        if split:
            return parts
        return sum(parts.values(), np.zeros((3, 3)))


class XC(AbstractFunctional):
    """``XC(xc='LDA')``: the reference's built-in Perdew-Zunger LDA
    (functional/xc/__init__.py:7-74, semilocal_xc.py:252-290)."""

    def __init__(self, xc=None, name=None, libxc=None, core_density=None, **kwargs):
        xc = xc or name
        if libxc and not xc:
            names = [libxc] if isinstance(libxc, str) else list(libxc)
            if sorted(n.lower() for n in names) == ["lda_c_pz", "lda_x"]:
                xc = "LDA"
        if xc is None or xc.upper() != "LDA":
            raise NotImplementedError("only the PZ LDA is implemented (got %r)" % (xc or libxc))
        self.type = "XC"
        self.name = "LDA"
        self.terms = _lib.TERM_XC
        self.options = kwargs
        # non-linear core correction (xc/__init__.py:49-74): XC of rho + core
        self.core_density = core_density
        self._core = {}

    def core_for(self, grid):
        """The core density as a contiguous device tensor in the precision of
        ``grid`` (the DirectField given, or an array of the grid's shape)."""
        c = self.core_density
        data = c.data if isinstance(c, DirectField) else torch.as_tensor(np.asarray(c))
        key = (grid.dtype, data.data_ptr() if data.is_cuda else id(c))
        if key not in self._core:
            t = data.to(device=grid.device, dtype=grid.dtype).contiguous()
            if tuple(t.shape) != tuple(int(n) for n in grid.nrR):
                raise ValueError("core density shape %s does not match the grid %s" % (tuple(t.shape),
                                                                                        tuple(grid.nrR)))
            self._core = {key: t}
        return self._core[key]

    def stress(self, density, **kwargs):
        """LDAStress (xc/semilocal_xc.py:308-322)."""
        return functional_stress([self], density)["XC"]


class Hartree(AbstractFunctional):
    """Hartree functional (functional/hartree.py:11-56)."""

    def __init__(self, **kwargs):
        self.type = "HARTREE"
        self.name = "HARTREE"
        self.terms = _lib.TERM_HARTREE
        self.options = kwargs

    def stress(self, density, **kwargs):
        """HartreeFunctionalStress (hartree.py:67-91)."""
        return functional_stress([self], density)["HARTREE"]


from .pseudo import LocalPseudo  # noqa: E402
from .total_functional import TotalFunctional  # noqa: E402

FunctionalTypeDict = {"KEDF": KEDF, "PSEUDO": LocalPseudo, "HARTREE": Hartree, "XC": XC}


def Functional(type=None, optional_kwargs=None, **kwargs):
    """Factory ``Functional(type='KEDF', name='WT')`` (functional/__init__.py:22-50)."""
    type = type.upper()
    if optional_kwargs is not None:
        kwargs.update(optional_kwargs)
    try:
        cls = FunctionalTypeDict[type]
    except KeyError:
        raise AttributeError("Unknown functional type: %s" % type)
    return cls(**kwargs)
