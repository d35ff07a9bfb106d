"""INI configuration of a run (config/config.py:1-149, config_entry.py:1-192
and the defaults of config/configentries.json of the reference), for the
sections the GPU path runs: JOB, PATH, MATH, PP, CELL, GRID, DENSITY, EXC,
KEDF, OPT, OUTPUT.

``DefaultOption()`` gives every key at its default; ``ReadConf(file)`` reads
an INI file over it and ``OptionFormat(conf)`` turns string values into
typed ones, as the reference does.  Numbers accept arithmetic
(``alpha = 5.0/6.0``), evaluated by a restricted parser, not ``eval``.
Keys of the reference that this package does not act on are accepted with
a warning.
"""
import ast
import configparser
import copy
import operator
import os
import re
import warnings

import numpy as np

from .constants import Units

# section -> key -> (type, default); names, types and defaults of the reference
ENTRIES = {
    "JOB": {"task": ("cstr", "Optdensity"), "calctype": ("cstrlist", "Energy")},
    "PATH": {"pppath": ("path", "./"), "cellpath": ("path", "./")},
    "MATH": {"linearii": ("bool", True), "linearie": ("bool", True), "twostep": ("bool", False),
             "multistep": ("int", 1), "reuse": ("bool", True)},
    "PP": {},
    "CELL": {"cellfile": ("str", "POSCAR"), "elename": ("cstrlist", None), "zval": ("floatlist", None),
             "format": ("str", None)},
    "GRID": {"ecut": ("float", 600), "spacing": ("float", None), "gfull": ("bool", False), "nr": ("intlist", None),
             "maxprime": ("int", 13), "scale": ("float", 0.99), "cplx": ("bool", False)},
    "DENSITY": {"nspin": ("int", 1), "magmom": ("float", 0), "densityini": ("lstr", "heg"),
                "densityfile": ("path", None), "densityoutput": ("path", None)},
    "EXC": {"xc": ("str", None), "libxc": ("lstrlist", "lda_x lda_c_pz"), "x_str": ("str", None),
            "c_str": ("str", None)},
    "KEDF": {"kedf": ("ustr", "WT"), "x": ("float", 1), "y": ("float", 1), "alpha": ("float", None),
             "beta": ("float", None), "sigma": ("float", None), "nsp": ("int", None), "interp": ("str", "hermite"),
             "kerneltype": ("str", None), "symmetrization": ("str", None), "lumpfactor": ("floatlist", None),
             "neta": ("int", 50000), "etamax": ("float", 50), "order": ("int", 3), "ratio": ("float", 1.15),
             "maxpoints": ("int", 1000), "delta": ("float", None), "kdd": ("int", 3), "rho0": ("float", None),
             "k_str": ("str", None), "params": ("floatlist", None), "kfmin": ("float", None),
             "kfmax": ("float", None), "rhomax": ("float", None), "ldw": ("float", None),
             "temperature": ("float", None), "temperature0": ("float", None)},
    "OPT": {"method": ("ustr", "CG-HS"), "algorithm": ("ustr", "EMM"), "vector": ("str", "Orthogonalization"),
            "c1": ("float", 1e-4), "c2": ("float", 0.2), "maxls": ("int", 10), "econv": ("float", 1e-6),
            "maxfun": ("int", 50), "maxiter": ("int", 100), "xtol": ("float", 1e-12), "h0": ("float", 1.0)},
    "OUTPUT": {"time": ("bool", True), "stress": ("bool", True), "electrostatic_potential": ("str", None)},
}

_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.Pow: operator.pow}
_UNOPS = {ast.UAdd: operator.pos, ast.USub: operator.neg}


def _arith(node):
    if isinstance(node, ast.Expression):
        return _arith(node.body)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return node.value
    if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
        return _BINOPS[type(node.op)](_arith(node.left), _arith(node.right))
    if isinstance(node, ast.UnaryOp) and type(node.op) in _UNOPS:
        return _UNOPS[type(node.op)](_arith(node.operand))
    raise ValueError("not an arithmetic expression")


def format_float(s):
    """A number or an arithmetic expression of numbers (+ - * / **)."""
    s = s.strip()
    try:
        return float(s)
    except ValueError:
        pass
    try:
        return float(_arith(ast.parse(s, mode="eval")))
    except (SyntaxError, ValueError, ZeroDivisionError) as e:
        raise ValueError("cannot read %r as a number" % s) from e


def format_bool(s):
    return s.strip().lower()[:1] not in ("n", "f", "0")


def format_path(s):
    return re.sub(r"\$\{(.*?)\}", lambda m: os.environ.get(m.group(1), ""), s.strip())


_FORMAT = {
    "bool": format_bool,
    "int": lambda s: int(s.strip()),
    "float": format_float,
    "str": lambda s: s.strip(),
    "lstr": lambda s: s.strip().lower(),
    "ustr": lambda s: s.strip().upper(),
    "cstr": lambda s: s.strip().capitalize(),
    "path": format_path,
    "intlist": lambda s: [int(x) for x in s.split()],
    "floatlist": lambda s: [format_float(x) for x in s.split()],
    "cstrlist": lambda s: s.title().split(),
    "lstrlist": lambda s: s.lower().split(),
}


def DefaultOption():
    """Every section and key at its default value (config/config.py:38-58)."""
    return {sec: {k: copy.deepcopy(d) for k, (t, d) in keys.items()} for sec, keys in ENTRIES.items()}


def _format_value(section, key, value):
    if not isinstance(value, str):
        return value
    typ = ENTRIES.get(section, {}).get(key, ("str", None))[0]
    value = re.split("[#!]", value)[0]
    if typ in ("str", "path", "cstrlist", "lstrlist", "floatlist", "intlist") and not value.strip():
        return None
    return _FORMAT[typ](value)


def ConfSpecialFormat(conf):
    """Grid spacing (A) <-> ecut (eV), Ecut = pi^2 / (2 h^2)
    (config/config.py:61-98); twostep means multistep = 2."""
    g = conf["GRID"]
    if g.get("spacing"):
        g["ecut"] = np.pi ** 2 / (2 * g["spacing"] ** 2) * Units.Ha * Units.Bohr ** 2
    else:
        g["spacing"] = np.sqrt(np.pi ** 2 / g["ecut"] * 0.5 * Units.Ha) * Units.Bohr
    lump = conf["KEDF"].get("lumpfactor")
    if lump and len(lump) == 1:
        conf["KEDF"]["lumpfactor"] = lump[0]
    if conf["MATH"]["twostep"] and conf["MATH"]["multistep"] == 1:
        conf["MATH"]["multistep"] = 2
    return conf


def OptionFormat(config):
    """Typed copy of a configuration whose values may be strings
    (config/config.py:131-149); PP keys are capitalised element names."""
    conf = {}
    for section, keys in config.items():
        conf[section] = {}
        for key, value in keys.items():
            if section == "PP":
                conf["PP"][key.capitalize()] = value
            else:
                conf[section][key] = _format_value(section, key, value)
    for section, keys in ENTRIES.items():
        conf.setdefault(section, {})
        for key, (t, d) in keys.items():
            conf[section].setdefault(key, copy.deepcopy(d))
    return ConfSpecialFormat(conf)


def ReadConf(infile):
    """Read an INI file over the defaults (config/config.py:115-129)."""
    if not os.path.isfile(str(infile)):
        raise FileNotFoundError(infile)
    parser = configparser.ConfigParser()
    parser.read(str(infile))
    conf = DefaultOption()
    for section in parser.sections():
        if section not in conf:
            warnings.warn("section [%s] is not used by this package" % section)
            conf[section] = {}
        for key in parser.options(section):
            if section == "PP":
                conf["PP"][key.capitalize()] = parser.get(section, key)
                continue
            if section in ENTRIES and key not in ENTRIES[section]:
                warnings.warn('"%s.%s" is not a known option' % (section, key))
            conf[section][key] = parser.get(section, key)
    return OptionFormat(conf)


def PrintConf(conf):
    import json
    text = json.dumps(conf, indent=4, sort_keys=True, default=str)
    print(text)
    return text
