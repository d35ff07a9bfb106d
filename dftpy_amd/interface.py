"""Configuration-driven runs (interface.py:21-245 of the reference):
``ConfigParser`` builds the grid, the functionals and the initial density
from a configuration, ``OptimizeDensityConf`` optimises the density and
computes the requested energies, forces and stress, ``evaluator2results``
prints and collects them.  ``RunJob`` chains the two for ``task =
Optdensity`` or ``Calculation``.

Every computation runs through libofdft_hip on the GPU; this module only
wires the pieces together.  Configuration features outside the GPU path
(multi-step grids, spin, OESCF, complex grids)
raise NotImplementedError.
"""
import os
import time

import numpy as np
import torch

from .config import ReadConf
from .constants import Units
from .density import DensityGenerator
from .field import DirectField
from .formats import io
from .functional import Functional, FunctionalOutput
from .functional.total_functional import TotalFunctional
from .grid import DirectGrid
from .math_utils import ecut2nr
from .optimization import Optimization

KEDF_KEYS = ("x", "y", "alpha", "beta", "sigma", "rho0")
OPT_KEYS = ("econv", "maxiter", "maxfun", "maxls", "c1", "c2", "xtol", "algorithm", "vector", "h0")
HA2EV = Units.Ha
# Ha/bohr^3 -> eV/A^3 -> GPa (1 eV/A^3 = e * 1e21 GPa with e in coulomb)
HA_BOHR3_TO_GPA = Units.Ha / Units.Bohr ** 3 * Units._e * 1.0e21


def _path(base, name):
    return name if os.path.isabs(str(name)) else os.path.join(str(base), str(name))


def _unsupported(config):
    if config["MATH"]["multistep"] > 1:
        raise NotImplementedError("multistep / twostep grids are not implemented")
    if config["DENSITY"]["nspin"] != 1:
        raise NotImplementedError("spin-polarised densities are not implemented")
    if config["GRID"]["cplx"]:
        raise NotImplementedError("complex grids are not implemented")
    if config["OPT"]["algorithm"] not in ("EMM",):
        raise NotImplementedError("OPT.algorithm %s is not implemented (EMM is)" % config["OPT"]["algorithm"])


def ConfigParser(config, ions=None, rhoini=None, pseudo=None, grid=None, dtype=torch.float64, device=None):
    """(config, others): others holds 'ions', 'field'/'rho' (initial density
    normalised to the ions' electron count), 'E_v_Evaluator' and 'nr2'."""
    if isinstance(config, (str, os.PathLike)):
        config = ReadConf(config)
    _unsupported(config)
    field = None
    if ions is None:
        ions, field, _ = io.read(_path(config["PATH"]["cellpath"], config["CELL"]["cellfile"]),
                                 format=config["CELL"]["format"], names=config["CELL"]["elename"], kind="all",
                                 dtype=dtype, device=device)
    lattice = ions.cell
    g = config["GRID"]
    if g["nr"] is not None:
        nr = np.asarray(g["nr"], dtype=np.int64)
    elif field is not None:
        nr = np.asarray(field.grid.nrR)
        rhoini = field
    else:
        nr = ecut2nr(lattice=lattice, spacing=g["spacing"] / Units.Bohr, max_prime=g["maxprime"], scale=g["scale"])
    print("The final grid size is ", nr)
    if grid is None:
        grid = DirectGrid(lattice, nr, full=g["gfull"], dtype=dtype, device=device)
    pp_list = {k.capitalize(): _path(config["PATH"]["pppath"], v) for k, v in config["PP"].items()}
    if pseudo is None:
        pseudo = Functional(type="PSEUDO", grid=grid, ions=ions, PP_list=pp_list, PME=config["MATH"]["linearie"])
    else:
        pseudo.restart(grid=grid, ions=ions)
    kc = config["KEDF"]
    ke = Functional(type="KEDF", name=kc["kedf"], **{k: kc[k] for k in KEDF_KEYS if kc.get(k) is not None})
    xc = Functional(type="XC", xc=config["EXC"]["xc"], libxc=config["EXC"]["libxc"])
    evaluator = TotalFunctional(KineticEnergyFunctional=ke, XCFunctional=xc, HARTREE=Functional(type="HARTREE"),
                                PSEUDO=pseudo)
    nelec = ions.get_ncharges()
    ini = config["DENSITY"]["densityini"]
    if rhoini is not None:
        src = rhoini
    elif ini == "heg":
        src = None
    elif ini == "read":
        src = io.read_density(config["DENSITY"]["densityfile"], dtype=dtype, device=device)
    elif ini == "atomic":
        # interface.py:123-125; recpot files carry no atomic density, so this is
        # the uniform gas unless the reader provides one
        src = DensityGenerator(pseudo=pseudo).guess_rho(ions, grid=grid)
    else:
        raise NotImplementedError("DENSITY.densityini = %s is not implemented (heg, read, atomic are)" % ini)
    if src is None:
        rho = DirectField(grid, data=torch.full(tuple(grid.nrR), nelec / grid.volume, dtype=grid.dtype,
                                                device=grid.device))
    else:
        if tuple(src.grid.nrR) != tuple(grid.nrR):
            raise ValueError("the initial density has grid %s, the run uses %s" % (tuple(src.grid.nrR),
                                                                                  tuple(grid.nrR)))
        rho = DirectField(grid, data=src.data.to(dtype=grid.dtype, device=grid.device).clone())
    # normalise the electron count (the cell may have changed)
    rho.data *= nelec / (float(rho.data.sum()) * grid.dV)
    others = {"ions": ions, "field": rho, "rho": rho, "E_v_Evaluator": evaluator, "nr2": nr.copy()}
    return config, others


def _calc_types(config):
    calc = set()
    ct = config["JOB"]["calctype"]
    ct = ct if isinstance(ct, (list, tuple)) else [ct]
    for word, flags in (("Both", "EV"), ("Potential", "V"), ("Energy", "E"), ("Force", "F"), ("Stress", "S")):
        if word in ct:
            calc.update(flags)
    return calc


def OptimizeDensityConf(config, ions=None, rho=None, E_v_Evaluator=None, nr2=None, **kwargs):
    """Optimise the density (task Optdensity), write the outputs the
    configuration names and return evaluator2results' dict."""
    _unsupported(config)
    task = config["JOB"]["task"]
    t0 = time.time()
    opt = None
    if "Optdensity" in task:
        oo = {k: config["OPT"][k] for k in OPT_KEYS}
        oo["econv"] = oo["econv"] * ions.nat
        opt = Optimization(EnergyEvaluator=E_v_Evaluator, optimization_options=oo,
                           optimization_method=config["OPT"]["method"])
        rho = opt.optimize_rho(guess_rho=rho)
    elif task != "Calculation":
        raise NotImplementedError("JOB.task %s is not implemented (Optdensity, Calculation are)" % task)
    print("-" * 80)
    if config["DENSITY"]["densityoutput"]:
        print("Write Density...")
        io.write(config["DENSITY"]["densityoutput"], rho, ions)
    if config["OUTPUT"]["electrostatic_potential"]:
        print("Write electrostatic potential...")
        v = get_electrostatic_potential(rho, E_v_Evaluator)
        io.write(config["OUTPUT"]["electrostatic_potential"], v, ions, data_type="potential")
    results = evaluator2results(E_v_Evaluator, rho=rho, calcType=_calc_types(config), ions=ions,
                                print_stress=config["OUTPUT"]["stress"])
    if opt is not None:
        results["optimization"] = opt.stats
    results["seconds"] = time.time() - t0
    return results


def get_electrostatic_potential(rho, evaluator):
    """Local pseudopotential plus Hartree potential
    (properties/potential.py:3-7)."""
    return (evaluator.funcDict["PSEUDO"](rho, calcType={"V"}).potential +
            evaluator.funcDict["HARTREE"](rho, calcType={"V"}).potential)


def evaluator2results(evaluator, rho=None, calcType={"E", "F"}, ions=None, print_stress=False, split=True, **kwargs):
    """Energies (split by part), forces and stress of ``evaluator`` at
    ``rho``, printed like interface.py:245-330 and returned as a dict with
    'density', 'energypotential', 'forces', 'stress', 'evaluator', 'pseudo'."""
    forces, stress = {}, {}
    ep = {"TOTAL": FunctionalOutput(name="TOTAL", energy=0.0)}
    calc_ev = {c for c in calcType if c in ("E", "V")}
    if calc_ev:
        print("Calculate Energy/Potential...")
        ep = evaluator.get_energy_potential(rho, calcType=calc_ev, split=split)
        if not split:
            ep = {"TOTAL": ep}
    if "E" in calcType:
        print(format("Energy information", "-^80"))
        ke = 0.0
        for key in sorted(ep):
            if key == "TOTAL":
                continue
            print("{:>10s} energy (eV): {:22.15E}".format(key, ep[key].energy * HA2EV))
            if key.startswith("KEDF"):
                ke += ep[key].energy
        etot = ep["TOTAL"].energy
        print("{:>10s} energy (eV): {:22.15E}".format("TOTAL", etot * HA2EV))
        print("-" * 80)
        fstr = "  {:<30s} : {:30.15f}"
        print(fstr.format("kedfs energy (a.u.)", ke))
        print(fstr.format("kedfs energy (eV)", ke * HA2EV))
        print(fstr.format("total energy (a.u.)", etot))
        print(fstr.format("total energy (eV)", etot * HA2EV))
        if ions is not None:
            print(fstr.format("total energy (eV/atom)", etot * HA2EV / ions.nat))
    if "F" in calcType:
        print("Calculate Force...")
        forces = evaluator.get_forces(rho, split=split)
        f = np.abs(forces["TOTAL"] if split else forces)
        for name, val in (("Max", f.max()), ("Min", f.min()), ("Ave", f.mean())):
            print(" " * 8 + "{0:>22s} : {1:<22.5f}".format("%s force (a.u.)" % name, val))
        print("-" * 80)
    if "S" in calcType:
        print("Calculate Stress...")
        stress = evaluator.get_stress(rho, split=split)
        st = stress if split else {"TOTAL": stress}
        row = " " * 16 + "{0[0]:12.5f} {0[1]:12.5f} {0[2]:12.5f}"
        if print_stress:
            for key in sorted(st):
                if key != "TOTAL":
                    print("{:>10s} stress (a.u.): ".format(key))
                    for i in range(3):
                        print(row.format(st[key][i]))
        print("{:>10s} stress (a.u.): ".format("TOTAL"))
        for i in range(3):
            print(row.format(st["TOTAL"][i]))
        print("{:>10s} stress (GPa): ".format("TOTAL"))
        for i in range(3):
            print(row.format(st["TOTAL"][i] * HA_BOHR3_TO_GPA))
        print("-" * 80)
    results = {"density": rho, "energypotential": ep, "forces": forces, "stress": stress, "evaluator": evaluator}
    if "PSEUDO" in getattr(evaluator, "funcDict", {}):
        results["pseudo"] = evaluator.funcDict["PSEUDO"]
    return results


def RunJob(config, **kwargs):
    """Read (if a path) and run a configuration: ConfigParser then
    OptimizeDensityConf."""
    config, others = ConfigParser(config, **kwargs)
    return OptimizeDensityConf(config, **others)
