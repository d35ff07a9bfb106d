"""INI configuration and the driver's set-up stage on the CPU
(config/config.py:115-149 and interface.py:21-147 of the reference): typed
parsing, unit conversions, the grid from the spacing, the functionals and
the normalised initial density.  The optimisation itself is a GPU test."""
import numpy as np
import pytest
import torch

from conftest import DATA, data
from dftpy_amd import config as C
from dftpy_amd.constants import Units

# the settings of examples/ofdft/config.ini of the reference, with this
# repository's compressed data files
INI = """
[JOB]
task = Optdensity
calctype = Both Force Stress

[PATH]
pppath = {path}
cellpath = {path}

[MATH]
linearie = T
linearii = T

[PP]
ga = Ga_lda.oe04.recpot.gz
As = As_lda.oe04.recpot.gz

[CELL]
cellfile = GaAs_random.vasp.gz
elename = Ga As
format = vasp

[GRID]
spacing = 0.25
gfull = F

[EXC]
xc = LDA

[KEDF]
kedf = WT
x = 1.0
y = 1.0
alpha = 5.0/6.0
beta  = 5.0/6.0  ! trailing comment

[OPT]
method = TN
c1 = 1e-4
c2 = 2e-1
maxls = 10
maxiter = 200
maxfun = 50
econv = 1e-6
h0 = 0.025
"""


@pytest.fixture
def ini(tmp_path):
    p = tmp_path / "config.ini"
    p.write_text(INI.format(path=DATA))
    return p


def test_defaults_match_reference():
    d = C.DefaultOption()
    assert d["OPT"]["method"] == "CG-HS" and d["OPT"]["maxiter"] == 100 and d["OPT"]["econv"] == 1e-6
    assert d["GRID"]["ecut"] == 600 and d["GRID"]["gfull"] is False and d["KEDF"]["kedf"] == "WT"
    assert d["DENSITY"]["densityini"] == "heg" and d["MATH"]["linearie"] is True and d["JOB"]["task"] == "Optdensity"


class _no_warning:
    def __enter__(self):
        import warnings
        self._w = warnings.catch_warnings(record=True)
        self.rec = self._w.__enter__()
        warnings.simplefilter("always")
        return self

    def __exit__(self, *a):
        self._w.__exit__(*a)
        assert not self.rec, [str(w.message) for w in self.rec]


def test_read_conf_types(ini):
    with _no_warning():
        conf = C.ReadConf(ini)
    assert conf["KEDF"]["alpha"] == pytest.approx(5.0 / 6.0, rel=1e-15)
    assert conf["KEDF"]["beta"] == pytest.approx(5.0 / 6.0, rel=1e-15)
    assert conf["CELL"]["elename"] == ["Ga", "As"] and conf["JOB"]["calctype"] == ["Both", "Force", "Stress"]
    assert set(conf["PP"]) == {"Ga", "As"}
    assert conf["MATH"]["linearie"] is True and conf["GRID"]["gfull"] is False
    assert conf["OPT"]["method"] == "TN" and conf["OPT"]["maxls"] == 10 and conf["OPT"]["h0"] == 0.025
    # Ecut = pi^2 / (2 h^2) with h in bohr, in eV
    h = 0.25 / Units.Bohr
    assert conf["GRID"]["ecut"] == pytest.approx(np.pi ** 2 / (2 * h * h) * Units.Ha, rel=1e-14)


def test_unknown_key_warns_and_arithmetic_is_restricted(tmp_path):
    p = tmp_path / "c.ini"
    p.write_text("[OPT]\nmethod = cg-hs\nfoo = 1\n[KEDF]\nsigma = 2**-3\n")
    with pytest.warns(UserWarning):
        conf = C.ReadConf(p)
    assert conf["OPT"]["method"] == "CG-HS" and conf["KEDF"]["sigma"] == 0.125
    for bad in ("__import__('os')", "1 +", "a*2"):
        with pytest.raises(ValueError):
            C.format_float(bad)
    assert C.format_bool("F") is False and C.format_bool("yes") is True


def test_option_format_from_dict():
    # the way examples/test/test_ase_opt.py:17-24 builds a configuration
    conf = C.DefaultOption()
    conf["PP"]["al"] = "al.lda.recpot"
    conf["JOB"]["calctype"] = "Energy Force Stress"
    conf["OPT"]["method"] = "TN"
    conf["GRID"]["spacing"] = "0.2"
    conf = C.OptionFormat(conf)
    assert conf["PP"] == {"Al": "al.lda.recpot"} and conf["JOB"]["calctype"] == ["Energy", "Force", "Stress"]
    assert conf["GRID"]["spacing"] == 0.2 and conf["GRID"]["ecut"] > 0


def test_config_parser_builds_the_run(ini):
    from dftpy_amd.interface import ConfigParser
    from dftpy_amd.math_utils import ecut2nr
    conf, others = ConfigParser(str(ini), device="cpu")
    ions, rho, ev = others["ions"], others["rho"], others["E_v_Evaluator"]
    assert ions.nat == 10 and ions.symbols.count("Ga") == 5
    assert list(rho.grid.nrR) == list(ecut2nr(lattice=ions.cell, spacing=0.25 / Units.Bohr))
    assert rho.grid.full is False and rho.data.device.type == "cpu"
    # uniform density holding the valence electrons (Ga 3, As 5)
    assert ions.get_ncharges() == 40.0
    assert float(rho.data.sum()) * rho.grid.dV == pytest.approx(40.0, rel=1e-13)
    assert float(rho.data.max()) == float(rho.data.min())
    assert set(ev.funcDict) == {"KineticEnergyFunctional", "XCFunctional", "HARTREE", "PSEUDO"}
    ke = ev.funcDict["KineticEnergyFunctional"]
    assert ke.name == "WT" and ke.options["alpha"] == pytest.approx(5 / 6)
    assert ev.funcDict["PSEUDO"].PME is True


def test_config_parser_reads_density_and_rejects_what_is_missing(ini, tmp_path):
    from dftpy_amd.formats import io
    from dftpy_amd.interface import ConfigParser
    conf = C.ReadConf(ini)
    conf["CELL"]["cellfile"] = data("GaAs_random.xsf")
    conf["CELL"]["format"] = None
    _, others = ConfigParser(conf, device="cpu")
    # the cell file carries a density and no GRID.nr is given: the file's
    # grid and density are the start, renormalised to 40 e
    src = io.read_density(data("GaAs_random.xsf"), device="cpu")
    r = others["rho"]
    scale = 40.0 / (float(src.data.sum()) * src.grid.dV)
    assert torch.allclose(r.data, src.data * scale, rtol=1e-14, atol=0)
    assert list(r.grid.nrR) == [16, 26, 18]
    # with GRID.nr the file's density is not used (interface.py:42-50)
    conf["GRID"]["nr"] = [16, 26, 18]
    _, others = ConfigParser(conf, device="cpu")
    assert float(others["rho"].data.max()) == float(others["rho"].data.min())
    for sec, key, val in (("MATH", "multistep", 2), ("DENSITY", "nspin", 2), ("OPT", "algorithm", "OESCF"),
                          ("DENSITY", "densityini", "random")):
        c2 = C.ReadConf(ini)
        c2[sec][key] = val
        with pytest.raises(NotImplementedError):
            ConfigParser(c2, device="cpu")
