"""Start densities on the GPU (density/density.py) against the oracle."""
import numpy as np
import pytest
import torch

from oracle import density as OD
from oracle import grid as OG

pytestmark = pytest.mark.gpu


def _c1():
    from dftpy_amd import DirectGrid, systems
    ions, nr = systems.c1_system()
    ions.set_charges({"Al": 3.0})  # what the Al pseudopotential sets
    return ions, DirectGrid(ions.cell, nr, full=False), OG.Grid(ions.cell, nr)


def test_gaussian_atoms_match_oracle():
    """gen_gaussian_density on the 4-atom cell (a = 7.60, 32^3): rcut = 5 bohr
    exceeds half the cell, so each box reaches its end planes twice."""
    from dftpy_amd import density as D
    ions, grid, og = _c1()
    opts = {"Al": {"rcut": 5.0, "sigma": 0.3, "scale": 3.0}}
    print("gaussian atoms", tuple(grid.nrR), flush=True)
    rho, nc = D.gen_gaussian_density(ions, grid, opts)
    ref, nref = OD.gen_gaussian_density(og, list(ions.symbols), ions.get_scaled_positions(), opts)
    assert nc == nref == 12.0
    assert np.max(np.abs(rho.numpy() - ref)) < 1e-12 * ref.max()
    # one more Gaussian added to that density at a cartesian position
    p = np.array([1.0, 2.5, 3.7])
    rho2 = D.build_pseudo_density(p, grid, scale=0.5, sigma=0.2, rcut=2.0, density=rho)
    ref2 = OD.build_pseudo_density(og, p @ np.linalg.inv(og.lattice), 0.5, 0.2, 2.0, ref.copy())
    assert np.max(np.abs(rho2.numpy() - ref2)) < 1e-12 * ref2.max()


def _radial(nonuniform):
    r = np.linspace(0.0, 1.0, 301) ** 2 * 6.0 if nonuniform else np.linspace(0.0, 6.0, 601)
    return r, 3.0 * np.exp(-r * r / 0.7) * (1.0 + 0.3 * r)


@pytest.mark.parametrize("nonuniform", [False, True])
def test_real_space_superposition_matches_oracle(nonuniform):
    """DensityGenerator(direct=True).guess_rho: splines of r, rho(r) within
    (r[0], min(max r, L/2)), scaled to the electron count."""
    from dftpy_amd import density as D
    ions, grid, og = _c1()
    r, v = _radial(nonuniform)
    r = r + 1e-3  # r[0] > 0: the lower cut is active
    print("real-space superposition", nonuniform, flush=True)
    rho = D.DensityGenerator(r={"Al": r}, arho={"Al": v}, direct=True).guess_rho(ions, grid)
    ref = OD.guess_rho_atom(og, list(ions.symbols), ions.get_scaled_positions(), {"Al": r}, {"Al": v},
                            ions.get_ncharges())
    assert np.max(np.abs(rho.numpy() - ref)) < 1e-12 * ref.max()
    assert np.isclose(rho.integral(), ions.get_ncharges(), rtol=1e-12)


@pytest.mark.parametrize("pme", [True, False])
@pytest.mark.parametrize("nonuniform", [False, True])
def test_recipe_matches_oracle(pme, nonuniform):
    """get_3d_value_recipe (structure factors): uniform tables are the
    reference's spline itself; non-uniform ones are resampled on 8x as many
    uniform points, within 1e-7 of the oracle."""
    from dftpy_amd import density as D
    ions, grid, og = _c1()
    q, v = _radial(nonuniform)
    q = q[q <= 5.0] + 0.02  # table starts above G = 0: zero there, as splev(ext=1)
    v = v[: len(q)]
    print("recipe", pme, nonuniform, flush=True)
    rho = D.get_3d_value_recipe({"Al": q}, {"Al": v}, ions, grid, dtol=1e-30, pme=pme)
    ref = OD.get_3d_value_recipe(og, list(ions.symbols), ions.positions, ions.get_scaled_positions(), {"Al": q},
                                 {"Al": v}, ions.get_ncharges(), dtol=1e-30, pme=pme)
    tol = 1e-7 if nonuniform else 1e-11
    assert np.max(np.abs(rho.numpy() - ref)) < tol * np.abs(ref).max()


def test_generator_list_file_and_fallbacks(tmp_path):
    """A .list file is read as two columns and placed like the dict form; a
    generator over recpot files (no atomic density) gives the uniform gas."""
    from dftpy_amd import density as D, systems
    from dftpy_amd.functional import Functional
    ions, grid, og = _c1()
    q, v = _radial(False)
    f = tmp_path / "al.list"
    np.savetxt(f, np.stack([q, v], axis=1))
    a = D.DensityGenerator(files={"Al": str(f)}).guess_rho(ions, grid)
    b = D.get_3d_value_recipe({"Al": q}, {"Al": v}, ions, grid, dtol=1e-30)
    assert torch.allclose(a.data, b.data, rtol=1e-13, atol=0.0)
    pp = Functional(type="PSEUDO", grid=grid, ions=ions, PP_list={"Al": systems.empty_core_recpot()})
    u = D.DensityGenerator(pseudo=pp).guess_rho(ions, grid)
    assert np.allclose(u.numpy(), ions.get_ncharges() / grid.volume, rtol=1e-12)
    h = D.guess_rho_heg(ions, grid)
    assert np.allclose(h.numpy(), ions.get_ncharges() / grid.volume, rtol=1e-15)
