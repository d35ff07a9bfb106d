"""File formats (CPU): round trips of every writer through its reader, the
unit conversions each format fixes, cross-format agreement on the
reference's own data files, and the io front end returning DirectFields.
Mirrors examples/test/test_snpy.py:19-36 and test_ase_pmg_io.py:11-17."""
import numpy as np
import pytest

from conftest import data
from dftpy_amd.constants import Units
from dftpy_amd.formats import io, qepp, snpy, vasp, xsf


@pytest.fixture(scope="module")
def al():
    return qepp.read_qepp_all(data("Al_fde_rho.pp"))


@pytest.fixture(scope="module")
def gaas():
    return xsf.read_xsf(data("GaAs_random.xsf"))


def _same_ions(a, b, rtol=1e-15):
    assert a.symbols == b.symbols
    assert np.allclose(a.cell, b.cell, rtol=rtol, atol=0)
    assert np.allclose(a.positions, b.positions, rtol=rtol, atol=1e-14)


def test_qepp_header_and_values(al):
    ions, rho, lat, info = al
    assert rho.shape == (36, 36, 72) and ions.nat == 2 and ions.symbols == ["Al", "Al"]
    assert info["ibrav"] == 0 and info["plot_num"] == 0 and info["ecut"] == 70.0
    assert np.isclose(ions.zval["Al"], 3.0)
    a = 7.64583193
    assert np.allclose(lat, a * np.array([[.5, .5, 0], [.5, 0, .5], [0, 1, 1]]), rtol=1e-15)
    assert np.allclose(ions.positions[1], a * np.array([0, .5, .5]))
    assert np.isclose(rho[0, 0, 0], 2.197668430e-01) and np.isclose(rho[1, 0, 0], 1.964315971e-01)
    # two Al with 3 valence electrons each
    assert abs(rho.mean() * abs(np.linalg.det(lat)) - 6.0) < 1e-3


def test_qepp_roundtrip(al, tmp_path):
    ions, rho, lat, info = al
    p = tmp_path / "rho.pp"
    qepp.write_qepp(p, ions, rho, information={k: info[k] for k in ("gcutm", "dual", "ecut")})
    ions2, rho2, lat2, info2 = qepp.read_qepp_all(p)
    assert np.array_equal(rho2, rho)  # 16 significant digits of values read from 16-digit text
    _same_ions(ions2, ions)
    assert np.allclose(lat2, lat, rtol=1e-15, atol=0) and ions2.zval == ions.zval
    assert all(info2[k] == info[k] for k in ("gcutm", "dual", "ecut", "ibrav"))
    # five "%22.15e" values per line, the remainder on the last line
    out = p.read_text().splitlines()
    assert len(out[-2]) == 5 * 22 + 4 and len(out[-1].split()) == rho.size % 5 == 2


def test_qepp_potential_is_rydberg(tmp_path, al):
    ions, rho, lat, _ = al
    p = tmp_path / "v.pp"
    qepp.write_qepp(p, ions, rho, data_type="potential")
    assert qepp.read_qepp_all(p)[3]["plot_num"] == 1
    np.testing.assert_array_equal(qepp.read_qepp(p)[1], 2.0 * rho)
    np.testing.assert_array_equal(qepp.read_qepp(p, data_type="potential")[1], rho)


def test_qepp_ibrav1(tmp_path):
    """ibrav = 1: the lattice is celldm(1) times the identity and tau are
    fractions of it (formats/qepp.py:37-40,80-83)."""
    ions = vasp.read_poscar(data("fcc.vasp"), names=["Al"])
    ions.set_charges(3.0)
    rho = np.random.default_rng(0).random((6, 5, 7))
    p = tmp_path / "c.pp"
    qepp.write_qepp(p, ions, rho, information={"ibrav": 1, "celldm": [ions.cell[0, 0], 0, 0, 0, 0, 0]})
    assert p.read_text().splitlines()[4].split()[1:] == ["Al", "3.0"]  # no lattice lines for ibrav != 0
    ions2, rho2, lat2 = qepp.read_qepp(p)
    assert np.allclose(lat2, ions.cell, rtol=1e-15) and np.allclose(rho2, rho, rtol=1e-15, atol=0)
    _same_ions(ions2, ions, 1e-14)


def test_xsf_reads_reference_file(gaas):
    ions, rho, lat = gaas
    assert ions.nat == 10 and ions.symbols[:5] == ["Ga"] * 5
    # datagrid 17 x 27 x 19 with periodic end points dropped
    assert rho.shape == (16, 26, 18)
    assert np.isclose(lat[0, 0], 4.1604539101422 / Units.Bohr)
    assert np.isclose(ions.positions[0, 0], 4.036460308754025 / Units.Bohr)
    # first value in 1/A^3 -> 1/bohr^3
    assert np.isclose(rho[0, 0, 0], 0.07031350690112871 * Units.Bohr ** 3, rtol=1e-15)


def test_xsf_roundtrip(gaas, tmp_path):
    ions, rho, lat = gaas
    p = tmp_path / "g.xsf"
    xsf.write_xsf(p, ions, rho)
    ions2, rho2, lat2 = xsf.read_xsf(p)
    _same_ions(ions2, ions, 1e-15)
    assert np.allclose(lat2, lat, rtol=1e-15, atol=0)
    assert np.max(np.abs(rho2 / rho - 1.0)) < 1e-15
    # the periodic padding repeats the first planes
    txt = p.read_text().split("BEGIN_DATAGRID_3D#0\n")[1].splitlines()
    assert txt[0].split() == ["17", "27", "19"]


def test_xsf_bohr_potential_and_components(gaas, tmp_path):
    ions, rho, lat = gaas
    p = tmp_path / "v.xsf"
    two = np.stack([rho, 2.0 * rho])
    xsf.write_xsf(p, ions, two, data_type="potential", units="bohr")
    ions2, v2, lat2 = xsf.read_xsf(p, units="bohr", data_type="potential")
    assert v2.shape == (2,) + rho.shape
    assert np.max(np.abs(v2 / two - 1.0)) < 1e-15
    _same_ions(ions2, ions, 0)
    # stored in eV: reading as a density gives Ha -> eV scaled values
    _, raw, _ = xsf.read_xsf(p, units="bohr")
    assert np.allclose(raw, two * Units.Ha, rtol=1e-15)


def test_pp_to_xsf_to_snpy_agree(al, tmp_path):
    ions, rho, lat, _ = al
    px, ps = tmp_path / "a.xsf", tmp_path / "a.snpy"
    xsf.write_xsf(px, ions, rho)
    ix, rx, lx = xsf.read_xsf(px)
    snpy.write_snpy(ps, ix, rx)
    i3, r3, info = snpy.read_snpy(ps)
    assert info is None and np.array_equal(r3, rx)
    assert np.max(np.abs(rx / rho - 1.0)) < 1e-15 and np.allclose(lx, lat, rtol=1e-15)
    assert list(i3.numbers) == [13, 13] and i3.symbols == ["Al", "Al"]


def test_snpy_serial_like_reference(tmp_path):
    # examples/test/test_snpy.py:19-36
    ions, field, _ = io.read_all(data("GaAs_random.xsf"), device="cpu")
    out = str(tmp_path / "serial.snpy")
    snpy.write_snpy(out, ions, field)
    ions2, field2, _ = io.read_all(out, device="cpu")
    assert np.allclose(ions.cell, ions2.cell) and np.allclose(ions.positions, ions2.positions)
    assert np.allclose(ions.numbers, ions2.numbers) and list(ions2.numbers[:6]) == [31] * 5 + [33]
    assert np.array_equal(field.numpy(), field2.numpy())
    assert field2.grid.nrR.tolist() == [16, 26, 18]


def test_snpy_information_and_gzip(al, tmp_path):
    ions, rho, lat, _ = al
    p = tmp_path / "i.snpy.gz"
    snpy.write_snpy(p, ions, rho, information=np.arange(4.0))
    i2, r2, info = snpy.read_snpy(p)
    assert np.array_equal(r2, rho) and np.array_equal(info, np.arange(4.0))
    assert snpy.read_snpy(p, kind="ions")[1] is None


def test_vasp_reference_file_and_roundtrip(tmp_path):
    # examples/test/test_ase_pmg_io.py:11-17 reads fcc.vasp (species given on
    # the first line only, so the names come from the caller here)
    ions = io.read(data("fcc.vasp"), format="vasp", names=["Al"])
    a = 4.05 / Units.Bohr
    assert np.allclose(ions.cell, a * np.eye(3))
    assert np.allclose(ions.positions, a * np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]]))
    mixed = vasp.read_poscar(data("GaAs_random.vasp"))
    for direct in (False, True):
        (tmp_path / str(direct)).mkdir()
        p = tmp_path / str(direct) / "POSCAR"
        vasp.write_vasp(p, mixed, direct=direct)
        back = io.read(p)
        order = np.argsort(np.asarray(mixed.symbols), kind="stable")
        assert back.symbols == [mixed.symbols[i] for i in order]
        assert np.allclose(back.positions, mixed.positions[order], rtol=0, atol=1e-13)
        assert np.allclose(back.cell, mixed.cell, rtol=1e-15)


def test_io_front_end(al, tmp_path):
    ions, rho, lat, _ = al
    f = io.read_density(data("Al_fde_rho.pp"), device="cpu")
    assert f.grid.nrR.tolist() == [36, 36, 72] and np.allclose(f.grid.lattice, lat)
    assert np.array_equal(f.numpy(), rho)
    i2, f2, info = io.read_all(data("Al_fde_rho.pp"), device="cpu")
    assert info["ibrav"] == 0 and i2.nat == 2
    assert io.read(data("Al_fde_rho.pp")).nat == 2  # kind='ions' by default
    for ext in ("pp", "xsf", "snpy"):
        p = str(tmp_path / ("o." + ext))
        io.write(p, f, ions)  # either argument order
        g = io.read_density(p, device="cpu")
        assert np.max(np.abs(g.numpy() / rho - 1.0)) < 1e-15
    io.write_potential(str(tmp_path / "v.pp"), f, ions)
    assert np.array_equal(io.read_potential(str(tmp_path / "v.pp"), device="cpu").numpy(), rho)
    with pytest.raises(ValueError):
        io.read_density(data("fcc.vasp"), format="vasp")
    with pytest.raises(ValueError):
        io.read(str(tmp_path / "x.cube"))
    with pytest.raises(ValueError):
        io.read_density(data("Al_fde_rho.pp"), device="cpu",
                        grid=io.read_density(data("GaAs_random.xsf"), device="cpu").grid)
