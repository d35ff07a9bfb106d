"""Host-side logic of dftpy_amd and the C ABI surface (CPU only: nothing
here launches a GPU kernel)."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.interpolate import splev, splrep

from conftest import ROOT, data
from oracle import dcsrch as D
from oracle import grid as OG
from oracle import io as OIO


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "ofdft_hip.h")).read()
    return sorted(set(re.findall(r"\b(ofdft_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from dftpy_amd import _lib
    syms = header_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(_lib.lib, s), s
        assert s in _lib.SIGNATURES, s
    assert _lib.lib.ofdft_version() == 1


def _cpp_line_search(fun, step0, c1=1e-4, c2=0.2, xtol=1e-12, maxiter=50):
    from dftpy_amd import _lib
    st = np.zeros(32)
    stp = ctypes.c_double(step0)
    task = ctypes.c_int(0)
    f, g = fun(0.0)
    trials = []
    for _ in range(maxiter):
        _lib.check(_lib.lib.ofdft_dcsrch(ctypes.byref(stp), f, g, c1, c2, xtol, 0.0, np.pi, ctypes.byref(task),
                                         _lib.dptr(st)))
        if task.value == 1:
            trials.append(stp.value)
            f, g = fun(stp.value)
        else:
            break
    return (stp.value if task.value == 2 else None), trials


@pytest.mark.parametrize("case", range(4))
def test_cpp_dcsrch_matches_oracle(case):
    funs = [lambda a: (np.cos(a + 0.3) ** 2 - a * 0.1, -2 * np.cos(a + 0.3) * np.sin(a + 0.3) - 0.1),
            lambda a: ((a - 1.3) ** 2, 2 * (a - 1.3)),
            lambda a: (-np.sin(3 * a) + a ** 4, -3 * np.cos(3 * a) + 4 * a ** 3),
            lambda a: (np.exp(-a) + 0.05 * a ** 3, -np.exp(-a) + 0.15 * a ** 2)]
    for s0 in (0.01, 0.1, 1.0, 2.5):
        a1, t1 = _cpp_line_search(funs[case], s0)
        tr = []

        def f(a):
            tr.append(a)
            return list(funs[case](a))

        a2, task, _, _ = D.line_search(f, s0, list(funs[case](0.0)), c1=1e-4, c2=0.2, amax=np.pi, amin=0.0,
                                       xtol=1e-12, maxiter=50)
        assert (a1 is None) == (a2 is None)
        assert np.allclose(t1, tr, rtol=1e-15, atol=0)


def test_cpp_dcsrch_rejects_uphill_start():
    from dftpy_amd import _lib
    st = np.zeros(32)
    stp = ctypes.c_double(0.1)
    task = ctypes.c_int(0)
    _lib.lib.ofdft_dcsrch(ctypes.byref(stp), 1.0, +0.5, 1e-4, 0.2, 1e-12, 0.0, np.pi, ctypes.byref(task),
                          _lib.dptr(st))
    assert task.value == 4


def test_cubic_table_matches_fitpack():
    from dftpy_amd.math_utils import cubic_table
    gp, vp, _ = OIO.read_recpot(data("al.lda.recpot"))
    c = cubic_table(gp[1:], vp[1:])
    tck = splrep(gp[1:], vp[1:], k=3)
    q = np.linspace(gp[1], gp[-1] * 0.999, 5001)
    dq = (gp[-1] - gp[1]) / (len(gp) - 2)
    i = np.clip(np.floor((q - gp[1]) / dq).astype(int), 0, c.shape[0] - 1)
    t = q - (gp[1] + i * dq)
    v = c[i, 0] + t * (c[i, 1] + t * (c[i, 2] + t * c[i, 3]))
    ref = splev(q, tck)
    assert np.max(np.abs(v - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_recpot_roundtrip_and_empty_core(tmp_path):
    from dftpy_amd import formats, systems
    q, v = systems.empty_core_table()
    assert len(q) == 4001 and np.isclose(q[1], 0.01)
    p = str(tmp_path / "ec.recpot")
    formats.write_recpot(p, q, v)
    q2, v2, z = formats.read_recpot(p)
    assert z == 3
    assert np.allclose(q2, q, rtol=1e-14) and np.allclose(v2, v, rtol=1e-14)
    q3, v3, z3 = OIO.read_recpot(p)
    assert np.array_equal(v3, v2) and z3 == 3


def test_readers_match_oracle():
    from dftpy_amd import formats
    i1, r1, l1 = formats.read_qepp(data("Al_fde_rho.pp"))
    i2, r2, l2 = OIO.read_qepp(data("Al_fde_rho.pp"))
    assert np.array_equal(r1, r2) and np.array_equal(l1, l2) and np.allclose(i1.positions, i2.positions)
    p1 = formats.read_poscar(data("GaAs_random.vasp"))
    p2 = OIO.read_poscar(data("GaAs_random.vasp"))
    assert np.allclose(p1.positions, p2.positions, rtol=1e-15) and p1.symbols == p2.symbols
    x1 = formats.read_xsf(data("GaAs_random.xsf"))
    x2 = OIO.read_xsf(data("GaAs_random.xsf"))
    assert np.array_equal(x1[1], x2[1]) and np.allclose(x1[2], x2[2])


@pytest.mark.parametrize("n", [7, 20, 33, 64, 77, 100, 157, 160, 250, 301])
def test_best_fft_size_matches_oracle(n):
    from dftpy_amd.math_utils import best_fft_size
    assert best_fft_size(n) == OG.best_fft_size(n)


def test_fft_sizes_are_supported_radices():
    from dftpy_amd.math_utils import best_fft_size
    for n in range(4, 400, 2):
        m = best_fft_size(n)
        for p in (2, 3, 5, 7, 11, 13):
            while m % p == 0:
                m //= p
        assert m == 1


def test_fcc_builder():
    from dftpy_amd import systems
    ions = systems.fcc_al((5, 5, 5), sigma=0.1, seed=7)
    assert ions.nat == 500
    assert np.allclose(np.diag(ions.cell), 38.0)
    ref = systems.fcc_al((5, 5, 5))
    d = ions.positions - ref.positions
    assert 0.08 < d.std() < 0.12


def test_uniform_response_kernel_is_the_oracle_linear_response():
    """K(q) of math_utils.uniform_response_kernel (the TN preconditioner's
    Hessian) equals the linear response of the oracle's total potential
    (TF + vW + WT + PZ LDA + Hartree) to rho0 + eps cos(G.r), by central
    differences, at three wave vectors and both LDA branches of rs."""
    from oracle import functionals as OFN
    from dftpy_amd.math_utils import uniform_response_kernel
    L, n = 9.0, 18
    g = OG.Grid(np.eye(3) * L, (n, n, n))
    x = np.arange(n) * (L / n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    for rho0 in (0.03, 0.002):  # rs < 1 and rs > 1
        kern = OFN.wt_kernel(g.q, rho0)
        for m in ((1, 0, 0), (2, 1, 0), (3, 2, 1)):
            G = 2.0 * np.pi / L * np.array(m, dtype=np.float64)
            c = np.cos(G[0] * X + G[1] * Y + G[2] * Z)
            eps = 1e-5 * rho0
            v = []
            for s in (1.0, -1.0):
                r = rho0 + s * eps * c
                _, vt, _ = OFN.total(r, g, np.zeros_like(r), kern, rho0=rho0, calc_e=False)
                v.append(vt)
            dv = (v[0] - v[1]) / (2.0 * eps)
            k_num = float(np.sum(dv * c) / np.sum(c * c))
            k = float(uniform_response_kernel(np.linalg.norm(G), rho0))
            assert abs(k_num - k) < 1e-6 * abs(k), (rho0, m, k_num, k)


def test_tn_preconditioner_table():
    """The radial table is 1 / (2 rho0 K(q)): zero at q = 0 (Hartree),
    2 / q^2 at large q (vW dominates), covering the grid's largest |G|."""
    from types import SimpleNamespace
    from dftpy_amd import _lib
    from dftpy_amd.math_utils import uniform_response_kernel
    from dftpy_amd.optimization import tn_preconditioner
    grid = SimpleNamespace(lattice=np.diag([7.6, 7.6, 9.0]), nrR=np.array([32, 32, 36]))
    terms = _lib.TERM_TF | _lib.TERM_VW | _lib.TERM_WT | _lib.TERM_XC | _lib.TERM_HARTREE
    kedf = dict(x=1.0, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0)
    tab, dq = tn_preconditioner(grid, 0.0273, terms, kedf)
    q = np.arange(tab.size) * dq
    assert tab[0] == 0.0
    assert q[-1] >= np.pi * np.sqrt((32 / 7.6) ** 2 * 2 + (36 / 9.0) ** 2)
    np.testing.assert_allclose(tab[1:], 1.0 / (2 * 0.0273 * uniform_response_kernel(q[1:], 0.0273)), rtol=1e-14)
    assert abs(tab[-1] * q[-1] ** 2 / 2.0 - 1.0) < 0.05
    # without Hartree the q = 0 value is the finite local stiffness
    tab2, _ = tn_preconditioner(grid, 0.0273, terms & ~_lib.TERM_HARTREE, kedf)
    assert np.isfinite(tab2[0]) and tab2[0] > 0


def test_ions_scaled_positions_and_species_index():
    """Fractional coordinates satisfy f @ cell = positions on a skewed cell;
    species indices follow the symbols (also after they change) and name a
    missing species."""
    from dftpy_amd import Ions
    rng = np.random.default_rng(3)
    cell = np.array([[7.0, 0.3, -0.2], [1.1, 6.5, 0.4], [-0.7, 0.9, 8.2]])
    pos = rng.uniform(-3, 12, (7, 3))
    ions = Ions(symbols=["Ga", "As", "Ga", "Al", "As", "Ga", "Al"], positions=pos, cell=cell)
    f = ions.get_scaled_positions()
    assert np.allclose(f @ cell, pos, rtol=0, atol=1e-12)
    assert np.allclose(f, np.linalg.solve(cell.T, pos.T).T, rtol=0, atol=1e-13)
    keys = ["Al", "As", "Ga"]
    assert ions.species_index(keys).tolist() == [2, 1, 2, 0, 1, 2, 0]
    assert ions.species_index(keys).tolist() == [2, 1, 2, 0, 1, 2, 0]  # kept
    ions.symbols[3] = "Ga"
    assert ions.species_index(keys).tolist() == [2, 1, 2, 2, 1, 2, 0]
    with pytest.raises(ValueError, match="Al"):
        ions.species_index(["As", "Ga"])


@pytest.mark.parametrize("terms", [(True, True, True, True, True), (True, False, True, True, False),
                                   (False, True, False, False, True)])
def test_uniform_response_coefs_reproduce_the_kernel(terms):
    """The coefficient form the device evaluates equals the host response
    kernel on a q grid (all terms, and subsets)."""
    from dftpy_amd.math_utils import lindhard, uniform_response_coefs, uniform_response_kernel
    tf, vw, wt, xc, h = terms
    rho0 = 0.0231
    q = np.linspace(0.0, 9.0, 777)[1:]
    kw = dict(tf=tf, vw=vw, wt=wt, xc=xc, hartree=h, x=0.9, y=0.8, alpha=5.0 / 6.0, beta=5.0 / 6.0)
    ref = uniform_response_kernel(q, rho0, **kw)
    c = uniform_response_coefs(rho0, **kw)
    got = c[0] + c[1] * q * q + c[2] * lindhard(q / c[3], c[4], 1.0) + c[5] * 4.0 * np.pi / (q * q)
    assert np.allclose(got, ref, rtol=1e-13, atol=0)
