"""ofdft_atomic_density_noalloc (static device buffer, no allocation) against
ofdft_atomic_density, through the C ABI: two calls in a row, the second with
fewer atoms, fewer species and shorter tables, for the Gaussian and the
spline kinds; and an input beyond the buffer rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(rng, nat, nspecies, nbreak, kind, lattice):
    """Host arrays for one placement call (kept alive by the caller)."""
    from dftpy_amd import density as D
    frac = np.ascontiguousarray(rng.uniform(0, 1, (nat, 3)))
    spec = np.ascontiguousarray(rng.integers(0, nspecies, nat), dtype=np.int32)
    rcut = np.ascontiguousarray(rng.uniform(1.5, 3.0, nspecies))
    rtol = np.zeros(nspecies)
    sigma = np.ascontiguousarray(rng.uniform(0.2, 0.5, nspecies))
    scale = np.ascontiguousarray(rng.uniform(1.0, 3.0, nspecies))
    keep = [frac, spec, rcut, rtol, sigma, scale]
    nb = br = co = None
    if kind == 0:
        tabs = []
        for k in range(nspecies):
            r = np.linspace(0.0, 3.5, nbreak)
            tabs.append(D._spline_ppoly(r, scale[k] * np.exp(-r * r / (0.5 + 0.1 * k))))
        nb = np.ascontiguousarray([len(t[0]) for t in tabs], dtype=np.int32)
        br = (ctypes.POINTER(ctypes.c_double) * nspecies)(*[t[0].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                             for t in tabs])
        co = (ctypes.POINTER(ctypes.c_double) * nspecies)(*[t[1].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                             for t in tabs])
        keep += [tabs, nb]
    return dict(frac=frac, spec=spec, rcut=rcut, rtol=rtol, sigma=sigma, scale=scale, nb=nb, br=br, co=co,
                keep=keep, nat=nat, ns=nspecies)


def _place(fn, plan, kind, a, out):
    from dftpy_amd import _lib
    return fn(plan.handle, kind, a["ns"], _lib.iptr(a["nb"]) if a["nb"] is not None else None, a["br"], a["co"],
              _lib.dptr(a["rcut"]), _lib.dptr(a["rtol"]), _lib.dptr(a["sigma"]), _lib.dptr(a["scale"]), a["nat"],
              _lib.dptr(a["frac"]), _lib.iptr(a["spec"]), ctypes.c_void_p(out.data_ptr()), _lib.stream())


@pytest.mark.parametrize("kind", [1, 0])
def test_noalloc_matches_allocating_entry_point(kind):
    from dftpy_amd import DirectGrid, _lib
    lat = np.array([[9.0, 0.0, 0.0], [0.5, 8.0, 0.0], [0.3, -0.4, 10.0]])
    grid = DirectGrid(lat, (36, 32, 40), full=False)
    plan = grid.plan(torch.float64)
    rng = np.random.default_rng(11 + kind)
    # second call: fewer atoms, species and breakpoints than the first
    for nat, ns, nbreak in ((300, 3, 120), (40, 2, 30)):
        print("atomic_density_noalloc kind %d nat %d species %d breaks %d" % (kind, nat, ns, nbreak), flush=True)
        a = _inputs(rng, nat, ns, nbreak, kind, lat)
        ref = torch.zeros(tuple(grid.nrR), dtype=torch.float64, device="cuda")
        got = torch.full_like(ref, 0.25)  # added to what the buffer holds
        _lib.check(_place(_lib.lib.ofdft_atomic_density, plan, kind, a, ref))
        _lib.check(_place(_lib.lib.ofdft_atomic_density_noalloc, plan, kind, a, got))
        ref += 0.25
        torch.cuda.synchronize()
        diff = float((got - ref).abs().max().item())
        # atomic additions in a different order: not bit-exact, FP64 rounding only
        assert diff <= 1e-13 * float(ref.abs().max().item()), diff
        assert float(ref.max().item()) > 0.5


def test_noalloc_rejects_inputs_beyond_its_buffer():
    from dftpy_amd import DirectGrid, _lib
    grid = DirectGrid(np.eye(3) * 8.0, (16, 16, 16), full=False)
    plan = grid.plan(torch.float64)
    a = _inputs(np.random.default_rng(3), 200000, 1, 0, 1, np.eye(3) * 8.0)  # 5.6 MB of positions
    out = torch.zeros(tuple(grid.nrR), dtype=torch.float64, device="cuda")
    rc = _place(_lib.lib.ofdft_atomic_density_noalloc, plan, 1, a, out)
    assert rc != 0
    msg = _lib.lib.ofdft_last_error().decode()
    assert "static device buffer" in msg, msg
    assert float(out.abs().max().item()) == 0.0
