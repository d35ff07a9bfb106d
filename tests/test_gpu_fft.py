"""HIP FFTs against numpy's pocketfft with DFTpy's normalisation
(field.py:373-390,739-764): forward * dV, inverse / dV."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (6, 10, 15), (20, 30, 22), (36, 36, 72), (7, 11, 13), (1, 6, 10), (5, 1, 9), (32, 32, 32),
          (27, 25, 49), (96, 96, 96), (128, 64, 56)]


def _grid(shape, dtype, full=False):
    from dftpy_amd import DirectGrid
    lat = np.array([[7.1, 0.2, 0.0], [0.3, 8.0, 0.1], [0.0, 0.5, 9.3]])
    return DirectGrid(lat, shape, full=full, dtype=dtype)


def _tol(dtype):
    return 1e-12 if dtype == torch.float64 else 2e-6


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_r2c_c2r(shape, dtype):
    from dftpy_amd import DirectField
    g = _grid(shape, dtype)
    rng = np.random.default_rng(hash(shape) % 2 ** 32)
    f = rng.standard_normal(shape)
    F = DirectField(g, data=f).fft()
    ref = np.fft.rfftn(f) * g.dV
    got = F.data.cpu().numpy()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= _tol(dtype) * np.max(np.abs(ref)) * 4
    back = F.ifft().data.cpu().numpy()
    assert np.max(np.abs(back - f)) <= _tol(dtype) * np.max(np.abs(f)) * 8
    # irfftn semantics on a spectrum that is not Hermitian at k=0 / Nyquist
    spec = np.fft.rfftn(f) * g.dV + 0.3j
    from dftpy_amd import ReciprocalField
    got = ReciprocalField(g.get_reciprocal(), data=torch.as_tensor(spec)).ifft().data.cpu().numpy()
    ref = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2)) / g.dV
    assert np.max(np.abs(got - ref)) <= _tol(dtype) * np.max(np.abs(ref)) * 8


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 10, 15), (16, 26, 18), (7, 11, 13), (32, 32, 32), (64, 48, 40)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_c2c(shape, dtype):
    from dftpy_amd import DirectField
    g = _grid(shape, dtype, full=True)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fld = DirectField(g, data=torch.as_tensor(f), cplx=True)
    F = fld.fft()
    ref = np.fft.fftn(f) * g.dV
    got = F.data.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= _tol(dtype) * np.max(np.abs(ref)) * 4
    back = F.ifft().data.cpu().numpy()
    assert np.max(np.abs(back - f)) <= _tol(dtype) * np.max(np.abs(f)) * 8


def test_gaussian_pair():
    # examples/test/test_field.py:247-285 (a 2/3/5-smooth grid instead of 51^3)
    from dftpy_amd import DirectField, DirectGrid
    nr = (50, 50, 50)
    g = DirectGrid(np.eye(3) * 10.0, nr, full=False)
    r = g.r.cpu().numpy()
    c = r[:, 25, 25, 25]
    s = 0.5
    a = r - c[:, None, None, None]
    df = (s * np.sqrt(2 * np.pi)) ** -3 * np.exp(-0.5 * np.einsum("lijk,lijk->ijk", a, a) / s ** 2)
    F = DirectField(g, data=df).fft()
    gv = g.get_reciprocal().g.cpu().numpy()
    gg = (gv ** 2).sum(0)
    ref = np.exp(-s ** 2 * gg / 2.0) * np.exp(-1j * np.einsum("lijk,l->ijk", gv, c))
    assert np.allclose(F.data.cpu().numpy(), ref, atol=1e-8)
    assert np.isclose(DirectField(g, data=df).integral(), 1.0)


def test_160_cube_fp32_fp64():
    from dftpy_amd import DirectField
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 3e-6)):
        g = _grid((160, 160, 160), dt)
        rng = np.random.default_rng(11)
        f = rng.random((160, 160, 160))
        F = DirectField(g, data=f).fft()
        ref = np.fft.rfftn(f) * g.dV
        assert np.max(np.abs(F.data.cpu().numpy() - ref)) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", [(160, 160, 160), (96, 72, 100), (48, 80, 64), (128, 120, 144), (32, 256, 240)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fixed_and_generic_kernels_agree(shape, dtype):
    """The size-specialised passes against pocketfft and against the generic
    mixed-radix kernels on the same input."""
    from dftpy_amd import DirectField, _lib
    g = _grid(shape, dtype)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(shape)
    fld = DirectField(g, data=f)
    plan = g.plan(dtype)
    _lib.check(_lib.lib.ofdft_plan_set_fft_kernels(plan.handle, 1))
    F1 = fld.fft().data.clone()
    b1 = DirectField(g, data=f).fft().ifft().data.clone()
    _lib.check(_lib.lib.ofdft_plan_set_fft_kernels(plan.handle, 0))
    F0 = fld.fft().data.clone()
    _lib.check(_lib.lib.ofdft_plan_set_fft_kernels(plan.handle, 1))
    ref = np.fft.rfftn(f) * g.dV
    tol = _tol(dtype) * np.max(np.abs(ref)) * 4
    assert np.max(np.abs(F1.cpu().numpy() - ref)) <= tol
    assert np.max(np.abs(F0.cpu().numpy() - ref)) <= tol
    assert np.max(np.abs(b1.cpu().numpy() - f)) <= _tol(dtype) * 8 * np.max(np.abs(f))


@pytest.mark.parametrize("shape", [(160, 160, 160), (96, 72, 100), (64, 48, 32)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_tile_variants_agree(shape, dtype):
    """Every variant of the size-specialised passes (the autotuner's search
    space) gives the same transforms and the same WT evaluation: the tile
    variants bit for bit, the reversed radix order of the column passes to
    rounding (1e-13 relative in FP64, 1e-5 in FP32, whose WT potential
    amplifies the transforms' rounding)."""
    from dftpy_amd import DirectField
    from dftpy_amd.functional import Functional
    g = _grid(shape, dtype)
    rng = np.random.default_rng(9)
    f = 0.02 + 0.01 * rng.random(shape)
    plan = g.plan(dtype)
    print("autotuned", plan.variants, plan.variant_ms.tolist())

    def run(v):
        plan.set_variants(v)
        F = DirectField(g, data=f).fft().data.cpu().numpy()
        o = Functional(type="KEDF", name="WT")(DirectField(g, data=f))
        return F, o.energy, o.potential.numpy()

    tiles = [run(v) for v in ([0, 0, 0, 0], [1, 1, 1, 1], [2, 2, 2, 2])]
    for F, e, p in tiles[1:]:
        assert np.array_equal(F, tiles[0][0])
        assert abs(e - tiles[0][1]) <= 1e-12 * abs(tiles[0][1])
        assert np.max(np.abs(p - tiles[0][2])) <= 1e-12 * np.max(np.abs(tiles[0][2]))
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    for v in ([0, 0, 3, 3], [1, 1, 4, 5], [2, 2, 5, 4], plan.variants):
        F, e, p = run(v)
        assert np.max(np.abs(F - tiles[0][0])) <= tol * np.max(np.abs(tiles[0][0]))
        assert abs(e - tiles[0][1]) <= tol * abs(tiles[0][1])
        assert np.max(np.abs(p - tiles[0][2])) <= tol * np.max(np.abs(tiles[0][2]))
