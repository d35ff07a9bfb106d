"""Warm-start network kernels (csrc/cnn.hip) against the FP64 oracle
(oracle/cnn.py) with the same bfloat16 rounding points.  The grid shape is
not a multiple of the kernels' point tiles, and dilations > 1 wrap around
the periodic cell."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPE = (12, 10, 15)
DIL = (1, 2, 1, 3, 2, 1)
# (12, 10, 15): every layer on the point-tiled kernel; (6, 8, 32),
# (4, 16, 64), (4, 8, 96) and (4, 8, 160): layers with ny % 4d == 0 on the
# row-tiled kernel (one instance per nz / 32 = 1, 2, 3, 5: 96 is the
# training grid, 160 the 500-atom cell's), the rest on the point-tiled one;
# (6, 12, 40): the inference forward's ring-staged kernel with a partial z
# block (40 = 32 + 8) and partial y groups (12 / d rows per residue class,
# d = 1, 2, 3)
SHAPES = [(12, 10, 15), (6, 8, 32), (4, 16, 64), (4, 8, 96), (4, 8, 160), (6, 12, 40)]


def _setup(seed=0, scale=1.5, shape=SHAPE):
    from oracle import cnn
    rng = np.random.default_rng(seed)
    p = cnn.init_params(seed + 10, scale=scale)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.2, p[k].shape)
    p = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}
    x0 = rng.normal(0, 1, (2,) + shape).astype(np.float32).astype(np.float64)
    target = rng.normal(0, 0.3, shape).astype(np.float32).astype(np.float64)
    return p, x0, target


def _net(p, dil=DIL):
    from dftpy_amd.learned import WarmStartNet
    net = WarmStartNet(dil=dil)
    net.set_params(p)
    return net


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("train", [False, True])
def test_forward_matches_oracle(train, shape):
    from oracle import cnn
    p, x0, _ = _setup(shape=shape)
    ref, _ = cnn.forward(p, x0, DIL, bf16=True)
    net = _net(p)
    x = torch.as_tensor(x0.reshape(2, -1), dtype=torch.float32, device="cuda")
    out = net.forward(x, shape, train=train).cpu().numpy().astype(np.float64).reshape(shape)
    print("forward rel L2 %.2e max %.2e (|ref| max %.2e)" % (_rel(out, ref), np.abs(out - ref).max(), np.abs(ref).max()))
    assert _rel(out, ref) < 2e-3
    assert np.abs(out - ref).max() < 2e-2 * np.abs(ref).max()


@pytest.mark.parametrize("dil", [(2, 1, 1, 1, 1, 1), (3, 1, 1, 1, 1, 2)])
def test_forward_first_layer_dilations(dil):
    """The row-staged first layer with d > 1 (its periodic z halo in LDS is d
    points each side) and the one-pass output layer with d = 2."""
    from oracle import cnn
    shape = (5, 24, 48)  # ny % 4d == 0 for d = 2, 3
    p, x0, _ = _setup(4, shape=shape)
    ref, _ = cnn.forward(p, x0, dil, bf16=True)
    net = _net(p, dil)
    x = torch.as_tensor(x0.reshape(2, -1), dtype=torch.float32, device="cuda")
    out = net.forward(x, shape).cpu().numpy().astype(np.float64).reshape(shape)
    print("forward rel L2 %.2e" % _rel(out, ref))
    assert _rel(out, ref) < 2e-3
    assert np.abs(out - ref).max() < 2e-2 * np.abs(ref).max()


@pytest.mark.parametrize("shape", SHAPES)
def test_backward_matches_oracle(shape):
    from oracle import cnn
    from dftpy_amd import DirectGrid
    p, x0, target = _setup(1, shape=shape)
    ref_out, cache = cnn.forward(p, x0, DIL, bf16=True)
    ref_loss, dout = cnn.loss(ref_out, target)
    gref = cnn.backward(p, cache, dout)
    net = _net(p)
    grid = DirectGrid(np.eye(3) * 5.0, shape, full=False, dtype=torch.float32)
    plan = grid.plan(torch.float32)
    x = torch.as_tensor(x0.reshape(2, -1), dtype=torch.float32, device="cuda")
    tgt = torch.as_tensor(target.reshape(-1), dtype=torch.float32, device="cuda")
    loss = net.train_step(plan, x, shape, tgt, lr=0.0)  # lr 0: gradients only
    assert abs(float(loss.item()) - ref_loss) < 1e-2 * ref_loss
    for k, g in gref.items():
        got = net.param(k, net.grad).cpu().numpy().astype(np.float64)
        print("%s rel %.2e" % (k, _rel(got, g)))
        assert _rel(got, g) < 1e-2, k


@pytest.mark.parametrize("shape", [(8, 16, 96), (8, 16, 160), (8, 20, 72)])
def test_shipped_dilations_match_oracle(shape):
    """The shipped network's dilations (1, 2, 4, 4, 2, 1) with nx and ny
    multiples of d for every layer: the inference forward runs all four
    middle layers on the ring-staged kernel, the training forward and the
    backward on the row-tiled instances of the 96^3 training grid
    (nz / 32 = 3) and of the 160^3 cell (nz / 32 = 5); (8, 20, 72) is a
    batch-cell z size (72 = 2 x 32 + 8) with y groups that end short for
    d = 4 (20 / 4 = 5 rows per residue class), its training forward and
    backward on the point-tiled kernel."""
    from oracle import cnn
    from dftpy_amd import DirectGrid
    from dftpy_amd.learned import WarmStartNet
    dil = WarmStartNet().dil
    assert dil == (1, 2, 4, 4, 2, 1)
    p, x0, target = _setup(3, shape=shape)
    ref, cache = cnn.forward(p, x0, dil, bf16=True)
    net = _net(p, dil)
    x = torch.as_tensor(x0.reshape(2, -1), dtype=torch.float32, device="cuda")
    for train in (False, True):
        out = net.forward(x, shape, train=train).cpu().numpy().astype(np.float64).reshape(shape)
        print("forward (train=%s) rel L2 %.2e" % (train, _rel(out, ref)))
        assert _rel(out, ref) < 2e-3
        assert np.abs(out - ref).max() < 2e-2 * np.abs(ref).max()
    ref_loss, dout = cnn.loss(ref, target)
    gref = cnn.backward(p, cache, dout)
    grid = DirectGrid(np.eye(3) * 5.0, shape, full=False, dtype=torch.float32)
    tgt = torch.as_tensor(target.reshape(-1), dtype=torch.float32, device="cuda")
    loss = net.train_step(grid.plan(torch.float32), x, shape, tgt, lr=0.0)
    assert abs(float(loss.item()) - ref_loss) < 1e-2 * ref_loss
    for k, g in gref.items():
        got = net.param(k, net.grad).cpu().numpy().astype(np.float64)
        print("%s rel %.2e" % (k, _rel(got, g)))
        assert _rel(got, g) < 1e-2, k


def test_ring_forward_matches_row_kernel_at_160():
    """At the 500-atom cell's full 160^3 grid (where the FP64 oracle is too
    slow to run), the inference forward (ring-staged middle layers) against
    the training forward (row-tiled middle layers, the instance the oracle
    tests above pin at nz = 160): the same network, the same bf16 rounding
    points, FP32 accumulation in a different order."""
    from dftpy_amd.learned import WarmStartNet
    from oracle import cnn
    shape = (160, 160, 160)
    p = cnn.init_params(21, scale=1.5)
    rng = np.random.default_rng(21)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.2, p[k].shape)
    net = WarmStartNet()
    net.set_params({k: v.astype(np.float32).astype(np.float64) for k, v in p.items()})
    x = torch.randn((2, int(np.prod(shape))), dtype=torch.float32, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(5))
    a = net.forward(x, shape, train=False).clone()
    b = net.forward(x, shape, train=True).clone()
    rel = float((a - b).norm() / b.norm())
    print("ring vs row kernel at 160^3: rel L2 %.2e, max %.2e" % (rel, float((a - b).abs().max())))
    assert rel < 1e-3
    assert float((a - b).abs().max()) < 1e-2 * float(b.abs().max())


@pytest.mark.parametrize("shape,dil", [((12, 10, 15), 1), ((4, 8, 160), 1), ((9, 6, 40), 2), ((8, 7, 300), 3),
                                       ((160, 160, 160), 1)])
def test_output_layer_one_pass_equals_tap_planes(shape, dil):
    """The output layer's one-pass form (work NULL: projections of a plane
    in LDS while the layer walks x) against the two-pass form through 27 FP16
    tap planes: the same FP16 projections summed in the same tap order, so the
    outputs are identical.  Shapes: walk cycles gcd(nx, d) > 1 (9, 6, 40; d 2
    is coprime to 9 and ny 6 = 3 y-blocks of 2 rows wrap), nz above one LDS
    block (300), and the 160^3 grid of the 500-atom cell."""
    from dftpy_amd import _lib
    n = int(np.prod(shape))
    g = torch.Generator(device="cuda").manual_seed(11)
    h = (torch.randn((n, 32), device="cuda", generator=g) * 2).to(torch.bfloat16)
    w = torch.randn((27, 32), device="cuda", generator=g) * 0.1
    b = torch.tensor([0.25], device="cuda")
    nr = np.ascontiguousarray(shape, dtype=np.int32)
    one = torch.full((n,), float("nan"), device="cuda")
    two = torch.full((n,), float("nan"), device="cuda")
    work = torch.empty(27 * n, device="cuda")
    s = _lib.stream()
    L = _lib.lib
    _lib.check(L.ofdft_cnn_conv_out(_lib.iptr(nr), dil, h.data_ptr(), w.data_ptr(), b.data_ptr(), one.data_ptr(),
                                    None, s))
    _lib.check(L.ofdft_cnn_conv_out(_lib.iptr(nr), dil, h.data_ptr(), w.data_ptr(), b.data_ptr(), two.data_ptr(),
                                    work.data_ptr(), s))
    assert bool(torch.isfinite(one).all())
    assert torch.equal(one, two)
    if n <= 20000:  # and both against the FP64 convolution of the FP16-rounded operands
        hh = h.float().cpu().numpy().astype(np.float64).reshape(shape + (32,))
        ww = w.half().double().cpu().numpy()
        ref = np.full(shape, 0.25)
        for t in range(27):
            dx, dy, dz = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
            ref += np.roll(hh @ ww[t], shift=(-dx * dil, -dy * dil, -dz * dil), axis=(0, 1, 2))
        err = np.abs(one.cpu().numpy().reshape(shape) - ref).max()
        print("one-pass vs FP64: max %.2e (|ref| max %.2e)" % (err, np.abs(ref).max()))
        assert err < 2e-3 * np.abs(ref).max()


def _oracle_adam(p, x0, tgt, lr, steps):
    from oracle import cnn
    q = {k: v.copy() for k, v in p.items()}
    m = {k: 0 * v for k, v in q.items()}
    v2 = {k: 0 * v for k, v in q.items()}
    losses = []
    for t in range(1, steps + 1):
        out, c = cnn.forward(q, x0, DIL, bf16=True)
        loss, d = cnn.loss(out, tgt)
        g = cnn.backward(q, c, d)
        losses.append(loss)
        for k in q:  # Adam, Kingma & Ba (2015), with bias correction
            m[k] = 0.9 * m[k] + 0.1 * g[k]
            v2[k] = 0.999 * v2[k] + 0.001 * g[k] ** 2
            q[k] -= lr * (m[k] / (1 - 0.9 ** t)) / (np.sqrt(v2[k] / (1 - 0.999 ** t)) + 1e-8)
    return losses


def test_training_trajectory_matches_oracle():
    """Sixty Adam steps on a learnable target (a linear map of the inputs)
    with the GPU forward, backward and Adam kernels follow the oracle's
    training trajectory and lower the loss."""
    from dftpy_amd import DirectGrid
    p, x0, _ = _setup(2, scale=1.0)
    tgt64 = 0.3 * x0[0] - 0.1 * x0[1]
    ref = _oracle_adam(p, x0, tgt64, 5e-4, 60)
    net = _net(p)
    grid = DirectGrid(np.eye(3) * 5.0, SHAPE, full=False, dtype=torch.float32)
    plan = grid.plan(torch.float32)
    x = torch.as_tensor(x0.reshape(2, -1), dtype=torch.float32, device="cuda")
    tgt = torch.as_tensor(tgt64.reshape(-1), dtype=torch.float32, device="cuda")
    losses = [float(net.train_step(plan, x, SHAPE, tgt, lr=5e-4).item()) for _ in range(60)]
    print(["%.4f" % v for v in losses[::6]])
    print(["%.4f" % v for v in ref[::6]])
    assert np.max(np.abs(np.array(losses) - ref) / np.array(ref)) < 0.02
    assert min(losses[-10:]) < 0.5 * losses[0]


def test_spectral_loss_matches_numpy():
    """Spectral (Hessian-metric) loss and its gradient against numpy's rfftn;
    w = 1 reduces to the mean squared error (Parseval)."""
    from dftpy_amd import DirectGrid
    from dftpy_amd import _lib
    rng = np.random.default_rng(4)
    grid = DirectGrid(np.eye(3) * 5.0, SHAPE, full=False, dtype=torch.float32)
    plan = grid.plan(torch.float32)
    pred = rng.normal(0, 1, SHAPE).astype(np.float32)
    tgt = rng.normal(0, 1, SHAPE).astype(np.float32)
    nzc = SHAPE[2] // 2 + 1
    for w in (np.ones(SHAPE[:2] + (nzc,), np.float32), rng.uniform(0, 3, SHAPE[:2] + (nzc,)).astype(np.float32)):
        d = pred.astype(np.float64) - tgt
        D = np.fft.rfftn(d)
        m = np.full(nzc, 2.0)
        m[0] = 1.0
        if SHAPE[2] % 2 == 0:
            m[-1] = 1.0
        n = d.size
        ref = float((m * w * np.abs(D) ** 2).sum() / n ** 2)
        gref = 2.0 / n * np.fft.irfftn(w * D, s=SHAPE)
        tp = torch.as_tensor(pred.reshape(-1), device="cuda")
        tt = torch.as_tensor(tgt.reshape(-1), device="cuda")
        tw = torch.as_tensor(w, device="cuda").contiguous()
        dp = torch.empty_like(tp)
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib.ofdft_cnn_loss_spectral(plan.handle, tp.data_ptr(), tt.data_ptr(), tw.data_ptr(),
                                                    dp.data_ptr(), loss.data_ptr(), _lib.stream()))
        assert abs(float(loss.item()) - ref) < 1e-5 * ref
        g = dp.cpu().numpy().astype(np.float64).reshape(SHAPE)
        assert np.abs(g - gref).max() < 1e-5 * np.abs(gref).max()
        if np.all(w == 1):
            assert abs(ref - float((d * d).mean())) < 1e-9 * ref
