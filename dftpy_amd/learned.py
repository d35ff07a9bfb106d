"""Learned warm start for the density optimizer.

The reference starts ``Optimization.optimize_rho`` from a guess density:
the uniform electron gas or a superposition of atomic densities
(``density/density.py:199-207,258-317``).  This module adds a start density
produced by a small periodic 3D CNN (the HIP kernels of ``csrc/cnn.hip``)
from two cheap fields:

* the superposition density rho_s: each atom carries the linear density
  response of the homogeneous electron gas at the mean density rho0 to its
  own local pseudopotential, rho_a(q) = -v_a(q) / K(q), where K(q) is the
  second functional derivative of the total functional at rho0 (TF, vW, WT
  and LDA from their definitions, plus Hartree 4 pi / q^2), with
  rho_a(0) = Z.  It is placed on the grid with the same structure-factor /
  PME machinery as the local pseudopotential (``ofdft_local_pp``).
* the local pseudopotential field vloc.

Network input x0 = [c (rho_s - rho0) / rho0, sv * vloc]; output delta; start
density rho = rho0 exp(x0[0] + delta), normalised to the electron count.
The optimizer then finishes from that density, so the plain method remains
both the fallback and the arbiter of the result.
"""
import os

import numpy as np
import torch

from . import _lib
from .field import DirectField
from .functional.pseudo import RadialTable
from .math_utils import lindhard, uniform_response_coefs

C = 32
NTAP = 27
LAYER_SHAPES = [("W1", (NTAP, 2, C)), ("b1", (C,))]
for _l in range(2, 6):
    LAYER_SHAPES += [("W%d" % _l, (NTAP, C, C)), ("b%d" % _l, (C,))]
LAYER_SHAPES += [("W6", (NTAP, C)), ("b6", (1,))]

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "warmstart_al_wt.npz")

CTF = 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)


# ------------------------------------------------------------ host-side tables
def _lda_dv(rho0):
    """d v_xc / d rho of the PZ LDA at a uniform density (semilocal_xc.py:291-303)."""
    a, c, d = 0.0311, 0.0020, -0.0116
    g, b1, b2 = -0.1423, 1.0529, 0.3334
    rc = np.cbrt(rho0)
    rs = np.cbrt(3.0 / (4.0 * np.pi)) / rc
    fx = -np.cbrt(3.0 / np.pi) / 3.0 * rc / rho0
    if rs < 1:
        fc = -a / 3.0 - (c / 9.0 * (np.log(rs) * 2.0 + 1.0) + d * 2.0 / 9.0) * rs
    else:
        ta, tb = b1 * np.sqrt(rs), b2 * rs
        fc = g / 36.0 * (5.0 * ta + 7.0 * ta * ta + 8.0 * tb + 16.0 * tb * tb + 21.0 * ta * tb) / (1 + ta + tb) ** 3
    return fx + fc / rho0


def response_kernel(q, rho0, x=1.0, y=1.0, alpha=5.0 / 6.0):
    """K(q) = d^2 E / d rho(q)^2 of TF + vW + WT + LDA + Hartree at the
    uniform density rho0 (alpha = beta)."""
    q = np.asarray(q, dtype=np.float64)
    tkf = 2.0 * (3.0 * rho0 * np.pi ** 2) ** (1.0 / 3.0)
    wtfac = 5.0 / (9.0 * alpha * alpha * rho0 ** (2 * alpha - 5.0 / 3.0)) * CTF
    k_tf = x * (10.0 / 9.0) * CTF * rho0 ** (-1.0 / 3.0)
    k_vw = y * q * q / (4.0 * rho0)
    k_wt = 2.0 * alpha * alpha * rho0 ** (2 * alpha - 2) * lindhard(q / tkf, x, 1.0) * wtfac
    with np.errstate(divide="ignore"):
        k_h = np.where(q > 0, 4.0 * np.pi / np.maximum(q, 1e-300) ** 2, np.inf)
    return k_tf + k_vw + k_wt + k_h + _lda_dv(rho0)


def response_table(gp, vp, z, rho0):
    """Radial table of one ion's screening density rho_a(q) = -v(q)/K(q), with
    rho_a(0) = Z, on the recpot q grid."""
    gp = np.asarray(gp, dtype=np.float64)
    vp = np.asarray(vp, dtype=np.float64)
    K = response_kernel(gp, rho0)
    ra = np.empty_like(vp)
    ra[1:] = -vp[1:] / K[1:]
    ra[0] = float(z)
    return RadialTable(gp, ra)


def metric_weights(grid, rho0):
    """w(q) = K(q) / K(q -> TF) on the half spectrum of ``grid`` (w(0) = 0):
    the Hessian of the total energy at the uniform density, so that the
    spectral training loss sum_q w |d(q)|^2 of a log-density error d weighs
    each wavelength by the energy it costs."""
    nr = [int(n) for n in grid.nrR]
    B = 2.0 * np.pi * np.linalg.inv(np.asarray(grid.lattice)).T
    kx = np.fft.fftfreq(nr[0], 1.0 / nr[0])
    ky = np.fft.fftfreq(nr[1], 1.0 / nr[1])
    kz = np.arange(nr[2] // 2 + 1, dtype=np.float64)
    K = np.stack(np.meshgrid(kx, ky, kz, indexing="ij"), axis=-1)
    q = np.linalg.norm(K @ B, axis=-1)
    qq = np.where(q > 0, q, 1.0)
    w = response_kernel(qq, rho0) / ((10.0 / 9.0) * CTF * rho0 ** (-1.0 / 3.0))
    w[0, 0, 0] = 0.0
    w = np.maximum(w, 0.0)
    return torch.as_tensor(w.astype(np.float32), device=grid.device).contiguous()


def superposition_density(grid, ions, readpp, rho0, pme=True, order=10):
    """rho_s on ``grid`` (device tensor in the grid's precision; the network's
    inputs take FP32): sum over atoms of the screening densities
    rho_a(q) = -v_a(q) / K(q) (the G = 0 term gives the electron count),
    evaluated on the device from each species' pseudopotential table and the
    closed-form uniform-gas response K at rho0 (response_table gives the same
    function as a radial table)."""
    plan = grid.plan(grid.dtype)
    keys = sorted(readpp.tables)
    ns = len(keys)
    # the pseudopotential tables' arguments, kept on the reader with their arrays
    cache = readpp.__dict__.get("_superposition_args")
    if cache is None or cache[0] != tuple(keys):
        tabs = [readpp.tables[k] for k in keys]
        coefs = [np.ascontiguousarray(t.coef.reshape(-1)) for t in tabs]
        arrs = (np.array([t.q0 for t in tabs]), np.array([t.dq for t in tabs]),
                np.array([t.coef.shape[0] for t in tabs], dtype=np.int32), np.array([t.qmax for t in tabs]),
                np.array([v for t in tabs for v in t.vsmall]), np.array([t.dqs for t in tabs]),
                np.array([float(readpp.zval[k]) for k in keys]))
        cptr = (_lib._DP * ns)(*[_lib.dptr(c) for c in coefs])
        ptrs = (_lib.dptr(arrs[0]), _lib.dptr(arrs[1]), _lib.iptr(arrs[2]), cptr, _lib.dptr(arrs[3]),
                _lib.dptr(arrs[4]), _lib.dptr(arrs[5]), _lib.dptr(arrs[6]))
        q12 = np.array([[readpp.gp[k][1], readpp.gp[k][2]] for k in keys])
        v12 = np.array([[readpp.vp[k][1], readpp.vp[k][2]] for k in keys])
        cache = (tuple(keys), arrs, coefs, ptrs, q12, v12)
        readpp._superposition_args = cache
    ptrs, q12, v12 = cache[3], cache[4], cache[5]
    # the response coefficients at rho0 and the screened values at the first
    # two table points (the small-q patch), kept for the last rho0
    rc = readpp.__dict__.get("_superposition_resp")
    if rc is None or rc[0] != float(rho0):
        K = response_kernel(q12.reshape(-1), rho0).reshape(q12.shape)
        r1 = np.ascontiguousarray(-v12[:, 0] / K[:, 0])
        r2 = np.ascontiguousarray(-v12[:, 1] / K[:, 1])
        resp = uniform_response_coefs(rho0)
        rc = (float(rho0), r1, r2, resp, _lib.dptr(r1), _lib.dptr(r2), _lib.dptr(resp))
        readpp._superposition_resp = rc
    pos = np.ascontiguousarray(ions.positions, dtype=np.float64)
    frac = np.ascontiguousarray(ions.get_scaled_positions(), dtype=np.float64)
    spec = ions.species_index(keys)
    out = torch.empty(tuple(int(n) for n in grid.nrR), dtype=grid.dtype, device=grid.device)
    _lib.check(_lib.lib.ofdft_screened_superposition(plan.handle, ns, *ptrs, rc[4], rc[5], rc[6], ions.nat,
                                                     _lib.dptr(pos), _lib.dptr(frac),
                                                     _lib.iptr(spec), int(bool(pme)), int(order), out.data_ptr(),
                                                     _lib.stream()))
    return out


# ------------------------------------------------------------------ network
class WarmStartNet:
    """The 6-layer periodic residual CNN (32 channels, 3x3x3 taps).

    Parameters live in one flat FP32 device vector (Adam updates it in one
    launch); the 32 -> 32 weights are also kept as bf16 copies for the MFMA
    kernels (forward and transposed/mirrored for backward-data)."""

    def __init__(self, dil=(1, 2, 4, 4, 2, 1), c=0.7, sv=1.0, seed=0, device="cuda"):
        _lib.require_gpu()
        self.dil = tuple(int(d) for d in dil)
        self.c, self.sv = float(c), float(sv)
        self.device = torch.device(device)
        self.offsets = {}
        off = 0
        for name, shp in LAYER_SHAPES:
            n = int(np.prod(shp))
            self.offsets[name] = (off, shp)
            off += n
        self.nparam = off
        self.flat = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.grad = torch.zeros_like(self.flat)
        self.adam_m = torch.zeros_like(self.flat)
        self.adam_v = torch.zeros_like(self.flat)
        self.step = 0
        self.wf = [torch.empty(NTAP * C * C, dtype=torch.bfloat16, device=self.device) for _ in range(4)]
        self.wb = [torch.empty_like(w) for w in self.wf]
        self._bufs = {}
        self._init(seed)

    # parameters ----------------------------------------------------------
    def param(self, name, which=None):
        off, shp = self.offsets[name]
        src = self.flat if which is None else which
        return src[off:off + int(np.prod(shp))].view(shp)

    def _init(self, seed):
        rng = np.random.default_rng(seed)
        p = {}
        p["W1"] = rng.normal(0, np.sqrt(2.0 / (NTAP * 2)), (NTAP, 2, C))
        for l in range(2, 6):
            p["W%d" % l] = rng.normal(0, 0.5 * np.sqrt(2.0 / (NTAP * C)), (NTAP, C, C))
        p["W6"] = rng.normal(0, 0.1 * np.sqrt(1.0 / (NTAP * C)), (NTAP, C))
        self.set_params(p)

    def set_params(self, p):
        for name, shp in LAYER_SHAPES:
            v = p.get(name)
            self.param(name).copy_(torch.zeros(shp) if v is None else torch.as_tensor(np.asarray(v, np.float32)))
        self._prep()

    def params(self):
        return {name: self.param(name).cpu().numpy().astype(np.float64) for name, _ in LAYER_SHAPES}

    def _prep(self):
        s = _lib.stream()
        for i, l in enumerate(range(2, 6)):
            _lib.check(_lib.lib.ofdft_cnn_prep_weights(self.param("W%d" % l).data_ptr(), self.wf[i].data_ptr(),
                                                       self.wb[i].data_ptr(), s))

    def save(self, path, **meta):
        np.savez(path, flat=self.flat.cpu().numpy(), dil=np.array(self.dil), c=self.c, sv=self.sv,
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device="cuda"):
        z = np.load(path, allow_pickle=False)
        net = cls(dil=tuple(int(d) for d in z["dil"]), c=float(z["c"]), sv=float(z["sv"]), device=device)
        flat = torch.as_tensor(z["flat"].astype(np.float32), device=net.device)
        if flat.numel() != net.nparam:
            raise ValueError("weight file has %d parameters, the network %d" % (flat.numel(), net.nparam))
        net.flat.copy_(flat)
        net._prep()
        return net

    # buffers ----------------------------------------------------------------
    def _buffers(self, nr, train):
        key = (tuple(int(n) for n in nr), bool(train))
        if key in self._bufs:
            return self._bufs[key]
        n = int(np.prod(nr))
        dev = self.device
        bf = lambda: torch.empty((n, C), dtype=torch.bfloat16, device=dev)  # noqa: E731
        b = {"out": torch.empty(n, dtype=torch.float32, device=dev)}
        if train:
            b["h"] = [bf() for _ in range(5)]  # h1..h5
            b["r"] = [bf() for _ in range(4)]  # r2..r5
            b["dh"] = torch.empty((n, C), dtype=torch.float32, device=dev)
            b["g"] = [bf(), bf()]
            b["dout"] = torch.empty(n, dtype=torch.float32, device=dev)
            b["nblk"] = 512
            kmax = max(int(_lib.lib.ofdft_cnn_wgrad_partials(layer, 512)) for layer in range(3))
            b["partials"] = torch.empty(kmax, dtype=torch.float32, device=dev)
            b["loss"] = torch.zeros(1, dtype=torch.float64, device=dev)
        else:
            b["h"] = [bf(), bf()]
        self._bufs = {k: v for k, v in self._bufs.items() if k[1] != train}  # one grid shape per mode
        self._bufs[key] = b
        return b

    # forward / backward -------------------------------------------------------
    def forward(self, x0, nr, train=False):
        """x0: FP32 device tensor (2, N).  Returns the output field (N,) (a
        buffer owned by the network, overwritten by the next call)."""
        nr_c = np.ascontiguousarray(nr, dtype=np.int32)
        nrp = _lib.iptr(nr_c)
        s = _lib.stream()
        b = self._buffers(nr, train)
        L = _lib.lib
        h = b["h"]
        _lib.check(L.ofdft_cnn_conv_in(nrp, self.dil[0], x0.data_ptr(), self.param("W1").data_ptr(),
                                       self.param("b1").data_ptr(), h[0].data_ptr(), s))
        for i, l in enumerate(range(2, 6)):
            if train:
                hin, hout, r = h[i], h[i + 1], b["r"][i].data_ptr()
            else:  # inference keeps no ReLU outputs
                hin, hout, r = h[i % 2], h[(i + 1) % 2], None
            _lib.check(L.ofdft_cnn_conv_mid(nrp, self.dil[l - 1], 0, hin.data_ptr(), self.wf[i].data_ptr(),
                                            self.param("b%d" % l).data_ptr(), None, None, hout.data_ptr(), r, s))
        hlast = h[4] if train else h[0]
        _lib.check(L.ofdft_cnn_conv_out(nrp, self.dil[5], hlast.data_ptr(), self.param("W6").data_ptr(),
                                        self.param("b6").data_ptr(), b["out"].data_ptr(), None, s))
        return b["out"]

    def backward(self, x0, nr, dout):
        """Gradients of sum(dout * out) for the last training forward; fills
        self.grad."""
        nr_c = np.ascontiguousarray(nr, dtype=np.int32)
        nrp = _lib.iptr(nr_c)
        s = _lib.stream()
        b = self._buffers(nr, True)
        L = _lib.lib
        h, r, g, dh, part, nblk = b["h"], b["r"], b["g"], b["dh"], b["partials"], b["nblk"]
        gp = lambda name: self.param(name, self.grad).data_ptr()  # noqa: E731
        _lib.check(L.ofdft_cnn_wgrad(nrp, self.dil[5], 2, dout.data_ptr(), h[4].data_ptr(), part.data_ptr(), nblk,
                                     gp("W6"), s))
        _lib.check(L.ofdft_cnn_out_backward(nrp, self.dil[5], dout.data_ptr(), self.param("W6").data_ptr(),
                                            r[3].data_ptr(), dh.data_ptr(), g[0].data_ptr(), s))
        cur = 0
        for l in range(5, 1, -1):
            i = l - 2
            _lib.check(L.ofdft_cnn_wgrad(nrp, self.dil[l - 1], 1, g[cur].data_ptr(), h[i].data_ptr(), part.data_ptr(),
                                         nblk, gp("W%d" % l), s))
            mask = r[i - 1] if l > 2 else h[0]
            _lib.check(L.ofdft_cnn_conv_mid(nrp, self.dil[l - 1], 1, g[cur].data_ptr(), self.wb[i].data_ptr(), None,
                                            dh.data_ptr(), mask.data_ptr(), dh.data_ptr(), g[1 - cur].data_ptr(), s))
            cur = 1 - cur
        _lib.check(L.ofdft_cnn_wgrad(nrp, self.dil[0], 0, g[cur].data_ptr(), x0.data_ptr(), part.data_ptr(), nblk,
                                     gp("W1"), s))

    def loss(self, plan, out, target, dout, loss, w=None):
        """Mean squared error (w None) or the spectral loss with weights w."""
        s = _lib.stream()
        if w is None:
            _lib.check(_lib.lib.ofdft_cnn_loss(plan.handle, out.data_ptr(), target.data_ptr(), dout.data_ptr(),
                                               loss.data_ptr(), s))
        else:
            _lib.check(_lib.lib.ofdft_cnn_loss_spectral(plan.handle, out.data_ptr(), target.data_ptr(), w.data_ptr(),
                                                        dout.data_ptr(), loss.data_ptr(), s))

    def train_step(self, plan, x0, nr, target, lr, b1=0.9, b2=0.999, eps=1e-8, w=None):
        """One Adam step on the loss of the output vs target (mean squared
        error, or the spectral loss with weights w); returns the loss as a
        device tensor (no host sync)."""
        b = self._buffers(nr, True)
        out = self.forward(x0, nr, train=True)
        s = _lib.stream()
        self.loss(plan, out, target, b["dout"], b["loss"], w)
        self.backward(x0, nr, b["dout"])
        self.step += 1
        _lib.check(_lib.lib.ofdft_cnn_adam(self.nparam, self.flat.data_ptr(), self.grad.data_ptr(),
                                           self.adam_m.data_ptr(), self.adam_v.data_ptr(), float(lr), b1, b2, eps,
                                           self.step, s))
        self._prep()
        return b["loss"]

    # start density ---------------------------------------------------------
    def features(self, plan, rho_s, vloc, rho0, out=None):
        n = rho_s.numel()
        x0 = out if out is not None else torch.empty((2, n), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.ofdft_cnn_features(plan.handle, rho_s.data_ptr(), vloc.data_ptr(), float(rho0), self.c,
                                               self.sv, x0.data_ptr(), _lib.stream()))
        return x0

    def start_density(self, grid, x0, rho0, ne, use_net=True):
        """rho0 exp(x0[0] + delta) normalised to ne, as a DirectField on grid."""
        plan = grid.plan(grid.dtype)
        delta = self.forward(x0, grid.nrR) if use_net else None
        rho = torch.empty(tuple(int(n) for n in grid.nrR), dtype=grid.dtype, device=grid.device)
        _lib.check(_lib.lib.ofdft_cnn_apply(plan.handle, x0.data_ptr(), None if delta is None else delta.data_ptr(),
                                            float(rho0), float(ne), rho.data_ptr(), _lib.stream()))
        return DirectField(grid, data=rho)


def learned_start(grid, ions, pseudo, net=None, use_net=True):
    """Start density for ``Optimization.optimize_rho`` on ``grid`` from the
    superposition features and the network (``use_net=False`` gives the
    superposition start alone).  ``pseudo`` is the LocalPseudo of the run."""
    net = net or WarmStartNet.load()
    ne = float(ions.get_ncharges())
    rho0 = ne / grid.volume
    # the features in FP32 on an FP32 target grid (the network's input
    # precision), else in FP64 on the pseudopotential's grid
    if grid.dtype == torch.float32:
        gs = grid
    else:
        gs = pseudo.grid if pseudo.grid.dtype == torch.float64 else pseudo.grid.with_dtype(torch.float64)
    rho_s = superposition_density(gs, ions, pseudo.readpp, rho0)
    vloc = pseudo.vreal_for(gs).data
    x0 = net.features(gs.plan(gs.dtype), rho_s, vloc, rho0)
    return net.start_density(grid, x0, rho0, ne, use_net=use_net)


def time_network(sizes=(96, 160), train_size=96, reps=5):
    """HIP-event timings of the network: inference at each size and one
    training step (forward, backward, weight gradients, Adam) at train_size,
    random inputs and weights.  Returns {name: milliseconds}."""
    out = {}
    net = WarmStartNet()
    for n in sizes:
        nr = (n, n, n)
        x = torch.randn(2, n ** 3, device=net.device)
        net.forward(x, nr)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            net.forward(x, nr)
        e1.record()
        torch.cuda.synchronize()
        out["inference_%d" % n] = e0.elapsed_time(e1) / reps
        if n == train_size:
            from .grid import DirectGrid
            g = DirectGrid(np.eye(3) * 0.24 * n, nr, full=False, dtype=torch.float32)
            plan = g.plan(torch.float32)
            t = torch.randn(n ** 3, device=net.device) * 0.1
            net.train_step(plan, x, nr, t, 1e-3)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                net.train_step(plan, x, nr, t, 1e-3)
            e1.record()
            torch.cuda.synchronize()
            out["train_step_%d" % n] = e0.elapsed_time(e1) / reps
    return out


if __name__ == "__main__":
    # python -m dftpy_amd.learned: timings of the network (see time_network)
    for k, ms in time_network().items():
        n = int(k.rsplit("_", 1)[1])
        flops = 2 * NTAP * n ** 3 * (2 * C + 4 * C * C + C) * (3 if k.startswith("train") else 1)
        print("%-16s %8.3f ms  %6.1f TFLOP/s" % (k, ms, flops / ms / 1e9), flush=True)
