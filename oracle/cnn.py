"""FP64 restatement of the learned warm-start network (test infrastructure).

The network has no counterpart in the reference: it supplies the start
density that ``Optimization.optimize_rho`` (``optimization.py:281-520``)
then refines, in place of the reference's guess densities
(``density/density.py:199-207,258-317``).  This module is the checker for
the HIP kernels of ``dftpy_amd/csrc/cnn.hip``.

Architecture (periodic 3D CNN on the density grid, channels-last fields
``(nx, ny, nz, C)``, C = 32, 3x3x3 taps with per-layer dilation d):

    h1 = relu(conv_d1(x0; W1) + b1)                 x0: 2 input channels
    h_l = h_{l-1} + relu(conv_dl(h_{l-1}; W_l) + b_l)   l = 2..5 (residual)
    out = conv_d6(h5; W6) + b6                      1 output channel

Tap t = (i*3 + j)*3 + k with (i, j, k) in {0,1,2}^3 reads the input at
p + d*(i-1, j-1, k-1) (periodic wrap).  W1[t, c, co], W_l[t, co, ci],
W6[t, ci].

``bf16=True`` rounds to bfloat16 (round to nearest even) exactly where the
HIP path stores bfloat16: the mid-layer weights, the hidden activations
h_l and the relu outputs r_l, and the gradients fed to the mid-layer
products; and to FP16 where the HIP path's forward uses FP16: the matrix
operands of the first layer (input features and W1) and of the output
projection (h5 and W6), and the 27 projected tap planes of the output layer.
Everything else is FP64 here (FP32 accumulation on the GPU).

Pinned by: central finite differences of the loss against ``backward``
in pure FP64 (``tests/test_oracle.py``).
"""
import numpy as np

C = 32
NTAP = 27


def to_bf16(a):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def to_f16(a):
    """Round to IEEE half precision (nearest even) and return as float64."""
    return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


def taps(d):
    return [(d * (i - 1), d * (j - 1), d * (k - 1)) for i in range(3) for j in range(3) for k in range(3)]


def shift(x, off):
    """x[p + off] with periodic wrap, for channels-last or scalar fields."""
    return np.roll(x, shift=(-off[0], -off[1], -off[2]), axis=(0, 1, 2))


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    p["W1"] = rng.normal(0, scale * np.sqrt(2.0 / (NTAP * 2)), (NTAP, 2, C))
    p["b1"] = np.zeros(C)
    for l in range(2, 6):
        p["W%d" % l] = rng.normal(0, scale * 0.5 * np.sqrt(2.0 / (NTAP * C)), (NTAP, C, C))
        p["b%d" % l] = np.zeros(C)
    p["W6"] = rng.normal(0, scale * 0.1 * np.sqrt(1.0 / (NTAP * C)), (NTAP, C))
    p["b6"] = np.zeros(1)
    return p


def forward(p, x0, dil=(1, 1, 1, 1, 1, 1), bf16=False):
    """x0: (2, nx, ny, nz).  Returns (out (nx,ny,nz), cache)."""
    rb = to_bf16 if bf16 else (lambda a: a)
    rh = to_f16 if bf16 else (lambda a: a)
    T1 = taps(dil[0])
    xin, W1 = rh(x0), rh(p["W1"])
    pre = np.zeros(x0.shape[1:] + (C,))
    for t, off in enumerate(T1):
        for c in range(2):
            pre += shift(xin[c], off)[..., None] * W1[t, c][None, None, None, :]
    h = rb(np.maximum(pre + p["b1"], 0.0))
    hs, rs = [h], [h]
    for l in range(2, 6):
        W = rb(p["W%d" % l])
        acc = np.zeros_like(h)
        for t, off in enumerate(taps(dil[l - 1])):
            acc += shift(h, off) @ W[t].T
        r = np.maximum(acc + p["b%d" % l], 0.0)
        h = rb(h + r)
        hs.append(h)
        rs.append(rb(r))
    out = np.zeros(x0.shape[1:])
    h16, W6 = rh(h), rh(p["W6"])
    for t, off in enumerate(taps(dil[5])):
        # tap plane u_t = h @ W6[t] at every point, then shifted
        out += shift(rh(h16 @ W6[t]), off)
    out += p["b6"][0]
    return out, dict(hs=hs, rs=rs, x0=x0, dil=dil, bf16=bf16)


def loss(out, target):
    n = out.size
    d = out - target
    return float((d * d).sum() / n), 2.0 * d / n


def backward(p, cache, dout):
    """Gradients of sum(dout * out) with respect to every parameter."""
    rb = to_bf16 if cache["bf16"] else (lambda a: a)
    hs, rs, x0, dil = cache["hs"], cache["rs"], cache["x0"], cache["dil"]
    g = {}
    h5 = hs[-1]
    T6 = taps(dil[5])
    g["W6"] = np.stack([np.einsum("xyz,xyzc->c", dout, shift(h5, off)) for off in T6])
    g["b6"] = np.array([dout.sum()])
    dh = np.zeros_like(h5)
    for t, off in enumerate(T6):
        dh += shift(dout, tuple(-o for o in off))[..., None] * p["W6"][t][None, None, None, :]
    for l in range(5, 1, -1):
        gl = rb(np.where(rs[l - 1] > 0, dh, 0.0))
        W = rb(p["W%d" % l])
        hprev = hs[l - 2]
        T = taps(dil[l - 1])
        g["W%d" % l] = np.stack([np.einsum("xyzo,xyzi->oi", gl, shift(hprev, off)) for off in T])
        g["b%d" % l] = gl.sum(axis=(0, 1, 2))
        acc = np.zeros_like(dh)
        for t, off in enumerate(T):
            acc += shift(gl, tuple(-o for o in off)) @ W[t]
        dh = dh + acc
    g1 = rb(np.where(hs[0] > 0, dh, 0.0))
    T1 = taps(dil[0])
    g["W1"] = np.stack([np.stack([np.einsum("xyzo,xyz->o", g1, shift(x0[c], off)) for c in range(2)]) for off in T1])
    g["b1"] = g1.sum(axis=(0, 1, 2))
    return g
