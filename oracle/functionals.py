"""Energy/potential functionals, FP64, restating the reference's KEDFs,
Hartree, PZ-LDA and the local pseudopotential energy term.

Every function takes a real density array ``rho`` (C-order, grid.nr) and a
:class:`oracle.grid.Grid`; it returns ``(energy, potential)``.
"""
import numpy as np

from .units import ZERO

CTF = 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)


# ---------------------------------------------------------------- TF
def tf(rho, grid, x=1.0):
    """``functional/kedf/tf.py:13-35,65-80``."""
    edens = np.cbrt(rho ** 5) * CTF           # PowerInt(rho, 5, 3)
    pot = (5.0 / 3.0) * CTF * np.cbrt(rho * rho)
    return edens.sum() * grid.dV * x, pot * x


# ---------------------------------------------------------------- vW
def vw(rho, grid, y=1.0, sigma=None):
    """``functional/kedf/vw.py:29-53,84-94,114-131``."""
    tol = 1e-30
    rhom = rho.copy()
    rhom[~(rho > 0)] = tol
    sq = np.sqrt(rhom)
    gg = grid.gg
    G = grid.fft(sq) * gg
    if sigma is not None:
        G = G * np.exp(-gg * sigma ** 2 / 4.0)
    a = grid.ifft(G) * 0.5
    sq[np.abs(sq) < tol] = tol
    pot = a / sq
    m = rho > 0
    ene = np.einsum("i,i->", rho[m], pot[m]) * grid.dV
    return ene * y, pot * y


# ---------------------------------------------------------------- Lindhard / WT
def lindhard(eta, lbda, mu, tol=1e-10):
    """Inverse Lindhard function with its four branches
    (``functional/kedf/kernel.py:13-136``).  Later branches win, in the
    reference's assignment order: normal < zero < one < large."""
    eta = np.asarray(eta, dtype=np.float64)
    tol2 = tol / 10
    out = eta.copy()
    c0 = (eta > tol2) & (np.abs(eta - 1.0) > tol2)
    c1 = eta < tol
    c2 = np.abs(eta - 1.0) < tol
    c3 = eta > 3.65
    e = eta[c0]
    out[c0] = 1.0 / (0.5 + 0.25 * (1.0 - e ** 2) * np.log((1.0 + e) / np.abs(1.0 - e)) / e) - 3.0 * mu * e ** 2 - lbda
    e = eta[c1]
    out[c1] = 1.0 - lbda + e ** 2 * (1.0 / 3.0 - 3.0 * mu)
    e = eta[c2]
    out[c2] = 2.0 - lbda - 3.0 * mu + 20.0 * (e - 1.0)
    e = eta[c3]
    out[c3] = lindhard_large(e, lbda, mu)
    return out


LARGE_COEFS = [-0.13714285714285712, -6.39999999999999875e-2, -3.77825602968460128e-2,
               -2.51824061652633074e-2, -1.80879839616166146e-2, -1.36715733124818332e-2,
               -1.07236045520990083e-2, -8.65192783339199453e-3, -7.1372762502456763e-3,
               -5.9945117538835746e-3, -5.10997527675418131e-3, -4.41060829979912465e-3,
               -3.84763737842981233e-3, -3.38745061493813488e-3, -3.00624946457977689e-3]


def lindhard_large(eta, lbda, mu):
    """Large-eta asymptotic series (``kernel.py:72-136``), Horner form."""
    inv = 1.0 / eta ** 2
    s = LARGE_COEFS[-1]
    for c in LARGE_COEFS[-2::-1]:
        s = c + inv * s
    return 3.0 * (1.0 - mu) * eta ** 2 - lbda - 0.6 + inv * s


def wt_kernel(q, rho0, x=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0):
    """``kernel.py:325-334``: Lindhard(q/2kF; x, 1) * 5 C_TF / (9 a b rho0^(a+b-5/3))."""
    tkf = 2.0 * (3.0 * rho0 * np.pi ** 2) ** (1.0 / 3.0)
    factor = 5.0 / (9.0 * alpha * beta * rho0 ** (alpha + beta - 5.0 / 3.0)) * CTF
    return lindhard(q / tkf, x, 1.0) * factor


def wt(rho, grid, rho0=None, x=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0, kernel=None, calc_e=True):
    """WT nonlocal part (``functional/kedf/wt.py:12-56,108-132``).  Returns
    (energy, potential, kernel).  alpha == beta: v = 2 alpha rho^(beta-1)
    IFFT(K FFT(rho^beta)) and E = sum v rho / (2 alpha) dV (``:18-22,124-125``);
    otherwise v = alpha rho^(alpha-1) IFFT(K FFT(rho^beta)) + beta rho^(beta-1)
    IFFT(K FFT(rho^alpha)) (``:23-25``) and E = sum rho^alpha IFFT(K
    FFT(rho^beta)) dV (``:44-56``), rho clamped at 1e-30 in both."""
    if rho0 is None:
        rho0 = rho.mean()
    if kernel is None:
        kernel = wt_kernel(grid.q, rho0, x, alpha, beta)
    r = rho.copy()
    r[r < ZERO] = ZERO
    rb = r ** beta
    cb = grid.ifft(grid.fft(rb) * kernel)
    if abs(beta - alpha) < 1e-9:
        pot = 2.0 * alpha * (rb / r) * cb
        ene = (pot * rho / (2 * alpha)).sum() * grid.dV if calc_e else 0.0
    else:
        ra = r ** alpha
        ca = grid.ifft(grid.fft(ra) * kernel)
        pot = alpha * r ** (alpha - 1.0) * cb + beta * r ** (beta - 1.0) * ca
        ene = (ra * cb).sum() * grid.dV if calc_e else 0.0
    return ene, pot, kernel


# ---------------------------------------------------------------- Hartree
def hartree(rho, grid):
    """``functional/hartree.py:22-45``."""
    v = grid.ifft(grid.fft(rho) * grid.invgg * 4 * np.pi)
    return np.einsum("ijk,ijk->", v, rho) * grid.dV / 2.0, v


# ---------------------------------------------------------------- LDA
PZ = dict(a=0.0311, b=-0.048, c=0.0020, d=-0.0116, gamma=-0.1423, beta1=1.0529, beta2=0.3334)


def lda(rho, grid):
    """Perdew-Zunger LDA, unpolarised (``functional/xc/semilocal_xc.py:252-290``)."""
    a, b, c, d = PZ["a"], PZ["b"], PZ["c"], PZ["d"]
    g, b1, b2 = PZ["gamma"], PZ["beta1"], PZ["beta2"]
    rc = np.cbrt(rho)
    rc[rc < 1e-30] = 1e-30
    Rs = np.cbrt(3.0 / (4.0 * np.pi)) / rc
    r1 = Rs < 1
    r2 = Rs >= 1
    sq = np.sqrt(Rs[r2])
    ex = -3.0 / 4.0 * np.cbrt(3.0 / np.pi) * rc
    ex[r1] += a * np.log(Rs[r1]) + b + c * Rs[r1] * np.log(Rs[r1]) + d * Rs[r1]
    ex[r2] += g / (1.0 + b1 * sq + b2 * Rs[r2])
    ene = np.einsum("ijk,ijk->", ex, rho) * grid.dV
    pot = np.cbrt(-3.0 / np.pi) * rc
    pot[r1] += np.log(Rs[r1]) * (a + 2.0 / 3 * c * Rs[r1]) + b - 1.0 / 3 * a + 1.0 / 3 * (2 * d - c) * Rs[r1]
    pot[r2] += (g + (7.0 / 6.0 * g * b1) * sq + (4.0 / 3.0 * g * b2 * Rs[r2])) / (1.0 + b1 * sq + b2 * Rs[r2]) ** 2
    return ene, pot


def lda_v2(rho):
    """d V_xc / d rho (``semilocal_xc.py:291-303``)."""
    a, c, d = PZ["a"], PZ["c"], PZ["d"]
    g, b1, b2 = PZ["gamma"], PZ["beta1"], PZ["beta2"]
    rc = np.cbrt(rho)
    Rs = np.cbrt(3.0 / (4.0 * np.pi)) / rc
    r1 = Rs < 1
    r2 = Rs >= 1
    fx = -np.cbrt(3.0 / np.pi) / 3.0 * rc / rho
    fc = np.empty(np.shape(rho))
    fc[r1] = -a / 3.0 - (c / 9.0 * (np.log(Rs[r1]) * 2.0 + 1.0) + d * 2.0 / 9.0) * Rs[r1]
    ta = b1 * np.sqrt(Rs[r2])
    tb = b2 * Rs[r2]
    de = 1.0 + ta + tb
    fc[r2] = g / 36.0 * (5.0 * ta + 7.0 * ta * ta + 8.0 * tb + 16.0 * tb * tb + 21.0 * ta * tb) / de ** 3
    fc /= rho
    return fx + fc


# ---------------------------------------------------------------- 'D' and 'V2'
def energy_density(rho, grid, terms=("TF", "vW", "NL"), x=1.0, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0,
                   kernel=None):
    """calcType 'D' of the KEDF terms: TF C rho^(5/3) (``kedf/tf.py:26-35,
    68-73``, without x), vW y v rho (``kedf/vw.py:125-126``), WT v rho / (2
    alpha) for alpha == beta (``kedf/wt.py:123-129``), else rho^alpha (K *
    rho^beta) (``:44-56``).  Hartree and the local PP have none."""
    d = np.zeros_like(rho)
    if "TF" in terms:
        d += np.cbrt(rho ** 5) * CTF
    if "vW" in terms:
        d += vw(rho, grid, y=y)[1] * rho
    if "NL" in terms:
        if abs(alpha - beta) < 1e-9:
            d += wt(rho, grid, alpha=alpha, beta=beta, kernel=kernel)[1] * rho / (2 * alpha)
        else:
            r = np.maximum(rho, ZERO)
            K = wt_kernel(grid.q, rho.mean(), alpha=alpha, beta=beta) if kernel is None else kernel
            d += r ** alpha * grid.ifft(grid.fft(r ** beta) * K)
    return d


def tf_v2(rho, x=1.0):
    """``kedf/tf.py:47-50,77-79``: x 10/9 C rho^(-1/3)."""
    return x * 10.0 / 9.0 * CTF / np.cbrt(rho)


# ---------------------------------------------------------------- total
def total(rho, grid, vreal, kernel, rho0=None, ewald=0.0, kedf="WT", calc_e=True):
    """TF + vW + WT(nl) + PZ-LDA + Hartree + local PP + Ewald
    (``functional/total_functional.py:115-146`` over the ``KEDF('WT')`` engine
    list ``functional/kedf/__init__.py:27-69``).  Returns (energy, potential,
    parts dict)."""
    parts = {}
    e, v = tf(rho, grid)
    parts["TF"] = e
    pot = v
    e, v = vw(rho, grid)
    parts["vW"] = e
    pot = pot + v
    if kedf == "WT":
        e, v, _ = wt(rho, grid, rho0=rho0, kernel=kernel, calc_e=calc_e)
        parts["NL"] = e
        pot = pot + v
    e, v = lda(rho, grid)
    parts["XC"] = e
    pot = pot + v
    e, v = hartree(rho, grid)
    parts["H"] = e
    pot = pot + v
    if vreal is not None:
        parts["PP"] = np.einsum("ijk,ijk->", vreal, rho) * grid.dV
        pot = pot + vreal
    parts["II"] = ewald
    ene = sum(parts.values())
    return ene, pot, parts
