"""Grid-size helpers (math_utils.py:243-254,337-365 and
mpi/mp_serial.py:17-90 of the reference) and the interpolating cubic spline
used for radial tables."""
import numpy as np


def best_fft_size(N, max_prime=13, scale=0.99, even=True):
    """Smallest 2^a 3^b 5^c 7^d (times at most one of 11..97 up to max_prime)
    size >= ceil(scale*N) - 1 (FFTW-friendly sizes, ``mp_serial.py:17-90``)."""
    one = [11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]
    a = int(np.ceil(np.log2(N))) + 1
    b = int(np.ceil(np.log(N) / np.log(3))) + 1
    c = int(np.ceil(np.log(N) / np.log(5))) + 1
    d = int(np.ceil(np.log(N) / np.log(7))) + 1
    lo = 1 if even else 0
    if max_prime == 2:
        cand = 2 ** np.arange(lo, a)
    elif max_prime == 3:
        e = np.mgrid[lo:a, :b].reshape(2, -1)
        cand = 2 ** e[0] * 3 ** e[1]
    elif max_prime == 5:
        e = np.mgrid[lo:a, :b, :c].reshape(3, -1)
        cand = 2 ** e[0] * 3 ** e[1] * 5 ** e[2]
    else:
        e = np.mgrid[lo:a, :b, :c, :d].reshape(4, -1)
        cand = 2 ** e[0] * 3 ** e[1] * 5 ** e[2] * 7 ** e[3]
    hi = 2 * N if N < 100 else 1.2 * N
    base = cand[(cand > N / (max_prime + 1)) & (cand < hi)]
    allc = [base] + [base * p for p in one if p <= max_prime]
    allc = np.concatenate(allc)
    return int(allc[allc > np.ceil(scale * N) - 1].min())


def smooth_fft_size(N, even=True):
    """Smallest 2/3/5-smooth size >= N (the sizes the HIP FFT supports)."""
    n = int(N)
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1 and (not even or n % 2 == 0):
            return n
        n += 1


def spacing2ecut(spacing):
    return np.pi ** 2 / (2 * spacing ** 2)


def ecut2spacing(ecut):
    return np.sqrt(np.pi ** 2 / ecut * 0.5)


def ecut2nr(ecut=None, lattice=None, optfft=True, spacing=None, max_prime=13, **kwargs):
    """``math_utils.py:349-365``; the HIP FFT has radices 2,3,4,5,7,8,11,13
    so every size best_fft_size returns for max_prime <= 13 is supported."""
    lattice = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
    if spacing is None:
        spacing = ecut2spacing(ecut)
    metric = lattice @ lattice.T
    nr = np.zeros(3, dtype=np.int64)
    for i in range(3):
        n = int(np.sqrt(metric[i, i]) / spacing)
        n += n % 2
        nr[i] = best_fft_size(n, max_prime=max_prime, **kwargs) if optfft else n
    return nr


def cubic_table(x, y):
    """Interpolating cubic spline with not-a-knot ends through (x, y) on a
    uniform grid, as per-interval polynomial coefficients
    c[i] = (c0, c1, c2, c3) of c0 + c1 t + c2 t^2 + c3 t^3, t = q - x[i].
    Same interpolant as FITPACK splrep(k=3, s=0), which places the interior
    knots at x[2:-2] (the reference's pseudo/__init__.py:641-658)."""
    from scipy.interpolate import CubicSpline
    cs = CubicSpline(x, y, bc_type="not-a-knot")
    c = cs.c  # (4, n-1), highest power first
    return np.ascontiguousarray(np.stack([c[3], c[2], c[1], c[0]], axis=1))


_LARGE_ETA_SERIES = [-0.13714285714285712, -6.39999999999999875e-2, -3.77825602968460128e-2, -2.51824061652633074e-2,
                     -1.80879839616166146e-2, -1.36715733124818332e-2, -1.07236045520990083e-2,
                     -8.65192783339199453e-3, -7.1372762502456763e-3, -5.9945117538835746e-3, -5.10997527675418131e-3,
                     -4.41060829979912465e-3, -3.84763737842981233e-3, -3.38745061493813488e-3,
                     -3.00624946457977689e-3]


def lindhard(eta, lbda=1.0, mu=1.0):
    """Inverse Lindhard function with the reference's four branches
    (functional/kedf/kernel.py:13-136), for host-side tables; the evaluation
    kernels compute it on the device."""
    eta = np.asarray(eta, dtype=np.float64)
    out = np.empty_like(eta)
    small = eta < 1e-10
    one = np.abs(eta - 1.0) < 1e-10
    large = eta > 3.65
    mid = ~(small | one | large)
    e = eta[mid]
    out[mid] = 1.0 / (0.5 + 0.25 * (1.0 - e * e) * np.log((1.0 + e) / np.abs(1.0 - e)) / e) - 3.0 * mu * e * e - lbda
    e = eta[small]
    out[small] = 1.0 - lbda + e * e * (1.0 / 3.0 - 3.0 * mu)
    e = eta[one]
    out[one] = 2.0 - lbda - 3.0 * mu + 20.0 * (e - 1.0)
    e = eta[large]
    inv = 1.0 / (e * e)
    acc = np.full_like(e, _LARGE_ETA_SERIES[-1])
    for c in _LARGE_ETA_SERIES[-2::-1]:
        acc = c + inv * acc
    out[large] = 3.0 * (1.0 - mu) * e * e - lbda - 0.6 + inv * acc
    return out


CTF = 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)


def lda_dv(rho0):
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


def uniform_response_coefs(rho0, tf=True, vw=True, wt=True, xc=True, hartree=True, x=1.0, y=1.0, alpha=5.0 / 6.0,
                           beta=5.0 / 6.0):
    """uniform_response_kernel as the coefficients the device evaluates:
    K(q) = c[0] + c[1] q^2 + c[2] G(q / c[3]; lambda = c[4]) + c[5] 4 pi / q^2
    (G the inverse Lindhard function)."""
    kc = 0.0
    if tf:
        kc += x * (10.0 / 9.0) * CTF * rho0 ** (-1.0 / 3.0)
    if xc:
        kc += lda_dv(rho0)
    tkf = 2.0 * (3.0 * rho0 * np.pi ** 2) ** (1.0 / 3.0)
    wtc = 0.0
    if wt:
        wtfac = 5.0 / (9.0 * alpha * beta * rho0 ** (alpha + beta - 5.0 / 3.0)) * CTF
        wtc = 2.0 * alpha * beta * rho0 ** (alpha + beta - 2) * wtfac
    return np.array([kc, y / (4.0 * rho0) if vw else 0.0, wtc, tkf, float(x), 1.0 if hartree else 0.0])


def uniform_response_kernel(q, rho0, tf=True, vw=True, wt=True, xc=True, hartree=True, x=1.0, y=1.0,
                            alpha=5.0 / 6.0, beta=5.0 / 6.0):
    """K(q) = d^2 E / d rho(q)^2 at the uniform density rho0 for the chosen
    terms: TF x (10/9) C_TF rho0^(-1/3), vW y q^2 / (4 rho0), WT
    2 alpha beta rho0^(alpha+beta-2) times the WT kernel (kernel.py:325-334),
    PZ LDA dv/drho, Hartree 4 pi / q^2 (infinite at q = 0)."""
    q = np.asarray(q, dtype=np.float64)
    k = np.zeros_like(q)
    if tf:
        k = k + x * (10.0 / 9.0) * CTF * rho0 ** (-1.0 / 3.0)
    if vw:
        k = k + y * q * q / (4.0 * rho0)
    if wt:
        tkf = 2.0 * (3.0 * rho0 * np.pi ** 2) ** (1.0 / 3.0)
        wtfac = 5.0 / (9.0 * alpha * beta * rho0 ** (alpha + beta - 5.0 / 3.0)) * CTF
        k = k + 2.0 * alpha * beta * rho0 ** (alpha + beta - 2) * lindhard(q / tkf, x, 1.0) * wtfac
    if hartree:
        with np.errstate(divide="ignore"):
            k = k + np.where(q > 0, 4.0 * np.pi / np.maximum(q, 1e-300) ** 2, np.inf)
    if xc:
        k = k + lda_dv(rho0)
    return k

