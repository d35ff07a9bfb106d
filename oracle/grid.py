"""Periodic grid and FFT conventions, restating ``grid.py`` and ``field.py``.

Real fields are (nx, ny, nz) C-order; their spectra are the R2C half grid
(nx, ny, nz//2+1) (``grid.py:45-47``).  Forward transform = rfftn * dV
(``field.py:373-390``), inverse = irfftn(s=nr) / dV (``field.py:739-764``).
"""
import numpy as np


def best_fft_size(N, max_prime=13, scale=0.99, even=True):
    """``mpi/mp_serial.py:17-90``."""
    prime_factors_one = [11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]
    a = int(np.ceil(np.log2(N))) + 1
    b = int(np.ceil(np.log(N) / np.log(3))) + 1
    c = int(np.ceil(np.log(N) / np.log(5))) + 1
    d = int(np.ceil(np.log(N) / np.log(7))) + 1
    istart = 1 if even else 0
    if max_prime == 2:
        arr0 = 2 ** np.arange(istart, a)
    elif max_prime == 3:
        m = np.mgrid[istart:a, :b].reshape(2, -1)
        arr0 = 2 ** m[0] * 3 ** m[1]
    elif max_prime == 5:
        m = np.mgrid[istart:a, :b, :c].reshape(3, -1)
        arr0 = 2 ** m[0] * 3 ** m[1] * 5 ** m[2]
    else:
        m = np.mgrid[istart:a, :b, :c, :d].reshape(4, -1)
        arr0 = 2 ** m[0] * 3 ** m[1] * 5 ** m[2] * 7 ** m[3]
    if N < 100:
        arr1 = arr0[np.logical_and(arr0 > N / (max_prime + 1), arr0 < 2 * N)]
    else:
        arr1 = arr0[np.logical_and(arr0 > N / (max_prime + 1), arr0 < 1.2 * N)]
    arr = list(arr1)
    for item in prime_factors_one:
        if max_prime >= item:
            arr.extend(arr1 * item)
    arr = np.asarray(arr)
    return int(np.min(arr[arr > np.ceil(scale * N) - 1]))


def ecut2nr(lattice, spacing):
    """``math_utils.py:349-365`` with a given spacing and optfft=True."""
    metric = np.dot(lattice, lattice.T)
    nr = np.zeros(3, dtype=np.int64)
    for i in range(3):
        n = int(np.sqrt(metric[i, i]) / spacing)
        n += n % 2
        nr[i] = best_fft_size(n)
    return nr


class Grid:
    """A periodic direct grid with its reciprocal (``grid.py:26-57,291-734``)."""

    def __init__(self, lattice, nr, full=False):
        self.full = bool(full)
        self.lattice = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
        self.nr = np.asarray(nr, dtype=np.int64).reshape(3)
        self.nnr = int(np.prod(self.nr))
        self.volume = abs(np.linalg.det(self.lattice))
        self.dV = self.volume / self.nnr
        # reciprocal lattice rows b_i, a_i.b_j = 2 pi delta_ij (grid.py:405-414)
        self.B = (2 * np.pi * np.linalg.inv(self.lattice)).T
        self._g = None
        self._gg = None
        self._mask = None
        self._r = None

    @property
    def nrG(self):
        if self.full:
            return self.nr.copy()
        return np.array([self.nr[0], self.nr[1], self.nr[2] // 2 + 1])

    def freqs(self, full=False):
        ax = [np.fft.fftfreq(self.nr[0], d=1.0 / self.nr[0]),
              np.fft.fftfreq(self.nr[1], d=1.0 / self.nr[1])]
        if full:
            ax.append(np.fft.fftfreq(self.nr[2], d=1.0 / self.nr[2]))
        else:
            ax.append(np.fft.rfftfreq(self.nr[2], d=1.0 / self.nr[2]))
        return ax

    def g_vectors(self, full=False):
        """``grid.py:584-620``: g = sum_j k_j b_j."""
        S = np.asarray(np.meshgrid(*self.freqs(full), indexing="ij"))
        return np.einsum("j...,jk->k...", S, self.B)

    @property
    def g(self):
        if self._g is None:
            self._g = self.g_vectors(self.full)
        return self._g

    @property
    def gg(self):
        if self._gg is None:
            self._gg = np.einsum("lijk,lijk->ijk", self.g, self.g)
        return self._gg

    @property
    def q(self):
        return np.sqrt(self.gg)

    @property
    def invgg(self):
        """``grid.py:531-541``: 1/G^2 with G=0 set to 0."""
        gg = self.gg.copy()
        gg[0, 0, 0] = 1.0
        inv = 1.0 / gg
        inv[0, 0, 0] = 0.0
        return inv

    @property
    def ggF(self):
        gF = self.g_vectors(True)
        return np.einsum("lijk,lijk->ijk", gF, gF)

    @property
    def mask(self):
        """Half-space deduplication mask (``grid.py:657-695``, serial form)."""
        if self._mask is None:
            nr = self.nr
            D = nr // 2
            Dm = nr % 2
            mask = np.ones(self.nrG, dtype=bool)
            if self.full:
                mask[:, :, D[2] + 1:] = False
            mask[0, D[1] + 1:, 0] = False
            mask[D[0] + 1:, :, 0] = False
            if Dm[2] == 0:
                mask[0, 0, D[2]] = False
                mask[0, D[1] + 1:, D[2]] = False
                mask[D[0] + 1:, :, D[2]] = False
                if Dm[1] == 0:
                    mask[0, D[1], D[2]] = False
                if Dm[0] == 0:
                    mask[D[0], 0, D[2]] = False
                    mask[D[0], D[1] + 1:, D[2]] = False
            if Dm[0] == 0:
                mask[D[0], D[1] + 1:, 0] = False
                if Dm[1] == 0:
                    mask[D[0], D[1], 0] = False
            if Dm[1] == 0:
                mask[0, D[1], 0] = False
            if np.all(Dm == 0):
                mask[D[0], D[1], D[2]] = False
            self._mask = mask
        return self._mask

    @property
    def r(self):
        """Cartesian coordinates of the grid points (``grid.py:355-372``)."""
        if self._r is None:
            ax = [np.linspace(0, 1, n, endpoint=False) for n in self.nr]
            S = np.asarray(np.meshgrid(*ax, indexing="ij"))
            self._r = np.einsum("j...,jk->k...", S, self.lattice)
        return self._r

    @property
    def spacings(self):
        return np.sqrt((self.lattice ** 2).sum(axis=1)) / self.nr

    def ecut(self):
        """``grid.py:456-459`` guess_ecut and ``math_utils.py:337-342``."""
        lengths = np.sqrt((self.lattice ** 2).sum(axis=1))
        sp2 = lengths / (self.nr - 1)
        sp = 0.5 * (self.spacings + sp2)
        return np.pi ** 2 / (2 * sp.max() ** 2)

    # --- transforms (field.py) ---
    # FFT threads: None = numpy's pocketfft (one thread, the reference's
    # default); an integer = scipy.fft with that many workers
    workers = None

    def fft(self, f):
        """R2C (full=False) or complex fftn (full=True, ``field.py:159-172``)."""
        if self.workers:
            import scipy.fft as sf
            if self.full:
                return sf.fftn(f, workers=self.workers) * self.dV
            return sf.rfftn(f, workers=self.workers) * self.dV
        if self.full:
            return np.fft.fftn(f) * self.dV
        return np.fft.rfftn(f) * self.dV

    def ifft(self, F):
        """Real field back; on a full grid the real part of ifftn is kept
        (``force_real``, ``field.py:739-764``)."""
        if self.workers:
            import scipy.fft as sf
            if self.full:
                return np.real(sf.ifftn(F, workers=self.workers)) / self.dV
            return sf.irfftn(F, s=tuple(self.nr), axes=(0, 1, 2), workers=self.workers) / self.dV
        if self.full:
            return np.real(np.fft.ifftn(F)) / self.dV
        return np.fft.irfftn(F, s=tuple(self.nr), axes=(0, 1, 2)) / self.dV

    def fft_c(self, f):
        return np.fft.fftn(f) * self.dV

    def ifft_c(self, F):
        return np.fft.ifftn(F) / self.dV

    def integral(self, f):
        return f.sum() * self.dV
