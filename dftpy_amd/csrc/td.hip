// TD-OFDFT Crank-Nicolson step (td/propagator/crank_nicholson.py:44-56,
// 147-188): solve (1 + i H dt/2) psi1 = (1 - i H dt/2) psi0 with the
// complex-symmetric CG of linear_solver.py:24-47 (sums without conjugation,
// norms with it), H psi = -1/2 lap_sigma psi + v psi (td/hamiltonian.py:70-114)
// where lap_sigma = IFFT(-G^2 exp(-G^2 sigma^2 / 2) FFT) on the full complex
// grid (field.py:343-350).  Inner products accumulate in FP64.
#include <cmath>
#include <complex>
#include <initializer_list>
#include <tuple>
#include <vector>

#include "internal.h"
#include "reduce.h"

namespace ofdft {

#define GRID_LOOP(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

template <int K> __device__ __forceinline__ void td_partials(double *acc, double *partials) {
    __shared__ double sm[16];
#pragma unroll
    for (int t = 0; t < K; ++t) {
        const double s = block_sum(acc[t], sm);
        if (threadIdx.x == 0) partials[(long long)blockIdx.x * K + t] = s;
    }
}

__device__ __forceinline__ int ffreq_td(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

// spec *= 1/2 G^2 exp(-G^2 sigma^2 / 2) / N   (full grid, fftfreq on all axes)
template <typename T>
__global__ void k_td_gmul(cpx<T> *spec, long long n, int nx, int ny, int nz, double B0, double B1, double B2,
                          double B3, double B4, double B5, double B6, double B7, double B8, double sigma,
                          double scale) {
    GRID_LOOP(i, n) {
        const int iz = (int)(i % nz);
        const long long r = i / nz;
        const int iy = (int)(r % ny), ix = (int)(r / ny);
        const int fx = ffreq_td(ix, nx), fy = ffreq_td(iy, ny), fz = ffreq_td(iz, nz);
        const double gx = fx * B0 + fy * B3 + fz * B6;
        const double gy = fx * B1 + fy * B4 + fz * B7;
        const double gz = fx * B2 + fy * B5 + fz * B8;
        const double gg = gx * gx + gy * gy + gz * gz;
        const double f = 0.5 * gg * exp(-gg * sigma * sigma / 2.0) * scale;
        cpx<T> v = spec[i];
        spec[i] = cpx<T>{(T)(v.x * f), (T)(v.y * f)};
    }
}

// batched form: element i of the stacked runs has grid index i % n
template <typename T>
__global__ void k_td_gmul_b(cpx<T> *spec, long long tot, long long n, int nx, int ny, int nz, double B0, double B1,
                            double B2, double B3, double B4, double B5, double B6, double B7, double B8, double sigma,
                            double scale) {
    GRID_LOOP(j, tot) {
        const long long i = j % n;
        const int iz = (int)(i % nz);
        const long long r = i / nz;
        const int iy = (int)(r % ny), ix = (int)(r / ny);
        const int fx = ffreq_td(ix, nx), fy = ffreq_td(iy, ny), fz = ffreq_td(iz, nz);
        const double gx = fx * B0 + fy * B3 + fz * B6;
        const double gy = fx * B1 + fy * B4 + fz * B7;
        const double gz = fx * B2 + fy * B5 + fz * B8;
        const double gg = gx * gx + gy * gy + gz * gz;
        const double f = 0.5 * gg * exp(-gg * sigma * sigma / 2.0) * scale;
        cpx<T> v = spec[j];
        spec[j] = cpx<T>{(T)(v.x * f), (T)(v.y * f)};
    }
}

// out = x + sgn * i (dt/2) (kin + v x)   with kin = -1/2 lap x already in `kin`
// sums: (mode 1) p.out complex (no conj)   (mode 2) |out|^2
template <typename T, int MODE>
__global__ void k_td_apply(const cpx<T> *x, const cpx<T> *kin, const T *v, double hdt, cpx<T> *out, long long n,
                           double *partials) {
    double acc[2] = {0, 0};
    GRID_LOOP(i, n) {
        const cpx<T> a = x[i], k = kin[i];
        const double vv = (double)v[i];
        const double hr = (double)k.x + vv * (double)a.x, hi = (double)k.y + vv * (double)a.y;
        // i * hdt * (hr + i hi) = -hdt hi + i hdt hr
        const double orr = (double)a.x - hdt * hi, oi = (double)a.y + hdt * hr;
        out[i] = cpx<T>{(T)orr, (T)oi};
        if (MODE == 1) {
            acc[0] += (double)a.x * orr - (double)a.y * oi;
            acc[1] += (double)a.x * oi + (double)a.y * orr;
        } else if (MODE == 2) {
            acc[0] += orr * orr + oi * oi;
        }
    }
    if (MODE == 1) td_partials<2>(acc, partials);
    if (MODE == 2) td_partials<1>(acc, partials);
}

// res = b - res (res holds A x0) ; p = res ; sums res.res (complex), |res|^2
template <typename T>
__global__ void k_td_init(const cpx<T> *b, cpx<T> *res, cpx<T> *p, long long n, double *partials) {
    double acc[3] = {0, 0, 0};
    GRID_LOOP(i, n) {
        const cpx<T> bb = b[i], ax = res[i];
        const cpx<T> r{bb.x - ax.x, bb.y - ax.y};
        res[i] = r;
        p[i] = r;
        acc[0] += (double)r.x * r.x - (double)r.y * r.y;
        acc[1] += 2.0 * (double)r.x * r.y;
        acc[2] += (double)r.x * r.x + (double)r.y * r.y;
    }
    td_partials<3>(acc, partials);
}

// x += a p ; res -= a v ; sums res.res (complex), |res|^2
template <typename T>
__global__ void k_td_update(cpx<T> *x, cpx<T> *res, const cpx<T> *p, const cpx<T> *v, double ar, double ai, long long n,
                            double *partials) {
    double acc[3] = {0, 0, 0};
    GRID_LOOP(i, n) {
        const cpx<T> pp = p[i], vv = v[i], xx = x[i], rr = res[i];
        x[i] = cpx<T>{(T)(xx.x + ar * pp.x - ai * pp.y), (T)(xx.y + ar * pp.y + ai * pp.x)};
        const double r0 = rr.x - (ar * vv.x - ai * vv.y), r1 = rr.y - (ar * vv.y + ai * vv.x);
        res[i] = cpx<T>{(T)r0, (T)r1};
        acc[0] += r0 * r0 - r1 * r1;
        acc[1] += 2.0 * r0 * r1;
        acc[2] += r0 * r0 + r1 * r1;
    }
    td_partials<3>(acc, partials);
}

// p = res + beta p
template <typename T> __global__ void k_td_pupd(const cpx<T> *res, cpx<T> *p, double br, double bi, long long n) {
    GRID_LOOP(i, n) {
        const cpx<T> pp = p[i], rr = res[i];
        p[i] = cpx<T>{(T)(rr.x + br * pp.x - bi * pp.y), (T)(rr.y + br * pp.y + bi * pp.x)};
    }
}

template <typename T> __global__ void k_td_copy(const cpx<T> *a, cpx<T> *b, long long n) {
    GRID_LOOP(i, n) b[i] = a[i];
}

#undef GRID_LOOP

template <typename T> struct CN {
    Plan &P;
    hipStream_t s;
    long long n;
    int nb;
    double *host;
    double sigma;
    cpx<T> *kin, *tmp;
    CN(Plan &p, hipStream_t st, double sg) : P(p), s(st), n(p.N), sigma(sg) {
        nb = std::min(reduce_blocks(n), 1024);
        host = P.pinned;
        kin = reinterpret_cast<cpx<T> *>(P.spec);
    }

    // the K sums, published by the reduction itself to the pinned buffer
    const double *fetch(int K) {
        std::vector<double> ones(K, 1.0);
        reduce_final(P.red, nb, K, P.scal + 64, ones.data(), s, nullptr, P.scal + 64, P.pinned_dev, K);
        stream_wait(s);
        return host;
    }
    // kin = -1/2 lap_sigma(x)
    void kinetic(const cpx<T> *x) {
        fft_c2c<T>(P, x, kin, false, 1, n, 1.0, s);
        hipLaunchKernelGGL((k_td_gmul<T>), dim3(std::min<long long>((n + 255) / 256, 4096)), dim3(256), 0, s, kin, n,
                           P.nr[0], P.nr[1], P.nr[2], P.B[0], P.B[1], P.B[2], P.B[3], P.B[4], P.B[5], P.B[6], P.B[7],
                           P.B[8], sigma, 1.0 / (double)n);
        OFDFT_HIP(hipGetLastError());
        fft_c2c<T>(P, kin, kin, true, 1, n, 1.0, s);
    }
    template <int MODE> void apply(const cpx<T> *x, const T *v, double hdt, cpx<T> *out) {
        kinetic(x);
        hipLaunchKernelGGL((k_td_apply<T, MODE>), dim3(nb), dim3(256), 0, s, x, kin, v, hdt, out, n, P.red);
        OFDFT_HIP(hipGetLastError());
    }

    void step(const cpx<T> *psi0, const T *v, double dt, double tol, int maxiter, cpx<T> *x, double *info) {
        cpx<T> *base = reinterpret_cast<cpx<T> *>(P.aux);
        cpx<T> *b = base, *res = base + n, *p = base + 2 * n, *Ap = base + 3 * n;
        // b = psi0 - i dt/2 H psi0, ||b||
        apply<2>(psi0, v, -0.5 * dt, b);
        const double bn = std::sqrt(fetch(1)[0]);
        const double atol = bn == 0.0 ? tol : tol * bn;
        // x = psi0 ; res = b - A x
        if (x != psi0) hipLaunchKernelGGL((k_td_copy<T>), dim3(nb), dim3(256), 0, s, psi0, x, n);
        apply<0>(x, v, 0.5 * dt, res);
        hipLaunchKernelGGL((k_td_init<T>), dim3(nb), dim3(256), 0, s, b, res, p, n, P.red);
        const double *h = fetch(3);
        double rr_r = h[0], rr_i = h[1];
        double rnorm = std::sqrt(h[2]);
        int status = 1, k = 0;
        if (rnorm < atol) {
            status = 0;
        } else {
            while (k < maxiter) {
                apply<1>(p, v, 0.5 * dt, Ap);
                h = fetch(2);
                const double pv_r = h[0], pv_i = h[1];
                const double den = pv_r * pv_r + pv_i * pv_i;
                const double ar = (rr_r * pv_r + rr_i * pv_i) / den, ai = (rr_i * pv_r - rr_r * pv_i) / den;
                hipLaunchKernelGGL((k_td_update<T>), dim3(nb), dim3(256), 0, s, x, res, p, Ap, ar, ai, n, P.red);
                h = fetch(3);
                const double nr_r = h[0], nr_i = h[1];
                rnorm = std::sqrt(h[2]);
                ++k;
                if (rnorm < atol) {
                    status = 0;
                    break;
                }
                const double d2 = rr_r * rr_r + rr_i * rr_i;
                const double br = (nr_r * rr_r + nr_i * rr_i) / d2, bi = (nr_i * rr_r - nr_r * rr_i) / d2;
                hipLaunchKernelGGL((k_td_pupd<T>), dim3(nb), dim3(256), 0, s, res, p, br, bi, n);
                rr_r = nr_r;
                rr_i = nr_i;
            }
        }
        stream_wait(s);
        if (info) {
            info[0] = status;
            info[1] = k;
            info[2] = rnorm;
            info[3] = atol;
        }
    }
};

// ---------------------------------------------------------------- Krylov solvers
// The other linear solvers of the Crank-Nicolson step
// (crank_nicholson.py:65-77): bicg and bicgstab as linear_solver.py:49-110
// writes them, and GMRES(restart) (reached there through scipy's gmres),
// on the same operator A = 1 + i H dt/2.  Complex vector kernels with FP64
// sums; the scalar recursions run on the host on FP64 values.
struct ZPairs {
    const void *a[4];
    const void *b[4];
    int conj[4];  // pair t sums conj(a) b (1) or a b (0)
    int k;
};

// up to 4 complex sums, columns (2t, 2t+1) = (re, im) of pair t
template <typename T> __global__ void k_zsums(ZPairs z, long long n, double *partials) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t < z.k) {
                const cpx<T> a = reinterpret_cast<const cpx<T> *>(z.a[t])[i];
                const cpx<T> b = reinterpret_cast<const cpx<T> *>(z.b[t])[i];
                const double ar = (double)a.x, ai = z.conj[t] ? -(double)a.y : (double)a.y;
                acc[2 * t] += ar * (double)b.x - ai * (double)b.y;
                acc[2 * t + 1] += ar * (double)b.y + ai * (double)b.x;
            }
        }
    }
    td_partials<8>(acc, partials);
}

// out = ca a + cb b (complex scalars; b may be null); out may alias a or b
template <typename T>
__global__ void k_zlin(const cpx<T> *a, double car, double cai, const cpx<T> *b, double cbr, double cbi, cpx<T> *out,
                       long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const cpx<T> x = a[i];
        double r = car * x.x - cai * x.y, m = car * x.y + cai * x.x;
        if (b) {
            const cpx<T> y = b[i];
            r += cbr * y.x - cbi * y.y;
            m += cbr * y.y + cbi * y.x;
        }
        out[i] = cpx<T>{(T)r, (T)m};
    }
}

constexpr int kZChunk = 8;  // basis vectors per projection launch
struct ZBasis {
    const void *v[kZChunk];
    double cr[kZChunk], ci[kZChunk];
    int nv;
};

// sums conj(v_j) w for the chunk's vectors, columns (2j, 2j+1)
template <typename T> __global__ void k_zproj(ZBasis zb, const cpx<T> *w, long long n, double *partials) {
    double acc[2 * kZChunk];
#pragma unroll
    for (int j = 0; j < 2 * kZChunk; ++j) acc[j] = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const cpx<T> x = w[i];
#pragma unroll
        for (int j = 0; j < kZChunk; ++j) {
            if (j < zb.nv) {
                const cpx<T> v = reinterpret_cast<const cpx<T> *>(zb.v[j])[i];
                acc[2 * j] += (double)v.x * x.x + (double)v.y * x.y;
                acc[2 * j + 1] += (double)v.x * x.y - (double)v.y * x.x;
            }
        }
    }
    td_partials<2 * kZChunk>(acc, partials);
}

// w += sum_j c_j v_j over the chunk's vectors
template <typename T> __global__ void k_zcomb(ZBasis zb, cpx<T> *w, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const cpx<T> x = w[i];
        double r = x.x, m = x.y;
#pragma unroll
        for (int j = 0; j < kZChunk; ++j) {
            if (j < zb.nv) {
                const cpx<T> v = reinterpret_cast<const cpx<T> *>(zb.v[j])[i];
                r += zb.cr[j] * v.x - zb.ci[j] * v.y;
                m += zb.cr[j] * v.y + zb.ci[j] * v.x;
            }
        }
        w[i] = cpx<T>{(T)r, (T)m};
    }
}

enum { CN_CG = 0, CN_BICG = 1, CN_BICGSTAB = 2, CN_GMRES = 3 };
constexpr int kMaxRestart = 64;

template <typename T> struct CNK : CN<T> {
    using C = cpx<T>;
    using CN<T>::P;
    using CN<T>::s;
    using CN<T>::n;
    using CN<T>::nb;
    char *ws;
    CNK(Plan &p, hipStream_t st, double sg, void *w) : CN<T>(p, st, sg), ws(reinterpret_cast<char *>(w)) {}
    C *field(int k) { return reinterpret_cast<C *>(ws) + (size_t)k * n; }
    typedef std::complex<double> Z;

    // complex sums of up to 4 pairs, fetched to the host
    std::vector<Z> sums(std::initializer_list<std::tuple<const C *, const C *, int>> pairs) {
        ZPairs z{};
        z.k = 0;
        for (auto &pr : pairs) {
            z.a[z.k] = std::get<0>(pr);
            z.b[z.k] = std::get<1>(pr);
            z.conj[z.k] = std::get<2>(pr);
            ++z.k;
        }
        hipLaunchKernelGGL((k_zsums<T>), dim3(nb), dim3(256), 0, s, z, n, P.red);
        OFDFT_HIP(hipGetLastError());
        const double *h = this->fetch(8);
        std::vector<Z> out(z.k);
        for (int t = 0; t < z.k; ++t) out[t] = Z(h[2 * t], h[2 * t + 1]);
        return out;
    }
    void lin(const C *a, Z ca, const C *b, Z cb, C *out) {
        hipLaunchKernelGGL((k_zlin<T>), dim3(nb), dim3(256), 0, s, a, ca.real(), ca.imag(), b, cb.real(), cb.imag(),
                           out, n);
        OFDFT_HIP(hipGetLastError());
    }
    double norm(const C *x) { return std::sqrt(sums({{x, x, 1}})[0].real()); }
    void A(const C *x, const T *v, double dt, C *out) { this->template apply<0>(x, v, 0.5 * dt, out); }

    // linear_solver.py:49-75 (sums conjugated)
    int bicg(const C *b, const T *v, double dt, double atol, int maxiter, C *x, int &k, double &rn) {
        C *res = field(1), *p = field(2), *w = field(3);
        A(x, v, dt, res);
        lin(b, 1.0, res, -1.0, res);
        rn = norm(res);
        k = 0;
        if (rn < atol) return 0;
        lin(res, 1.0, nullptr, 0.0, p);
        Z rr = sums({{res, res, 1}})[0];
        while (k < maxiter) {
            A(p, v, dt, w);
            const Z alpha = rr / sums({{p, w, 1}})[0];
            lin(x, 1.0, p, alpha, x);
            lin(res, 1.0, w, -alpha, res);
            const Z rn2 = sums({{res, res, 1}})[0];
            rn = std::sqrt(rn2.real());
            ++k;
            if (rn < atol) return 0;
            lin(res, 1.0, p, rn2 / rr, p);
            rr = rn2;
        }
        return 1;
    }

    // linear_solver.py:78-110 (sums not conjugated; res and r follow the
    // same recursion, so one field carries both)
    int bicgstab(const C *b, const T *v, double dt, double atol, int maxiter, C *x, int &k, double &rn) {
        C *r0 = field(1), *r = field(2), *p = field(3), *w = field(4), *sv = field(5), *t = field(6);
        A(x, v, dt, r0);
        lin(b, 1.0, r0, -1.0, r0);
        rn = norm(r0);
        k = 0;
        if (rn < atol) return 0;
        lin(r0, 1.0, nullptr, 0.0, r);
        hipLaunchKernelGGL((k_zlin<T>), dim3(nb), dim3(256), 0, s, r0, 0.0, 0.0, (const C *)nullptr, 0.0, 0.0, p, n);
        hipLaunchKernelGGL((k_zlin<T>), dim3(nb), dim3(256), 0, s, r0, 0.0, 0.0, (const C *)nullptr, 0.0, 0.0, w, n);
        Z rho_old = 1.0, omega = 1.0, alpha = 1.0;
        while (k < maxiter) {
            const Z rho = sums({{r0, r, 0}})[0];
            const Z beta = rho / rho_old * alpha / omega;
            lin(p, 1.0, w, -omega, p);  // p - omega v
            lin(r, 1.0, p, beta, p);    // r + beta (p - omega v)
            A(p, v, dt, w);
            alpha = rho / sums({{r0, w, 0}})[0];
            lin(x, 1.0, p, alpha, x);  // h = x + alpha p
            lin(r, 1.0, w, -alpha, sv);  // s = r - alpha v (= res)
            rn = norm(sv);
            if (rn < atol) {
                ++k;
                return 0;
            }
            A(sv, v, dt, t);
            const std::vector<Z> ts = sums({{t, sv, 0}, {t, t, 0}});
            omega = ts[0] / ts[1];
            lin(x, 1.0, sv, omega, x);
            lin(sv, 1.0, t, -omega, r);
            rn = norm(r);
            ++k;
            if (rn < atol) return 0;
            rho_old = rho;
        }
        return 1;
    }

    // GMRES(m): Arnoldi by classical Gram-Schmidt applied twice, Givens
    // rotations on the host; the residual estimate |g_{j+1}| against atol,
    // the true residual recomputed at each restart.  maxiter counts Arnoldi
    // steps (operator applications).
    void project(int nv, C *w, std::vector<Z> &h) {
        h.assign(nv, Z(0.0, 0.0));
        for (int c0 = 0; c0 < nv; c0 += kZChunk) {
            ZBasis zb{};
            zb.nv = std::min(kZChunk, nv - c0);
            for (int j = 0; j < zb.nv; ++j) zb.v[j] = field(4 + c0 + j);
            hipLaunchKernelGGL((k_zproj<T>), dim3(nb), dim3(256), 0, s, zb, (const C *)w, n, P.red);
            OFDFT_HIP(hipGetLastError());
            const double *hh = this->fetch(2 * kZChunk);
            for (int j = 0; j < zb.nv; ++j) h[c0 + j] = Z(hh[2 * j], hh[2 * j + 1]);
        }
    }
    void combine(int nv, const std::vector<Z> &c, double sign, C *w) {
        for (int c0 = 0; c0 < nv; c0 += kZChunk) {
            ZBasis zb{};
            zb.nv = std::min(kZChunk, nv - c0);
            for (int j = 0; j < zb.nv; ++j) {
                zb.v[j] = field(4 + c0 + j);
                zb.cr[j] = sign * c[c0 + j].real();
                zb.ci[j] = sign * c[c0 + j].imag();
            }
            hipLaunchKernelGGL((k_zcomb<T>), dim3(nb), dim3(256), 0, s, zb, w, n);
            OFDFT_HIP(hipGetLastError());
        }
    }
    int gmres(const C *b, const T *v, double dt, double atol, int maxiter, int m, C *x, int &k, double &rn) {
        C *r = field(1), *w = field(2);
        k = 0;
        std::vector<std::vector<Z>> H(m + 1, std::vector<Z>(m, Z(0, 0)));
        std::vector<Z> g(m + 1), cs(m), sn(m), h, h2;
        while (true) {
            A(x, v, dt, r);
            lin(b, 1.0, r, -1.0, r);
            const double beta = norm(r);
            rn = beta;
            if (beta < atol) return 0;
            if (k >= maxiter) return 1;
            lin(r, 1.0 / beta, nullptr, 0.0, field(4));
            std::fill(g.begin(), g.end(), Z(0, 0));
            g[0] = beta;
            int j = 0;
            bool done = false;
            for (; j < m && k < maxiter; ++j) {
                A(field(4 + j), v, dt, w);
                ++k;
                project(j + 1, w, h);
                combine(j + 1, h, -1.0, w);
                project(j + 1, w, h2);
                combine(j + 1, h2, -1.0, w);
                for (int i = 0; i <= j; ++i) H[i][j] = h[i] + h2[i];
                const double hn = norm(w);
                H[j + 1][j] = hn;
                for (int i = 0; i < j; ++i) {  // earlier rotations
                    const Z a = H[i][j], c = H[i + 1][j];
                    H[i][j] = std::conj(cs[i]) * a + std::conj(sn[i]) * c;
                    H[i + 1][j] = -sn[i] * a + cs[i] * c;
                }
                const Z a = H[j][j], c = H[j + 1][j];
                const double den = std::sqrt(std::norm(a) + std::norm(c));
                cs[j] = den > 0 ? a / den : Z(1, 0);
                sn[j] = den > 0 ? c / den : Z(0, 0);
                H[j][j] = den;
                H[j + 1][j] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = std::conj(cs[j]) * g[j];
                rn = std::abs(g[j + 1]);
                if (rn < atol || hn == 0.0) {
                    ++j;
                    done = true;
                    break;
                }
                if (j + 1 < m) lin(w, 1.0 / hn, nullptr, 0.0, field(4 + j + 1));
            }
            // y = H^-1 g (upper triangular j x j), x += V y
            std::vector<Z> y(j);
            for (int i = j - 1; i >= 0; --i) {
                Z acc = g[i];
                for (int l = i + 1; l < j; ++l) acc -= H[i][l] * y[l];
                y[i] = acc / H[i][i];
            }
            combine(j, y, 1.0, x);
            if (done) {
                // the estimate met the tolerance: confirm with the true residual
                A(x, v, dt, r);
                lin(b, 1.0, r, -1.0, r);
                rn = norm(r);
                if (rn < atol) return 0;
            }
        }
    }

    void solve(int solver, int restart, const C *psi0, const T *v, double dt, double tol, int maxiter, C *x,
               double *info) {
        C *b = field(0);
        this->template apply<2>(psi0, v, -0.5 * dt, b);
        const double bn = std::sqrt(this->fetch(1)[0]);
        const double atol = bn == 0.0 ? tol : tol * bn;
        if (x != psi0) lin(psi0, 1.0, nullptr, 0.0, x);
        int k = 0, st = 1;
        double rn = 0.0;
        if (solver == CN_BICG)
            st = bicg(b, v, dt, atol, maxiter, x, k, rn);
        else if (solver == CN_BICGSTAB)
            st = bicgstab(b, v, dt, atol, maxiter, x, k, rn);
        else
            st = gmres(b, v, dt, atol, maxiter, restart, x, k, rn);
        stream_wait(s);
        if (info) {
            info[0] = st;
            info[1] = k;
            info[2] = rn;
            info[3] = atol;
        }
    }
};

// ---------------------------------------------------------------- batched
// nb independent runs (each its own psi, v and CG scalars) advance in lock
// step: every Hamiltonian application is one batched C2C transform pair and
// one launch over all runs (blockIdx.y = run); per-run FP64 partial sums are
// laid out partials[(run * nblk + block) * K + t].  Runs that have converged
// stop updating; the loop ends when all have.
constexpr int kMaxRuns = 16;

struct RunScalars {
    double ar[kMaxRuns], ai[kMaxRuns];
    int active[kMaxRuns];
};

template <int K> __device__ __forceinline__ void run_partials(double *acc, double *partials) {
    __shared__ double sm[16];
#pragma unroll
    for (int t = 0; t < K; ++t) {
        const double s = block_sum(acc[t], sm);
        if (threadIdx.x == 0) partials[((long long)blockIdx.y * gridDim.x + blockIdx.x) * K + t] = s;
    }
}

#define RUN_LOOP(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

template <typename T, int MODE>
__global__ void k_tdb_apply(const cpx<T> *x, const cpx<T> *kin, const T *v, double hdt, cpx<T> *out, long long n,
                            double *partials) {
    const long long o = (long long)blockIdx.y * n;
    x += o;
    kin += o;
    v += o;
    out += o;
    double acc[2] = {0, 0};
    RUN_LOOP(i, n) {
        const cpx<T> a = x[i], k = kin[i];
        const double vv = (double)v[i];
        const double hr = (double)k.x + vv * (double)a.x, hi = (double)k.y + vv * (double)a.y;
        const double orr = (double)a.x - hdt * hi, oi = (double)a.y + hdt * hr;
        out[i] = cpx<T>{(T)orr, (T)oi};
        if (MODE == 1) {
            acc[0] += (double)a.x * orr - (double)a.y * oi;
            acc[1] += (double)a.x * oi + (double)a.y * orr;
        } else if (MODE == 2) {
            acc[0] += orr * orr + oi * oi;
        }
    }
    if (MODE == 1) run_partials<2>(acc, partials);
    if (MODE == 2) run_partials<1>(acc, partials);
}

template <typename T>
__global__ void k_tdb_init(const cpx<T> *b, cpx<T> *res, cpx<T> *p, long long n, double *partials) {
    const long long o = (long long)blockIdx.y * n;
    double acc[3] = {0, 0, 0};
    RUN_LOOP(i, n) {
        const cpx<T> bb = b[o + i], ax = res[o + i];
        const cpx<T> r{bb.x - ax.x, bb.y - ax.y};
        res[o + i] = r;
        p[o + i] = r;
        acc[0] += (double)r.x * r.x - (double)r.y * r.y;
        acc[1] += 2.0 * (double)r.x * r.y;
        acc[2] += (double)r.x * r.x + (double)r.y * r.y;
    }
    run_partials<3>(acc, partials);
}

template <typename T>
__global__ void k_tdb_update(cpx<T> *x, cpx<T> *res, const cpx<T> *p, const cpx<T> *v, RunScalars rs, long long n,
                             double *partials) {
    const int run = blockIdx.y;
    const long long o = (long long)run * n;
    double acc[3] = {0, 0, 0};
    if (rs.active[run]) {
        const double ar = rs.ar[run], ai = rs.ai[run];
        RUN_LOOP(i, n) {
            const cpx<T> pp = p[o + i], vv = v[o + i], xx = x[o + i], rr = res[o + i];
            x[o + i] = cpx<T>{(T)(xx.x + ar * pp.x - ai * pp.y), (T)(xx.y + ar * pp.y + ai * pp.x)};
            const double r0 = rr.x - (ar * vv.x - ai * vv.y), r1 = rr.y - (ar * vv.y + ai * vv.x);
            res[o + i] = cpx<T>{(T)r0, (T)r1};
            acc[0] += r0 * r0 - r1 * r1;
            acc[1] += 2.0 * r0 * r1;
            acc[2] += r0 * r0 + r1 * r1;
        }
    }
    run_partials<3>(acc, partials);
}

template <typename T> __global__ void k_tdb_pupd(const cpx<T> *res, cpx<T> *p, RunScalars rs, long long n) {
    const int run = blockIdx.y;
    if (!rs.active[run]) return;
    const long long o = (long long)run * n;
    const double br = rs.ar[run], bi = rs.ai[run];
    RUN_LOOP(i, n) {
        const cpx<T> pp = p[o + i], rr = res[o + i];
        p[o + i] = cpx<T>{(T)(rr.x + br * pp.x - bi * pp.y), (T)(rr.y + br * pp.y + bi * pp.x)};
    }
}

// out[run * K + t] = sum_b partials[(run * nblk + b) * K + t]
// one workgroup per (run, term); pub (the device view of pinned host memory,
// or null) receives a copy for the host
__global__ void k_tdb_reduce(const double *partials, int nblk, int K, double *out, double *pub) {
    __shared__ double sm[16];
    const int run = blockIdx.x / K, t = blockIdx.x % K;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) acc += partials[((long long)run * nblk + b) * K + t];
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) {
        out[blockIdx.x] = acc;
        if (pub) pub[blockIdx.x] = acc;
    }
}

#undef RUN_LOOP

template <typename T> struct CNBatch {
    Plan &P;
    hipStream_t s;
    long long n;
    int nb, nblk;
    double *host;
    double sigma;
    cpx<T> *kin, *b, *res, *p, *Ap;
    CNBatch(Plan &pl, hipStream_t st, double sg, int nruns, void *work)
        : P(pl), s(st), n(pl.N), nb(nruns), sigma(sg) {
        nblk = std::max(1, std::min(reduce_blocks(n), 256));
        OFDFT_REQUIRE((long long)nb * nblk * 3 <= 4096LL * 16, "cn_step_batch: too many runs for the partials");
        host = P.pinned;
        cpx<T> *w = reinterpret_cast<cpx<T> *>(work);
        kin = w;
        b = w + (long long)nb * n;
        res = w + 2LL * nb * n;
        p = w + 3LL * nb * n;
        Ap = w + 4LL * nb * n;
    }

    const double *fetch(int K) {
        OFDFT_REQUIRE(nb * K <= 128, "cn_step_batch: read-back larger than the pinned buffer");
        hipLaunchKernelGGL(k_tdb_reduce, dim3(nb * K), dim3(256), 0, s, P.red, nblk, K, P.scal + 128, P.pinned_dev);
        OFDFT_HIP(hipGetLastError());
        stream_wait(s);
        return host;
    }
    void kinetic(const cpx<T> *x) {
        fft_c2c<T>(P, x, kin, false, nb, n, 1.0, s);
        const long long tot = (long long)nb * n;
        hipLaunchKernelGGL((k_td_gmul_b<T>), dim3(std::min<long long>((tot + 255) / 256, 8192)), dim3(256), 0, s, kin,
                           tot, n, P.nr[0], P.nr[1], P.nr[2], P.B[0], P.B[1], P.B[2], P.B[3], P.B[4], P.B[5], P.B[6],
                           P.B[7], P.B[8], sigma, 1.0 / (double)n);
        OFDFT_HIP(hipGetLastError());
        fft_c2c<T>(P, kin, kin, true, nb, n, 1.0, s);
    }
    template <int MODE> void apply(const cpx<T> *x, const T *v, double hdt, cpx<T> *out) {
        kinetic(x);
        hipLaunchKernelGGL((k_tdb_apply<T, MODE>), dim3(nblk, nb), dim3(256), 0, s, x, kin, v, hdt, out, n, P.red);
        OFDFT_HIP(hipGetLastError());
    }
    void step(const cpx<T> *psi0, const T *v, double dt, double tol, int maxiter, cpx<T> *x, double *info) {
        apply<2>(psi0, v, -0.5 * dt, b);
        const double *h = fetch(1);
        double atol[kMaxRuns], rr_r[kMaxRuns], rr_i[kMaxRuns], rnorm[kMaxRuns];
        int status[kMaxRuns], iters[kMaxRuns];
        for (int r = 0; r < nb; ++r) {
            const double bn = std::sqrt(h[r]);
            atol[r] = bn == 0.0 ? tol : tol * bn;
        }
        OFDFT_HIP(hipMemcpyAsync(x, psi0, (size_t)nb * n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
        apply<0>(x, v, 0.5 * dt, res);
        hipLaunchKernelGGL((k_tdb_init<T>), dim3(nblk, nb), dim3(256), 0, s, b, res, p, n, P.red);
        OFDFT_HIP(hipGetLastError());
        h = fetch(3);
        RunScalars sc{};
        int nactive = 0;
        for (int r = 0; r < nb; ++r) {
            rr_r[r] = h[3 * r];
            rr_i[r] = h[3 * r + 1];
            rnorm[r] = std::sqrt(h[3 * r + 2]);
            iters[r] = 0;
            status[r] = rnorm[r] < atol[r] ? 0 : 1;
            sc.active[r] = status[r];
            nactive += status[r];
        }
        for (int k = 0; k < maxiter && nactive > 0; ++k) {
            apply<1>(p, v, 0.5 * dt, Ap);
            h = fetch(2);
            for (int r = 0; r < nb; ++r) {
                if (!sc.active[r]) continue;
                const double pv_r = h[2 * r], pv_i = h[2 * r + 1];
                const double den = pv_r * pv_r + pv_i * pv_i;
                sc.ar[r] = (rr_r[r] * pv_r + rr_i[r] * pv_i) / den;
                sc.ai[r] = (rr_i[r] * pv_r - rr_r[r] * pv_i) / den;
            }
            hipLaunchKernelGGL((k_tdb_update<T>), dim3(nblk, nb), dim3(256), 0, s, x, res, p, Ap, sc, n, P.red);
            OFDFT_HIP(hipGetLastError());
            h = fetch(3);
            RunScalars bs = sc;
            for (int r = 0; r < nb; ++r) {
                if (!sc.active[r]) continue;
                const double nr_r = h[3 * r], nr_i = h[3 * r + 1];
                rnorm[r] = std::sqrt(h[3 * r + 2]);
                iters[r] = k + 1;
                if (rnorm[r] < atol[r]) {
                    status[r] = 0;
                    bs.active[r] = 0;
                    continue;
                }
                const double d2 = rr_r[r] * rr_r[r] + rr_i[r] * rr_i[r];
                bs.ar[r] = (nr_r * rr_r[r] + nr_i * rr_i[r]) / d2;
                bs.ai[r] = (nr_i * rr_r[r] - nr_r * rr_i[r]) / d2;
                rr_r[r] = nr_r;
                rr_i[r] = nr_i;
            }
            hipLaunchKernelGGL((k_tdb_pupd<T>), dim3(nblk, nb), dim3(256), 0, s, res, p, bs, n);
            OFDFT_HIP(hipGetLastError());
            nactive = 0;
            for (int r = 0; r < nb; ++r) {
                sc.active[r] = bs.active[r];
                nactive += bs.active[r];
            }
        }
        stream_wait(s);
        if (info)
            for (int r = 0; r < nb; ++r) {
                info[4 * r] = status[r];
                info[4 * r + 1] = iters[r];
                info[4 * r + 2] = rnorm[r];
                info[4 * r + 3] = atol[r];
            }
    }
};

}  // namespace ofdft

using namespace ofdft;

extern "C" size_t ofdft_cn_batch_workspace_bytes(const ofdft_plan *plan, int nb) {
    const Plan *P = reinterpret_cast<const Plan *>(plan);
    if (!P || nb < 1) return 0;
    return (size_t)5 * nb * P->N * (P->prec == OFDFT_FP64 ? 16 : 8);
}

extern "C" int ofdft_cn_step_batch(ofdft_plan *plan, int nb, const void *psi0, const void *v, double dt,
                                   double sigma, double tol, int maxiter, void *psi1, double *info, void *work,
                                   size_t work_bytes, void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        OFDFT_REQUIRE(P && psi0 && v && psi1 && work, "cn_step_batch: bad arguments");
        OFDFT_REQUIRE(nb >= 1 && nb <= kMaxRuns, "cn_step_batch: 1 to 16 runs");
        OFDFT_REQUIRE(psi1 != psi0, "cn_step_batch: psi1 must not alias psi0");
        OFDFT_REQUIRE(work_bytes >= ofdft_cn_batch_workspace_bytes(plan, nb), "cn_step_batch: workspace too small");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (P->prec == OFDFT_FP64) {
            CNBatch<double> c(*P, s, sigma, nb, work);
            c.step((const double2 *)psi0, (const double *)v, dt, tol, maxiter, (double2 *)psi1, info);
        } else {
            CNBatch<float> c(*P, s, sigma, nb, work);
            c.step((const float2 *)psi0, (const float *)v, dt, tol, maxiter, (float2 *)psi1, info);
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}

extern "C" size_t ofdft_cn_solver_workspace_bytes(const ofdft_plan *plan, int solver, int restart) {
    const Plan *P = reinterpret_cast<const Plan *>(plan);
    if (!P) return 0;
    const size_t f = (size_t)P->N * (P->prec == OFDFT_FP64 ? 16 : 8);
    if (solver == CN_GMRES) return (size_t)(4 + std::max(restart, 1)) * f;
    return (size_t)7 * f;
}

extern "C" int ofdft_cn_step_solver(ofdft_plan *plan, int solver, int restart, const void *psi0, const void *v,
                                    double dt, double sigma, double tol, int maxiter, void *psi1, double *info,
                                    void *work, size_t work_bytes, void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        OFDFT_REQUIRE(P && psi0 && v && psi1, "cn_step_solver: bad arguments");
        OFDFT_REQUIRE(solver >= CN_CG && solver <= CN_GMRES, "cn_step_solver: solver is 0 cg, 1 bicg, 2 bicgstab, "
                                                             "3 gmres");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (solver == CN_CG) {
            OFDFT_REQUIRE(P->aux_bytes >= (size_t)4 * P->N * (P->prec == OFDFT_FP64 ? 16 : 8), "cn_step: workspace");
            if (P->prec == OFDFT_FP64) {
                CN<double> c(*P, s, sigma);
                c.step((const double2 *)psi0, (const double *)v, dt, tol, maxiter, (double2 *)psi1, info);
            } else {
                CN<float> c(*P, s, sigma);
                c.step((const float2 *)psi0, (const float *)v, dt, tol, maxiter, (float2 *)psi1, info);
            }
            return 0;
        }
        OFDFT_REQUIRE(solver != CN_GMRES || (restart >= 1 && restart <= kMaxRestart),
                      "cn_step_solver: restart must be in [1, 64]");
        OFDFT_REQUIRE(work != nullptr && work_bytes >= ofdft_cn_solver_workspace_bytes(plan, solver, restart),
                      "cn_step_solver: workspace too small");
        OFDFT_REQUIRE(psi1 != psi0, "cn_step_solver: psi1 must not alias psi0");
        if (P->prec == OFDFT_FP64) {
            CNK<double> c(*P, s, sigma, work);
            c.solve(solver, restart, (const double2 *)psi0, (const double *)v, dt, tol, maxiter, (double2 *)psi1,
                    info);
        } else {
            CNK<float> c(*P, s, sigma, work);
            c.solve(solver, restart, (const float2 *)psi0, (const float *)v, dt, tol, maxiter, (float2 *)psi1, info);
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}

extern "C" int ofdft_cn_step(ofdft_plan *plan, const void *psi0, const void *v, double dt, double sigma, double tol,
                             int maxiter, void *psi1, double *info, void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        OFDFT_REQUIRE(P && psi0 && v && psi1, "cn_step: bad arguments");
        OFDFT_REQUIRE(P->aux_bytes >= (size_t)4 * P->N * (P->prec == OFDFT_FP64 ? 16 : 8), "cn_step: workspace");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (P->prec == OFDFT_FP64) {
            CN<double> c(*P, s, sigma);
            c.step((const double2 *)psi0, (const double *)v, dt, tol, maxiter, (double2 *)psi1, info);
        } else {
            CN<float> c(*P, s, sigma);
            c.step((const float2 *)psi0, (const float *)v, dt, tol, maxiter, (float2 *)psi1, info);
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}
