// G-space kernels of the evaluation pipeline, shared by the stand-alone
// multiply (eval.hip) and the x-pass that fuses forward FFT, multiply and
// inverse FFT (fft_fixed.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "fastmath.h"

namespace ofdft {

// ------------------------------------------------------------------ Lindhard
// Inverse Lindhard function with the reference's branch priority
// normal < zero (eta < tol) < one (|eta-1| < tol) < large (eta > 3.65).
// T = double for the FP64 pipeline, float for the FP32 one
template <typename T> __device__ __forceinline__ T lindhard_t(T eta, T lbda, T mu) {
    const T tol = (T)1e-10;
    if (eta > (T)3.65) {
        const T inv = fm::rcp(eta * eta);
        T s = (T)-3.00624946457977689e-3;
        s = (T)-3.38745061493813488e-3 + inv * s;
        s = (T)-3.84763737842981233e-3 + inv * s;
        s = (T)-4.41060829979912465e-3 + inv * s;
        s = (T)-5.10997527675418131e-3 + inv * s;
        s = (T)-5.9945117538835746e-3 + inv * s;
        s = (T)-7.1372762502456763e-3 + inv * s;
        s = (T)-8.65192783339199453e-3 + inv * s;
        s = (T)-1.07236045520990083e-2 + inv * s;
        s = (T)-1.36715733124818332e-2 + inv * s;
        s = (T)-1.80879839616166146e-2 + inv * s;
        s = (T)-2.51824061652633074e-2 + inv * s;
        s = (T)-3.77825602968460128e-2 + inv * s;
        s = (T)-6.39999999999999875e-2 + inv * s;
        s = (T)-0.13714285714285712 + inv * s;
        return (T)3 * ((T)1 - mu) * eta * eta - lbda - (T)0.6 + inv * s;
    }
    if (fabs(eta - (T)1) < tol) return (T)2 - lbda - (T)3 * mu + (T)20 * (eta - (T)1);
    if (eta < tol) return (T)1 - lbda + eta * eta * ((T)(1.0 / 3.0) - (T)3 * mu);
    // the logarithm stays the library's: its argument tends to 1 as eta -> 0
    const T lg = log(fm::div((T)1 + eta, fabs((T)1 - eta)));
    return fm::rcp((T)0.5 + fm::div((T)0.25 * ((T)1 - eta * eta) * lg, eta)) - (T)3 * mu * eta * eta - lbda;
}

__device__ __forceinline__ double lindhard(double eta, double lbda, double mu) { return lindhard_t<double>(eta, lbda, mu); }

// K(q) = d^2E/drho(q)^2 of the functional's terms at a uniform density:
// kc (TF and LDA) + vw q^2 + wt G(q / tkf) + 4 pi / q^2 (Hartree; infinite
// at q = 0), the host giving the rho0-dependent coefficients
struct RespK {
    double kc, vw, wt, tkf, lbda;
    int hartree;
};

// in CT (double, or float with the hardware math for FP32 fields)
template <typename CT> __device__ __forceinline__ CT resp_k(const RespK &r, CT q) {
    CT k = (CT)r.kc + (CT)r.vw * q * q;
    if (r.wt != 0.0) k += (CT)r.wt * lindhard_t<CT>(q * (CT)(1.0 / r.tkf), (CT)r.lbda, (CT)1);
    if (r.hartree) k += q > (CT)0 ? fm::div((CT)(4.0 * M_PI), q * q) : (CT)INFINITY;
    return k;
}

struct GParams {
    double B[9];
    int kind[4];  // per spectrum slot: 0 vW, 1 WT, 2 Hartree, 3 radial table, -1 none
    int nx, ny, nz, nzc;
    int pitch;  // row pitch of the spectra (>= nzc)
    int full;
    double scale;
    int s_vw, s_wt, s_h;
    int s_wt2 = -1;  // WT with alpha != beta: the rho^alpha spectrum (same kernel)
    double sigma;
    double lbda, alpha, beta;
    double rho0;             // used when rho0_dev == nullptr
    const double *rho0_dev;  // cached rho0 on device
    // with rho0_dev: the WT constants (tkf, wtfac) of the cached rho0, which
    // the kernel that updates the cache computes once (null: per thread)
    const double *wt_dev = nullptr;
    long long bs;
    // kind 3: a radial table of |G| (linear interpolation, constant past the
    // end), e.g. the optimizer's preconditioner
    const double *tab;
    int tab_n;
    double tab_inv_dq;
};

__device__ __forceinline__ double radial_table(const GParams &p, double g) {
    const double t = g * p.tab_inv_dq;
    if (t >= (double)(p.tab_n - 1)) return p.tab[p.tab_n - 1];
    const int i = (int)t;
    const double f = t - i;
    return p.tab[i] * (1.0 - f) + p.tab[i + 1] * f;
}

__device__ __forceinline__ int ffreq(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }
__device__ __forceinline__ int wrapf(int f, int n) { return ((n & 1) == 0 && f == n / 2) ? -n / 2 : f; }

__device__ __forceinline__ double gsq(const double *B, int fx, int fy, int fz) {
    const double gx = fx * B[0] + fy * B[3] + fz * B[6];
    const double gy = fx * B[1] + fy * B[4] + fz * B[7];
    const double gz = fx * B[2] + fy * B[5] + fz * B[8];
    return gx * gx + gy * gy + gz * gz;
}

struct KVals {
    double vw, wt, h;
};

__device__ __forceinline__ KVals kernel_vals(const GParams &p, double gg, bool g0, double tkf, double wtfac) {
    KVals k;
    k.vw = 0.5 * gg;
    if (p.sigma > 0) k.vw *= exp(-gg * p.sigma * p.sigma / 4.0);
    k.wt = lindhard(sqrt(gg) / tkf, p.lbda, 1.0) * wtfac;
    k.h = g0 ? 0.0 : 4.0 * M_PI / gg;
    return k;
}


// value of one kernel (kind 0 vW, 1 WT, 2 Hartree) at |G|^2 = gg; the WT
// Lindhard function is evaluated in T
template <typename T>
__device__ __forceinline__ double kernel_val(const GParams &p, int kind, double gg, bool g0, double tkf, double wtfac) {
    if (kind == 0) {
        double v = 0.5 * gg;
        if (p.sigma > 0) v *= exp(-gg * p.sigma * p.sigma / 4.0);
        return v;
    }
    if (kind == 1) return (double)lindhard_t<T>((T)(sqrt(gg) / tkf), (T)p.lbda, (T)1) * wtfac;
    if (kind == 3) return radial_table(p, sqrt(gg));
    return g0 ? 0.0 : 4.0 * M_PI / gg;
}

// multiplier of spectrum kind `kind` at frequency indices (ix, iy, iz) of the
// half grid, including the full-grid Nyquist averaging and the 1/N scale
template <typename T>
__device__ __forceinline__ double gfactor(const GParams &p, int kind, int ix, int iy, int iz, double tkf,
                                          double wtfac) {
    int fx = ffreq(ix, p.nx), fy = ffreq(iy, p.ny), fz = iz;
    const bool g0 = (ix == 0 && iy == 0 && iz == 0);
    double f;
    if (p.full) {
        fz = wrapf(fz, p.nz);
        const bool nyq = ((p.nx & 1) == 0 && ix == p.nx / 2) || ((p.ny & 1) == 0 && iy == p.ny / 2) ||
                         ((p.nz & 1) == 0 && iz == p.nz / 2);
        f = kernel_val<T>(p, kind, gsq(p.B, fx, fy, fz), g0, tkf, wtfac);
        if (nyq)
            f = 0.5 * (f + kernel_val<T>(p, kind, gsq(p.B, wrapf(-fx, p.nx), wrapf(-fy, p.ny), wrapf(-fz, p.nz)), g0,
                                         tkf, wtfac));
    } else {
        f = kernel_val<T>(p, kind, gsq(p.B, fx, fy, fz), g0, tkf, wtfac);
    }
    return f * p.scale;
}

// 2 k_F of rho0 and the WT kernel prefactor (wt.py:66, kernel.py:325-334)
__device__ __forceinline__ void wt_consts_at(double rho0, double alpha, double beta, double &tkf, double &wtfac) {
    tkf = 2.0 * cbrt(3.0 * rho0 * M_PI * M_PI);
    const double ctf = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
    wtfac = 5.0 / (9.0 * alpha * beta * pow(rho0, alpha + beta - 5.0 / 3.0)) * ctf;
}

__device__ __forceinline__ void wt_consts(const GParams &p, double &tkf, double &wtfac) {
    if (p.wt_dev) {
        tkf = p.wt_dev[0];
        wtfac = p.wt_dev[1];
        return;
    }
    wt_consts_at(p.rho0_dev ? *p.rho0_dev : p.rho0, p.alpha, p.beta, tkf, wtfac);
}

}  // namespace ofdft
