// LMGP nonlocal kinetic functional: a kernel that depends on the local Fermi
// wavevector k_F(r) = 2 (3 pi^2 rho(r))^(1/3), evaluated as in
// functional/kedf/lwt.py:85-341 (LWTPotentialEnergy, linear interpolation in
// k_F, kdd = 3) with the MGP kernel table of kernel.py:266-322.
//
// For a geometric ladder of k_F samples kf_j (ratio 1.15, lwt.py:187-192) the
// pipeline streams over j, holding two samples at a time like the reference:
//   p_j = IFFT(K(|G|/kf_j) FFT(rho^beta)),  m_j = IFFT(K'(|G|/kf_j)/kf_j FFT(rho^beta))
//   pot1, pot3 = linear interpolation of p, m in k_F between the samples
//                bracketing k_F(r)
//   pot2 = IFFT(sum_j K(|G|/kf_j) FFT(rho^alpha w_j)), w_j the hat weight of j
// and the final point-wise kernel forms the energy and potential with the
// low-density factor of lwt.py:307-317.  The kernel table K(eta) and
// -eta dK/deta are piecewise cubics on a uniform eta grid (splines of
// lwt.py:541-557); the table itself is computed on the GPU by
// ofdft_mgp_kernel_table.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gspace.h"
#include "internal.h"
#include "reduce.h"

namespace ofdft {

struct LmgpState {
    int ntab = 0;
    double etamax = 50.0, deta = 0.0;
    double *kcoef = nullptr, *dcoef = nullptr;  // device, 4 (ntab - 1) each
    double klimit = 0.0, dlimit = 0.0;          // values at etamax
    double ratio = 1.15;
    double kfmin_saved = -1.0, kfmax_saved = -1.0;  // < 0: none yet
    double *kfl = nullptr;                          // device, ladder
    int kfl_cap = 0;
    void *ws = nullptr;  // workspace in the plan precision
    size_t ws_bytes = 0;
    int last_nsp = 0;
};

void lmgp_free(Plan &P) {
    LmgpState *L = P.lmgp;
    if (!L) return;
    if (L->kcoef) (void)hipFree(L->kcoef);
    if (L->dcoef) (void)hipFree(L->dcoef);
    if (L->kfl) (void)hipFree(L->kfl);
    if (L->ws) (void)hipFree(L->ws);
    delete L;
    P.lmgp = nullptr;
}

#define GRID_LOOP(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

// ---------------------------------------------------------------- MGP table
// K(eta_k) = coe sum_{i=1..M} G(x) / t_i^(1/6), x = eta_k / t_i^(1/3), t_i = i / M,
// coe = 4/5 C_TF 5/6 / M.  G is the cubic spline of the inverse Lindhard
// function (lambda = mu = 1) on the log grid gp[0..ng), with x clamped to
// [gp[0], gp[ng-1]] (splev ext=3), as the reference evaluates it.
__device__ __forceinline__ double gspline(double x, const double *gp, const double *gc, int ng, double lg0,
                                          double dlg) {
    x = x < gp[0] ? gp[0] : (x > gp[ng - 1] ? gp[ng - 1] : x);
    int k = (int)((log10(x) - lg0) / dlg);
    k = k < 0 ? 0 : (k > ng - 2 ? ng - 2 : k);
    while (k > 0 && x < gp[k]) --k;
    while (k < ng - 2 && x >= gp[k + 1]) ++k;
    const double t = x - gp[k];
    const double *c = gc + 4 * k;
    return c[0] + t * (c[1] + t * (c[2] + t * c[3]));
}

__global__ void k_mgp_table(int neta, double deta, int maxpoints, const double *gp, const double *gc, int ng,
                            double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= neta) return;
    const double eta = k * deta;
    const double dt = 1.0 / maxpoints;
    const double lg0 = log10(gp[0]), dlg = (log10(gp[ng - 1]) - lg0) / (ng - 1);
    double s = 0.0;
    for (int i = 1; i <= maxpoints; ++i) {
        const double t = i * dt;
        s += gspline(eta / cbrt(t), gp, gc, ng, lg0, dlg) / pow(t, 1.0 / 6.0);
    }
    const double ctf = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
    out[k] = s * (4.0 / 5.0 * ctf * 5.0 / 6.0 * dt);
}

// ---------------------------------------------------------------- kernels
struct LmgpGeo {
    double B[9];
    int nx, ny, nz, nzc, full;
    long long nc;
};

__device__ __forceinline__ double tab_eval(const double *coef, int ntab, double deta, double etamax, double limit,
                                           double eta) {
    if (!(eta < etamax)) return limit;
    int k = (int)(eta / deta);
    if (k > ntab - 2) k = ntab - 2;
    const double t = eta - k * deta;
    const double *c = coef + 4 * k;
    return c[0] + t * (c[1] + t * (c[2] + t * c[3]));
}

// K and K'/kf at one half-grid point, with the full-grid Nyquist averaging
__device__ __forceinline__ void kvals(const LmgpGeo &g, long long i, const double *kc, const double *dc, int ntab,
                                      double deta, double etamax, double klim, double dlim, double kf, double &kv,
                                      double &dv) {
    const int iz = (int)(i % g.nzc);
    const long long r = i / g.nzc;
    const int iy = (int)(r % g.ny), ix = (int)(r / g.ny);
    if (ix == 0 && iy == 0 && iz == 0) {
        kv = dv = 0.0;
        return;
    }
    int fx = ffreq(ix, g.nx), fy = ffreq(iy, g.ny), fz = iz;
    if (g.full) fz = wrapf(fz, g.nz);
    double q = sqrt(gsq(g.B, fx, fy, fz));
    kv = tab_eval(kc, ntab, deta, etamax, klim, q / kf);
    dv = tab_eval(dc, ntab, deta, etamax, dlim, q / kf) / kf;
    if (g.full && (((g.nx & 1) == 0 && ix == g.nx / 2) || ((g.ny & 1) == 0 && iy == g.ny / 2) ||
                   ((g.nz & 1) == 0 && iz == g.nz / 2))) {
        q = sqrt(gsq(g.B, wrapf(-fx, g.nx), wrapf(-fy, g.ny), wrapf(-fz, g.nz)));
        kv = 0.5 * (kv + tab_eval(kc, ntab, deta, etamax, klim, q / kf));
        dv = 0.5 * (dv + tab_eval(dc, ntab, deta, etamax, dlim, q / kf) / kf);
    }
}

// PM[0] = K S / N, PM[1] = (K'/kf) S / N
template <typename T>
__global__ void k_lmgp_gmul(const cpx<T> *S, cpx<T> *PM, LmgpGeo g, const double *kc, const double *dc, int ntab,
                            double deta, double etamax, double klim, double dlim, double kf, double scale) {
    GRID_LOOP(i, g.nc) {
        double kv, dv;
        kvals(g, i, kc, dc, ntab, deta, etamax, klim, dlim, kf, kv, dv);
        const cpx<T> s = S[i];
        PM[i] = cpx<T>{(T)(s.x * kv * scale), (T)(s.y * kv * scale)};
        PM[g.nc + i] = cpx<T>{(T)(s.x * dv * scale), (T)(s.y * dv * scale)};
    }
}

// pot2G (=|+=) K W / N
template <typename T>
__global__ void k_lmgp_acc(const cpx<T> *W, cpx<T> *acc, int first, LmgpGeo g, const double *kc, int ntab,
                           double deta, double etamax, double klim, double kf, double scale) {
    GRID_LOOP(i, g.nc) {
        double kv, dv;
        kvals(g, i, kc, kc, ntab, deta, etamax, klim, klim, kf, kv, dv);
        const cpx<T> w = W[i];
        cpx<T> a = first ? cpx<T>{(T)0, (T)0} : acc[i];
        a.x = (T)((double)a.x + w.x * kv * scale);
        a.y = (T)((double)a.y + w.y * kv * scale);
        acc[i] = a;
    }
}

__device__ __forceinline__ double kf_of(double r, double kfmax) {
    const double kf = 2.0 * cbrt(3.0 * M_PI * M_PI * r);
    return kf > kfmax ? kfmax : kf;
}

// interval i of k_F between kfl[i] and kfl[i+1] (lwt.py:226-294, linear):
// pot1 = p0 (1-t) + p1 t, pot3 = m0 (1-t) + m1 t; for i == 0 also the points
// below the ladder (pot1 = p0 t, t = (kf - kfl[0]) / kfl[0]).
template <typename T>
__global__ void k_lmgp_interp(const T *rho, long long n, double kfmax, double lo, double hi, int first, const T *p0,
                              const T *m0, const T *p1, const T *m1, T *pot1, T *pot3) {
    GRID_LOOP(i, n) {
        const double kf = kf_of((double)rho[i], kfmax);
        if (kf > lo && kf < hi + 1e-18) {
            const double t = (kf - lo) / (hi - lo), t0 = 1.0 - t;
            pot1[i] = (T)((double)p0[i] * t0 + (double)p1[i] * t);
            pot3[i] = (T)((double)m0[i] * t0 + (double)m1[i] * t);
        } else if (first && kf < lo + 1e-18) {
            pot1[i] = (T)((double)p0[i] * (kf - lo) / lo);
        }
    }
}

// w_j rho^alpha (rho clamped to 1e-30 where not positive): hat weight of
// sample j, plus the below-ladder weight for j == 0
template <typename T>
__global__ void k_lmgp_wfield(const T *rho, long long n, double kfmax, const double *kfl, int nsp, int j, double alpha,
                              int b56, T *out) {
    const double lo = kfl[j];
    const double left = j > 0 ? kfl[j - 1] : 0.0;
    const double right = j + 1 < nsp ? kfl[j + 1] : 0.0;
    GRID_LOOP(i, n) {
        const double r = (double)rho[i];
        const double kf = kf_of(r, kfmax);
        double w = 0.0;
        if (j + 1 < nsp && kf > lo && kf < right + 1e-18) w += 1.0 - (kf - lo) / (right - lo);
        if (j > 0 && kf > left && kf < lo + 1e-18) w += (kf - left) / (lo - left);
        if (j == 0 && kf < lo + 1e-18) w += (kf - lo) / lo;
        const double rc = r > 0.0 ? r : 1e-30;
        const double ra = b56 ? sqrt(rc) * cbrt(rc) : pow(rc, alpha);
        out[i] = (T)(ra * w);
    }
}

// final point-wise pass (lwt.py:307-338): energy partials and potential
template <typename T>
__global__ void k_lmgp_final(const T *rho, long long n, double kfmax, double rhomax, double alpha, double beta,
                             double ldw, int b56, const T *pot1, const T *pot2, const T *pot3, T *pot, int want_v,
                             double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    const double ld = ldw > 0.1 ? ldw : 0.1;
    GRID_LOOP(i, n) {
        const double r = (double)rho[i];
        double f = 1.0;
        if (r < 1e-6) f = pow(fabs(r), ld) / pow(rhomax, ld);
        if (r < 0.0) f = 0.0;
        const double rc = r > 0.0 ? r : 1e-30;
        const double ra = b56 ? sqrt(rc) * cbrt(rc) : pow(rc, alpha);
        const double rb = b56 ? ra : pow(rc, beta);
        const double p1 = (double)pot1[i] * f;
        acc += ra * p1;
        if (want_v) {
            const double kf = kf_of(r, kfmax);
            const double v = p1 * alpha * ra / rc + (double)pot2[i] * f * beta * rb / rc +
                             (double)pot3[i] * f * (kf / 3.0) * ra / rc;
            pot[i] = (T)((double)pot[i] + v);
        }
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// min and max of rho over the grid (block partials [2 per block])
template <typename T> __global__ void k_minmax(const T *x, long long n, double *partials) {
    __shared__ double smin[16], smax[16];
    double lo = 1e300, hi = -1e300;
    GRID_LOOP(i, n) {
        const double v = (double)x[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        smin[wid] = lo;
        smax[wid] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) {
            lo = fmin(lo, smin[w]);
            hi = fmax(hi, smax[w]);
        }
        partials[2 * blockIdx.x] = lo;
        partials[2 * blockIdx.x + 1] = hi;
    }
}

__global__ void k_minmax_final(const double *partials, int nb, double *out) {
    __shared__ double smin[16], smax[16];
    double lo = 1e300, hi = -1e300;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        lo = fmin(lo, partials[2 * b]);
        hi = fmax(hi, partials[2 * b + 1]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        smin[wid] = lo;
        smax[wid] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) {
            lo = fmin(lo, smin[w]);
            hi = fmax(hi, smax[w]);
        }
        out[0] = lo;
        out[1] = hi;
    }
}

// energies[E_NL] = e dV, energies[E_SUM] += e dV
__global__ void __launch_bounds__(kFinalThreads)
    k_lmgp_energy(const double *partials, int nb, double dV, double *energies) {
    __shared__ double sm[kFinalThreads];
    sum_columns(partials, nb, 1, sm);
    if (threadIdx.x == 0) {
        energies[OFDFT_E_NL] = sm[0] * dV;
        energies[OFDFT_E_SUM] += sm[0] * dV;
    }
}

#undef GRID_LOOP

// guess_kf_bound (lwt.py:36-82) with the bounds saved across calls
static void kf_bounds(LmgpState &L, double kf_l, double kf_r, double &kfmin, double &kfmax) {
    const double kftol = 1e-3;
    double kmin_prev = L.kfmin_saved, kmax_prev = L.kfmax_saved;
    bool have_min = false, have_max = false;
    if (kmin_prev < 0) {
        kmin_prev = 10.0;
        kfmin = kf_l;
        have_min = true;
    }
    if (kmax_prev < 0) {
        kmax_prev = 1.0;
        kfmax = kf_r;
        have_max = true;
    }
    if (!have_min || kfmin > kmin_prev) kfmin = kmin_prev;
    if (!have_max || kfmax < kmax_prev) kfmax = kmax_prev;
    if (kfmin > kf_l) {
        const double kfl[10] = {1e-5, 1e-4, 5e-3, 1e-3, 5e-3, 1e-2, 5e-2, 0.1, 0.5, 1.0};
        const double ratio = kf_l / kfmin;
        for (int i = 9; i > 0; --i)
            if (ratio > kfl[i]) {
                kfmin *= kfl[i];
                break;
            }
    }
    if (kfmin < kftol) kfmin = kftol;
    if (kfmax < kf_r) kfmax += (std::nearbyint((kf_r - kfmax) / 0.5) + 1.0) * 0.5;
    L.kfmin_saved = kfmin;
    L.kfmax_saved = kfmax;
}

size_t lmgp_ws_bytes(const Plan &P) {
    const size_t rs = P.prec == OFDFT_FP64 ? 8 : 4;
    // spectra: S, PM (2), W, pot2G ; fields: p0, m0, p1, m1, wf, pot1, pot3, pot2
    return (size_t)5 * P.Nc * 2 * rs + (size_t)8 * P.N * rs;
}

template <typename T> void lmgp_eval(Plan &P, const T *rho, T *pot, double *energies, hipStream_t s) {
    OFDFT_REQUIRE(std::fabs(P.alpha - P.beta) < 1e-9, "LMGP with alpha != beta is not supported");
    LmgpState &L = *P.lmgp;
    const long long n = P.N, nc = P.Nc;
    // rho range -> k_F range (k_F is monotone in rho)
    const int nbm = std::min(reduce_blocks(n), 1024);
    hipLaunchKernelGGL((k_minmax<T>), dim3(nbm), dim3(256), 0, s, rho, n, P.red);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, P.red, nbm, P.scal + 220);
    OFDFT_HIP(hipGetLastError());
    OFDFT_HIP(hipMemcpyAsync(P.pinned + 100, P.scal + 220, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    OFDFT_HIP(hipStreamSynchronize(s));
    const double rmin = P.pinned[100], rmax = P.pinned[101];
    const double kf_l = 2.0 * std::cbrt(3.0 * M_PI * M_PI * rmin), kf_r = 2.0 * std::cbrt(3.0 * M_PI * M_PI * rmax);
    if (std::fabs(kf_r - kf_l) < 1e-8) {  // homogeneous gas: no nonlocal term (lwt.py:112-120)
        if (energies) {
            OFDFT_HIP(hipMemsetAsync(energies + OFDFT_E_NL, 0, sizeof(double), s));
        }
        return;
    }
    double kfmin, kfmax;
    kf_bounds(L, kf_l, kf_r, kfmin, kfmax);
    const int nsp = (int)std::ceil(std::log(kfmax / kfmin) / std::log(L.ratio)) + 1;
    std::vector<double> kfl(nsp);
    for (int j = 0; j < nsp; ++j) kfl[j] = kfmin * std::pow(L.ratio, (double)j);
    kfl[0] -= 1e-16;
    kfl[nsp - 1] += 1e-16;
    OFDFT_REQUIRE(nsp >= 2 && nsp <= L.kfl_cap, "LMGP: k_F ladder out of range");
    OFDFT_HIP(hipMemcpyAsync(L.kfl, kfl.data(), nsp * sizeof(double), hipMemcpyHostToDevice, s));
    L.last_nsp = nsp;
    // workspace
    cpx<T> *S = reinterpret_cast<cpx<T> *>(L.ws);
    cpx<T> *PM = S + nc;
    cpx<T> *W = S + 3 * nc;
    cpx<T> *acc = S + 4 * nc;
    T *f0 = reinterpret_cast<T *>(S + 5 * nc);
    T *p0 = f0, *m0 = f0 + n, *p1 = f0 + 2 * n, *m1 = f0 + 3 * n, *wf = f0 + 4 * n, *pot1 = f0 + 5 * n,
      *pot3 = f0 + 6 * n, *pot2 = f0 + 7 * n;
    LmgpGeo g;
    for (int i = 0; i < 9; ++i) g.B[i] = P.B[i];
    g.nx = P.nr[0];
    g.ny = P.nr[1];
    g.nz = P.nr[2];
    g.nzc = P.nzc;
    g.full = P.full;
    g.nc = nc;
    const double scale = 1.0 / (double)n;
    const int b56 = std::fabs(P.beta - 5.0 / 6.0) < 1e-14 && std::fabs(P.alpha - 5.0 / 6.0) < 1e-14;
    const dim3 gb(std::min<long long>((nc + 255) / 256, 4096)), gr(std::min<long long>((n + 255) / 256, 4096));
    // S = FFT(rho^beta) (rho clamped to 1e-30 where not positive)
    Prologue pro{};
    pro.code[0] = OP_POW_BETA;
    pro.beta = P.beta;
    pro.b56 = b56;
    fft_r2c<T>(P, rho, S, 1, n, nc, 1.0, &pro, nullptr, s);
    OFDFT_HIP(hipMemsetAsync(pot1, 0, 2 * n * sizeof(T), s));  // pot1 and pot3 (adjacent fields)
    auto sample = [&](int j, T *pj, T *mj) {
        hipLaunchKernelGGL((k_lmgp_gmul<T>), gb, dim3(256), 0, s, S, PM, g, L.kcoef, L.dcoef, L.ntab, L.deta, L.etamax,
                           L.klimit, L.dlimit, kfl[j], scale);
        OFDFT_HIP(hipGetLastError());
        // p_j, m_j: two C2R transforms of PM (pj and mj are consecutive fields)
        fft_c2r<T>(P, PM, PM, pj, 2, nc, nc, mj - pj, 1.0, s);
    };
    sample(0, p0, m0);
    for (int j = 0; j < nsp; ++j) {
        if (j + 1 < nsp) {
            sample(j + 1, p1, m1);
            hipLaunchKernelGGL((k_lmgp_interp<T>), gr, dim3(256), 0, s, rho, n, kfmax, kfl[j], kfl[j + 1], j == 0 ? 1 : 0,
                               p0, m0, p1, m1, pot1, pot3);
            OFDFT_HIP(hipGetLastError());
        }
        // pot2G += K_j FFT(rho^alpha w_j)
        hipLaunchKernelGGL((k_lmgp_wfield<T>), gr, dim3(256), 0, s, rho, n, kfmax, L.kfl, nsp, j, P.alpha, b56, wf);
        OFDFT_HIP(hipGetLastError());
        fft_r2c<T>(P, wf, W, 1, n, nc, 1.0, nullptr, nullptr, s);
        hipLaunchKernelGGL((k_lmgp_acc<T>), gb, dim3(256), 0, s, W, acc, j == 0 ? 1 : 0, g, L.kcoef, L.ntab, L.deta,
                           L.etamax, L.klimit, kfl[j], scale);
        OFDFT_HIP(hipGetLastError());
        std::swap(p0, p1);
        std::swap(m0, m1);
    }
    fft_c2r<T>(P, acc, acc, pot2, 1, nc, nc, n, 1.0, s);
    const int nb = std::min(reduce_blocks(n), 1024);
    hipLaunchKernelGGL((k_lmgp_final<T>), dim3(nb), dim3(256), 0, s, rho, n, kfmax, rmax, P.alpha, P.beta, 1.0 / 6.0,
                       b56, pot1, pot2, pot3, pot, pot != nullptr ? 1 : 0, P.red);
    OFDFT_HIP(hipGetLastError());
    if (energies) hipLaunchKernelGGL(k_lmgp_energy, dim3(1), dim3(kFinalThreads), 0, s, P.red, nb, P.dV, energies);
    OFDFT_HIP(hipGetLastError());
}

template void lmgp_eval<float>(Plan &, const float *, float *, double *, hipStream_t);
template void lmgp_eval<double>(Plan &, const double *, double *, double *, hipStream_t);

}  // namespace ofdft

using namespace ofdft;

struct ofdft_plan : public Plan {};

extern "C" {

int ofdft_mgp_kernel_table(int neta, double etamax, int maxpoints, int ng, const double *gp, const double *gcoef,
                           double *out) {
    double *d = nullptr, *dgp = nullptr, *dgc = nullptr;
    OFDFT_LEGACY_LOCK;
    try {
        OFDFT_REQUIRE(neta >= 4 && etamax > 0 && maxpoints >= 1 && ng >= 4 && gp && gcoef && out,
                      "mgp_kernel_table: bad arguments");
        const double deta = etamax / (neta - 1);
        OFDFT_HIP(hipMalloc(&d, neta * sizeof(double)));
        OFDFT_HIP(hipMalloc(&dgp, ng * sizeof(double)));
        OFDFT_HIP(hipMalloc(&dgc, (size_t)4 * (ng - 1) * sizeof(double)));
        OFDFT_HIP(hipMemcpy(dgp, gp, ng * sizeof(double), hipMemcpyHostToDevice));
        OFDFT_HIP(hipMemcpy(dgc, gcoef, (size_t)4 * (ng - 1) * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mgp_table, dim3((neta + 127) / 128), dim3(128), 0, 0, neta, deta, maxpoints, dgp, dgc, ng,
                           d);
        OFDFT_HIP(hipGetLastError());
        OFDFT_HIP(hipMemcpy(out, d, neta * sizeof(double), hipMemcpyDeviceToHost));
        (void)hipFree(d);
        (void)hipFree(dgp);
        (void)hipFree(dgc);
        return 0;
    } catch (const std::exception &ex) {
        if (d) (void)hipFree(d);
        if (dgp) (void)hipFree(dgp);
        if (dgc) (void)hipFree(dgc);
        set_error(ex.what());
        return 1;
    }
}

int ofdft_set_lmgp(ofdft_plan *plan, int ntab, double etamax, const double *kcoef, const double *dcoef,
                   double ratio) {
    try {
        OFDFT_REQUIRE(plan != nullptr, "set_lmgp: null plan");
        OFDFT_LEGACY_LOCK;
        Plan &P = *plan;
        if (ntab <= 0) {
            lmgp_free(P);
            return 0;
        }
        OFDFT_REQUIRE(ntab >= 4 && kcoef && dcoef && etamax > 0 && ratio > 1.0, "set_lmgp: bad arguments");
        if (!P.lmgp) P.lmgp = new LmgpState();
        LmgpState &L = *P.lmgp;
        const size_t cb = (size_t)4 * (ntab - 1) * sizeof(double);
        if (L.ntab != ntab) {
            if (L.kcoef) (void)hipFree(L.kcoef);
            if (L.dcoef) (void)hipFree(L.dcoef);
            L.kcoef = L.dcoef = nullptr;
            OFDFT_HIP(hipMalloc(&L.kcoef, cb));
            OFDFT_HIP(hipMalloc(&L.dcoef, cb));
        }
        OFDFT_HIP(hipMemcpy(L.kcoef, kcoef, cb, hipMemcpyHostToDevice));
        OFDFT_HIP(hipMemcpy(L.dcoef, dcoef, cb, hipMemcpyHostToDevice));
        L.ntab = ntab;
        // the k_F ladder at its largest (4096 samples): evaluations allocate nothing
        if (L.kfl_cap < 4096) {
            if (L.kfl) (void)hipFree(L.kfl);
            L.kfl = nullptr;
            L.kfl_cap = 0;
            OFDFT_HIP(hipMalloc(&L.kfl, 4096 * sizeof(double)));
            L.kfl_cap = 4096;
        }
        L.etamax = etamax;
        L.deta = etamax / (ntab - 1);
        // values at the last knot: the end of the last cubic piece
        const double h = L.deta;
        const double *c = kcoef + 4 * (ntab - 2), *cd = dcoef + 4 * (ntab - 2);
        L.klimit = c[0] + h * (c[1] + h * (c[2] + h * c[3]));
        L.dlimit = cd[0] + h * (cd[1] + h * (cd[2] + h * cd[3]));
        L.ratio = ratio;
        L.kfmin_saved = L.kfmax_saved = -1.0;
        const size_t need = lmgp_ws_bytes(P);
        if (L.ws_bytes < need) {
            if (L.ws) (void)hipFree(L.ws);
            L.ws = nullptr;
            OFDFT_HIP(hipMalloc(&L.ws, need));
            L.ws_bytes = need;
        }
        return 0;
    } catch (const std::exception &ex) {
        set_error(ex.what());
        return 1;
    }
}

int ofdft_lmgp_reset(ofdft_plan *plan) {
    if (!plan || !plan->lmgp) return 0;
    plan->lmgp->kfmin_saved = plan->lmgp->kfmax_saved = -1.0;
    return 0;
}

}  // extern "C"
