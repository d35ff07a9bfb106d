// One TotalFunctional evaluation (total_functional.py:115-146) for the terms
// TF + vW + WT + PZ-LDA + Hartree + local PP.
//
// Pipeline (3 batched spectra for vW, WT and Hartree):
//   1. z pass R2C with the point-wise prologue applied on load
//      (sqrt(rhom) for vW, max(rho,1e-30)^beta for WT, rho for Hartree),
//      also summing rho for rho0 = mean(rho);
//   2. y and x passes;
//   3. G-space multiply: vW 1/2 G^2 (vw.py:45-49), WT Lindhard kernel
//      evaluated from |G| (kernel.py:13-136,325-334), Hartree 4 pi / G^2
//      (hartree.py:34-36), times 1/N so the unscaled forward+inverse pair
//      equals DFTpy's (rfftn * dV, irfftn / dV);
//   4. inverse x, y passes and z pass C2R;
//   5. post kernel: the total potential and the FP64 energy sums of every
//      term (tf.py:13-80, vw.py:29-94, wt.py:12-28,124-125,
//      semilocal_xc.py:252-290, hartree.py:38, pseudo/__init__.py:228).
#include <cmath>

#include "gspace.h"
#include "internal.h"
#include "pointwise.h"
#include "stream.h"
#include "reduce.h"

namespace ofdft {

template <typename T>
__global__ void k_gspace(cpx<T> *__restrict__ spec, long long nc, GParams p, const double *gate) {
    if (gate_off(gate)) return;
    // nc counts the pitched layout; columns iz >= nzc are padding
    const double rho0 = p.rho0_dev ? *p.rho0_dev : p.rho0;
    const double tkf = 2.0 * cbrt(3.0 * rho0 * M_PI * M_PI);
    const double ctf = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
    const double wtfac = 5.0 / (9.0 * p.alpha * p.beta * pow(rho0, p.alpha + p.beta - 5.0 / 3.0)) * ctf;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        const int iz = (int)(i % p.pitch);
        if (iz >= p.nzc) continue;
        const long long rest = i / p.pitch;
        const int iy = (int)(rest % p.ny);
        const int ix = (int)(rest / p.ny);
        int fx = ffreq(ix, p.nx), fy = ffreq(iy, p.ny), fz = iz;
        const bool g0 = (ix == 0 && iy == 0 && iz == 0);
        KVals k;
        if (p.full) {
            // full-grid convention: z Nyquist is negative; Nyquist aliases average
            fz = wrapf(fz, p.nz);
            const bool nyq = ((p.nx & 1) == 0 && ix == p.nx / 2) || ((p.ny & 1) == 0 && iy == p.ny / 2) ||
                             ((p.nz & 1) == 0 && iz == p.nz / 2);
            k = kernel_vals(p, gsq(p.B, fx, fy, fz), g0, tkf, wtfac);
            if (nyq) {
                const KVals k2 = kernel_vals(
                    p, gsq(p.B, wrapf(-fx, p.nx), wrapf(-fy, p.ny), wrapf(-fz, p.nz)), g0, tkf, wtfac);
                k.vw = 0.5 * (k.vw + k2.vw);
                k.wt = 0.5 * (k.wt + k2.wt);
                k.h = 0.5 * (k.h + k2.h);
            }
        } else {
            k = kernel_vals(p, gsq(p.B, fx, fy, fz), g0, tkf, wtfac);
        }
        if (p.s_vw >= 0) {
            cpx<T> &v = spec[p.s_vw * p.bs + i];
            const double f = k.vw * p.scale;
            v.x = (T)(v.x * f);
            v.y = (T)(v.y * f);
        }
        if (p.s_wt >= 0) {
            cpx<T> &v = spec[p.s_wt * p.bs + i];
            const double f = k.wt * p.scale;
            v.x = (T)(v.x * f);
            v.y = (T)(v.y * f);
        }
        if (p.s_h >= 0) {
            cpx<T> &v = spec[p.s_h * p.bs + i];
            const double f = k.h * p.scale;
            v.x = (T)(v.x * f);
            v.y = (T)(v.y * f);
        }
        if (p.s_wt2 >= 0) {
            cpx<T> &v = spec[p.s_wt2 * p.bs + i];
            const double f = k.wt * p.scale;
            v.x = (T)(v.x * f);
            v.y = (T)(v.y * f);
        }
    }
}

// rho0 = sum / N with the kernel-reuse threshold of wt.py:64, and the WT
// constants (tkf, wtfac) of the cached rho0 at rho0_cache[kWtSlot], [kWtSlot + 1]
constexpr int kWtSlot = 120;  // offset from the rho0 cache (P.scal[0]) in the plan's device scalars
__global__ void __launch_bounds__(kFinalThreads)
    k_rho0_update(const double *partials, int nb, double inv_n, double *rho0_cache, double alpha, double beta,
                  const double *gate) {
    if (gate_off(gate)) return;
    __shared__ double sm[kFinalThreads];
    sum_columns(partials, nb, 1, sm);
    if (threadIdx.x == 0) {
        const double r = sm[0] * inv_n;
        if (fabs(*rho0_cache - r) > 1e-6) *rho0_cache = r;
        wt_consts_at(*rho0_cache, alpha, beta, rho0_cache[kWtSlot], rho0_cache[kWtSlot + 1]);
    }
}

static int grid_blocks(long long n, int per = 256 * 4) {
    long long b = (n + per - 1) / per;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (int)b;
}

template <typename T>
bool xfused_fixed(int n, int var, cpx<T> *data, int batch, long long bs, long long ncols, long long pitch,
                  long long vp, const void *tw, const GParams &gp, hipStream_t s, const double *gate);

// the WT exponents take the single-spectrum form of wt.py:18-22
static bool wt_alpha_is_beta(const Plan &P) { return std::fabs(P.alpha - P.beta) < 1e-9; }

// the x pass can run fused with the G-space multiply (size-specialised x kernel)
static bool x_fusable(const Plan &P) { return P.use_fixed && P.nr[0] > 1 && P.nr[1] * P.nr[2] > 1 && fixed_size(P.nr[0]); }

// G-space multiply of the ns pipeline spectra: fused into the x pass when
// `fused` (the caller skipped the x passes of its R2C / C2R), else the
// stand-alone kernel between complete 3D transforms
template <typename T>
static void gspace_apply(Plan &P, cpx<T> *spec, int ns, long long ncp, long long pitch, GParams &gp, bool fused,
                         hipStream_t s) {
    for (int i = 0; i < 4; ++i) gp.kind[i] = -1;
    if (gp.s_vw >= 0) gp.kind[gp.s_vw] = 0;
    if (gp.s_wt >= 0) gp.kind[gp.s_wt] = 1;
    if (gp.s_h >= 0) gp.kind[gp.s_h] = 2;
    if (gp.s_wt2 >= 0) gp.kind[gp.s_wt2] = 1;
    if (fused) {
        const long long ncols = (long long)P.nr[1] * pitch;
        OFDFT_REQUIRE(xfused_fixed<T>(P.nr[0], P.var_cols[0], spec, ns, ncp, ncols, pitch, P.nzc, P.tw[0], gp, s,
                                      P.gate),
                      "fused x pass unavailable");
    } else {
        hipLaunchKernelGGL((k_gspace<T>), dim3(grid_blocks(ncp)), dim3(256), 0, s, spec, ncp, gp, P.gate);
    }
    OFDFT_HIP(hipGetLastError());
}

// Energy-only evaluation: the spectral terms as sums over the half spectrum,
// E = sum_G w(kz) f(G) |F(G)|^2 / N (Parseval; w = 2 for the kz columns whose
// mirror is not stored), with f the same multiplier the potential path
// applies.  A flat grid stride over the stored points.  The
// sums land in the k_post partial columns of vW (times y), WT and Hartree
// (times 1/2), which k_post leaves at zero when it has no fields.
template <typename T>
__global__ void __launch_bounds__(256) k_gspace_energy(const cpx<T> *__restrict__ spec, int ns, GParams p,
                                                       double y, double *partials, const double *gate) {
    if (gate_off(gate)) return;
    __shared__ double sm[16];
    double tkf = 0.0, wtfac = 0.0;
    if (p.s_wt >= 0) wt_consts(p, tkf, wtfac);
    double acc_vw = 0.0, acc_wt = 0.0, acc_h = 0.0;
    // flat grid-stride over the stored half spectrum (every lane busy; rows
    // are shorter than a workgroup), 32-bit index arithmetic (the half
    // spectrum of a plan has far fewer than 2^31 points); |G|^2 once per
    // point for all spectra
    const unsigned nzc = (unsigned)p.nzc, ny = (unsigned)p.ny;
    const unsigned total = (unsigned)p.nx * ny * nzc;
    for (unsigned idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const unsigned row = idx / nzc, iz = idx - row * nzc;
        const unsigned ix = row / ny, iy = row - ix * ny;
        const int fx = ffreq((int)ix, p.nx), fy = ffreq((int)iy, p.ny);
        const bool g0 = (idx == 0);
        double gg, gg2 = 0.0;
        bool nyq = false;
        if (p.full) {
            const int fz = wrapf((int)iz, p.nz);
            gg = gsq(p.B, fx, fy, fz);
            nyq = ((p.nx & 1) == 0 && (int)ix == p.nx / 2) || ((p.ny & 1) == 0 && (int)iy == p.ny / 2) ||
                  ((p.nz & 1) == 0 && (int)iz == p.nz / 2);
            if (nyq) gg2 = gsq(p.B, wrapf(-fx, p.nx), wrapf(-fy, p.ny), wrapf(-fz, p.nz));
        } else {
            gg = gsq(p.B, fx, fy, (int)iz);
        }
        const double w = (iz == 0 || ((p.nz & 1) == 0 && (int)iz == p.nz / 2)) ? 1.0 : 2.0;
        const long long e = (long long)row * p.pitch + iz;
        for (int sl = 0; sl < ns; ++sl) {
            const int kind = p.kind[sl];
            const cpx<T> v = spec[sl * p.bs + e];
            const double a2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
            double f = kernel_val<T>(p, kind, gg, g0, tkf, wtfac);
            if (nyq) f = 0.5 * (f + kernel_val<T>(p, kind, gg2, g0, tkf, wtfac));
            const double t = w * (f * p.scale) * a2;
            if (kind == 0)
                acc_vw += t;
            else if (kind == 1)
                acc_wt += t;
            else
                acc_h += t;
        }
    }
    const double sums[3] = {block_sum(acc_vw, sm), block_sum(acc_wt, sm), block_sum(acc_h, sm)};
    if (threadIdx.x == 0) {
        partials[(long long)blockIdx.x * 7 + 1] = y * sums[0];
        partials[(long long)blockIdx.x * 7 + 2] = sums[1];
        partials[(long long)blockIdx.x * 7 + 4] = 0.5 * sums[2];
    }
}

struct PostParams {
    int s_vw, s_wt, s_h;
    int s_wt2 = -1;  // alpha != beta: IFFT(K FFT(rho^alpha))
    const void *core = nullptr;  // NLCC: the XC term sees rho + core
    int b56;  // beta == 5/6
    unsigned terms;
    double x, y, alpha, beta;
    long long n;
    long long bs;  // work batch stride (reals)
    int want_v, want_e;
};

// PZ LDA constants (semilocal_xc.py:257-263)
constexpr double PZ_A = 0.0311, PZ_B = -0.048, PZ_C = 0.0020, PZ_D = -0.0116;
constexpr double PZ_G = -0.1423, PZ_B1 = 1.0529, PZ_B2 = 0.3334;

// Reciprocal and reciprocal square root: exact division for FP64; the
// hardware approximations (about 1 ulp) for the FP32-bulk pipeline, whose
// point-wise work is otherwise dominated by IEEE division sequences.
template <typename CT> __device__ __forceinline__ CT rcp_p(CT x) { return fm::rcp(x); }
template <typename CT> __device__ __forceinline__ CT rsq_p(CT x) { return fm::rsq(x); }

template <typename T, typename CT, int VEC>
__global__ void k_post(const T *__restrict__ rho, const T *__restrict__ work, const T *__restrict__ vloc,
                       T *__restrict__ pot, double *partials, PostParams p, const double *gate) {
    if (gate_off(gate)) return;
    const CT ctf = (CT)(0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0));
    const CT cbrt34pi = (CT)cbrt(3.0 / (4.0 * M_PI));
    const CT cx_e = (CT)(-3.0 / 4.0 * cbrt(3.0 / M_PI));
    const CT cx_v = (CT)cbrt(-3.0 / M_PI);
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const bool xc_core = (p.terms & OFDFT_TERM_XC) && p.core;
    pw_groups<VEC>(p.n, [&](long long i0) {
        // this group's inputs, one vector access per field present
        const PwVec<VEC, T> rv = pw_ld<VEC>(rho, i0);
        PwVec<VEC, T> av{}, cv{}, c2v{}, hv{}, lv{}, corev{}, out{};
        if (p.s_vw >= 0) av = pw_ld<VEC>(work + p.s_vw * p.bs, i0);
        if (p.s_wt >= 0) cv = pw_ld<VEC>(work + p.s_wt * p.bs, i0);
        if (p.s_wt >= 0 && p.s_wt2 >= 0) c2v = pw_ld<VEC>(work + p.s_wt2 * p.bs, i0);
        if (p.s_h >= 0) hv = pw_ld<VEC>(work + p.s_h * p.bs, i0);
        if (vloc) lv = pw_ld<VEC>(vloc, i0);
        if (xc_core) corev = pw_ld<VEC>(reinterpret_cast<const T *>(p.core), i0);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const CT r = (CT)rv.v[k];
            CT v = (CT)0;
            // one cube root and one square root serve every term (for beta = 5/6)
            const CT cr = fm::cbrt(r);
            const CT rc = r < (CT)1e-30 ? (CT)1e-30 : r;
            const CT c_rc = r < (CT)1e-30 ? (CT)1e-10 : cr;  // cbrt(max(rho, 1e-30))
            if (p.terms & OFDFT_TERM_TF) {
                const CT c2 = cr * cr;  // rho^(2/3)
                if (p.want_e) acc[0] += (double)(ctf * c2 * r);
                v += (CT)p.x * ((CT)(5.0 / 3.0) * ctf * c2);
        }
        if (p.s_vw >= 0) {
            const CT a = (CT)av.v[k];
            // 1 / sqrt(rho > 0 ? rho : 1e-30)
            const CT isq = rsq_p<CT>((r > (CT)0 && r < (CT)1e-30) ? r : rc);
            const CT vw = (CT)p.y * (a * isq);
            if (p.want_e && r > (CT)0) acc[1] += (double)(r * vw);
            v += vw;
        }
        if (p.s_wt >= 0 && p.s_wt2 < 0) {
            const CT c = (CT)cv.v[k];
            const CT rb_rc = p.b56 ? c_rc * rsq_p<CT>(rc) : fm::pow(rc, (CT)p.beta) * rcp_p<CT>(rc);  // rho^beta / rho
            const CT vnl = (CT)(2.0 * p.alpha) * rb_rc * c;
            if (p.want_e) acc[2] += (double)(vnl * r) / (2.0 * p.alpha);
            v += vnl;
        } else if (p.s_wt >= 0) {
            // alpha != beta (wt.py:23-25,44-56): cb = K * rho^beta, ca = K * rho^alpha
            const CT cb = (CT)cv.v[k];
            const CT ca = (CT)c2v.v[k];
            const CT ra = fm::pow(rc, (CT)p.alpha), rb = fm::pow(rc, (CT)p.beta);
            const CT irc = rcp_p<CT>(rc);
            if (p.want_e) acc[2] += (double)(ra * cb);
            v += (CT)p.alpha * ra * irc * cb + (CT)p.beta * rb * irc * ca;
        }
        if (p.terms & OFDFT_TERM_XC) {
            // xc/__init__.py:59-74: with a core density the XC term is
            // evaluated at rho + core (energy sum_r eps(rho+core)(rho+core) dV)
            CT rx = r, rc3 = cr;
            if (p.core) {
                rx = r + (CT)corev.v[k];
                rc3 = fm::cbrt(rx);
            }
            if (rc3 < (CT)1e-30) rc3 = (CT)1e-30;
            const CT rs = cbrt34pi * rcp_p<CT>(rc3);
            CT ec, vc;
            if (rs < (CT)1) {
                const CT lr = fm::log(rs);
                ec = (CT)PZ_A * lr + (CT)PZ_B + (CT)PZ_C * rs * lr + (CT)PZ_D * rs;
                vc = lr * ((CT)PZ_A + (CT)(2.0 / 3.0 * PZ_C) * rs) + (CT)PZ_B - (CT)(1.0 / 3.0 * PZ_A) +
                     (CT)(1.0 / 3.0 * (2 * PZ_D - PZ_C)) * rs;
            } else {
                const CT sr = fm::sqrt(rs);
                const CT iden = rcp_p<CT>((CT)1 + (CT)PZ_B1 * sr + (CT)PZ_B2 * rs);
                ec = (CT)PZ_G * iden;
                vc = ((CT)PZ_G + (CT)(7.0 / 6.0 * PZ_G * PZ_B1) * sr + (CT)(4.0 / 3.0 * PZ_G * PZ_B2) * rs) *
                     (iden * iden);
            }
            if (p.want_e) acc[3] += (double)((cx_e * rc3 + ec) * rx);
            v += cx_v * rc3 + vc;
        }
        if (p.s_h >= 0) {
            const CT vh = (CT)hv.v[k];
            if (p.want_e) acc[4] += 0.5 * (double)vh * (double)r;
            v += vh;
        }
        if (vloc) {
            const CT vl = (CT)lv.v[k];
            if (p.want_e) acc[5] += (double)vl * (double)r;
            v += vl;
        }
        acc[6] += (double)r;
        out.v[k] = (T)v;
        }
        if (p.want_v) pw_st<VEC>(pot, i0, out);
    });
    if (p.want_e) block_sums<7>(acc, partials);
}

// k_post with the element width the field length allows (stream.h)
template <typename T>
static void launch_post(int nb, hipStream_t s, const T *rho, const T *work, const T *vloc, T *pot, double *partials,
                        const PostParams &pp, const double *gate) {
    if (pw_vec(pp.n) == 4 && pw_vec(pp.bs) == 4 && pw_aligned({rho, work, vloc, pot, pp.core}))
        hipLaunchKernelGGL((k_post<T, T, 4>), dim3(nb), dim3(256), 0, s, rho, work, vloc, pot, partials, pp, gate);
    else
        hipLaunchKernelGGL((k_post<T, T, 1>), dim3(nb), dim3(256), 0, s, rho, work, vloc, pot, partials, pp, gate);
}

// energies[0..5] = term sums * dV (TF also * x), [6] = their sum + E_ion,
// [7] = electron count
__global__ void __launch_bounds__(kFinalThreads)
    k_energy_final(const double *__restrict__ partials, int nb, double *out, double dV, double x, double eion,
                   const double *gate) {
    if (gate_off(gate)) return;
    __shared__ double sm[kFinalThreads];
    sum_columns(partials, nb, 7, sm);
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int t = 0; t < 6; ++t) {
            const double e = sm[t] * dV * (t == 0 ? x : 1.0);
            out[t] = e;
            tot += e;
        }
        out[6] = tot + eion;
        out[7] = sm[6] * dV;
    }
}

// E only: forward transforms of the pipeline spectra, their energies summed
// in G space (k_gspace_energy) and the point-wise terms by k_post without
// fields; no inverse transforms and no potential
template <typename T>
static void eval_energy(Plan &P, const T *rho, int ns, int s_vw, int s_wt, int s_h, double *energies,
                        hipStream_t s) {
    cpx<T> *spec = reinterpret_cast<cpx<T> *>(P.spec);
    double *rho0_cache = P.scal + 0;
    Prologue pro{};
    if (s_vw >= 0) pro.code[s_vw] = OP_SQRT_RHOM;
    if (s_wt >= 0) pro.code[s_wt] = OP_POW_BETA;
    if (s_h >= 0) pro.code[s_h] = OP_IDENTITY;
    pro.beta = P.beta;
    pro.b56 = beta_is_5_6(P.beta);
    const bool auto_rho0 = (s_wt >= 0) && !(P.rho0 > 0);
    const long long pitch = pipe_pitch(P);
    const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
    fft_r2c<T>(P, rho, spec, ns, 0, ncp, 1.0, &pro, auto_rho0 ? P.red : nullptr, s, pitch, false);
    if (auto_rho0) {
        const int nb0 = r2c_num_partials(P, ns);
        hipLaunchKernelGGL(k_rho0_update, dim3(1), dim3(kFinalThreads), 0, s, P.red, nb0, 1.0 / (double)P.N, rho0_cache,
                           P.alpha, P.beta, P.gate);
        OFDFT_HIP(hipGetLastError());
    }
    GParams gp;
    for (int i = 0; i < 9; ++i) gp.B[i] = P.B[i];
    for (int i = 0; i < 4; ++i) gp.kind[i] = -1;
    if (s_vw >= 0) gp.kind[s_vw] = 0;
    if (s_wt >= 0) gp.kind[s_wt] = 1;
    if (s_h >= 0) gp.kind[s_h] = 2;
    gp.nx = P.nr[0];
    gp.ny = P.nr[1];
    gp.nz = P.nr[2];
    gp.nzc = P.nzc;
    gp.full = P.full;
    gp.scale = 1.0 / (double)P.N;
    gp.s_vw = s_vw;
    gp.s_wt = s_wt;
    gp.s_h = s_h;
    gp.sigma = P.sigma;
    gp.lbda = P.x;
    gp.alpha = P.alpha;
    gp.beta = P.beta;
    gp.rho0 = P.rho0 > 0 ? P.rho0 : 1.0;
    gp.rho0_dev = auto_rho0 ? rho0_cache : nullptr;
    gp.wt_dev = auto_rho0 ? rho0_cache + kWtSlot : nullptr;
    gp.bs = ncp;
    gp.pitch = (int)pitch;
    PostParams pp;
    pp.s_vw = pp.s_wt = pp.s_h = pp.s_wt2 = -1;
    pp.core = (P.terms & OFDFT_TERM_XC) ? P.core : nullptr;
    pp.terms = P.terms;
    pp.x = P.x;
    pp.y = P.y;
    pp.alpha = P.alpha;
    pp.beta = P.beta;
    pp.b56 = beta_is_5_6(P.beta);
    pp.n = P.N;
    pp.bs = P.N;
    pp.want_v = 0;
    pp.want_e = 1;
    const T *vloc = (P.terms & OFDFT_TERM_PP) ? reinterpret_cast<const T *>(P.vloc) : nullptr;
    const int nb = std::min(reduce_blocks(P.N), 1024);
    launch_post<T>(nb, s, rho, nullptr, vloc, nullptr, P.red, pp, P.gate);
    OFDFT_HIP(hipGetLastError());
    OFDFT_REQUIRE((long long)P.nr[0] * P.nr[1] * P.nzc < (1LL << 31), "energy pass: grid too large");
    hipLaunchKernelGGL((k_gspace_energy<T>), dim3(nb), dim3(256), 0, s, spec, ns, gp, P.y, P.red, P.gate);
    OFDFT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kFinalThreads), 0, s, P.red, nb, energies, P.dV, P.x, P.eion,
                       P.gate);
    OFDFT_HIP(hipGetLastError());
}

// G-space parameters of the pipeline's spectra (slots s_vw, s_wt, s_h of ns)
static GParams pipeline_gparams(const Plan &P, int s_vw, int s_wt, int s_h, bool auto_rho0, double *rho0_cache,
                                long long ncp, long long pitch, int s_wt2 = -1) {
    GParams gp;
    gp.s_wt2 = s_wt2;
    for (int i = 0; i < 9; ++i) gp.B[i] = P.B[i];
    gp.nx = P.nr[0];
    gp.ny = P.nr[1];
    gp.nz = P.nr[2];
    gp.nzc = P.nzc;
    gp.full = P.full;
    gp.scale = 1.0 / (double)P.N;
    gp.s_vw = s_vw;
    gp.s_wt = s_wt;
    gp.s_h = s_h;
    gp.sigma = P.sigma;
    gp.lbda = P.x;  // WT kernel lambda = x, mu = 1 (wt.py:66, kernel.py:331-334)
    gp.alpha = P.alpha;
    gp.beta = P.beta;
    gp.rho0 = P.rho0 > 0 ? P.rho0 : 1.0;
    gp.rho0_dev = auto_rho0 ? rho0_cache : nullptr;
    gp.wt_dev = auto_rho0 ? rho0_cache + kWtSlot : nullptr;
    gp.bs = ncp;
    gp.pitch = (int)pitch;
    return gp;
}

static PostParams pipeline_post(const Plan &P, int s_vw, int s_wt, int s_h, bool want_v, bool want_e,
                                int s_wt2 = -1) {
    PostParams pp;
    pp.s_wt2 = s_wt2;
    pp.core = (P.terms & OFDFT_TERM_XC) ? P.core : nullptr;
    pp.s_vw = s_vw;
    pp.s_wt = s_wt;
    pp.s_h = s_h;
    pp.terms = P.terms;
    pp.x = P.x;
    pp.y = P.y;
    pp.alpha = P.alpha;
    pp.beta = P.beta;
    pp.b56 = beta_is_5_6(P.beta);
    pp.n = P.N;
    pp.bs = P.N;
    pp.want_v = want_v;
    pp.want_e = want_e;
    return pp;
}

template <typename T>
void eval(Plan &P, const T *rho, unsigned calc, T *pot, double *energies, hipStream_t s) {
    const unsigned terms = P.terms;
    int ns = 0;
    const int s_vw = (terms & OFDFT_TERM_VW) ? ns++ : -1;
    const int s_wt = (terms & OFDFT_TERM_WT) ? ns++ : -1;
    const int s_h = (terms & OFDFT_TERM_HARTREE) ? ns++ : -1;
    // WT with alpha != beta transforms rho^alpha as a fourth spectrum
    const int s_wt2 = (s_wt >= 0 && !wt_alpha_is_beta(P)) ? ns : -1;
    cpx<T> *spec = reinterpret_cast<cpx<T> *>(P.spec);
    T *work = reinterpret_cast<T *>(P.work);
    double *rho0_cache = P.scal + 0;
    const bool energy_only = (calc & OFDFT_CALC_E) && energies != nullptr && !((calc & OFDFT_CALC_V) && pot) &&
                             !(terms & OFDFT_TERM_LMGP) && s_wt2 < 0;
    if (energy_only && ns > 0) {
        eval_energy<T>(P, rho, ns, s_vw, s_wt, s_h, energies, s);
        return;
    }
    if (ns > 0) {
        Prologue pro{};
        if (s_vw >= 0) pro.code[s_vw] = OP_SQRT_RHOM;
        if (s_wt >= 0) pro.code[s_wt] = OP_POW_BETA;
        if (s_h >= 0) pro.code[s_h] = OP_IDENTITY;
        pro.beta = P.beta;
        pro.b56 = beta_is_5_6(P.beta);
        const bool auto_rho0 = (s_wt >= 0) && !(P.rho0 > 0);
        const long long pitch = pipe_pitch(P);
        const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
        const bool fused = x_fusable(P);
        fft_r2c<T>(P, rho, spec, ns, 0, ncp, 1.0, &pro, auto_rho0 ? P.red : nullptr, s, pitch, fused);
        if (auto_rho0) {
            const int nb = r2c_num_partials(P, ns);
            hipLaunchKernelGGL(k_rho0_update, dim3(1), dim3(kFinalThreads), 0, s, P.red, nb, 1.0 / (double)P.N,
                               rho0_cache, P.alpha, P.beta, P.gate);
            OFDFT_HIP(hipGetLastError());
        }
        int nsa = ns;
        if (s_wt2 >= 0) {
            Prologue pa{};
            pa.code[0] = OP_POW_BETA;  // max(rho, 1e-30)^alpha
            pa.beta = P.alpha;
            pa.b56 = beta_is_5_6(P.alpha);
            fft_r2c<T>(P, rho, spec + (long long)s_wt2 * ncp, 1, 0, ncp, 1.0, &pa, nullptr, s, pitch, fused);
            nsa = ns + 1;
        }
        GParams gp = pipeline_gparams(P, s_vw, s_wt, s_h, auto_rho0, rho0_cache, ncp, pitch, s_wt2);
        gspace_apply<T>(P, spec, nsa, ncp, pitch, gp, fused, s);
        fft_c2r<T>(P, spec, spec, work, nsa, ncp, ncp, P.N, 1.0, s, pitch, fused);
    }
    PostParams pp = pipeline_post(P, s_vw, s_wt, s_h, (calc & OFDFT_CALC_V) && pot != nullptr,
                                  (calc & OFDFT_CALC_E) && energies != nullptr, s_wt2);
    const T *vloc = (terms & OFDFT_TERM_PP) ? reinterpret_cast<const T *>(P.vloc) : nullptr;
    const int nb = std::min(reduce_blocks(P.N), 1024);
    launch_post<T>(nb, s, rho, work, vloc, pot, P.red, pp, P.gate);
    OFDFT_HIP(hipGetLastError());
    if (pp.want_e) {
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kFinalThreads), 0, s, P.red, nb, energies, P.dV, P.x, P.eion,
                           P.gate);
        OFDFT_HIP(hipGetLastError());
    }
    if (terms & OFDFT_TERM_LMGP) {
        OFDFT_REQUIRE(P.lmgp != nullptr, "LMGP term without tables (ofdft_set_lmgp)");
        OFDFT_REQUIRE(P.gate == nullptr, "gated launch of the LMGP evaluation");
        lmgp_eval<T>(P, rho, pp.want_v ? pot : nullptr, pp.want_e ? energies : nullptr, s);
    }
}

template void eval<float>(Plan &, const float *, unsigned, float *, double *, hipStream_t);
template void eval<double>(Plan &, const double *, unsigned, double *, double *, hipStream_t);

// Energy density ('D') and second derivative ('V2') fields of the terms that
// have them in the reference, from rho and the inverse transforms an E+V
// evaluation left in the work slots: TF C rho^(5/3) (tf.py:26-35,68-73,
// unscaled by x as there) and x 10/9 C rho^(-1/3) (tf.py:47-50,77-79); vW
// y v_vW rho (vw.py:125-126); WT rho^alpha (K * rho^beta) (wt.py:44-56,
// 123-129, the alpha == beta form v rho / (2 alpha)); PZ-LDA d v_xc / d rho
// (semilocal_xc.py:291-303).  Hartree and the local PP have neither.
struct FieldParams {
    int s_vw, s_wt, s_wt2;
    const void *core;  // XC at rho + core
    unsigned terms;
    double x, y, alpha, beta;
    long long n, bs;
};

template <typename T>
__global__ void k_fields(const T *__restrict__ rho, const T *__restrict__ work, T *__restrict__ edens,
                         T *__restrict__ v2, FieldParams p) {
    const double ctf = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (long long)gridDim.x * blockDim.x) {
        const double r = (double)rho[i];
        const double rc = r < 1e-30 ? 1e-30 : r;
        double ed = 0.0, f2 = 0.0;
        if (p.terms & OFDFT_TERM_TF) {
            const double c = cbrt(r);
            ed += ctf * r * c * c;
            f2 += p.x * 10.0 / 9.0 * ctf / c;
        }
        if (p.s_vw >= 0) {
            const double isq = 1.0 / sqrt((r > 0.0 && r < 1e-30) ? r : rc);
            ed += p.y * (double)work[p.s_vw * p.bs + i] * isq * r;
        }
        if (p.s_wt >= 0) {
            const double cb = (double)work[p.s_wt * p.bs + i];
            if (p.s_wt2 < 0)
                ed += pow(rc, p.beta - 1.0) * cb * r;
            else
                ed += pow(rc, p.alpha) * cb;
        }
        if (p.terms & OFDFT_TERM_XC) {
            const double rx = p.core ? r + (double)reinterpret_cast<const T *>(p.core)[i] : r;
            const double cr = cbrt(rx);
            const double rs = cbrt(3.0 / (4.0 * M_PI)) / (cr < 1e-30 ? 1e-30 : cr);
            const double fx = -cbrt(3.0 / M_PI) / 3.0 * cr / rx;
            double fc;
            if (rs < 1.0) {
                fc = -PZ_A / 3.0 - (PZ_C / 9.0 * (log(rs) * 2.0 + 1.0) + PZ_D * 2.0 / 9.0) * rs;
            } else {
                const double ta = PZ_B1 * sqrt(rs), tb = PZ_B2 * rs, de = 1.0 + ta + tb;
                fc = PZ_G / 36.0 * (5.0 * ta + 7.0 * ta * ta + 8.0 * tb + 16.0 * tb * tb + 21.0 * ta * tb) /
                     (de * de * de);
            }
            f2 += fx + fc / rx;
        }
        if (edens) edens[i] = (T)ed;
        if (v2) v2[i] = (T)f2;
    }
}

template <typename T>
void eval_fields(Plan &P, const T *rho, unsigned calc, T *pot, double *energies, T *edens, T *v2, hipStream_t s) {
    OFDFT_REQUIRE(!(P.terms & OFDFT_TERM_LMGP) || edens == nullptr,
                  "the energy density of the LMGP functional is not implemented");
    // the full pipeline (inverse transforms in the work slots), the potential
    // into the caller's buffer or the plan's scratch field
    T *vout = pot ? pot : reinterpret_cast<T *>(P.aux);
    eval<T>(P, rho, calc | OFDFT_CALC_V, vout, energies, s);
    FieldParams fp;
    int ns = 0;
    fp.s_vw = (P.terms & OFDFT_TERM_VW) ? ns++ : -1;
    fp.s_wt = (P.terms & OFDFT_TERM_WT) ? ns++ : -1;
    if (P.terms & OFDFT_TERM_HARTREE) ns++;
    fp.s_wt2 = (fp.s_wt >= 0 && !wt_alpha_is_beta(P)) ? ns : -1;
    fp.terms = P.terms;
    fp.core = (P.terms & OFDFT_TERM_XC) ? P.core : nullptr;
    fp.x = P.x;
    fp.y = P.y;
    fp.alpha = P.alpha;
    fp.beta = P.beta;
    fp.n = P.N;
    fp.bs = P.N;
    OFDFT_REQUIRE(P.gate == nullptr, "gated launch of the field evaluation");
    hipLaunchKernelGGL((k_fields<T>), dim3(grid_blocks(P.N)), dim3(256), 0, s, rho,
                       reinterpret_cast<const T *>(P.work), edens, v2, fp);
    OFDFT_HIP(hipGetLastError());
}
template void eval_fields<float>(Plan &, const float *, unsigned, float *, double *, float *, float *, hipStream_t);
template void eval_fields<double>(Plan &, const double *, unsigned, double *, double *, double *, double *,
                                  hipStream_t);

// One stage of the E+V pipeline on the plan's workspace, for timing (the
// spectra and work fields hold whatever the last evaluation left): stage 0 =
// the x pass with the G-space multiply of every spectrum, in place; stage 1 =
// the point-wise post pass (potential into pot, energy partials).
template <typename T> bool bench_xfused(Plan &P, hipStream_t s) {
    if (!x_fusable(P)) return false;
    const long long pitch = pipe_pitch(P);
    const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
    GParams gp = pipeline_gparams(P, 0, 1, 2, false, nullptr, ncp, pitch);
    gspace_apply<T>(P, reinterpret_cast<cpx<T> *>(P.spec), 3, ncp, pitch, gp, true, s);
    return true;
}
template bool bench_xfused<float>(Plan &, hipStream_t);
template bool bench_xfused<double>(Plan &, hipStream_t);

template <typename T> void bench_eval_stage(Plan &P, int stage, const T *rho, T *pot, hipStream_t s) {
    const unsigned terms = P.terms;
    int ns = 0;
    const int s_vw = (terms & OFDFT_TERM_VW) ? ns++ : -1;
    const int s_wt = (terms & OFDFT_TERM_WT) ? ns++ : -1;
    const int s_h = (terms & OFDFT_TERM_HARTREE) ? ns++ : -1;
    const long long pitch = pipe_pitch(P);
    const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
    if (stage == 0) {
        OFDFT_REQUIRE(ns > 0, "bench: no spectra in the pipeline");
        GParams gp = pipeline_gparams(P, s_vw, s_wt, s_h, false, nullptr, ncp, pitch);
        gspace_apply<T>(P, reinterpret_cast<cpx<T> *>(P.spec), ns, ncp, pitch, gp, x_fusable(P), s);
    } else {
        PostParams pp = pipeline_post(P, s_vw, s_wt, s_h, true, true);
        const T *vloc = (terms & OFDFT_TERM_PP) ? reinterpret_cast<const T *>(P.vloc) : nullptr;
        const int nb = std::min(reduce_blocks(P.N), 1024);
        launch_post<T>(nb, s, rho, reinterpret_cast<const T *>(P.work), vloc, pot, P.red, pp, P.gate);
    }
    OFDFT_HIP(hipGetLastError());
}
template void bench_eval_stage<float>(Plan &, int, const float *, float *, hipStream_t);
template void bench_eval_stage<double>(Plan &, int, const double *, double *, hipStream_t);

// ------------------------------------------------------------------ linear response
// dV[drho] of the same functional at rho: the FFT terms are linear in their
// transformed inputs, so the spectra of d sqrt(rho) = drho / (2 sqrt(rho)),
// d rho^beta = beta rho^(beta-1) drho and drho go through the identical
// G-space multiply; the local terms use dV/drho point-wise.
template <typename T>
void eval_linear(Plan &P, const T *rho, const T *phi, const T *p, hipStream_t s) {
    const unsigned terms = P.terms;
    int ns = 0;
    const int s_vw = (terms & OFDFT_TERM_VW) ? ns++ : -1;
    const int s_wt = (terms & OFDFT_TERM_WT) ? ns++ : -1;
    const int s_h = (terms & OFDFT_TERM_HARTREE) ? ns++ : -1;
    if (ns == 0) return;
    cpx<T> *spec = reinterpret_cast<cpx<T> *>(P.spec);
    T *work = reinterpret_cast<T *>(P.work);
    Prologue pro{};
    if (s_vw >= 0) pro.code[s_vw] = OP_DSQRT;
    if (s_wt >= 0) pro.code[s_wt] = OP_DPOW_BETA;
    if (s_h >= 0) pro.code[s_h] = OP_IDENTITY;
    pro.beta = P.beta;
    pro.b56 = beta_is_5_6(P.beta);
    pro.aux = rho;
    pro.phi = phi;  // transforms of drho = 2 phi p
    const long long pitch = pipe_pitch(P);
    const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
    const bool fused = x_fusable(P);
    fft_r2c<T>(P, p, spec, ns, 0, ncp, 1.0, &pro, nullptr, s, pitch, fused);
    GParams gp;
    for (int i = 0; i < 9; ++i) gp.B[i] = P.B[i];
    gp.nx = P.nr[0];
    gp.ny = P.nr[1];
    gp.nz = P.nr[2];
    gp.nzc = P.nzc;
    gp.full = P.full;
    gp.scale = 1.0 / (double)P.N;
    gp.s_vw = s_vw;
    gp.s_wt = s_wt;
    gp.s_h = s_h;
    gp.sigma = P.sigma;
    gp.lbda = P.x;
    gp.alpha = P.alpha;
    gp.beta = P.beta;
    gp.rho0 = P.rho0 > 0 ? P.rho0 : 1.0;
    gp.rho0_dev = (s_wt >= 0 && !(P.rho0 > 0)) ? P.scal + 0 : nullptr;
    gp.wt_dev = gp.rho0_dev ? P.scal + kWtSlot : nullptr;
    gp.bs = ncp;
    gp.pitch = (int)pitch;
    gspace_apply<T>(P, spec, ns, ncp, pitch, gp, fused, s);
    fft_c2r<T>(P, spec, spec, work, ns, ncp, ncp, P.N, 1.0, s, pitch, fused);
}

// Ap = (V - mu) p + phi dV[2 phi p]  and  sum p.Ap.
// base_a = 1/2 L(sqrt(rho)) and base_c = C(rho^beta) of the evaluation at rho.
template <typename T, typename CT, int VEC>
__global__ void k_hvp(const T *__restrict__ rho, const T *__restrict__ phi,
                      const T *__restrict__ p, const T *__restrict__ pot, const T *__restrict__ work,
                      const T *__restrict__ base_a, const T *__restrict__ base_c, const double *__restrict__ mu,
                      T *__restrict__ Ap, double *partials, PostParams pp, const double *gate) {
    if (gate_off(gate)) return;
    const CT ctf = (CT)(0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0));
    const CT cbrt34pi = (CT)cbrt(3.0 / (4.0 * M_PI));
    const CT cx3 = (CT)(-cbrt(3.0 / M_PI) / 3.0);
    const double m = *mu;
    double acc[1] = {0.0};
    pw_groups<VEC>(pp.n, [&](long long i0) {
        const PwVec<VEC, T> phv = pw_ld<VEC>(phi, i0), pv = pw_ld<VEC>(p, i0), rv = pw_ld<VEC>(rho, i0),
                            potv = pw_ld<VEC>(pot, i0);
        PwVec<VEC, T> bav{}, avw{}, bcv{}, cwt{}, hv{}, out{};
        if (pp.s_vw >= 0) {
            bav = pw_ld<VEC>(base_a, i0);
            avw = pw_ld<VEC>(work + pp.s_vw * pp.bs, i0);
        }
        if (pp.s_wt >= 0) {
            bcv = pw_ld<VEC>(base_c, i0);
            cwt = pw_ld<VEC>(work + pp.s_wt * pp.bs, i0);
        }
        if (pp.s_h >= 0) hv = pw_ld<VEC>(work + pp.s_h * pp.bs, i0);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const T phi_i = phv.v[k], p_i = pv.v[k];
            const CT r = (CT)rv.v[k], dr = (CT)(T)(2.0 * (double)phi_i * (double)p_i);  // drho = 2 phi p
            const CT rp = r > (CT)1e-30 ? r : (CT)1e-30;
            const CT c_rp = fm::cbrt(rp);
            const CT ic = rcp_p<CT>(c_rp);  // rho^(-1/3)
            const CT irp = ic * ic * ic;    // 1 / rho
            CT dv = (CT)0;
            if (pp.terms & OFDFT_TERM_TF) dv += (CT)pp.x * (CT)(10.0 / 9.0) * ctf * ic * dr;  // tf.py:38-40
            if (pp.s_vw >= 0) {
                const CT isq = rsq_p<CT>((r > (CT)0 && r < (CT)1e-30) ? r : rp);  // 1 / sqrt(rho > 0 ? rho : 1e-30)
                const CT ds = (CT)0.5 * dr * isq;
                const CT vwb = (CT)pp.y * (CT)bav.v[k] * isq;
                dv += ((CT)pp.y * (CT)avw.v[k] - vwb * ds) * isq;
        }
        if (pp.s_wt >= 0) {
            const CT rb1 = pp.b56 ? c_rp * rsq_p<CT>(rp) : fm::pow(rp, (CT)pp.beta - (CT)1);  // rho^(beta-1)
            dv += (CT)(2.0 * pp.alpha) * (((CT)pp.beta - (CT)1) * rb1 * irp * dr * (CT)bcv.v[k] +
                                           rb1 * (CT)cwt.v[k]);
        }
        if (pp.terms & OFDFT_TERM_XC) {
            // semilocal_xc.py:291-303
            const CT rc3 = c_rp;
            const CT rs = cbrt34pi * ic;
            CT fc;
            if (rs < (CT)1) {
                fc = -(CT)(PZ_A / 3.0) -
                     ((CT)(PZ_C / 9.0) * (fm::log(rs) * (CT)2 + (CT)1) + (CT)(PZ_D * 2.0 / 9.0)) * rs;
            } else {
                const CT ta = (CT)PZ_B1 * fm::sqrt(rs), tb = (CT)PZ_B2 * rs;
                const CT ide = rcp_p<CT>((CT)1 + ta + tb);
                fc = (CT)(PZ_G / 36.0) *
                     ((CT)5 * ta + (CT)7 * ta * ta + (CT)8 * tb + (CT)16 * tb * tb + (CT)21 * ta * tb) *
                     (ide * ide * ide);
            }
            dv += (cx3 * rc3 + fc) * irp * dr;
        }
        if (pp.s_h >= 0) dv += (CT)hv.v[k];
        const double a = ((double)potv.v[k] - m) * (double)p_i + (double)phi_i * (double)dv;
        out.v[k] = (T)a;
        acc[0] += (double)p_i * (double)(T)a;
        }
        pw_st<VEC>(Ap, i0, out);
    });
    block_sums<1>(acc, partials);
}

template <typename T>
void hvp(Plan &P, const T *rho, const T *phi, const T *p, const T *pot, const T *base_a, const T *base_c,
         const double *mu, T *Ap, int nb, hipStream_t s) {
    OFDFT_REQUIRE(!(P.terms & OFDFT_TERM_LMGP), "the analytic Hessian product does not cover LMGP; use the "
                                                 "finite-difference one");
    OFDFT_REQUIRE(!(P.terms & OFDFT_TERM_WT) || wt_alpha_is_beta(P),
                  "the analytic Hessian product covers WT with alpha == beta only; use the finite-difference one");
    OFDFT_REQUIRE(!((P.terms & OFDFT_TERM_XC) && P.core),
                  "the analytic Hessian product does not cover a core density in XC; use the finite-difference one");
    eval_linear<T>(P, rho, phi, p, s);
    const unsigned terms = P.terms;
    int ns = 0;
    PostParams pp;
    pp.s_vw = (terms & OFDFT_TERM_VW) ? ns++ : -1;
    pp.s_wt = (terms & OFDFT_TERM_WT) ? ns++ : -1;
    pp.s_h = (terms & OFDFT_TERM_HARTREE) ? ns++ : -1;
    pp.terms = terms;
    pp.x = P.x;
    pp.y = P.y;
    pp.alpha = P.alpha;
    pp.beta = P.beta;
    pp.b56 = beta_is_5_6(P.beta);
    pp.n = P.N;
    pp.bs = P.N;
    pp.want_v = 1;
    pp.want_e = 0;
    if (pw_vec(pp.n) == 4 && pw_aligned({rho, phi, p, pot, P.work, base_a, base_c, Ap}))
        hipLaunchKernelGGL((k_hvp<T, T, 4>), dim3(nb), dim3(256), 0, s, rho, phi, p, pot,
                           reinterpret_cast<const T *>(P.work), base_a, base_c, mu, Ap, P.red, pp, P.gate);
    else
        hipLaunchKernelGGL((k_hvp<T, T, 1>), dim3(nb), dim3(256), 0, s, rho, phi, p, pot,
                           reinterpret_cast<const T *>(P.work), base_a, base_c, mu, Ap, P.red, pp, P.gate);
    OFDFT_HIP(hipGetLastError());
}

template void hvp<float>(Plan &, const float *, const float *, const float *, const float *, const float *,
                         const float *, const double *, float *, int, hipStream_t);
template void hvp<double>(Plan &, const double *, const double *, const double *, const double *, const double *,
                          const double *, const double *, double *, int, hipStream_t);

// slots of the FFT terms in P.work after an evaluation: (vW, WT, Hartree), -1 if absent
void work_slots(const Plan &P, int *slot) {
    int ns = 0;
    slot[0] = (P.terms & OFDFT_TERM_VW) ? ns++ : -1;
    slot[1] = (P.terms & OFDFT_TERM_WT) ? ns++ : -1;
    slot[2] = (P.terms & OFDFT_TERM_HARTREE) ? ns++ : -1;
}

}  // namespace ofdft
