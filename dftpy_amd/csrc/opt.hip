// Density optimizer (Optimization.optimize_rho, optimization.py:281-539) run
// natively: direct minimisation over phi = sqrt(rho) with norm-conserving
// rotations phi cos(theta) + p sin(theta), a truncated-Newton (or CG)
// direction and the Moré-Thuente line search on theta in [0, pi].
//
// Field updates are fused HIP kernels with FP64 partial sums; the scalar
// control flow (TN inner CG, line search, convergence test) runs on the host
// on FP64 scalars fetched once per step.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dcsrch.h"
#include "gspace.h"
#include "internal.h"
#include "reduce.h"
#include "stream.h"

namespace ofdft {

enum AuxField {
    F_PHI, F_RES, F_P, F_DIR, F_BEST, F_TRES, F_TP, F_PHI1, F_RHO1, F_POT1, F_AP, F_RES_PREV, F_DIR_PREV,
    F_POT, F_RHO, F_NPHI, F_NRHO, F_NPOT, F_BASE_A, F_BASE_C, F_Z, F_COUNT
};

void plan_aux_setup(Plan &P) {
    const size_t rs = P.prec == OFDFT_FP64 ? 8 : 4;
    P.aux_bytes = (size_t)F_COUNT * P.N * rs;
    OFDFT_HIP(hipMalloc(&P.aux, P.aux_bytes));
}

// the cached superstep graph of the device-driven TN iteration
struct TnGraphs {
    std::vector<char> key;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;  // capture stream
    int captures = 0, updates = 0;
};

void plan_aux_free(Plan &P) {
    if (P.aux) (void)hipFree(P.aux);
    P.aux = nullptr;
    if (P.tn_graphs) {
        if (P.tn_graphs->exec) (void)hipGraphExecDestroy(P.tn_graphs->exec);
        if (P.tn_graphs->cap) (void)hipStreamDestroy(P.tn_graphs->cap);
        delete P.tn_graphs;
        P.tn_graphs = nullptr;
    }
}

template <typename T> static T *fld(Plan &P, int f) { return reinterpret_cast<T *>(P.aux) + (size_t)f * P.N; }


template <typename T>
bool xfused_fixed(int n, int var, cpx<T> *data, int batch, long long bs, long long ncols, long long pitch,
                  long long vp, const void *tw, const GParams &gp, hipStream_t s, const double *gate);

// ------------------------------------------------------------------ kernels
// Streaming kernels: template VEC (4 consecutive elements per thread when the
// field length allows, else 1), see stream.h.  The reducing ones write one
// row of FP64 partial sums per workgroup.
#define PW_BODY(i) [&](long long i)

// phi = rho > 0 ? sqrt(rho) : 1e-300 ; sums rho and phi^2 (optimization.py:305-315)
template <typename T, int VEC> __global__ void k_init_phi(const double *gate,
                                                          const T *__restrict__ rho, T *__restrict__ phi,
                                                          long long n, double *partials) {
    if (gate_off(gate)) return;
    double acc[2] = {0, 0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const T r = rho[i];
        const T f = r > (T)0 ? sqrt(r) : (T)1e-300;
        phi[i] = f;
        acc[0] += (double)r;
        acc[1] += (double)f * (double)f;
    });
    block_sums<2>(acc, partials);
}

// phi *= sqrt(s[0] / s[1]) ; rho = phi^2 (the normalisation factor from
// device sums, optimization.py:305-316)
template <typename T, int VEC> __global__ void k_scale_square(const double *gate,
                                                              T *__restrict__ phi, T *__restrict__ rho,
                                                              const double *s, long long n) {
    if (gate_off(gate)) return;
    const double f = sqrt(s[0] / s[1]);
    pw_loop<VEC>(n, PW_BODY(i) {
        const T v = (T)((double)phi[i] * f);
        phi[i] = v;
        rho[i] = v * v;
    });
}

// sums  V*rho  and  rho  (mu = int V rho / N, optimization.py:537-539)
template <typename T, int VEC> __global__ void k_mu_sums(const double *gate,
                                                         const T *__restrict__ pot, const T *__restrict__ rho,
                                                         long long n, double *partials) {
    if (gate_off(gate)) return;
    double acc[2] = {0, 0};
    pw_loop<VEC>(n, PW_BODY(i) {
        acc[0] += (double)pot[i] * (double)rho[i];
        acc[1] += (double)rho[i];
    });
    block_sums<2>(acc, partials);
}

// scal[dst] = scal[a] / scal[b] (a device-side ratio, so the next kernel can
// read it without a host round trip)
__global__ void k_ratio(const double *gate, double *scal, int dst, int a, int b) {
    if (gate_off(gate)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) scal[dst] = scal[a] / scal[b];
}

// res = (V - mu) phi ; sum res^2
template <typename T, int VEC>
__global__ void k_residual(const double *gate,
                           const T *__restrict__ pot, const T *__restrict__ phi, T *__restrict__ res,
                           const double *mu, long long n, double *partials) {
    if (gate_off(gate)) return;
    const double m = *mu;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const T r = (T)(((double)pot[i] - m) * (double)phi[i]);
        res[i] = r;
        acc[0] += (double)r * (double)r;
    });
    block_sums<1>(acc, partials);
}

// TN start: tres = -res, tp = tres, dir = 0 ; sum tres^2
template <typename T, int VEC>
__global__ void k_tn_start(const double *gate,
                           const T *__restrict__ res, T *__restrict__ tres, T *__restrict__ tp, T *__restrict__ dir,
                           long long n, double *partials) {
    if (gate_off(gate)) return;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const T r = -res[i];
        tres[i] = r;
        tp[i] = r;
        dir[i] = (T)0;
        acc[0] += (double)r * (double)r;
    });
    block_sums<1>(acc, partials);
}

// phi1 = phi + eps p ; rho1 = phi1^2
template <typename T, int VEC>
__global__ void k_tn_shift(const double *gate,
                           const T *__restrict__ phi, const T *__restrict__ p, double eps, T *__restrict__ phi1,
                           T *__restrict__ rho1, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) {
        const T v = (T)((double)phi[i] + eps * (double)p[i]);
        phi1[i] = v;
        rho1[i] = v * v;
    });
}

// Ap = ((V1 - mu) phi1 - res0) / eps ; sum p.Ap   (optimization.py:105-114)
template <typename T, int VEC>
__global__ void k_tn_ap(const double *gate,
                        const T *__restrict__ pot1, const T *__restrict__ phi1, const T *__restrict__ res0,
                        const T *__restrict__ p, const double *mu, double eps, T *__restrict__ Ap,
                        long long n, double *partials) {
    if (gate_off(gate)) return;
    const double m = *mu;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const double a = (((double)pot1[i] - m) * (double)phi1[i] - (double)res0[i]) / eps;
        Ap[i] = (T)a;
        acc[0] += (double)p[i] * (double)(T)a;
    });
    block_sums<1>(acc, partials);
}

// dir += alpha p ; tres -= alpha Ap ; sum tres^2, with alpha = r0 / pAp taken
// from the device (pAp reduced by the previous launch), so the inner CG needs one host round trip per iteration;
// negative curvature (pAp < 0) leaves dir and tres untouched for the host.
template <typename T, int VEC>
__global__ void k_tn_update_dev(const double *gate,
                                T *__restrict__ dir, T *__restrict__ tres, const T *__restrict__ p,
                                const T *__restrict__ Ap, const double *rzp, const double *pap,
                                long long n, double *partials) {
    if (gate_off(gate)) return;
    const double pa = *pap, r0 = *rzp;
    double acc[1] = {0};
    if (!(pa < 0.0)) {
        const double alpha = r0 / pa;
        pw_loop<VEC>(n, PW_BODY(i) {
            dir[i] = (T)((double)dir[i] + alpha * (double)p[i]);
            const T r = (T)((double)tres[i] - alpha * (double)Ap[i]);
            tres[i] = r;
            acc[0] += (double)r * (double)r;
        });
    }
    block_sums<1>(acc, partials);
}

// y = a x + b y
template <typename T, int VEC> __global__ void k_axpby(const double *gate,
                                                       double a, const T *__restrict__ x, double b,
                                                       T *__restrict__ y, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) { y[i] = (T)(a * (double)x[i] + b * (double)y[i]); });
}

template <typename T, int VEC> __global__ void k_copy(const double *gate,
                                                      const T *__restrict__ x, T *__restrict__ y, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) { y[i] = x[i]; });
}

template <typename T, int VEC> __global__ void k_fill(const double *gate, T *__restrict__ y, T v, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) { y[i] = v; });
}

// sums p*phi and phi*phi
template <typename T, int VEC> __global__ void k_pphi(const double *gate,
                                                      const T *__restrict__ p, const T *__restrict__ phi,
                                                      long long n, double *partials) {
    if (gate_off(gate)) return;
    double acc[2] = {0, 0};
    pw_loop<VEC>(n, PW_BODY(i) {
        acc[0] += (double)p[i] * (double)phi[i];
        acc[1] += (double)phi[i] * (double)phi[i];
    });
    block_sums<2>(acc, partials);
}

// p -= f phi ; sum p^2
template <typename T, int VEC>
__global__ void k_orth(const double *gate,
                       T *__restrict__ p, const T *__restrict__ phi, const double *f, long long n, double *partials) {
    if (gate_off(gate)) return;
    const double ff = *f;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const T v = (T)((double)p[i] - ff * (double)phi[i]);
        p[i] = v;
        acc[0] += (double)v * (double)v;
    });
    block_sums<1>(acc, partials);
}

// sum V phi p
template <typename T, int VEC>
__global__ void k_dot3(const double *gate,
                       const T *__restrict__ a, const T *__restrict__ b, const T *__restrict__ c, long long n,
                       double *partials) {
    if (gate_off(gate)) return;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) { acc[0] += (double)a[i] * (double)b[i] * (double)c[i]; });
    block_sums<1>(acc, partials);
}

// newphi = phi cos + p sin ; newrho = newphi^2   (optimization.py:236-238)
template <typename T, int VEC>
__global__ void k_rotate(const double *gate, const T *__restrict__ phi, const T *__restrict__ p, double c, double s,
                         const double *cs, T *__restrict__ nphi, T *__restrict__ nrho, long long n) {
    if (gate_off(gate)) return;
    if (cs) {  // (cos, sin) from the device
        c = cs[0];
        s = cs[1];
    }
    pw_loop<VEC>(n, PW_BODY(i) {
        const T v = (T)((double)phi[i] * c + (double)p[i] * s);
        nphi[i] = v;
        nrho[i] = v * v;
    });
}

// sum V * newphi * (p cos - phi sin)   (optimization.py:270-276)
template <typename T, int VEC>
__global__ void k_theta_grad(const double *gate, const T *__restrict__ pot, const T *__restrict__ nphi,
                             const T *__restrict__ p, const T *__restrict__ phi, double c, double s, const double *cs,
                             long long n, double *partials) {
    if (gate_off(gate)) return;
    if (cs) {
        c = cs[0];
        s = cs[1];
    }
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const T p2 = (T)((double)p[i] * c - (double)phi[i] * s);
        acc[0] += (double)pot[i] * (double)nphi[i] * (double)p2;
    });
    block_sums<1>(acc, partials);
}

// p = -res (steepest descent)
template <typename T, int VEC> __global__ void k_neg(const double *gate,
                                                     const T *__restrict__ x, T *__restrict__ y, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) { y[i] = -x[i]; });
}

// CG directions (math_utils.py:395-423): sums r1(r1 - r0), d(r1 - r0), r1^2,
// r0^2, d r0
template <typename T, int VEC>
__global__ void k_cg_sums(const double *gate,
                          const T *__restrict__ r1, const T *__restrict__ r0, const T *__restrict__ d, long long n,
                          double *partials) {
    if (gate_off(gate)) return;
    double acc[5] = {0, 0, 0, 0, 0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const double a = r1[i], b = r0[i], dd = d[i];
        acc[0] += a * (a - b);
        acc[1] += dd * (a - b);
        acc[2] += a * a;
        acc[3] += b * b;
        acc[4] += dd * b;
    });
    block_sums<5>(acc, partials);
}

// CG-CD as math_utils.py:411-412 writes it: beta = -(r1 ** 2) / sum(d r0) is a
// field (the numerator is not summed), p = -r1 + beta d
template <typename T, int VEC> __global__ void k_cg_cd(const double *gate,
                                                       const T *__restrict__ r1, const T *__restrict__ d,
                                                       double sdr0, T *__restrict__ p, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) {
        const double r = r1[i];
        const double beta = -(r * r) / sdr0;
        p[i] = (T)(-r + beta * (double)d[i]);
    });
}

// L-BFGS (math_utils.py:374-455).  A pair: s = phi - phi_old,
// y = res - res_old, and the partial sums of y s
template <typename T, int VEC>
__global__ void k_lbfgs_pair(const double *gate,
                             const T *__restrict__ phi, const T *__restrict__ phi_old, const T *__restrict__ res,
                             const T *__restrict__ res_old, T *__restrict__ s, T *__restrict__ y, long long n,
                             double *partials) {
    if (gate_off(gate)) return;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const T a = phi[i] - phi_old[i], b = res[i] - res_old[i];
        s[i] = a;
        y[i] = b;
        acc[0] += (double)b * (double)a;
    });
    block_sums<1>(acc, partials);
}

// scal[k] = 1 / scal[k]
__global__ void k_recip(const double *gate, double *scal, int k) {
    if (gate_off(gate)) return;
    if (threadIdx.x == 0) scal[k] = 1.0 / scal[k];
}

// first loop: alpha = rho_i sum(s_i q) (scal[rho] * scal[dot]), q -= alpha y_i;
// alpha is kept at scal[alpha_out] for the second loop
template <typename T, int VEC>
__global__ void k_lbfgs_q(const double *gate,
                          T *__restrict__ q, const T *__restrict__ y, double *scal, int rho, int dot, int alpha_out,
                          long long n) {
    if (gate_off(gate)) return;
    const double alpha = scal[rho] * scal[dot];
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[alpha_out] = alpha;
    pw_loop<VEC>(n, PW_BODY(i) { q[i] = (T)((double)q[i] - alpha * (double)y[i]); });
}

// second loop: beta = rho_i sum(y_i d), d += s_i (alpha_i - beta); alpha_in < 0
// stands for alpha = 0 (the oldest pair, which the first loop does not visit)
template <typename T, int VEC>
__global__ void k_lbfgs_d(const double *gate,
                          T *__restrict__ d, const T *__restrict__ s, const double *scal, int rho, int dot,
                          int alpha_in, long long n) {
    if (gate_off(gate)) return;
    const double beta = scal[rho] * scal[dot];
    const double a = (alpha_in < 0 ? 0.0 : scal[alpha_in]) - beta;
    pw_loop<VEC>(n, PW_BODY(i) { d[i] = (T)((double)d[i] + (double)s[i] * a); });
}

// sums s y and y y (gamma = s.y / y.y of the newest pair)
template <typename T, int VEC> __global__ void k_sy_yy(const double *gate,
                                                       const T *__restrict__ s, const T *__restrict__ y, long long n,
                                                       double *partials) {
    if (gate_off(gate)) return;
    double acc[2] = {0, 0};
    pw_loop<VEC>(n, PW_BODY(i) {
        const double a = s[i], b = y[i];
        acc[0] += a * b;
        acc[1] += b * b;
    });
    block_sums<2>(acc, partials);
}

// x *= scal[a] / scal[b]
template <typename T, int VEC> __global__ void k_scale_ratio(const double *gate,
                                                             T *__restrict__ x, const double *scal, int a, int b,
                                                             long long n) {
    if (gate_off(gate)) return;
    const double f = scal[a] / scal[b];
    pw_loop<VEC>(n, PW_BODY(i) { x[i] = (T)(f * (double)x[i]); });
}

// x *= f
template <typename T, int VEC> __global__ void k_scale(const double *gate, T *__restrict__ x, double f, long long n) {
    if (gate_off(gate)) return;
    pw_loop<VEC>(n, PW_BODY(i) { x[i] = (T)(f * (double)x[i]); });
}

// p *= sqrt(scal[a] / scal[b])  (normalise p to sum p^2 = scal[a] with scal[b] = sum p^2)
template <typename T, int VEC> __global__ void k_scale_norm(const double *gate,
                                                            T *__restrict__ p, const double *scal, int a, int b,
                                                            long long n) {
    if (gate_off(gate)) return;
    const double f = sqrt(scal[a] / scal[b]);
    pw_loop<VEC>(n, PW_BODY(i) { p[i] = (T)(f * (double)p[i]); });
}

// sum a.b
template <typename T, int VEC> __global__ void k_dot2(const double *gate,
                                                      const T *__restrict__ a, const T *__restrict__ b, long long n,
                                                      double *partials) {
    if (gate_off(gate)) return;
    double acc[1] = {0};
    pw_loop<VEC>(n, PW_BODY(i) { acc[0] += (double)a[i] * (double)b[i]; });
    block_sums<1>(acc, partials);
}

#undef PW_BODY

struct PrecondParams {
    double B[9];
    const double *tab;
    int n;
    double inv_dq, scale;
};

// half spectrum (pitched rows, one workgroup per (ix, iy) row) times
// M^-1(|G|), linearly interpolated in the radial table, and the 1/N of the
// inverse transform
template <typename T>
__global__ void k_precond(const double *gate, cpx<T> *spec, long long pitch, int nzc, int nx, int ny,
                          PrecondParams pp) {
    if (gate_off(gate)) return;
    const int row = blockIdx.x;
    const int ix = row / ny, iy = row - ix * ny;
    const int fx = ix <= (nx - 1) / 2 ? ix : ix - nx;
    const int fy = iy <= (ny - 1) / 2 ? iy : iy - ny;
    const double cx = fx * pp.B[0] + fy * pp.B[3], cy = fx * pp.B[1] + fy * pp.B[4], cz = fx * pp.B[2] + fy * pp.B[5];
    cpx<T> *r = spec + (long long)row * pitch;
    for (int iz = threadIdx.x; iz < nzc; iz += blockDim.x) {
        const double gx = cx + iz * pp.B[6], gy = cy + iz * pp.B[7], gz = cz + iz * pp.B[8];
        const double t = sqrt(gx * gx + gy * gy + gz * gz) * pp.inv_dq;
        double m;
        if (t >= (double)(pp.n - 1)) {
            m = pp.tab[pp.n - 1];
        } else {
            const int i = (int)t;
            const double f = t - i;
            m = pp.tab[i] * (1.0 - f) + pp.tab[i + 1] * f;
        }
        const T c = (T)(m * pp.scale);
        cpx<T> v = r[iz];
        v.x *= c;
        v.y *= c;
        r[iz] = v;
    }
}


// ------------------------------------------------------------------ device-driven TN
// The TN iteration of Opt::run with every scalar decision taken on the
// device: the inner-CG stopping tests (optimization.py:113-145), the
// orthonormalisation's uphill restart, the line search (dcsrch_core.h, the
// same arithmetic as the host loop) and the convergence test
// (optimization.py:480-508).  The host queues a fixed superstep of kernels
// (CG start, G inner-CG iterations, orthonormalisation, L line-search
// evaluations, accept), each section behind a gate that the control kernels
// open and close, so one superstep is one hipGraph launch and the host reads
// back a few scalars once per superstep instead of once per step.
//
// State: doubles at P.scal + DT.  Gates hold 0 (skip) or 1 (run).
enum DtSlot : int {
    DT = 256,
    G_CGSTART = DT, G_CG, G_ORTH, G_UPHILL, G_LS, G_ACCEPT, G_END,
    // the update after inner-CG iteration k of a superstep runs when G_AFTER
    // holds k + 1; T_AFTER holds its AfterOp bits
    G_AFTER = DT + 8, T_AFTER,
    // inner CG
    T_CGIT = DT + 16, T_R0, T_RZ, T_RCONV, T_RMIN, T_BESTZERO, T_RZCUR, T_NEG0_A, T_NEG0_B, T_PUPD_A, T_PUPD_B,
    // outer iteration (T_IT .. T_NEV are read back after every superstep)
    T_IT = DT + 32, T_STATUS, T_ENERGY, T_NV, T_NEV, T_THETA, T_STP, T_C, T_S, T_F, T_LSTASK, T_LSCALLS,
    T_LSSINGLE, T_UPDONE,
    // options
    T_ECONV = DT + 56, T_C1, T_C2, T_XTOL, T_DV, T_MAXITER, T_MAXFUN, T_MAXLS, T_NCHECK, T_FP32, T_PRE,
    // line-search state (15 doubles)
    T_LS = DT + 80,
    // energies of the accepted iterations (ring of DT_HCAP)
    T_HIST = DT + 128
};
constexpr int DT_HCAP = kScalSlots - T_HIST;
constexpr int DT_READ = T_NEV - T_IT + 1;

struct DtOpts {
    double econv, c1, c2, xtol, dv;
    int maxiter, maxfun, maxls, ncheck, fp32, pre;
};

enum DtStage { DS_CGSTART, DS_CGITER, DS_ORTH, DS_LS, DS_ACCEPT };

// the updates that end an inner-CG iteration (optimization.py:114-145), in
// this order: dir = (r.z / pAp) p at a negative curvature on the first
// iteration; best = dir; dir = 0 or dir = best on a stall; p = z + beta p
enum AfterOp { AF_NEG0 = 1, AF_BEST = 2, AF_FBZERO = 4, AF_FBBEST = 8, AF_PUPD = 16 };

__device__ static void dt_finish(double *S, int status) {
    S[T_STATUS] = status;
    for (int g = G_CGSTART; g < G_END; ++g) S[g] = 0.0;
}

__device__ static void dt_trial(double *S) {
    S[T_C] = cos(S[T_STP]);
    S[T_S] = sin(S[T_STP]);
}

// the line search's first trial of the host-driven FP32 TN, queued before
// the orthonormalisation's results are read: theta = min(theta0, previous
// step) with theta0 = sqrt(scal[23] / scal[22]) as the host computes it, and
// its cosine and sine for the rotation
__global__ void k_trial_theta(const double *gate, double *S, double theta_prev) {
    if (gate_off(gate)) return;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double theta0 = sqrt(S[23] / S[22]);
    S[T_STP] = theta0 < theta_prev ? theta0 : theta_prev;
    dt_trial(S);
}

// after the starting evaluation and residual (run eagerly, not captured)
__global__ void k_tn_init(double *S, DtOpts o) {
    if (threadIdx.x != 0) return;
    S[T_ECONV] = o.econv;
    S[T_C1] = o.c1;
    S[T_C2] = o.c2;
    S[T_XTOL] = o.xtol;
    S[T_DV] = o.dv;
    S[T_MAXITER] = o.maxiter;
    S[T_MAXFUN] = o.maxfun;
    S[T_MAXLS] = o.maxls;
    S[T_NCHECK] = o.ncheck;
    S[T_FP32] = o.fp32;
    S[T_PRE] = o.pre;
    S[T_IT] = 0.0;
    S[T_STATUS] = -1.0;
    S[T_ENERGY] = S[32 + OFDFT_E_SUM];
    S[T_HIST] = S[T_ENERGY];
    S[T_THETA] = 0.1;
    S[T_NV] = 0.0;
    S[T_NEV] = 0.0;
    for (int g = G_CGSTART; g < G_END; ++g) S[g] = 0.0;
    S[G_AFTER] = 0.0;
    if (o.maxiter > 1)
        S[G_CGSTART] = 1.0;
    else
        S[T_STATUS] = 2.0;
}

// one control step; k: the inner-CG iteration's index within the superstep
__device__ static void dt_ctrl(double *S, int stage, int k) {
    const bool pre = S[T_PRE] != 0.0;
    switch (stage) {
    case DS_CGSTART: {  // |r|^2 at scal[25], r.z at scal[26]
        const double r0 = S[25];
        S[T_CGIT] = 0.0;
        S[T_R0] = r0;
        S[T_RZ] = S[T_RZCUR] = pre ? S[26] : r0;
        S[T_RCONV] = r0 * 0.1;
        S[T_RMIN] = 1e10;
        S[T_BESTZERO] = 1.0;
        S[G_CGSTART] = 0.0;
        S[G_AFTER] = 0.0;
        if (S[T_MAXFUN] > 0) {
            S[G_CG] = 1.0;
        } else {
            S[G_ORTH] = 1.0;
            S[T_UPDONE] = 0.0;
        }
        break;
    }
    case DS_CGITER: {  // pAp at scal[16], r1 at scal[17], r1.z1 at scal[18]
        const int it = (int)S[T_CGIT];
        S[T_NV] += 1.0;
        const double pAp = S[16], r1 = S[17], rz1 = pre ? S[18] : S[17];
        bool done = false;
        int ops = 0;
        if (pAp < 0.0) {
            if (it == 0) {
                S[T_NEG0_A] = S[T_RZ] / pAp;
                S[T_NEG0_B] = 0.0;
                ops |= AF_NEG0;
            }
            done = true;
        } else {
            const double rmin = S[T_RMIN];  // min over the residuals so far
            if (r1 < rmin) {
                ops |= AF_BEST;
                S[T_BESTZERO] = 0.0;
                S[T_RMIN] = r1;
            }
            const double r0 = S[T_R0];
            if (r1 < S[T_RCONV]) {
                done = true;
            } else if (r1 > 1000.0 * rmin || (it > 10 && fabs(r0 - r1) < 0.1 * r0)) {
                ops |= S[T_BESTZERO] != 0.0 ? AF_FBZERO : AF_FBBEST;
                done = true;
            } else {
                S[T_PUPD_A] = 1.0;
                S[T_PUPD_B] = rz1 / S[T_RZ];
                S[T_R0] = r1;
                S[T_RZ] = S[T_RZCUR] = rz1;
                ops |= AF_PUPD;
            }
        }
        S[T_AFTER] = ops;
        S[G_AFTER] = ops ? k + 1 : 0;
        S[T_CGIT] = it + 1;
        if (!done && it + 1 >= (int)S[T_MAXFUN]) done = true;
        if (done) {
            S[G_CG] = 0.0;
            S[G_ORTH] = 1.0;
            S[T_UPDONE] = 0.0;
        }
        break;
    }
    case DS_ORTH: {  // sum phi^2 at scal[22], sum p^2 at scal[23], sum V phi p at scal[24]
        const double theta0 = sqrt(S[23] / S[22]), vphip = S[24];
        if (vphip > 0.0 && S[T_UPDONE] == 0.0) {  // uphill: once more from -res
            S[G_UPHILL] = 1.0;
            S[T_UPDONE] = 1.0;
            break;
        }
        S[G_UPHILL] = 0.0;
        S[G_ORTH] = 0.0;
        S[G_AFTER] = 0.0;
        const double theta = ls::dmin(theta0, S[T_THETA]);
        S[T_THETA] = theta;
        const double f = S[T_ENERGY], g = 2.0 * vphip * S[T_DV];
        S[T_STP] = theta;
        S[T_LSCALLS] = 0.0;
        if (S[T_FP32] != 0.0 && fabs(g) * theta < 0.1 * S[T_ECONV]) {
            S[T_LSSINGLE] = 1.0;  // one evaluation at the trial step (the host loop's FP32 rule)
            dt_trial(S);
            S[G_LS] = 1.0;
            break;
        }
        S[T_LSSINGLE] = 0.0;
        if (S[T_MAXLS] <= 0) {
            dt_finish(S, 1);
            break;
        }
        double stp = theta;
        const LsTask t = ls::dcsrch_step(stp, f, g, S[T_C1], S[T_C2], S[T_XTOL], 0.0, M_PI, LS_START, S + T_LS);
        S[T_LSCALLS] = 1.0;
        S[T_STP] = stp;
        S[T_LSTASK] = (double)t;
        if (t == LS_FG) {
            dt_trial(S);
            S[G_LS] = 1.0;
        } else {
            dt_finish(S, 1);
        }
        break;
    }
    case DS_LS: {  // energies at scal[32..], sum V nphi p' at scal[32 + E_COUNT]
        S[T_NEV] += 1.0;
        const double f = S[32 + OFDFT_E_SUM], g = 2.0 * S[32 + OFDFT_E_COUNT] * S[T_DV];
        if (S[T_LSSINGLE] != 0.0) {
            S[T_F] = f;
            S[G_LS] = 0.0;
            S[G_ACCEPT] = 1.0;
            break;
        }
        if (S[T_LSCALLS] >= S[T_MAXLS]) {  // maxls calls, the last one asked for this evaluation
            dt_finish(S, 1);
            break;
        }
        double stp = S[T_STP];
        const LsTask t = ls::dcsrch_step(stp, f, g, S[T_C1], S[T_C2], S[T_XTOL], 0.0, M_PI, (LsTask)(int)S[T_LSTASK],
                                         S + T_LS);
        S[T_LSCALLS] += 1.0;
        S[T_STP] = stp;
        S[T_LSTASK] = (double)t;
        if (t == LS_FG) {
            dt_trial(S);
        } else if (t == LS_CONVERGENCE) {
            S[T_F] = f;
            S[G_LS] = 0.0;
            S[G_ACCEPT] = 1.0;
        } else {
            dt_finish(S, 1);
        }
        break;
    }
    case DS_ACCEPT: {
        S[T_THETA] = S[T_STP];
        const double e = S[T_F];
        S[T_ENERGY] = e;
        const int it = (int)S[T_IT] + 1;
        S[T_IT] = it;
        S[T_HIST + it % DT_HCAP] = e;
        S[G_ACCEPT] = 0.0;
        const int ncheck = (int)S[T_NCHECK];
        if (it >= ncheck) {
            bool conv = true;
            for (int i = 0; i < ncheck; ++i)
                if (fabs(S[T_HIST + (it - 1 - i) % DT_HCAP] - e) > S[T_ECONV]) conv = false;
            if (conv) {
                dt_finish(S, 0);
                break;
            }
        }
        if (it + 1 >= (int)S[T_MAXITER]) {
            dt_finish(S, 2);
            break;
        }
        S[G_CGSTART] = 1.0;
        break;
    }
    default: break;
    }
}

// out[t] = sum of the partial column t (k_reduce_final with unit
// coefficients, the same order of additions), then a control step
__global__ void __launch_bounds__(kFinalThreads)
    k_red_ctrl(const double *gate, const double *__restrict__ partials, int nb, int nterms, double *out, double *S,
               int stage, int k) {
    if (gate_off(gate)) return;
    __shared__ double sm[kFinalThreads];
    sum_columns(partials, nb, nterms, sm);
    if (threadIdx.x == 0) {
        for (int t = 0; t < nterms; ++t) out[t] = sm[t] * 1.0;
        dt_ctrl(S, stage, k);
    }
}

// the updates after inner-CG iteration k (AfterOp bits of T_AFTER), each
// element as the separate kernels of the host-driven loop compute it
template <typename T, int VEC>
__global__ void k_cg_after(const double *S, int k, T *__restrict__ dir, T *__restrict__ best, T *__restrict__ tp,
                           const T *__restrict__ x, long long n) {
    if (S[G_AFTER] != (double)(k + 1)) return;
    const int ops = (int)S[T_AFTER];
    const double a0 = S[T_NEG0_A], b0 = S[T_NEG0_B], a1 = S[T_PUPD_A], b1 = S[T_PUPD_B];
    pw_loop<VEC>(n, [&](long long i) {
        T d = dir[i];
        if (ops & AF_NEG0) d = (T)(a0 * (double)tp[i] + b0 * (double)d);
        if (ops & AF_BEST) best[i] = d;
        if (ops & AF_FBZERO) d = (T)0;
        if (ops & AF_FBBEST) d = best[i];
        if (ops & (AF_NEG0 | AF_FBZERO | AF_FBBEST)) dir[i] = d;
        if (ops & AF_PUPD) tp[i] = (T)(a1 * (double)x[i] + b1 * (double)tp[i]);
    });
}

// ------------------------------------------------------------------ host helpers
struct Ctx {
    Plan &P;
    hipStream_t s;
    int nb;
    double *host;  // pinned
    int vec;  // element width of the streaming kernels (stream.h)
    int nbw;  // workgroups of the non-reducing streaming kernels
    Ctx(Plan &p, hipStream_t st) : P(p), s(st) {
        nb = std::min(reduce_blocks(P.N), 1024);
        vec = pw_vec(P.N);
        nbw = pw_blocks(P.N);
        host = P.pinned;
    }

    dim3 grid() const { return dim3(nb); }
    // reduce K partial columns into scal[slot..slot+K) and copy them to the host
    void wait() const { stream_wait(s); }
    const double *fetch(int K, int slot = 16) {
        reduce_final_cols(K, slot);
        OFDFT_HIP(hipMemcpyAsync(host, P.scal + slot, K * sizeof(double), hipMemcpyDeviceToHost, s));
        wait();
        return host;
    }
    // pub_n > 0: the same launch then publishes scal[pub_slot..pub_slot+pub_n)
    // to the pinned host buffer, read by published()
    void reduce_final_cols(int K, int slot, int pub_n = 0, int pub_slot = 0, int pub_off = 0) {
        // partials laid out [block][K]
        std::vector<double> ones(K, 1.0);
        reduce_final(P.red, nb, K, P.scal + slot, ones.data(), s, P.gate, pub_n ? P.scal + pub_slot : nullptr,
                     pub_n ? P.pinned_dev + pub_off : nullptr, pub_n);
    }
    // the scalars the last publishing reduction sent, once the stream is done
    const double *published() const {
        wait();
        return host;
    }
    // copy scal[slot..slot+K) to the host
    const double *to_host(int K, int slot) {
        OFDFT_HIP(hipMemcpyAsync(host, P.scal + slot, K * sizeof(double), hipMemcpyDeviceToHost, s));
        wait();
        return host;
    }
};

// Launch the streaming kernel template K<T, VEC> with the context's element
// width: the reducing kernels on the C.nb partial-sum workgroups, the others
// on C.nbw workgroups.
#define PW_LAUNCH_(K, NB, ...)                                                                \
    do {                                                                                      \
        if (C.vec == 4)                                                                       \
            hipLaunchKernelGGL((K<T, 4>), dim3(NB), dim3(256), 0, C.s, P.gate, __VA_ARGS__);  \
        else                                                                                  \
            hipLaunchKernelGGL((K<T, 1>), dim3(NB), dim3(256), 0, C.s, P.gate, __VA_ARGS__);  \
    } while (0)
#define PW_RED(K, ...) PW_LAUNCH_(K, C.nb, __VA_ARGS__)
#define PW_STREAM(K, ...) PW_LAUNCH_(K, C.nbw, __VA_ARGS__)

template <typename T> struct Opt {
    Plan &P;
    Ctx C;
    const long long n;
    int n_ev = 0, n_v = 0;
    double energies_h[OFDFT_E_COUNT];
    double *energies_d;
    int map[F_COUNT];
    Opt(Plan &p, hipStream_t s) : P(p), C(p, s), n(p.N) {
        energies_d = P.scal + 32;
        for (int i = 0; i < F_COUNT; ++i) map[i] = i;
    }

    // fields are addressed through a permutation so accepting a trial point
    // swaps pointers instead of copying
    T *F(int f) { return fld<T>(P, map[f]); }

    void eval_v(const T *rho, T *pot) {
        eval<T>(P, rho, OFDFT_CALC_V, pot, nullptr, C.s);
        ++n_v;
    }
    // mu = sum(V rho) / sum(rho) (optimization.py:537-539) kept on device at
    // scal[2] (sums at scal[18], scal[19]), then res = (V - mu) phi.  No host
    // round trip: the host needs mu only for the final report.
    void mu_and_residual(const T *pot, const T *rho, const T *phi, T *res) {
        PW_RED(k_mu_sums, pot, rho, n, P.red);
        C.reduce_final_cols(2, 18);
        hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, C.s, P.gate, P.scal, 2, 18, 19);
        PW_RED(k_residual, pot, phi, res, P.scal + 2, n, P.red);
        C.reduce_final_cols(1, 20);
    }

    // out = IFFT[M^-1(|G|) FFT[in]] with the plan's radial preconditioner table;
    // for size-specialised x lengths the multiply runs inside the x pass
    // (forward x, multiply, inverse x in one kernel), as in the evaluation
    void apply_precond(const T *in, T *out) {
        const long long pitch = pipe_pitch(P);
        const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
        cpx<T> *spec = reinterpret_cast<cpx<T> *>(P.spec);
        const bool fused = P.use_fixed && P.nr[0] > 1 && P.nr[1] * P.nr[2] > 1 && fixed_size(P.nr[0]);
        if (fused) {
            GParams gp{};
            for (int i = 0; i < 9; ++i) gp.B[i] = P.B[i];
            gp.kind[0] = 3;
            gp.kind[1] = gp.kind[2] = gp.kind[3] = -1;
            gp.nx = P.nr[0];
            gp.ny = P.nr[1];
            gp.nz = P.nr[2];
            gp.nzc = P.nzc;
            gp.pitch = (int)pitch;
            gp.full = P.full;
            gp.scale = 1.0 / (double)P.N;
            gp.s_vw = gp.s_wt = gp.s_h = -1;
            gp.bs = ncp;
            gp.tab = P.tn_pre;
            gp.tab_n = P.tn_pre_n;
            gp.tab_inv_dq = 1.0 / P.tn_pre_dq;
            fft_r2c<T>(P, in, spec, 1, 0, ncp, 1.0, nullptr, nullptr, C.s, pitch, true);
            OFDFT_REQUIRE(xfused_fixed<T>(P.nr[0], P.var_cols[0], spec, 1, ncp, (long long)P.nr[1] * pitch, pitch,
                                          P.nzc, P.tw[0], gp, C.s, P.gate),
                          "fused x pass unavailable");
            OFDFT_HIP(hipGetLastError());
            fft_c2r<T>(P, spec, spec, out, 1, ncp, ncp, P.N, 1.0, C.s, pitch, true);
            return;
        }
        fft_r2c<T>(P, in, spec, 1, 0, ncp, 1.0, nullptr, nullptr, C.s, pitch);
        PrecondParams pp;
        for (int i = 0; i < 9; ++i) pp.B[i] = P.B[i];
        pp.tab = P.tn_pre;
        pp.n = P.tn_pre_n;
        pp.inv_dq = 1.0 / P.tn_pre_dq;
        pp.scale = 1.0 / (double)P.N;
        hipLaunchKernelGGL((k_precond<T>), dim3((unsigned)(P.nr[0] * P.nr[1])), dim3(128), 0, C.s, P.gate, spec, pitch,
                           P.nzc, P.nr[0], P.nr[1], pp);
        OFDFT_HIP(hipGetLastError());
        fft_c2r<T>(P, spec, spec, out, 1, ncp, ncp, P.N, 1.0, C.s, pitch);
    }

    // get_direction_TN (optimization.py:91-147); returns number of V evaluations.
    // hvp = 0: forward difference with eps (the reference); hvp = 1: exact
    // linear response (needs the last E+V evaluation's work slots at phi).
    int direction_tn(int maxfun, double eps, int hvp_mode) {
        T *dir = F(F_DIR), *best = F(F_BEST), *tres = F(F_TRES), *tp = F(F_TP);
        T *phi1 = F(F_PHI1), *rho1 = F(F_RHO1), *pot1 = F(F_POT1), *Ap = F(F_AP);
        if (hvp_mode) {
            int sl[3];
            work_slots(P, sl);
            const T *work = reinterpret_cast<const T *>(P.work);
            if (sl[0] >= 0) PW_STREAM(k_copy, work + sl[0] * n, F(F_BASE_A), n);
            if (sl[1] >= 0) PW_STREAM(k_copy, work + sl[1] * n, F(F_BASE_C), n);
        }
        // preconditioned CG (when the plan holds a preconditioner): z = M^-1 r,
        // alpha = r.z / p.Ap, p <- z + beta p with beta = r1.z1 / r0.z0
        const bool pre = P.tn_pre_n > 0;
        T *z = F(F_Z);
        auto precond_dot = [&]() {
            apply_precond(tres, z);
            PW_RED(k_dot2, tres, z, n, P.red);
        };
        // |r|^2 -> scal[25] (stopping tests), r.z -> scal[26]; the host reads
        // both with the first iteration's scalars
        PW_RED(k_tn_start, F(F_RES), tres, tp, dir, n, P.red);
        C.reduce_final_cols(1, 25);
        if (pre) {
            precond_dot();
            C.reduce_final_cols(1, 26);
            swap_fields(F_Z, F_TP);  // p0 = z0 (z is recomputed before its next use)
            tp = F(F_TP);
            z = F(F_Z);
        }
        double r0 = 0.0, rz = 0.0, rconv = 0.0;
        bool best_zero = true;
        std::vector<double> rl{1e10};
        int it = 0;
        for (it = 0; it < maxfun; ++it) {
            if (hvp_mode) {
                hvp<T>(P, F(F_RHO), F(F_PHI), tp, F(F_POT), F(F_BASE_A), F(F_BASE_C), P.scal + 2, Ap, C.nb, C.s);
                ++n_v;
            } else {
                PW_STREAM(k_tn_shift, F(F_PHI), tp, eps, phi1, rho1, n);
                eval_v(rho1, pot1);
                PW_RED(k_tn_ap, pot1, phi1, F(F_RES), tp, P.scal + 2, eps,
                                   Ap, n, P.red);
            }
            // pAp -> scal[16]; the update takes alpha = r.z / pAp from the
            // device (r.z of the start, then of the previous iteration);
            // r1 -> scal[17]; r1.z1 -> scal[18]
            C.reduce_final_cols(1, 16);
            const int rz_slot = it == 0 ? (pre ? 26 : 25) : (pre ? 18 : 17);
            PW_RED(k_tn_update_dev, dir, tres, tp, Ap, P.scal + rz_slot,
                               P.scal + 16, n, P.red);
            const int nh = it == 0 ? 11 : (pre ? 3 : 2);  // scal[16..16+nh) to the host
            C.reduce_final_cols(1, 17, pre ? 0 : nh, 16);
            if (pre) {
                precond_dot();
                C.reduce_final_cols(1, 18, nh, 16);
            }
            const double *hh = C.published();
            if (it == 0) {
                r0 = hh[25 - 16];
                rz = pre ? hh[26 - 16] : r0;
                rconv = r0 * 0.1;
            }
            const double pAp = hh[0], r1 = hh[1], rz1 = pre ? hh[2] : hh[1];
            if (pAp < 0.0) {
                if (it == 0) PW_STREAM(k_axpby, rz / pAp, tp, 0.0, dir, n);
                break;
            }
            double rmin = rl[0];
            for (double v : rl) rmin = std::min(rmin, v);
            if (r1 < rmin) {
                PW_STREAM(k_copy, dir, best, n);
                best_zero = false;
            }
            rl.push_back(r1);
            if (r1 < rconv) break;
            double rmin_prev = rl[0];
            for (size_t i = 0; i + 1 < rl.size(); ++i) rmin_prev = std::min(rmin_prev, rl[i]);
            if (r1 > 1000.0 * rmin_prev || (it > 10 && std::fabs(r0 - r1) < 0.1 * r0)) {
                // fall back to the best direction seen
                if (best_zero)
                    PW_STREAM(k_fill, dir, (T)0, n);
                else
                    PW_STREAM(k_copy, best, dir, n);
                break;
            }
            const double beta = rz1 / rz;
            r0 = r1;
            rz = rz1;
            PW_STREAM(k_axpby, 1.0, pre ? z : tres, beta, tp, n);
        }
        return std::min(it + 1, maxfun);
    }

    // OrthogonalNormalization (optimization.py:204-225) of F_P against phi;
    // returns theta0.
    // OrthogonalNormalization (optimization.py:204-225): p -= (sum p phi / sum phi^2) phi,
    // then scaled to sum p^2 = sum phi^2, with the projection and the scale taken
    // from device scalars.  Returns theta0 = sqrt(|p|^2 / N) before scaling and
    // sum(V phi p) after it (the line search's slope at theta = 0 is 2 dV times
    // that), in one host round trip.
    // The same copy also brings back the energies of the last E+V
    // evaluation (scal[32..]), which the first line search needs.
    void orthonormalize(double &theta0, double &vphip, double *energies = nullptr) {
        orth_queue(energies != nullptr);
        orth_parse(C.published(), theta0, vphip, energies);
    }
    void orth_queue(bool energies) {
        T *p = F(F_P), *phi = F(F_PHI);
        PW_RED(k_pphi, p, phi, n, P.red);
        C.reduce_final_cols(2, 21);  // scal[21] = sum p phi, scal[22] = sum phi^2
        hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, C.s, P.gate, P.scal, 3, 21, 22);
        PW_RED(k_orth, p, phi, P.scal + 3, n, P.red);
        C.reduce_final_cols(1, 23);  // scal[23] = sum p^2
        PW_STREAM(k_scale_norm, p, P.scal, 22, 23, n);
        PW_RED(k_dot3, F(F_POT), phi, p, n, P.red);
        // scal[24] = sum V phi p; scal[21..] to the host
        C.reduce_final_cols(1, 24, energies ? 32 + OFDFT_E_COUNT - 21 : 4, 21);
    }
    static void orth_parse(const double *h, double &theta0, double &vphip, double *energies) {
        theta0 = std::sqrt(h[2] / h[1]);
        vphip = h[3];
        if (energies)
            for (int i = 0; i < OFDFT_E_COUNT; ++i) energies[i] = h[32 - 21 + i];
    }

    // the first trial of the line search at theta = min(theta0, theta_prev),
    // queued behind the orthonormalisation (cosine and sine on the device);
    // its energies and gradient sum are published at pinned[kTrialPub..]
    static constexpr int kTrialPub = 64;
    void trial_queue(double theta_prev) {
        hipLaunchKernelGGL(k_trial_theta, dim3(1), dim3(64), 0, C.s, P.gate, P.scal, theta_prev);
        OFDFT_HIP(hipGetLastError());
        PW_STREAM(k_rotate, F(F_PHI), F(F_P), 0.0, 0.0, P.scal + T_C, F(F_NPHI), F(F_NRHO), n);
        eval<T>(P, F(F_NRHO), OFDFT_CALC_E | OFDFT_CALC_V, F(F_NPOT), energies_d, C.s);
        ++n_ev;
        PW_RED(k_theta_grad, F(F_NPOT), F(F_NPHI), F(F_P), F(F_PHI), 0.0, 0.0, P.scal + T_C, n, P.red);
        C.reduce_final_cols(1, 32 + OFDFT_E_COUNT, OFDFT_E_COUNT + 1, 32, kTrialPub);
    }

    // value and derivative at theta (ValueAndDerivative, optimization.py:227-279);
    // the energies and the gradient sum come back in one copy
    void value_deriv(double th, double &e, double &g) {
        const double c = std::cos(th), s = std::sin(th);
        PW_STREAM(k_rotate, F(F_PHI), F(F_P), c, s, (const double *)nullptr, F(F_NPHI), F(F_NRHO), n);
        eval<T>(P, F(F_NRHO), OFDFT_CALC_E | OFDFT_CALC_V, F(F_NPOT), energies_d, C.s);
        ++n_ev;
        PW_RED(k_theta_grad, F(F_NPOT), F(F_NPHI), F(F_P), F(F_PHI), c, s, (const double *)nullptr, n, P.red);
        C.reduce_final_cols(1, 32 + OFDFT_E_COUNT, OFDFT_E_COUNT + 1, 32);
        const double *h = C.published();
        e = h[OFDFT_E_SUM];
        g = 2.0 * h[OFDFT_E_COUNT] * P.dV;
    }

    // b takes a's contents (a becomes scratch)
    void swap_fields(int a, int b) { std::swap(map[a], map[b]); }

    // L-BFGS memory (math_utils.py:374-392): up to bound + 1 (s, y) pairs in
    // slots of one device buffer, oldest first in `order`; rho_i = 1 / (y_i s_i)
    // of slot k sits at scal[LB_RHO + k], the first loop's alpha_i at
    // scal[LB_ALPHA + i], dot products at scal[LB_DOT..]
    static constexpr int LB_RHO = 160, LB_ALPHA = 176, LB_DOT = 192, LB_MAX_PAIRS = 16;
    struct Lbfgs {
        T *buf = nullptr;  // the plan's pair memory (ofdft_plan_reserve_lbfgs)
        int bound = 0;
        double h0 = 0.0;
        std::vector<int> order, free_slots;
    };
    T *lb_s(Lbfgs &L, int k) { return L.buf + (size_t)(2 * k) * n; }
    T *lb_y(Lbfgs &L, int k) { return L.buf + (size_t)(2 * k + 1) * n; }

    void lbfgs_update(Lbfgs &L) {
        if ((int)L.order.size() > L.bound) {
            L.free_slots.push_back(L.order.front());
            L.order.erase(L.order.begin());
        }
        const int k = L.free_slots.back();
        L.free_slots.pop_back();
        // phi_old is F_NPHI and res_old F_RES_PREV after the accept swaps
        PW_RED(k_lbfgs_pair, F(F_PHI), F(F_NPHI), F(F_RES),
                           F(F_RES_PREV), lb_s(L, k), lb_y(L, k), n, P.red);
        C.reduce_final_cols(1, LB_RHO + k);
        hipLaunchKernelGGL(k_recip, dim3(1), dim3(64), 0, C.s, P.gate, P.scal, LB_RHO + k);
        L.order.push_back(k);
    }

    // two-loop recursion (math_utils.py:431-455) into F_P, scalars on the device
    void lbfgs_direction(Lbfgs &L) {
        T *q = F(F_P);
        PW_STREAM(k_neg, F(F_RES), q, n);
        const int m = (int)L.order.size();
        for (int i = m - 1; i >= 1; --i) {  // the reference's first loop stops before the oldest pair
            const int k = L.order[i];
            PW_RED(k_dot2, lb_s(L, k), q, n, P.red);
            C.reduce_final_cols(1, LB_DOT);
            PW_STREAM(k_lbfgs_q, q, lb_y(L, k), P.scal, LB_RHO + k,
                               LB_DOT, LB_ALPHA + i, n);
        }
        if (L.h0 == 0.0) {
            if (m >= 1) {  // gamma = s.y / y.y of the newest pair; gamma = 1 without pairs
                const int k = L.order[m - 1];
                PW_RED(k_sy_yy, lb_s(L, k), lb_y(L, k), n, P.red);
                C.reduce_final_cols(2, LB_DOT + 1);
                PW_STREAM(k_scale_ratio, q, P.scal, LB_DOT + 1,
                                   LB_DOT + 2, n);
            }
        } else {
            PW_STREAM(k_scale, q, L.h0, n);
        }
        for (int i = 0; i < m; ++i) {
            const int k = L.order[i];
            PW_RED(k_dot2, lb_y(L, k), q, n, P.red);
            C.reduce_final_cols(1, LB_DOT);
            PW_STREAM(k_lbfgs_d, q, lb_s(L, k), P.scal, LB_RHO + k,
                               LB_DOT, i == 0 ? -1 : LB_ALPHA + i, n);
        }
    }

    // ---------------------------------------------------------------- device-driven TN
    // the following launches run behind state slot g (g < 0: ungated)
    void gate(int g) { P.gate = g >= 0 ? P.scal + g : nullptr; }
    // reduce K partial columns into scal[slot..slot+K) and take control step
    // `stage`, in one kernel behind gate g
    void red_ctrl(int K, int slot, int stage, int g, int k = 0) {
        hipLaunchKernelGGL(k_red_ctrl, dim3(1), dim3(kFinalThreads), 0, C.s, P.scal + g, P.red, C.nb, K, P.scal + slot,
                           P.scal, stage, k);
        OFDFT_HIP(hipGetLastError());
    }

    // One superstep: CG start, G inner-CG iterations, orthonormalisation, L
    // line-search evaluations and the accept, each behind its gate (DtSlot).
    // The kernels and their order are those of direction_tn, orthonormalize,
    // value_deriv and the accept of run(); the direction buffer is p itself
    // and the accept copies instead of swapping, so every pointer is fixed.
    void enqueue_superstep(int hvp_mode, int G, int L) {
        const bool pre = P.tn_pre_n > 0;
        T *dir = F(F_DIR), *best = F(F_BEST), *tres = F(F_TRES), *tp = F(F_TP), *z = F(F_Z);
        T *phi1 = F(F_PHI1), *rho1 = F(F_RHO1), *pot1 = F(F_POT1), *Ap = F(F_AP);
        const double eps = 1.0e-9;
        gate(G_CGSTART);
        if (hvp_mode) {
            int sl[3];
            work_slots(P, sl);
            const T *work = reinterpret_cast<const T *>(P.work);
            if (sl[0] >= 0) PW_STREAM(k_copy, work + sl[0] * n, F(F_BASE_A), n);
            if (sl[1] >= 0) PW_STREAM(k_copy, work + sl[1] * n, F(F_BASE_C), n);
        }
        PW_RED(k_tn_start, F(F_RES), tres, tp, dir, n, P.red);
        if (pre) {
            C.reduce_final_cols(1, 25);
            apply_precond(tres, z);
            PW_RED(k_dot2, tres, z, n, P.red);
            PW_STREAM(k_copy, z, tp, n);  // p0 = z0
            red_ctrl(1, 26, DS_CGSTART, G_CGSTART);
        } else {
            red_ctrl(1, 25, DS_CGSTART, G_CGSTART);
        }
        for (int k = 0; k < G; ++k) {
            gate(G_CG);
            if (hvp_mode) {
                hvp<T>(P, F(F_RHO), F(F_PHI), tp, F(F_POT), F(F_BASE_A), F(F_BASE_C), P.scal + 2, Ap, C.nb, C.s);
            } else {
                PW_STREAM(k_tn_shift, F(F_PHI), tp, eps, phi1, rho1, n);
                eval<T>(P, rho1, OFDFT_CALC_V, pot1, nullptr, C.s);
                PW_RED(k_tn_ap, pot1, phi1, F(F_RES), tp, P.scal + 2, eps, Ap, n, P.red);
            }
            C.reduce_final_cols(1, 16);
            PW_RED(k_tn_update_dev, dir, tres, tp, Ap, P.scal + T_RZCUR, P.scal + 16, n, P.red);
            if (pre) {
                C.reduce_final_cols(1, 17);
                apply_precond(tres, z);
                PW_RED(k_dot2, tres, z, n, P.red);
                red_ctrl(1, 18, DS_CGITER, G_CG, k);
            } else {
                red_ctrl(1, 17, DS_CGITER, G_CG, k);
            }
            const T *x = pre ? z : tres;
            if (C.vec == 4)
                hipLaunchKernelGGL((k_cg_after<T, 4>), dim3(C.nbw), dim3(256), 0, C.s, P.scal, k, dir, best, tp, x, n);
            else
                hipLaunchKernelGGL((k_cg_after<T, 1>), dim3(C.nbw), dim3(256), 0, C.s, P.scal, k, dir, best, tp, x, n);
            OFDFT_HIP(hipGetLastError());
        }
        gate(G_UPHILL);
        PW_STREAM(k_neg, F(F_RES), dir, n);
        gate(G_ORTH);
        PW_RED(k_pphi, dir, F(F_PHI), n, P.red);
        C.reduce_final_cols(2, 21);
        hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, C.s, P.gate, P.scal, 3, 21, 22);
        PW_RED(k_orth, dir, F(F_PHI), P.scal + 3, n, P.red);
        C.reduce_final_cols(1, 23);
        PW_STREAM(k_scale_norm, dir, P.scal, 22, 23, n);
        PW_RED(k_dot3, F(F_POT), F(F_PHI), dir, n, P.red);
        red_ctrl(1, 24, DS_ORTH, G_ORTH);
        for (int k = 0; k < L; ++k) {
            gate(G_LS);
            PW_STREAM(k_rotate, F(F_PHI), dir, 0.0, 0.0, P.scal + T_C, F(F_NPHI), F(F_NRHO), n);
            eval<T>(P, F(F_NRHO), OFDFT_CALC_E | OFDFT_CALC_V, F(F_NPOT), energies_d, C.s);
            PW_RED(k_theta_grad, F(F_NPOT), F(F_NPHI), dir, F(F_PHI), 0.0, 0.0, P.scal + T_C, n, P.red);
            red_ctrl(1, 32 + OFDFT_E_COUNT, DS_LS, G_LS);
        }
        gate(G_ACCEPT);
        PW_STREAM(k_copy, F(F_NPHI), F(F_PHI), n);
        PW_STREAM(k_copy, F(F_NRHO), F(F_RHO), n);
        PW_STREAM(k_copy, F(F_NPOT), F(F_POT), n);
        // mu_and_residual with the accept's control step in its last reduction
        PW_RED(k_mu_sums, F(F_POT), F(F_RHO), n, P.red);
        C.reduce_final_cols(2, 18);
        hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, C.s, P.gate, P.scal, 2, 18, 19);
        PW_RED(k_residual, F(F_POT), F(F_PHI), F(F_RES), P.scal + 2, n, P.red);
        red_ctrl(1, 20, DS_ACCEPT, G_ACCEPT);
        gate(-1);
    }

    // everything a captured superstep bakes into its kernel arguments
    std::vector<char> graph_key(int hvp_mode, int G, int L) const {
        std::vector<char> k;
        auto put = [&](const void *p, size_t b) {
            const char *c = static_cast<const char *>(p);
            k.insert(k.end(), c, c + b);
        };
        const int ints[] = {hvp_mode, G, L, (int)sizeof(T), C.vec, C.nb, C.nbw, P.nr[0], P.nr[1], P.nr[2], P.nzc,
                            P.full, (int)P.use_fixed, P.var_r2c, P.var_c2r, P.var_cols[0], P.var_cols[1],
                            P.var_cols[2], (int)P.terms, P.tn_pre_n};
        put(ints, sizeof(ints));
        const double dbl[] = {P.x, P.y, P.alpha, P.beta, P.rho0, P.sigma, P.eion, P.dV, P.tn_pre_dq};
        put(dbl, sizeof(dbl));
        put(P.B, sizeof(P.B));
        const void *ptrs[] = {P.spec, P.work, P.red, P.scal, P.aux, P.vloc, P.core, P.tn_pre, P.tw[0], P.tw[1],
                              P.tw[2], P.lmgp};
        put(ptrs, sizeof(ptrs));
        return k;
    }

    // the superstep's executable graph: captured on the plan's own capture
    // stream (non-blocking, default priority) and cached on the plan; a new
    // capture with other kernel arguments updates the cached executable in
    // place when the topology allows
    hipGraphExec_t superstep_graph(int hvp_mode, int G, int L) {
        std::vector<char> key = graph_key(hvp_mode, G, L);
        if (!P.tn_graphs) P.tn_graphs = new TnGraphs();
        TnGraphs &tg = *P.tn_graphs;
        if (tg.exec && tg.key == key) return tg.exec;
        if (!tg.cap) OFDFT_HIP(hipStreamCreateWithFlags(&tg.cap, hipStreamNonBlocking));
        OFDFT_LEGACY_LOCK;
        const hipStream_t user = C.s;
        C.s = tg.cap;
        hipGraph_t g = nullptr;
        OFDFT_HIP(hipStreamBeginCapture(C.s, hipStreamCaptureModeRelaxed));
        try {
            enqueue_superstep(hvp_mode, G, L);
        } catch (...) {
            (void)hipStreamEndCapture(C.s, &g);
            if (g) (void)hipGraphDestroy(g);
            C.s = user;
            P.gate = nullptr;
            throw;
        }
        C.s = user;
        OFDFT_HIP(hipStreamEndCapture(tg.cap, &g));
        bool updated = false;
        if (tg.exec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res;
            updated = hipGraphExecUpdate(tg.exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            (void)hipGetLastError();
            if (updated) ++tg.updates;
            if (!updated) {
                (void)hipGraphExecDestroy(tg.exec);
                tg.exec = nullptr;
            }
        }
        if (!updated) {
            const hipError_t e = hipGraphInstantiate(&tg.exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                tg.exec = nullptr;
                (void)hipGraphDestroy(g);
                OFDFT_HIP(e);
            }
            ++tg.captures;
        }
        (void)hipGraphDestroy(g);
        tg.key = key;
        return tg.exec;
    }

    // Opt::run's TN with the device-driven iteration: one graph launch and
    // one read-back of DT_READ scalars per superstep
    int run_graph(T *rho_io, T *pot_out, int hvp_mode, const double *opts, int G, int L, double *stats,
                  double *history, int hcap) {
        OFDFT_REQUIRE(G >= 1 && G <= 64 && L >= 1 && L <= 64, "optimize_tn_graph: steps per superstep out of range");
        OFDFT_REQUIRE(!(P.terms & OFDFT_TERM_LMGP), "optimize_tn_graph: LMGP is not supported");
        for (int d = 0; d < 3; ++d)
            OFDFT_REQUIRE(P.use_fixed && (P.nr[d] == 1 || fixed_size(P.nr[d])),
                          "optimize_tn_graph: every axis needs a size-specialised FFT pass (2/3/5-smooth sizes of "
                          "the fixed table)");
        const double econv = opts[0];
        const int maxiter = (int)opts[1], maxfun = (int)opts[2], maxls = (int)opts[3];
        const auto t0 = std::chrono::steady_clock::now();
        auto secs = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
        struct GateReset {
            Plan &P;
            ~GateReset() { P.gate = nullptr; }
        } reset{P};
        // start (optimization.py:305-316), as in run()
        PW_RED(k_init_phi, rho_io, F(F_PHI), n, P.red);
        C.reduce_final_cols(2, 16);
        PW_STREAM(k_scale_square, F(F_PHI), F(F_RHO), P.scal + 16, n);
        eval<T>(P, F(F_RHO), OFDFT_CALC_E | OFDFT_CALC_V, F(F_POT), energies_d, C.s);
        mu_and_residual(F(F_POT), F(F_RHO), F(F_PHI), F(F_RES));
        DtOpts o{econv, opts[4], opts[5], opts[6], P.dV, maxiter, maxfun, maxls, (int)opts[7],
                 P.prec == OFDFT_FP32 ? 1 : 0, P.tn_pre_n > 0 ? 1 : 0};
        hipLaunchKernelGGL(k_tn_init, dim3(1), dim3(64), 0, C.s, P.scal, o);
        OFDFT_HIP(hipGetLastError());
        std::vector<double> times{secs()};
        int steps = 0, status = maxiter > 1 ? -1 : 2, iters = 0;
        double nv = 0.0, nev = 0.0, energy = 0.0;
        if (status < 0) {
            const hipGraphExec_t ge = superstep_graph(hvp_mode, G, L);
            // every superstep moves one section on (DtSlot), so this bounds a run
            const long long cap =
                (long long)maxiter * ((maxfun + G - 1) / G + (maxls + L - 1) / L + 4) + 4;
            while (status < 0) {
                OFDFT_REQUIRE(++steps <= cap, "optimize_tn_graph: no progress (internal error)");
                OFDFT_HIP(hipGraphLaunch(ge, C.s));
                const double *h = C.to_host(DT_READ, T_IT);
                const int it = (int)h[0];
                while ((int)times.size() <= it) times.push_back(secs());
                status = (int)h[T_STATUS - T_IT];
            }
        }
        PW_STREAM(k_copy, F(F_RHO), rho_io, n);
        if (pot_out) PW_STREAM(k_copy, F(F_POT), pot_out, n);
        const double *h = C.to_host(DT_READ, T_IT);
        iters = (int)h[0];
        energy = h[T_ENERGY - T_IT];
        nv = h[T_NV - T_IT];
        nev = h[T_NEV - T_IT];
        const double mu = C.to_host(1, 2)[0];
        std::vector<double> hist(std::min(iters + 1, DT_HCAP));
        OFDFT_HIP(hipMemcpyAsync(hist.data(), P.scal + T_HIST, hist.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 C.s));
        C.wait();
        if (stats) {
            for (int i = 0; i < 16; ++i) stats[i] = 0.0;
            stats[0] = status;
            stats[1] = iters;
            stats[2] = energy;
            stats[3] = mu;
            stats[4] = nv;
            stats[5] = nev + 1.0;  // and the starting evaluation
            stats[6] = secs();
            stats[7] = steps;
            stats[8] = P.tn_graphs ? P.tn_graphs->captures : 0;
            stats[9] = P.tn_graphs ? P.tn_graphs->updates : 0;
        }
        if (history)
            for (int i = 0; i < hcap && i <= iters && i < (int)hist.size(); ++i) {
                history[i] = hist[i];
                history[hcap + i] = i < (int)times.size() ? times[i] : times.back();
            }
        return status;
    }

    int run(T *rho_io, T *pot_out, int method, int hvp, const double *opts, double *stats, double *history,
            int hcap) {
        const double econv = opts[0];
        const int maxiter = (int)opts[1], maxfun = (int)opts[2], maxls = (int)opts[3];
        const double c1 = opts[4], c2 = opts[5], xtol = opts[6];
        const int ncheck = (int)opts[7];
        Lbfgs lb;
        if (method == OFDFT_OPT_LBFGS) {
            lb.bound = (int)opts[8];
            lb.h0 = opts[9];
            OFDFT_REQUIRE(lb.bound >= 0 && lb.bound + 1 <= LB_MAX_PAIRS, "optimize: maxcor must be in [0, 15]");
            OFDFT_REQUIRE(P.lbfgs_buf != nullptr && P.lbfgs_pairs >= lb.bound + 1,
                          "optimize: L-BFGS needs the plan's pair memory; call ofdft_plan_reserve_lbfgs(plan, maxcor) "
                          "first");
            lb.buf = reinterpret_cast<T *>(P.lbfgs_buf);
            for (int k = lb.bound; k >= 0; --k) lb.free_slots.push_back(k);
        }
        const auto t0 = std::chrono::steady_clock::now();
        // initial phi (optimization.py:305-316), normalised on the device; the
        // starting energy reaches the host with the first orthonormalisation
        PW_RED(k_init_phi, rho_io, F(F_PHI), n, P.red);
        C.reduce_final_cols(2, 16);
        PW_STREAM(k_scale_square, F(F_PHI), F(F_RHO), P.scal + 16, n);
        eval<T>(P, F(F_RHO), OFDFT_CALC_E | OFDFT_CALC_V, F(F_POT), energies_d, C.s);
        ++n_ev;
        mu_and_residual(F(F_POT), F(F_RHO), F(F_PHI), F(F_RES));
        bool energy_pending = true;
        double energy = 0.0;
        std::vector<double> hist{energy};
        std::vector<double> times{std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()};
        double theta = 0.1;
        int converged = 2, it = 0, nres = 0;
        for (it = 1; it < maxiter; ++it) {
            int nd = 1;
            if (method == OFDFT_OPT_TN) {
                nd = direction_tn(maxfun, 1.0e-9, hvp);
                swap_fields(F_DIR, F_P);  // p takes the direction (dir is reset by the next TN start)
            } else if (method == OFDFT_OPT_LBFGS) {
                lbfgs_direction(lb);
            } else {
                // CG (math_utils.py:395-423); first step and missing history: -res;
                // GD (math_utils.py:426-428) always -res
                if (nres < 1 || method == OFDFT_OPT_GD) {
                    PW_STREAM(k_neg, F(F_RES), F(F_P), n);
                } else {
                    PW_RED(k_cg_sums, F(F_RES), F(F_RES_PREV),
                                       F(F_DIR_PREV), n, P.red);
                    const double *q = C.fetch(5);
                    if (method == OFDFT_OPT_CG_CD) {
                        PW_STREAM(k_cg_cd, F(F_RES), F(F_DIR_PREV), q[4],
                                           F(F_P), n);
                    } else {
                        double beta;
                        if (method == OFDFT_OPT_CG_HS)
                            beta = q[0] / q[1];
                        else if (method == OFDFT_OPT_CG_PR)
                            beta = std::max(q[0] / q[3], 0.0);
                        else if (method == OFDFT_OPT_CG_DY)
                            beta = q[2] / q[1];
                        else if (method == OFDFT_OPT_CG_LS)
                            beta = q[0] / q[4];
                        else
                            beta = q[2] / q[3];
                        PW_STREAM(k_copy, F(F_DIR_PREV), F(F_P), n);
                        PW_STREAM(k_axpby, -1.0, F(F_RES), beta, F(F_P), n);
                    }
                }
            }
            double theta0, vphip;
            // FP32 TN: the line search's first trial (its step follows from
            // the orthonormalisation's sums) is queued behind them, and both
            // come back in one round trip
            bool have_trial = false;
            double trial_f = 0.0, trial_g = 0.0;
            if (P.prec == OFDFT_FP32 && method == OFDFT_OPT_TN) {
                orth_queue(energy_pending);
                trial_queue(theta);
                const double *h = C.published();
                orth_parse(h, theta0, vphip, energy_pending ? energies_h : nullptr);
                trial_f = h[kTrialPub + OFDFT_E_SUM];
                trial_g = 2.0 * h[kTrialPub + OFDFT_E_COUNT] * P.dV;
                have_trial = true;
            } else {
                orthonormalize(theta0, vphip, energy_pending ? energies_h : nullptr);
            }
            if (energy_pending) {
                energy = hist[0] = energies_h[OFDFT_E_SUM];
                energy_pending = false;
            }
            if (vphip > 0) {  // uphill: restart from the steepest descent direction
                have_trial = false;
                PW_STREAM(k_neg, F(F_RES), F(F_P), n);
                orthonormalize(theta0, vphip);
            }
            theta = std::min(theta0, theta);
            // line search (math_utils.py:76-106) with func0 at theta = 0, where
            // dE/dtheta = 2 dV sum(V phi p)
            double f = energy, g = 2.0 * vphip * P.dV;
            double st[16];
            LsTask task = LS_START;
            double stp = theta;
            bool ok = false, evaluated = false;
            if (P.prec == OFDFT_FP32 && std::fabs(g) * theta < 0.1 * econv) {
                // FP32: the first-order energy change of the trial step is
                // below a tenth of the stop threshold, so the Wolfe tests
                // would compare energy differences at single-precision
                // resolution and could only exhaust maxls; the trial (for TN
                // the Newton) step is taken as it is, with one evaluation
                if (have_trial) {
                    f = trial_f;
                    g = trial_g;
                } else {
                    value_deriv(stp, f, g);
                }
                ok = evaluated = true;
            } else {
                for (int i = 0; i < maxls; ++i) {
                    task = dcsrch(stp, f, g, c1, c2, xtol, 0.0, M_PI, task, st);
                    if (task == LS_FG) {
                        if (have_trial && stp == theta) {  // the queued trial is dcsrch's first
                            f = trial_f;
                            g = trial_g;
                        } else {
                            value_deriv(stp, f, g);
                        }
                        have_trial = false;
                        evaluated = true;
                    } else {
                        ok = (task == LS_CONVERGENCE);
                        break;
                    }
                }
            }
            if (!ok || !evaluated) {
                converged = 1;  // optimization.py:440-444
                break;
            }
            theta = stp;
            // accept: phi <- newphi, rho <- newrho, pot <- newpot
            swap_fields(F_NPHI, F_PHI);
            swap_fields(F_NRHO, F_RHO);
            swap_fields(F_NPOT, F_POT);
            energy = f;
            if (method != OFDFT_OPT_TN) {
                swap_fields(F_RES, F_RES_PREV);
                swap_fields(F_P, F_DIR_PREV);
                ++nres;
            }
            mu_and_residual(F(F_POT), F(F_RHO), F(F_PHI), F(F_RES));
            if (method == OFDFT_OPT_LBFGS) lbfgs_update(lb);
            hist.push_back(energy);
            times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            (void)nd;
            if ((int)hist.size() - 1 >= ncheck) {
                bool conv = true;
                for (int i = 0; i < ncheck; ++i)
                    if (std::fabs(hist[hist.size() - 2 - i] - energy) > econv) conv = false;
                if (conv) {
                    converged = 0;
                    break;
                }
            }
        }
        PW_STREAM(k_copy, F(F_RHO), rho_io, n);
        if (pot_out) PW_STREAM(k_copy, F(F_POT), pot_out, n);
        // mu of the final density (scal[2], set by the last residual) and, when
        // no iteration ran, the starting energy
        const double *hm = C.to_host(32 + OFDFT_E_COUNT - 2, 2);
        const double mu = hm[0];
        if (energy_pending) energy = hist[0] = hm[32 - 2 + OFDFT_E_SUM];
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) {
            for (int i = 0; i < 16; ++i) stats[i] = 0.0;
            stats[0] = converged;
            stats[1] = (double)(hist.size() - 1);
            stats[2] = energy;
            stats[3] = mu;
            stats[4] = n_v;
            stats[5] = n_ev;
            stats[6] = secs;
        }
        if (history)
            for (int i = 0; i < hcap && i < (int)hist.size(); ++i) {
                history[i] = hist[i];
                history[hcap + i] = times[i];
            }
        return converged;
    }
};

#undef PW_STREAM
#undef PW_RED
#undef PW_LAUNCH_

// The TN preconditioner table M^-1(q_i), M = 2 rho0 K(q_i) at q_i = i dq, K
// the uniform-gas response of the functional's terms; 0 where M is infinite
// (Hartree at q = 0)
__global__ void k_tn_precond_fill(double *tab, int n, double dq, double rho0, RespK r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double M = 2.0 * rho0 * resp_k<double>(r, i * dq);
    tab[i] = isinf(M) ? 0.0 : 1.0 / fmax(M, 1e-8);
}

void tn_precond_fill(Plan &P, int n, double dq, double rho0, const double *resp, hipStream_t s) {
    const RespK r{resp[0], resp[1], resp[2], resp[3], resp[4], resp[5] != 0.0 ? 1 : 0};
    hipLaunchKernelGGL(k_tn_precond_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P.tn_pre, n, dq, rho0,
                       r);
    OFDFT_HIP(hipGetLastError());
}

}  // namespace ofdft

using namespace ofdft;

extern "C" int ofdft_optimize_tn_graph(ofdft_plan *plan, void *rho, void *pot, int hvp, const double *opts,
                                      int cg_steps, int ls_steps, double *stats, double *history, int hcap,
                                      void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        OFDFT_REQUIRE(P && rho && opts, "optimize_tn_graph: bad arguments");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (P->prec == OFDFT_FP64) {
            Opt<double> o(*P, s);
            o.run_graph((double *)rho, (double *)pot, hvp, opts, cg_steps, ls_steps, stats, history, hcap);
        } else {
            Opt<float> o(*P, s);
            o.run_graph((float *)rho, (float *)pot, hvp, opts, cg_steps, ls_steps, stats, history, hcap);
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}

extern "C" int ofdft_optimize(ofdft_plan *plan, void *rho, void *pot, int method, int hvp, const double *opts,
                              double *stats, double *history, int hcap, void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        OFDFT_REQUIRE(P && rho && opts, "optimize: bad arguments");
        OFDFT_REQUIRE(method >= OFDFT_OPT_TN && method <= OFDFT_OPT_LBFGS, "optimize: unknown method");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (P->prec == OFDFT_FP64) {
            Opt<double> o(*P, s);
            o.run((double *)rho, (double *)pot, method, hvp, opts, stats, history, hcap);
        } else {
            Opt<float> o(*P, s);
            o.run((float *)rho, (float *)pot, method, hvp, opts, stats, history, hcap);
        }
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}
