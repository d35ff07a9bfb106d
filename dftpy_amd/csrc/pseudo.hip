// Local pseudopotential and Ewald ion-ion energy on the grid (setup, FP64).
//
//  * vloc(|G|) from a per-species piecewise-cubic table (the interpolating
//    cubic spline of pseudo/__init__.py:641-658) with the quartic patch
//    below the first table step and zero beyond the table
//    (pseudo/__init__.py:350-375).
//  * exact structure factor exp(-i G.R) (ions.py:121-124,
//    pseudo/__init__.py:377-387) or order-n cardinal B-spline PME
//    (ewald.py:11-194, pseudo/__init__.py:389-404).
//  * Ewald: real-space erfc sum over periodic images (ewald.py:280-320),
//    reciprocal sum on the G grid with the half-space mask
//    (grid.py:657-695; ewald.py:448-465, 786-808), correction
//    (ewald.py:467-482).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gspace.h"
#include "internal.h"
#include "pointwise.h"
#include "reduce.h"

namespace ofdft {

constexpr int kMaxSpecies = 8;
constexpr int kMaxOrder = 16;

struct VlocTable {
    const double *coef[kMaxSpecies];
    double q0[kMaxSpecies], dq[kMaxSpecies], qmax[kMaxSpecies];
    int ncoef[kMaxSpecies];
    double v0[kMaxSpecies], v1[kMaxSpecies], v2[kMaxSpecies], dqs[kMaxSpecies];
};

__device__ __forceinline__ double vloc_eval(const VlocTable &t, int s, double q) {
    if (q < t.dqs[s]) {
        // quartic through (v2, v1, v0, v1, v2) (math_utils.py:457-467), odd terms vanish
        const double f0 = t.v2[s], f1 = t.v1[s], f2 = t.v0[s];
        const double t2 = -f0 + 16.0 * f1 - 30.0 * f2 + 16.0 * f1 - f0;
        const double t4 = f0 - 4.0 * f1 + 6.0 * f2 - 4.0 * f1 + f0;
        const double dx = q / t.dqs[s];
        return f2 + dx * (dx * (t2 + dx * (dx * t4))) / 24.0;
    }
    if (!(q < t.qmax[s])) return 0.0;
    int i = (int)floor((q - t.q0[s]) / t.dq[s]);
    if (i < 0) i = 0;
    if (i >= t.ncoef[s]) i = t.ncoef[s] - 1;
    const double x = q - (t.q0[s] + i * t.dq[s]);
    const double *c = t.coef[s] + 4 * (long long)i;
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
}

// Screening densities of the ions, rho_a(q) = -v_a(q) / K(q) (the linear
// response of the uniform gas to each local pseudopotential, rho_a(0) = Z):
// below the first table step the quartic through (r2, r1, Z, r1, r2), the
// screened values at the first two table points; above it the PP table
// divided by K on the fly
struct ScreenParams {
    RespK k;
    double z[kMaxSpecies], r1[kMaxSpecies], r2[kMaxSpecies];
};

// the response in T (FP32 fields: FP32 hardware math)
template <bool SCREEN, typename T = double>
__device__ __forceinline__ double vloc_or_screened(const VlocTable &t, const ScreenParams &sp, int s, double q) {
    if constexpr (!SCREEN) {
        return vloc_eval(t, s, q);
    } else {
        if (q < t.dqs[s]) {
            const double f0 = sp.r2[s], f1 = sp.r1[s], f2 = sp.z[s];
            const double t2 = -f0 + 16.0 * f1 - 30.0 * f2 + 16.0 * f1 - f0;
            const double t4 = f0 - 4.0 * f1 + 6.0 * f2 - 4.0 * f1 + f0;
            const double dx = q / t.dqs[s];
            return f2 + dx * (dx * (t2 + dx * (dx * t4))) / 24.0;
        }
        const double v = vloc_eval(t, s, q);
        return v == 0.0 ? 0.0 : -v / (double)resp_k<T>(sp.k, (T)q);
    }
}

__device__ __forceinline__ int ffreq_(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }
__device__ __forceinline__ int wrapf_(int f, int n) { return ((n & 1) == 0 && f == n / 2) ? -n / 2 : f; }

struct GGeom {
    double B[9];
    int nx, ny, nz, nzc, full;
};

__device__ __forceinline__ void gvec(const GGeom &g, int fx, int fy, int fz, double *G) {
    G[0] = fx * g.B[0] + fy * g.B[3] + fz * g.B[6];
    G[1] = fx * g.B[1] + fy * g.B[4] + fz * g.B[7];
    G[2] = fx * g.B[2] + fy * g.B[5] + fz * g.B[8];
}

// integer frequencies of half-grid point (ix,iy,iz) in the grid's convention
// and, on a full grid at a Nyquist point, those of the alias partner.
__device__ __forceinline__ bool freqs_of(const GGeom &g, int ix, int iy, int iz, int *f, int *fp) {
    f[0] = ffreq_(ix, g.nx);
    f[1] = ffreq_(iy, g.ny);
    f[2] = iz;
    if (!g.full) return false;
    f[2] = wrapf_(iz, g.nz);
    const bool nyq = ((g.nx & 1) == 0 && ix == g.nx / 2) || ((g.ny & 1) == 0 && iy == g.ny / 2) ||
                     ((g.nz & 1) == 0 && iz == g.nz / 2);
    if (!nyq) return false;
    fp[0] = wrapf_(-f[0], g.nx);
    fp[1] = wrapf_(-f[1], g.ny);
    fp[2] = wrapf_(-f[2], g.nz);
    return true;
}

// v(G) = sum_a vloc_s(a)(|G|) exp(-i G.R_a) on the half grid, scaled (FP64
// per point, stored in the plan's precision T).
template <typename T, bool SCREEN>
__global__ void k_vloc_exact(cpx<T> *__restrict__ out, long long nc, GGeom g, VlocTable t, ScreenParams sp,
                             int nat, const double *__restrict__ pos, const int *__restrict__ spec, double scale) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        const int iz = (int)(i % g.nzc);
        const long long rest = i / g.nzc;
        const int iy = (int)(rest % g.ny), ix = (int)(rest / g.ny);
        int f[3], fp[3];
        const bool two = freqs_of(g, ix, iy, iz, f, fp);
        double G1[3], G2[3];
        gvec(g, f[0], f[1], f[2], G1);
        double vs1[kMaxSpecies], vs2[kMaxSpecies];
        const double q1 = sqrt(G1[0] * G1[0] + G1[1] * G1[1] + G1[2] * G1[2]);
        for (int s = 0; s < kMaxSpecies; ++s)
            vs1[s] = t.coef[s] ? vloc_or_screened<SCREEN, T>(t, sp, s, q1) : 0.0;
        if (two) {
            // partner enters as f(-G_partner)
            gvec(g, fp[0], fp[1], fp[2], G2);
            G2[0] = -G2[0];
            G2[1] = -G2[1];
            G2[2] = -G2[2];
            const double q2 = sqrt(G2[0] * G2[0] + G2[1] * G2[1] + G2[2] * G2[2]);
            for (int s = 0; s < kMaxSpecies; ++s)
                vs2[s] = t.coef[s] ? vloc_or_screened<SCREEN, T>(t, sp, s, q2) : 0.0;
        }
        double re = 0.0, im = 0.0;
        for (int a = 0; a < nat; ++a) {
            const double *R = pos + 3 * a;
            double sn, cs;
            sincos(G1[0] * R[0] + G1[1] * R[1] + G1[2] * R[2], &sn, &cs);
            double w = two ? 0.5 : 1.0;
            re += w * vs1[spec[a]] * cs;
            im -= w * vs1[spec[a]] * sn;
            if (two) {
                sincos(G2[0] * R[0] + G2[1] * R[1] + G2[2] * R[2], &sn, &cs);
                re += 0.5 * vs2[spec[a]] * cs;
                im -= 0.5 * vs2[spec[a]] * sn;
            }
        }
        out[i] = cpx<T>{(T)(re * scale), (T)(im * scale)};
    }
}

// Cardinal B-spline values M_n(x + k), k = 0..order-1 -> m[k] (ewald.py:66-86)
__device__ __forceinline__ void calc_mn(double x, int order, double *m /* order+1 */) {
    for (int i = 0; i <= order; ++i) m[i] = 0.0;
    m[1] = x;
    m[2] = 1.0 - x;
    for (int i = 3; i <= order; ++i)
        for (int j = 0; j < i; ++j) {
            const int n = i - j;
            m[n] = ((x + n - 1) * m[n] + (j + 1 - x) * m[n - 1]) / (i - 1);
        }
}

// Q[floor(u)+1-i] += w M(frac(u)+i) per axis, one workgroup per atom.
template <typename T>
__global__ void k_pme_spread(T *__restrict__ Q, int nx, int ny, int nz, int order, int nat,
                             const double *__restrict__ frac, const double *__restrict__ w, const int *__restrict__ spec,
                             int want_species) {
    const int a = blockIdx.x;
    if (a >= nat) return;
    if (want_species >= 0 && spec[a] != want_species) return;
    __shared__ double mx[kMaxOrder + 1], my[kMaxOrder + 1], mz[kMaxOrder + 1];
    __shared__ int fl[3];
    const int nr[3] = {nx, ny, nz};
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        const double u = frac[3 * a + d] * nr[d];
        const double f = floor(u);
        fl[d] = (int)f;
        double m[kMaxOrder + 1];
        calc_mn(u - f, order, m);
        for (int k = 0; k <= order; ++k) (d == 0 ? mx : (d == 1 ? my : mz))[k] = m[k];
    }
    __syncthreads();
    const double wa = w ? w[a] : 1.0;
    const int o3 = order * order * order;
    for (int e = threadIdx.x; e < o3; e += blockDim.x) {
        const int i = e / (order * order), j = (e / order) % order, k = e % order;
        int lx = (1 + fl[0] - i) % nx;
        if (lx < 0) lx += nx;
        int ly = (1 + fl[1] - j) % ny;
        if (ly < 0) ly += ny;
        int lz = (1 + fl[2] - k) % nz;
        if (lz < 0) lz += nz;
        atomicAdd(&Q[((long long)lx * ny + ly) * nz + lz], (T)(wa * mx[i + 1] * my[j + 1] * mz[k + 1]));
    }
}

struct BTab {
    const double2 *bx, *by, *bz;
};

// acc(G) += vloc_s(|G|)_eff * S_s(G) * B(G)  (S_s = rfftn of the species Q array)
template <typename T, bool SCREEN>
__global__ void k_vloc_pme_acc(cpx<T> *__restrict__ acc, const cpx<T> *__restrict__ sq, long long nc, GGeom g,
                               VlocTable t, ScreenParams sp, int s, BTab b, int first) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        const int iz = (int)(i % g.nzc);
        const long long rest = i / g.nzc;
        const int iy = (int)(rest % g.ny), ix = (int)(rest / g.ny);
        int f[3], fp[3];
        const bool two = freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        double v = vloc_or_screened<SCREEN, T>(t, sp, s, sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]));
        if (two) {
            gvec(g, fp[0], fp[1], fp[2], G);
            v = 0.5 * (v + vloc_or_screened<SCREEN, T>(t, sp, s, sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2])));
        }
        const double2 bx = b.bx[ix], by = b.by[iy], bz = b.bz[iz];
        double2 bb{bx.x * by.x - bx.y * by.y, bx.x * by.y + bx.y * by.x};
        bb = double2{bb.x * bz.x - bb.y * bz.y, bb.x * bz.y + bb.y * bz.x};
        const cpx<T> q = sq[i];
        const double2 val{v * ((double)q.x * bb.x - (double)q.y * bb.y), v * ((double)q.x * bb.y + (double)q.y * bb.x)};
        if (first)
            acc[i] = cpx<T>{(T)val.x, (T)val.y};
        else
            acc[i] = cpx<T>{(T)((double)acc[i].x + val.x), (T)((double)acc[i].y + val.y)};
    }
}

// half-space deduplication mask of grid.py:657-695 (serial form)
__device__ __forceinline__ bool gmask(int ix, int iy, int iz, int nx, int ny, int nz, int full) {
    const int D0 = nx / 2, D1 = ny / 2, D2 = nz / 2;
    const bool e0 = (nx & 1) == 0, e1 = (ny & 1) == 0, e2 = (nz & 1) == 0;
    if (full && iz > D2) return false;
    if (ix == 0 && iy > D1 && iz == 0) return false;
    if (ix > D0 && iz == 0) return false;
    if (e2) {
        if (ix == 0 && iy == 0 && iz == D2) return false;
        if (ix == 0 && iy > D1 && iz == D2) return false;
        if (ix > D0 && iz == D2) return false;
        if (e1 && ix == 0 && iy == D1 && iz == D2) return false;
        if (e0 && ix == D0 && iy == 0 && iz == D2) return false;
        if (e0 && ix == D0 && iy > D1 && iz == D2) return false;
    }
    if (e0) {
        if (ix == D0 && iy > D1 && iz == 0) return false;
        if (e1 && ix == D0 && iy == D1 && iz == 0) return false;
    }
    if (e1 && ix == 0 && iy == D1 && iz == 0) return false;
    if (e0 && e1 && e2 && ix == D0 && iy == D1 && iz == D2) return false;
    return true;
}

// sum over masked G != 0 of |S(G)|^2 exp(-G^2/(4 eta)) / G^2 with S either
// given (PME: rfftn(Q) * B) or computed from the charges (exact).
__global__ void k_ewald_rec(const double2 *__restrict__ sq, long long nc, GGeom g, double eta, BTab b, int nat,
                            const double *__restrict__ pos, const double *__restrict__ q, double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        const int iz = (int)(i % g.nzc);
        const long long rest = i / g.nzc;
        const int iy = (int)(rest % g.ny), ix = (int)(rest / g.ny);
        if (!gmask(ix, iy, iz, g.nx, g.ny, g.nz, g.full)) continue;
        if (ix == 0 && iy == 0 && iz == 0) continue;
        int f[3], fp[3];
        freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        const double gg = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        double s2;
        if (sq) {
            const double2 bx = b.bx[ix], by = b.by[iy], bz = b.bz[iz];
            double2 bb{bx.x * by.x - bx.y * by.y, bx.x * by.y + bx.y * by.x};
            bb = double2{bb.x * bz.x - bb.y * bz.y, bb.x * bz.y + bb.y * bz.x};
            const double2 v = sq[i];
            const double re = v.x * bb.x - v.y * bb.y, im = v.x * bb.y + v.y * bb.x;
            s2 = re * re + im * im;
        } else {
            double re = 0.0, im = 0.0;
            for (int a = 0; a < nat; ++a) {
                double sn, cs;
                sincos(G[0] * pos[3 * a] + G[1] * pos[3 * a + 1] + G[2] * pos[3 * a + 2], &sn, &cs);
                re += q[a] * cs;
                im -= q[a] * sn;
            }
            s2 = re * re + im * im;
        }
        acc += s2 * exp(-gg / (4.0 * eta)) / gg;
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// real-space sum over atom pairs and images: q_i q_j erfc(sqrt(eta) d) / d
__global__ void k_ewald_real(int nat, const double *__restrict__ pos, const double *__restrict__ q, double L00,
                             double L01, double L02, double L10, double L11, double L12, double L20, double L21,
                             double L22, int N0, int N1, int N2, double rmax, double seta, double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    const long long npair = (long long)nat * nat;
    for (long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x; pidx < npair;
         pidx += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(pidx / nat), j = (int)(pidx % nat);
        const double dx0 = pos[3 * i] - pos[3 * j], dy0 = pos[3 * i + 1] - pos[3 * j + 1],
                     dz0 = pos[3 * i + 2] - pos[3 * j + 2];
        double s = 0.0;
        for (int a = -N0; a <= N0; ++a)
            for (int bb = -N1; bb <= N1; ++bb)
                for (int c = -N2; c <= N2; ++c) {
                    const double rx = a * L00 + bb * L10 + c * L20;
                    const double ry = a * L01 + bb * L11 + c * L21;
                    const double rz = a * L02 + bb * L12 + c * L22;
                    const double dx = dx0 + rx, dy = dy0 + ry, dz = dz0 + rz;
                    const double d = sqrt(dx * dx + dy * dy + dz * dz);
                    if (d < rmax && d > 0.001) s += erfc(seta * d) / d;
                }
        acc += q[i] * q[j] * s;
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ void k_fill_zero(double *__restrict__ x, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// ------------------------------------------------------------------ host
static GGeom geom_of(const Plan &P) {
    GGeom g;
    for (int i = 0; i < 9; ++i) g.B[i] = P.B[i];
    g.nx = P.nr[0];
    g.ny = P.nr[1];
    g.nz = P.nr[2];
    g.nzc = P.nzc;
    g.full = P.full;
    return g;
}

static void host_calc_mn(double x, int order, double *m) {
    for (int i = 0; i <= order; ++i) m[i] = 0.0;
    m[1] = x;
    m[2] = 1.0 - x;
    for (int i = 3; i <= order; ++i)
        for (int j = 0; j < i; ++j) {
            const int n = i - j;
            m[n] = ((x + n - 1) * m[n] + (j + 1 - x) * m[n - 1]) / (i - 1);
        }
}

// Euler exponential spline factors b(m) (ewald.py:88-104), per axis
static std::vector<double> bspline_bm(int n, int len, int order) {
    double M1[kMaxOrder + 1];
    host_calc_mn(1.0, order, M1);
    std::vector<double> out(2 * len);
    for (int m = 0; m < len; ++m) {
        const double q = 2.0 * M_PI * m / n;
        double fr = 0.0, fi = 0.0;
        for (int k = 1; k < order; ++k) {
            fr += M1[k] * std::cos(k * q);
            fi -= M1[k] * std::sin(k * q);
        }
        const double tr = std::cos((order - 1.0) * q), ti = -std::sin((order - 1.0) * q);
        const double den = fr * fr + fi * fi;
        out[2 * m] = (tr * fr + ti * fi) / den;
        out[2 * m + 1] = (ti * fr - tr * fi) / den;
    }
    return out;
}

// Host inputs of one setup call go through the plan's staging pair: filled
// in pinned host memory and sent with one asynchronous copy.  An event
// recorded after each upload guards the host buffer: the next call waits for
// it (normally long complete) before refilling, so a routine that only
// produces device results need not wait for its stream.
static char *stage_host(Plan &P, size_t bytes) {
    if (P.stg_evt) {
        hipError_t e;
        while ((e = hipEventQuery(P.stg_evt)) == hipErrorNotReady) {
        }
        OFDFT_HIP(e);
    }
    if (bytes > P.stg_cap) {
        OFDFT_LEGACY_LOCK;
        // kernels of an earlier call may still read the device copy
        if (P.stg_d) OFDFT_HIP(hipDeviceSynchronize());
        if (P.stg_h) OFDFT_HIP(hipHostFree(P.stg_h));
        if (P.stg_d) OFDFT_HIP(hipFree(P.stg_d));
        P.stg_h = P.stg_d = nullptr;
        P.stg_cap = 0;
        const size_t cap = std::max<size_t>(bytes, 64 * 1024);
        OFDFT_HIP(hipHostMalloc((void **)&P.stg_h, cap));
        OFDFT_HIP(hipMalloc((void **)&P.stg_d, cap));
        P.stg_cap = cap;
    }
    return P.stg_h;
}

static char *stage_upload(Plan &P, size_t bytes, hipStream_t s) {
    OFDFT_HIP(hipMemcpyAsync(P.stg_d, P.stg_h, bytes, hipMemcpyHostToDevice, s));
    if (!P.stg_evt) OFDFT_HIP(hipEventCreateWithFlags(&P.stg_evt, hipEventDisableTiming));
    OFDFT_HIP(hipEventRecord(P.stg_evt, s));
    return P.stg_d;
}

// byte offsets of consecutive staged arrays (8-byte aligned)
struct StageLayout {
    size_t size = 0;
    size_t add(size_t bytes) {
        const size_t o = size;
        size += (bytes + 7) & ~(size_t)7;
        return o;
    }
};

void plan_setup_free(Plan &P) {
    OFDFT_LEGACY_LOCK;
    if (P.stg_evt) (void)hipEventSynchronize(P.stg_evt);
    if (P.stg_evt) (void)hipEventDestroy(P.stg_evt);
    if (P.stg_h) (void)hipHostFree(P.stg_h);
    if (P.stg_d) (void)hipFree(P.stg_d);
    if (P.pme_b) (void)hipFree(P.pme_b);
    if (P.fs_ws) (void)hipFree(P.fs_ws);
    P.fs_ws = nullptr;
    P.fs_cap = 0;
    P.stg_h = P.stg_d = nullptr;
    P.stg_cap = 0;
    P.pme_b = nullptr;
    P.pme_order = 0;
}

// B-spline moduli of the plan's grid for `order`, computed once per plan
static BTab pme_tab(Plan &P, int order) {
    if (P.pme_b == nullptr || P.pme_order != order) {
        OFDFT_LEGACY_LOCK;
        if (P.pme_b) OFDFT_HIP(hipFree(P.pme_b));
        P.pme_b = nullptr;
        auto x = bspline_bm(P.nr[0], P.nr[0], order);
        auto y = bspline_bm(P.nr[1], P.nr[1], order);
        auto z = bspline_bm(P.nr[2], P.nzc, order);
        std::vector<double> all(x);
        all.insert(all.end(), y.begin(), y.end());
        all.insert(all.end(), z.begin(), z.end());
        OFDFT_HIP(hipMalloc(&P.pme_b, all.size() * sizeof(double)));
        OFDFT_HIP(hipMemcpy(P.pme_b, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
        P.pme_order = order;
    }
    const double2 *b = reinterpret_cast<const double2 *>(P.pme_b);
    return BTab{b, b + P.nr[0], b + P.nr[0] + P.nr[1]};
}

static int nblk(long long n) {
    long long b = (n + 255) / 256;
    return (int)std::max(1LL, std::min(b, 4096LL));
}

// the local PP on the grid: the real-space potential, or any radial recipe
// summed over atoms; per-point math in FP64, fields and transforms in the
// plan's precision T
template <typename T, bool SCREEN>
static void local_pp_t(Plan &P, int nspecies, const VlocTable &t, const ScreenParams &scr, int nat, const double *dpos,
                       const double *dfrac, const int *dspec, int pme, int order, T *vreal, hipStream_t s) {
    const GGeom g = geom_of(P);
    cpx<T> *spec = reinterpret_cast<cpx<T> *>(P.spec);  // slot 0: accumulated v(G), slot 1: species S(G)
    T *Q = reinterpret_cast<T *>(P.work);
    if (!pme) {
        // vreal = irfftn(v) / dV
        hipLaunchKernelGGL((k_vloc_exact<T, SCREEN>), dim3(nblk(P.Nc)), dim3(256), 0, s, spec, P.Nc, g, t, scr, nat,
                           dpos, dspec, 1.0);
        OFDFT_HIP(hipGetLastError());
    } else {
        const BTab bt = pme_tab(P, order);
        for (int sp = 0; sp < nspecies; ++sp) {
            OFDFT_HIP(hipMemsetAsync(Q, 0, sizeof(T) * P.N, s));
            hipLaunchKernelGGL(k_pme_spread<T>, dim3(nat), dim3(256), 0, s, Q, P.nr[0], P.nr[1], P.nr[2], order, nat,
                               dfrac, nullptr, dspec, sp);
            OFDFT_HIP(hipGetLastError());
            fft_r2c<T>(P, Q, spec + P.Nc, 1, 0, P.Nc, 1.0, nullptr, nullptr, s);
            // v = sum_s vloc_s * rfftn(Q_s) * dV * B * N / V = sum_s vloc_s rfftn(Q_s) B
            hipLaunchKernelGGL((k_vloc_pme_acc<T, SCREEN>), dim3(nblk(P.Nc)), dim3(256), 0, s, spec, spec + P.Nc, P.Nc,
                               g, t, scr, sp, bt, sp == 0 ? 1 : 0);
            OFDFT_HIP(hipGetLastError());
        }
    }
    fft_c2r<T>(P, spec, spec, vreal, 1, P.Nc, P.Nc, P.N, 1.0 / ((double)P.N * P.dV), s);
}

// screen (null: the potential itself): per species Z, the screened values at
// the first two table points, and K(q) = resp[0] + resp[1] q^2 + resp[2]
// G(q / resp[3]; lambda = resp[4]) + (resp[5] != 0) 4 pi / q^2
void local_pp(Plan &P, int nspecies, const double *q0, const double *dq, const int *ncoef,
              const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
              const double *pos_cart, const double *pos_frac, const int *species, int pme, int order, void *vreal,
              hipStream_t s, const double *screen_z, const double *screen_r1, const double *screen_r2,
              const double *resp) {
    OFDFT_REQUIRE(nspecies >= 1 && nspecies <= kMaxSpecies, "1..8 species supported");
    OFDFT_REQUIRE(order >= 3 && order <= kMaxOrder, "B-spline order must be 3..16");
    VlocTable t;
    std::memset(&t, 0, sizeof(t));
    StageLayout lay;
    size_t coff[kMaxSpecies];
    for (int sp = 0; sp < nspecies; ++sp) coff[sp] = lay.add(sizeof(double) * 4 * ncoef[sp]);
    const size_t opos = lay.add(sizeof(double) * 3 * nat), ofrac = lay.add(sizeof(double) * 3 * nat),
                 ospec = lay.add(sizeof(int) * nat);
    char *h = stage_host(P, lay.size);
    for (int sp = 0; sp < nspecies; ++sp) std::memcpy(h + coff[sp], coef[sp], sizeof(double) * 4 * ncoef[sp]);
    std::memcpy(h + opos, pos_cart, sizeof(double) * 3 * nat);
    std::memcpy(h + ofrac, pos_frac, sizeof(double) * 3 * nat);
    std::memcpy(h + ospec, species, sizeof(int) * nat);
    char *d = stage_upload(P, lay.size, s);
    for (int sp = 0; sp < nspecies; ++sp) {
        t.coef[sp] = reinterpret_cast<const double *>(d + coff[sp]);
        t.q0[sp] = q0[sp];
        t.dq[sp] = dq[sp];
        t.qmax[sp] = qmax[sp];
        t.ncoef[sp] = ncoef[sp];
        t.v0[sp] = vsmall[3 * sp];
        t.v1[sp] = vsmall[3 * sp + 1];
        t.v2[sp] = vsmall[3 * sp + 2];
        t.dqs[sp] = dqs[sp];
    }
    const double *dpos = reinterpret_cast<const double *>(d + opos);
    const double *dfrac = reinterpret_cast<const double *>(d + ofrac);
    const int *dspec = reinterpret_cast<const int *>(d + ospec);
    ScreenParams scr;
    std::memset(&scr, 0, sizeof(scr));
    const bool screen = screen_z != nullptr;
    if (screen) {
        OFDFT_REQUIRE(screen_r1 && screen_r2 && resp && resp[3] > 0.0, "screened superposition: bad response");
        scr.k = RespK{resp[0], resp[1], resp[2], resp[3], resp[4], resp[5] != 0.0 ? 1 : 0};
        for (int sp = 0; sp < nspecies; ++sp) {
            scr.z[sp] = screen_z[sp];
            scr.r1[sp] = screen_r1[sp];
            scr.r2[sp] = screen_r2[sp];
        }
    }
    if (P.prec == OFDFT_FP64) {
        if (screen)
            local_pp_t<double, true>(P, nspecies, t, scr, nat, dpos, dfrac, dspec, pme, order, (double *)vreal, s);
        else
            local_pp_t<double, false>(P, nspecies, t, scr, nat, dpos, dfrac, dspec, pme, order, (double *)vreal, s);
    } else {
        if (screen)
            local_pp_t<float, true>(P, nspecies, t, scr, nat, dpos, dfrac, dspec, pme, order, (float *)vreal, s);
        else
            local_pp_t<float, false>(P, nspecies, t, scr, nat, dpos, dfrac, dspec, pme, order, (float *)vreal, s);
    }
}

static double erfcinv_host(double y) {
    // Newton iterations on erfc(x) = y from a log-based start (y small)
    double x = std::sqrt(-std::log(y / 2.0));
    for (int it = 0; it < 100; ++it) {
        const double f = std::erfc(x) - y;
        const double df = -2.0 / std::sqrt(M_PI) * std::exp(-x * x);
        const double dx = f / df;
        x -= dx;
        if (std::fabs(dx) < 1e-15 * std::fabs(x)) break;
    }
    return x;
}

void ewald(Plan &P, int nat, const double *pos_cart, const double *pos_frac, const double *charges, int pme,
           int order, double precision, double *out, hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "ewald needs an FP64 plan");
    // gmax = sqrt(2 ecut), ecut from the grid spacing (grid.py:456-459)
    double lengths[3], sp_max = 0.0;
    for (int d = 0; d < 3; ++d) {
        lengths[d] = std::sqrt(P.lat[3 * d] * P.lat[3 * d] + P.lat[3 * d + 1] * P.lat[3 * d + 1] +
                               P.lat[3 * d + 2] * P.lat[3 * d + 2]);
        const double sp1 = lengths[d] / P.nr[d], sp2 = lengths[d] / (P.nr[d] - 1);
        sp_max = std::max(sp_max, 0.5 * (sp1 + sp2));
    }
    const double ecut = M_PI * M_PI / (2.0 * sp_max * sp_max);
    const double gmax = std::sqrt(2.0 * ecut);
    double q2 = 0.0, qsum = 0.0;
    for (int a = 0; a < nat; ++a) {
        q2 += charges[a] * charges[a];
        qsum += charges[a];
    }
    double eta = 1.6;
    while (true) {
        const double ub = 4.0 * M_PI * nat * q2 * std::sqrt(eta / M_PI) * std::erfc(gmax / 2.0 * std::sqrt(1.0 / eta));
        if (ub < precision) break;
        eta -= 0.01;
        OFDFT_REQUIRE(eta > 0.0, "ewald: no eta satisfies the precision");
    }
    const double e_corr = -std::sqrt(eta / M_PI) * q2 - 4.0 * M_PI * (1.0 / (4.0 * eta * P.volume) / 2.0) * qsum * qsum;
    // real space
    const double rmax = erfcinv_host(precision / 3.0) / std::sqrt(eta);
    int Nimg[3];
    for (int d = 0; d < 3; ++d) Nimg[d] = (int)std::ceil(rmax / lengths[d]);
    StageLayout lay;
    const size_t opos = lay.add(sizeof(double) * 3 * nat), ofrac = lay.add(sizeof(double) * 3 * nat),
                 oq = lay.add(sizeof(double) * nat);
    char *hs = stage_host(P, lay.size);
    std::memcpy(hs + opos, pos_cart, sizeof(double) * 3 * nat);
    std::memcpy(hs + ofrac, pos_frac, sizeof(double) * 3 * nat);
    std::memcpy(hs + oq, charges, sizeof(double) * nat);
    char *ds = stage_upload(P, lay.size, s);
    const double *dpos = reinterpret_cast<const double *>(ds + opos);
    const double *dfrac = reinterpret_cast<const double *>(ds + ofrac);
    const double *dq = reinterpret_cast<const double *>(ds + oq);
    double *res = P.scal + 8;
    const long long npair = (long long)nat * nat;
    const int nbr = nblk(npair);
    const double *L = P.lat;
    hipLaunchKernelGGL(k_ewald_real, dim3(nbr), dim3(256), 0, s, nat, dpos, dq, L[0], L[1],
                       L[2], L[3], L[4], L[5], L[6], L[7], L[8], Nimg[0], Nimg[1], Nimg[2], rmax, std::sqrt(eta),
                       P.red);
    OFDFT_HIP(hipGetLastError());
    reduce_final(P.red, nbr, 1, res, nullptr, s);
    // reciprocal space
    const GGeom g = geom_of(P);
    const int nbg = std::min(nblk(P.Nc), kRedBlocks);
    double2 *spec = reinterpret_cast<double2 *>(P.spec);
    if (pme) {
        const BTab bt = pme_tab(P, order);
        double *Q = reinterpret_cast<double *>(P.work);
        OFDFT_HIP(hipMemsetAsync(Q, 0, sizeof(double) * P.N, s));
        hipLaunchKernelGGL(k_pme_spread<double>, dim3(nat), dim3(256), 0, s, Q, P.nr[0], P.nr[1], P.nr[2], order,
                           nat, dfrac, dq, nullptr, -1);
        OFDFT_HIP(hipGetLastError());
        fft_r2c<double>(P, Q, spec, 1, 0, P.Nc, 1.0, nullptr, nullptr, s);
        hipLaunchKernelGGL(k_ewald_rec, dim3(nbg), dim3(256), 0, s, spec, P.Nc, g, eta, bt, nat, nullptr,
                           nullptr, P.red);
        OFDFT_HIP(hipGetLastError());
        const double c = 4.0 * M_PI / P.volume;
        reduce_final(P.red, nbg, 1, res + 1, &c, s);
    } else {
        BTab none{nullptr, nullptr, nullptr};
        hipLaunchKernelGGL(k_ewald_rec, dim3(nbg), dim3(256), 0, s, nullptr, P.Nc, g, eta, none, nat,
                           dpos, dq, P.red);
        OFDFT_HIP(hipGetLastError());
        const double c = 4.0 * M_PI / P.volume;
        reduce_final(P.red, nbg, 1, res + 1, &c, s);
    }
    // both sums in one copy to the plan's pinned buffer, waited for by polling
    OFDFT_HIP(hipMemcpyAsync(P.pinned, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    stream_wait(s);
    const double e_real = P.pinned[0] / 2.0, e_rec = P.pinned[1];
    out[0] = e_corr + e_real + e_rec;
    out[1] = e_corr;
    out[2] = e_real;
    out[3] = e_rec;
    out[4] = eta;
}


// ================================================================== forces and stress
// Local-PP and Ewald forces and stress with exact structure factors, and the
// stress of the TF, vW, WT, PZ-LDA and Hartree terms (ewald.py:541-760,
// functional/pseudo/__init__.py:414-470, kedf/tf.py:52-62, kedf/vw.py:97-111,
// kedf/wt.py:75-105, kedf/kernel.py:197-211, hartree.py:67-91,
// xc/semilocal_xc.py:308-322).  G-space sums run over the half-space mask
// with the reference's factors.  exp(-i G.R_a) is the product of per-axis
// tables exp(-2 pi i k f_{a,d}) (G.R = 2 pi k.f), so no sincos is evaluated
// per (atom, G) pair.

// phase tables, per atom [x: nx | y: ny | z: nzc], frequencies as gvec uses them
__global__ void k_phase_tables(int nat, const double *__restrict__ frac, int nx, int ny, int nz, int nzc, int full,
                               double2 *__restrict__ tab) {
    const int len = nx + ny + nzc;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)nat * len;
         e += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(e / len), k = (int)(e - (long long)a * len);
        int d, kk;
        if (k < nx) {
            d = 0;
            kk = ffreq_(k, nx);
        } else if (k < nx + ny) {
            d = 1;
            kk = ffreq_(k - nx, ny);
        } else {
            d = 2;
            kk = k - nx - ny;
            if (full) kk = wrapf_(kk, nz);
        }
        double sn, cs;
        sincospi(2.0 * kk * frac[3 * a + d], &sn, &cs);
        tab[e] = double2{cs, -sn};
    }
}

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

__device__ __forceinline__ void half_index(long long i, int ny, int nzc, int &ix, int &iy, int &iz) {
    iz = (int)(i % nzc);
    const long long rest = i / nzc;
    iy = (int)(rest % ny);
    ix = (int)(rest / ny);
}

// S(G) = sum_a w_a exp(-i G.R_a) on the half grid
__global__ void k_strf_w(long long nc, int nx, int ny, int nzc, int nat, const double2 *__restrict__ tab,
                         const double *__restrict__ w, double2 *__restrict__ S) {
    const int len = nx + ny + nzc;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, ny, nzc, ix, iy, iz);
        double re = 0.0, im = 0.0;
        for (int a = 0; a < nat; ++a) {
            const double wa = w[a];
            if (wa == 0.0) continue;
            const double2 *t = tab + (long long)a * len;
            const double2 ph = cmul2(cmul2(t[ix], t[nx + iy]), t[nx + ny + iz]);
            re += wa * ph.x;
            im += wa * ph.y;
        }
        S[i] = double2{re, im};
    }
}

// d vloc / dq of the cubic table (no small-q patch: the spline's own
// derivative, extrapolated below q0), zero beyond qmax
__device__ __forceinline__ double vloc_deriv(const VlocTable &t, int s, double q) {
    if (!(q < t.qmax[s])) return 0.0;
    int i = (int)floor((q - t.q0[s]) / t.dq[s]);
    if (i < 0) i = 0;
    if (i >= t.ncoef[s]) i = t.ncoef[s] - 1;
    const double x = q - (t.q0[s] + i * t.dq[s]);
    const double *c = t.coef[s] + 4 * (long long)i;
    return c[1] + x * (2.0 * c[2] + x * 3.0 * c[3]);
}

// X(G) = cp vloc_s(|G|) rho(G) + ce exp(-G^2/(4 eta)) / G^2 S_Z(G)  (G != 0)
__global__ void k_force_x(long long nc, GGeom g, VlocTable t, int s, double cp, const double2 *__restrict__ rhoG,
                          double ce, double eta, const double2 *__restrict__ SZ, double2 *__restrict__ X) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, g.ny, g.nzc, ix, iy, iz);
        int f[3], fp[3];
        freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        const double gg = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        double2 x{0.0, 0.0};
        if (rhoG) {
            const double v = cp * vloc_eval(t, s, sqrt(gg));
            x.x += v * rhoG[i].x;
            x.y += v * rhoG[i].y;
        }
        if (SZ && gg > 0.0) {
            const double w = ce * exp(-gg / (4.0 * eta)) / gg;
            x.x += w * SZ[i].x;
            x.y += w * SZ[i].y;
        }
        X[i] = x;
    }
}

// partials[(a * gridDim.x + blk) * 3 + d] = sum over masked G of G_d Im(exp(iG.R_a) X(G))
// for atom alist[blockIdx.y]
__global__ void k_force_sum(long long nc, GGeom g, const double2 *__restrict__ tab, const double2 *__restrict__ X,
                            const int *__restrict__ alist, double *__restrict__ partials) {
    __shared__ double sm[16];
    const int a = alist[blockIdx.y];
    const int len = g.nx + g.ny + g.nzc;
    const double2 *t = tab + (long long)a * len;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, g.ny, g.nzc, ix, iy, iz);
        if (!gmask(ix, iy, iz, g.nx, g.ny, g.nz, g.full)) continue;
        int f[3], fp[3];
        freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        const double2 ph = cmul2(cmul2(t[ix], t[g.nx + iy]), t[g.nx + g.ny + iz]);  // exp(-iG.R)
        const double2 x = X[i];
        const double im = ph.x * x.y - ph.y * x.x;  // Im(conj(ph) x)
        acc[0] += G[0] * im;
        acc[1] += G[1] * im;
        acc[2] += G[2] * im;
    }
    for (int d = 0; d < 3; ++d) {
        const double v = block_sum(acc[d], sm);
        if (threadIdx.x == 0) partials[((long long)blockIdx.y * gridDim.x + blockIdx.x) * 3 + d] = v;
    }
}

// eta dG/deta of the inverse Lindhard function (kernel.py:197-211)
__device__ __forceinline__ double lindhard_deriv(double eta, double mu) {
    const double atol = 1.0e-10;
    double d;
    if (eta < atol) {
        d = -2.0 * eta * (1.0 / 3.0 - 3.0 * mu);
    } else if (fabs(eta - 1.0) < atol) {
        d = -48.0;
    } else {
        const double A = log(fabs((1.0 + eta) / (1.0 - eta)));
        const double den = 0.5 + 0.25 * (1.0 - eta * eta) / eta * A;
        d = (0.5 / eta - 0.25 * (eta * eta + 1.0) / (eta * eta) * A) / (den * den) + 6.0 * eta * mu;
    }
    return d * eta;
}

struct StressGIn {
    const double2 *rhoG;  // Hartree and PP (null: neither)
    const double2 *rA, *rB;  // FFT(rho^alpha), FFT(rho^beta) (null: no WT)
    double itkf, mu;
    const double2 *Ss;  // nsp species structure factors, stride nc (null: no PP)
    int nsp;
    VlocTable t;
    const double2 *SZ;  // charge structure factor (null: no Ewald)
    double eta;
};

// per masked G != 0, the six g_i g_j sums of: Hartree |rho|^2/G^4,
// PP Re(rho v')/|G| (v' = sum_s dvloc_s/dq conj(S_s)), WT (g_i g_j/G^2 -
// delta/3) eta G'(eta) Re(rA conj(rB)), Ewald |S_Z|^2 exp(-G^2/4eta)/G^2
// (2 g_i g_j (1 + G^2/4eta)/G^2 - delta); partials [block][24]
__global__ void k_stress_g(long long nc, GGeom g, StressGIn in, double *__restrict__ partials) {
    __shared__ double sm[16];
    double acc[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) acc[k] = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, g.ny, g.nzc, ix, iy, iz);
        if (!gmask(ix, iy, iz, g.nx, g.ny, g.nz, g.full)) continue;
        if (ix == 0 && iy == 0 && iz == 0) continue;
        int f[3], fp[3];
        freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        const double gg = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        const double q = sqrt(gg), ig = 1.0 / gg;
        double h = 0.0, pp = 0.0, wt = 0.0, ew = 0.0, ewd = 0.0;
        if (in.rhoG) {
            const double2 r = in.rhoG[i];
            h = (r.x * r.x + r.y * r.y) * ig * ig;
            if (in.Ss) {
                double vx = 0.0, vy = 0.0;
                for (int s = 0; s < in.nsp; ++s) {
                    const double d = vloc_deriv(in.t, s, q);
                    const double2 S = in.Ss[(long long)s * nc + i];
                    vx += d * S.x;
                    vy -= d * S.y;
                }
                pp = (r.x * vx - r.y * vy) / q;
            }
        }
        if (in.rA) {
            const double2 a = in.rA[i], b = in.rB[i];
            wt = lindhard_deriv(q * in.itkf, in.mu) * (a.x * b.x + a.y * b.y);
        }
        if (in.SZ) {
            const double2 S = in.SZ[i];
            const double e = (S.x * S.x + S.y * S.y) * exp(-gg / (4.0 * in.eta)) * ig;
            ew = e * 2.0 * ig * (1.0 + gg / (4.0 * in.eta));
            ewd = e;
        }
        int k = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++k) {
                const double gab = G[a] * G[b];
                acc[k] += gab * h;
                acc[6 + k] += gab * pp;
                acc[12 + k] += (gab * ig - (a == b ? 1.0 / 3.0 : 0.0)) * wt;
                acc[18 + k] += gab * ew - (a == b ? ewd : 0.0);
            }
    }
    for (int k = 0; k < 24; ++k) {
        const double v = block_sum(acc[k], sm);
        if (threadIdx.x == 0) partials[(long long)blockIdx.x * 24 + k] = v;
    }
}

// spectra i G_d rho(G), d = 0..2 (dense, stride nc)
__global__ void k_grad_spectra(long long nc, GGeom g, const double2 *__restrict__ rhoG, double2 *__restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, g.ny, g.nzc, ix, iy, iz);
        int f[3], fp[3];
        freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        const double2 r = rhoG[i];
        for (int d = 0; d < 3; ++d) out[(long long)d * nc + i] = double2{-G[d] * r.y, G[d] * r.x};
    }
}

// [block][7]: six sums of d_i d_j / rho, and the sum of v_xc rho
__global__ void k_real_stress(long long n, const double *__restrict__ rho, const double *__restrict__ d,
                              double *__restrict__ partials) {
    __shared__ double sm[16];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double r = rho[i];
        const double g0 = d[i], g1 = d[n + i], g2 = d[2 * n + i];
        const double ir = 1.0 / r;
        acc[0] += g0 * g0 * ir;
        acc[1] += g0 * g1 * ir;
        acc[2] += g0 * g2 * ir;
        acc[3] += g1 * g1 * ir;
        acc[4] += g1 * g2 * ir;
        acc[5] += g2 * g2 * ir;
        double e, v;
        pz_lda<double>(r < 1e-30 ? 1e-10 : cbrt(r), e, v);
        acc[6] += v * r;
    }
    for (int k = 0; k < 7; ++k) {
        const double v = block_sum(acc[k], sm);
        if (threadIdx.x == 0) partials[(long long)blockIdx.x * 7 + k] = v;
    }
}

// Ewald real-space forces and stress, one workgroup per atom i over the
// atoms j and the lattice images within rmax: partials[i][0..2] force,
// [3..8] the six sums of q_i q_j (2 sqrt(eta/pi) exp(-eta d^2) + erfc(sqrt(eta) d)/d) R_a R_b / d^2
__global__ void k_ewald_real_fs(int nat, const double *__restrict__ pos, const double *__restrict__ q, GGeom lat,
                                int N0, int N1, int N2, double rmax, double eta, double *__restrict__ partials) {
    __shared__ double sm[16];
    const int i = blockIdx.x;
    const double se = sqrt(eta);
    const int n0 = 2 * N0 + 1, n1 = 2 * N1 + 1, n2 = 2 * N2 + 1;
    const long long nimg = (long long)n0 * n1 * n2;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (long long e = threadIdx.x; e < nimg * nat; e += blockDim.x) {
        const int j = (int)(e % nat);
        const long long im = e / nat;
        const int c = (int)(im % n2) - N2, b = (int)((im / n2) % n1) - N1, a = (int)(im / ((long long)n1 * n2)) - N0;
        // R = r_i - (r_j - T), T = a L0 + b L1 + c L2 (lat.B holds the lattice rows here)
        double R[3];
        for (int d = 0; d < 3; ++d)
            R[d] = pos[3 * i + d] - pos[3 * j + d] + a * lat.B[d] + b * lat.B[3 + d] + c * lat.B[6 + d];
        const double dd = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        if (!(dd < rmax && dd > 0.001)) continue;
        const double x = se * dd;
        const double ex = exp(-x * x), ec = erfc(x);
        const double qq = q[i] * q[j];
        const double ff = qq * (ec / (x * x * x) + 2.0 / sqrt(M_PI) * ex / (x * x)) * se * se * se;
        acc[0] += ff * R[0];
        acc[1] += ff * R[1];
        acc[2] += ff * R[2];
        const double ss = qq * (2.0 * se / sqrt(M_PI) * ex + ec / dd) / (dd * dd);
        int k = 3;
        for (int u = 0; u < 3; ++u)
            for (int v = u; v < 3; ++v, ++k) acc[k] += ss * R[u] * R[v];
    }
    for (int k = 0; k < 9; ++k) {
        const double v = block_sum(acc[k], sm);
        if (threadIdx.x == 0) partials[(long long)i * 9 + k] = v;
    }
}

// ------------------------------------------------------------------ host
static double *fs_workspace(Plan &P, size_t bytes) {
    if (bytes > P.fs_cap) {
        OFDFT_LEGACY_LOCK;
        if (P.fs_ws) OFDFT_HIP(hipFree(P.fs_ws));
        P.fs_ws = nullptr;
        P.fs_cap = 0;
        OFDFT_HIP(hipMalloc(&P.fs_ws, bytes));
        P.fs_cap = bytes;
    }
    return reinterpret_cast<double *>(P.fs_ws);
}

// the host inputs of one force / stress call, staged in one upload
struct IonInputs {
    VlocTable t;
    const double *pos = nullptr, *frac = nullptr, *w = nullptr;  // w: nw weight vectors of nat
    const int *spec = nullptr, *alist = nullptr;                // alist: atom indices grouped by species
};

static IonInputs stage_ions(Plan &P, int nspecies, const double *q0, const double *dq, const int *ncoef,
                            const double *const *coef, const double *qmax, const double *vsmall, const double *dqs,
                            int nat, const double *pos_cart, const double *pos_frac, const int *species,
                            const std::vector<double> &weights, const std::vector<int> &alist, hipStream_t s) {
    StageLayout lay;
    size_t coff[kMaxSpecies] = {};
    for (int sp = 0; sp < nspecies; ++sp) coff[sp] = lay.add(sizeof(double) * 4 * ncoef[sp]);
    const size_t opos = lay.add(sizeof(double) * 3 * nat), ofrac = lay.add(sizeof(double) * 3 * nat),
                 ospec = lay.add(sizeof(int) * nat), ow = lay.add(sizeof(double) * weights.size()),
                 oal = lay.add(sizeof(int) * alist.size());
    char *h = stage_host(P, lay.size);
    for (int sp = 0; sp < nspecies; ++sp) std::memcpy(h + coff[sp], coef[sp], sizeof(double) * 4 * ncoef[sp]);
    std::memcpy(h + opos, pos_cart, sizeof(double) * 3 * nat);
    std::memcpy(h + ofrac, pos_frac, sizeof(double) * 3 * nat);
    if (species) std::memcpy(h + ospec, species, sizeof(int) * nat);
    if (!weights.empty()) std::memcpy(h + ow, weights.data(), sizeof(double) * weights.size());
    if (!alist.empty()) std::memcpy(h + oal, alist.data(), sizeof(int) * alist.size());
    char *d = stage_upload(P, lay.size, s);
    IonInputs in;
    std::memset(&in.t, 0, sizeof(in.t));
    for (int sp = 0; sp < nspecies; ++sp) {
        in.t.coef[sp] = reinterpret_cast<const double *>(d + coff[sp]);
        in.t.q0[sp] = q0[sp];
        in.t.dq[sp] = dq[sp];
        in.t.qmax[sp] = qmax[sp];
        in.t.ncoef[sp] = ncoef[sp];
        in.t.v0[sp] = vsmall[3 * sp];
        in.t.v1[sp] = vsmall[3 * sp + 1];
        in.t.v2[sp] = vsmall[3 * sp + 2];
        in.t.dqs[sp] = dqs[sp];
    }
    in.pos = reinterpret_cast<const double *>(d + opos);
    in.frac = reinterpret_cast<const double *>(d + ofrac);
    in.spec = reinterpret_cast<const int *>(d + ospec);
    in.w = reinterpret_cast<const double *>(d + ow);
    in.alist = reinterpret_cast<const int *>(d + oal);
    return in;
}

static void phase_tables(Plan &P, int nat, const double *frac, double2 *tab, hipStream_t s) {
    const long long n = (long long)nat * (P.nr[0] + P.nr[1] + P.nzc);
    hipLaunchKernelGGL(k_phase_tables, dim3(nblk(n)), dim3(256), 0, s, nat, frac, P.nr[0], P.nr[1], P.nr[2], P.nzc,
                       P.full, tab);
    OFDFT_HIP(hipGetLastError());
}

// Ewald parameters of the reference (ewald.py:256-279 eta, 280-320 cut-off)
struct EwaldParams {
    double eta, rmax, gmax;
    int nimg[3];
};

static EwaldParams ewald_params(const Plan &P, int nat, const double *charges, double precision) {
    double lengths[3], sp_max = 0.0;
    for (int d = 0; d < 3; ++d) {
        lengths[d] = std::sqrt(P.lat[3 * d] * P.lat[3 * d] + P.lat[3 * d + 1] * P.lat[3 * d + 1] +
                               P.lat[3 * d + 2] * P.lat[3 * d + 2]);
        const double sp1 = lengths[d] / P.nr[d], sp2 = lengths[d] / (P.nr[d] - 1);
        sp_max = std::max(sp_max, 0.5 * (sp1 + sp2));
    }
    EwaldParams e;
    e.gmax = std::sqrt(2.0 * (M_PI * M_PI / (2.0 * sp_max * sp_max)));
    double q2 = 0.0;
    for (int a = 0; a < nat; ++a) q2 += charges[a] * charges[a];
    e.eta = 1.6;
    while (true) {
        const double ub =
            4.0 * M_PI * nat * q2 * std::sqrt(e.eta / M_PI) * std::erfc(e.gmax / 2.0 * std::sqrt(1.0 / e.eta));
        if (ub < precision) break;
        e.eta -= 0.01;
        OFDFT_REQUIRE(e.eta > 0.0, "ewald: no eta satisfies the precision");
    }
    e.rmax = erfcinv_host(precision / 3.0) / std::sqrt(e.eta);
    for (int d = 0; d < 3; ++d) e.nimg[d] = (int)std::ceil(e.rmax / lengths[d]);
    return e;
}

// sum the force partials [na][nb][3] of k_force_sum into F[alist[k]]
static void reduce_forces(const double *dpart, int na, int nb, const std::vector<int> &alist, double scale, double *F,
                          hipStream_t s) {
    std::vector<double> h((size_t)na * nb * 3);
    OFDFT_HIP(hipMemcpyAsync(h.data(), dpart, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OFDFT_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < na; ++k)
        for (int d = 0; d < 3; ++d) {
            double acc = 0.0;
            for (int b = 0; b < nb; ++b) acc += h[((size_t)k * nb + b) * 3 + d];
            F[3 * alist[k] + d] += scale * acc;
        }
}

static std::vector<double> reduce_cols(const double *dpart, int nb, int K, hipStream_t s) {
    std::vector<double> h((size_t)nb * K), out(K, 0.0);
    OFDFT_HIP(hipMemcpyAsync(h.data(), dpart, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OFDFT_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < K; ++k) out[k] += h[(size_t)b * K + k];
    return out;
}

static void sym_add(double *S, const double *v6, double scale) {
    int k = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b, ++k) {
            S[3 * a + b] += scale * v6[k];
            if (a != b) S[3 * b + a] += scale * v6[k];
        }
}

constexpr int kForceChunks = 64;

// F (host, nat x 3) += local-PP forces (pseudo/__init__.py:456-470)
void pp_forces(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
               const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
               const double *pos_cart, const double *pos_frac, const int *species, double *F, hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "forces need an FP64 plan");
    std::vector<int> alist;
    std::vector<int> cnt(nspecies, 0);
    for (int sp = 0; sp < nspecies; ++sp)
        for (int a = 0; a < nat; ++a)
            if (species[a] == sp) {
                alist.push_back(a);
                ++cnt[sp];
            }
    IonInputs in = stage_ions(P, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac, species,
                              {}, alist, s);
    const long long nc = P.Nc;
    const int len = P.nr[0] + P.nr[1] + P.nzc;
    double *ws = fs_workspace(P, sizeof(double) * (2 * nc * 2 + (size_t)nat * len * 2 + (size_t)nat * kForceChunks * 3));
    double2 *rhoG = reinterpret_cast<double2 *>(ws);
    double2 *X = rhoG + nc;
    double2 *tab = X + nc;
    double *part = reinterpret_cast<double *>(tab + (size_t)nat * len);
    fft_r2c<double>(P, rho, rhoG, 1, P.N, nc, P.dV, nullptr, nullptr, s);
    phase_tables(P, nat, in.frac, tab, s);
    const GGeom g = geom_of(P);
    int off = 0;
    for (int sp = 0; sp < nspecies; ++sp) {
        if (cnt[sp] == 0) continue;
        hipLaunchKernelGGL(k_force_x, dim3(nblk(nc)), dim3(256), 0, s, nc, g, in.t, sp, 2.0 / P.volume, rhoG, 0.0,
                           1.0, (const double2 *)nullptr, X);
        OFDFT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_force_sum, dim3(kForceChunks, cnt[sp]), dim3(256), 0, s, nc, g, tab, X,
                           in.alist + off, part + (size_t)off * kForceChunks * 3);
        OFDFT_HIP(hipGetLastError());
        off += cnt[sp];
    }
    std::vector<int> order(alist.begin(), alist.end());
    reduce_forces(part, nat, kForceChunks, order, 1.0, F, s);
}

// S (host, 3 x 3) = local-PP stress (pseudo/__init__.py:414-454)
void pp_stress(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
               const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
               const double *pos_cart, const double *pos_frac, const int *species, double energy, double *S,
               hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "stress needs an FP64 plan");
    std::vector<double> w((size_t)nspecies * nat, 0.0);
    for (int a = 0; a < nat; ++a) w[(size_t)species[a] * nat + a] = 1.0;
    IonInputs in = stage_ions(P, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac, species, w,
                              {}, s);
    const long long nc = P.Nc;
    const int len = P.nr[0] + P.nr[1] + P.nzc;
    const int nb = 1024;
    double *ws = fs_workspace(P, sizeof(double) * ((1 + nspecies) * nc * 2 + (size_t)nat * len * 2 + nb * 24));
    double2 *rhoG = reinterpret_cast<double2 *>(ws);
    double2 *Ss = rhoG + nc;
    double2 *tab = Ss + (size_t)nspecies * nc;
    double *part = reinterpret_cast<double *>(tab + (size_t)nat * len);
    fft_r2c<double>(P, rho, rhoG, 1, P.N, nc, P.dV, nullptr, nullptr, s);
    phase_tables(P, nat, in.frac, tab, s);
    for (int sp = 0; sp < nspecies; ++sp) {
        hipLaunchKernelGGL(k_strf_w, dim3(nblk(nc)), dim3(256), 0, s, nc, P.nr[0], P.nr[1], P.nzc, nat, tab,
                           in.w + (size_t)sp * nat, Ss + (size_t)sp * nc);
        OFDFT_HIP(hipGetLastError());
    }
    StressGIn si;
    std::memset(&si, 0, sizeof(si));
    si.rhoG = rhoG;
    si.Ss = Ss;
    si.nsp = nspecies;
    si.t = in.t;
    hipLaunchKernelGGL(k_stress_g, dim3(nb), dim3(256), 0, s, nc, geom_of(P), si, part);
    OFDFT_HIP(hipGetLastError());
    const std::vector<double> v = reduce_cols(part, nb, 24, s);
    for (int k = 0; k < 9; ++k) S[k] = 0.0;
    sym_add(S, v.data() + 6, -2.0 / (P.volume * P.volume));
    for (int d = 0; d < 3; ++d) S[4 * d] -= energy / P.volume;
}

// Ewald forces (F, nat x 3, optional) and stress (S, 3 x 3, optional), real
// space over images plus exact reciprocal sums (ewald.py:541-760)
void ewald_fs(Plan &P, int nat, const double *pos_cart, const double *pos_frac, const double *charges,
              double precision, double *F, double *S, hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "Ewald forces need an FP64 plan");
    const EwaldParams ep = ewald_params(P, nat, charges, precision);
    std::vector<double> w(charges, charges + nat);
    std::vector<int> alist(nat);
    for (int a = 0; a < nat; ++a) alist[a] = a;
    IonInputs in = stage_ions(P, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nat, pos_cart,
                              pos_frac, nullptr, w, alist, s);
    const long long nc = P.Nc;
    const int len = P.nr[0] + P.nr[1] + P.nzc;
    const int nb = 1024;
    double *ws = fs_workspace(P, sizeof(double) * (2 * nc * 2 + (size_t)nat * len * 2 +
                                                   std::max((size_t)nat * kForceChunks * 3, (size_t)nb * 24) +
                                                   (size_t)nat * 9));
    double2 *SZ = reinterpret_cast<double2 *>(ws);
    double2 *X = SZ + nc;
    double2 *tab = X + nc;
    double *part = reinterpret_cast<double *>(tab + (size_t)nat * len);
    double *rpart = part + std::max((size_t)nat * kForceChunks * 3, (size_t)nb * 24);
    phase_tables(P, nat, in.frac, tab, s);
    hipLaunchKernelGGL(k_strf_w, dim3(nblk(nc)), dim3(256), 0, s, nc, P.nr[0], P.nr[1], P.nzc, nat, tab, in.w, SZ);
    OFDFT_HIP(hipGetLastError());
    const GGeom g = geom_of(P);
    // real space
    GGeom lat = g;
    for (int k = 0; k < 9; ++k) lat.B[k] = P.lat[k];
    hipLaunchKernelGGL(k_ewald_real_fs, dim3(nat), dim3(256), 0, s, nat, in.pos, in.w, lat, ep.nimg[0], ep.nimg[1],
                       ep.nimg[2], ep.rmax, ep.eta, rpart);
    OFDFT_HIP(hipGetLastError());
    std::vector<double> hr((size_t)nat * 9);
    OFDFT_HIP(hipMemcpyAsync(hr.data(), rpart, hr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OFDFT_HIP(hipStreamSynchronize(s));
    if (F) {
        for (int a = 0; a < nat; ++a)
            for (int d = 0; d < 3; ++d) F[3 * a + d] += hr[(size_t)a * 9 + d];
        // reciprocal: F_a = (8 pi / V) Z_a sum_mask G exp(-G^2/4eta)/G^2 Im(exp(iG.R_a) S_Z(G))
        hipLaunchKernelGGL(k_force_x, dim3(nblk(nc)), dim3(256), 0, s, nc, g, VlocTable{}, 0, 0.0,
                           (const double2 *)nullptr, 8.0 * M_PI / P.volume, ep.eta, SZ, X);
        OFDFT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_force_sum, dim3(kForceChunks, nat), dim3(256), 0, s, nc, g, tab, X, in.alist, part);
        OFDFT_HIP(hipGetLastError());
        std::vector<double> Fr((size_t)nat * 3, 0.0);
        reduce_forces(part, nat, kForceChunks, alist, 1.0, Fr.data(), s);
        for (int a = 0; a < nat; ++a)
            for (int d = 0; d < 3; ++d) F[3 * a + d] += charges[a] * Fr[3 * a + d];
    }
    if (S) {
        for (int k = 0; k < 9; ++k) S[k] = 0.0;
        std::vector<double> r6(6, 0.0);
        for (int a = 0; a < nat; ++a)
            for (int k = 0; k < 6; ++k) r6[k] += hr[(size_t)a * 9 + 3 + k];
        sym_add(S, r6.data(), -0.5 / P.volume);
        StressGIn si;
        std::memset(&si, 0, sizeof(si));
        si.SZ = SZ;
        si.eta = ep.eta;
        hipLaunchKernelGGL(k_stress_g, dim3(nb), dim3(256), 0, s, nc, g, si, part);
        OFDFT_HIP(hipGetLastError());
        const std::vector<double> v = reduce_cols(part, nb, 24, s);
        sym_add(S, v.data() + 18, 4.0 * M_PI / (P.volume * P.volume));
        double qsum = 0.0;
        for (int a = 0; a < nat; ++a) qsum += charges[a];
        for (int d = 0; d < 3; ++d) S[4 * d] += qsum * qsum * 4.0 * M_PI * (1.0 / (4.0 * ep.eta * P.volume * P.volume) / 2.0);
    }
}

// ------------------------------------------------------------------ PME forms
// Forces and stress with PME structure factors (pseudo/__init__.py:475-563,
// ewald.py:810-864,915-960).  The stress is the exact form with each
// structure factor replaced by B(G) rfftn(Q)(G); the forces contract the
// derivative of each atom's B-spline stencil with a smoothed field.

// S *= B (dense half grid)
__global__ void k_bmul(double2 *__restrict__ S, long long nc, GGeom g, BTab b) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, g.ny, g.nzc, ix, iy, iz);
        const double2 bb = cmul2(cmul2(b.bx[ix], b.by[iy]), b.bz[iz]);
        S[i] = cmul2(S[i], bb);
    }
}

// X = conj(B) vloc_s(|G|) rhoG (mode 0, local PP) or |B|^2 exp(-G^2/4eta)/G^2 SQ
// (mode 1, Ewald); X(0) = 0
__global__ void k_pme_x(long long nc, GGeom g, VlocTable t, int s, const double2 *__restrict__ in, double eta, BTab b,
                        int mode, double2 *__restrict__ X) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        int ix, iy, iz;
        half_index(i, g.ny, g.nzc, ix, iy, iz);
        if (ix == 0 && iy == 0 && iz == 0) {
            X[i] = double2{0.0, 0.0};
            continue;
        }
        int f[3], fp[3];
        freqs_of(g, ix, iy, iz, f, fp);
        double G[3];
        gvec(g, f[0], f[1], f[2], G);
        const double gg = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        const double2 bb = cmul2(cmul2(b.bx[ix], b.by[iy]), b.bz[iz]);
        const double2 v = in[i];
        if (mode == 0) {
            const double w = vloc_eval(t, s, sqrt(gg));
            const double2 cb{bb.x, -bb.y};
            const double2 r = cmul2(cb, v);
            X[i] = double2{w * r.x, w * r.y};
        } else {
            const double w = (bb.x * bb.x + bb.y * bb.y) * exp(-gg / (4.0 * eta)) / gg;
            X[i] = double2{w * v.x, w * v.y};
        }
    }
}

// F_a = -scale_a (sum_l grad_frac Q_a(l) field(l)) . Ainv (the reference's
// matmul(Q_derivative^T, cell_inv)); one workgroup per atom alist[blockIdx.x]
__global__ void k_pme_force(const double *__restrict__ field, int nx, int ny, int nz, int order,
                            const double *__restrict__ frac, const int *__restrict__ alist,
                            const double *__restrict__ scale, GGeom ainv, double *__restrict__ F) {
    __shared__ double mv[3][kMaxOrder + 1], dv[3][kMaxOrder + 1];
    __shared__ int fl[3];
    __shared__ double sm[16];
    const int a = alist[blockIdx.x];
    const int nr[3] = {nx, ny, nz};
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        const double u = frac[3 * a + d] * nr[d];
        const double f = floor(u);
        fl[d] = (int)f;
        double m[kMaxOrder + 1], m2[kMaxOrder + 1];
        calc_mn(u - f, order, m);
        calc_mn(u - f, order - 1, m2);
        m2[order] = 0.0;
        for (int k = 0; k < order; ++k) {
            mv[d][k] = m[k + 1];
            dv[d][k] = nr[d] * (m2[k + 1] - m2[k]);
        }
    }
    __syncthreads();
    double acc[3] = {0.0, 0.0, 0.0};
    const int o3 = order * order * order;
    for (int e = threadIdx.x; e < o3; e += blockDim.x) {
        const int i = e / (order * order), j = (e / order) % order, k = e % order;
        int lx = (1 + fl[0] - i) % nx;
        if (lx < 0) lx += nx;
        int ly = (1 + fl[1] - j) % ny;
        if (ly < 0) ly += ny;
        int lz = (1 + fl[2] - k) % nz;
        if (lz < 0) lz += nz;
        const double v = field[((long long)lx * ny + ly) * nz + lz];
        acc[0] += dv[0][i] * mv[1][j] * mv[2][k] * v;
        acc[1] += mv[0][i] * dv[1][j] * mv[2][k] * v;
        acc[2] += mv[0][i] * mv[1][j] * dv[2][k] * v;
    }
    double t[3];
    for (int d = 0; d < 3; ++d) t[d] = block_sum(acc[d], sm);
    if (threadIdx.x == 0) {
        const double sc = scale ? scale[a] : 1.0;
        for (int k = 0; k < 3; ++k)
            F[3 * a + k] = -sc * (t[0] * ainv.B[k] + t[1] * ainv.B[3 + k] + t[2] * ainv.B[6 + k]);
    }
}

// inverse of the lattice matrix (rows = lattice vectors), row-major in B[9]
static GGeom inverse_lattice(const Plan &P) {
    const double *A = P.lat;
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                       A[2] * (A[3] * A[7] - A[4] * A[6]);
    GGeom g = geom_of(P);
    g.B[0] = (A[4] * A[8] - A[5] * A[7]) / det;
    g.B[1] = (A[2] * A[7] - A[1] * A[8]) / det;
    g.B[2] = (A[1] * A[5] - A[2] * A[4]) / det;
    g.B[3] = (A[5] * A[6] - A[3] * A[8]) / det;
    g.B[4] = (A[0] * A[8] - A[2] * A[6]) / det;
    g.B[5] = (A[2] * A[3] - A[0] * A[5]) / det;
    g.B[6] = (A[3] * A[7] - A[4] * A[6]) / det;
    g.B[7] = (A[1] * A[6] - A[0] * A[7]) / det;
    g.B[8] = (A[0] * A[4] - A[1] * A[3]) / det;
    return g;
}

// field = irfftn(X) into `field` (N doubles), X on the dense half grid
static void pme_field(Plan &P, double2 *X, double *field, hipStream_t s) {
    fft_c2r<double>(P, X, X, field, 1, P.Nc, P.Nc, P.N, 1.0 / (double)P.N, s);
}

// F (host, nat x 3) += local-PP forces, PME form (pseudo/__init__.py:475-520)
void pp_forces_pme(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
                   const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
                   const double *pos_cart, const double *pos_frac, const int *species, int order, double *F,
                   hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "forces need an FP64 plan");
    OFDFT_REQUIRE(order >= 3 && order <= kMaxOrder, "B-spline order must be 3..16");
    std::vector<int> alist;
    std::vector<int> cnt(nspecies, 0);
    for (int sp = 0; sp < nspecies; ++sp)
        for (int a = 0; a < nat; ++a)
            if (species[a] == sp) {
                alist.push_back(a);
                ++cnt[sp];
            }
    IonInputs in = stage_ions(P, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac, species,
                              {}, alist, s);
    const long long nc = P.Nc;
    double *ws = fs_workspace(P, sizeof(double) * (2 * nc * 2 + P.N + (size_t)3 * nat));
    double2 *rhoG = reinterpret_cast<double2 *>(ws);
    double2 *X = rhoG + nc;
    double *field = reinterpret_cast<double *>(X + nc);
    double *Fd = field + P.N;
    const BTab bt = pme_tab(P, order);
    const GGeom g = geom_of(P);
    const GGeom ai = inverse_lattice(P);
    fft_r2c<double>(P, rho, rhoG, 1, P.N, nc, 1.0, nullptr, nullptr, s);
    int off = 0;
    for (int sp = 0; sp < nspecies; ++sp) {
        if (cnt[sp] == 0) continue;
        hipLaunchKernelGGL(k_pme_x, dim3(nblk(nc)), dim3(256), 0, s, nc, g, in.t, sp, (const double2 *)rhoG, 0.0, bt,
                           0, X);
        OFDFT_HIP(hipGetLastError());
        pme_field(P, X, field, s);
        hipLaunchKernelGGL(k_pme_force, dim3(cnt[sp]), dim3(256), 0, s, (const double *)field, P.nr[0], P.nr[1],
                           P.nr[2], order, in.frac, in.alist + off, (const double *)nullptr, ai, Fd);
        OFDFT_HIP(hipGetLastError());
        off += cnt[sp];
    }
    std::vector<double> h((size_t)3 * nat);
    OFDFT_HIP(hipMemcpyAsync(h.data(), Fd, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OFDFT_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 3 * nat; ++k) F[k] += h[k];
}

// S (host, 3 x 3) = local-PP stress, PME form (pseudo/__init__.py:522-563)
void pp_stress_pme(Plan &P, const double *rho, int nspecies, const double *q0, const double *dq, const int *ncoef,
                   const double *const *coef, const double *qmax, const double *vsmall, const double *dqs, int nat,
                   const double *pos_cart, const double *pos_frac, const int *species, int order, double energy,
                   double *S, hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "stress needs an FP64 plan");
    OFDFT_REQUIRE(order >= 3 && order <= kMaxOrder, "B-spline order must be 3..16");
    IonInputs in = stage_ions(P, nspecies, q0, dq, ncoef, coef, qmax, vsmall, dqs, nat, pos_cart, pos_frac, species,
                              {}, {}, s);
    const long long nc = P.Nc;
    const int nb = 1024;
    double *ws = fs_workspace(P, sizeof(double) * ((1 + nspecies) * nc * 2 + P.N + nb * 24));
    double2 *rhoG = reinterpret_cast<double2 *>(ws);
    double2 *Ss = rhoG + nc;
    double *Q = reinterpret_cast<double *>(Ss + (size_t)nspecies * nc);
    double *part = Q + P.N;
    const BTab bt = pme_tab(P, order);
    const GGeom g = geom_of(P);
    fft_r2c<double>(P, rho, rhoG, 1, P.N, nc, P.dV, nullptr, nullptr, s);
    for (int sp = 0; sp < nspecies; ++sp) {
        hipLaunchKernelGGL(k_fill_zero, dim3(nblk(P.N)), dim3(256), 0, s, Q, P.N);
        hipLaunchKernelGGL(k_pme_spread<double>, dim3(nat), dim3(256), 0, s, Q, P.nr[0], P.nr[1], P.nr[2], order,
                           nat, in.frac, (const double *)nullptr, in.spec, sp);
        OFDFT_HIP(hipGetLastError());
        fft_r2c<double>(P, Q, Ss + (size_t)sp * nc, 1, P.N, nc, 1.0, nullptr, nullptr, s);
        hipLaunchKernelGGL(k_bmul, dim3(nblk(nc)), dim3(256), 0, s, Ss + (size_t)sp * nc, nc, g, bt);
        OFDFT_HIP(hipGetLastError());
    }
    StressGIn si;
    std::memset(&si, 0, sizeof(si));
    si.rhoG = rhoG;
    si.Ss = Ss;
    si.nsp = nspecies;
    si.t = in.t;
    hipLaunchKernelGGL(k_stress_g, dim3(nb), dim3(256), 0, s, nc, g, si, part);
    OFDFT_HIP(hipGetLastError());
    const std::vector<double> v = reduce_cols(part, nb, 24, s);
    for (int k = 0; k < 9; ++k) S[k] = 0.0;
    sym_add(S, v.data() + 6, -2.0 / (P.volume * P.volume));
    for (int d = 0; d < 3; ++d) S[4 * d] -= energy / P.volume;
}

// Ewald forces and stress with the PME reciprocal forms (ewald.py:810-864,
// 915-960) and the real-space sums of ewald_fs
void ewald_fs_pme(Plan &P, int nat, const double *pos_cart, const double *pos_frac, const double *charges,
                  double precision, int order, double *F, double *S, hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "Ewald forces need an FP64 plan");
    OFDFT_REQUIRE(order >= 3 && order <= kMaxOrder, "B-spline order must be 3..16");
    const EwaldParams ep = ewald_params(P, nat, charges, precision);
    std::vector<double> w(charges, charges + nat);
    std::vector<int> alist(nat);
    for (int a = 0; a < nat; ++a) alist[a] = a;
    IonInputs in = stage_ions(P, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nat, pos_cart,
                              pos_frac, nullptr, w, alist, s);
    const long long nc = P.Nc;
    const int nb = 1024;
    double *ws = fs_workspace(P, sizeof(double) * (2 * nc * 2 + P.N + (size_t)3 * nat + (size_t)nb * 24 +
                                                   (size_t)nat * 9 + nat));
    double2 *SZ = reinterpret_cast<double2 *>(ws);
    double2 *X = SZ + nc;
    double *Q = reinterpret_cast<double *>(X + nc);
    double *Fd = Q + P.N;
    double *part = Fd + 3 * nat;
    double *rpart = part + (size_t)nb * 24;
    double *qs = rpart + (size_t)nat * 9;
    const BTab bt = pme_tab(P, order);
    const GGeom g = geom_of(P);
    hipLaunchKernelGGL(k_fill_zero, dim3(nblk(P.N)), dim3(256), 0, s, Q, P.N);
    hipLaunchKernelGGL(k_pme_spread<double>, dim3(nat), dim3(256), 0, s, Q, P.nr[0], P.nr[1], P.nr[2], order, nat,
                       in.frac, in.w, (const int *)nullptr, -1);
    OFDFT_HIP(hipGetLastError());
    fft_r2c<double>(P, Q, SZ, 1, P.N, nc, 1.0, nullptr, nullptr, s);  // rfftn(Q_Z)
    // real space, as ewald_fs
    GGeom lat = g;
    for (int k = 0; k < 9; ++k) lat.B[k] = P.lat[k];
    hipLaunchKernelGGL(k_ewald_real_fs, dim3(nat), dim3(256), 0, s, nat, in.pos, in.w, lat, ep.nimg[0], ep.nimg[1],
                       ep.nimg[2], ep.rmax, ep.eta, rpart);
    OFDFT_HIP(hipGetLastError());
    std::vector<double> hr((size_t)nat * 9);
    OFDFT_HIP(hipMemcpyAsync(hr.data(), rpart, hr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OFDFT_HIP(hipStreamSynchronize(s));
    if (F) {
        for (int a = 0; a < nat; ++a)
            for (int d = 0; d < 3; ++d) F[3 * a + d] += hr[(size_t)a * 9 + d];
        // F_a = -q_a (4 pi / dV) stencil(irfftn(rfftn(Q) |B|^2 exp(-G^2/4eta)/G^2))
        std::vector<double> sc(nat);
        for (int a = 0; a < nat; ++a) sc[a] = charges[a] * 4.0 * M_PI / P.dV;
        OFDFT_HIP(hipMemcpyAsync(qs, sc.data(), sizeof(double) * nat, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_pme_x, dim3(nblk(nc)), dim3(256), 0, s, nc, g, VlocTable{}, 0, (const double2 *)SZ, ep.eta,
                           bt, 1, X);
        OFDFT_HIP(hipGetLastError());
        pme_field(P, X, Q, s);
        hipLaunchKernelGGL(k_pme_force, dim3(nat), dim3(256), 0, s, (const double *)Q, P.nr[0], P.nr[1], P.nr[2],
                           order, in.frac, in.alist, (const double *)qs, inverse_lattice(P), Fd);
        OFDFT_HIP(hipGetLastError());
        std::vector<double> h((size_t)3 * nat);
        OFDFT_HIP(hipMemcpyAsync(h.data(), Fd, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        OFDFT_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < 3 * nat; ++k) F[k] += h[k];
    }
    if (S) {
        for (int k = 0; k < 9; ++k) S[k] = 0.0;
        std::vector<double> r6(6, 0.0);
        for (int a = 0; a < nat; ++a)
            for (int k = 0; k < 6; ++k) r6[k] += hr[(size_t)a * 9 + 3 + k];
        sym_add(S, r6.data(), -0.5 / P.volume);
        hipLaunchKernelGGL(k_bmul, dim3(nblk(nc)), dim3(256), 0, s, SZ, nc, g, bt);  // B rfftn(Q_Z)
        OFDFT_HIP(hipGetLastError());
        StressGIn si;
        std::memset(&si, 0, sizeof(si));
        si.SZ = SZ;
        si.eta = ep.eta;
        hipLaunchKernelGGL(k_stress_g, dim3(nb), dim3(256), 0, s, nc, g, si, part);
        OFDFT_HIP(hipGetLastError());
        const std::vector<double> v = reduce_cols(part, nb, 24, s);
        sym_add(S, v.data() + 18, 4.0 * M_PI / (P.volume * P.volume));
        double qsum = 0.0;
        for (int a = 0; a < nat; ++a) qsum += charges[a];
        for (int d = 0; d < 3; ++d)
            S[4 * d] += qsum * qsum * 4.0 * M_PI * (1.0 / (4.0 * ep.eta * P.volume * P.volume) / 2.0);
    }
}

// stress of TF, vW, WT (nonlocal part), PZ LDA and Hartree at rho, out [5][9]
// in that order; energies are the evaluation's (OFDFT_E_* order)
void functional_stress(Plan &P, const double *rho, const double *energies, double *out, hipStream_t s) {
    OFDFT_REQUIRE(P.prec == OFDFT_FP64, "stress needs an FP64 plan");
    const long long nc = P.Nc, n = P.N;
    const int nb = 1024;
    for (int k = 0; k < 45; ++k) out[k] = 0.0;
    double *ws = fs_workspace(P, sizeof(double) * (3 * nc * 2 + (size_t)nb * 24));
    double2 *rhoG = reinterpret_cast<double2 *>(ws);
    double2 *rA = rhoG + nc, *rB = rA + nc;
    double *part = reinterpret_cast<double *>(rB + nc);
    fft_r2c<double>(P, rho, rhoG, 1, n, nc, P.dV, nullptr, nullptr, s);
    const GGeom g = geom_of(P);
    const double V = P.volume;
    // rho0 = mean(rho) for the WT kernel (wt.py:79)
    double rho0 = 0.0;
    {
        double2 h0;
        OFDFT_HIP(hipMemcpyAsync(&h0, rhoG, sizeof(h0), hipMemcpyDeviceToHost, s));
        OFDFT_HIP(hipStreamSynchronize(s));
        rho0 = h0.x / V;
    }
    const bool wt = (P.terms & OFDFT_TERM_WT) != 0;
    StressGIn si;
    std::memset(&si, 0, sizeof(si));
    si.rhoG = (P.terms & OFDFT_TERM_HARTREE) ? rhoG : nullptr;
    const double tkf = 2.0 * std::cbrt(3.0 * rho0 * M_PI * M_PI);
    if (wt) {
        Prologue pa{}, pb{};
        pa.code[0] = OP_POW_BETA;
        pa.beta = P.alpha;
        pa.b56 = beta_is_5_6(P.alpha);
        fft_r2c<double>(P, rho, rA, 1, n, nc, P.dV, &pa, nullptr, s);
        si.rA = rA;
        si.rB = rA;
        if (std::fabs(P.alpha - P.beta) > 1e-14) {
            pb.code[0] = OP_POW_BETA;
            pb.beta = P.beta;
            pb.b56 = beta_is_5_6(P.beta);
            fft_r2c<double>(P, rho, rB, 1, n, nc, P.dV, &pb, nullptr, s);
            si.rB = rB;
        }
        si.itkf = 1.0 / tkf;
        si.mu = P.y;
    }
    if (si.rhoG || si.rA) {
        hipLaunchKernelGGL(k_stress_g, dim3(nb), dim3(256), 0, s, nc, g, si, part);
        OFDFT_HIP(hipGetLastError());
        const std::vector<double> v = reduce_cols(part, nb, 24, s);
        if (si.rhoG) {
            sym_add(out + 36, v.data(), 8.0 * M_PI / (V * V));
            for (int d = 0; d < 3; ++d) out[36 + 4 * d] -= energies[OFDFT_E_H] / V;
        }
        if (wt) {
            const double a = P.alpha, b = P.beta;
            sym_add(out + 18, v.data() + 12, M_PI * M_PI / (a * b * std::pow(rho0, a + b - 2) * tkf / 2.0) / (V * V));
            for (int d = 0; d < 3; ++d) out[18 + 4 * d] -= 2.0 / 3.0 * energies[OFDFT_E_NL] / V;
        }
    }
    if (P.terms & OFDFT_TERM_TF)
        for (int d = 0; d < 3; ++d) out[4 * d] = -2.0 / 3.0 * energies[OFDFT_E_TF] / V;
    if (P.terms & (OFDFT_TERM_VW | OFDFT_TERM_XC)) {
        // gradients of rho: C2R of i G_d rho(G) into the work fields
        double2 *gs = reinterpret_cast<double2 *>(P.spec);
        double *d = reinterpret_cast<double *>(P.work);
        hipLaunchKernelGGL(k_grad_spectra, dim3(nblk(nc)), dim3(256), 0, s, nc, g, rhoG, gs);
        OFDFT_HIP(hipGetLastError());
        fft_c2r<double>(P, gs, gs, d, 3, nc, nc, n, 1.0 / ((double)n * P.dV), s);
        hipLaunchKernelGGL(k_real_stress, dim3(nb), dim3(256), 0, s, n, rho, d, part);
        OFDFT_HIP(hipGetLastError());
        const std::vector<double> v = reduce_cols(part, nb, 7, s);
        if (P.terms & OFDFT_TERM_VW) sym_add(out + 9, v.data(), -0.25 / V * P.dV * P.y);
        if (P.terms & OFDFT_TERM_XC)
            for (int d2 = 0; d2 < 3; ++d2) out[27 + 4 * d2] = (energies[OFDFT_E_XC] - v[6] * P.dV) / V;
    }
}

}  // namespace ofdft
