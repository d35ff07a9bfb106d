// Start densities from atomic radial functions (density/density.py): the
// real-space placement of AtomicDensity.step (:18-67) used by guess_rho_atom
// (:209-249) and build_pseudo_density (:365-406), and the clamp + electron
// count normalisation that follows it and get_3d_value_recipe (:300-311).
// The reciprocal recipe itself runs through ofdft_local_pp with the density's
// radial table.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"
#include "reduce.h"

namespace ofdft {

struct PlaceSpecies {
    double rcut, rtol, sigma, scale;
    int border[3];
    int nbreak;
    long long off;  // first breakpoint in the packed breaks; coefficients at 4 * (off - species index)
};

// piecewise cubic c0 + c1 t + c2 t^2 + c3 t^3, t = d - x[i], x[i] <= d < x[i+1];
// the end pieces extend beyond the breakpoints
__device__ __forceinline__ double ppoly_eval(const double *x, const double *c, int nb, double d) {
    int lo = 0, hi = nb - 2;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (x[mid] <= d)
            lo = mid;
        else
            hi = mid - 1;
    }
    const double t = d - x[lo];
    const double *p = c + 4 * (long long)lo;
    return p[0] + t * (p[1] + t * (p[2] + t * p[3]));
}

// one workgroup per atom: the box of 2 border + 1 points per axis around the
// grid point below the atom; values are added atomically (atoms overlap)
__global__ void k_place_atoms(double *__restrict__ rho, int kind, int nat, const double *__restrict__ frac,
                              const int *__restrict__ spec, const PlaceSpecies *__restrict__ sp,
                              const double *__restrict__ breaks, const double *__restrict__ coefs, int nx, int ny,
                              int nz, const double *__restrict__ lat) {
    const int a = blockIdx.x;
    if (a >= nat) return;
    const PlaceSpecies S = sp[spec[a]];
    const int nr[3] = {nx, ny, nz};
    int ip[3];
    double px[3];
    for (int d = 0; d < 3; ++d) {
        const double u = frac[3 * a + d] * nr[d];
        const double f = floor(u + 1e-8);
        ip[d] = (int)f;
        px[d] = u - f;
    }
    const int bx = S.border[0], by = S.border[1], bz = S.border[2];
    const int wx = 2 * bx + 1, wy = 2 * by + 1, wz = 2 * bz + 1;
    // a box as wide as the grid reaches its first and last plane twice; the
    // reference's fancy-index += keeps the later (+border) value there
    const bool dupx = 2 * bx == nx, dupy = 2 * by == ny, dupz = 2 * bz == nz;
    const long long total = (long long)wx * wy * wz;
    const double g = 1.0 / (sqrt(2.0 * M_PI) * S.sigma);
    const double gnorm = S.scale * g * g * g;
    const double *x = breaks + S.off;
    const double *c = coefs + 4 * (S.off - spec[a]);
    for (long long t = threadIdx.x; t < total; t += blockDim.x) {
        const int k = (int)(t % wz);
        const long long r = t / wz;
        const int j = (int)(r % wy), i = (int)(r / wy);
        const int ix = i - bx, iy = j - by, iz = k - bz;
        if ((dupx && ix == -bx) || (dupy && iy == -by) || (dupz && iz == -bz)) continue;
        const double s0 = (ix + px[0]) / nx, s1 = (iy + px[1]) / ny, s2 = (iz + px[2]) / nz;
        // displacement as density.py:55 forms it: lattice[k][j] * s_j
        const double d0 = lat[0] * s0 + lat[1] * s1 + lat[2] * s2;
        const double d1 = lat[3] * s0 + lat[4] * s1 + lat[5] * s2;
        const double d2 = lat[6] * s0 + lat[7] * s1 + lat[8] * s2;
        const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        if (!(dist < S.rcut)) continue;
        if (S.rtol > 0.0 && !(dist > S.rtol)) continue;
        const double v = kind == 0 ? ppoly_eval(x, c, S.nbreak, dist) : gnorm * exp(-0.5 * (dist / S.sigma) * (dist / S.sigma));
        int gx = (ip[0] - ix) % nx, gy = (ip[1] - iy) % ny, gz = (ip[2] - iz) % nz;
        gx += gx < 0 ? nx : 0;
        gy += gy < 0 ? ny : 0;
        gz += gz < 0 ? nz : 0;
        atomicAdd(rho + ((long long)gx * ny + gy) * nz + gz, v);
    }
}

// rho < dtol -> dtol (optional); partial sums of rho
__global__ void k_clamp_sum(double *__restrict__ rho, long long n, int clamp, double dtol, double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double v = rho[i];
        if (clamp && v < dtol) {
            v = dtol;
            rho[i] = v;
        }
        acc += v;
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// rho *= ncharge / (sum dV), the sum at *s
__global__ void k_norm_scale(double *__restrict__ rho, long long n, const double *s, double ncharge, double dV) {
    const double f = ncharge / (s[0] * dV);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        rho[i] *= f;
}

constexpr int kMaxPlaceSpecies = 16;
constexpr int kNormSlot = 240;

// Species records and the upload layout shared by both placement entry
// points: species records, breakpoints, coefficients, scaled positions,
// species indices and the lattice, each part 256-byte aligned.
struct PlaceUpload {
    PlaceSpecies sp[kMaxPlaceSpecies];
    long long nbk = 0;
    size_t b_sp = 0, b_x = 0, b_c = 0, b_f = 0, b_s = 0, b_l = 0;
    size_t o_x = 0, o_c = 0, o_f = 0, o_s = 0, o_l = 0, total = 0;
};

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// validates the arguments and fills U (no device work)
static void place_prepare(const Plan *P, int kind, int nspecies, const int *nbreak, const double *const *breaks,
                          const double *const *coef, const double *rcut, const double *rtol, const double *sigma,
                          const double *scale, int nat, const double *pos_frac, const int *species, const void *rho,
                          PlaceUpload &U) {
    OFDFT_REQUIRE(P && rho && rcut && scale && pos_frac && species, "atomic_density: bad arguments");
    OFDFT_REQUIRE(P->prec == OFDFT_FP64, "atomic_density: needs an FP64 plan");
    OFDFT_REQUIRE(kind == 0 || kind == 1, "atomic_density: kind is 0 (radial spline) or 1 (Gaussian)");
    OFDFT_REQUIRE(nspecies >= 1 && nspecies <= kMaxPlaceSpecies, "atomic_density: 1..16 species");
    OFDFT_REQUIRE(nat >= 0, "atomic_density: nat < 0");
    if (kind == 0) OFDFT_REQUIRE(nbreak && breaks && coef, "atomic_density: spline tables missing");
    if (kind == 1) OFDFT_REQUIRE(sigma, "atomic_density: sigma missing");
    const int nr[3] = {P->nr[0], P->nr[1], P->nr[2]};
    U.nbk = 0;
    for (int k = 0; k < nspecies; ++k) {
        PlaceSpecies &S = U.sp[k];
        S.rcut = rcut[k];
        S.rtol = rtol ? rtol[k] : 0.0;
        S.sigma = kind == 1 ? sigma[k] : 1.0;
        S.scale = scale[k];
        OFDFT_REQUIRE(S.rcut > 0.0, "atomic_density: rcut must be positive");
        if (kind == 1) OFDFT_REQUIRE(S.sigma > 0.0, "atomic_density: sigma must be positive");
        for (int d = 0; d < 3; ++d) {
            // border = min(int(rcut / spacing) + 1, nr // 2), spacing = |a_d| / nr_d
            const double len = std::sqrt(P->lat[3 * d] * P->lat[3 * d] + P->lat[3 * d + 1] * P->lat[3 * d + 1] +
                                         P->lat[3 * d + 2] * P->lat[3 * d + 2]);
            const int b = (int)(S.rcut / (len / nr[d])) + 1;
            S.border[d] = std::min(b, nr[d] / 2);
        }
        S.nbreak = kind == 0 ? nbreak[k] : 2;
        if (kind == 0) OFDFT_REQUIRE(S.nbreak >= 2 && breaks[k] && coef[k], "atomic_density: bad spline table");
        S.off = U.nbk;
        U.nbk += S.nbreak;
    }
    for (int a = 0; a < nat; ++a)
        OFDFT_REQUIRE(species[a] >= 0 && species[a] < nspecies, "atomic_density: species index out of range");
    U.b_sp = sizeof(PlaceSpecies) * nspecies;
    U.b_x = sizeof(double) * U.nbk;
    U.b_c = sizeof(double) * 4 * (U.nbk - nspecies);
    U.b_f = sizeof(double) * 3 * (size_t)nat;
    U.b_s = sizeof(int) * (size_t)nat;
    U.b_l = sizeof(double) * 9;
    U.o_x = align256(U.b_sp);
    U.o_c = U.o_x + align256(U.b_x);
    U.o_f = U.o_c + align256(U.b_c);
    U.o_s = U.o_f + align256(U.b_f);
    U.o_l = U.o_s + align256(U.b_s);
    U.total = U.o_l + align256(U.b_l);
}

static void place_launch(const Plan *P, int kind, int nat, const PlaceUpload &U, const char *d, double *rho,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_place_atoms, dim3(nat), dim3(256), 0, s, rho, kind, nat,
                       reinterpret_cast<const double *>(d + U.o_f), reinterpret_cast<const int *>(d + U.o_s),
                       reinterpret_cast<const PlaceSpecies *>(d), reinterpret_cast<const double *>(d + U.o_x),
                       reinterpret_cast<const double *>(d + U.o_c), P->nr[0], P->nr[1], P->nr[2],
                       reinterpret_cast<const double *>(d + U.o_l));
    OFDFT_HIP(hipGetLastError());
}

// Device memory of ofdft_atomic_density_noalloc: its upload lives here
// instead of in a per-call allocation.  Each GPU holds its own copy until the
// process ends; every call on a GPU reuses that GPU's copy.
constexpr size_t kPlaceScratchBytes = size_t(4) << 20;
static __device__ __align__(256) char scratch[kPlaceScratchBytes];
constexpr int kMaxScratchDevices = 64;

// device address of `scratch` on the current device, looked up once per
// device (the symbol's host-side name is no device address)
static char *place_scratch() {
    static std::mutex mtx;
    static char *addr[kMaxScratchDevices] = {};
    int dev = -1;
    OFDFT_HIP(hipGetDevice(&dev));
    OFDFT_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "atomic_density_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(mtx);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        OFDFT_REQUIRE(e == hipSuccess && p != nullptr,
                      std::string("atomic_density_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
        addr[dev] = reinterpret_cast<char *>(p);
    }
    return addr[dev];
}

}  // namespace ofdft

using namespace ofdft;

extern "C" int ofdft_atomic_density(ofdft_plan *plan, int kind, int nspecies, const int *nbreak,
                                    const double *const *breaks, const double *const *coef, const double *rcut,
                                    const double *rtol, const double *sigma, const double *scale, int nat,
                                    const double *pos_frac, const int *species, void *rho, void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    OFDFT_LEGACY_LOCK;  // the upload buffer is allocated and freed per call
    struct DevBuf {
        void *p = nullptr;
        ~DevBuf() {
            if (p) (void)hipFree(p);
        }
    } buf;
    try {
        PlaceUpload U;
        place_prepare(P, kind, nspecies, nbreak, breaks, coef, rcut, rtol, sigma, scale, nat, pos_frac, species, rho,
                      U);
        if (nat == 0) return 0;
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        // one upload from one host image of the layout
        std::vector<char> h(U.total, 0);
        char *o = h.data();
        std::memcpy(o, U.sp, U.b_sp);
        if (kind == 0)
            for (int k = 0; k < nspecies; ++k) {
                std::memcpy(o + U.o_x + sizeof(double) * U.sp[k].off, breaks[k], sizeof(double) * U.sp[k].nbreak);
                std::memcpy(o + U.o_c + sizeof(double) * 4 * (U.sp[k].off - k), coef[k],
                            sizeof(double) * 4 * (U.sp[k].nbreak - 1));
            }
        std::memcpy(o + U.o_f, pos_frac, U.b_f);
        std::memcpy(o + U.o_s, species, U.b_s);
        std::memcpy(o + U.o_l, P->lat, U.b_l);
        OFDFT_HIP(hipMalloc(&buf.p, U.total));
        char *d = reinterpret_cast<char *>(buf.p);
        OFDFT_HIP(hipMemcpyAsync(d, h.data(), U.total, hipMemcpyHostToDevice, s));
        place_launch(P, kind, nat, U, d, reinterpret_cast<double *>(rho), s);
        // the upload buffer is freed (after the stream drains) when buf goes out of scope
        OFDFT_HIP(hipStreamSynchronize(s));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}

extern "C" int ofdft_atomic_density_noalloc(ofdft_plan *plan, int kind, int nspecies, const int *nbreak,
                                            const double *const *breaks, const double *const *coef,
                                            const double *rcut, const double *rtol, const double *sigma,
                                            const double *scale, int nat, const double *pos_frac, const int *species,
                                            void *rho, void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        PlaceUpload U;
        place_prepare(P, kind, nspecies, nbreak, breaks, coef, rcut, rtol, sigma, scale, nat, pos_frac, species, rho,
                      U);
        OFDFT_REQUIRE(U.total <= kPlaceScratchBytes,
                      "atomic_density_noalloc: tables and positions need " + std::to_string(U.total) +
                          " bytes, more than the " + std::to_string(kPlaceScratchBytes) +
                          " bytes of its static device buffer; use ofdft_atomic_density");
        if (nat == 0) return 0;
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        char *d = place_scratch();
        // the parts go straight from the caller's arrays into the static
        // buffer (every byte the kernel reads is written first, so nothing is
        // cleared); the host waits below, so the arrays may be released after
        // the call
        OFDFT_HIP(hipMemcpyAsync(d, U.sp, U.b_sp, hipMemcpyHostToDevice, s));
        if (kind == 0)
            for (int k = 0; k < nspecies; ++k) {
                OFDFT_HIP(hipMemcpyAsync(d + U.o_x + sizeof(double) * U.sp[k].off, breaks[k],
                                         sizeof(double) * U.sp[k].nbreak, hipMemcpyHostToDevice, s));
                OFDFT_HIP(hipMemcpyAsync(d + U.o_c + sizeof(double) * 4 * (U.sp[k].off - k), coef[k],
                                         sizeof(double) * 4 * (U.sp[k].nbreak - 1), hipMemcpyHostToDevice, s));
            }
        OFDFT_HIP(hipMemcpyAsync(d + U.o_f, pos_frac, U.b_f, hipMemcpyHostToDevice, s));
        OFDFT_HIP(hipMemcpyAsync(d + U.o_s, species, U.b_s, hipMemcpyHostToDevice, s));
        OFDFT_HIP(hipMemcpyAsync(d + U.o_l, P->lat, U.b_l, hipMemcpyHostToDevice, s));
        place_launch(P, kind, nat, U, d, reinterpret_cast<double *>(rho), s);
        OFDFT_HIP(hipStreamSynchronize(s));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}

extern "C" int ofdft_clamp_normalize(ofdft_plan *plan, void *rho, int clamp, double dtol, double ncharge,
                                     void *stream) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    try {
        OFDFT_REQUIRE(P && rho, "clamp_normalize: bad arguments");
        OFDFT_REQUIRE(P->prec == OFDFT_FP64, "clamp_normalize: needs an FP64 plan");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        const int nb = std::min(reduce_blocks(P->N), P->red_blocks);
        hipLaunchKernelGGL(k_clamp_sum, dim3(nb), dim3(256), 0, s, reinterpret_cast<double *>(rho), P->N, clamp, dtol,
                           P->red);
        if (ncharge > 1e-6) {
            const double one = 1.0;
            reduce_final(P->red, nb, 1, P->scal + kNormSlot, &one, s);
            hipLaunchKernelGGL(k_norm_scale, dim3(nb), dim3(256), 0, s, reinterpret_cast<double *>(rho), P->N,
                               P->scal + kNormSlot, ncharge, P->dV);
        }
        OFDFT_HIP(hipGetLastError());
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    }
}
