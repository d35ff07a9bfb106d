// Mixed-radix (2/3/4/5/8, and 7/11/13) 3D FFT for gfx950.
//
// A 3D transform is three axis passes.  The z pass works on contiguous rows
// (R2C/C2R pack two real rows into one complex FFT); the y and x passes work
// on strided pencils, staging a tile of `tile` adjacent columns x n points in
// LDS so every global access is a contiguous run of tile*sizeof(complex)
// bytes.  Inside LDS each pencil is transformed by a Stockham autosort
// sequence of radix-R stages (one butterfly per thread per stage), twiddles
// read from an LDS copy of exp(-2 pi i m / n).
//
// Normalisation is left to the caller through `scale`, applied on the final
// store of the last pass.  The DFTpy conventions (forward * dV, inverse
// / (N dV), field.py:373-390,739-764) are applied by the C ABI layer.
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "dft.h"
#include "internal.h"
#include "pointwise.h"

namespace ofdft {

// One Stockham stage on `npen` pencils of length n held in LDS.  Element e of
// pencil p lives at buf[p * pstride + e * estride].
template <int R, bool INV, bool PEN_FAST, typename C>
__device__ __forceinline__ void stage(C *buf, const C *tw, int n, int Ns, int npen, int pstride, int estride) {
    const int nb = n / R;
    const int total = nb * npen;
    const int b = threadIdx.x;
    const bool act = b < total;
    C v[R];
    int pen = 0, j = 0;
    if (act) {
        if (PEN_FAST) {
            pen = b % npen;
            j = b / npen;
        } else {
            pen = b / nb;
            j = b - pen * nb;
        }
        const C *base = buf + pen * pstride;
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = base[(j + r * nb) * estride];
        if (Ns > 1) {
            const int k = j % Ns;
            const int step = n / (Ns * R);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[r * k * step];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
    }
    __syncthreads();
    if (act) {
        DFT<R, INV, C>::run(v);
        const int k = j % Ns;
        const int d = (j / Ns) * Ns * R + k;
        C *base = buf + pen * pstride;
#pragma unroll
        for (int r = 0; r < R; ++r) base[(d + r * Ns) * estride] = v[r];
    }
    __syncthreads();
}

template <bool INV, bool PEN_FAST, typename C>
__device__ void lds_fft(C *buf, const C *tw, const AxisPlan &ap, int npen, int pstride, int estride) {
    int Ns = 1;
    for (int s = 0; s < ap.nst; ++s) {
        const int R = ap.rad[s];
        switch (R) {
            case 2: stage<2, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            case 3: stage<3, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            case 4: stage<4, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            case 5: stage<5, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            case 7: stage<7, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            case 11: stage<11, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            case 13: stage<13, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
            default: stage<8, INV, PEN_FAST>(buf, tw, ap.n, Ns, npen, pstride, estride); break;
        }
        Ns *= R;
    }
}

// ------------------------------------------------------------------ prologue
template <typename T>
__device__ __forceinline__ T prologue(int code, T x, double beta, bool b56, const T *aux, long long idx) {
    switch (code) {
        case OP_SQRT_RHOM: return sqrt(x > (T)0 ? x : (T)1e-30);
        case OP_POW_BETA: return pow_beta<T>(x < (T)1e-30 ? (T)1e-30 : x, (T)beta, b56);
        case OP_DSQRT: {
            T r = aux[idx];
            T s = sqrt(r > (T)0 ? r : (T)1e-30);
            return x / ((T)2 * s);
        }
        case OP_DPOW_BETA: {
            T r = aux[idx];
            T rc = r < (T)1e-30 ? (T)1e-30 : r;
            return (T)beta * pow_beta_m1<T>(rc, (T)beta, b56) * x;
        }
        default: return x;
    }
}

// ------------------------------------------------------------------ kernels
// Columns pass: data viewed as [outer][n][M]; transform along n for a tile of
// `tile` adjacent columns.  in may equal out.
template <typename T, bool INV>
__global__ void k_cols(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, long long bs, int n,
                       long long M, int tiles, AxisPlan ap, const cpx<T> *__restrict__ twg, int tile, T scale) {
    using C = cpx<T>;
    extern __shared__ unsigned char smem_raw[];
    C *tw = reinterpret_cast<C *>(smem_raw);
    C *buf = tw + n;
    const int t = blockIdx.x % tiles;
    const long long o = blockIdx.x / tiles;
    const long long off = (long long)blockIdx.y * bs + o * n * M + (long long)t * tile;
    const C *src = in + off;
    C *dst = out + off;
    const int mv = (int)min((long long)tile, M - (long long)t * tile);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    const int tot = n * tile;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int i = e / tile, c = e - i * tile;
        buf[e] = (c < mv) ? src[(long long)i * M + c] : C{(T)0, (T)0};
    }
    __syncthreads();
    lds_fft<INV, true>(buf, tw, ap, tile, 1, tile);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int i = e / tile, c = e - i * tile;
        if (c < mv) {
            C v = buf[e];
            dst[(long long)i * M + c] = C{v.x * scale, v.y * scale};
        }
    }
}

// Rows pass, real -> half complex.  Rows 2p and 2p+1 are packed into one
// complex FFT z = a + i b and separated afterwards:
// A_k = (Z_k + conj Z_{-k}) / 2, B_k = (Z_k - conj Z_{-k}) / (2i).
// Optionally accumulates the raw row sum of batch slot `sum_slot` in FP64
// partials[blockIdx.x] (used for rho0 = mean(rho)).
template <typename T>
__global__ void k_rows_r2c(const T *__restrict__ in, cpx<T> *__restrict__ out, long long in_bs, long long out_bs,
                           long long nrows, int n, int nzc, AxisPlan ap, const cpx<T> *__restrict__ twg, int rows,
                           T scale, Prologue pro, double *partials, int sum_slot) {
    using C = cpx<T>;
    extern __shared__ unsigned char smem_raw[];
    C *tw = reinterpret_cast<C *>(smem_raw);
    C *buf = tw + n;
    const int slot = blockIdx.y;
    const long long row0 = (long long)blockIdx.x * rows;
    const T *src = in + (long long)slot * in_bs + row0 * n;
    C *dst = out + (long long)slot * out_bs + row0 * nzc;
    const int nv = (int)min((long long)rows, nrows - row0);
    const int code = pro.code[slot & 3];
    const T *aux = reinterpret_cast<const T *>(pro.aux);
    const T *phi = reinterpret_cast<const T *>(pro.phi);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    const int tot = rows * n;
    double acc = 0.0;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / n;
        T x = (T)0;
        if (r < nv) {
            x = src[e];
            if (phi) x = (T)(2.0 * (double)phi[row0 * n + e] * (double)x);
            acc += (double)x;
            x = prologue<T>(code, x, pro.beta, pro.b56, aux, row0 * n + e);
        }
        T *pb = reinterpret_cast<T *>(&buf[(r >> 1) * n + (e - r * n)]);
        pb[r & 1] = x;
    }
    if (partials != nullptr && slot == sum_slot) {
        // block sum of the raw input (deterministic tree order)
        __shared__ double red[16];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) red[wid] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) s += red[w];
            partials[blockIdx.x] = s;
        }
    }
    __syncthreads();
    lds_fft<false, false>(buf, tw, ap, rows >> 1, n, 1);
    const int toto = rows * nzc;
    for (int e = threadIdx.x; e < toto; e += blockDim.x) {
        const int r = e / nzc, k = e - r * nzc;
        if (r < nv) {
            const int p = r >> 1;
            C z = buf[p * n + k];
            C zc = buf[p * n + (k == 0 ? 0 : n - k)];
            zc.y = -zc.y;
            C o;
            if ((r & 1) == 0) {
                o = C{(z.x + zc.x) * (T)0.5, (z.y + zc.y) * (T)0.5};
            } else {
                o = C{(z.y - zc.y) * (T)0.5, -(z.x - zc.x) * (T)0.5};
            }
            dst[e] = C{o.x * scale, o.y * scale};
        }
    }
}

// Rows pass, half complex -> real (irfft semantics: the imaginary parts of
// the k = 0 and Nyquist entries are ignored).
template <typename T>
__global__ void k_rows_c2r(const cpx<T> *__restrict__ in, T *__restrict__ out, long long in_bs, long long out_bs,
                           long long nrows, int n, int nzc, AxisPlan ap, const cpx<T> *__restrict__ twg, int rows,
                           T scale) {
    using C = cpx<T>;
    extern __shared__ unsigned char smem_raw[];
    C *tw = reinterpret_cast<C *>(smem_raw);
    C *stg = tw + n;
    C *buf = stg + rows * nzc;
    const long long row0 = (long long)blockIdx.x * rows;
    const C *src = in + (long long)blockIdx.y * in_bs + row0 * nzc;
    T *dst = out + (long long)blockIdx.y * out_bs + row0 * n;
    const int nv = (int)min((long long)rows, nrows - row0);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    const int tot = rows * nzc;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nzc, k = e - r * nzc;
        C v = (r < nv) ? src[e] : C{(T)0, (T)0};
        if (k == 0 || 2 * k == n) v.y = (T)0;
        stg[e] = v;
    }
    __syncthreads();
    const int npen = rows >> 1;
    const int totb = npen * n;
    for (int e = threadIdx.x; e < totb; e += blockDim.x) {
        const int p = e / n, k = e - p * n;
        C A, B;
        if (k < nzc) {
            A = stg[(2 * p) * nzc + k];
            B = stg[(2 * p + 1) * nzc + k];
        } else {
            A = stg[(2 * p) * nzc + n - k];
            B = stg[(2 * p + 1) * nzc + n - k];
            A.y = -A.y;
            B.y = -B.y;
        }
        buf[e] = C{A.x - B.y, A.y + B.x};
    }
    __syncthreads();
    lds_fft<true, false>(buf, tw, ap, npen, n, 1);
    const int toto = rows * n;
    for (int e = threadIdx.x; e < toto; e += blockDim.x) {
        const int r = e / n;
        if (r < nv) {
            const C z = buf[(r >> 1) * n + (e - r * n)];
            dst[e] = ((r & 1) ? z.y : z.x) * scale;
        }
    }
}

// Rows pass, complex -> complex on contiguous rows of length n.
template <typename T, bool INV>
__global__ void k_rows_c2c(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, long long bs, long long nrows,
                           int n, AxisPlan ap, const cpx<T> *__restrict__ twg, int rows, T scale) {
    using C = cpx<T>;
    extern __shared__ unsigned char smem_raw[];
    C *tw = reinterpret_cast<C *>(smem_raw);
    C *buf = tw + n;
    const long long row0 = (long long)blockIdx.x * rows;
    const C *src = in + (long long)blockIdx.y * bs + row0 * n;
    C *dst = out + (long long)blockIdx.y * bs + row0 * n;
    const int nv = (int)min((long long)rows, nrows - row0);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    const int tot = rows * n;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) buf[e] = (e / n < nv) ? src[e] : C{(T)0, (T)0};
    __syncthreads();
    lds_fft<INV, false>(buf, tw, ap, rows, n, 1);
    for (int e = threadIdx.x; e < tot; e += blockDim.x)
        if (e / n < nv) {
            C v = buf[e];
            dst[e] = C{v.x * scale, v.y * scale};
        }
}

// ------------------------------------------------------------------ host side
bool fixed_size(int n);
template <typename T>
bool cols_fixed(int n, int var, bool inv, const cpx<T> *in, cpx<T> *out, int batch, long long bs, long long ncols,
                long long P, long long Q, long long S, long long vp, const void *tw, double scale, hipStream_t s,
                const double *gate);
template <typename T>
bool r2c_fixed(int n, int var, const T *in, cpx<T> *out, int batch, long long in_bs, long long out_bs,
               long long nrows, long long opitch, const void *tw, double scale, const Prologue &pro, double *partials,
               int sum_slot, hipStream_t s, const double *gate);
template <typename T>
bool c2r_fixed(int n, int var, const cpx<T> *in, T *out, int batch, long long in_bs, long long out_bs,
               long long nrows, long long ipitch, const void *tw, double scale, hipStream_t s, const double *gate);
int r2c_fixed_tiles(int n, int var, int prec, int batch, long long nrows);

int r2c_num_partials(const Plan &P, int batch) {
    const long long nrows = (long long)P.nr[0] * P.nr[1];
    if (P.use_fixed && fixed_size(P.nr[2])) return r2c_fixed_tiles(P.nr[2], P.var_r2c, P.prec, batch, nrows);
    return (int)((nrows + P.rows_r2c.tile - 1) / P.rows_r2c.tile);
}

long long pipe_pitch(const Plan &P) {
    if (P.use_fixed && fixed_size(P.nr[0]) && fixed_size(P.nr[1]) && fixed_size(P.nr[2])) return P.nzp;
    return P.nzc;
}

static AxisPlan make_axis(int n) {
    AxisPlan a;
    std::memset(&a, 0, sizeof(a));
    a.n = n;
    int m = n, e2 = 0;
    std::vector<int> r;
    while (m % 2 == 0) {
        m /= 2;
        ++e2;
    }
    if (e2 % 3 == 0) {
        for (int i = 0; i < e2 / 3; ++i) r.push_back(8);
    } else if (e2 % 3 == 1) {
        if (e2 >= 4) {
            for (int i = 0; i < e2 / 3 - 1; ++i) r.push_back(8);
            r.push_back(4);
            r.push_back(4);
        } else {
            r.push_back(2);
        }
    } else {
        for (int i = 0; i < e2 / 3; ++i) r.push_back(8);
        r.push_back(4);
    }
    while (m % 5 == 0) {
        m /= 5;
        r.push_back(5);
    }
    while (m % 3 == 0) {
        m /= 3;
        r.push_back(3);
    }
    for (int pr : {7, 11, 13})
        while (m % pr == 0) {
            m /= pr;
            r.push_back(pr);
        }
    OFDFT_REQUIRE(m == 1, "FFT length " + std::to_string(n) + " has a prime factor above 13");
    OFDFT_REQUIRE((int)r.size() <= kMaxStages, "FFT length too large");
    a.nst = (int)r.size();
    a.rmin = 64;
    for (size_t i = 0; i < r.size(); ++i) {
        a.rad[i] = r[i];
        a.rmin = std::min(a.rmin, r[i]);
    }
    if (a.nst == 0) a.rmin = 1;
    return a;
}

static int round64(int t) { return ((t + 63) / 64) * 64; }

static PassShape cols_shape(const AxisPlan &a, long long M, size_t cs) {
    const int tpp = a.n / a.rmin;
    int tile = cs == 8 ? 32 : 16;  // 256-byte contiguous runs
    if ((long long)tile > M) tile = (int)M;
    const size_t lds_cap = 64 * 1024;
    while (tile > 1 && (tile * tpp > 1024 || (size_t)(tile + 1) * a.n * cs > lds_cap)) tile = (tile + 1) / 2;
    OFDFT_REQUIRE(tile * tpp <= 1024 && (size_t)(tile + 1) * a.n * cs <= lds_cap, "FFT axis too long");
    PassShape p;
    p.tile = tile;
    p.threads = std::min(1024, std::max(64, round64(tile * tpp)));
    p.lds = (int)((size_t)(tile + 1) * a.n * cs);
    return p;
}

static PassShape rows_shape(const AxisPlan &a, int nzc, size_t cs, int mode /*0 r2c,1 c2r,2 c2c*/) {
    const int tpp = a.n / a.rmin;
    int npen = 16;
    const size_t lds_cap = 64 * 1024;
    auto lds_of = [&](int np) -> size_t {
        if (mode == 0) return (size_t)(np * a.n + a.n) * cs;
        if (mode == 1) return (size_t)(2 * np * nzc + np * a.n + a.n) * cs;
        return (size_t)(np * a.n + a.n) * cs;
    };
    while (npen > 1 && (npen * tpp > 1024 || lds_of(npen) > lds_cap)) --npen;
    OFDFT_REQUIRE(npen * tpp <= 1024 && lds_of(npen) <= lds_cap, "FFT row too long");
    PassShape p;
    p.tile = mode == 2 ? npen : 2 * npen;
    p.threads = std::min(1024, std::max(64, round64(npen * tpp)));
    p.lds = (int)lds_of(npen);
    return p;
}

void plan_fft_setup(Plan &P) {
    const size_t cs = P.prec == OFDFT_FP64 ? 16 : 8;
    for (int d = 0; d < 3; ++d) {
        P.ax[d] = make_axis(P.nr[d]);
        const int n = P.nr[d];
        std::vector<double> tw(2 * n);
        for (int m = 0; m < n; ++m) {
            // exp(-2 pi i m / n) with the angle reduced to the first octant
            long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
            tw[2 * m] = (double)cosl(ang);
            tw[2 * m + 1] = (double)-sinl(ang);
        }
        OFDFT_HIP(hipMalloc(&P.tw[d], n * cs));
        if (P.prec == OFDFT_FP64) {
            OFDFT_HIP(hipMemcpy(P.tw[d], tw.data(), n * cs, hipMemcpyHostToDevice));
        } else {
            std::vector<float> twf(2 * n);
            for (int i = 0; i < 2 * n; ++i) twf[i] = (float)tw[i];
            OFDFT_HIP(hipMemcpy(P.tw[d], twf.data(), n * cs, hipMemcpyHostToDevice));
        }
    }
    P.rows_r2c = rows_shape(P.ax[2], P.nzc, cs, 0);
    P.rows_c2r = rows_shape(P.ax[2], P.nzc, cs, 1);
    P.rows_c2c = rows_shape(P.ax[2], P.nzc, cs, 2);
    P.cols_y = cols_shape(P.ax[1], P.nzc, cs);
    P.cols_x = cols_shape(P.ax[0], (long long)P.nr[1] * P.nzc, cs);
    P.cols_yc = cols_shape(P.ax[1], P.nr[2], cs);
    P.cols_xc = cols_shape(P.ax[0], (long long)P.nr[1] * P.nr[2], cs);
}

void plan_fft_free(Plan &P) {
    for (int d = 0; d < 3; ++d)
        if (P.tw[d]) {
            (void)hipFree(P.tw[d]);
            P.tw[d] = nullptr;
        }
}

template <typename T>
void fft_cols(Plan &P, int axis, bool full_grid, bool inverse, const cpx<T> *in, cpx<T> *out, int batch,
              long long bs, double scale, hipStream_t s, long long pitch) {
    const int n = P.nr[axis];
    const long long zlen = full_grid ? P.nr[2] : P.nzc;
    if (pitch <= 0) pitch = zlen;
    if (n == 1 && in == out && scale == 1.0) return;
    if (P.use_fixed && fixed_size(n)) {
        // flattened pencils: base(c) = (c / Pp) * Qq + c % Pp, stride Ss; columns
        // at or beyond zlen within a pitched row are padding and skipped
        long long ncols, Pp, Qq, Ss;
        if (axis == 1) {
            ncols = (long long)P.nr[0] * pitch;
            Pp = pitch;
            Qq = (long long)P.nr[1] * pitch;
            Ss = pitch;
        } else {
            ncols = (long long)P.nr[1] * pitch;
            Pp = Qq = pitch;
            Ss = ncols;
        }
        cols_fixed<T>(n, P.var_cols[axis], inverse, in, out, batch, bs, ncols, Pp, Qq, Ss, zlen, P.tw[axis], scale, s,
                      P.gate);
        OFDFT_HIP(hipGetLastError());
        return;
    }
    OFDFT_REQUIRE(pitch == zlen, "generic column pass needs an unpadded layout");
    OFDFT_REQUIRE(P.gate == nullptr, "gated launch of a generic FFT pass");
    long long M, outer;
    PassShape sh;
    if (axis == 1) {
        M = zlen;
        outer = P.nr[0];
        sh = full_grid ? P.cols_yc : P.cols_y;
    } else {
        M = (long long)P.nr[1] * zlen;
        outer = 1;
        sh = full_grid ? P.cols_xc : P.cols_x;
    }
    const int tiles = (int)((M + sh.tile - 1) / sh.tile);
    dim3 grid((unsigned)(outer * tiles), (unsigned)batch);
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(P.tw[axis]);
    if (inverse)
        hipLaunchKernelGGL((k_cols<T, true>), grid, dim3(sh.threads), sh.lds, s, in, out, bs, n, M, tiles,
                           P.ax[axis], tw, sh.tile, (T)scale);
    else
        hipLaunchKernelGGL((k_cols<T, false>), grid, dim3(sh.threads), sh.lds, s, in, out, bs, n, M, tiles,
                           P.ax[axis], tw, sh.tile, (T)scale);
    OFDFT_HIP(hipGetLastError());
}

template <typename T>
void fft_r2c(Plan &P, const T *in, cpx<T> *out, int batch, long long in_bs, long long out_bs, double scale,
             const Prologue *pro, double *sum_partials, hipStream_t s, long long pitch, bool skip_x) {
    if (pitch <= 0) pitch = P.nzc;
    OFDFT_REQUIRE(!skip_x || scale == 1.0, "r2c without the x pass cannot scale");
    Prologue pr;
    std::memset(&pr, 0, sizeof(pr));
    int sum_slot = -1;
    if (pro) pr = *pro;
    for (int b = 0; b < 4; ++b)
        OFDFT_REQUIRE(pr.aux != nullptr || (pr.code[b] != OP_DSQRT && pr.code[b] != OP_DPOW_BETA),
                      "R2C prologue: a linear-response op needs rho");
    if (sum_partials) {
        sum_slot = 0;
        for (int b = 0; b < batch; ++b)
            if (pr.code[b & 3] == OP_IDENTITY) sum_slot = b;
    }
    const long long nrows = (long long)P.nr[0] * P.nr[1];
    const PassShape sh = P.rows_r2c;
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(P.tw[2]);
    const bool last_z = (P.nr[0] == 1 && P.nr[1] == 1);
    if (!(P.use_fixed && r2c_fixed<T>(P.nr[2], P.var_r2c, in, out, batch, in_bs, out_bs, nrows, pitch, P.tw[2],
                                      last_z ? scale : 1.0, pr, sum_partials, sum_slot, s, P.gate))) {
        OFDFT_REQUIRE(pitch == P.nzc, "generic R2C pass needs an unpadded layout");
        OFDFT_REQUIRE(P.gate == nullptr, "gated launch of a generic FFT pass");
        dim3 grid((unsigned)((nrows + sh.tile - 1) / sh.tile), (unsigned)batch);
        hipLaunchKernelGGL((k_rows_r2c<T>), grid, dim3(sh.threads), sh.lds, s, in, out, in_bs, out_bs, nrows, P.nr[2],
                           P.nzc, P.ax[2], tw, sh.tile, (T)(last_z ? scale : 1.0), pr, sum_partials, sum_slot);
    }
    OFDFT_HIP(hipGetLastError());
    if (last_z) return;
    fft_cols<T>(P, 1, false, false, out, out, batch, out_bs, 1.0, s, pitch);
    if (!skip_x) fft_cols<T>(P, 0, false, false, out, out, batch, out_bs, scale, s, pitch);
}

template <typename T>
void fft_c2r(Plan &P, const cpx<T> *in, cpx<T> *tmp, T *out, int batch, long long in_bs, long long tmp_bs,
             long long out_bs, double scale, hipStream_t s, long long pitch, bool skip_x) {
    if (pitch <= 0) pitch = P.nzc;
    const cpx<T> *src = in;
    long long src_bs = in_bs;
    if (!(P.nr[0] == 1 && P.nr[1] == 1)) {
        OFDFT_REQUIRE(in_bs == tmp_bs, "c2r: batch strides of input and scratch differ");
        if (!skip_x) fft_cols<T>(P, 0, false, true, in, tmp, batch, tmp_bs, 1.0, s, pitch);
        fft_cols<T>(P, 1, false, true, skip_x ? in : tmp, tmp, batch, tmp_bs, 1.0, s, pitch);
        src = tmp;
        src_bs = tmp_bs;
    }
    const long long nrows = (long long)P.nr[0] * P.nr[1];
    const PassShape sh = P.rows_c2r;
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(P.tw[2]);
    if (!(P.use_fixed && c2r_fixed<T>(P.nr[2], P.var_c2r, src, out, batch, src_bs, out_bs, nrows, pitch, P.tw[2],
                                      scale, s, P.gate))) {
        OFDFT_REQUIRE(pitch == P.nzc, "generic C2R pass needs an unpadded layout");
        OFDFT_REQUIRE(P.gate == nullptr, "gated launch of a generic FFT pass");
        dim3 grid((unsigned)((nrows + sh.tile - 1) / sh.tile), (unsigned)batch);
        hipLaunchKernelGGL((k_rows_c2r<T>), grid, dim3(sh.threads), sh.lds, s, src, out, src_bs, out_bs, nrows,
                           P.nr[2], P.nzc, P.ax[2], tw, sh.tile, (T)scale);
    }
    OFDFT_HIP(hipGetLastError());
}

template <typename T>
void fft_c2c(Plan &P, const cpx<T> *in, cpx<T> *out, bool inverse, int batch, long long bs, double scale,
             hipStream_t s) {
    const long long nrows = (long long)P.nr[0] * P.nr[1];
    const PassShape sh = P.rows_c2c;
    dim3 grid((unsigned)((nrows + sh.tile - 1) / sh.tile), (unsigned)batch);
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(P.tw[2]);
    const bool last_z = (P.nr[0] == 1 && P.nr[1] == 1);
    const T sz = (T)(last_z ? scale : 1.0);
    OFDFT_REQUIRE(P.gate == nullptr, "gated launch of a generic FFT pass");
    if (inverse)
        hipLaunchKernelGGL((k_rows_c2c<T, true>), grid, dim3(sh.threads), sh.lds, s, in, out, bs, nrows, P.nr[2],
                           P.ax[2], tw, sh.tile, sz);
    else
        hipLaunchKernelGGL((k_rows_c2c<T, false>), grid, dim3(sh.threads), sh.lds, s, in, out, bs, nrows, P.nr[2],
                           P.ax[2], tw, sh.tile, sz);
    OFDFT_HIP(hipGetLastError());
    if (last_z) return;
    fft_cols<T>(P, 1, true, inverse, out, out, batch, bs, 1.0, s, P.nr[2]);
    fft_cols<T>(P, 0, true, inverse, out, out, batch, bs, scale, s, P.nr[2]);
}

template <typename T>
void bench_rows_r2c(Plan &P, const T *in, cpx<T> *out, int batch, const Prologue *pro, hipStream_t s) {
    const long long nrows = (long long)P.nr[0] * P.nr[1];
    const long long pitch = pipe_pitch(P);
    if (P.use_fixed && r2c_fixed<T>(P.nr[2], P.var_r2c, in, out, batch, 0LL, (long long)P.nr[0] * P.nr[1] * pitch, nrows, pitch,
                                    P.tw[2], 1.0, *pro, nullptr, -1, s, nullptr)) {
        OFDFT_HIP(hipGetLastError());
        return;
    }
    const PassShape sh = P.rows_r2c;
    dim3 grid((unsigned)((nrows + sh.tile - 1) / sh.tile), (unsigned)batch);
    hipLaunchKernelGGL((k_rows_r2c<T>), grid, dim3(sh.threads), sh.lds, s, in, out, 0LL, P.Nc, nrows, P.nr[2],
                       P.nzc, P.ax[2], reinterpret_cast<const cpx<T> *>(P.tw[2]), sh.tile, (T)1, *pro,
                       (double *)nullptr, -1);
    OFDFT_HIP(hipGetLastError());
}

template <typename T> void bench_rows_c2r(Plan &P, const cpx<T> *in, T *out, int batch, hipStream_t s) {
    const long long nrows = (long long)P.nr[0] * P.nr[1];
    const long long pitch = pipe_pitch(P);
    if (P.use_fixed && c2r_fixed<T>(P.nr[2], P.var_c2r, in, out, batch, (long long)P.nr[0] * P.nr[1] * pitch, P.N, nrows, pitch,
                                    P.tw[2], 1.0, s, nullptr)) {
        OFDFT_HIP(hipGetLastError());
        return;
    }
    const PassShape sh = P.rows_c2r;
    dim3 grid((unsigned)((nrows + sh.tile - 1) / sh.tile), (unsigned)batch);
    hipLaunchKernelGGL((k_rows_c2r<T>), grid, dim3(sh.threads), sh.lds, s, in, out, P.Nc, P.N, nrows, P.nr[2],
                       P.nzc, P.ax[2], reinterpret_cast<const cpx<T> *>(P.tw[2]), sh.tile, (T)1);
    OFDFT_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ autotuner
// Times every tile variant of the size-specialised passes on the plan's own
// shapes (3 spectra, the evaluation pipeline's batch) and keeps the fastest.
// Runs on the plan's workspace; the data are whatever it holds.
template <typename F> static float time_launches(hipStream_t s, int iters, F &&f) {
    hipEvent_t e0, e1;
    OFDFT_HIP(hipEventCreate(&e0));
    OFDFT_HIP(hipEventCreate(&e1));
    f();  // warm-up
    OFDFT_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < iters; ++i) f();
    OFDFT_HIP(hipEventRecord(e1, s));
    OFDFT_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    OFDFT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / iters;
}

template <typename T> static void autotune_t(Plan &P, hipStream_t s, float *times) {
    const long long pitch = pipe_pitch(P);
    const long long ncp = (long long)P.nr[0] * P.nr[1] * pitch;
    cpx<T> *spec = reinterpret_cast<cpx<T> *>(P.spec);
    T *work = reinterpret_cast<T *>(P.work);
    const int iters = 4;
    // times: kColVariants per slot; the row passes have kTileVariants
    auto pick = [&](int *var, int slot, int nvar, auto &&run) {
        float best = 1e30f;
        int bv = 0;
        for (int v = 0; v < kColVariants; ++v)
            if (times) times[slot * kColVariants + v] = -1.0f;
        for (int v = 0; v < nvar; ++v) {
            *var = v;
            const float t = time_launches(s, iters, run);
            if (times) times[slot * kColVariants + v] = t;
            if (t < best) {
                best = t;
                bv = v;
            }
        }
        *var = bv;
    };
    if (fixed_size(P.nr[2]) && P.nr[0] * P.nr[1] > 1) {
        Prologue pro{};
        pro.code[0] = OP_SQRT_RHOM;
        pro.code[1] = OP_POW_BETA;
        pro.beta = P.beta;
        pro.b56 = 1;
        pick(&P.var_r2c, 0, kTileVariants, [&] { bench_rows_r2c<T>(P, work, spec, 3, &pro, s); });
        pick(&P.var_c2r, 1, kTileVariants, [&] { bench_rows_c2r<T>(P, spec, work, 3, s); });
    }
    // the x variant is timed on the pass the evaluation runs: the fused
    // forward x, G-space multiply and inverse x where the axis has one
    for (int axis = 0; axis < 2; ++axis)
        if (fixed_size(P.nr[axis]))
            pick(&P.var_cols[axis], 2 + axis, kColVariants, [&] {
                if (axis != 0 || !bench_xfused<T>(P, s))
                    fft_cols<T>(P, axis, false, false, spec, spec, 3, ncp, 1.0, s, pitch);
            });
}

// The variants chosen for a grid shape and precision are kept for the rest of
// the process: every later plan of that shape takes them without timing, so
// all plans of a shape transform identically (the radix order changes the
// rounding) whichever thread or stream created them.
void plan_autotune(Plan &P, hipStream_t s, float *times) {
    if (!P.use_fixed) return;
    using Key = std::array<int, 5>;
    static std::mutex mu;
    static std::map<Key, std::array<int, 5>> chosen;
    const Key key{P.nr[0], P.nr[1], P.nr[2], P.prec, P.full};
    // held through the timing, so plans of one shape created concurrently
    // wait for the first one's choice
    std::lock_guard<std::mutex> lk(mu);
    auto it = chosen.find(key);
    if (it != chosen.end()) {
        P.var_r2c = it->second[0];
        P.var_c2r = it->second[1];
        for (int d = 0; d < 3; ++d) P.var_cols[d] = it->second[2 + d];
        if (times)
            for (int i = 0; i < 4 * kColVariants; ++i) times[i] = -1.0f;
        return;
    }
    if (P.prec == OFDFT_FP64)
        autotune_t<double>(P, s, times);
    else
        autotune_t<float>(P, s, times);
    chosen.emplace(key, std::array<int, 5>{P.var_r2c, P.var_c2r, P.var_cols[0], P.var_cols[1], P.var_cols[2]});
}

template void bench_rows_r2c<float>(Plan &, const float *, float2 *, int, const Prologue *, hipStream_t);
template void bench_rows_r2c<double>(Plan &, const double *, double2 *, int, const Prologue *, hipStream_t);
template void bench_rows_c2r<float>(Plan &, const float2 *, float *, int, hipStream_t);
template void bench_rows_c2r<double>(Plan &, const double2 *, double *, int, hipStream_t);

#define INST(T)                                                                                               \
    template void fft_cols<T>(Plan &, int, bool, bool, const cpx<T> *, cpx<T> *, int, long long, double,       \
                              hipStream_t, long long);                                                        \
    template void fft_r2c<T>(Plan &, const T *, cpx<T> *, int, long long, long long, double, const Prologue *, \
                             double *, hipStream_t, long long, bool);                                               \
    template void fft_c2r<T>(Plan &, const cpx<T> *, cpx<T> *, T *, int, long long, long long, long long,      \
                             double, hipStream_t, long long, bool);                                                 \
    template void fft_c2c<T>(Plan &, const cpx<T> *, cpx<T> *, bool, int, long long, double, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace ofdft
