// Size-specialised FFT passes for the common 2/3/5-smooth lengths (kernels and
// launch templates; instantiated per precision in fft_fixed_f32.hip and
// fft_fixed_f64.hip).
//
// The length N, the tile and the radix sequence are compile-time, so all
// index arithmetic folds to shifts and constant multiplies.  The first radix
// stage reads its inputs straight from HBM into registers and the last stage
// (column passes, C2R rows) writes straight back, both coalesced across the
// `TILE` adjacent pencils of a workgroup; intermediate stages exchange
// through LDS.  Column passes address a flattened set of pencils
// c in [0, ncols): base(c) = (c / P) * Q + c % P, element stride S, which
// covers the y pass (P = nzc, Q = ny*nzc, S = nzc) and the x pass
// (P = Q = S = ny*nzc) without partial tiles at plane boundaries.
#pragma once
#include <algorithm>
#include <type_traits>

#include "dft.h"
#include "gspace.h"
#include "internal.h"
#include "pointwise.h"

namespace ofdft {

struct StageList {
    int nst;
    int rad[kMaxStages];
    int rmin;
};

// radices 8/4/2 for the power of two (8s first, 4 4 instead of 8 2), then 5, 3
constexpr StageList factorize(int n) {
    StageList s{};
    s.nst = 0;
    s.rmin = 64;
    int m = n, e2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++e2;
    }
    int r8 = 0, r4 = 0, r2 = 0;
    if (e2 % 3 == 0) {
        r8 = e2 / 3;
    } else if (e2 % 3 == 1) {
        if (e2 >= 4) {
            r8 = e2 / 3 - 1;
            r4 = 2;
        } else {
            r2 = 1;
        }
    } else {
        r8 = e2 / 3;
        r4 = 1;
    }
    for (int i = 0; i < r8; ++i) s.rad[s.nst++] = 8;
    for (int i = 0; i < r4; ++i) s.rad[s.nst++] = 4;
    for (int i = 0; i < r2; ++i) s.rad[s.nst++] = 2;
    while (m % 5 == 0) {
        m /= 5;
        s.rad[s.nst++] = 5;
    }
    while (m % 3 == 0) {
        m /= 3;
        s.rad[s.nst++] = 3;
    }
    for (int i = 0; i < s.nst; ++i) s.rmin = s.rad[i] < s.rmin ? s.rad[i] : s.rmin;
    if (m != 1) s.nst = -1;
    return s;
}

// radix order ORD: 0 as factorize lists it (8s first); 1 reversed (the odd
// radices first), a variant of the column passes the autotuner times
constexpr StageList factorize_ord(int n, int ord) {
    StageList s = factorize(n);
    if (ord == 1)
        for (int i = 0, k = s.nst - 1; i < k; ++i, --k) {
            const int t = s.rad[i];
            s.rad[i] = s.rad[k];
            s.rad[k] = t;
        }
    return s;
}

template <int N, int ORD = 0> struct SL {
    static constexpr StageList s = factorize_ord(N, ORD);
    static constexpr int nst = s.nst;
    static constexpr int rmin = s.rmin;
    static constexpr int tpp = N / s.rmin;  // threads per pencil
    template <int I> static constexpr int rad() { return s.rad[I]; }
};

template <typename C> struct Real;
template <> struct Real<float2> { using type = float; };
template <> struct Real<double2> { using type = double; };

// ---------------------------------------------------------------- column pass
// One stage for TILE pencils, pencil element e of pencil p at buf[e*TILE + p].
template <typename C, int N, int TILE, int R, int NS, bool INV, bool FIRST, bool LAST>
__device__ __forceinline__ void cstage(C *buf, const C *tw, const C *__restrict__ src, C *__restrict__ dst,
                                       long long S, bool valid, int pen, int j, typename Real<C>::type scale) {
    using T = typename Real<C>::type;
    constexpr int NB = N / R;
    const bool act = j < NB;
    C v[R];
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (FIRST)
                v[r] = valid ? src[(long long)(j + r * NB) * S] : C{(T)0, (T)0};
            else
                v[r] = buf[(j + r * NB) * TILE + pen];
        }
        if (NS > 1) {
            const int k = j % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[(r * k * (N / (NS * R))) % N];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
    }
    if (!FIRST || LAST) __syncthreads();
    if (act) {
        DFT<R, INV, C>::run(v);
        const int k = j % NS;
        const int d = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (LAST) {
                if (valid) dst[(long long)(d + r * NS) * S] = C{v[r].x * scale, v[r].y * scale};
            } else {
                buf[(d + r * NS) * TILE + pen] = v[r];
            }
        }
    }
    if (!LAST) __syncthreads();
}

template <typename C, int N, int TILE, bool INV, int I, int NS, int ORD>
__device__ __forceinline__ void cstages(C *buf, const C *tw, const C *src, C *dst, long long S, bool valid, int pen,
                                        int j, typename Real<C>::type scale) {
    if constexpr (I < SL<N, ORD>::nst) {
        constexpr int R = SL<N, ORD>::template rad<I>();
        cstage<C, N, TILE, R, NS, INV, I == 0, I == SL<N, ORD>::nst - 1>(buf, tw, src, dst, S, valid, pen, j, scale);
        cstages<C, N, TILE, INV, I + 1, NS * R, ORD>(buf, tw, src, dst, S, valid, pen, j, scale);
    }
}

template <typename T, int N, int TILE, bool INV, int ORD>
__global__ void __launch_bounds__(TILE *SL<N>::tpp)
    k_cols_f(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, long long bs, long long ncols, long long P,
             long long Q, long long S, long long vp, const cpx<T> *__restrict__ twg, T scale, const double *gate) {
    if (gate_off(gate)) return;
    using C = cpx<T>;
    __shared__ C tw[N];
    __shared__ C buf[N * TILE];
    const int pen = threadIdx.x % TILE;
    const int j = threadIdx.x / TILE;
    for (int i = threadIdx.x; i < N; i += TILE * SL<N>::tpp) tw[i] = twg[i];
    const long long c = (long long)blockIdx.x * TILE + pen;
    const long long cp = c % P;
    const bool valid = c < ncols && cp < vp;  // pad columns of a pitched layout are skipped
    const long long base = valid ? (c / P) * Q + cp : 0;
    const long long off = (long long)blockIdx.y * bs + base;
    cstages<C, N, TILE, INV, 0, 1, ORD>(buf, tw, in + off, out + off, S, valid, pen, j, scale);
}

// Stages with a choice of source and destination: GIN reads the first stage
// from HBM (else from LDS), GOUT writes the last stage to HBM (else to LDS in
// natural order).
template <typename C, int N, int TILE, bool INV, int I, int NS, bool GIN, bool GOUT, int ORD>
__device__ __forceinline__ void cstages_io(C *buf, const C *tw, const C *src, C *dst, long long S, bool valid,
                                           int pen, int j, typename Real<C>::type scale) {
    if constexpr (I < SL<N, ORD>::nst) {
        constexpr int R = SL<N, ORD>::template rad<I>();
        cstage<C, N, TILE, R, NS, INV, GIN && I == 0, GOUT && I == SL<N, ORD>::nst - 1>(buf, tw, src, dst, S, valid,
                                                                                         pen, j, scale);
        cstages_io<C, N, TILE, INV, I + 1, NS * R, GIN, GOUT, ORD>(buf, tw, src, dst, S, valid, pen, j, scale);
    }
}

// x pass of the evaluation pipeline with the G-space multiply fused in: the
// forward x transform of a column tile ends in LDS, each element is scaled by
// its kernel value (gfactor: vW, WT or Hartree by spectrum slot), and the
// inverse x transform starts from LDS and writes back in place.  Saves one
// HBM round trip of every spectrum on each side of the multiply.
template <typename T, int N, int TILE, int ORD>
__global__ void __launch_bounds__(TILE *SL<N>::tpp)
    k_xfused_f(cpx<T> *__restrict__ data, long long bs, long long ncols, long long pitch, long long vp,
               const cpx<T> *__restrict__ twg, GParams gp, const double *gate) {
    if (gate_off(gate)) return;
    using C = cpx<T>;
    __shared__ C tw[N];
    __shared__ C buf[N * TILE];
    constexpr int NT = TILE * SL<N>::tpp;
    const int pen = threadIdx.x % TILE;
    const int j = threadIdx.x / TILE;
    for (int i = threadIdx.x; i < N; i += NT) tw[i] = twg[i];
    const long long c = (long long)blockIdx.x * TILE + pen;
    const bool valid = c < ncols && c % pitch < vp;
    C *col = data + (long long)blockIdx.y * bs + (valid ? c : 0);
    const long long S = ncols;  // x stride of the (nx, ny, pitch) layout
    double tkf = 1.0, wtfac = 1.0;
    const int kind = gp.kind[blockIdx.y];
    if (kind == 1) wt_consts(gp, tkf, wtfac);
    cstages_io<C, N, TILE, false, 0, 1, true, false, ORD>(buf, tw, col, nullptr, S, valid, pen, j, (T)1);
    // multiply: this thread's pencil (pen) is one (iy, iz) column, its
    // elements are e = j, j + tpp, ...; |G|^2 is built from the column's
    // fixed part plus fx * b1 in the plan precision
    if (valid) {
        const int iy = (int)(c / pitch), iz = (int)(c - (long long)iy * pitch);
        constexpr int TPP = SL<N>::tpp;
        constexpr int PER = N / TPP;
        if (gp.full) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int e = j + k * TPP;
                const double f = gfactor<T>(gp, kind, e, iy, iz, tkf, wtfac);
                C v = buf[e * TILE + pen];
                v.x = (T)(v.x * f);
                v.y = (T)(v.y * f);
                buf[e * TILE + pen] = v;
            }
        } else {
            const int fy = ffreq(iy, gp.ny);
            const T cx = (T)(fy * gp.B[3] + iz * gp.B[6]);
            const T cy = (T)(fy * gp.B[4] + iz * gp.B[7]);
            const T cz = (T)(fy * gp.B[5] + iz * gp.B[8]);
            const T bx = (T)gp.B[0], by = (T)gp.B[1], bz = (T)gp.B[2];
            const T sc = (T)gp.scale;
            const T itkf = (T)(1.0 / tkf), wsc = (T)(wtfac * gp.scale), lb = (T)gp.lbda;
            const T sig = (T)(gp.sigma * gp.sigma / 4.0);
            const bool yz0 = iy == 0 && iz == 0;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int e = j + k * TPP;
                const T fx = (T)ffreq(e, N);
                const T gx = fx * bx + cx, gy = fx * by + cy, gz = fx * bz + cz;
                const T gg = gx * gx + gy * gy + gz * gz;
                T f;
                if (kind == 0) {
                    f = (T)0.5 * gg * sc;
                    if (gp.sigma > 0) f *= exp(-gg * sig);
                } else if (kind == 1) {
                    f = lindhard_t<T>(fm::sqrt(gg) * itkf, lb, (T)1) * wsc;
                } else if (kind == 3) {
                    f = (T)(radial_table(gp, sqrt((double)gg)) * gp.scale);
                } else {
                    f = (yz0 && e == 0) ? (T)0 : fm::div((T)(4.0 * M_PI) * sc, gg);
                }
                C v = buf[e * TILE + pen];
                v.x *= f;
                v.y *= f;
                buf[e * TILE + pen] = v;
            }
        }
    }
    __syncthreads();
    cstages_io<C, N, TILE, true, 0, 1, false, true, ORD>(buf, tw, nullptr, col, S, valid, pen, j, (T)1);
}

// ---------------------------------------------------------------- rows passes
// LDS stage for contiguous rows: pencil p element e at buf[p*LD + e].
template <typename C, int N, int LD, int NPEN, int R, int NS, bool INV>
__device__ __forceinline__ void rstage(C *buf, const C *tw, int tid) {
    constexpr int NB = N / R;
    const int p = tid / NB;  // pencil
    const int j = tid % NB;
    const bool act = p < NPEN;
    C v[R];
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = buf[p * LD + j + r * NB];
        if (NS > 1) {
            const int k = j % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[(r * k * (N / (NS * R))) % N];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
    }
    __syncthreads();
    if (act) {
        DFT<R, INV, C>::run(v);
        const int k = j % NS;
        const int d = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[p * LD + d + r * NS] = v[r];
    }
    __syncthreads();
}

template <typename C, int N, int LD, int NPEN, bool INV, int I, int NS>
__device__ __forceinline__ void rstages(C *buf, const C *tw, int tid) {
    if constexpr (I < SL<N>::nst) {
        constexpr int R = SL<N>::template rad<I>();
        rstage<C, N, LD, NPEN, R, NS, INV>(buf, tw, tid);
        rstages<C, N, LD, NPEN, INV, I + 1, NS * R>(buf, tw, tid);
    }
}

// Prologue of one spectrum of the R2C pass.  AUX and B56 are fixed per
// launch (block-uniform branches select the instance), so the compiler never
// evaluates the linear-response cases or the general power speculatively for
// the element-wise ops that do not need them.  Without AUX the
// linear-response codes (which need rho) cannot occur: the host rejects them.
template <typename T, bool AUX, bool B56> __device__ __forceinline__ T prologue_t(int code, T x, T beta, T r) {
    switch (code) {
        case OP_SQRT_RHOM: return fm::sqrt(x > (T)0 ? x : (T)1e-30);
        case OP_POW_BETA: {
            const T xc = x < (T)1e-30 ? (T)1e-30 : x;
            if constexpr (B56)
                return fm::sqrt(xc) * fm::cbrt(xc);
            else
                return fm::pow(xc, beta);
        }
        case OP_DSQRT:
            if constexpr (AUX)
                return fm::div(x, (T)2 * fm::sqrt(r > (T)0 ? r : (T)1e-30));
            else
                return x;
        case OP_DPOW_BETA:
            if constexpr (AUX) {
                const T rc = r < (T)1e-30 ? (T)1e-30 : r;
                if constexpr (B56)
                    return beta * fm::div(fm::cbrt(rc), fm::sqrt(rc)) * x;
                else
                    return beta * fm::pow(rc, beta - (T)1) * x;
            } else {
                return x;
            }
        default: return x;
    }
}

// Rows R2C for NF spectra at once: NPEN packed pairs of rows per spectrum,
// so the NF*NPEN pencils run through the same LDS stages.  When all spectra
// transform the same input (in_bs == 0, each with its own prologue) the input
// rows are read once.  Output rows have pitch opitch >= N/2+1.
//
// Every load of the tile is issued at once, before the twiddles, with
// addresses clamped into the tile instead of branches (the rows past the end
// of a partial tile are zeroed afterwards), so a workgroup waits for one
// memory round trip rather than one per element.  Consecutive lanes take
// consecutive elements: the loads and the LDS writes stay unit-stride.
// partials[blockIdx.y] receives the FP64 sum of slot sum_slot over the tile.
template <typename T, int N, int NPEN, int NF>
__global__ void __launch_bounds__(NF *NPEN *SL<N>::tpp)
    k_rows_r2c_f(const T *__restrict__ in, cpx<T> *__restrict__ out, long long in_bs, long long out_bs, long long nrows,
                 long long opitch, const cpx<T> *__restrict__ twg, T scale, Prologue pro, double *partials,
                 int sum_slot, const double *gate) {
    if (gate_off(gate)) return;
    using C = cpx<T>;
    constexpr int NT = NF * NPEN * SL<N>::tpp;
    constexpr int ROWS = 2 * NPEN;
    constexpr int NZC = N / 2 + 1;
    constexpr int LD = N + 1;
    constexpr int NE = ROWS * N;
    constexpr int PER = (NE + NT - 1) / NT;
    __shared__ C tw[N];
    __shared__ C buf[NF * NPEN * LD];
    __shared__ double red[16];
    const int slot0 = blockIdx.x * NF;
    const long long row0 = (long long)blockIdx.y * ROWS;
    const int nv = (int)min((long long)ROWS, nrows - row0);
    const int lim = nv * N;
    const T beta = (T)pro.beta;
    const bool b56 = pro.b56 != 0;
    const T *aux = reinterpret_cast<const T *>(pro.aux);
    const T *phi = reinterpret_cast<const T *>(pro.phi);
    const bool sep = in_bs != 0;
    const bool do_sum = partials != nullptr && sum_slot >= slot0 && sum_slot < slot0 + NF;
    const T *src = in + slot0 * in_bs + row0 * N;
    T x0[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = threadIdx.x + i * NT;
        x0[i] = src[e < lim ? e : lim - 1];
    }
    C twv = (threadIdx.x < N) ? twg[threadIdx.x] : C{};
    T pv[PER], av[PER];
    if (phi) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * NT;
            pv[i] = phi[row0 * N + (e < lim ? e : lim - 1)];
        }
    }
    if (aux) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * NT;
            av[i] = aux[row0 * N + (e < lim ? e : lim - 1)];
        }
    }
    for (int i = threadIdx.x + NT; i < N; i += NT) tw[i] = twg[i];
    if (threadIdx.x < N) tw[threadIdx.x] = twv;
    double acc = 0.0;
    auto write = [&](auto aux_c, auto b56_c) {
        constexpr bool AUX = decltype(aux_c)::value, B56 = decltype(b56_c)::value;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = threadIdx.x + i * NT;
            if (NE % NT == 0 || e < NE) {
                const int r = e / N;
                const int pos = e - r * N;
                const bool live = e < lim;
                T xa = live ? x0[i] : (T)0;
                if (live && phi) xa = (T)(2.0 * (double)pv[i] * (double)xa);
                const T a = AUX ? av[i] : (T)0;
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    T x = xa;
                    if (f > 0 && sep && live) x = in[(long long)(slot0 + f) * in_bs + row0 * N + e];
                    if (do_sum && slot0 + f == sum_slot) acc += (double)x;
                    if (live) x = prologue_t<T, AUX, B56>(pro.code[(slot0 + f) & 3], x, beta, a);
                    reinterpret_cast<T *>(&buf[(f * NPEN + (r >> 1)) * LD + pos])[r & 1] = x;
                }
            }
        }
    };
    if (aux) {
        if (b56)
            write(std::true_type{}, std::true_type{});
        else
            write(std::true_type{}, std::false_type{});
    } else {
        if (b56)
            write(std::false_type{}, std::true_type{});
        else
            write(std::false_type{}, std::false_type{});
    }
    if (do_sum) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (do_sum && threadIdx.x == 0) {
        double sm = 0.0;
        for (int w = 0; w < (NT + 63) / 64; ++w) sm += red[w];
        partials[blockIdx.y] = sm;
    }
    rstages<C, N, LD, NF * NPEN, false, 0, 1>(buf, tw, (int)threadIdx.x);
    for (int e = threadIdx.x; e < NF * ROWS * NZC; e += NT) {
        const int fr = e / NZC, k = e - fr * NZC;
        const int f = fr / ROWS, r = fr - f * ROWS;
        if (r < nv) {
            const int p = f * NPEN + (r >> 1);
            const C z = buf[p * LD + k];
            C zc = buf[p * LD + (k == 0 ? 0 : N - k)];
            zc.y = -zc.y;
            C o;
            if ((r & 1) == 0)
                o = C{(z.x + zc.x) * (T)0.5, (z.y + zc.y) * (T)0.5};
            else
                o = C{(z.y - zc.y) * (T)0.5, -(z.x - zc.x) * (T)0.5};
            out[(long long)(slot0 + f) * out_bs + (row0 + r) * opitch + k] = C{o.x * scale, o.y * scale};
        }
    }
}

// Rows C2R (irfft semantics: imaginary parts of k = 0 and Nyquist ignored).
// Each thread loads column k of a row pair (A of row 2p, B of row 2p+1) and
// writes A + iB to column k and its Hermitian image conj(A) + i conj(B) to
// column N - k of the packed pencil, so the half spectrum goes from registers
// to the pencil in one LDS pass.  All loads of the tile are issued at once
// with clamped addresses.
template <typename T, int N, int NPEN>
__global__ void __launch_bounds__(NPEN *SL<N>::tpp)
    k_rows_c2r_f(const cpx<T> *__restrict__ in, T *__restrict__ out, long long in_bs, long long out_bs,
                 long long nrows, long long ipitch, const cpx<T> *__restrict__ twg, T scale, const double *gate) {
    if (gate_off(gate)) return;
    using C = cpx<T>;
    constexpr int NT = NPEN * SL<N>::tpp;
    constexpr int ROWS = 2 * NPEN;
    constexpr int NZC = N / 2 + 1;
    constexpr int LD = N + 1;
    constexpr int NW = NPEN * NZC;
    constexpr int PER = (NW + NT - 1) / NT;
    __shared__ C tw[N];
    __shared__ C buf[NPEN * LD];
    const long long row0 = (long long)blockIdx.y * ROWS;
    const C *src = in + (long long)blockIdx.x * in_bs + row0 * ipitch;
    T *dst = out + (long long)blockIdx.x * out_bs + row0 * N;
    const int nv = (int)min((long long)ROWS, nrows - row0);
    C ra[PER], rb[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int w = threadIdx.x + i * NT;
        const int wc = w < NW ? w : NW - 1;
        const int p = wc / NZC, k = wc - p * NZC;
        const int r0 = 2 * p < nv ? 2 * p : nv - 1;
        const int r1 = 2 * p + 1 < nv ? 2 * p + 1 : nv - 1;
        ra[i] = src[r0 * ipitch + k];
        rb[i] = src[r1 * ipitch + k];
    }
    C twv = (threadIdx.x < N) ? twg[threadIdx.x] : C{};
    for (int i = threadIdx.x + NT; i < N; i += NT) tw[i] = twg[i];
    if (threadIdx.x < N) tw[threadIdx.x] = twv;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int w = threadIdx.x + i * NT;
        if (NW % NT == 0 || w < NW) {
            const int p = w / NZC, k = w - p * NZC;
            C A = 2 * p < nv ? ra[i] : C{};
            C B = 2 * p + 1 < nv ? rb[i] : C{};
            if (k == 0 || 2 * k == N) {
                A.y = (T)0;
                B.y = (T)0;
            }
            buf[p * LD + k] = C{A.x - B.y, A.y + B.x};
            if (k > 0 && 2 * k < N) buf[p * LD + N - k] = C{A.x + B.y, B.x - A.y};
        }
    }
    __syncthreads();
    rstages<C, N, LD, NPEN, true, 0, 1>(buf, tw, (int)threadIdx.x);
    for (int e = threadIdx.x; e < nv * N; e += NT) {
        const int r = e / N;
        const C z = buf[(r >> 1) * LD + (e - r * N)];
        dst[e] = ((r & 1) ? z.y : z.x) * scale;
    }
}

// ---------------------------------------------------------------- dispatch
// Tile of the column pass: as many pencils as fit 1024 threads, power of 2,
// at most 32 (FP32) / 16 (FP64), LDS <= 64 KB; variant v halves it v times.
template <typename T, int N> constexpr int col_tile() {
    int t = sizeof(T) == 4 ? 32 : 16;
    while (t > 1 && (t * SL<N>::tpp > 1024 || (t + 1) * N * 2 * (int)sizeof(T) > 64 * 1024)) t /= 2;
    return t;
}
template <typename T, int N, int V> constexpr int col_tile_v() {
    const int t = col_tile<T, N>() >> V;
    return t < 1 ? 1 : t;
}
template <typename T, int N> constexpr int row_pens() {
    int p = 16;
    while (p > 1 && (p * SL<N>::tpp > 1024 || (p * (N + 1) + N) * 2 * (int)sizeof(T) > 64 * 1024)) --p;
    return p;
}
template <typename T, int N, int V> constexpr int row_pens_v() {
    const int p = row_pens<T, N>() >> V;
    return p < 1 ? 1 : p;
}
template <typename T, int N, int NF> constexpr int r2c_pens() {
    int p = 8;
    while (p > 1 && (NF * p * SL<N>::tpp > 1024 || (NF * p * (N + 1) + N) * 2 * (int)sizeof(T) + 128 > 64 * 1024))
        --p;
    return p;
}
template <typename T, int N, int NF, int V> constexpr int r2c_pens_v() {
    const int p = r2c_pens<T, N, NF>() >> V;
    return p < 1 ? 1 : p;
}

#define OFDFT_VARIANT(var, CALL) \
    switch (var) {               \
        case 1: CALL(1); break;  \
        case 2: CALL(2); break;  \
        default: CALL(0); break; \
    }
#define OFDFT_COL_VARIANT(var, CALL) \
    switch (var) {                   \
        case 1: CALL(1, 0); break;   \
        case 2: CALL(2, 0); break;   \
        case 3: CALL(0, 1); break;   \
        case 4: CALL(1, 1); break;   \
        case 5: CALL(2, 1); break;   \
        default: CALL(0, 0); break;  \
    }

template <typename T, int N>
static void launch_cols(int var, bool inv, const cpx<T> *in, cpx<T> *out, int batch, long long bs, long long ncols,
                        long long P, long long Q, long long S, long long vp, const cpx<T> *tw, T scale, hipStream_t s,
                        const double *gate) {
#define CALL(V, O)                                                                                                 \
    {                                                                                                              \
        constexpr int TILE = col_tile_v<T, N, V>();                                                                \
        dim3 grid((unsigned)((ncols + TILE - 1) / TILE), (unsigned)batch);                                         \
        if (inv)                                                                                                   \
            hipLaunchKernelGGL((k_cols_f<T, N, TILE, true, O>), grid, dim3(TILE * SL<N>::tpp), 0, s, in, out, bs,   \
                               ncols, P, Q, S, vp, tw, scale, gate);                                               \
        else                                                                                                       \
            hipLaunchKernelGGL((k_cols_f<T, N, TILE, false, O>), grid, dim3(TILE * SL<N>::tpp), 0, s, in, out, bs,  \
                               ncols, P, Q, S, vp, tw, scale, gate);                                               \
    }
    OFDFT_COL_VARIANT(var, CALL)
#undef CALL
}

template <typename T, int N, int NF>
static void launch_r2c_nf(int var, const T *in, cpx<T> *out, int nblk, long long in_bs, long long out_bs,
                          long long nrows, long long opitch, const cpx<T> *tw, T scale, const Prologue &pro,
                          double *partials, int sum_slot, hipStream_t s, const double *gate) {
#define CALL(V)                                                                                                \
    {                                                                                                          \
        constexpr int NP = r2c_pens_v<T, N, NF, V>();                                                          \
        const long long tiles = (nrows + 2 * NP - 1) / (2 * NP);                                               \
        OFDFT_REQUIRE(tiles <= 65535, "too many rows for the R2C pass");                                       \
        dim3 grid((unsigned)nblk, (unsigned)tiles);                                                            \
        hipLaunchKernelGGL((k_rows_r2c_f<T, N, NP, NF>), grid, dim3(NF * NP * SL<N>::tpp), 0, s, in, out, in_bs, \
                           out_bs, nrows, opitch, tw, scale, pro, partials, sum_slot, gate);                   \
    }
    OFDFT_VARIANT(var, CALL)
#undef CALL
}

template <typename T, int N>
static void launch_r2c(int var, const T *in, cpx<T> *out, int batch, long long in_bs, long long out_bs,
                       long long nrows, long long opitch, const cpx<T> *tw, T scale, const Prologue &pro,
                       double *partials, int sum_slot, hipStream_t s, const double *gate) {
    if (batch == 3)
        launch_r2c_nf<T, N, 3>(var, in, out, 1, in_bs, out_bs, nrows, opitch, tw, scale, pro, partials, sum_slot, s,
                               gate);
    else if (batch == 2)
        launch_r2c_nf<T, N, 2>(var, in, out, 1, in_bs, out_bs, nrows, opitch, tw, scale, pro, partials, sum_slot, s,
                               gate);
    else
        launch_r2c_nf<T, N, 1>(var, in, out, batch, in_bs, out_bs, nrows, opitch, tw, scale, pro, partials, sum_slot,
                               s, gate);
}

template <typename T, int N>
static void launch_c2r(int var, const cpx<T> *in, T *out, int batch, long long in_bs, long long out_bs,
                       long long nrows, long long ipitch, const cpx<T> *tw, T scale, hipStream_t s,
                       const double *gate) {
#define CALL(V)                                                                                                \
    {                                                                                                          \
        constexpr int NP = row_pens_v<T, N, V>();                                                              \
        const long long tiles = (nrows + 2 * NP - 1) / (2 * NP);                                               \
        OFDFT_REQUIRE(tiles <= 65535, "too many rows for the C2R pass");                                       \
        dim3 grid((unsigned)batch, (unsigned)tiles);                                                           \
        hipLaunchKernelGGL((k_rows_c2r_f<T, N, NP>), grid, dim3(NP * SL<N>::tpp), 0, s, in, out, in_bs, out_bs, \
                           nrows, ipitch, tw, scale, gate);                                                    \
    }
    OFDFT_VARIANT(var, CALL)
#undef CALL
}

template <typename T, int N>
static void launch_xfused(int var, cpx<T> *data, int batch, long long bs, long long ncols, long long pitch,
                          long long vp, const cpx<T> *tw, const GParams &gp, hipStream_t s, const double *gate) {
#define CALL(V, O)                                                                                                 \
    {                                                                                                              \
        constexpr int TILE = col_tile_v<T, N, V>();                                                                \
        dim3 grid((unsigned)((ncols + TILE - 1) / TILE), (unsigned)batch);                                         \
        hipLaunchKernelGGL((k_xfused_f<T, N, TILE, O>), grid, dim3(TILE * SL<N>::tpp), 0, s, data, bs, ncols, pitch, \
                           vp, tw, gp, gate);                                                                      \
    }
    OFDFT_COL_VARIANT(var, CALL)
#undef CALL
}

#define OFDFT_FIXED_SIZES(X) X(32) X(48) X(64) X(72) X(80) X(96) X(100) X(120) X(128) X(144) X(160) X(192) X(200) X(240) X(256)

template <typename T>
bool xfused_fixed(int n, int var, cpx<T> *data, int batch, long long bs, long long ncols, long long pitch,
                  long long vp, const void *tw, const GParams &gp, hipStream_t s, const double *gate) {
    switch (n) {
#define CASE(N)                                                                                                       \
    case N:                                                                                                           \
        launch_xfused<T, N>(var, data, batch, bs, ncols, pitch, vp, reinterpret_cast<const cpx<T> *>(tw), gp, s,       \
                            gate);                                                                                    \
        return true;
        OFDFT_FIXED_SIZES(CASE)
#undef CASE
        default: return false;
    }
}

template <typename T>
bool cols_fixed(int n, int var, bool inv, const cpx<T> *in, cpx<T> *out, int batch, long long bs, long long ncols,
                long long P, long long Q, long long S, long long vp, const void *tw, double scale, hipStream_t s,
                const double *gate) {
    switch (n) {
#define CASE(N)                                                                                                        \
    case N:                                                                                                            \
        launch_cols<T, N>(var, inv, in, out, batch, bs, ncols, P, Q, S, vp, reinterpret_cast<const cpx<T> *>(tw),      \
                          (T)scale, s, gate);                                                                         \
        return true;
        OFDFT_FIXED_SIZES(CASE)
#undef CASE
        default: return false;
    }
}

template <typename T>
bool r2c_fixed(int n, int var, const T *in, cpx<T> *out, int batch, long long in_bs, long long out_bs,
               long long nrows, long long opitch, const void *tw, double scale, const Prologue &pro, double *partials,
               int sum_slot, hipStream_t s, const double *gate) {
    switch (n) {
#define CASE(N)                                                                                                       \
    case N:                                                                                                           \
        launch_r2c<T, N>(var, in, out, batch, in_bs, out_bs, nrows, opitch, reinterpret_cast<const cpx<T> *>(tw),     \
                         (T)scale, pro, partials, sum_slot, s, gate);                                                 \
        return true;
        OFDFT_FIXED_SIZES(CASE)
#undef CASE
        default: return false;
    }
}

template <typename T>
bool c2r_fixed(int n, int var, const cpx<T> *in, T *out, int batch, long long in_bs, long long out_bs,
               long long nrows, long long ipitch, const void *tw, double scale, hipStream_t s, const double *gate) {
    switch (n) {
#define CASE(N)                                                                                                       \
    case N:                                                                                                           \
        launch_c2r<T, N>(var, in, out, batch, in_bs, out_bs, nrows, ipitch, reinterpret_cast<const cpx<T> *>(tw),     \
                         (T)scale, s, gate);                                                                          \
        return true;
        OFDFT_FIXED_SIZES(CASE)
#undef CASE
        default: return false;
    }
}

#define OFDFT_FIXED_INST(T)                                                                                            \
    template bool xfused_fixed<T>(int, int, cpx<T> *, int, long long, long long, long long, long long, const void *,   \
                                  const GParams &, hipStream_t, const double *);                                       \
    template bool cols_fixed<T>(int, int, bool, const cpx<T> *, cpx<T> *, int, long long, long long, long long,        \
                                long long, long long, long long, const void *, double, hipStream_t, const double *);   \
    template bool r2c_fixed<T>(int, int, const T *, cpx<T> *, int, long long, long long, long long, long long,         \
                               const void *, double, const Prologue &, double *, int, hipStream_t, const double *);    \
    template bool c2r_fixed<T>(int, int, const cpx<T> *, T *, int, long long, long long, long long, long long,         \
                               const void *, double, hipStream_t, const double *);

}  // namespace ofdft
