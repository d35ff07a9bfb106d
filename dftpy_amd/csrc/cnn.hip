// Learned warm start: a periodic residual 3D CNN on the density grid.
//
// No counterpart in the reference.  The network maps a cheap start density
// (superposition features) to a log-density correction; the optimizer of
// optimization.py:281-520 then runs from the corrected density.
//
// Layout: hidden fields are channels-last, (N points) x 32 channels, bfloat16
// (64 bytes per point); the two input channels and the output are planar FP32
// fields.  Tap t = (i*3 + j)*3 + k reads the input at p + d*(i-1, j-1, k-1)
// with periodic wrap (d = dilation of the layer).
//
//   h1 = relu(conv(x0; W1) + b1)                          FP16 MFMA (K = 54)
//   h_l = h_{l-1} + relu(conv(h_{l-1}; W_l) + b_l), l=2..5  BF16 MFMA, FP32 acc
//   out = conv(h5; W6) + b6                               FP16 MFMA projection
//                                                         onto the 27 taps,
//                                                         summed per point
//
// The 32 -> 32 layers are an implicit GEMM per tap: D[p][co] += X[p+off][ci]
// W[co][ci] with v_mfma_f32_16x16x32_bf16 (16 points x 16 channels x 32 input
// channels).  Backward-data is the same kernel with the weights transposed and
// the taps mirrored; the weight gradient is a second MFMA kernel reducing over
// points (K = points) through LDS-transposed tiles.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "internal.h"
#include "reduce.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

namespace ofdft {

constexpr int CH = 32;
constexpr int NTAP = 27;
constexpr int KMID = NTAP * CH * CH;  // weights of one 32 -> 32 layer
constexpr int KIN = NTAP * 2 * CH;
constexpr int KOUT = NTAP * CH;

struct Geo {
    int nx, ny, nz, d;
    long long n;
};

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// index of p + d*(i-1, j-1, k-1) for the point (x, y, z)
struct Nbr {
    int bx[3], by[3], bz[3];
    __device__ __forceinline__ void init(const Geo &g, int x, int y, int z) {
        const int nyz = g.ny * g.nz;
        bx[0] = wrapi(x - g.d, g.nx) * nyz;
        bx[1] = x * nyz;
        bx[2] = wrapi(x + g.d, g.nx) * nyz;
        by[0] = wrapi(y - g.d, g.ny) * g.nz;
        by[1] = y * g.nz;
        by[2] = wrapi(y + g.d, g.ny) * g.nz;
        bz[0] = wrapi(z - g.d, g.nz);
        bz[1] = z;
        bz[2] = wrapi(z + g.d, g.nz);
    }
    __device__ __forceinline__ int at(int t) const { return bx[t / 9] + by[(t / 3) % 3] + bz[t % 3]; }
};

__device__ __forceinline__ void decompose(const Geo &g, long long p, int &x, int &y, int &z) {
    z = (int)(p % g.nz);
    const long long r = p / g.nz;
    y = (int)(r % g.ny);
    x = (int)(r / g.ny);
}

typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(_Float16)))) _Float16 f16x4;

// ------------------------------------------------------------ layer 1 (2 -> 32)
// MFMA form: K = 27 taps x 2 channels = 54 (zero-padded to 64), FP16
// operands (the inputs are FP32 features; FP16 keeps 11 bits, finer than the
// bf16 the output is stored in) and FP32 accumulation.  The weights are the A
// operand (rows = output channels) and the gathered neighbours the B operand
// (columns = 16 consecutive points), so each lane ends with 4 consecutive
// channels of one point and stores them as 8 bytes.  A wave strides over
// 16-point tiles; the weights stay in registers.
constexpr int KIN_PAD = 64;

__global__ void __launch_bounds__(256) k_cnn_in_mfma(Geo g, const float *__restrict__ x0, const float *__restrict__ w,
                                                     const float *__restrict__ b, __bf16 *__restrict__ h,
                                                     long long ntiles) {
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    // A operands: W1[t][c][channel] at k = 2 t + c, rows = channels nt*16 + m
    f16x8 wa[2][2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = kb * 32 + kq * 8 + e, kc = min(k, 2 * NTAP - 1);
                const float v = w[(kc >> 1) * 64 + (kc & 1) * 32 + nt * 16 + m];  // unconditional load
                wa[kb][nt][e] = k < 2 * NTAP ? (_Float16)v : (_Float16)0.0f;
            }
    float bl[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) bl[nt][j] = b[nt * 16 + kq * 4 + j];
    for (long long tile = wave0; tile < ntiles; tile += nwaves) {
        const long long p = tile * 16 + m;
        const long long pc = p < g.n ? p : g.n - 1;  // tail lanes compute a copy and store nothing
        int x, y, z;
        decompose(g, pc, x, y, z);
        Nbr nb;
        nb.init(g, x, y, z);
        f16x8 xb[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = kb * 32 + kq * 8 + e;
                xb[kb][e] = k < 2 * NTAP ? (_Float16)x0[(k & 1) * g.n + nb.at(k >> 1)] : (_Float16)0.0f;
            }
        floatx4 acc[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wa[kb][nt], xb[kb], acc[nt], 0, 0, 0);
        if (p < g.n) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                bf16x4 o;
#pragma unroll
                for (int j = 0; j < 4; ++j) o[j] = (__bf16)fmaxf(acc[nt][j] + bl[nt][j], 0.0f);
                *reinterpret_cast<bf16x4 *>(h + p * CH + nt * 16 + kq * 4) = o;
            }
        }
    }
}

// Row-staged form for nz % 16 == 0 and ny % (4 d) == 0: a workgroup owns 4
// output rows y0 + k d of one x plane (as the middle-layer kernel) and stages
// the 3 planes x 6 rows x 2 channels of features it reads as FP16 in LDS, so
// the 54 neighbour values of a point come from LDS and no per-point index
// division is needed.
constexpr int IN_ROWS_THREADS = 256;
constexpr int IN_ROWS_MAX_NZ = 512;

// CHUNKS: 16-byte row chunks per lane (nz <= 256 * CHUNKS)
template <int CHUNKS>
__global__ void __launch_bounds__(IN_ROWS_THREADS)
    k_cnn_in_rows(Geo g, const float *__restrict__ x0, const float *__restrict__ w, const float *__restrict__ b,
                  __bf16 *__restrict__ h) {
    extern __shared__ _Float16 xs[];  // [3 planes][6 rows][2 channels][rs]: z = -d .. nz + d - 1 at lp - d ..
    const int nz = g.nz, d = g.d;
    const int lp = (d + 3) & ~3, rs = nz + 2 * lp;  // 8-byte aligned rows with a periodic halo of d each side
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, m = lane & 15, kq = lane >> 4;
    const int x = blockIdx.y, gy = blockIdx.x;
    const int y0 = (gy / d) * 4 * d + gy % d;
    const long long nyz = (long long)g.ny * nz;
    // stage: row q = (i * 6 + r) * 2 + c holds xs[q * nz + z]; wave w takes
    // rows w, w + 4, ..., 16 bytes per lane per load (nz % 16 == 0).  Every
    // load (and the weights' below) is issued before any is used, from a
    // clamped address, so a workgroup waits for one memory round trip
    // instead of one per row
    constexpr int NW = IN_ROWS_THREADS / 64;
    constexpr int RPW = 36 / NW;  // staged rows per wave
    const int nz4 = nz >> 2;
    float4 stv[RPW][CHUNKS];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        const int q = wave + NW * u;
        const int c = q & 1, r = (q >> 1) % 6, i = q / 12;
        const int xp = wrapi(x + (i - 1) * d, g.nx), yr = wrapi(y0 + (r - 1) * d, g.ny);
        const float4 *src = reinterpret_cast<const float4 *>(x0 + (long long)c * g.n + (long long)xp * nyz +
                                                             (long long)yr * nz);
#pragma unroll
        for (int ch = 0; ch < CHUNKS; ++ch) stv[u][ch] = src[min(lane + 64 * ch, nz4 - 1)];
    }
    float wl[2][2][8];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = min(kb * 32 + kq * 8 + e, 2 * NTAP - 1);
                wl[kb][nt][e] = w[(k >> 1) * 64 + (k & 1) * 32 + nt * 16 + m];
            }
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        f16x4 *dst = reinterpret_cast<f16x4 *>(xs + (wave + NW * u) * rs + lp);
#pragma unroll
        for (int ch = 0; ch < CHUNKS; ++ch) {
            const int z4 = lane + 64 * ch;
            const float4 v = stv[u][ch];
            if (z4 < nz4) dst[z4] = f16x4{(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
        }
    }
    f16x8 wa[2][2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int e = 0; e < 8; ++e)
                wa[kb][nt][e] = kb * 32 + kq * 8 + e < 2 * NTAP ? (_Float16)wl[kb][nt][e] : (_Float16)0.0f;
    float bl[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) bl[nt][j] = b[nt * 16 + kq * 4 + j];
    // this lane's 16 operand slots k = 2 t + c: LDS offset of (plane i, row
    // j, channel c, z shift) relative to output row 0 and point z = 0; the
    // 10 padding slots (zero weights) read the centre tap
    int off[2][8];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int kk = kb * 32 + kq * 8 + e;
            const int t = kk < 2 * NTAP ? kk >> 1 : 13, c = kk < 2 * NTAP ? kk & 1 : 0;
            off[kb][e] = (((t / 9) * 6 + (t / 3) % 3) * 2 + c) * rs + lp + (t % 3 - 1) * d;
        }
    // per-wave staging of one output tile (16 points x 32 channels, bf16)
    // for coalesced 16-byte stores; the wave-scope fences around the stage
    // order its LDS accesses only ("local"), so a tile's global stores stay
    // in flight while the next tile computes
    bf16x4 *ot = reinterpret_cast<bf16x4 *>(xs + 36 * rs) + wave * 128;
    __syncthreads();
    // periodic halo: z = -d .. -1 and nz .. nz + d - 1 of every staged row
    for (int q = threadIdx.x; q < 36 * 2 * d; q += IN_ROWS_THREADS) {
        const int row = q / (2 * d), e = q - row * 2 * d;
        _Float16 *r = xs + row * rs + lp;
        if (e < d)
            r[e - d] = r[nz - d + e];
        else
            r[nz + e - d] = r[e - d];
    }
    __syncthreads();
    const int tiles_per_row = nz / 16, ntile = 4 * tiles_per_row;
    for (int tile = wave; tile < ntile; tile += IN_ROWS_THREADS / 64) {
        const int k = tile / tiles_per_row, zt = tile - k * tiles_per_row;
        const _Float16 *xk = xs + k * 2 * rs + zt * 16 + m;
        f16x8 xb[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) xb[kb][e] = xk[off[kb][e]];
        floatx4 acc[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wa[kb][nt], xb[kb], acc[nt], 0, 0, 0);
        // lane (m, kq) holds channels nt*16 + kq*4 .. +3 of point m
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            bf16x4 o;
#pragma unroll
            for (int j = 0; j < 4; ++j) o[j] = (__bf16)fmaxf(acc[nt][j] + bl[nt][j], 0.0f);
            ot[m * 8 + nt * 4 + kq] = o;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        // lane l stores 16 bytes: point l / 4, channels (l % 4) * 8 .. +7
        const long long p0 = (long long)x * nyz + (long long)(y0 + k * d) * nz + zt * 16;
        const bf16x8 v = reinterpret_cast<const bf16x8 *>(ot)[lane];
        reinterpret_cast<bf16x8 *>(h)[(p0 + (lane >> 2)) * 4 + (lane & 3)] = v;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
}


// ------------------------------------------------------- layers 2..5 (MFMA)
// MODE 0 (forward):  r = relu(conv(x) + bias); out0 (bf16) = x + r; out1 (bf16) = r
// MODE 1 (backward): v = conv(x) + resid (fp32); out0 (fp32) = v;
//                    out1 (bf16) = mask > 0 ? v : 0
// Workgroup: 8 waves x 64 points; each wave 4 M-tiles of 16 points x 2 N-tiles
// of 16 output channels.  The weights of the layer (27 x 32 x 32 bf16,
// 55,296 bytes) are staged in LDS.
constexpr int MID_THREADS = 512;
constexpr int MID_POINTS = MID_THREADS;  // 64 points per wave

template <int MODE>
__global__ void __launch_bounds__(MID_THREADS)
    k_cnn_mid(Geo g, const __bf16 *__restrict__ xin, const __bf16 *__restrict__ w, const float *__restrict__ bias,
              const float *__restrict__ resid, const __bf16 *__restrict__ mask, void *__restrict__ out0,
              __bf16 *__restrict__ out1) {
    __shared__ bf16x8 wl[NTAP * CH * 4];
    const bf16x8 *wg = reinterpret_cast<const bf16x8 *>(w);
    for (int i = threadIdx.x; i < NTAP * CH * 4; i += MID_THREADS) wl[i] = wg[i];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const long long pbase = (long long)blockIdx.x * MID_POINTS + wave * 64;
    Nbr nb[4];
    bool valid[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const long long p = pbase + mt * 16 + m;
        valid[mt] = p < g.n;
        int x, y, z;
        decompose(g, valid[mt] ? p : 0, x, y, z);
        nb[mt].init(g, x, y, z);
    }
    __syncthreads();
    floatx4 acc[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt][0] = acc[mt][1] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 *xv = reinterpret_cast<const bf16x8 *>(xin);
    const bf16x8 zero8 = {};
#pragma unroll
    for (int t = 0; t < NTAP; ++t) {
        const bf16x8 b0 = wl[(t * CH + m) * 4 + kq];
        const bf16x8 b1 = wl[(t * CH + 16 + m) * 4 + kq];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const bf16x8 a = valid[mt] ? xv[(long long)nb[mt].at(t) * 4 + kq] : zero8;
            acc[mt][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[mt][0], 0, 0, 0);
            acc[mt][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[mt][1], 0, 0, 0);
        }
    }
    // D[row][col]: lane holds rows kq*4 + i (points), column m (channel)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int co = nt * 16 + m;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long p = pbase + mt * 16 + kq * 4 + i;
                if (p >= g.n) continue;
                const long long e = p * CH + co;
                float v = acc[mt][nt][i];
                if (MODE == 0) {
                    v += bias[co];
                    const float r = fmaxf(v, 0.0f);
                    reinterpret_cast<__bf16 *>(out0)[e] = (__bf16)((float)xin[e] + r);
                    if (out1) out1[e] = (__bf16)r;
                } else {
                    v += resid[e];
                    reinterpret_cast<float *>(out0)[e] = v;
                    out1[e] = (float)mask[e] > 0.0f ? (__bf16)v : (__bf16)0.0f;
                }
            }
        }
    }
}

// Row-tiled variant for nz % 32 == 0, nz <= 160 and ny % (4 d) == 0.  A work
// item is 4 output rows y0 + k d (k = 0..3) of one x plane, all of z.  For
// each x offset (a stage) the 6 input rows y0 + (r-1) d of plane x + dx are
// staged in LDS (6 nz points x 64 bytes) with the 9 taps' weights, so each
// input point is read from HBM/L2 about 4.5 times per layer instead of 27.
// Rows in a residue class mod d keep the dilated y taps inside the staged
// rows.
//
// One wave per 16-point z tile, owning that tile in all 4 output rows: an A
// operand read from staged row r serves the output rows k = r - dy of every
// dy tap (up to 3 rows), so each A fragment feeds up to 6 MFMAs and the
// B operands of one dz column (3 dy x 2 channel halves) stay in registers.
//
// The kernel is persistent (one workgroup per CU) and double-buffers the
// stages in LDS: the global loads of the next stage are in flight while the
// current one is computed, then land in the other buffer.  Work items are
// dealt to the 8 XCDs in contiguous chunks (workgroup b runs on XCD b % 8),
// so the planes an XCD re-reads stay in its own L2.
constexpr int ROWS_MAX_MT = 5;                       // nz = 32 MT <= 160
constexpr int ROWS_MAX_THREADS = 128 * ROWS_MAX_MT;  // 64 lanes x nz / 16 waves
constexpr int ROWS_XCDS = 8;

template <int MODE, int MT>
__global__ void __launch_bounds__(ROWS_MAX_THREADS)
    k_cnn_mid_rows(Geo g, const __bf16 *__restrict__ xin, const __bf16 *__restrict__ w,
                   const float *__restrict__ bias, const float *__restrict__ resid, const __bf16 *__restrict__ mask,
                   void *__restrict__ out0, __bf16 *__restrict__ out1) {
    extern __shared__ bf16x8 lds[];  // 2 x ([6][nz][4] rows + [9][32][4] weights)
    constexpr int NZ = 32 * MT;      // the launcher picks MT = nz / 32
    constexpr int NT = 128 * MT;     // threads: one wave per z tile
    constexpr int NLOAD = 6 * NZ * 4;
    constexpr int NWL = 9 * CH * 4;
    constexpr int BUF = NLOAD + NWL;
    constexpr int PER = (NLOAD + NT - 1) / NT, PERW = (NWL + NT - 1) / NT;
    const int nz = NZ, d = g.d;
    const int lane = threadIdx.x & 63, zt = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const long long nyz = (long long)g.ny * nz;
    const bf16x8 *xv = reinterpret_cast<const bf16x8 *>(xin);
    const bf16x8 *wv = reinterpret_cast<const bf16x8 *>(w);
    // this workgroup's items: slot s of its XCD's chunk, then every
    // (workgroups per XCD)-th item of the chunk
    const int ngroups = g.ny / 4;
    const long long items = (long long)g.nx * ngroups;
    const int nwg = gridDim.x, per_xcd = nwg / ROWS_XCDS;
    const int xcd = blockIdx.x % ROWS_XCDS, slot = blockIdx.x / ROWS_XCDS;
    const long long c0 = items * xcd / ROWS_XCDS, c1 = items * (xcd + 1) / ROWS_XCDS;
    const long long first = c0 + slot;
    const int nmine = first < c1 ? (int)((c1 - 1 - first) / per_xcd + 1) : 0;
    const int nq = 3 * nmine;  // stages: x offsets -d, +d, 0 of each item (centre last)
    auto item_of = [&](int q, int &x, int &y0, int &xp, int &ix) {
        const long long it = first + (long long)(q / 3) * per_xcd;
        const int st = q % 3;
        x = (int)(it / ngroups);
        const int gy = (int)(it - (long long)x * ngroups);  // residue gy % d, block gy / d
        y0 = (gy / d) * 4 * d + gy % d;
        ix = st == 0 ? 0 : (st == 1 ? 2 : 1);
        xp = wrapi(x + (ix - 1) * d, g.nx);
    };
    bf16x8 st_r[PER], st_w[PERW];
    auto fetch = [&](int q) {
        int x, y0, xp, ix;
        item_of(q, x, y0, xp, ix);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + j * NT;
            if (e < NLOAD) {
                const int r = e / (NZ * 4), rem = e - r * NZ * 4;
                const int yr = wrapi(y0 + (r - 1) * d, g.ny);
                st_r[j] = xv[((long long)xp * nyz + (long long)yr * NZ) * 4 + rem];
            }
        }
#pragma unroll
        for (int j = 0; j < PERW; ++j) {
            const int e = threadIdx.x + j * NT;
            if (e < NWL) st_w[j] = wv[ix * NWL + e];
        }
    };
    auto land = [&](int buf) {
        bf16x8 *dst = lds + buf * BUF;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + j * NT;
            if (e < NLOAD) dst[e] = st_r[j];
        }
#pragma unroll
        for (int j = 0; j < PERW; ++j) {
            const int e = threadIdx.x + j * NT;
            if (e < NWL) dst[NLOAD + e] = st_w[j];
        }
    };
    if (nq > 0) {
        fetch(0);
        land(0);
    }
    // the forward bias of this lane's two output channels
    const float bias_lane[2] = {MODE == 0 ? bias[m] : 0.0f, MODE == 0 ? bias[16 + m] : 0.0f};
    __syncthreads();
    floatx4 acc[4][2];
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k][0] = acc[k][1] = floatx4{0.f, 0.f, 0.f, 0.f};
    // the 16-byte stores of a finished item (forward): issued at the top of
    // the next stage, before its loads, so the wait that lands those loads in
    // LDS does not also wait for fresh stores
    auto copy_out = [&](int q) {
        int x, y0, xp, ix;
        item_of(q, x, y0, xp, ix);
        bf16x8 *ov = reinterpret_cast<bf16x8 *>(out0);
        const bf16x8 *src = lds + (q & 1) * BUF;
        for (int e = threadIdx.x; e < 4 * NZ * 4; e += NT) {
            const int k = e / (NZ * 4), rem = e - k * NZ * 4;
            ov[((long long)x * nyz + (long long)(y0 + k * d) * NZ) * 4 + rem] = src[(k + 1) * NZ * 4 + rem];
        }
    };
#pragma unroll 1
    for (int q = 0; q < nq; ++q) {
        const int cur = q & 1;
        if (MODE == 0 && q % 3 == 0 && q > 0) copy_out(q - 1);
        if (q + 1 < nq) fetch(q + 1);
        const bf16x8 *rows = lds + cur * BUF;
        const bf16x8 *wl = rows + NLOAD;
#pragma unroll
        for (int iz = 0; iz < 3; ++iz) {
            bf16x8 b[3][2];
#pragma unroll
            for (int iy = 0; iy < 3; ++iy) {
                const int tt = iy * 3 + iz;
                b[iy][0] = wl[(tt * CH + m) * 4 + kq];
                b[iy][1] = wl[(tt * CH + 16 + m) * 4 + kq];
            }
            int z = zt * 16 + m + (iz - 1) * d;
            z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                const bf16x8 a = rows[(r * nz + z) * 4 + kq];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int iy = r - k;  // output row k reads staged row k + dy
                    if (iy >= 0 && iy < 3) {
                        acc[k][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[iy][0], acc[k][0], 0, 0, 0);
                        acc[k][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[iy][1], acc[k][1], 0, 0, 0);
                    }
                }
            }
        }
        if (q % 3 == 2) {  // the item's last stage (centre plane): write it out
            int x, y0, xp, ix;
            item_of(q, x, y0, xp, ix);
            if (MODE == 0) {
                // forward: out = x + relu(acc + b).  The residual x at an
                // output point is LDS row k + 1 of this (centre) stage; each
                // output replaces its own residual in LDS, then the 4 rows go
                // out as coalesced 16-byte stores
                __syncthreads();  // every wave is done reading the stage
                __bf16 *rs = reinterpret_cast<__bf16 *>(lds + cur * BUF);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        const int co = nt * 16 + m;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int z = zt * 16 + kq * 4 + j;
                            const int li = ((k + 1) * nz + z) * CH + co;
                            const float r = fmaxf(acc[k][nt][j] + bias_lane[nt], 0.0f);
                            if (out1) {
                                const long long e =
                                    ((long long)x * nyz + (long long)(y0 + k * d) * nz + z) * CH + co;
                                out1[e] = (__bf16)r;
                            }
                            rs[li] = (__bf16)((float)rs[li] + r);
                        }
                    }
                }
                // the rows go out at the top of the next stage (copy_out)
            } else {
                // backward-data: out0 = acc + resid (FP32), out1 = out0 where
                // the mask is positive (bf16); per row, resid and mask are
                // loaded before the stores
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const long long rowbase = (long long)x * nyz + (long long)(y0 + k * d) * nz + zt * 16;
                    float rv[2][4];
                    bool keep[2][4];
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const long long e = (rowbase + kq * 4 + j) * CH + nt * 16 + m;
                            rv[nt][j] = resid[e];
                            keep[nt][j] = (float)mask[e] > 0.0f;
                        }
                    }
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const long long e = (rowbase + kq * 4 + j) * CH + nt * 16 + m;
                            const float v = acc[k][nt][j] + rv[nt][j];
                            reinterpret_cast<float *>(out0)[e] = v;
                            out1[e] = keep[nt][j] ? (__bf16)v : (__bf16)0.0f;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k][0] = acc[k][1] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        if (q + 1 < nq) {
            // the buffer landed now is the one copy_out read at the top of
            // this stage: every wave must be past those reads
            if (MODE == 0 && q % 3 == 0 && q > 0) __syncthreads();
            land(cur ^ 1);
        }
        __syncthreads();
    }
    if (MODE == 0 && nq > 0) copy_out(nq - 1);
}

// Ring-staged forward (inference: no ReLU output kept) for nz >= 36,
// ny % d == 0, ny >= 4 d, nx % d == 0 and d <= 4.  A step computes 4 output
// rows y0 + k d (k = 0..3) of one x plane for one 32-point z block; the rows
// past ny of a residue class's last group and the points past nz of the last
// z block are computed from wrapped inputs and not written (the staged rows
// and points wrap once: ny >= 4 d, nz >= 36).  The steps of
// a walk visit x = c, c + d, c + 2d, ... (one residue class mod d, nx / d
// planes), so the input planes x - d, x, x + d of step j are planes j - 1,
// j, j + 1 of the walk: a ring of 4 LDS plane slots (6 staged rows y0 +
// (r - 1) d, 32 + 2 DMAX z points with the dz halo, 32 channels) holds the
// three in use and receives the next one.  The global loads of a plane are
// issued two steps before it is needed, at the top of step j for plane j + 3,
// into one of two register sets; the other set (plane j + 2, issued a step
// earlier) lands in the free slot after the step's MFMAs.  Each input row is staged
// 1.5 (y) x 1.25 (z halo) times per layer instead of 4.5 times, and the
// weights never pass through LDS: wave (z tile, channel half) keeps the B
// operands of all 27 taps for its 16 output channels in registers.
//
// The concatenated walks (XCD k takes the k-th eighth of them, its
// workgroups contiguous sub-ranges) are split into equal contiguous step
// ranges, one per workgroup; a range start or a walk boundary loads the
// walk's first three planes before computing.  Results go through an LDS
// output stage (two, alternating) and leave as 16-byte stores at the top of
// the next step.
constexpr int RING_Z = 32;                               // output z points per step
constexpr int RING_DMAX = 4;                             // largest dilation
constexpr int RING_ZS = RING_Z + 2 * RING_DMAX;          // staged z points
constexpr int RING_PLANE = 6 * RING_ZS * 4;              // bf16x8 per plane slot
constexpr int RING_OUT = 4 * RING_Z * 4;                 // bf16x8 per output stage
constexpr int RING_THREADS = 256;                        // 2 z tiles x 2 channel halves
constexpr int RING_PER = (RING_PLANE + RING_THREADS - 1) / RING_THREADS;
constexpr size_t RING_LDS = (size_t)(4 * RING_PLANE + 2 * RING_OUT) * 16 + CH * sizeof(float);  // + the bias

__global__ void __launch_bounds__(RING_THREADS, 2)
    k_cnn_mid_ring(Geo g, const __bf16 *__restrict__ xin, const __bf16 *__restrict__ w,
                   const float *__restrict__ bias, __bf16 *__restrict__ out) {
    extern __shared__ bf16x8 lds[];
    bf16x8 *ostage = lds + 4 * RING_PLANE;
    const int d = g.d, nx = g.nx, ny = g.ny, nz = g.nz;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int zt = wv & 1, hh = wv >> 1;
    const int m = lane & 15, kq = lane >> 4;
    const long long nyz = (long long)ny * nz;
    const bf16x8 *xv = reinterpret_cast<const bf16x8 *>(xin);
    const bf16x8 *wvv = reinterpret_cast<const bf16x8 *>(w);
    bf16x8 bw[NTAP];
#pragma unroll
    for (int t = 0; t < NTAP; ++t) bw[t] = wvv[(t * CH + hh * 16 + m) * 4 + kq];
    // the accumulators are D[co][point] (weights as the A operand): lane
    // (m, kq) ends with channels hh*16 + 4kq .. +3 of point zt*16 + m
    const int co0 = hh * 16 + kq * 4;
    // the bias lives in LDS after the output stages (read once per step)
    float *bias_s = reinterpret_cast<float *>(ostage + 2 * RING_OUT);
    if (tid < CH) bias_s[tid] = bias[tid];
    // y groups: per residue class mod d, ceil((ny / d) / 4) groups of 4 rows
    const int ngroups = d * ((ny / d + 3) / 4), nzb = (nz + RING_Z - 1) / RING_Z, walk = nx / d;
    const long long total = (long long)ngroups * nzb * d * walk;
    const int nwg = gridDim.x, per_xcd = nwg / ROWS_XCDS;
    const int xcd = blockIdx.x % ROWS_XCDS, slotw = blockIdx.x / ROWS_XCDS;
    const long long xa = total * xcd / ROWS_XCDS, xb = total * (xcd + 1) / ROWS_XCDS;
    const long long s_begin = xa + (xb - xa) * slotw / per_xcd, s_end = xa + (xb - xa) * (slotw + 1) / per_xcd;
    // position of a step: walk (c, zb, gy), its first row y0, plane j; the
    // divisions happen once, at the range start, later steps advance it
    struct Step {
        int c, zb, gy, y0, j;
    };
    auto step_of = [&](long long s) {
        Step st;
        const long long wk = s / walk;
        st.j = (int)(s - wk * walk);
        st.gy = (int)(wk % ngroups);
        const long long r = wk / ngroups;
        st.zb = (int)(r % nzb);
        st.c = (int)(r / nzb);
        st.y0 = (st.gy / d) * 4 * d + st.gy % d;
        return st;
    };
    auto advance = [&](Step &st) {
        if (++st.j == walk) {
            st.j = 0;
            if (++st.gy == ngroups) {
                st.gy = 0;
                if (++st.zb == nzb) {
                    st.zb = 0;
                    ++st.c;
                }
            }
            st.y0 = (st.gy / d) * 4 * d + st.gy % d;
        }
    };
    // this thread's offsets within an x plane of the walk's staged elements
    // (row, z, 16-byte chunk; the last element repeated past the plane's end),
    // set at each walk start
    int in_off[RING_PER];
    auto walk_setup = [&](const Step &st) {
        const int z0 = st.zb * RING_Z - RING_DMAX;
#pragma unroll
        for (int q = 0; q < RING_PER; ++q) {
            const int e = min(tid + q * RING_THREADS, RING_PLANE - 1);
            const int r = e / (RING_ZS * 4), rem = e - r * (RING_ZS * 4);
            const int yr = wrapi(st.y0 + (r - 1) * d, ny);
            int z = z0 + (rem >> 2);
            z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
            in_off[q] = (yr * nz + z) * 4 + (rem & 3);
        }
    };
    auto fetch = [&](const Step &st, int i, bf16x8 *stg) {
        // plane i in [-1, walk + 2]: x = c + d i wrapped into [0, nx)
        int xp = st.c + d * i;
        if (xp >= nx) xp -= nx;
        if (xp >= nx) xp -= nx;
        if (xp < 0) xp += nx;
        const bf16x8 *src = xv + (long long)xp * nyz * 4;
#pragma unroll
        for (int q = 0; q < RING_PER; ++q) stg[q] = src[in_off[q]];
    };
    // LDS layout of a staged point (z index zs): its four 16-byte channel
    // chunks rotated by zs / 4, so that the 16 points of an MFMA operand read
    // (64-byte stride) fall on distinct banks
    auto land = [&](int slot, const bf16x8 *stg) {
        bf16x8 *dst = lds + slot * RING_PLANE;
#pragma unroll
        for (int q = 0; q < RING_PER; ++q) {
            const int e = tid + q * RING_THREADS;
            const int pnt = e >> 2, zs = pnt % RING_ZS;
            if (e < RING_PLANE) dst[(pnt << 2) | (((e & 3) + (zs >> 2)) & 3)] = stg[q];
        }
    };
    // output stage element e of this thread: row k, 16-byte chunk rem; its
    // offset from the step's first output point, and its row and z offsets
    // (the output stage rotates the chunks of a point by z / 4 like the planes)
    int oc[RING_OUT / RING_THREADS], ok[RING_OUT / RING_THREADS], oz[RING_OUT / RING_THREADS];
#pragma unroll
    for (int q = 0; q < RING_OUT / RING_THREADS; ++q) {
        const int e = tid + q * RING_THREADS;
        const int k = e / (RING_Z * 4), rem = e - k * (RING_Z * 4);
        oc[q] = k * d * nz * 4 + rem;
        ok[q] = k * d;
        oz[q] = rem >> 2;
    }
    auto copy_out = [&](const Step &st, int buf) {
        const int x = st.c + d * st.j;
        const bf16x8 *src = ostage + buf * RING_OUT;
        bf16x8 *ov =
            reinterpret_cast<bf16x8 *>(out) + ((long long)x * nyz + (long long)st.y0 * nz + st.zb * RING_Z) * 4;
        const int ylim = ny - st.y0, zlim = nz - st.zb * RING_Z;  // rows and points of the step inside the grid
#pragma unroll
        for (int q = 0; q < RING_OUT / RING_THREADS; ++q)
            if (ok[q] < ylim && oz[q] < zlim) {
                const int e = tid + q * RING_THREADS;
                ov[oc[q]] = src[(e & ~3) | (((e & 3) + (oz[q] >> 2)) & 3)];
            }
    };
    int head = 0;  // slot of walk plane j - 1; j at head + 1, j + 1 at head + 2, j + 2 lands at head + 3
    bf16x8 sa[RING_PER], sb[RING_PER];
    // one step: `landing` holds plane j + 2 (issued a step earlier, or at the
    // walk's start), `issuing` receives plane j + 3
    auto do_step = [&](long long s, const Step &st, const Step &prev, bf16x8 *landing, bf16x8 *issuing) {
        if (s > s_begin) copy_out(prev, (int)((s - 1) & 1));
        if (s == s_begin || st.j == 0) {
            walk_setup(st);
            // the walk's first three planes, then plane j + 2 into the landing set
#pragma unroll 1
            for (int i = -1; i <= 1; ++i) {
                fetch(st, st.j + i, issuing);
                land((head + 1 + i) & 3, issuing);
            }
            fetch(st, st.j + 2, landing);
            __syncthreads();
        }
        fetch(st, st.j + 3, issuing);
        floatx4 acc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] = floatx4{0.f, 0.f, 0.f, 0.f};
        bf16x8 av[2][6];
        auto load_a = [&](int set, int t) {
            const int dx = t / 3, dz = t % 3;
            const bf16x8 *base = lds + ((head + dx) & 3) * RING_PLANE;
            const int zi = RING_DMAX + zt * 16 + m + (dz - 1) * d;
            const int ch = (kq + (zi >> 2)) & 3;  // the rotated chunk
#pragma unroll
            for (int r = 0; r < 6; ++r) av[set][r] = base[(r * RING_ZS + zi) * 4 + ch];
        };
        load_a(0, 0);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int cur = t & 1;
            if (t + 1 < 9) load_a(cur ^ 1, t + 1);
            const int dx = t / 3, dz = t % 3;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[(dx * 3 + dy) * 3 + dz], av[cur][k + dy],
                                                                     acc[k], 0, 0, 0);
            }
        }
        {
            const __bf16 *rs = reinterpret_cast<const __bf16 *>(lds + ((head + 1) & 3) * RING_PLANE);
            __bf16 *os = reinterpret_cast<__bf16 *>(ostage + (int)(s & 1) * RING_OUT);
            const int z = zt * 16 + m;
            const float4 bco = *reinterpret_cast<const float4 *>(bias_s + co0);
            // channels co0 .. co0 + 3 lie in chunk co0 / 8 of a point, rotated by the point's z / 4
            const int zs = RING_DMAX + z;
            const int rin = (((co0 >> 3) + (zs >> 2)) & 3) * 8 + (co0 & 7);
            const int rout = (((co0 >> 3) + (z >> 2)) & 3) * 8 + (co0 & 7);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                // 4 channels of one point: one 8-byte LDS read and write
                const bf16x4 x4 = *reinterpret_cast<const bf16x4 *>(rs + ((k + 1) * RING_ZS + zs) * CH + rin);
                bf16x4 o4;
                o4[0] = (__bf16)((float)x4[0] + fmaxf(acc[k][0] + bco.x, 0.0f));
                o4[1] = (__bf16)((float)x4[1] + fmaxf(acc[k][1] + bco.y, 0.0f));
                o4[2] = (__bf16)((float)x4[2] + fmaxf(acc[k][2] + bco.z, 0.0f));
                o4[3] = (__bf16)((float)x4[3] + fmaxf(acc[k][3] + bco.w, 0.0f));
                *reinterpret_cast<bf16x4 *>(os + (k * RING_Z + z) * CH + rout) = o4;
            }
        }
        land((head + 3) & 3, landing);
        __syncthreads();
        head = (head + 1) & 3;
    };
    // unrolled by two so the register sets swap roles statically; with every
    // global load unconditional the compiler waits only for the set landing
    Step st = step_of(s_begin), prev = st;
#pragma unroll 1
    for (long long s = s_begin; s < s_end; s += 2) {
        do_step(s, st, prev, sa, sb);
        prev = st;
        advance(st);
        if (s + 1 >= s_end) break;
        do_step(s + 1, st, prev, sb, sa);
        prev = st;
        advance(st);
    }
    if (s_end > s_begin) copy_out(prev, (int)((s_end - 1) & 1));
}

// ------------------------------------------------------------ layer 6 (32 -> 1)
// MFMA projection: u[t][q] = sum_c W6[t][c] h[q][c] as 27 FP16 planes (the
// hidden state converts to FP16 exactly in its range, W6 is rounded to FP16).
// The hidden points are the A operand (rows = 16 consecutive points), W6 the
// B operand (columns = taps).  A wave takes 64 consecutive points (4 tiles),
// gathers the 27 x 64 results in LDS and writes each tap plane's 128 bytes
// with 16-byte stores.  W6 stays in registers.
constexpr int PROJ_THREADS = 256;

__global__ void __launch_bounds__(PROJ_THREADS) k_cnn_out_proj16(Geo g, const __bf16 *__restrict__ h,
                                                                 const float *__restrict__ w,
                                                                 _Float16 *__restrict__ u, long long ngroups) {
    __shared__ f16x8 stage[PROJ_THREADS / 64][NTAP * 8];  // [wave][tap][8 x 8 points]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, m = lane & 15, kq = lane >> 4;
    const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    f16x8 wb[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = nt * 16 + m;
            const float v = w[min(t, NTAP - 1) * CH + kq * 8 + e];  // unconditional: all in flight together
            wb[nt][e] = t < NTAP ? (_Float16)v : (_Float16)0.0f;
        }
    const bf16x8 *hv = reinterpret_cast<const bf16x8 *>(h);
    _Float16 *st = reinterpret_cast<_Float16 *>(stage[wave]);
    const bool vec = (g.n & 7) == 0;
    for (long long grp = wave0; grp < ngroups; grp += nwaves) {
        const long long p0 = grp * 64;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long p = p0 + i * 16 + m;
            const bf16x8 a8 = hv[(p < g.n ? p : g.n - 1) * 4 + kq];
            f16x8 a;
#pragma unroll
            for (int e = 0; e < 8; ++e) a[e] = (_Float16)(float)a8[e];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const floatx4 d =
                    __builtin_amdgcn_mfma_f32_16x16x32_f16(a, wb[nt], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                const int t = nt * 16 + m;  // lane holds points i*16 + kq*4 + j of tap t
                if (t < NTAP)
                    *reinterpret_cast<f16x4 *>(st + t * 64 + i * 16 + kq * 4) =
                        f16x4{(_Float16)d[0], (_Float16)d[1], (_Float16)d[2], (_Float16)d[3]};
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        // chunk c = tap * 8 + k8 holds points k8*8 .. +7 of that tap
        for (int c = lane; c < NTAP * 8; c += 64) {
            const int t = c >> 3, k8 = c & 7;
            const long long q = p0 + k8 * 8;
            const f16x8 v = stage[wave][c];
            _Float16 *dst = u + (long long)t * g.n + q;
            if (vec && q + 7 < g.n) {
                *reinterpret_cast<f16x8 *>(dst) = v;
            } else {
                for (int j = 0; j < 8; ++j)
                    if (q + j < g.n) dst[j] = v[j];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
}

// out[p] = b + sum_t u[t][p + off_t]: one workgroup per (x, y) row, threads
// along z, so the neighbour rows are workgroup-uniform
__global__ void __launch_bounds__(256) k_cnn_out_gather16(Geo g, const _Float16 *__restrict__ u,
                                                          const float *__restrict__ b, float *__restrict__ out) {
    const int y = blockIdx.x, x = blockIdx.y;
    const long long nyz = (long long)g.ny * g.nz;
    const long long bx[3] = {wrapi(x - g.d, g.nx) * nyz, x * nyz, wrapi(x + g.d, g.nx) * nyz};
    const long long by[3] = {(long long)wrapi(y - g.d, g.ny) * g.nz, (long long)y * g.nz,
                             (long long)wrapi(y + g.d, g.ny) * g.nz};
    const float b0 = b[0];
    for (int z = threadIdx.x; z < g.nz; z += blockDim.x) {
        const int bz[3] = {wrapi(z - g.d, g.nz), z, wrapi(z + g.d, g.nz)};
        float acc = b0;
#pragma unroll
        for (int t = 0; t < NTAP; ++t)
            acc += (float)u[(long long)t * g.n + bx[t / 9] + by[(t / 3) % 3] + bz[t % 3]];
        out[bx[1] + by[1] + z] = acc;
    }
}

// The projection form in one pass, without the tap planes.  A workgroup owns
// TY output rows (y) x zt output points (z) and walks x along x_k = c + k d
// (mod nx), so the planes x - d, x, x + d of an output plane are consecutive
// steps.  A step projects one input plane's rows y0 - d .. y0 + TY - 1 + d and
// points z0 - d .. z0 + zt - 1 + d onto the 27 taps (FP16 MFMA with the
// operands of k_cnn_out_proj16) into LDS, then adds each output point's 9
// taps of that plane to one of three accumulators: taps 0-8 start plane
// x_{k+1} (from the bias), taps 9-17 go to x_k, taps 18-26 complete x_{k-1},
// which is written.  Every output is summed in the tap order of
// k_cnn_out_gather16, so the two forms agree bit for bit.  The next plane's
// loads are issued before the tap sums of the current one.
constexpr int OUTW_THREADS = 512;
constexpr int OUTW_NW = OUTW_THREADS / 64;
constexpr int OUTW_MAXT = 8;                        // 16-point MFMA tiles per wave and step
constexpr int OUTW_NP = OUTW_NW * OUTW_MAXT * 16;  // input points per step
constexpr int OUTW_TS = OUTW_NP + 4;               // LDS tap stride: 16 taps' stores hit distinct banks

struct OutWalk {
    int zt, cols, npts, ntiles;  // output points per row of a block; input points per row, per step; MFMA tiles
    int nyt;                     // blocks along y
    int cyc, len, xc;            // gcd(nx, d) walk cycles of nx / cyc planes; xc output planes per block
};

template <int TY>
__global__ void __launch_bounds__(OUTW_THREADS) k_cnn_out_walk(Geo g, OutWalk wk, const __bf16 *__restrict__ h,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ b, float *__restrict__ out) {
    __shared__ _Float16 P[NTAP * OUTW_TS];  // [tap][input point of the step]
    constexpr int OPT = 2;                  // TY * zt < OUTW_NP: at most 2 output points per thread
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, m = lane & 15, kq = lane >> 4;
    const int yb = blockIdx.x % wk.nyt, zb = blockIdx.x / wk.nyt;
    const int y0 = yb * TY, z0 = zb * wk.zt;
    const int cyc = blockIdx.y % wk.cyc;
    const int ka = (blockIdx.y / wk.cyc) * wk.xc, kb = min(wk.len, ka + wk.xc);
    const int nsteps = kb - ka + 2;
    const int nyz = g.ny * g.nz, d = g.d, cols = wk.cols;
    f16x8 wb[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = nt * 16 + m;
            const float v = w[min(t, NTAP - 1) * CH + kq * 8 + e];
            wb[nt][e] = t < NTAP ? (_Float16)v : (_Float16)0.0f;
        }
    const float b0 = b[0];
    // in-plane offsets of this lane's input points (clamped into the step's
    // points), kept in LDS: each thread reads back only its own entries
    __shared__ int offs[OUTW_MAXT * OUTW_THREADS];
#pragma unroll
    for (int k = 0; k < OUTW_MAXT; ++k) {
        const int p = min((wave + k * OUTW_NW) * 16 + m, wk.npts - 1);
        const int r = p / cols, c = p - r * cols;
        const int y = ((y0 - d + r) % g.ny + g.ny) % g.ny, z = ((z0 - d + c) % g.nz + g.nz) % g.nz;
        offs[k * OUTW_THREADS + threadIdx.x] = y * g.nz + z;
    }
    // this thread's output points: LDS index of the centre input point, in-plane output offset
    int oin[OPT], oout[OPT];
    bool ov[OPT];
#pragma unroll
    for (int i = 0; i < OPT; ++i) {
        const int o = min((int)threadIdx.x + i * OUTW_THREADS, TY * wk.zt - 1);
        const int r = o / wk.zt, c = o - r * wk.zt;
        ov[i] = (int)threadIdx.x + i * OUTW_THREADS < TY * wk.zt && y0 + r < g.ny && z0 + c < g.nz;
        oin[i] = (r + d) * cols + c + d;
        oout[i] = (y0 + r) * g.nz + z0 + c;
    }
    const int dr = d * cols;
    auto plane = [&](int k) -> long long {  // x_k (k taken modulo the cycle length)
        const int kk = (k % wk.len + wk.len) % wk.len;
        return ((long long)cyc + (long long)kk * d) % g.nx;
    };
    const bf16x8 *hv = reinterpret_cast<const bf16x8 *>(h);
    bf16x8 cur[OUTW_MAXT];
    auto load = [&](int k) {
        const long long base = plane(k) * nyz;
#pragma unroll
        for (int j = 0; j < OUTW_MAXT; ++j) cur[j] = hv[(base + offs[j * OUTW_THREADS + threadIdx.x]) * 4 + kq];
    };
    float s0[OPT], s1[OPT], s2[OPT];
#pragma unroll
    for (int i = 0; i < OPT; ++i) s0[i] = s1[i] = s2[i] = 0.0f;
    load(ka - 1);
    auto step = [&](int s, float(&ai)[OPT], float(&am)[OPT], float(&af)[OPT]) {
#pragma unroll
        for (int k = 0; k < OUTW_MAXT; ++k) {
            const int tile = wave + k * OUTW_NW;
            if (tile < wk.ntiles) {
                f16x8 a;
#pragma unroll
                for (int e = 0; e < 8; ++e) a[e] = (_Float16)(float)cur[k][e];
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    const floatx4 dd =
                        __builtin_amdgcn_mfma_f32_16x16x32_f16(a, wb[nt], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    const int t = nt * 16 + m;  // lane holds points tile*16 + kq*4 + j of tap t
                    if (t < NTAP)
                        *reinterpret_cast<f16x4 *>(P + t * OUTW_TS + tile * 16 + kq * 4) =
                            f16x4{(_Float16)dd[0], (_Float16)dd[1], (_Float16)dd[2], (_Float16)dd[3]};
                }
            }
        }
        if (s + 1 < nsteps) load(ka + s);
        __syncthreads();
        const long long xo = plane(ka - 2 + s) * nyz;  // plane completed by this step
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const _Float16 *q = P + oin[i];
            float a0 = b0;
#pragma unroll
            for (int t = 0; t < 9; ++t) a0 += (float)q[t * OUTW_TS + (t / 3 - 1) * dr + (t % 3 - 1) * d];
            float a1 = am[i];
#pragma unroll
            for (int t = 9; t < 18; ++t) a1 += (float)q[t * OUTW_TS + ((t / 3) % 3 - 1) * dr + (t % 3 - 1) * d];
            float a2 = af[i];
#pragma unroll
            for (int t = 18; t < 27; ++t) a2 += (float)q[t * OUTW_TS + ((t / 3) % 3 - 1) * dr + (t % 3 - 1) * d];
            ai[i] = a0;
            am[i] = a1;
            if (s >= 2 && ov[i]) out[xo + oout[i]] = a2;
        }
        __syncthreads();
    };
#pragma unroll 1
    for (int s = 0; s < nsteps; s += 3) {
        step(s, s0, s2, s1);
        if (s + 1 < nsteps) step(s + 1, s1, s0, s2);
        if (s + 2 < nsteps) step(s + 2, s2, s1, s0);
    }
}

// dh[q][c] = sum_t W6[t][c] dout[q - off_t];  g = mask > 0 ? dh : 0 (bf16)
__global__ void __launch_bounds__(256) k_cnn_out_bwd(Geo g, const float *__restrict__ dout,
                                                     const float *__restrict__ w, const __bf16 *__restrict__ mask,
                                                     float *__restrict__ dh, __bf16 *__restrict__ gout) {
    __shared__ float ws[KOUT];
    for (int i = threadIdx.x; i < KOUT; i += blockDim.x) ws[i] = w[i];
    __syncthreads();
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n) return;
    int x, y, z;
    decompose(g, p, x, y, z);
    Nbr nb;
    nb.init(g, x, y, z);
    float acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0f;
#pragma unroll 3
    for (int t = 0; t < NTAP; ++t) {
        const float dv = dout[nb.at(NTAP - 1 - t)];  // q - off_t = q + off_{26-t}
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] += ws[t * CH + c] * dv;
    }
    float4 *dd = reinterpret_cast<float4 *>(dh + p * CH);
#pragma unroll
    for (int q = 0; q < 8; ++q) dd[q] = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
    const bf16x8 *mv = reinterpret_cast<const bf16x8 *>(mask + p * CH);
    bf16x8 *gv = reinterpret_cast<bf16x8 *>(gout + p * CH);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bf16x8 mk = mv[q];
        bf16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = (float)mk[e] > 0.0f ? (__bf16)acc[q * 8 + e] : (__bf16)0.0f;
        gv[q] = o;
    }
}

// --------------------------------------------------------- weight gradients
constexpr int WG_CHUNK = 256;  // points per chunk

// dW6[t][c] = sum_p dout[p] h[p + off_t][c]; db6 = sum_p dout[p].
// 128 threads: thread j < 108 owns (t = j / 4, channels 8*(j%4)..+8); j == 108
// the bias.  Grid-stride over chunks; partials[block][KOUT + 1].
__global__ void __launch_bounds__(128) k_cnn_wgrad_out(Geo g, const float *__restrict__ dout,
                                                       const __bf16 *__restrict__ h, float *__restrict__ partials) {
    __shared__ float ds[WG_CHUNK];
    const int j = threadIdx.x;
    const int t = j / 4, cg = j % 4;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float accb = 0.0f;
    const bf16x8 *hv = reinterpret_cast<const bf16x8 *>(h);
    const long long nchunk = (g.n + WG_CHUNK - 1) / WG_CHUNK;
    for (long long ck = blockIdx.x; ck < nchunk; ck += gridDim.x) {
        const long long p0 = ck * WG_CHUNK;
        __syncthreads();
        for (int i = j; i < WG_CHUNK; i += 128) ds[i] = (p0 + i < g.n) ? dout[p0 + i] : 0.0f;
        __syncthreads();
        const int np = (int)min((long long)WG_CHUNK, g.n - p0);
        if (j < 108) {
            int x, y, z;
            decompose(g, p0, x, y, z);
            for (int i = 0; i < np; ++i) {
                Nbr nb;
                nb.init(g, x, y, z);
                const bf16x8 v = hv[(long long)nb.at(t) * 4 + cg];
                const float dv = ds[i];
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[e] += dv * (float)v[e];
                if (++z == g.nz) {
                    z = 0;
                    if (++y == g.ny) {
                        y = 0;
                        ++x;
                    }
                }
            }
        } else if (j == 108) {
            for (int i = 0; i < np; ++i) accb += ds[i];
        }
    }
    float *pp = partials + (long long)blockIdx.x * (KOUT + 1);
    if (j < 108) {
#pragma unroll
        for (int e = 0; e < 8; ++e) pp[t * CH + cg * 8 + e] = acc[e];
    } else if (j == 108) {
        pp[KOUT] = accb;
    }
}

// dW1[t][c][co] = sum_p g[p][co] x0[c][p + off_t]; db1[co] = sum_p g[p][co].
// 256 threads: j < 216 owns (t = j / 8, c = (j / 4) % 2, co group j % 4);
// 216 <= j < 220 the bias of co group j - 216.  partials[block][KIN + 32].
__global__ void __launch_bounds__(256) k_cnn_wgrad_in(Geo g, const __bf16 *__restrict__ gin,
                                                      const float *__restrict__ x0, float *__restrict__ partials) {
    __shared__ bf16x8 gs[WG_CHUNK * 4];
    const int j = threadIdx.x;
    const int t = j / 8, c = (j / 4) % 2, cg = j % 4;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bf16x8 *gv = reinterpret_cast<const bf16x8 *>(gin);
    const long long nchunk = (g.n + WG_CHUNK - 1) / WG_CHUNK;
    const bf16x8 zero8 = {};
    for (long long ck = blockIdx.x; ck < nchunk; ck += gridDim.x) {
        const long long p0 = ck * WG_CHUNK;
        __syncthreads();
        for (int i = j; i < WG_CHUNK * 4; i += 256) gs[i] = (p0 + i / 4 < g.n) ? gv[p0 * 4 + i] : zero8;
        __syncthreads();
        const int np = (int)min((long long)WG_CHUNK, g.n - p0);
        if (j < 216) {
            int x, y, z;
            decompose(g, p0, x, y, z);
            const float *xc = x0 + (long long)c * g.n;
            for (int i = 0; i < np; ++i) {
                Nbr nb;
                nb.init(g, x, y, z);
                const float xv = xc[nb.at(t)];
                const bf16x8 gg = gs[i * 4 + cg];
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[e] += xv * (float)gg[e];
                if (++z == g.nz) {
                    z = 0;
                    if (++y == g.ny) {
                        y = 0;
                        ++x;
                    }
                }
            }
        } else if (j < 220) {
            for (int i = 0; i < np; ++i) {
                const bf16x8 gg = gs[i * 4 + (j - 216)];
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[e] += (float)gg[e];
            }
        }
    }
    float *pp = partials + (long long)blockIdx.x * (KIN + CH);
    if (j < 216) {
#pragma unroll
        for (int e = 0; e < 8; ++e) pp[(t * 2 + c) * CH + cg * 8 + e] = acc[e];
    } else if (j < 220) {
#pragma unroll
        for (int e = 0; e < 8; ++e) pp[KIN + (j - 216) * 8 + e] = acc[e];
    }
}

// Mid-layer weight gradient with MFMA: dW[t][co][ci] = sum_p g[p][co] x[p+off_t][ci].
// 4 waves; wave w owns the 16x16 block (co block w>>1, ci block w&1) of all 27
// taps.  Per chunk of 256 points: g is staged transposed in LDS (gT[co][p])
// once, and for each tap the shifted x rows are staged transposed (xT[ci][p]);
// the MFMA K dimension runs over points.  partials[block][KMID + 32].
constexpr int WG_PITCH = WG_CHUNK + 8;  // bf16 elements per LDS row (16-byte aligned)

__global__ void __launch_bounds__(256) k_cnn_wgrad_mid(Geo g, const __bf16 *__restrict__ gin,
                                                       const __bf16 *__restrict__ xin, float *__restrict__ partials) {
    __shared__ __attribute__((aligned(16))) __bf16 gT[CH * WG_PITCH];
    __shared__ __attribute__((aligned(16))) __bf16 xT[CH * WG_PITCH];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const int mt = wave >> 1, nt = wave & 1;
    floatx4 acc[NTAP];
#pragma unroll
    for (int t = 0; t < NTAP; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    float dbacc = 0.0f;
    const bf16x8 *gv = reinterpret_cast<const bf16x8 *>(gin);
    const bf16x8 *xv = reinterpret_cast<const bf16x8 *>(xin);
    const long long nchunk = (g.n + WG_CHUNK - 1) / WG_CHUNK;
    const bf16x8 zero8 = {};
    for (long long ck = blockIdx.x; ck < nchunk; ck += gridDim.x) {
        const long long p = ck * WG_CHUNK + tid;  // this thread's point
        const bool valid = p < g.n;
        int x, y, z;
        decompose(g, valid ? p : 0, x, y, z);
        Nbr nb;
        nb.init(g, x, y, z);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bf16x8 v = valid ? gv[p * 4 + q] : zero8;
#pragma unroll
            for (int e = 0; e < 8; ++e) gT[(q * 8 + e) * WG_PITCH + tid] = v[e];
        }
        __syncthreads();
        {  // bias gradient: thread (c = tid / 8, part = tid % 8) sums 32 points of channel c
            const int c = tid >> 3, part = tid & 7;
            float s = 0.0f;
#pragma unroll 8
            for (int i = 0; i < 32; ++i) s += (float)gT[c * WG_PITCH + part * 32 + i];
            dbacc += s;
        }
        bf16x8 a[WG_CHUNK / 32];
#pragma unroll
        for (int ks = 0; ks < WG_CHUNK / 32; ++ks)
            a[ks] = *reinterpret_cast<const bf16x8 *>(&gT[(mt * 16 + m) * WG_PITCH + ks * 32 + kq * 8]);
        bf16x8 xr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xr[q] = valid ? xv[(long long)nb.at(0) * 4 + q] : zero8;
#pragma unroll
        for (int t = 0; t < NTAP; ++t) {
            __syncthreads();  // previous tap's readers are done with xT
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) xT[(q * 8 + e) * WG_PITCH + tid] = xr[q][e];
            __syncthreads();
            if (t + 1 < NTAP) {
#pragma unroll
                for (int q = 0; q < 4; ++q) xr[q] = valid ? xv[(long long)nb.at(t + 1) * 4 + q] : zero8;
            }
#pragma unroll
            for (int ks = 0; ks < WG_CHUNK / 32; ++ks) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&xT[(nt * 16 + m) * WG_PITCH + ks * 32 + kq * 8]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], b, acc[t], 0, 0, 0);
            }
        }
    }
    float *pp = partials + (long long)blockIdx.x * (KMID + CH);
#pragma unroll
    for (int t = 0; t < NTAP; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) pp[t * CH * CH + (mt * 16 + kq * 4 + i) * CH + nt * 16 + m] = acc[t][i];
    // bias: 8 partial sums per channel, reduced in LDS
    __syncthreads();
    float *red = reinterpret_cast<float *>(gT);
    red[tid] = dbacc;
    __syncthreads();
    if (tid < CH) {
        float s = 0.0f;
        for (int part = 0; part < 8; ++part) s += red[tid * 8 + part];
        pp[KMID + tid] = s;
    }
}

// out[j] = sum_b partials[b * K + j]  (FP64 accumulation)
__global__ void k_cnn_reduce(const float *__restrict__ partials, int nblk, int K, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= K) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += (double)partials[(long long)b * K + j];
    out[j] = (float)s;
}

// wf[t][co][ci] = bf16(W[t][co][ci]);  wb[t][ci][co] = bf16(W[26-t][co][ci])
__global__ void k_cnn_prep(const float *__restrict__ w, __bf16 *__restrict__ wf, __bf16 *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= KMID) return;
    const int t = i / (CH * CH), co = (i / CH) % CH, ci = i % CH;
    wf[i] = (__bf16)w[i];
    wb[((NTAP - 1 - t) * CH + ci) * CH + co] = (__bf16)w[i];
}

// d = pred - target; dpred = 2 d / n; partial sums of d^2
__global__ void k_cnn_loss(long long n, const float *__restrict__ pred, const float *__restrict__ target,
                           float *__restrict__ dpred, double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    const float s = 2.0f / (float)n;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = pred[i] - target[i];
        dpred[i] = s * d;
        acc += (double)d * (double)d;
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// d = pred - target
__global__ void k_cnn_diff(long long n, const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ d) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = a[i] - b[i];
}

// Spectral loss: D *= w (half spectrum, dense nzc rows), partial sums of
// m(kz) w |D|^2 with m = 1 at kz = 0 and the Nyquist column, else 2
__global__ void k_cnn_metric(long long nc, int nzc, int nz, float2 *__restrict__ D, const float *__restrict__ w,
                             double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(i % nzc);
        const float wi = w[i];
        const float2 v = D[i];
        const double m = (kz == 0 || 2 * kz == nz) ? 1.0 : 2.0;
        acc += m * (double)wi * ((double)v.x * v.x + (double)v.y * v.y);
        D[i] = make_float2(v.x * wi, v.y * wi);
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction, FP32 master weights
__global__ void k_cnn_adam(long long n, float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float g = gr[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// Input features: x0[0] = c (rho_s - rho0) / rho0, x0[1] = sv * vloc
template <typename T>
__global__ void k_cnn_features(long long n, const T *__restrict__ rho_s, const T *__restrict__ vloc, double rho0,
                               double c, double sv, float *__restrict__ x0) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        x0[i] = (float)(c * ((double)rho_s[i] - rho0) / rho0);
        x0[n + i] = (float)(sv * (double)vloc[i]);
    }
}

// Training target: log(rho / rho0) - x0[0]
template <typename T>
__global__ void k_cnn_target(long long n, const T *__restrict__ rho, const float *__restrict__ x0, double rho0,
                             float *__restrict__ target) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        target[i] = (float)(log((double)rho[i] / rho0) - (double)x0[i]);
}

// rho = rho0 exp(x0[0] + delta), with the FP64 sum of rho in partials
template <typename T>
__global__ void k_cnn_apply(long long n, const float *__restrict__ x0, const float *__restrict__ delta, double rho0,
                            T *__restrict__ rho, double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double r = rho0 * exp((double)x0[i] + (double)(delta ? delta[i] : 0.0f));
        rho[i] = (T)r;
        acc += r;
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// rho *= ne / (sum * dV)
template <typename T> __global__ void k_cnn_normalise(long long n, T *rho, const double *sum, double ne, double dV) {
    const double f = ne / (*sum * dV);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        rho[i] = (T)((double)rho[i] * f);
}

static Geo make_geo(const int *nr, int dil) {
    OFDFT_REQUIRE(nr != nullptr && nr[0] > 0 && nr[1] > 0 && nr[2] > 0, "bad grid shape");
    OFDFT_REQUIRE(dil >= 1 && dil < nr[0] && dil < nr[1] && dil < nr[2], "dilation must be in [1, min(nr))");
    Geo g;
    g.nx = nr[0];
    g.ny = nr[1];
    g.nz = nr[2];
    g.d = dil;
    g.n = (long long)nr[0] * nr[1] * nr[2];
    OFDFT_REQUIRE(g.n < (1LL << 31), "grid too large for 32-bit point indices");
    return g;
}

static unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

static int cnn_guard_impl(void (*f)(void *), void *ctx) {
    try {
        f(ctx);
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("unknown error");
        return 1;
    }
}

template <typename F> static int cnn_guard(F &&f) {
    return cnn_guard_impl([](void *c) { (*reinterpret_cast<F *>(c))(); }, &f);
}

static hipStream_t St(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace ofdft

using namespace ofdft;

struct ofdft_plan : public Plan {};

extern "C" {

int ofdft_cnn_conv_in(const int nr[3], int dil, const float *x0, const float *w, const float *b, void *h,
                      void *stream) {
    return cnn_guard([&] {
        const Geo g = make_geo(nr, dil);
        OFDFT_REQUIRE(x0 && w && b && h, "null pointer");
        if (g.nz % 16 == 0 && g.nz <= IN_ROWS_MAX_NZ && g.ny % (4 * g.d) == 0 && g.d <= 64 &&
            (reinterpret_cast<uintptr_t>(x0) & 15) == 0) {
            // staged features (16-byte row loads) + one 1 KB output tile per wave
            const size_t lds = (size_t)36 * (g.nz + 2 * ((g.d + 3) & ~3)) * sizeof(_Float16) +
                               (size_t)(IN_ROWS_THREADS / 64) * 1024;
            if (g.nz <= 256)
                hipLaunchKernelGGL(k_cnn_in_rows<1>, dim3((unsigned)(g.ny / 4), (unsigned)g.nx), dim3(IN_ROWS_THREADS),
                                   lds, St(stream), g, x0, w, b, reinterpret_cast<__bf16 *>(h));
            else
                hipLaunchKernelGGL(k_cnn_in_rows<2>, dim3((unsigned)(g.ny / 4), (unsigned)g.nx), dim3(IN_ROWS_THREADS),
                                   lds, St(stream), g, x0, w, b, reinterpret_cast<__bf16 *>(h));
        } else {
            const long long ntiles = (g.n + 15) / 16;
            // 8 tiles per wave amortise the weight loads
            const unsigned blocks =
                (unsigned)std::max<long long>(1, std::min<long long>((ntiles + 31) / 32, 1 << 20));
            hipLaunchKernelGGL(k_cnn_in_mfma, dim3(blocks), dim3(256), 0, St(stream), g, x0, w, b,
                               reinterpret_cast<__bf16 *>(h), ntiles);
        }
        OFDFT_HIP(hipGetLastError());
    });
}

// workgroups of the persistent row kernel: one per CU, a multiple of the 8
// XCDs
static int rows_workgroups() {
    int dev = 0, cus = 0;
    OFDFT_HIP(hipGetDevice(&dev));
    OFDFT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(ROWS_XCDS, cus / ROWS_XCDS * ROWS_XCDS);
}

// allow a row-tiled launch above the default 64 KB of dynamic LDS
static void rows_lds_limit(const void *kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        OFDFT_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

int ofdft_cnn_conv_mid(const int nr[3], int dil, int mode, const void *x, const void *w, const float *bias,
                       const float *resid, const void *mask, void *out0, void *out1, void *stream) {
    return cnn_guard([&] {
        const Geo g = make_geo(nr, dil);
        OFDFT_REQUIRE(x && w && out0 && (out1 || mode == 0), "null pointer");
        const __bf16 *xb = reinterpret_cast<const __bf16 *>(x);
        const __bf16 *wb = reinterpret_cast<const __bf16 *>(w);
        OFDFT_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (forward) or 1 (backward-data)");
        OFDFT_REQUIRE(mode == 1 || bias, "forward mode needs the bias");
        OFDFT_REQUIRE(mode == 0 || (resid && mask), "backward mode needs resid and mask");
        if (mode == 0 && out1 == nullptr && g.nz >= RING_Z + RING_DMAX && g.ny % g.d == 0 && g.ny >= 4 * g.d &&
            g.nx % g.d == 0 && g.d <= RING_DMAX) {
            rows_lds_limit((const void *)&k_cnn_mid_ring, RING_LDS);
            hipLaunchKernelGGL(k_cnn_mid_ring, dim3((unsigned)(2 * rows_workgroups())), dim3(RING_THREADS), RING_LDS,
                               St(stream), g, xb, wb, bias, reinterpret_cast<__bf16 *>(out0));
            OFDFT_HIP(hipGetLastError());
            return;
        }
        if (g.nz % 32 == 0 && g.nz <= 32 * ROWS_MAX_MT && g.ny % (4 * g.d) == 0) {
            // two stage buffers of 6 rows and 9 taps' weights each
            const size_t lds = 2 * ((size_t)6 * g.nz * CH * sizeof(__bf16) + (size_t)9 * CH * CH * sizeof(__bf16));
            const int mt = g.nz / 32;
            const dim3 block((unsigned)(4 * g.nz));  // one wave per 16-point z tile
            const dim3 grid((unsigned)rows_workgroups());
#define ROWS_LAUNCH(MTV)                                                                                            \
    do {                                                                                                            \
        rows_lds_limit((const void *)&k_cnn_mid_rows<0, MTV>, lds);                                                \
        rows_lds_limit((const void *)&k_cnn_mid_rows<1, MTV>, lds);                                                \
        if (mode == 0)                                                                                              \
            hipLaunchKernelGGL((k_cnn_mid_rows<0, MTV>), grid, block, lds, St(stream), g, xb, wb, bias, nullptr,      \
                               nullptr, out0, reinterpret_cast<__bf16 *>(out1));                                    \
        else                                                                                                        \
            hipLaunchKernelGGL((k_cnn_mid_rows<1, MTV>), grid, block, lds, St(stream), g, xb, wb, nullptr, resid,     \
                               reinterpret_cast<const __bf16 *>(mask), out0, reinterpret_cast<__bf16 *>(out1));      \
    } while (0)
            switch (mt) {
                case 1: ROWS_LAUNCH(1); break;
                case 2: ROWS_LAUNCH(2); break;
                case 3: ROWS_LAUNCH(3); break;
                case 4: ROWS_LAUNCH(4); break;
                default: ROWS_LAUNCH(5); break;
            }
#undef ROWS_LAUNCH
            OFDFT_HIP(hipGetLastError());
            return;
        }
        const dim3 grid(blocks_for(g.n, MID_POINTS));
        if (mode == 0) {
            OFDFT_REQUIRE(bias, "forward mode needs the bias");
            hipLaunchKernelGGL(k_cnn_mid<0>, grid, dim3(MID_THREADS), 0, St(stream), g, xb, wb, bias, nullptr, nullptr,
                               out0, reinterpret_cast<__bf16 *>(out1));
        } else if (mode == 1) {
            OFDFT_REQUIRE(resid && mask, "backward mode needs resid and mask");
            hipLaunchKernelGGL(k_cnn_mid<1>, grid, dim3(MID_THREADS), 0, St(stream), g, xb, wb, nullptr, resid,
                               reinterpret_cast<const __bf16 *>(mask), out0, reinterpret_cast<__bf16 *>(out1));
        } else {
            OFDFT_REQUIRE(false, "mode must be 0 (forward) or 1 (backward-data)");
        }
        OFDFT_HIP(hipGetLastError());
    });
}

// one-pass output layer: blocks of TY rows x zt points, about 512 workgroups
// (two per CU) walking at least one output plane each
static void launch_out_walk(const Geo &g, const void *h, const float *w, const float *b, float *out, hipStream_t st) {
    constexpr int TY = 8;
    OutWalk wk;
    const int rows = TY + 2 * g.d;
    const int ztmax = OUTW_NP / rows - 2 * g.d;
    OFDFT_REQUIRE(ztmax >= 1, "dilation too large for the one-pass output layer");
    const int nzt = (g.nz + ztmax - 1) / ztmax;
    wk.zt = (g.nz + nzt - 1) / nzt;
    wk.cols = wk.zt + 2 * g.d;
    wk.npts = rows * wk.cols;  // <= OUTW_NP, and TY * zt < OUTW_NP
    wk.ntiles = (wk.npts + 15) / 16;
    wk.nyt = (g.ny + TY - 1) / TY;
    wk.cyc = std::gcd(g.nx, g.d);
    wk.len = g.nx / wk.cyc;
    const long long byz = (long long)wk.nyt * nzt;
    int nch = (int)std::min<long long>(wk.len, std::max<long long>(1, (512 + byz * wk.cyc - 1) / (byz * wk.cyc)));
    wk.xc = (wk.len + nch - 1) / nch;
    nch = (wk.len + wk.xc - 1) / wk.xc;
    OFDFT_REQUIRE(byz < (1LL << 31) && (long long)wk.cyc * nch <= 65535, "grid too large for the output layer");
    hipLaunchKernelGGL(k_cnn_out_walk<TY>, dim3((unsigned)byz, (unsigned)(wk.cyc * nch)), dim3(OUTW_THREADS), 0, st,
                       g, wk, reinterpret_cast<const __bf16 *>(h), w, b, out);
}

int ofdft_cnn_conv_out(const int nr[3], int dil, const void *h, const float *w, const float *b, float *out,
                       float *work, void *stream) {
    return cnn_guard([&] {
        const Geo g = make_geo(nr, dil);
        OFDFT_REQUIRE(h && w && b && out, "null pointer");
        if (work) {
            // the work buffer (27 N floats) holds the 27 FP16 tap planes
            _Float16 *u = reinterpret_cast<_Float16 *>(work);
            const long long ngroups = (g.n + 63) / 64;  // 64-point groups, 4 per wave on average
            const unsigned blocks = (unsigned)std::max<long long>(1, std::min<long long>((ngroups + 15) / 16, 1 << 20));
            hipLaunchKernelGGL(k_cnn_out_proj16, dim3(blocks), dim3(PROJ_THREADS), 0, St(stream), g,
                               reinterpret_cast<const __bf16 *>(h), w, u, ngroups);
            OFDFT_HIP(hipGetLastError());
            const unsigned threads = (unsigned)std::min(256, (g.nz + 63) / 64 * 64);
            hipLaunchKernelGGL(k_cnn_out_gather16, dim3((unsigned)g.ny, (unsigned)g.nx), dim3(threads), 0, St(stream),
                               g, u, b, out);
        } else {
            launch_out_walk(g, h, w, b, out, St(stream));
        }
        OFDFT_HIP(hipGetLastError());
    });
}

int ofdft_cnn_out_backward(const int nr[3], int dil, const float *dout, const float *w, const void *mask, float *dh,
                           void *gout, void *stream) {
    return cnn_guard([&] {
        const Geo g = make_geo(nr, dil);
        OFDFT_REQUIRE(dout && w && mask && dh && gout, "null pointer");
        hipLaunchKernelGGL(k_cnn_out_bwd, dim3(blocks_for(g.n, 256)), dim3(256), 0, St(stream), g, dout, w,
                           reinterpret_cast<const __bf16 *>(mask), dh, reinterpret_cast<__bf16 *>(gout));
        OFDFT_HIP(hipGetLastError());
    });
}

size_t ofdft_cnn_wgrad_partials(int layer, int nblk) {
    const int K = layer == 0 ? KIN + CH : (layer == 1 ? KMID + CH : KOUT + 1);
    return (size_t)nblk * K;
}

int ofdft_cnn_wgrad(const int nr[3], int dil, int layer, const void *g_in, const void *x, float *partials, int nblk,
                    float *dw, void *stream) {
    return cnn_guard([&] {
        const Geo g = make_geo(nr, dil);
        OFDFT_REQUIRE(g_in && x && partials && dw, "null pointer");
        OFDFT_REQUIRE(nblk >= 1 && nblk <= 65535, "nblk out of range");
        int K = 0;
        if (layer == 0) {
            K = KIN + CH;
            hipLaunchKernelGGL(k_cnn_wgrad_in, dim3(nblk), dim3(256), 0, St(stream), g,
                               reinterpret_cast<const __bf16 *>(g_in), reinterpret_cast<const float *>(x), partials);
        } else if (layer == 1) {
            K = KMID + CH;
            hipLaunchKernelGGL(k_cnn_wgrad_mid, dim3(nblk), dim3(256), 0, St(stream), g,
                               reinterpret_cast<const __bf16 *>(g_in), reinterpret_cast<const __bf16 *>(x), partials);
        } else if (layer == 2) {
            K = KOUT + 1;
            hipLaunchKernelGGL(k_cnn_wgrad_out, dim3(nblk), dim3(128), 0, St(stream), g,
                               reinterpret_cast<const float *>(g_in), reinterpret_cast<const __bf16 *>(x), partials);
        } else {
            OFDFT_REQUIRE(false, "layer must be 0 (input), 1 (32->32) or 2 (output)");
        }
        OFDFT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_cnn_reduce, dim3(blocks_for(K, 256)), dim3(256), 0, St(stream), partials, nblk, K, dw);
        OFDFT_HIP(hipGetLastError());
    });
}

int ofdft_cnn_prep_weights(const float *w, void *wf, void *wb, void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(w && wf && wb, "null pointer");
        hipLaunchKernelGGL(k_cnn_prep, dim3(blocks_for(KMID, 256)), dim3(256), 0, St(stream), w,
                           reinterpret_cast<__bf16 *>(wf), reinterpret_cast<__bf16 *>(wb));
        OFDFT_HIP(hipGetLastError());
    });
}

int ofdft_cnn_loss(ofdft_plan *plan, const float *pred, const float *target, float *dpred, double *loss,
                   void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(plan && pred && target && dpred && loss, "null pointer");
        const int nb = reduce_blocks(plan->N);
        hipLaunchKernelGGL(k_cnn_loss, dim3(nb), dim3(256), 0, St(stream), plan->N, pred, target, dpred, plan->red);
        OFDFT_HIP(hipGetLastError());
        const double s = 1.0 / (double)plan->N;
        reduce_final(plan->red, nb, 1, loss, &s, St(stream));
    });
}

int ofdft_cnn_loss_spectral(ofdft_plan *plan, const float *pred, const float *target, const float *w, float *dpred,
                            double *loss, void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(plan && pred && target && w && dpred && loss, "null pointer");
        OFDFT_REQUIRE(plan->prec == OFDFT_FP32, "spectral loss needs an FP32 plan");
        Plan &P = *plan;
        const hipStream_t s = St(stream);
        float *d = reinterpret_cast<float *>(P.work);
        float2 *D = reinterpret_cast<float2 *>(P.spec);
        const int nb = reduce_blocks(P.N);
        hipLaunchKernelGGL(k_cnn_diff, dim3(std::min<unsigned>(blocks_for(P.N, 256), 4096u)), dim3(256), 0, s, P.N,
                           pred, target, d);
        OFDFT_HIP(hipGetLastError());
        fft_r2c<float>(P, d, D, 1, P.N, P.Nc, 1.0, nullptr, nullptr, s, 0);
        const int nbc = reduce_blocks(P.Nc);
        hipLaunchKernelGGL(k_cnn_metric, dim3(nbc), dim3(256), 0, s, P.Nc, P.nzc, P.nr[2], D, w, P.red);
        OFDFT_HIP(hipGetLastError());
        const double inv_n2 = 1.0 / ((double)P.N * (double)P.N);
        reduce_final(P.red, nbc, 1, loss, &inv_n2, s);
        // d loss / d pred = (2 / N^2) sum_q w D e^{iqr}: unnormalised C2R scaled by 2 / N^2
        fft_c2r<float>(P, D, D, dpred, 1, P.Nc, P.Nc, P.N, 2.0 * inv_n2, s, 0);
        (void)nb;
    });
}

int ofdft_cnn_adam(long long n, float *p, const float *g, float *m, float *v, double lr, double b1, double b2,
                   double eps, int step, void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(n > 0 && p && g && m && v && step >= 1, "bad Adam arguments");
        const float bc1 = (float)(1.0 - std::pow(b1, step)), bc2 = (float)(1.0 - std::pow(b2, step));
        hipLaunchKernelGGL(k_cnn_adam, dim3(std::min<unsigned>(blocks_for(n, 256), 1024u)), dim3(256), 0, St(stream),
                           n, p, g, m, v, (float)lr, (float)b1, (float)b2, (float)eps, bc1, bc2);
        OFDFT_HIP(hipGetLastError());
    });
}

int ofdft_cnn_features(ofdft_plan *plan, const void *rho_s, const void *vloc, double rho0, double c, double sv,
                       float *x0, void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(plan && rho_s && vloc && x0, "null pointer");
        const unsigned nb = std::min<unsigned>(blocks_for(plan->N, 256), 4096u);
        if (plan->prec == OFDFT_FP64)
            hipLaunchKernelGGL(k_cnn_features<double>, dim3(nb), dim3(256), 0, St(stream), plan->N,
                               reinterpret_cast<const double *>(rho_s), reinterpret_cast<const double *>(vloc), rho0,
                               c, sv, x0);
        else
            hipLaunchKernelGGL(k_cnn_features<float>, dim3(nb), dim3(256), 0, St(stream), plan->N,
                               reinterpret_cast<const float *>(rho_s), reinterpret_cast<const float *>(vloc), rho0, c,
                               sv, x0);
        OFDFT_HIP(hipGetLastError());
    });
}

int ofdft_cnn_target(ofdft_plan *plan, const void *rho, const float *x0, double rho0, float *target, void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(plan && rho && x0 && target && rho0 > 0, "bad arguments");
        const unsigned nb = std::min<unsigned>(blocks_for(plan->N, 256), 4096u);
        if (plan->prec == OFDFT_FP64)
            hipLaunchKernelGGL(k_cnn_target<double>, dim3(nb), dim3(256), 0, St(stream), plan->N,
                               reinterpret_cast<const double *>(rho), x0, rho0, target);
        else
            hipLaunchKernelGGL(k_cnn_target<float>, dim3(nb), dim3(256), 0, St(stream), plan->N,
                               reinterpret_cast<const float *>(rho), x0, rho0, target);
        OFDFT_HIP(hipGetLastError());
    });
}

int ofdft_cnn_apply(ofdft_plan *plan, const float *x0, const float *delta, double rho0, double ne, void *rho,
                    void *stream) {
    return cnn_guard([&] {
        OFDFT_REQUIRE(plan && x0 && rho, "null pointer");
        const int nb = reduce_blocks(plan->N);
        double *sum = plan->scal + 200;
        const double one = 1.0;
        if (plan->prec == OFDFT_FP64) {
            double *r = reinterpret_cast<double *>(rho);
            hipLaunchKernelGGL(k_cnn_apply<double>, dim3(nb), dim3(256), 0, St(stream), plan->N, x0, delta, rho0, r,
                               plan->red);
            OFDFT_HIP(hipGetLastError());
            reduce_final(plan->red, nb, 1, sum, &one, St(stream));
            hipLaunchKernelGGL(k_cnn_normalise<double>, dim3(nb), dim3(256), 0, St(stream), plan->N, r, sum, ne,
                               plan->dV);
        } else {
            float *r = reinterpret_cast<float *>(rho);
            hipLaunchKernelGGL(k_cnn_apply<float>, dim3(nb), dim3(256), 0, St(stream), plan->N, x0, delta, rho0, r,
                               plan->red);
            OFDFT_HIP(hipGetLastError());
            reduce_final(plan->red, nb, 1, sum, &one, St(stream));
            hipLaunchKernelGGL(k_cnn_normalise<float>, dim3(nb), dim3(256), 0, St(stream), plan->N, r, sum, ne,
                               plan->dV);
        }
        OFDFT_HIP(hipGetLastError());
    });
}

}  // extern "C"
