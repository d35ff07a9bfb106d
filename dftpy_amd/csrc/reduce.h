// Wave64 / workgroup reduction helpers (FP64).
#pragma once
#include <hip/hip_runtime.h>

namespace ofdft {

constexpr int kRedBlocks = 1024;  // max blocks of a grid-stride reduction
constexpr int kMaxTerms = 16;     // max simultaneous reduced quantities

// Sum over the workgroup; the result is valid in thread 0.  sm: 16 doubles.
__device__ __forceinline__ double block_sum(double v, double *sm) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) s += sm[w];
    }
    __syncthreads();
    return s;
}

constexpr int kFinalThreads = 1024;  // workgroup size of the final (one-workgroup) reductions

// Sums of the `nterms` interleaved columns of `p` (element b * nterms + t,
// b < nb), by one workgroup of kFinalThreads threads.  With M the largest
// power of two such that M * nterms <= kFinalThreads, thread i < M * nterms
// reads the flat elements i, i + M * nterms, ... (coalesced, four loads in
// flight per step, all of them belonging to term i % nterms); the M values of
// each term are then combined by an LDS tree.  The sums are left in
// sm[0 .. nterms-1], valid for every thread after the call.  sm holds
// kFinalThreads doubles.  The order of the additions depends only on nb and
// nterms, so the result is reproducible.
__device__ __forceinline__ void sum_columns(const double *__restrict__ p, int nb, int nterms, double *sm) {
    int m = kFinalThreads / nterms;
    m = 1 << (31 - __clz(m));
    const int nth = m * nterms;
    const long long total = (long long)nb * nterms;
    const int tid = threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (tid < nth) {
        long long i = tid;
        for (; i + 3LL * nth < total; i += 4LL * nth) {
            a0 += p[i];
            a1 += p[i + nth];
            a2 += p[i + 2LL * nth];
            a3 += p[i + 3LL * nth];
        }
        for (; i < total; i += nth) a0 += p[i];
    }
    sm[tid] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    for (int h = m >> 1; h >= 1; h >>= 1) {
        if (tid < h * nterms) sm[tid] += sm[tid + h * nterms];
        __syncthreads();
    }
}

void reduce_final_sum(const double *partials, int nblocks, int nterms, double *out, const double *coef,
                      double add, hipStream_t s);

}  // namespace ofdft
