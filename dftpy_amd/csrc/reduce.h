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

void reduce_final_sum(const double *partials, int nblocks, int nterms, double *out, const double *coef,
                      double add, hipStream_t s);

}  // namespace ofdft
