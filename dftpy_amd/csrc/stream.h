// Element loops and block reductions of the point-wise (streaming) kernels.
//
// A streaming kernel reads a few grid fields at the same index and writes a
// few; it is bound by how many bytes each wave keeps in flight.  With VEC = 4
// (n % 4 == 0) a thread takes 4 consecutive elements per step, which the
// compiler turns into one 16-byte load or store per field (two for FP64), so
// a wave moves 1 KB (2 KB) per field access instead of 256 B (512 B).
// VEC = 1 is the plain grid-stride loop for other lengths.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <initializer_list>

namespace ofdft {

template <int VEC, typename F> __device__ __forceinline__ void pw_loop(long long n, F &&body) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (VEC == 4) {
        const long long ng = n >> 2;
        for (long long g = t0; g < ng; g += stride) {
#pragma unroll
            for (int k = 0; k < 4; ++k) body(4 * g + k);
        }
    } else {
        for (long long i = t0; i < n; i += stride) body(i);
    }
}

// Workgroup sums of K accumulators in one pass (wave shuffles, one LDS
// exchange, one barrier): partials[blockIdx.x * K + t].  The order of the
// additions is fixed by the launch shape.
template <int K> __device__ __forceinline__ void block_sums(const double (&acc)[K], double *partials) {
    __shared__ double sm[K][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < K; ++t) {
        double v = acc[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) sm[t][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int nw = (blockDim.x + 63) >> 6;
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += sm[threadIdx.x][w];
        partials[(long long)blockIdx.x * K + threadIdx.x] = s;
    }
}

// VEC consecutive elements starting at i0 (a multiple of VEC) of a field, as
// one 16-byte access (FP32) or two (FP64) when VEC = 4.  The host launches
// VEC = 4 only for fields whose element i0 is 16-byte aligned (pw_aligned).
template <int VEC, typename T> struct PwVec {
    T v[VEC];
};
template <int VEC, typename T> __device__ __forceinline__ PwVec<VEC, T> pw_ld(const T *__restrict__ p, long long i0) {
    PwVec<VEC, T> r;
    if constexpr (VEC == 4 && sizeof(T) == 4) {
        const float4 a = *reinterpret_cast<const float4 *>(p + i0);
        r.v[0] = a.x;
        r.v[1] = a.y;
        r.v[2] = a.z;
        r.v[3] = a.w;
    } else if constexpr (VEC == 4 && sizeof(T) == 8) {
        const double2 a = *reinterpret_cast<const double2 *>(p + i0);
        const double2 b = *reinterpret_cast<const double2 *>(p + i0 + 2);
        r.v[0] = a.x;
        r.v[1] = a.y;
        r.v[2] = b.x;
        r.v[3] = b.y;
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) r.v[k] = p[i0 + k];
    }
    return r;
}
template <int VEC, typename T>
__device__ __forceinline__ void pw_st(T *__restrict__ p, long long i0, const PwVec<VEC, T> &x) {
    if constexpr (VEC == 4 && sizeof(T) == 4) {
        *reinterpret_cast<float4 *>(p + i0) = float4{x.v[0], x.v[1], x.v[2], x.v[3]};
    } else if constexpr (VEC == 4 && sizeof(T) == 8) {
        *reinterpret_cast<double2 *>(p + i0) = double2{x.v[0], x.v[1]};
        *reinterpret_cast<double2 *>(p + i0 + 2) = double2{x.v[2], x.v[3]};
    } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) p[i0 + k] = x.v[k];
    }
}

// loop over groups of VEC elements: body(i0) with i0 the group's first element
template <int VEC, typename F> __device__ __forceinline__ void pw_groups(long long n, F &&body) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long ng = VEC == 4 ? (n >> 2) : n;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += stride) body(g * VEC);
}

// host side: element width of a launch and the block count of a
// non-reducing streaming kernel of 256 threads
inline int pw_vec(long long n) { return (n & 3) == 0 ? 4 : 1; }
// every (possibly null) pointer 16-byte aligned
inline bool pw_aligned(std::initializer_list<const void *> ps) {
    for (const void *p : ps)
        if (reinterpret_cast<uintptr_t>(p) & 15) return false;
    return true;
}
inline int pw_blocks(long long n) {
    const long long per = 256LL * pw_vec(n);
    return (int)std::max<long long>(1, std::min<long long>((n + per - 1) / per, 16384));
}

}  // namespace ofdft
