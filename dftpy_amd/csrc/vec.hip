// FP64-accumulated reductions over grid fields (MP.einsum / MP.asum,
// mpi/mpi.py:199-246, serial semantics) and small vector kernels used by the
// optimizer and propagator.
#include "internal.h"
#include "reduce.h"

namespace ofdft {

int reduce_blocks(long long n) {
    long long b = (n + 256 * 8 - 1) / (256 * 8);
    if (b < 1) b = 1;
    if (b > kRedBlocks) b = kRedBlocks;
    return (int)b;
}

// partials[blockIdx.x] = sum a*b*c over a grid-stride share (b, c optional)
template <typename T>
__global__ void k_dot(const T *__restrict__ a, const T *__restrict__ b, const T *__restrict__ c, long long n,
                      double *partials) {
    __shared__ double sm[16];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double v = (double)a[i];
        if (b) v *= (double)b[i];
        if (c) v *= (double)c[i];
        acc += v;
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

struct Coefs {
    double c[kMaxTerms];
    double add;  // added to out[nterms] (sum of the scaled terms) when sum_out
    int sum_out;
};

// out[t] = coef[t] * sum_b partials[b * nterms + t]; optionally out[nterms] =
// sum_t out[t] + add.
__global__ void __launch_bounds__(kFinalThreads)
    k_reduce_final(const double *__restrict__ partials, int nblocks, int nterms, double *out, Coefs co,
                   const double *gate, const double *pub_src, double *pub_dst, int pub_n) {
    if (gate_off(gate)) return;
    __shared__ double sm[kFinalThreads];
    sum_columns(partials, nblocks, nterms, sm);
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int t = 0; t < nterms; ++t) {
            out[t] = sm[t] * co.c[t];
            tot += sm[t] * co.c[t];
        }
        if (co.sum_out) out[nterms] = tot + co.add;
    }
    if (pub_n > 0) {
        __syncthreads();  // thread 0's outputs are visible to the workgroup
        if ((int)threadIdx.x < pub_n) pub_dst[threadIdx.x] = pub_src[threadIdx.x];
    }
}

void reduce_final(const double *partials, int nblocks, int nterms, double *out, const double *coef,
                  hipStream_t s, const double *gate, const double *pub_src, double *pub_dst, int pub_n) {
    OFDFT_REQUIRE(pub_n >= 0 && pub_n <= kFinalThreads && (pub_n == 0 || (pub_src && pub_dst)),
                  "reduce_final: bad publish range");
    Coefs co{};
    for (int t = 0; t < nterms && t < kMaxTerms; ++t) co.c[t] = coef ? coef[t] : 1.0;
    co.sum_out = 0;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kFinalThreads), 0, s, partials, nblocks, nterms, out, co, gate,
                       pub_src, pub_dst, pub_n);
    OFDFT_HIP(hipGetLastError());
}

void reduce_final_sum(const double *partials, int nblocks, int nterms, double *out, const double *coef,
                      double add, hipStream_t s) {
    Coefs co{};
    for (int t = 0; t < nterms && t < kMaxTerms; ++t) co.c[t] = coef ? coef[t] : 1.0;
    co.sum_out = 1;
    co.add = add;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kFinalThreads), 0, s, partials, nblocks, nterms, out, co,
                       (const double *)nullptr, (const double *)nullptr, (double *)nullptr, 0);
    OFDFT_HIP(hipGetLastError());
}

template <typename T>
void dot(Plan &P, const T *a, const T *b, const T *c, long long n, double *out, double scale, hipStream_t s) {
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL((k_dot<T>), dim3(nb), dim3(256), 0, s, a, b, c, n, P.red);
    OFDFT_HIP(hipGetLastError());
    reduce_final(P.red, nb, 1, out, &scale, s);
}

template void dot<float>(Plan &, const float *, const float *, const float *, long long, double *, double,
                         hipStream_t);
template void dot<double>(Plan &, const double *, const double *, const double *, long long, double *, double,
                          hipStream_t);

}  // namespace ofdft
