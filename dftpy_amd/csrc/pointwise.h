// Point-wise powers of the density shared by the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace ofdft {

// x^beta and x^(beta-1) for x > 0.  For the WT default beta = 5/6 these are
// sqrt(x) cbrt(x) and cbrt(x) / sqrt(x): one square root and one cube root
// instead of a general pow (log, multiply, exp).
template <typename T> __device__ __forceinline__ T pow_beta(T x, T beta, bool b56) {
    return b56 ? sqrt(x) * cbrt(x) : pow(x, beta);
}
template <typename T> __device__ __forceinline__ T pow_beta_m1(T x, T beta, bool b56) {
    return b56 ? cbrt(x) / sqrt(x) : pow(x, beta - (T)1);
}

inline bool beta_is_5_6(double beta) { return std::fabs(beta - 5.0 / 6.0) < 1e-14; }

}  // namespace ofdft
