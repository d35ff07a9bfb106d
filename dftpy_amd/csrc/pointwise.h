// Point-wise powers of the density shared by the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "fastmath.h"

namespace ofdft {

// x^beta and x^(beta-1) for x > 0.  For the WT default beta = 5/6 these are
// sqrt(x) cbrt(x) and cbrt(x) / sqrt(x): one square root and one cube root
// instead of a general pow (log, multiply, exp).
template <typename T> __device__ __forceinline__ T pow_beta(T x, T beta, bool b56) {
    return b56 ? fm::sqrt(x) * fm::cbrt(x) : fm::pow(x, beta);
}
template <typename T> __device__ __forceinline__ T pow_beta_m1(T x, T beta, bool b56) {
    return b56 ? fm::div(fm::cbrt(x), fm::sqrt(x)) : fm::pow(x, beta - (T)1);
}

inline bool beta_is_5_6(double beta) { return std::fabs(beta - 5.0 / 6.0) < 1e-14; }

// PZ LDA (semilocal_xc.py:252-305, unpolarised) of rho from its cube root
// rc3 = cbrt(max(rho, 1e-30)): energy per electron e and potential v.
template <typename T> __device__ __forceinline__ void pz_lda(T rc3, T &e, T &v) {
    constexpr double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    constexpr double G = -0.1423, B1 = 1.0529, B2 = 0.3334;
    const T rs = fm::div((T)cbrt(3.0 / (4.0 * M_PI)), rc3);
    T ec, vc;
    if (rs < (T)1) {
        const T lr = fm::log(rs);
        ec = (T)A * lr + (T)B + (T)C * rs * lr + (T)D * rs;
        vc = lr * ((T)A + (T)(2.0 / 3.0 * C) * rs) + (T)B - (T)(1.0 / 3.0 * A) + (T)(1.0 / 3.0 * (2 * D - C)) * rs;
    } else {
        const T sr = fm::sqrt(rs);
        const T den = (T)1 + (T)B1 * sr + (T)B2 * rs;
        ec = fm::div((T)G, den);
        vc = fm::div((T)G + (T)(7.0 / 6.0 * G * B1) * sr + (T)(4.0 / 3.0 * G * B2) * rs, den * den);
    }
    e = (T)(-3.0 / 4.0 * cbrt(3.0 / M_PI)) * rc3 + ec;
    v = (T)cbrt(-3.0 / M_PI) * rc3 + vc;
}

}  // namespace ofdft
