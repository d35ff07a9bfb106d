// Complex helpers and in-register small DFTs (radix 2, 3, 4, 5, 7, 8, 11, 13)
// shared by the FFT kernels.  Forward = exp(-2 pi i jk/R); INV conjugates.
#pragma once
#include <hip/hip_runtime.h>

namespace ofdft {

// ------------------------------------------------------------------ complex
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C mul_mi(C a) {
    return INV ? C{-a.y, a.x} : C{a.y, -a.x};
}

// ------------------------------------------------------------------ small DFTs
template <bool INV, typename C> __device__ __forceinline__ void dft2(C &a, C &b) {
    C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV, typename C> __device__ __forceinline__ void dft4(C &v0, C &v1, C &v2, C &v3) {
    C t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = mul_mi<INV>(csub(v1, v3));
    v0 = cadd(t0, t2);
    v2 = csub(t0, t2);
    v1 = cadd(t1, t3);
    v3 = csub(t1, t3);
}

template <int R, bool INV, typename C> struct DFT;

template <bool INV, typename C> struct DFT<2, INV, C> {
    __device__ __forceinline__ static void run(C *v) { dft2<INV>(v[0], v[1]); }
};
template <bool INV, typename C> struct DFT<4, INV, C> {
    __device__ __forceinline__ static void run(C *v) { dft4<INV>(v[0], v[1], v[2], v[3]); }
};
template <bool INV, typename C> struct DFT<3, INV, C> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(v[0].x);
        const T s = (T)0.86602540378443864676;  // sin(2 pi / 3)
        C t = cadd(v[1], v[2]);
        C m = C{v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y};
        C d = csub(v[1], v[2]);
        C sd = mul_mi<INV>(C{s * d.x, s * d.y});  // -i s d (forward)
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, sd);
        v[2] = csub(m, sd);
    }
};
template <bool INV, typename C> struct DFT<5, INV, C> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(v[0].x);
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        C a1 = C{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
        C a2 = C{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
        C b1 = mul_mi<INV>(C{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
        C b2 = mul_mi<INV>(C{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
        v[0] = C{v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y};
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    }
};
template <bool INV, typename C> struct DFT<8, INV, C> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(v[0].x);
        const T h = (T)0.70710678118654752440;
        C a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
        C b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
        dft4<INV>(a0, a1, a2, a3);
        dft4<INV>(b0, b1, b2, b3);
        // b_k *= W8^k (forward W8 = e^{-i pi/4}), conjugated for the inverse
        if (!INV) {
            b1 = C{h * (b1.x + b1.y), h * (b1.y - b1.x)};
            b2 = C{b2.y, -b2.x};
            b3 = C{h * (b3.y - b3.x), -h * (b3.x + b3.y)};
        } else {
            b1 = C{h * (b1.x - b1.y), h * (b1.x + b1.y)};
            b2 = C{-b2.y, b2.x};
            b3 = C{-h * (b3.x + b3.y), h * (b3.x - b3.y)};
        }
        v[0] = cadd(a0, b0);
        v[4] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[5] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[6] = csub(a2, b2);
        v[3] = cadd(a3, b3);
        v[7] = csub(a3, b3);
    }
};

// Odd-prime radix by the symmetric direct DFT: for k = 1..(R-1)/2
// a_k = x0 + sum_j cos(2 pi jk/R)(x_j + x_{R-j}), b_k = sum_j sin(2 pi jk/R)(x_j - x_{R-j}),
// y_k = a_k - i b_k, y_{R-k} = a_k + i b_k (forward; conjugate for inverse).
template <int R> struct OddTab;
template <> struct OddTab<7> {
    static constexpr double c[3] = {6.23489801858733594386e-01, -2.22520933956314337365e-01, -9.00968867902419034976e-01};
    static constexpr double s[3] = {7.81831482468029803634e-01, 9.74927912181823619342e-01, 4.33883739117558231424e-01};
};
template <> struct OddTab<11> {
    static constexpr double c[5] = {8.41253532831181205509e-01, 4.15415013001886435084e-01, -1.42314838273285004799e-01, -6.54860733945284989588e-01, -9.59492973614497368651e-01};
    static constexpr double s[5] = {5.40640817455597555430e-01, 9.09631995354518330110e-01, 9.89821441880932795243e-01, 7.55749574354258268905e-01, 2.81732556841429671035e-01};
};
template <> struct OddTab<13> {
    static constexpr double c[6] = {8.85456025653209910509e-01, 5.68064746731155922888e-01, 1.20536680255323006006e-01, -3.54604887042535454889e-01, -7.48510748171101192305e-01, -9.70941817426052011797e-01};
    static constexpr double s[6] = {4.64723172043768506523e-01, 8.22983865893656352242e-01, 9.92708874098053972723e-01, 9.35016242685414833424e-01, 6.63122658240795193052e-01, 2.39315664287557683387e-01};
};

template <int R, bool INV, typename C> struct DFTOdd {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(v[0].x);
        constexpr int H = (R - 1) / 2;
        C sp[H], sm[H];
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            sp[j - 1] = cadd(v[j], v[R - j]);
            sm[j - 1] = csub(v[j], v[R - j]);
        }
        C y0 = v[0];
#pragma unroll
        for (int j = 0; j < H; ++j) y0 = cadd(y0, sp[j]);
        C out[R];
        out[0] = y0;
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            C a = v[0], b = C{(T)0, (T)0};
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const int m = (j * k) % R;  // cos/sin of 2 pi m / R with symmetry
                const int mm = m <= H ? m : R - m;
                const T cj = (T)OddTab<R>::c[mm - 1];
                const T sj = (T)(m <= H ? OddTab<R>::s[mm - 1] : -OddTab<R>::s[mm - 1]);
                a = C{a.x + cj * sp[j - 1].x, a.y + cj * sp[j - 1].y};
                b = C{b.x + sj * sm[j - 1].x, b.y + sj * sm[j - 1].y};
            }
            const C ib = mul_mi<INV>(b);  // -i b (forward)
            out[k] = cadd(a, ib);
            out[R - k] = csub(a, ib);
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = out[k];
    }
};
template <bool INV, typename C> struct DFT<7, INV, C> : DFTOdd<7, INV, C> {};
template <bool INV, typename C> struct DFT<11, INV, C> : DFTOdd<11, INV, C> {};
template <bool INV, typename C> struct DFT<13, INV, C> : DFTOdd<13, INV, C> {};

}  // namespace ofdft
