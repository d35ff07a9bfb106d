// Size table of the size-specialised FFT passes and the tile count of the
// R2C pass (its number of FP64 partial sums).
#include "fft_fixed.h"

namespace ofdft {

bool fixed_size(int n) {
    switch (n) {
#define CASE(N) case N:
        OFDFT_FIXED_SIZES(CASE)
#undef CASE
        return true;
        default: return false;
    }
}

// number of row tiles of the fixed R2C kernel for `batch` spectra and variant var
int r2c_fixed_tiles(int n, int var, int prec, int batch, long long nrows) {
    int np = 1;
    switch (n) {
#define PENS(T, N, NF) \
    (var == 1 ? r2c_pens_v<T, N, NF, 1>() : (var == 2 ? r2c_pens_v<T, N, NF, 2>() : r2c_pens_v<T, N, NF, 0>()))
#define CASE(N)                                                                      \
    case N:                                                                          \
        if (batch == 3)                                                              \
            np = prec == OFDFT_FP64 ? PENS(double, N, 3) : PENS(float, N, 3);        \
        else if (batch == 2)                                                         \
            np = prec == OFDFT_FP64 ? PENS(double, N, 2) : PENS(float, N, 2);        \
        else                                                                         \
            np = prec == OFDFT_FP64 ? PENS(double, N, 1) : PENS(float, N, 1);        \
        break;
        OFDFT_FIXED_SIZES(CASE)
#undef CASE
#undef PENS
        default: return -1;
    }
    return (int)((nrows + 2 * np - 1) / (2 * np));
}

}  // namespace ofdft
