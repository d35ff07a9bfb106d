// Size-specialised FFT passes instantiated for double.
#include "fft_fixed.h"

namespace ofdft {
OFDFT_FIXED_INST(double)
}  // namespace ofdft
