// Size-specialised FFT passes instantiated for float.
#include "fft_fixed.h"

namespace ofdft {
OFDFT_FIXED_INST(float)
}  // namespace ofdft
