// Register two-level FFT passes, complex64 instantiations.
#define HASTY_FFT2_V float2
#include "fft_reg_dispatch.hpp"
