// Register two-level FFT passes, complex128 instantiations.
#define HASTY_FFT2_V double2
#include "fft_reg_dispatch.hpp"
