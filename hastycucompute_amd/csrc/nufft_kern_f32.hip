// Spread / interp kernel instantiations, complex64 (widths 2..9).
#include "nufft_kernels.hpp"

namespace hasty {

template <> void* nufft_kernel_ptr<float2>(int which, int w) {
    switch (w) {
#define HASTY_CASE(Wv)                                                                      \
    case Wv:                                                                                \
        return which == 0   ? reinterpret_cast<void*>(&interp_kernel<float2, Wv>)           \
               : which == 2 ? reinterpret_cast<void*>(&interp_gm_kernel<float2, Wv>)        \
               : which == 3 ? reinterpret_cast<void*>(&spread_gm_kernel<float2, Wv>)        \
                            : reinterpret_cast<void*>(&coord_grad_kernel<float2, Wv>);
        HASTY_CASE(2) HASTY_CASE(3) HASTY_CASE(4) HASTY_CASE(5) HASTY_CASE(6) HASTY_CASE(7)
        HASTY_CASE(8) HASTY_CASE(9)
#undef HASTY_CASE
    }
    throw Error(HASTY_ERR_INVALID, "complex64 plans support kernel widths 2..9 (tol >= 1e-8)");
}

template <> void* nufft_spread_ptr<float2>(int cg) {
    switch (cg) {
        case 1: return reinterpret_cast<void*>(&spread_kernel<float2, 1>);
        case 2: return reinterpret_cast<void*>(&spread_kernel<float2, 2>);
        case 4: return reinterpret_cast<void*>(&spread_kernel<float2, 4>);
        case 8: return reinterpret_cast<void*>(&spread_kernel<float2, 8>);
        case 16: return reinterpret_cast<void*>(&spread_kernel<float2, 16>);
        case 32: return reinterpret_cast<void*>(&spread_kernel<float2, 32>);
    }
    throw Error(HASTY_ERR_INTERNAL, "no spreading kernel for this transform group");
}

}  // namespace hasty
