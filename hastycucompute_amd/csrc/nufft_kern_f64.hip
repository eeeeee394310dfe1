// Spread / interp kernel instantiations, complex128 (widths 2..14).
#include "nufft_kernels.hpp"

namespace hasty {

template <> void* nufft_kernel_ptr<double2>(int which, int w) {
    switch (w) {
#define HASTY_CASE(Wv)                                                                      \
    case Wv:                                                                                \
        return which == 0   ? reinterpret_cast<void*>(&interp_kernel<double2, Wv>)           \
               : which == 2 ? reinterpret_cast<void*>(&interp_gm_kernel<double2, Wv>)        \
               : which == 3 ? reinterpret_cast<void*>(&spread_gm_kernel<double2, Wv>)        \
                            : reinterpret_cast<void*>(&coord_grad_kernel<double2, Wv>);
        HASTY_CASE(2) HASTY_CASE(3) HASTY_CASE(4) HASTY_CASE(5) HASTY_CASE(6) HASTY_CASE(7)
        HASTY_CASE(8) HASTY_CASE(9) HASTY_CASE(10) HASTY_CASE(11) HASTY_CASE(12) HASTY_CASE(13)
        HASTY_CASE(14)
#undef HASTY_CASE
    }
    throw Error(HASTY_ERR_INVALID, "complex128 plans support kernel widths 2..14 (tol >= 1e-13)");
}

template <> void* nufft_spread_ptr<double2>(int cg) {
    switch (cg) {
        case 1: return reinterpret_cast<void*>(&spread_kernel<double2, 1>);
        case 2: return reinterpret_cast<void*>(&spread_kernel<double2, 2>);
        case 4: return reinterpret_cast<void*>(&spread_kernel<double2, 4>);
        case 8: return reinterpret_cast<void*>(&spread_kernel<double2, 8>);
        case 16: return reinterpret_cast<void*>(&spread_kernel<double2, 16>);
    }
    throw Error(HASTY_ERR_INTERNAL, "no spreading kernel for this transform group");
}

}  // namespace hasty
