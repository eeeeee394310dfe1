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

template <> void* nufft_interp_fixed_ptr<double2>(int dim, int w) {
    if (dim == 3) {
        switch (w) {
            case 2: return reinterpret_cast<void*>(&interp_kernel<double2, 2, 16, 16, 16>);
            case 3: return reinterpret_cast<void*>(&interp_kernel<double2, 3, 16, 16, 16>);
            case 4: return reinterpret_cast<void*>(&interp_kernel<double2, 4, 16, 16, 16>);
            case 5: return reinterpret_cast<void*>(&interp_kernel<double2, 5, 16, 16, 16>);
            case 6: return reinterpret_cast<void*>(&interp_kernel<double2, 6, 16, 16, 16>);
            case 7: return reinterpret_cast<void*>(&interp_kernel<double2, 7, 16, 16, 16>);
            case 8: return reinterpret_cast<void*>(&interp_kernel<double2, 8, 16, 16, 16>);
            case 9: return reinterpret_cast<void*>(&interp_kernel<double2, 9, 16, 16, 16>);
            case 10: return reinterpret_cast<void*>(&interp_kernel<double2, 10, 16, 16, 16>);
            case 11: return reinterpret_cast<void*>(&interp_kernel<double2, 11, 16, 16, 16>);
            case 12: return reinterpret_cast<void*>(&interp_kernel<double2, 12, 16, 16, 16>);
            case 13: return reinterpret_cast<void*>(&interp_kernel<double2, 13, 16, 16, 16>);
            case 14: return reinterpret_cast<void*>(&interp_kernel<double2, 14, 16, 16, 16>);
        }
    } else {
        switch (w) {
            case 2: return reinterpret_cast<void*>(&interp_kernel<double2, 2, 1, 32, 32>);
            case 3: return reinterpret_cast<void*>(&interp_kernel<double2, 3, 1, 32, 32>);
            case 4: return reinterpret_cast<void*>(&interp_kernel<double2, 4, 1, 32, 32>);
            case 5: return reinterpret_cast<void*>(&interp_kernel<double2, 5, 1, 32, 32>);
            case 6: return reinterpret_cast<void*>(&interp_kernel<double2, 6, 1, 32, 32>);
            case 7: return reinterpret_cast<void*>(&interp_kernel<double2, 7, 1, 32, 32>);
            case 8: return reinterpret_cast<void*>(&interp_kernel<double2, 8, 1, 32, 32>);
            case 9: return reinterpret_cast<void*>(&interp_kernel<double2, 9, 1, 32, 32>);
            case 10: return reinterpret_cast<void*>(&interp_kernel<double2, 10, 1, 32, 32>);
            case 11: return reinterpret_cast<void*>(&interp_kernel<double2, 11, 1, 32, 32>);
            case 12: return reinterpret_cast<void*>(&interp_kernel<double2, 12, 1, 32, 32>);
            case 13: return reinterpret_cast<void*>(&interp_kernel<double2, 13, 1, 32, 32>);
            case 14: return reinterpret_cast<void*>(&interp_kernel<double2, 14, 1, 32, 32>);
        }
    }
    return nullptr;
}

}  // namespace hasty
