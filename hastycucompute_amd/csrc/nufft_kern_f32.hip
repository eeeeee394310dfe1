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
        case 32: return reinterpret_cast<void*>(&spread_kernel_w6<float2, 32>);
    }
    throw Error(HASTY_ERR_INTERNAL, "no spreading kernel for this transform group");
}

void* nufft_spread_lanes_ptr(int w, int group) {
    switch (w * 64 + group) {
#define HASTY_CASE(Wv)                                                                          \
    case Wv * 64 + 16: return reinterpret_cast<void*>(&spread_lanes_kernel<Wv, 16>);           \
    case Wv * 64 + 32: return reinterpret_cast<void*>(&spread_lanes_kernel<Wv, 32>);
        HASTY_CASE(2) HASTY_CASE(3) HASTY_CASE(4) HASTY_CASE(5) HASTY_CASE(6) HASTY_CASE(7)
        HASTY_CASE(8) HASTY_CASE(9)
#undef HASTY_CASE
    }
    return nullptr;
}

void* nufft_interp_pair_ptr(int dim, int w) {
    // 3-D only: a 2-D window is one z plane, whose 16-B reads the compiler
    // hoists together past the register budget
    if (dim == 3) {
        switch (w) {
            case 2: return reinterpret_cast<void*>(&interp_pair_kernel<2, 16, 16, 16>);
            case 3: return reinterpret_cast<void*>(&interp_pair_kernel<3, 16, 16, 16>);
            case 4: return reinterpret_cast<void*>(&interp_pair_kernel<4, 16, 16, 16>);
            case 5: return reinterpret_cast<void*>(&interp_pair_kernel<5, 16, 16, 16>);
            case 6: return reinterpret_cast<void*>(&interp_pair_kernel<6, 16, 16, 16>);
        }
    }
    return nullptr;
}

template <> void* nufft_interp_fixed_ptr<float2>(int dim, int w) {
    if (dim == 3) {
        switch (w) {
            case 2: return reinterpret_cast<void*>(&interp_kernel<float2, 2, 16, 16, 16>);
            case 3: return reinterpret_cast<void*>(&interp_kernel<float2, 3, 16, 16, 16>);
            case 4: return reinterpret_cast<void*>(&interp_kernel<float2, 4, 16, 16, 16>);
            case 5: return reinterpret_cast<void*>(&interp_kernel<float2, 5, 16, 16, 16>);
            case 6: return reinterpret_cast<void*>(&interp_kernel<float2, 6, 16, 16, 16>);
            case 7: return reinterpret_cast<void*>(&interp_kernel<float2, 7, 16, 16, 16>);
            case 8: return reinterpret_cast<void*>(&interp_kernel<float2, 8, 16, 16, 16>);
            case 9: return reinterpret_cast<void*>(&interp_kernel<float2, 9, 16, 16, 16>);
        }
    } else {
        switch (w) {
            case 2: return reinterpret_cast<void*>(&interp_kernel<float2, 2, 1, 32, 32>);
            case 3: return reinterpret_cast<void*>(&interp_kernel<float2, 3, 1, 32, 32>);
            case 4: return reinterpret_cast<void*>(&interp_kernel<float2, 4, 1, 32, 32>);
            case 5: return reinterpret_cast<void*>(&interp_kernel<float2, 5, 1, 32, 32>);
            case 6: return reinterpret_cast<void*>(&interp_kernel<float2, 6, 1, 32, 32>);
            case 7: return reinterpret_cast<void*>(&interp_kernel<float2, 7, 1, 32, 32>);
            case 8: return reinterpret_cast<void*>(&interp_kernel<float2, 8, 1, 32, 32>);
            case 9: return reinterpret_cast<void*>(&interp_kernel<float2, 9, 1, 32, 32>);
        }
    }
    return nullptr;
}

}  // namespace hasty
