// Host dispatch of the register two-level FFT pass for one precision.
// Included by exactly one translation unit per precision.
#pragma once

#include <atomic>
#include <cstdlib>
#include <mutex>

#include "fft_reg.hpp"

namespace hasty {

namespace {

template <class V, int L1, int L2, int SIGN, int MODE>
void launch_one(const PassArgs<V>& a, const char* tag, hipStream_t s) {
    constexpr int T = RegShape<L1, L2>::T;
    const int64_t nb_blocks = (MODE == PM_ROW_STORE && a.sum_batch) ? 1
                              : MODE == PM_COL_PSF                ? ceil_div(a.nbatch, a.bchunk)
                                                                  : int64_t(a.nbatch);
    const int64_t blocks = ceil_div(a.ni, a.tx) * int64_t(a.no) * nb_blocks;
    if (blocks == 0) return;
    const size_t lds = (size_t(a.tx) * size_t(a.ldp) + size_t(L1 * L2)) * sizeof(V);
    auto kern = fft2_pass_kernel<V, L1, L2, SIGN, MODE>;
    if (lds > 64 * 1024) {
        static std::once_flag once;
        std::call_once(once, [&] {
            HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                160 * 1024));
        });
    }
    HASTY_PROF(tag, s);
    hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(unsigned(a.tx * T)), lds, s, a);
    HASTY_CHECK_HIP(hipGetLastError());
}

template <class V, int L1, int L2, int SIGN>
void launch_mode(const PassArgs<V>& a, int mode, const char* tag, hipStream_t s) {
    switch (mode) {
        case PM_COL: launch_one<V, L1, L2, SIGN, PM_COL>(a, tag, s); break;
        case PM_COL_PSF: launch_one<V, L1, L2, SIGN, PM_COL_PSF>(a, tag, s); break;
        case PM_ROW_LOAD: launch_one<V, L1, L2, SIGN, PM_ROW_LOAD>(a, tag, s); break;
        default: launch_one<V, L1, L2, SIGN, PM_ROW_STORE>(a, tag, s); break;
    }
}

template <class V, int L1, int L2>
void launch_shape(PassArgs<V> a, int sign, int mode, const char* tag, hipStream_t s) {
    constexpr int T = RegShape<L1, L2>::T;
    constexpr int L = L1 * L2;
    a.ldp = L + T + 1;
    const bool rows = mode == PM_ROW_LOAD || mode == PM_ROW_STORE;
    HASTY_REQUIRE(!(mode == PM_ROW_STORE && a.sum_batch) ||
                      a.out_len <= ((L / 2 + T - 1) / T) * T,
                  "coil-sum pass supports outputs up to half the line");
    // HASTY_COL_TX / HASTY_PSF_TX: lines per block of the column passes
    // (tuning; default 16 columns = 128 contiguous bytes of c64 per row)
    static const int env_col_tx = std::getenv("HASTY_COL_TX") ? std::atoi(std::getenv("HASTY_COL_TX")) : 0;
    static const int env_psf_tx = std::getenv("HASTY_PSF_TX") ? std::atoi(std::getenv("HASTY_PSF_TX")) : 0;
    if (rows) {
        a.tx = std::max(1, 256 / T);
    } else if (mode == PM_COL_PSF && env_psf_tx > 0) {
        a.tx = env_psf_tx;
    } else if (env_col_tx > 0) {
        a.tx = env_col_tx;
    } else if (a.tx <= 0) {
        a.tx = 16;
    }
    while (a.tx > 1 && (size_t(a.tx) * a.ldp + size_t(L)) * sizeof(V) > 160 * 1024) a.tx /= 2;
    while (a.tx * T > 512) a.tx /= 2;
    if (mode == PM_COL_PSF) {
        // batch elements per block: as many as keep >= 4096 blocks in the
        // launch, so the kernel columns are read from HBM once per chunk
        const int64_t line_blocks = ceil_div(a.ni, a.tx) * int64_t(a.no);
        const int64_t nchunks = std::min<int64_t>(a.nbatch, std::max<int64_t>(1, ceil_div(4096, line_blocks)));
        a.bchunk = int(ceil_div(a.nbatch, nchunks));
    } else {
        a.bchunk = 1;
    }
    if (sign < 0) launch_mode<V, L1, L2, -1>(a, mode, tag, s);
    else launch_mode<V, L1, L2, 1>(a, mode, tag, s);
}

}  // namespace

template <> bool launch_fft2_pass<HASTY_FFT2_V>(PassArgs<HASTY_FFT2_V> a, int L, int sign, int rows,
                                                 const char* tag, hipStream_t s) {
    using V = HASTY_FFT2_V;
    switch (L) {
        case 16: launch_shape<V, 4, 4>(a, sign, rows, tag, s); return true;
        case 20: launch_shape<V, 5, 4>(a, sign, rows, tag, s); return true;
        case 24: launch_shape<V, 6, 4>(a, sign, rows, tag, s); return true;
        case 32: launch_shape<V, 8, 4>(a, sign, rows, tag, s); return true;
        case 40: launch_shape<V, 8, 5>(a, sign, rows, tag, s); return true;
        case 48: launch_shape<V, 8, 6>(a, sign, rows, tag, s); return true;
        case 64: launch_shape<V, 8, 8>(a, sign, rows, tag, s); return true;
        case 80: launch_shape<V, 10, 8>(a, sign, rows, tag, s); return true;
        case 96: launch_shape<V, 12, 8>(a, sign, rows, tag, s); return true;
        case 128: launch_shape<V, 16, 8>(a, sign, rows, tag, s); return true;
        case 160: launch_shape<V, 16, 10>(a, sign, rows, tag, s); return true;
        case 192: launch_shape<V, 16, 12>(a, sign, rows, tag, s); return true;
        case 256: launch_shape<V, 16, 16>(a, sign, rows, tag, s); return true;
        case 320: launch_shape<V, 20, 16>(a, sign, rows, tag, s); return true;
        case 384: launch_shape<V, 24, 16>(a, sign, rows, tag, s); return true;
        case 512: launch_shape<V, 32, 16>(a, sign, rows, tag, s); return true;
        case 640: launch_shape<V, 32, 20>(a, sign, rows, tag, s); return true;
        default: return false;
    }
}

}  // namespace hasty
