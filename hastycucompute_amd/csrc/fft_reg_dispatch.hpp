// Host dispatch of the register two-level FFT pass for one precision.
// Included by exactly one translation unit per precision.
#pragma once

#include <atomic>
#include <cstdlib>
#include <mutex>

#include "fft_reg.hpp"

namespace hasty {

namespace {

template <class V, int L1, int L2, int SIGN, int MODE, int SPEC>
void launch_one(const PassArgs<V>& a, const char* tag, hipStream_t s) {
    constexpr int T = RegShape<L1, L2>::T;
    const int64_t nb_blocks = ceil_div(a.nbatch, a.bchunk);
    const int64_t blocks = ceil_div(a.ni, a.tx) * int64_t(a.no) * nb_blocks;
    if (blocks == 0) return;
    const size_t lds = (size_t(a.tx) * size_t(a.ldp) + size_t(2 * L1 * L2)) * sizeof(V);
    auto kern = fft2_pass_kernel<V, L1, L2, SIGN, MODE, SPEC>;
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
    if (MODE == PM_ROW_STORE && a.sum_batch && a.groups > 1) {
        const int64_t n = a.ni * int64_t(a.out_len);
        hipLaunchKernelGGL(sum_partials_kernel<V>, dim3(unsigned(std::min<int64_t>(ceil_div(n, 256), 4096))),
                           dim3(256), 0, s, static_cast<const V*>(a.partial), a.out, int64_t(a.ni), a.out_len,
                           a.out_i, a.groups, a.accumulate, int64_t(0));
        HASTY_CHECK_HIP(hipGetLastError());
    }
}

// The half-line specialisations are built for the direction the Toeplitz
// passes use them in (pad + forward transform, inverse transform + crop);
// any other combination runs the general kernel.
template <class V, int L1, int L2, int SIGN>
void launch_mode(const PassArgs<V>& a, int mode, int spec, const char* tag, hipStream_t s) {
    constexpr int F = -1, I = 1;   // forward / inverse signs
    switch (mode) {
        case PM_COL:
            if (SIGN == F && spec == PS_HALF_IN) return launch_one<V, L1, L2, SIGN, PM_COL, PS_HALF_IN>(a, tag, s);
            if (SIGN == I && spec == PS_HALF_OUT) return launch_one<V, L1, L2, SIGN, PM_COL, PS_HALF_OUT>(a, tag, s);
            if (SIGN == I && spec == PS_HALF_OUT0) return launch_one<V, L1, L2, SIGN, PM_COL, PS_HALF_OUT0>(a, tag, s);
            return launch_one<V, L1, L2, SIGN, PM_COL, PS_NONE>(a, tag, s);
        case PM_COL_PSF:
            if (SIGN == F && spec == (PS_HALF_IN | PS_HALF_OUT))
                return launch_one<V, L1, L2, SIGN, PM_COL_PSF, PS_HALF_IN | PS_HALF_OUT>(a, tag, s);
            return launch_one<V, L1, L2, SIGN, PM_COL_PSF, PS_NONE>(a, tag, s);
        case PM_ROW_LOAD:
            if (SIGN == F && spec == PS_HALF_IN) return launch_one<V, L1, L2, SIGN, PM_ROW_LOAD, PS_HALF_IN>(a, tag, s);
            return launch_one<V, L1, L2, SIGN, PM_ROW_LOAD, PS_NONE>(a, tag, s);
        default:
            if (SIGN == I && spec == PS_HALF_OUT) return launch_one<V, L1, L2, SIGN, PM_ROW_STORE, PS_HALF_OUT>(a, tag, s);
            if (SIGN == I && spec == PS_HALF_OUT0) return launch_one<V, L1, L2, SIGN, PM_ROW_STORE, PS_HALF_OUT0>(a, tag, s);
            return launch_one<V, L1, L2, SIGN, PM_ROW_STORE, PS_NONE>(a, tag, s);
    }
}

// Does launch_mode have a kernel for this (mode, sign, spec)?  Otherwise the
// general kernel runs, which must then get neither the spec nor a twist.
inline bool spec_built(int mode, int sign, int spec) {
    if (spec == PS_NONE) return true;
    switch (mode) {
        case PM_COL:
            return (sign < 0 && spec == PS_HALF_IN) || (sign > 0 && (spec == PS_HALF_OUT || spec == PS_HALF_OUT0));
        case PM_COL_PSF: return sign < 0 && spec == (PS_HALF_IN | PS_HALF_OUT);
        case PM_ROW_LOAD: return sign < 0 && spec == PS_HALF_IN;
        default: return sign > 0 && (spec == PS_HALF_OUT || spec == PS_HALF_OUT0);
    }
}

// Largest byte offset a thread forms inside one (batch, o) slab: the kernel
// addresses memory as a 64-bit slab base plus a 32-bit offset.
inline bool offsets_fit_32(int64_t ni, int64_t line_stride, int64_t len, int64_t elem_stride, size_t vb) {
    if (ni <= 0 || len <= 0) return true;
    const int64_t maxoff = ((ni - 1) * line_stride + (len - 1) * elem_stride + 1) * int64_t(vb);
    return line_stride >= 0 && elem_stride >= 0 && maxoff < (int64_t(1) << 32);
}

template <class V, int L1, int L2>
bool launch_shape(PassArgs<V> a, int sign, int mode, const char* tag, hipStream_t s) {
    constexpr int T = RegShape<L1, L2>::T;
    constexpr int L = L1 * L2;
    a.ldp = L + T + 1;
    const bool rows = mode == PM_ROW_LOAD || mode == PM_ROW_STORE;
    HASTY_REQUIRE(!(mode == PM_ROW_STORE && a.sum_batch) ||
                      a.out_len <= ((L / 2 + T - 1) / T) * T,
                  "coil-sum pass supports outputs up to half the line");
    const size_t vb = sizeof(V);
    bool fit = offsets_fit_32(a.ni, a.in_i, a.in_len, a.in_s, vb) &&
               offsets_fit_32(a.ni, a.out_i, a.out_len, a.out_s, vb);
    if (a.psf) fit = fit && offsets_fit_32(a.ni, a.psf_i, L, a.psf_s, vb);
    for (const DiagOp<V>* d : {&a.lm1, &a.lm2, &a.sm1, &a.sm2, &a.sio})
        if (d->ptr) fit = fit && offsets_fit_32(a.ni, d->sr, L, 1, vb);
    if (!fit) return false;   // the Stockham kernels take 64-bit offsets
    a.in_ib = uint32_t(a.in_i * int64_t(vb));
    a.in_sb = uint32_t(a.in_s * int64_t(vb));
    a.out_ib = uint32_t(a.out_i * int64_t(vb));
    a.out_sb = uint32_t(a.out_s * int64_t(vb));
    a.psf_ib = uint32_t(a.psf_i * int64_t(vb));
    a.psf_sb = uint32_t(a.psf_s * int64_t(vb));
    a.lm1_rb = uint32_t(a.lm1.sr * int64_t(vb));
    a.lm2_rb = uint32_t(a.lm2.sr * int64_t(vb));
    a.sm1_rb = uint32_t(a.sm1.sr * int64_t(vb));
    a.sm2_rb = uint32_t(a.sm2.sr * int64_t(vb));
    a.sio_rb = uint32_t(a.sio.sr * int64_t(vb));
    int spec = PS_NONE;
    a.twist_in = a.twist_out = 0;
    // a half line at shift -L/4 (positions [3L/4, L) and [0, L/4)) runs as the
    // half line at shift 0 with a quarter-turn twist on the other side
    const int64_t in_sh = ((int64_t(a.in_shift) % L) + L) % L, out_sh = ((int64_t(a.out_shift) % L) + L) % L;
    if (L % 2 == 0 && a.in_shift == 0 && a.in_len == L / 2) {
        spec |= PS_HALF_IN;
    } else if (L % 4 == 0 && L1 % 4 == 0 && a.in_len == L / 2 && in_sh == 3 * L / 4) {
        spec |= PS_HALF_IN;
        a.twist_out = 1;
    }
    if (L % 2 == 0 && a.out_shift == L / 2 - 1 && a.out_len == L / 2) {
        spec |= PS_HALF_OUT;
    } else if (L % 4 == 0 && L2 % 4 == 0 && a.out_len == L / 2 && out_sh == 3 * L / 4) {
        spec |= PS_HALF_OUT0;
        a.twist_in = 1;
    }
    if (!spec_built(mode, sign, spec) || (a.twist_in && a.twist_out)) {
        spec = PS_NONE;
        a.twist_in = a.twist_out = 0;
    }
    // HASTY_COL_TX / HASTY_PSF_TX: lines per block of the column passes
    // (tuning; defaults measured at 128^3 x 32 coils: 32 columns = 256
    // contiguous bytes of c64 per row for the plain column passes, 16 for
    // the PSF pass, whose larger register and LDS footprint per line would
    // leave one 512-thread block per CU)
    static const int env_col_tx = std::getenv("HASTY_COL_TX") ? std::atoi(std::getenv("HASTY_COL_TX")) : 0;
    static const int env_psf_tx = std::getenv("HASTY_PSF_TX") ? std::atoi(std::getenv("HASTY_PSF_TX")) : 0;
    // HASTY_PSF_BLOCKS: block count the PSF / row-load batch chunks aim for (tuning)
    static const int64_t env_blocks = std::getenv("HASTY_PSF_BLOCKS") ? std::atoll(std::getenv("HASTY_PSF_BLOCKS")) : 0;
    if (rows) {
        a.tx = std::max(1, 256 / T);
    } else if (mode == PM_COL_PSF && env_psf_tx > 0) {
        a.tx = env_psf_tx;
    } else if (env_col_tx > 0) {
        a.tx = env_col_tx;
    } else if (a.tx <= 0) {
        a.tx = mode == PM_COL_PSF ? 16 : 32;
    }
    while (a.tx > 1 && (size_t(a.tx) * a.ldp + size_t(2 * L)) * sizeof(V) > 160 * 1024) a.tx /= 2;
    while (a.tx * T > 512) a.tx /= 2;
    // batch elements per block: the PSF pass and the half-line row load
    // take as many as keep >= 4096 blocks in the launch (16 per CU) and read
    // per-line data (the PSF kernel column, broadcast diagonals) once per
    // chunk; the other passes run one element per block
    a.groups = 1;
    if (mode == PM_ROW_STORE && a.sum_batch) {
        // a batch sum: each block sums the elements of one group of its rows;
        // several groups (partials in the scratch, added by a second kernel)
        // when one block per row tile would leave the GPU under-filled
        const int64_t row_blocks = ceil_div(a.ni, a.tx);
        int64_t groups = std::min<int64_t>(a.nbatch, std::max<int64_t>(1, ceil_div(2048, row_blocks)));
        const int64_t per = ceil_div(a.nbatch, groups);
        groups = ceil_div(a.nbatch, per);
        if (groups > 1 && a.partial && groups * a.ni * int64_t(a.out_len) <= a.partial_elems) {
            a.groups = int(groups);
            a.bchunk = int(per);
        } else {
            a.bchunk = a.nbatch;
        }
    } else if (!(mode == PM_COL_PSF || (mode == PM_ROW_LOAD && sign < 0 && (spec & PS_HALF_IN)))) {
        a.bchunk = 1;
    } else {
        const int64_t line_blocks = ceil_div(a.ni, a.tx) * int64_t(a.no);
        const int64_t target = env_blocks > 0 ? env_blocks : 4096;
        const int64_t nchunks = std::min<int64_t>(a.nbatch, std::max<int64_t>(1, ceil_div(target, line_blocks)));
        a.bchunk = int(std::max<int64_t>(1, ceil_div(a.nbatch, nchunks)));
        // frames: a chunk must not straddle two frames' kernels
        if (mode == PM_COL_PSF && a.psf_div > 0) {
            int c = std::min(a.bchunk, a.psf_div);
            while (a.psf_div % c != 0) --c;
            a.bchunk = c;
        }
    }
    if (sign < 0) launch_mode<V, L1, L2, -1>(a, mode, spec, tag, s);
    else launch_mode<V, L1, L2, 1>(a, mode, spec, tag, s);
    return true;
}

}  // namespace

template <> bool launch_fft2_pass<HASTY_FFT2_V>(PassArgs<HASTY_FFT2_V> a, int L, int sign, int rows,
                                                 const char* tag, hipStream_t s) {
    using V = HASTY_FFT2_V;
    switch (L) {
        case 16: return launch_shape<V, 4, 4>(a, sign, rows, tag, s);
        case 20: return launch_shape<V, 5, 4>(a, sign, rows, tag, s);
        case 24: return launch_shape<V, 6, 4>(a, sign, rows, tag, s);
        case 32: return launch_shape<V, 8, 4>(a, sign, rows, tag, s);
        case 40: return launch_shape<V, 8, 5>(a, sign, rows, tag, s);
        case 48: return launch_shape<V, 8, 6>(a, sign, rows, tag, s);
        case 64: return launch_shape<V, 8, 8>(a, sign, rows, tag, s);
        case 80: return launch_shape<V, 10, 8>(a, sign, rows, tag, s);
        case 96: return launch_shape<V, 12, 8>(a, sign, rows, tag, s);
        case 128: return launch_shape<V, 16, 8>(a, sign, rows, tag, s);
        case 160: return launch_shape<V, 16, 10>(a, sign, rows, tag, s);
        case 192: return launch_shape<V, 16, 12>(a, sign, rows, tag, s);
        case 256: return launch_shape<V, 16, 16>(a, sign, rows, tag, s);
        case 320: return launch_shape<V, 20, 16>(a, sign, rows, tag, s);
        case 384: return launch_shape<V, 24, 16>(a, sign, rows, tag, s);
        case 512: return launch_shape<V, 32, 16>(a, sign, rows, tag, s);
        case 640: return launch_shape<V, 32, 20>(a, sign, rows, tag, s);
        default: return false;
    }
}

}  // namespace hasty
