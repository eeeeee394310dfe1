// Fused Toeplitz passes of the SENSE normal operator on gfx950, for planes of
// 128 x 128 (2n = 256 per axis):
//   toep_pad_fft_kernel   x and y zero-padded forward FFTs (replaces P1 + P2)
//   toep_crop_sum_kernel  y and x inverse FFTs, crops, output diagonals and
//                         the coil sum (replaces P4 + P5)
// so neither [B, nz, ny, 2nx] intermediate goes through HBM.  Between them
// the z pass (P3) runs unchanged on the grid, reading the kernel through the
// permutation below.
//
// The x axis of the grid is class-major: frequency k = 4m + r sits at
// position r L/4 + m, so class r of a row is 64 contiguous values.  The
// 256-point DFT along x splits by residue classes:
//     X[4m + r] = sum_{n'<64} e^{-2 pi i r n'/256} (x[n'] + (-i)^r x[n'+64]) e^{-2 pi i m n'/64}
//     x[n]      = sum_r e^{+2 pi i r n / 256} inner_r(n mod 64),
//     inner_r(n') = sum_m X[4m + r] e^{+2 pi i m n' / 64},
// so each class costs one 64-point DFT (the same work as one 256-point DFT
// over the four classes).  One block owns a plane z and a group of coils b;
// per (b, r) the inverse pass:
//   A  reads the 64 class columns of the (2ny, L/4) slab (512 B per row),
//      inverse-transforms them along y (two-level 16 x 16, one LDS exchange)
//      and writes the crop [ny-1, 2ny-1) to LDS rows;
//   B  runs each row's 64-point inverse DFT (8 x 8, one exchange), giving
//      inner_r, and accumulates it into the crop [nx-1, 2nx-1): n = n' + 128
//      for every n', and n = n' + 192 for n' < 63 or n = 127 for n' = 63
//      (16 outputs a thread).
// After a coil's four classes the thread scales its 16 pixels, applies the
// output diagonals and adds them to the coil sum.  The forward pass runs the
// same two phases in the other order (rows, then columns).  Traffic per
// apply: the (2ny, 2nx) planes written once and read once, the maps read
// twice, the image (or partial sums) written once.
#include "toeplitz_fused.hpp"

#include <algorithm>

#include "fft_reg.hpp"

namespace hasty {

namespace {
constexpr int kNY = 128, kNX = 128;         // plane shape (ny, nx)
constexpr int kLY = 2 * kNY, kLX = 2 * kNX;  // 256, 256
constexpr int kM = kLX / 4;                 // 64 columns per class
#ifndef HASTY_TOEP_FUSED_H
#define HASTY_TOEP_FUSED_H 2
#endif
constexpr int kH = HASTY_TOEP_FUSED_H;       // lines per thread in both phases of the last pass
// lines per thread of the first pass (one: 96 VGPRs at 1024 threads, 4 waves per SIMD)
#ifndef HASTY_TOEP_F12_H
#define HASTY_TOEP_F12_H 1
#endif
constexpr int kH12 = HASTY_TOEP_F12_H;
constexpr int kThreads = 1024 / kH;
constexpr int kRS = kM + 8;                 // LDS row stride of the class rows
// (the 8 x 8 exchange of a row uses stride 9 within it: the step-1 writes and
// step-2 reads of a wave then touch each bank twice, the minimum for 512 B)
static_assert(7 * 9 + 7 < kRS, "8 x 8 exchange fits in a class row");
constexpr int kE = 16 * 16 * kM;            // y exchange, [k1][p2][m]
constexpr int kR = kNY * kRS;               // one set of class rows
constexpr int kLdsElems = (kE > 2 * kR ? kE : 2 * kR) + kLX;
static_assert(kH * kThreads == 16 * kM && kH * kThreads == 8 * kNY, "thread layout");
}  // namespace

// output pixel of slot i of lane s: slot k2 < 8 -> x = s + 8 k2 + 1, slot
// 8 + k2 -> x = s + 8 k2 + 65 (n' = s + 8 k2 < 63) or x = 0 (n' = 63)
__device__ inline int crop_px(int s, int i) {
    const int np = s + 8 * (i & 7);
    return i < 8 ? np + 1 : (np == 63 ? 0 : np + 65);
}

// v (.) d at element (b, row, x) for the 16 pixels of lane s
__device__ inline void crop_sum_diag(float2* v, const DiagOp<float2>& d, int64_t b, int64_t row, int s) {
    if (!d.ptr) return;
    const float2* base = d.ptr + b * d.sb + row * d.sr;
    const float sgn = d.type == HASTY_MULT_CONJ ? -1.f : 1.f;
    // two rounds of 8 loads (the next columns are in flight in registers)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float2 w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = base[crop_px(s, 8 * h + i)];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[8 * h + i] = cmul_sgn(v[8 * h + i], w[i], sgn);
    }
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 4 / kH)))
toep_crop_sum_kernel(ToepCropSumArgs a, const float2* tw) {
    using V = float2;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    V* E = lds;              // phase A exchange
    V* R1 = lds;             // class rows (after the exchange is read)
    V* R2 = lds + kR;        // phase B step-1 results
    V* twl = lds + (kLdsElems - kLX);
    const int tid = threadIdx.x;
    for (int k = tid; k < kLX; k += kThreads) twl[k] = tw[k];

    const int z = int(blockIdx.x % unsigned(a.nz));
    const int g = int(blockIdx.x / unsigned(a.nz));
    const int cpb = (a.nbatch + a.groups - 1) / a.groups;
    const int b0 = g * cpb;
    const int b1 = min(a.nbatch, b0 + cpb);

    // phase A roles: column m; step-1 lines p2 = pa, pa + 8 and step-2 lines
    // k1 = pa, pa + 8 (pa = wave)
    const int m0 = tid & (kM - 1);
    const int pa0 = tid >> 6;
    // phase B roles: rows y = yb, yb + 64, lane s of the 8 x 8 DFT
    const int yb0 = tid >> 3;
    const int s0 = tid & 7;
    V tot[kH][16];
#pragma unroll
    for (int h = 0; h < kH; ++h)
#pragma unroll
        for (int i = 0; i < 16; ++i) tot[h][i] = make_float2(0.f, 0.f);
    __syncthreads();

    // (coil, class) iterations it = (b - b0) 4 + r; the first half of the loads
    // of iteration it + 1 is issued once iteration it has moved its columns to
    // LDS, so it is in flight during the row transforms and the coil epilogue
    // (the second half waits for the next iteration: registers)
    const int nit = 4 * max(0, b1 - b0);
    V v[kH][16];
    auto load_class = [&](int it, int h0, int h1) {
        int m = m0, pa = pa0;
        asm volatile("" : "+v"(m), "+v"(pa));
        const int bb = b0 + (it >> 2), rr = it & 3;
        const char* plane =
            reinterpret_cast<const char*>(a.grid + int64_t(bb) * a.grid_b + int64_t(z) * (kLY * kLX));
        const uint32_t off = uint32_t((pa * kLX + rr * kM + m) * int(sizeof(V)));
#pragma unroll
        for (int h = 0; h < kH; ++h)
            if (h >= h0 && h < h1)
#pragma unroll
                for (int p1 = 0; p1 < 16; ++p1)
                    v[h][p1] =
                        load_at<V, true>(plane, off + uint32_t((16 * p1 + (16 / kH) * h) * kLX * int(sizeof(V))));
    };
    constexpr int kPre = kH > 1 ? kH / 2 : kH;   // lines per thread prefetched a phase early
    if (nit > 0) load_class(0, 0, kPre);
    V acc[kH][16];
#pragma unroll 1
    for (int it = 0; it < nit; ++it) {
        const int b = b0 + (it >> 2), r = it & 3;
        // per-iteration copies of the lane indices: otherwise every LDS and
        // twiddle address derived from them is hoisted out of the loop and
        // stays live in registers
        int m = m0, pa = pa0, yb = yb0, s = s0;
        asm volatile("" : "+v"(m), "+v"(pa), "+v"(yb), "+v"(s));
        if (r == 0) {
#pragma unroll
            for (int h = 0; h < kH; ++h)
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[h][i] = make_float2(0.f, 0.f);
        }
        if (kPre < kH) load_class(it, kPre, kH);
        // ---- A: y inverse DFT of the class columns ----
#pragma unroll
        for (int h = 0; h < kH; ++h) {
            const int p2 = pa + (16 / kH) * h;
            RegDFT<16, +1, V>::run(v[h]);
#pragma unroll
            for (int k1 = 1; k1 < 16; ++k1) v[h][k1] = twiddle_mul<V, +1>(v[h][k1], twl, (p2 * k1) & (kLX - 1));
        }
        __syncthreads();   // the previous iteration's phase B reads of R1 / R2 are done
#pragma unroll
        for (int h = 0; h < kH; ++h)
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) E[(k1 * 16 + pa + (16 / kH) * h) * kM + m] = v[h][k1];
        __syncthreads();
#pragma unroll
        for (int h = 0; h < kH; ++h) {
            const int k1 = pa + (16 / kH) * h;
#pragma unroll
            for (int p2 = 0; p2 < 16; ++p2) v[h][p2] = E[(k1 * 16 + p2) * kM + m];
            RegDFT<16, +1, V>::run(v[h]);   // v[h][k2] = X[k1 + 16 k2]
        }
        __syncthreads();   // E is read; R1 overwrites it
#pragma unroll
        for (int h = 0; h < kH; ++h) {
#pragma unroll
            for (int k2 = 0; k2 < 16; ++k2) {
                const int j = pa + (16 / kH) * h + 16 * k2;
                if (j >= kNY - 1 && j < 2 * kNY - 1) R1[(j - (kNY - 1)) * kRS + m] = v[h][k2];
            }
        }
        if (it + 1 < nit) load_class(it + 1, 0, kPre);
        __syncthreads();
        // ---- B: 64-point inverse DFT of each class row, 8 x 8 ----
        V w[kH][8];
#pragma unroll
        for (int h = 0; h < kH; ++h) {
            const int y = yb + (128 / kH) * h;
#pragma unroll
            for (int p1 = 0; p1 < 8; ++p1) w[h][p1] = R1[y * kRS + 8 * p1 + s];
            RegDFT<8, +1, V>::run(w[h]);
#pragma unroll
            for (int k1 = 1; k1 < 8; ++k1) w[h][k1] = twiddle_mul<V, +1>(w[h][k1], twl, (4 * s * k1) & (kLX - 1));
#pragma unroll
            for (int k1 = 0; k1 < 8; ++k1) R2[y * kRS + s * 9 + k1] = w[h][k1];
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < kH; ++h) {
            const int y = yb + (128 / kH) * h;
#pragma unroll
            for (int p2 = 0; p2 < 8; ++p2) w[h][p2] = R2[y * kRS + p2 * 9 + s];
            RegDFT<8, +1, V>::run(w[h]);   // w[h][k2] = inner_r(s + 8 k2)
            // combine: acc += e^{+2 pi i r n / 256} inner_r(n mod 64)
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) {
                const int np = s + 8 * k2;
                const int na = np + 128, nb = np == 63 ? 127 : np + 192;
                acc[h][k2] = cadd(acc[h][k2], twiddle_mul<V, +1>(w[h][k2], twl, (r * na) & (kLX - 1)));
                acc[h][8 + k2] = cadd(acc[h][8 + k2], twiddle_mul<V, +1>(w[h][k2], twl, (r * nb) & (kLX - 1)));
            }
        }
        if (r == 3) {
            // ---- coil epilogue: scale, output diagonals, coil sum ----
#pragma unroll
            for (int h = 0; h < kH; ++h) {
                const int64_t row = int64_t(z) * kNY + yb + (128 / kH) * h;
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[h][i] = cscale(acc[h][i], a.scale);
                crop_sum_diag(acc[h], a.m1, b, row, s);
                crop_sum_diag(acc[h], a.m2, b, row, s);
                crop_sum_diag(acc[h], a.io, b, row, s);
#pragma unroll
                for (int i = 0; i < 16; ++i) tot[h][i] = cadd(tot[h][i], acc[h][i]);
            }
        }
    }
    const int yb = yb0, s = s0;
#pragma unroll
    for (int h = 0; h < kH; ++h) {
        const int64_t row = int64_t(z) * kNY + yb + (128 / kH) * h;
        if (a.groups > 1) {
            V* part = a.partial + (int64_t(g) * a.nz * kNY + row) * kNX;
#pragma unroll
            for (int i = 0; i < 16; ++i) part[crop_px(s, i)] = tot[h][i];
        } else {
            V* o = a.out + row * kNX;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int x = crop_px(s, i);
                o[x] = a.accumulate ? cadd(tot[h][i], o[x]) : tot[h][i];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// First two passes fused: x and y zero-padded forward DFTs, x class-major.
// Per plane z and coil b the thread loads its two rows' 16 + 16 inputs once
// (x positions n' and n' + 64, times the diagonals); per class r:
//   B' each row's class: u[n'] = e^{-2 pi i r n' / 256} (x[n'] + (-i)^r
//      x[n' + 64]), a 64-point DFT (8 x 8, one exchange) -> X[4m + r], rows to LDS;
//   A' each class column: 128 inputs zero-padded to 256 (16 x 16 with half
//      the first-step inputs zero), stored at (ky, r L/4 + m): 512 B per wave.
// The next coil's inputs are loaded once the last class has folded them.
// ---------------------------------------------------------------------------
template <int H>
__device__ inline void pad_fft_diag(float2 (*xs)[16], const DiagOp<float2>& d, int64_t b, int z, int yb, int s) {
    if (!d.ptr) return;
    const float sgn = d.type == HASTY_MULT_CONJ ? -1.f : 1.f;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int64_t row = int64_t(z) * kNY + yb + (128 / H) * h;
        const float2* base = d.ptr + b * d.sb + row * d.sr + s;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            float2 w[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) w[i] = base[8 * i + 64 * q];
#pragma unroll
            for (int i = 0; i < 8; ++i) xs[h][8 * q + i] = cmul_sgn(xs[h][8 * q + i], w[i], sgn);
        }
    }
}

template <int H>
__global__ void __launch_bounds__(1024 / H) __attribute__((amdgpu_waves_per_eu(1, 4 / H)))
toep_pad_fft_kernel(ToepPadFftArgs a, const float2* tw, int groups) {
    using V = float2;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    V* E = lds;              // column exchange
    V* R1 = lds;             // class rows (read before the exchange is written)
    V* R2 = lds + kR;        // row 8 x 8 exchange
    V* twl = lds + (kLdsElems - kLX);
    const int tid = threadIdx.x;
    for (int k = tid; k < kLX; k += (1024 / H)) twl[k] = tw[k];

    const int z = int(blockIdx.x % unsigned(a.nz));
    const int g = int(blockIdx.x / unsigned(a.nz));
    const int cpb = (a.nbatch + groups - 1) / groups;
    const int b0 = g * cpb;
    const int b1 = min(a.nbatch, b0 + cpb);
    const int m0 = tid & (kM - 1), pa0 = tid >> 6, yb0 = tid >> 3, s0 = tid & 7;

    // xs[h][i]: row yb + (128 / H) h, x = 8 i + s (i < 8) or 8 (i - 8) + 64 + s
    V xs[H][16];
    auto load_in = [&](int b) {
        int yb = yb0, s = s0;
        asm volatile("" : "+v"(yb), "+v"(s));
        const char* pl = reinterpret_cast<const char*>(a.in + int64_t(b) * a.in_b + int64_t(z) * (kNY * kNX));
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const uint32_t off = uint32_t(((yb + (128 / H) * h) * kNX + s) * int(sizeof(V)));
#pragma unroll
            for (int i = 0; i < 16; ++i)
                xs[h][i] = load_at<V>(pl, off + uint32_t((8 * (i & 7) + 64 * (i >> 3)) * int(sizeof(V))));
        }
    };
    if (b0 < b1) load_in(b0);
    // a broadcast first diagonal (the image of a SENSE apply) is read once
    // per block, not once per coil
    const bool m1_bcast = a.m1.ptr && a.m1.sb == 0;
    V d1[H][16];
    if (m1_bcast) {
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int i = 0; i < 16; ++i) d1[h][i] = make_float2(1.f, 0.f);
        pad_fft_diag<H>(d1, a.m1, 0, z, yb0, s0);
    }
    __syncthreads();

#pragma unroll 1
    for (int b = b0; b < b1; ++b) {
        {
            int yb = yb0, s = s0;
            asm volatile("" : "+v"(yb), "+v"(s));
            if (m1_bcast) {
#pragma unroll
                for (int h = 0; h < H; ++h)
#pragma unroll
                    for (int i = 0; i < 16; ++i) xs[h][i] = cmul(xs[h][i], d1[h][i]);
            } else {
                pad_fft_diag<H>(xs, a.m1, b, z, yb, s);
            }
            pad_fft_diag<H>(xs, a.m2, b, z, yb, s);
        }
        char* gp = reinterpret_cast<char*>(a.grid + int64_t(b) * a.grid_b + int64_t(z) * (kLY * kLX));
#pragma unroll 1
        for (int r = 0; r < 4; ++r) {
            int m = m0, pa = pa0, yb = yb0, s = s0;
            asm volatile("" : "+v"(m), "+v"(pa), "+v"(yb), "+v"(s));
            // ---- B': fold, 64-point DFT of each row's class ----
            V u[H][8];
#pragma unroll
            for (int h = 0; h < H; ++h) {
#pragma unroll
                for (int p1 = 0; p1 < 8; ++p1) {
                    const int np = 8 * p1 + s;
                    V t = cadd(xs[h][p1], quarter_turn(xs[h][8 + p1], 4 - r));
                    u[h][p1] = r ? twiddle_mul<V, -1>(t, twl, (r * np) & (kLX - 1)) : t;
                }
            }
            if (r == 3 && b + 1 < b1) load_in(b + 1);   // this coil's inputs are folded
#pragma unroll
            for (int h = 0; h < H; ++h) {
                RegDFT<8, -1, V>::run(u[h]);
#pragma unroll
                for (int k1 = 1; k1 < 8; ++k1) u[h][k1] = twiddle_mul<V, -1>(u[h][k1], twl, (4 * s * k1) & (kLX - 1));
            }
            __syncthreads();   // the previous class's exchange reads are done (E aliases R2)
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int y = yb + (128 / H) * h;
#pragma unroll
                for (int k1 = 0; k1 < 8; ++k1) R2[y * kRS + s * 9 + k1] = u[h][k1];
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int y = yb + (128 / H) * h;
#pragma unroll
                for (int p2 = 0; p2 < 8; ++p2) u[h][p2] = R2[y * kRS + p2 * 9 + s];
                RegDFT<8, -1, V>::run(u[h]);   // u[h][k2] = X_r[y][s + 8 k2]
#pragma unroll
                for (int k2 = 0; k2 < 8; ++k2) R1[y * kRS + s + 8 * k2] = u[h][k2];
            }
            __syncthreads();
            // ---- A': y DFT of the class columns, input rows [0, ny) ----
            V v[H][16];
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int p2 = pa + (16 / H) * h;
#pragma unroll
                for (int p1 = 0; p1 < 16; ++p1)
                    v[h][p1] = p1 < 8 ? R1[(16 * p1 + p2) * kRS + m] : make_float2(0.f, 0.f);
                RegDFT<16, -1, V>::run(v[h]);
#pragma unroll
                for (int k1 = 1; k1 < 16; ++k1) v[h][k1] = twiddle_mul<V, -1>(v[h][k1], twl, (p2 * k1) & (kLX - 1));
            }
            __syncthreads();   // R1 is read; the exchange overwrites it
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int k1 = 0; k1 < 16; ++k1) E[(k1 * 16 + pa + (16 / H) * h) * kM + m] = v[h][k1];
            __syncthreads();
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int k1 = pa + (16 / H) * h;
#pragma unroll
                for (int p2 = 0; p2 < 16; ++p2) v[h][p2] = E[(k1 * 16 + p2) * kM + m];
                RegDFT<16, -1, V>::run(v[h]);   // v[h][k2] = column output ky = k1 + 16 k2
                const uint32_t off = uint32_t((k1 * kLX + r * kM + m) * int(sizeof(V)));
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2)
                    store_at<V, true>(gp, off + uint32_t(16 * k2 * kLX * int(sizeof(V))), v[h][k2]);
            }
        }
    }
}

void launch_toep_pad_fft(const ToepPadFftArgs& a, hipStream_t s) {
    HASTY_REQUIRE(a.nz >= 1 && a.nbatch >= 1, "fused pad pass: empty batch or plane count");
    HASTY_REQUIRE(a.grid_b >= int64_t(a.nz) * kLY * kLX, "fused pad pass: grid stride too small");
    HASTY_REQUIRE(a.in_b >= int64_t(a.nz) * kNY * kNX || a.nbatch == 1, "fused pad pass: input stride too small");
    const size_t bytes = size_t(kLdsElems) * sizeof(float2);
    static const bool lds_ok = [&] {
        HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(toep_pad_fft_kernel<kH12>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
        return true;
    }();
    (void)lds_ok;
    const int groups = int(std::max<int64_t>(1, std::min<int64_t>((256 + a.nz - 1) / a.nz, a.nbatch)));
    const float2* tw = reinterpret_cast<const float2*>(twiddles(kLX, false));
    HASTY_PROF("toep_F12_xy_pad_fft", s);
    hipLaunchKernelGGL(toep_pad_fft_kernel<kH12>, dim3(unsigned(int64_t(a.nz) * groups)), dim3(1024 / kH12), bytes, s,
                       a, tw, groups);
    HASTY_CHECK_HIP(hipGetLastError());
}

bool toep_crop_sum_supported(int64_t ny, int64_t nx) { return ny == kNY && nx == kNX; }

void launch_toep_crop_sum(const ToepCropSumArgs& a, hipStream_t s) {
    HASTY_REQUIRE(a.nz >= 1 && a.nbatch >= 1, "fused crop pass: empty batch or plane count");
    HASTY_REQUIRE(a.groups >= 1 && a.groups <= a.nbatch, "fused crop pass: invalid coil group count");
    HASTY_REQUIRE(a.grid_b >= int64_t(a.nz) * kLY * kLX, "fused crop pass: grid stride too small");
    const int64_t nimg = int64_t(a.nz) * kNY * kNX;
    if (a.groups > 1)
        HASTY_REQUIRE(a.partial && a.partial_elems >= int64_t(a.groups) * nimg,
                      "fused crop pass: partial-sum scratch too small");
    const size_t bytes = size_t(kLdsElems) * sizeof(float2);
    static const bool lds_ok = [&] {
        HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(toep_crop_sum_kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
        return true;
    }();
    (void)lds_ok;
    const float2* tw = reinterpret_cast<const float2*>(twiddles(kLX, false));
    {
        HASTY_PROF("toep_F45_yx_ifft_crop_sum", s);
        hipLaunchKernelGGL(toep_crop_sum_kernel, dim3(unsigned(int64_t(a.nz) * a.groups)), dim3(kThreads), bytes,
                           s, a, tw);
        HASTY_CHECK_HIP(hipGetLastError());
    }
    if (a.groups > 1) {
        HASTY_PROF("toep_F45 sum_partials", s);
        const int64_t n = nimg;
        hipLaunchKernelGGL(sum_partials_kernel<float2>, dim3(unsigned(std::min<int64_t>(ceil_div(n, 256), 4096))),
                           dim3(256), 0, s, static_cast<const float2*>(a.partial), a.out, int64_t(a.nz) * kNY, kNX,
                           int64_t(kNX), a.groups, a.accumulate);
        HASTY_CHECK_HIP(hipGetLastError());
    }
}

}  // namespace hasty
