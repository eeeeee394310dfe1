// Fused Toeplitz passes of the SENSE normal operator on gfx950, for planes of
// N x N with N = 64, 96 or 128 (2N = 128, 192, 256 per axis; written out
// below for N = 128):
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
// Plane shape N x N.  Column DFTs of length 2N run as A1 x A2 with A2 = 16
// step-1 lanes of A1 = N/8 values (A1 step-2 lanes of 16 values); the class
// rows (M = N/2) run as B1 x B2 with B2 = 8 step-1 lanes of B1 = N/16 values
// (B1 step-2 lanes of 8 values).  One line per lane needs 8N threads.
template <int N> struct FShape {
    static constexpr int L = 2 * N;            // 2n per axis
    static constexpr int M = N / 2;            // columns per class (L / 4)
    static constexpr int A2 = 16, A1 = L / 16;
    static constexpr int B2 = 8, B1 = M / 8;
    static constexpr int RS = M + 8;           // LDS row stride of the class rows
    static constexpr int XS = B1 + 1;          // row-exchange stride within a row
    static constexpr int E = A1 * A2 * M;      // column exchange, [k1][p2][m]
    static constexpr int R = N * RS;           // one set of class rows
    static constexpr int LDS = (E > 2 * R ? E : 2 * R) + L;
    static_assert(A1 * A2 == L && B1 * B2 == M && A1 <= 16 && B1 <= 8, "plane factorisation");
    static_assert(7 * XS + B1 - 1 < RS, "row exchange fits in a class row");
};
// lines per thread of the last pass (N = 128: two, to fit 512 threads in 256
// VGPRs; the others one) and of the first pass (one)
#ifndef HASTY_TOEP_FUSED_H
#define HASTY_TOEP_FUSED_H 2
#endif
template <int N> constexpr int kH45 = N == 128 ? HASTY_TOEP_FUSED_H : 1;
#ifndef HASTY_TOEP_F12_H
#define HASTY_TOEP_F12_H 1
#endif
constexpr int kH12 = HASTY_TOEP_F12_H;
}  // namespace

// x mod L for x >= 0 (unsigned: a mask when L is a power of two)
template <int L>
__device__ inline int mod_l(int x) {
    return int(unsigned(x) % unsigned(L));
}

// Orders a wave's LDS writes before its own later LDS reads of them: the
// row exchanges, where each row belongs to the 8 lanes (tid >> 3) of one
// wave, need no block barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// output pixel of slot i of row lane s (s < B1): slot k2 < 8 -> x = n' + 1,
// slot 8 + k2 -> x = n' + M + 1 (n' < M - 1) or x = 0 (n' = M - 1), with
// n' = s + B1 k2
template <int N>
__device__ inline int crop_px(int s, int i) {
    using S = FShape<N>;
    const int np = s + S::B1 * (i & 7);
    return i < 8 ? np + 1 : (np == S::M - 1 ? 0 : np + S::M + 1);
}

// v (.) d at element (b, row, x) for the 16 pixels of lane s
template <int N>
__device__ inline void crop_sum_diag(float2* v, const DiagOp<float2>& d, int64_t b, int64_t row, int s) {
    if (!d.ptr) return;
    const float2* base = d.ptr + b * d.sb + row * d.sr;
    const float sgn = d.type == HASTY_MULT_CONJ ? -1.f : 1.f;
    // two rounds of 8 loads (the next columns are in flight in registers)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float2 w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = base[crop_px<N>(s, 8 * h + i)];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[8 * h + i] = cmul_sgn(v[8 * h + i], w[i], sgn);
    }
}

template <int N, int H>
__global__ void __launch_bounds__(8 * N / H) __attribute__((amdgpu_waves_per_eu(1, 8 * N / H / 256)))
toep_crop_sum_kernel(ToepCropSumArgs a, const float2* tw) {
    using V = float2;
    using S = FShape<N>;
    constexpr int T = 8 * N / H, L = S::L, M = S::M, A1 = S::A1, B1 = S::B1, RS = S::RS, XS = S::XS;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    V* E = lds;              // phase A exchange
    V* R1 = lds;             // class rows (after the exchange is read)
    V* R2 = lds + S::R;      // phase B step-1 results
    V* twl = lds + (S::LDS - L);
    const int tid = threadIdx.x;
    {   // this block's frame
        const int64_t fr = blockIdx.y;
        a.grid += fr * a.grid_f;
        a.out += fr * a.out_f;
        a.partial += fr * a.partial_f;
        if (a.m1.ptr) a.m1.ptr += fr * a.m1.sf;
        if (a.m2.ptr) a.m2.ptr += fr * a.m2.sf;
        if (a.io.ptr) a.io.ptr += fr * a.io.sf;
    }
    for (int k = tid; k < L; k += T) twl[k] = tw[k];

    const int z = int(blockIdx.x % unsigned(a.nz));
    const int g = int(blockIdx.x / unsigned(a.nz));
    const int cpb = (a.nbatch + a.groups - 1) / a.groups;
    const int b0 = g * cpb;
    const int b1 = min(a.nbatch, b0 + cpb);

    // phase A roles: column m; step-1 lines p2 = pa + (16 / H) h, step-2 lines
    // k1 = pa + (16 / H) h (those below A1)
    const int m0 = int(unsigned(tid) % unsigned(M));
    const int pa0 = int(unsigned(tid) / unsigned(M));
    // phase B roles: rows y = yb + (N / H) h, lane s of the B1 x 8 DFT
    const int yb0 = tid >> 3;
    const int s0 = tid & 7;
    V tot[H][16];
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int i = 0; i < 16; ++i) tot[h][i] = make_float2(0.f, 0.f);
    __syncthreads();

    // (coil, class) iterations it = (b - b0) 4 + r; the first half of the loads
    // of iteration it + 1 is issued once iteration it has moved its columns to
    // LDS, so it is in flight during the row transforms and the coil epilogue
    // (the second half waits for the next iteration: registers)
    const int nit = 4 * max(0, b1 - b0);
    V v[H][16];
    auto load_class = [&](int it, int h0, int h1) {
        int m = m0, pa = pa0;
        asm volatile("" : "+v"(m), "+v"(pa));
        const int bb = b0 + (it >> 2), rr = it & 3;
        const char* plane =
            reinterpret_cast<const char*>(a.grid + int64_t(bb) * a.grid_b + int64_t(z) * (L * L));
        const uint32_t off = uint32_t((pa * L + rr * M + m) * int(sizeof(V)));
#pragma unroll
        for (int h = 0; h < H; ++h)
            if (h >= h0 && h < h1)
#pragma unroll
                for (int p1 = 0; p1 < A1; ++p1)
                    v[h][p1] = load_at<V, true>(plane, off + uint32_t((16 * p1 + (16 / H) * h) * L * int(sizeof(V))));
    };
    constexpr int kPre = H > 1 ? H / 2 : H;   // lines per thread prefetched a phase early
    if (nit > 0) load_class(0, 0, kPre);
    V acc[H][16];
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
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[h][i] = make_float2(0.f, 0.f);
        }
        if (kPre < H) load_class(it, kPre, H);
        // ---- A: y inverse DFT of the class columns ----
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int p2 = pa + (16 / H) * h;
            RegDFT<A1, +1, V>::run(v[h]);
#pragma unroll
            for (int k1 = 1; k1 < A1; ++k1) v[h][k1] = twiddle_mul<V, +1>(v[h][k1], twl, mod_l<L>(p2 * k1));
        }
        __syncthreads();   // the previous iteration's phase B reads of R1 / R2 are done
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int k1 = 0; k1 < A1; ++k1) E[(k1 * 16 + pa + (16 / H) * h) * M + m] = v[h][k1];
        __syncthreads();
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int k1 = pa + (16 / H) * h;
            if (A1 == 16 || k1 < A1) {   // every line when A1 == 16
#pragma unroll
                for (int p2 = 0; p2 < 16; ++p2) v[h][p2] = E[(k1 * 16 + p2) * M + m];
                RegDFT<16, +1, V>::run(v[h]);   // v[h][k2] = X[k1 + A1 k2]
            }
        }
        __syncthreads();   // E is read; R1 overwrites it
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int k1 = pa + (16 / H) * h;
            if (A1 == 16 || k1 < A1) {   // every line when A1 == 16
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    const int j = k1 + A1 * k2;
                    if (j >= N - 1 && j < 2 * N - 1) R1[(j - (N - 1)) * RS + m] = v[h][k2];
                }
            }
        }
        if (it + 1 < nit) load_class(it + 1, 0, kPre);
        __syncthreads();
        // ---- B: M-point inverse DFT of each class row, B1 x 8 ----
        V w[H][8];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int y = yb + (N / H) * h;
#pragma unroll
            for (int p1 = 0; p1 < B1; ++p1) w[h][p1] = R1[y * RS + 8 * p1 + s];
            RegDFT<B1, +1, V>::run(w[h]);
#pragma unroll
            for (int k1 = 1; k1 < B1; ++k1) w[h][k1] = twiddle_mul<V, +1>(w[h][k1], twl, mod_l<L>(4 * s * k1));
#pragma unroll
            for (int k1 = 0; k1 < B1; ++k1) R2[y * RS + s * XS + k1] = w[h][k1];
        }
        wave_lds_sync();   // (the row's 8 lanes are one wave's: 0.660 -> 0.648 ms)
        if (B1 == 8 || s < B1) {   // all 8 lanes when B1 == 8
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int y = yb + (N / H) * h;
#pragma unroll
                for (int p2 = 0; p2 < 8; ++p2) w[h][p2] = R2[y * RS + p2 * XS + s];
                RegDFT<8, +1, V>::run(w[h]);   // w[h][k2] = inner_r(s + B1 k2)
                // combine: acc += e^{+2 pi i r n / L} inner_r(n mod M)
#pragma unroll
                for (int k2 = 0; k2 < 8; ++k2) {
                    const int np = s + B1 * k2;
                    const int na = np + N, nb = np == M - 1 ? N - 1 : np + 3 * M;
                    acc[h][k2] = cadd(acc[h][k2], twiddle_mul<V, +1>(w[h][k2], twl, mod_l<L>(r * na)));
                    acc[h][8 + k2] = cadd(acc[h][8 + k2], twiddle_mul<V, +1>(w[h][k2], twl, mod_l<L>(r * nb)));
                }
            }
            if (r == 3) {
                // ---- coil epilogue: scale, output diagonals, coil sum ----
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const int64_t row = int64_t(z) * N + yb + (N / H) * h;
#pragma unroll
                    for (int i = 0; i < 16; ++i) acc[h][i] = cscale(acc[h][i], a.scale);
                    crop_sum_diag<N>(acc[h], a.m1, b, row, s);
                    crop_sum_diag<N>(acc[h], a.m2, b, row, s);
                    crop_sum_diag<N>(acc[h], a.io, b, row, s);
#pragma unroll
                    for (int i = 0; i < 16; ++i) tot[h][i] = cadd(tot[h][i], acc[h][i]);
                }
            }
        }
    }
    const int yb = yb0, s = s0;
    if (B1 == 8 || s < B1) {   // all 8 lanes when B1 == 8
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int64_t row = int64_t(z) * N + yb + (N / H) * h;
            if (a.groups > 1) {
                V* part = a.partial + (int64_t(g) * a.nz * N + row) * N;
#pragma unroll
                for (int i = 0; i < 16; ++i) part[crop_px<N>(s, i)] = tot[h][i];
            } else {
                V* o = a.out + row * N;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int x = crop_px<N>(s, i);
                    o[x] = a.accumulate ? cadd(tot[h][i], o[x]) : tot[h][i];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// First two passes fused: x and y zero-padded forward DFTs, x class-major.
// Per plane z and coil b the thread loads its rows' inputs once (x positions
// n' and n' + M, times the diagonals); per class r:
//   B' each row's class: u[n'] = e^{-2 pi i r n' / L} (x[n'] + (-i)^r
//      x[n' + M]), an M-point DFT (B1 x 8, one exchange) -> X[4m + r], rows to LDS;
//   A' each class column: N inputs zero-padded to 2N (A1 x 16 with half the
//      first-step inputs zero), stored at (ky, r M + m): M values per row.
// The next coil's inputs are loaded once the last class has folded them.
// ---------------------------------------------------------------------------
template <int N, int H>
__device__ inline void pad_fft_diag(float2 (*xs)[2 * FShape<N>::B1], const DiagOp<float2>& d, int64_t b, int z,
                                    int yb, int s) {
    using S = FShape<N>;
    if (!d.ptr) return;
    const float sgn = d.type == HASTY_MULT_CONJ ? -1.f : 1.f;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int64_t row = int64_t(z) * N + yb + (N / H) * h;
        const float2* base = d.ptr + b * d.sb + row * d.sr + s;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            float2 w[S::B1];
#pragma unroll
            for (int i = 0; i < S::B1; ++i) w[i] = base[8 * i + S::M * q];
#pragma unroll
            for (int i = 0; i < S::B1; ++i) xs[h][S::B1 * q + i] = cmul_sgn(xs[h][S::B1 * q + i], w[i], sgn);
        }
    }
}

template <int N, int H>
__global__ void __launch_bounds__(8 * N / H) __attribute__((amdgpu_waves_per_eu(1, 8 * N / H / 256)))
toep_pad_fft_kernel(ToepPadFftArgs a, const float2* tw, int groups) {
    using V = float2;
    using S = FShape<N>;
    constexpr int T = 8 * N / H, L = S::L, M = S::M, A1 = S::A1, B1 = S::B1, RS = S::RS, XS = S::XS;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    V* E = lds;              // column exchange
    V* R1 = lds;             // class rows (read before the exchange is written)
    V* R2 = lds + S::R;      // row exchange
    V* twl = lds + (S::LDS - L);
    const int tid = threadIdx.x;
    {   // this block's frame
        const int64_t fr = blockIdx.y;
        a.in += fr * a.in_f;
        a.grid += fr * a.grid_f;
        if (a.m1.ptr) a.m1.ptr += fr * a.m1.sf;
        if (a.m2.ptr) a.m2.ptr += fr * a.m2.sf;
    }
    for (int k = tid; k < L; k += T) twl[k] = tw[k];

    const int z = int(blockIdx.x % unsigned(a.nz));
    const int g = int(blockIdx.x / unsigned(a.nz));
    const int cpb = (a.nbatch + groups - 1) / groups;
    const int b0 = g * cpb;
    const int b1 = min(a.nbatch, b0 + cpb);
    const int m0 = int(unsigned(tid) % unsigned(M)), pa0 = int(unsigned(tid) / unsigned(M)), yb0 = tid >> 3,
              s0 = tid & 7;

    // xs[h][i]: row yb + (N / H) h, x = 8 i + s (i < B1) or 8 (i - B1) + M + s
    V xs[H][2 * B1];
    auto load_in = [&](int b) {
        int yb = yb0, s = s0;
        asm volatile("" : "+v"(yb), "+v"(s));
        const char* pl = reinterpret_cast<const char*>(a.in + int64_t(b) * a.in_b + int64_t(z) * (N * N));
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const uint32_t off = uint32_t(((yb + (N / H) * h) * N + s) * int(sizeof(V)));
#pragma unroll
            for (int i = 0; i < 2 * B1; ++i)
                xs[h][i] = load_at<V>(pl, off + uint32_t((8 * (i % B1) + M * (i / B1)) * int(sizeof(V))));
        }
    };
    if (b0 < b1) load_in(b0);
    // a broadcast first diagonal (the image of a SENSE apply) is read once
    // per block, not once per coil
    const bool m1_bcast = a.m1.ptr && a.m1.sb == 0;
    V d1[H][2 * B1];
    if (m1_bcast) {
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int i = 0; i < 2 * B1; ++i) d1[h][i] = make_float2(1.f, 0.f);
        pad_fft_diag<N, H>(d1, a.m1, 0, z, yb0, s0);
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
                    for (int i = 0; i < 2 * B1; ++i) xs[h][i] = cmul(xs[h][i], d1[h][i]);
            } else {
                pad_fft_diag<N, H>(xs, a.m1, b, z, yb, s);
            }
            pad_fft_diag<N, H>(xs, a.m2, b, z, yb, s);
        }
        char* gp = reinterpret_cast<char*>(a.grid + int64_t(b) * a.grid_b + int64_t(z) * (L * L));
#pragma unroll 1
        for (int r = 0; r < 4; ++r) {
            int m = m0, pa = pa0, yb = yb0, s = s0;
            asm volatile("" : "+v"(m), "+v"(pa), "+v"(yb), "+v"(s));
            // ---- B': fold, M-point DFT of each row's class ----
            V u[H][8];
#pragma unroll
            for (int h = 0; h < H; ++h) {
#pragma unroll
                for (int p1 = 0; p1 < B1; ++p1) {
                    const int np = 8 * p1 + s;
                    V t = cadd(xs[h][p1], quarter_turn(xs[h][B1 + p1], 4 - r));
                    u[h][p1] = r ? twiddle_mul<V, -1>(t, twl, mod_l<L>(r * np)) : t;
                }
            }
            if (r == 3 && b + 1 < b1) load_in(b + 1);   // this coil's inputs are folded
#pragma unroll
            for (int h = 0; h < H; ++h) {
                RegDFT<B1, -1, V>::run(u[h]);
#pragma unroll
                for (int k1 = 1; k1 < B1; ++k1) u[h][k1] = twiddle_mul<V, -1>(u[h][k1], twl, mod_l<L>(4 * s * k1));
            }
            __syncthreads();   // the previous class's exchange reads are done (E aliases R2)
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int y = yb + (N / H) * h;
#pragma unroll
                for (int k1 = 0; k1 < B1; ++k1) R2[y * RS + s * XS + k1] = u[h][k1];
            }
            // (a block barrier, though the row's 8 lanes are one wave's: the
            // wave-level sync measured 0.616 against 0.613 ms here)
            __syncthreads();
            if (B1 == 8 || s < B1) {   // all 8 lanes when B1 == 8
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const int y = yb + (N / H) * h;
#pragma unroll
                    for (int p2 = 0; p2 < 8; ++p2) u[h][p2] = R2[y * RS + p2 * XS + s];
                    RegDFT<8, -1, V>::run(u[h]);   // u[h][k2] = X_r[y][s + B1 k2]
#pragma unroll
                    for (int k2 = 0; k2 < 8; ++k2) R1[y * RS + s + B1 * k2] = u[h][k2];
                }
            }
            __syncthreads();
            // ---- A': y DFT of the class columns, input rows [0, N) ----
            V v[H][16];
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int p2 = pa + (16 / H) * h;
#pragma unroll
                for (int p1 = 0; p1 < A1; ++p1)
                    v[h][p1] = 2 * p1 < A1 ? R1[(16 * p1 + p2) * RS + m] : make_float2(0.f, 0.f);
                RegDFT<A1, -1, V>::run(v[h]);
#pragma unroll
                for (int k1 = 1; k1 < A1; ++k1) v[h][k1] = twiddle_mul<V, -1>(v[h][k1], twl, mod_l<L>(p2 * k1));
            }
            __syncthreads();   // R1 is read; the exchange overwrites it
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int k1 = 0; k1 < A1; ++k1) E[(k1 * 16 + pa + (16 / H) * h) * M + m] = v[h][k1];
            __syncthreads();
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int k1 = pa + (16 / H) * h;
                if (A1 == 16 || k1 < A1) {   // every line when A1 == 16
#pragma unroll
                    for (int p2 = 0; p2 < 16; ++p2) v[h][p2] = E[(k1 * 16 + p2) * M + m];
                    RegDFT<16, -1, V>::run(v[h]);   // v[h][k2] = column output ky = k1 + A1 k2
                    const uint32_t off = uint32_t((k1 * L + r * M + m) * int(sizeof(V)));
#pragma unroll
                    for (int k2 = 0; k2 < 16; ++k2)
                        store_at<V, true>(gp, off + uint32_t(A1 * k2 * L * int(sizeof(V))), v[h][k2]);
                }
            }
        }
    }
}

namespace {
template <int N, class K> void allow_fused_lds(K kernel) {
    static const bool ok = [&] {
        HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            int(size_t(FShape<N>::LDS) * sizeof(float2))));
        return true;
    }();
    (void)ok;
}

template <int N> void pad_fft_launch(const ToepPadFftArgs& a, hipStream_t s) {
    using S = FShape<N>;
    HASTY_REQUIRE(a.grid_b >= int64_t(a.nz) * S::L * S::L, "fused pad pass: grid stride too small");
    HASTY_REQUIRE(a.in_b >= int64_t(a.nz) * N * N || a.nbatch == 1, "fused pad pass: input stride too small");
    auto kern = toep_pad_fft_kernel<N, kH12>;
    allow_fused_lds<N>(kern);
    const int groups = int(std::max<int64_t>(1, std::min<int64_t>((256 + a.nz - 1) / a.nz, a.nbatch)));
    const float2* tw = reinterpret_cast<const float2*>(twiddles(S::L, false));
    HASTY_PROF("toep_F12_xy_pad_fft", s);
    hipLaunchKernelGGL(kern, dim3(unsigned(int64_t(a.nz) * groups), unsigned(a.nframes)), dim3(8 * N / kH12),
                       size_t(S::LDS) * sizeof(float2), s, a, tw, groups);
    HASTY_CHECK_HIP(hipGetLastError());
}

template <int N> void crop_sum_launch(const ToepCropSumArgs& a, hipStream_t s) {
    using S = FShape<N>;
    HASTY_REQUIRE(a.grid_b >= int64_t(a.nz) * S::L * S::L, "fused crop pass: grid stride too small");
    const int64_t nimg = int64_t(a.nz) * N * N;
    if (a.groups > 1) {
        HASTY_REQUIRE(a.partial_f >= int64_t(a.groups) * nimg || a.nframes == 1,
                      "fused crop pass: partial-sum frame stride too small");
        HASTY_REQUIRE(a.partial && a.partial_elems >= int64_t(a.nframes - 1) * a.partial_f + int64_t(a.groups) * nimg,
                      "fused crop pass: partial-sum scratch too small");
    }
    auto kern = toep_crop_sum_kernel<N, kH45<N>>;
    allow_fused_lds<N>(kern);
    const float2* tw = reinterpret_cast<const float2*>(twiddles(S::L, false));
    {
        HASTY_PROF("toep_F45_yx_ifft_crop_sum", s);
        hipLaunchKernelGGL(kern, dim3(unsigned(int64_t(a.nz) * a.groups), unsigned(a.nframes)), dim3(8 * N / kH45<N>),
                           size_t(S::LDS) * sizeof(float2), s, a, tw);
        HASTY_CHECK_HIP(hipGetLastError());
    }
    if (a.groups > 1) {
        HASTY_PROF("toep_F45 sum_partials", s);
        // (frame f's partials sit at partial + f * groups * nimg)
        HASTY_REQUIRE(a.nframes == 1 || a.partial_f == int64_t(a.groups) * nimg,
                      "fused crop pass: frames' partial sums must be packed");
        hipLaunchKernelGGL(sum_partials_kernel<float2>,
                           dim3(unsigned(std::min<int64_t>(ceil_div(nimg, 256), 4096)), unsigned(a.nframes)),
                           dim3(256), 0, s, static_cast<const float2*>(a.partial), a.out, int64_t(a.nz) * N, N,
                           int64_t(N), a.groups, a.accumulate, a.out_f);
        HASTY_CHECK_HIP(hipGetLastError());
    }
}
}  // namespace

bool toep_crop_sum_supported(int64_t ny, int64_t nx) { return ny == nx && (nx == 64 || nx == 96 || nx == 128); }

void launch_toep_pad_fft(const ToepPadFftArgs& a, hipStream_t s) {
    HASTY_REQUIRE(a.nz >= 1 && a.nbatch >= 1, "fused pad pass: empty batch or plane count");
    HASTY_REQUIRE(a.nframes >= 1 && a.nframes <= 65535, "fused pad pass: invalid frame count");
    switch (a.n) {
        case 64: return pad_fft_launch<64>(a, s);
        case 96: return pad_fft_launch<96>(a, s);
        case 128: return pad_fft_launch<128>(a, s);
        default: throw Error(HASTY_ERR_INVALID, "fused pad pass: plane size must be 64, 96 or 128");
    }
}

void launch_toep_crop_sum(const ToepCropSumArgs& a, hipStream_t s) {
    HASTY_REQUIRE(a.nz >= 1 && a.nbatch >= 1, "fused crop pass: empty batch or plane count");
    HASTY_REQUIRE(a.nframes >= 1 && a.nframes <= 65535, "fused crop pass: invalid frame count");
    HASTY_REQUIRE(a.groups >= 1 && a.groups <= a.nbatch, "fused crop pass: invalid coil group count");
    switch (a.n) {
        case 64: return crop_sum_launch<64>(a, s);
        case 96: return crop_sum_launch<96>(a, s);
        case 128: return crop_sum_launch<128>(a, s);
        default: throw Error(HASTY_ERR_INVALID, "fused crop pass: plane size must be 64, 96 or 128");
    }
}

}  // namespace hasty
