// Fused x and y passes of the NUFFT type-2 (forward) fine-grid transform on
// gfx950 for 3-D c64 plans with n_y = n_x = 128 on a 256 x 256 fine grid
// (upsampling 2): the pruned x pass and y pass in one kernel, so the
// [B, n_z, n_y, 256] intermediate between them never goes through HBM.
//
// The image holds centred modes (index j <-> mode j - 64, fine-grid position
// j - 64 mod 256) times the deapodisation dz dy dx and the load diagonals.
// The y transform splits by residue classes of ky = 4m + r, so that x stays
// in natural order for the interpolation that reads the grid:
//     X[4m + r] = sum_{n'<64} e^{s 2 pi i r n'/256} (g[n' + 64] + c_r g[n'])
//                 e^{s 2 pi i m n'/64},   c_r = e^{-s pi i r / 2},
// (s the plan's sign; g[j] the column, j = n' + 64 the modes n' >= 0 and
// j = n' the modes n' - 64 < 0).  Per plane z, coil b and class r:
//   Y  each column x: fold, 64-point DFT (8 x 8, one LDS exchange) -> rows m;
//   X  each row m: half-input 256-point DFT along x (16 x 16, one
//      exchange), times e^{-s pi i kx / 2} for the -64 shift; the second
//      step runs 16 lanes per row (lane k1 holds kx = k1 + 16 k2), so each
//      store instruction writes 128-B runs of 4 fine-grid rows ky = 4m + r.
#include <algorithm>

#include "fft_reg.hpp"
#include "toeplitz_fused.hpp"

namespace hasty {

namespace {
constexpr int kN = 128, kL = 256, kM = 64;
constexpr int kT = 1024;
constexpr int kR2 = 72 * kN;           // y exchange [s 9 + k1][x]
constexpr int kRS1 = kN + 1;           // class rows [m][x], odd stride
constexpr int kR1 = kM * kRS1;
constexpr int kES = 17;                // x exchange [p2][m][k1], k1 padded to 17
constexpr int kE = 16 * kM * kES;
constexpr int kLdsElems = std::max(kR2 + kR1, kE) + kL;
}  // namespace

template <int SIGN>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(1, 4)))
nufft_fwd_xy_kernel(NufftFwdXyArgs a, const float2* tw) {
    using V = float2;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    V* R2 = lds;               // y exchange
    V* R1 = lds + kR2;         // class rows
    V* E = lds;                // x exchange (R1 / R2 are read by then)
    V* twl = lds + (kLdsElems - kL);
    const int tid = threadIdx.x;
    for (int k = tid; k < kL; k += kT) twl[k] = tw[k];

    const int z = int(blockIdx.x % unsigned(a.nz));
    const int g = int(blockIdx.x / unsigned(a.nz));
    const int cpb = (a.nbatch + a.groups - 1) / a.groups;
    const int b0 = g * cpb;
    const int b1 = min(a.nbatch, b0 + cpb);
    // phase Y roles: column x, lane s of the 8 x 8 DFT (s wave-uniform)
    const int x0 = tid & (kN - 1), s0 = tid >> 7;
    // phase X roles: row m, line p2 / k1 of the 16 x 16 DFT (wave-uniform)
    const int mm0 = tid & (kM - 1), pa0 = tid >> 6;

    // g[h][p1]: column x0, rows j = 8 p1 + s (h = 0) and 8 p1 + s + 64 (h = 1)
    V gx[2][8];
    auto load_in = [&](int b) {
        int x = x0, s = s0;
        asm volatile("" : "+v"(x), "+v"(s));
        const char* pl = reinterpret_cast<const char*>(a.in + int64_t(b) * a.in_b + int64_t(z) * (kN * kN));
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p1 = 0; p1 < 8; ++p1)
                gx[h][p1] = load_at<V>(pl, uint32_t(((8 * p1 + s + 64 * h) * kN + x) * int(sizeof(V))));
    };
    auto apply = [&](const DiagOp<V>& d, int b) {
        if (!d.ptr) return;
        int x = x0, s = s0;
        asm volatile("" : "+v"(x), "+v"(s));
        const float sgn = d.type == HASTY_MULT_CONJ ? -1.f : 1.f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            V w[8];
#pragma unroll
            for (int p1 = 0; p1 < 8; ++p1)
                w[p1] = d.ptr[int64_t(b) * d.sb + (int64_t(z) * kN + 8 * p1 + s + 64 * h) * d.sr + x];
#pragma unroll
            for (int p1 = 0; p1 < 8; ++p1) gx[h][p1] = cmul_sgn(gx[h][p1], w[p1], sgn);
        }
    };
    if (b0 < b1) load_in(b0);
    __syncthreads();

#pragma unroll 1
    for (int b = b0; b < b1; ++b) {
        apply(a.m1, b);
        apply(a.m2, b);
        if (a.dx) {   // deapodisation dz[z] dy[j] dx[x]
            int x = x0, s = s0;
            asm volatile("" : "+v"(x), "+v"(s));
            const float zx = a.dz[z] * a.dx[x];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int p1 = 0; p1 < 8; ++p1) gx[h][p1] = cscale(gx[h][p1], zx * a.dy[8 * p1 + s + 64 * h]);
        }
        char* gp = reinterpret_cast<char*>(a.grid + int64_t(b) * a.grid_b + int64_t(z) * (kL * kL));
#pragma unroll 1
        for (int r = 0; r < 4; ++r) {
            int x = x0, s = s0, mm = mm0, pa = pa0;
            asm volatile("" : "+v"(x), "+v"(s), "+v"(mm), "+v"(pa));
            // ---- Y: fold the centred column by class r, 64-point DFT ----
            V u[8];
#pragma unroll
            for (int p1 = 0; p1 < 8; ++p1) {
                const int np = 8 * p1 + s;
                // c_r = e^{-s pi i r / 2} = i^{-s r}
                V t = cadd(gx[1][p1], quarter_turn(gx[0][p1], SIGN < 0 ? r : 4 - r));
                u[p1] = r ? twiddle_mul<V, SIGN>(t, twl, (r * np) & (kL - 1)) : t;
            }
            if (r == 3 && b + 1 < b1) load_in(b + 1);   // this coil's inputs are folded
            RegDFT<8, SIGN, V>::run(u);
#pragma unroll
            for (int k1 = 1; k1 < 8; ++k1) u[k1] = twiddle_mul<V, SIGN>(u[k1], twl, (4 * s * k1) & (kL - 1));
            __syncthreads();   // the previous class's exchange reads are done
#pragma unroll
            for (int k1 = 0; k1 < 8; ++k1) R2[(s * 9 + k1) * kN + x] = u[k1];
            __syncthreads();
#pragma unroll
            for (int p2 = 0; p2 < 8; ++p2) u[p2] = R2[(p2 * 9 + s) * kN + x];
            RegDFT<8, SIGN, V>::run(u);   // u[k2] = X_r[m = s + 8 k2][x]
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) R1[(s + 8 * k2) * kRS1 + x] = u[k2];
            __syncthreads();
            // ---- X: half-input 256-point DFT of each class row ----
            V v[16];
#pragma unroll
            for (int p1 = 0; p1 < 16; ++p1) v[p1] = p1 < 8 ? R1[mm * kRS1 + 16 * p1 + pa] : make_float2(0.f, 0.f);
            RegDFT<16, SIGN, V>::run(v);
#pragma unroll
            for (int k1 = 1; k1 < 16; ++k1) v[k1] = twiddle_mul<V, SIGN>(v[k1], twl, (pa * k1) & (kL - 1));
            __syncthreads();   // R1 is read; E overwrites it
            // (lanes over rows mm: stride 17 keeps the 8-B writes conflict-free)
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) E[(pa * kM + mm) * kES + k1] = v[k1];
            __syncthreads();
            {
                // second step: lane k1 of row mr gathers p2 = 0..15
                int k1 = tid & 15, mr = tid >> 4;
                asm volatile("" : "+v"(k1), "+v"(mr));
#pragma unroll
                for (int p2 = 0; p2 < 16; ++p2) v[p2] = E[(p2 * kM + mr) * kES + k1];
                RegDFT<16, SIGN, V>::run(v);   // v[k2] = kx = k1 + 16 k2, before the shift
                // the -64 shift: output kx times e^{-s pi i kx / 2} = i^{-s (kx mod 4)}, kx mod 4 = k1 mod 4
                // (lane-varying: a multiply by i, then a negation, as selects)
                const int q = SIGN < 0 ? (k1 & 3) : ((4 - (k1 & 3)) & 3);
                const uint32_t row = uint32_t(((4 * mr + r) * kL + k1) * int(sizeof(V)));
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    float re = v[k2].x, im = v[k2].y;
                    const float t = re;
                    re = (q & 1) ? -im : re;
                    im = (q & 1) ? t : im;
                    re = (q & 2) ? -re : re;
                    im = (q & 2) ? -im : im;
                    store_at<V, true>(gp, row + uint32_t(16 * k2 * int(sizeof(V))), make_float2(re, im));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused y and x passes of the NUFFT type-1 (adjoint) fine-grid transform for
// the same plans: grid [B, nz, 256, 256] (after the z pass) -> img
// [B, nz, 128, 128], the output index j reading line position j - 64 mod 256
// along y and x.  The y transform splits by residue classes of ky = 4m + r:
//     g[q] = sum_r e^{s 2 pi i r q / 256} inner_r(q mod 64),
//     inner_r(n') = sum_{m<64} X[4m + r] e^{s 2 pi i m n' / 64},
// and the kept y <-> q = y - 64 mod 256, so inner_r(n') feeds y = n'
// (q = n' + 192) and y = n' + 64 (q = n').  Per plane z, coil b and class r:
//   X  each fine-grid row ky = 4m + r, read as 128-B runs (16 lanes per row,
//      lane kl taking kx = 16 p1 + kl): 256-point DFT along x (16 x 16, one
//      exchange) of X[kx] i^{-s kx} (the -64 shift), keeping x < 128; rows
//      m to LDS;
//   Y  each column x: 64-point DFT over m (8 x 8, one exchange), then the
//      twiddled class sum into the thread's 16 outputs.
// After a coil's four classes the thread applies the row pass's epilogue
// (scale, store diagonals, deapodisation dz dy dx, accumulate) and stores
// its 16 pixels of column x (lanes over x: 512-B runs).
// ---------------------------------------------------------------------------
namespace {
constexpr int kIES = 17;                  // x exchange [m][kl][x_lo], x_lo padded to 17
constexpr int kIE = kM * 16 * kIES;
constexpr int kIRS = kN + 16;             // class rows [m][x]: stride 144 keeps 4-row reads conflict-free
constexpr int kIR1 = kM * kIRS;
constexpr int kIR2 = 72 * kN;             // y exchange [s 9 + k1][x]
constexpr int kInvLds = std::max(kIE, kIR1 + kIR2) + kL;

// v * i^q, q in 0..3 lane-varying (a multiply by i, then a negation, as selects)
__device__ inline float2 quarter_turn_sel(float2 v, int q) {
    float re = v.x, im = v.y;
    const float t = re;
    re = (q & 1) ? -im : re;
    im = (q & 1) ? t : im;
    re = (q & 2) ? -re : re;
    im = (q & 2) ? -im : im;
    return make_float2(re, im);
}
}  // namespace

template <int SIGN>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(1, 4)))
nufft_inv_xy_kernel(NufftInvXyArgs a, const float2* tw) {
    using V = float2;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    V* E = lds;               // x exchange
    V* R1 = lds;              // class rows (E is read by then)
    V* R2 = lds + kIR1;       // y exchange
    V* twl = lds + (kInvLds - kL);
    const int tid = threadIdx.x;
    for (int k = tid; k < kL; k += kT) twl[k] = tw[k];

    const int z = int(blockIdx.x % unsigned(a.nz));
    const int g = int(blockIdx.x / unsigned(a.nz));
    const int cpb = (a.nbatch + a.groups - 1) / a.groups;
    const int b0 = g * cpb;
    const int b1 = min(a.nbatch, b0 + cpb);
    // phase X roles: row m, lane kl (x_lo in the second step)
    const int kl0 = tid & 15, m0 = tid >> 4;
    // phase Y roles: column x, lane s of the 8 x 8 DFT (s wave-uniform)
    const int x0 = tid & (kN - 1), s0 = tid >> 7;
    const int nit = 4 * max(0, b1 - b0);
    V v[16];
    // fine-grid row ky = 4 m + r of iteration it's coil: kx = 16 p1 + kl
    auto load_rows = [&](int it) {
        int kl = kl0, m = m0;
        asm volatile("" : "+v"(kl), "+v"(m));
        const int bb = b0 + (it >> 2), rr = it & 3;
        const char* pl = reinterpret_cast<const char*>(a.in + int64_t(bb) * a.in_b + int64_t(z) * (kL * kL));
        const uint32_t off = uint32_t(((4 * m + rr) * kL + kl) * int(sizeof(V)));
#pragma unroll
        for (int p1 = 0; p1 < 16; ++p1) v[p1] = load_at<V, true>(pl, off + uint32_t(16 * p1 * int(sizeof(V))));
    };
    __syncthreads();
    if (nit > 0) load_rows(0);
    V acc[16];
#pragma unroll 1
    for (int it = 0; it < nit; ++it) {
        const int b = b0 + (it >> 2), r = it & 3;
        // per-iteration copies of the lane indices (otherwise the LDS and
        // twiddle addresses derived from them are hoisted and stay live)
        int kl = kl0, m = m0, x = x0, s = s0;
        asm volatile("" : "+v"(kl), "+v"(m), "+v"(x), "+v"(s));
        if (r == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = make_float2(0.f, 0.f);
        }
        // ---- X step 1: the -64 shift (i^{-s kx}, kx mod 4 = kl mod 4), DFT over p1 ----
        const int qi = SIGN > 0 ? ((4 - (kl & 3)) & 3) : (kl & 3);
#pragma unroll
        for (int p1 = 0; p1 < 16; ++p1) v[p1] = quarter_turn_sel(v[p1], qi);
        RegDFT<16, SIGN, V>::run(v);   // v[x_lo]
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) v[k1] = twiddle_mul<V, SIGN>(v[k1], twl, (kl * k1) & (kL - 1));
        __syncthreads();   // the previous class's Y reads are done (E overlaps R1 and R2)
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) E[(m * 16 + kl) * kIES + k1] = v[k1];
        __syncthreads();
        // ---- X step 2: lane x_lo = kl of row m gathers kl' = 0..15 ----
#pragma unroll
        for (int p2 = 0; p2 < 16; ++p2) v[p2] = E[(m * 16 + p2) * kIES + kl];
        RegDFT<16, SIGN, V>::run(v);   // v[x_hi]: x = kl + 16 x_hi, kept for x_hi < 8
        __syncthreads();   // E is read; R1 overwrites it
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) R1[m * kIRS + kl + 16 * k2] = v[k2];
        if (it + 1 < nit) load_rows(it + 1);   // in flight during Y
        __syncthreads();
        // ---- Y: 64-point DFT over m of column x (8 x 8), class sum ----
        V u[8];
#pragma unroll
        for (int p1 = 0; p1 < 8; ++p1) u[p1] = R1[(8 * p1 + s) * kIRS + x];
        RegDFT<8, SIGN, V>::run(u);   // u[n_lo]
#pragma unroll
        for (int k1 = 1; k1 < 8; ++k1) u[k1] = twiddle_mul<V, SIGN>(u[k1], twl, (4 * s * k1) & (kL - 1));
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) R2[(s * 9 + k1) * kN + x] = u[k1];
        __syncthreads();
#pragma unroll
        for (int p2 = 0; p2 < 8; ++p2) u[p2] = R2[(p2 * 9 + s) * kN + x];
        RegDFT<8, SIGN, V>::run(u);   // u[n_hi] = inner_r(n' = s + 8 n_hi)
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
            const int np = s + 8 * k2;
            acc[k2] = cadd(acc[k2], r ? twiddle_mul<V, SIGN>(u[k2], twl, (r * (np + 192)) & (kL - 1)) : u[k2]);
            acc[8 + k2] = cadd(acc[8 + k2], r ? twiddle_mul<V, SIGN>(u[k2], twl, (r * np) & (kL - 1)) : u[k2]);
        }
        if (r == 3) {
            // ---- epilogue: pixels (y, x), y = s + 8 k2 and s + 8 k2 + 64 ----
            char* ob = reinterpret_cast<char*>(a.out + int64_t(b) * a.out_b);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int y = s + 8 * (i & 7) + 64 * (i >> 3);
                const int64_t row = int64_t(z) * kN + y;
                V val = cscale(acc[i], a.scale);
                const DiagOp<V>* ds[3] = {&a.m1, &a.m2, &a.io};
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (ds[d]->ptr) {
                        const float sgn = ds[d]->type == HASTY_MULT_CONJ ? -1.f : 1.f;
                        val = cmul_sgn(val, ds[d]->ptr[int64_t(b) * ds[d]->sb + row * ds[d]->sr + x], sgn);
                    }
                }
                if (a.dx) val = cscale(val, (a.dz[z] * a.dy[y]) * a.dx[x]);
                const uint32_t off = uint32_t((row * kN + x) * int64_t(sizeof(V)));
                store_at<V>(ob, off, a.accumulate ? cadd(val, load_at<V>(ob, off)) : val);
            }
        }
    }
}

bool nufft_inv_xy_supported(int64_t ny, int64_t nx, int64_t Ly, int64_t Lx, int shift_y, int shift_x) {
    return ny == kN && nx == kN && Ly == kL && Lx == kL && shift_y == -kN / 2 && shift_x == -kN / 2;
}

void launch_nufft_inv_xy(const NufftInvXyArgs& a, hipStream_t s) {
    HASTY_REQUIRE(a.nz >= 1 && a.nbatch >= 1, "fused inverse fine-grid pass: empty batch or plane count");
    HASTY_REQUIRE(a.sign == 1 || a.sign == -1, "fused inverse fine-grid pass: sign must be +1 or -1");
    HASTY_REQUIRE(a.in_b >= int64_t(a.nz) * kL * kL && a.out_b >= int64_t(a.nz) * kN * kN,
                  "fused inverse fine-grid pass: batch stride too small");
    const size_t bytes = size_t(kInvLds) * sizeof(float2);
    auto kern = a.sign < 0 ? nufft_inv_xy_kernel<-1> : nufft_inv_xy_kernel<1>;
    static const bool lds_ok = [&] {
        for (auto k : {nufft_inv_xy_kernel<-1>, nufft_inv_xy_kernel<1>})
            HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
        return true;
    }();
    (void)lds_ok;
    const int groups = int(std::max<int64_t>(1, std::min<int64_t>((256 + a.nz - 1) / a.nz, a.nbatch)));
    NufftInvXyArgs b = a;
    b.groups = groups;
    const float2* tw = reinterpret_cast<const float2*>(twiddles(kL, false));
    HASTY_PROF("inv_xy_fft_crop", s);
    hipLaunchKernelGGL(kern, dim3(unsigned(int64_t(a.nz) * groups)), dim3(kT), bytes, s, b, tw);
    HASTY_CHECK_HIP(hipGetLastError());
}

bool nufft_fwd_xy_supported(int64_t ny, int64_t nx, int64_t Ly, int64_t Lx, int shift_y, int shift_x) {
    return ny == kN && nx == kN && Ly == kL && Lx == kL && shift_y == -kN / 2 && shift_x == -kN / 2;
}

void launch_nufft_fwd_xy(const NufftFwdXyArgs& a, hipStream_t s) {
    HASTY_REQUIRE(a.nz >= 1 && a.nbatch >= 1, "fused fine-grid pass: empty batch or plane count");
    HASTY_REQUIRE(a.sign == 1 || a.sign == -1, "fused fine-grid pass: sign must be +1 or -1");
    HASTY_REQUIRE(a.grid_b >= int64_t(a.nz) * kL * kL, "fused fine-grid pass: grid stride too small");
    const size_t bytes = size_t(kLdsElems) * sizeof(float2);
    auto kern = a.sign < 0 ? nufft_fwd_xy_kernel<-1> : nufft_fwd_xy_kernel<1>;
    static const bool lds_ok = [&] {
        for (auto k : {nufft_fwd_xy_kernel<-1>, nufft_fwd_xy_kernel<1>})
            HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
        return true;
    }();
    (void)lds_ok;
    const int groups = int(std::max<int64_t>(1, std::min<int64_t>((256 + a.nz - 1) / a.nz, a.nbatch)));
    NufftFwdXyArgs b = a;
    b.groups = groups;
    const float2* tw = reinterpret_cast<const float2*>(twiddles(kL, false));
    HASTY_PROF("fwd_xy_pad_fft", s);
    hipLaunchKernelGGL(kern, dim3(unsigned(int64_t(a.nz) * groups)), dim3(kT), bytes, s, b, tw);
    HASTY_CHECK_HIP(hipGetLastError());
}

}  // namespace hasty
