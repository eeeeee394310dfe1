// Toeplitz-embedded normal operator on gfx950.
//
// One apply of T on a batch [B, n...] is five Stockham passes (three in 2-D):
//   P1 row  : load in[b] (.) diagonals, zero-pad x to 2n_x, FFT      -> A [B, nz, ny, 2nx]
//   P2 col y: zero-pad y to 2n_y, FFT                                -> Bf[B, nz, 2ny, 2nx]
//   P3 col z: zero-pad z to 2n_z, FFT, * K, IFFT, keep [nz-1, 2nz-1) -> Bf (in place)
//   P4 col y: IFFT, keep [ny-1, 2ny-1)                               -> A
//   P5 row  : IFFT, keep [nx-1, 2nx-1), * 1/(2n)^d, (.) diagonals,
//             (+ out), optionally summed over the batch (coil combine) -> out
// The kernel K is read once per (ky, kx) column tile per batch element and
// the (2n)^d grid never exists: P1/P2 skip the all-zero lines (pruning) and
// P4/P5 only produce what the crop keeps.
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <vector>

#include "nufft.hpp"
#include "toeplitz_fused.hpp"

namespace hasty {

struct ToeplitzDims {
    int dim;
    int64_t n[3];     // slots (z, y, x); n[0] = 1 in 2-D
    int64_t nimg() const { return n[0] * n[1] * n[2]; }
    int64_t nA() const { return n[0] * n[1] * 2 * n[2]; }
    int64_t nB() const { return n[0] * (2 * n[1]) * 2 * n[2]; }
    int64_t nK() const { return (dim == 3 ? 2 * n[0] : 1) * 2 * n[1] * 2 * n[2]; }
};

static ToeplitzDims toeplitz_dims(int dim, const int64_t* n) {
    HASTY_REQUIRE(dim == 2 || dim == 3, "Toeplitz dim must be 2 or 3");
    HASTY_REQUIRE(n != nullptr, "shape is NULL");
    ToeplitzDims d{};
    d.dim = dim;
    d.n[0] = d.n[1] = d.n[2] = 1;
    for (int a = 0; a < dim; ++a) {
        HASTY_REQUIRE(n[a] >= 1, "Toeplitz shape must be positive");
        HASTY_REQUIRE(is_smooth(2 * n[a]), "Toeplitz grid 2n must be 2^a 3^b 5^c");
        d.n[3 - dim + a] = n[a];
    }
    return d;
}

// The fused passes (toeplitz_fused.hip) run coil-summed 3-D applies of at
// least two coils on n x n planes (n = 64, 96, 128): P1 + P2 in one kernel writing x
// class-major, P3 reading the kernel through the same permutation, P4 + P5
// in one kernel with the coil sum.  HASTY_TOEP_FUSED=0 turns them off, =2
// keeps only the last one (P1 then stores class-major itself);
// HASTY_TOEP_GROUPS sets the coil groups of the last one (default: about 256
// blocks).
static int toep_fused_mode() {
    static const int v = [] {
        const char* e = std::getenv("HASTY_TOEP_FUSED");
        return e ? (e[0] == '0' ? 0 : (e[0] == '2' ? 2 : 1)) : 1;
    }();
    return v;
}
static int toep_fused_groups(int64_t nz, int64_t B) {
    const char* e = std::getenv("HASTY_TOEP_GROUPS");   // read per apply (tests set it)
    const int env = e ? std::atoi(e) : 0;
    int64_t g = env > 0 ? env : std::max<int64_t>(1, (256 + nz - 1) / nz);
    return int(std::max<int64_t>(1, std::min<int64_t>(g, B)));
}

// Frames of a batched apply: frame f (f < n) reads input + f*input_f and
// kernel + f*kernel_f, writes output + f*output_f, and takes its diagonals at
// ptr + f*sf (DiagOp::sf); each frame has its own B batch elements in the
// buffers (bufA, bufB hold n*B elements).
struct ToepFrames {
    int n = 1;
    int64_t input_f = 0, kernel_f = 0, output_f = 0;
};

static DiagOp<float2> diag_at_frame(DiagOp<float2> d, int f) {
    if (d.ptr) d.ptr += int64_t(f) * d.sf;
    return d;
}

// Core Toeplitz apply on one batch (see file comment).  With frames.n > 1
// and the fused passes, each pass is one launch over all frames (the frame
// is the second grid axis of the fused passes; the z pass runs the frames'
// n*B lines with each frame's kernel); otherwise the frames run one after
// another.
static void toeplitz_core(const ToeplitzDims& d, int64_t B, const float2* input, float2* output,
                          const float2* kernel, float2* bufA, float2* bufB, DiagOp<float2> in1,
                          DiagOp<float2> in2, DiagOp<float2> out1, DiagOp<float2> out2,
                          DiagOp<float2> io, int accumulate, int sum_batch, hipStream_t s,
                          const ToepFrames& fr = ToepFrames{}) {
    const int64_t nz = d.n[0], ny = d.n[1], nx = d.n[2];
    const int64_t Lz = d.dim == 3 ? 2 * nz : 1, Ly = 2 * ny, Lx = 2 * nx;
    // (one coil per launch: a plane per block leaves half the CUs idle at
    // nz = 128, and P4 + P5 are faster there)
    const bool fused = sum_batch && d.dim == 3 && B >= 2 && toep_crop_sum_supported(ny, nx) &&
                       toep_fused_mode() != 0 && !fft_force_stockham();
    const bool fused_first = fused && toep_fused_mode() == 1;
    HASTY_REQUIRE(fr.n >= 1, "frame count must be >= 1");
    if (fr.n > 1 && !fused_first) {
        for (int f = 0; f < fr.n; ++f)
            toeplitz_core(d, B, input + f * fr.input_f, output + f * fr.output_f, kernel + f * fr.kernel_f, bufA,
                          bufB, diag_at_frame(in1, f), diag_at_frame(in2, f), diag_at_frame(out1, f),
                          diag_at_frame(out2, f), diag_at_frame(io, f), accumulate, sum_batch, s);
        return;
    }
    if (fused_first) {
        ToepPadFftArgs f{};
        f.in = input; f.in_b = d.nimg();
        f.m1 = in1; f.m2 = in2;
        f.grid = bufB; f.grid_b = d.nB();
        f.nbatch = int(B); f.nz = int(nz); f.n = int(nx);
        f.nframes = fr.n; f.in_f = fr.input_f; f.grid_f = B * d.nB();
        launch_toep_pad_fft(f, s);
    }
    // P1
    RowLoadArgs<float2> r{};
    r.in = input; r.out = bufA;
    r.nrows = nz * ny; r.nbatch = int(B);
    r.in_b = d.nimg(); r.in_r = nx;
    r.out_b = nz * ny * Lx; r.out_r = Lx;
    r.L = int(Lx); r.LP = int(Lx) + 1; r.in_len = int(nx); r.in_shift = 0;
    r.sign = -1; r.nl = rows_per_block<float2>(int(Lx), int(Lx));
    r.m1 = in1; r.m2 = in2;
    r.dz = r.dy = r.dx = nullptr; r.ny = int(ny);
    r.plan = make_fft_plan(int(Lx));
    r.tw = reinterpret_cast<const float2*>(twiddles(int(Lx), false));
    r.tag = "toep_P1_x_pad_fft";
    r.perm4 = fused ? 1 : 0;
    if (!fused_first) launch_row_load<float2>(r, s);
    if (d.dim == 3) {
        // P2
        ColArgs<float2> c = make_col_args<float2>(int(Ly));
        c.in = bufA; c.ni = int(Lx); c.no = int(nz); c.nbatch = int(B);
        c.in_s = Lx; c.in_o = ny * Lx; c.in_b = nz * ny * Lx; c.in_len = int(ny); c.in_shift = 0;
        c.out = bufB; c.out_s = Lx; c.out_o = Ly * Lx; c.out_b = nz * Ly * Lx;
        c.out_len = int(Ly); c.out_shift = 0; c.sign = -1;
    c.tag = "toep_P2_y_pad_fft";
        if (!fused_first) launch_col_pass<float2>(c, s);
        // P3 (in place on the first nz planes of bufB)
        ColArgs<float2> z = make_col_args<float2>(int(Lz));
        z.in = bufB; z.ni = int(Lx); z.no = int(Ly); z.nbatch = int(B * fr.n);
        if (fr.n > 1) { z.psf_f = fr.kernel_f; z.psf_div = int(B); }
        z.in_s = Ly * Lx; z.in_o = Lx; z.in_b = nz * Ly * Lx; z.in_len = int(nz); z.in_shift = 0;
        z.out = bufB; z.out_s = Ly * Lx; z.out_o = Lx; z.out_b = nz * Ly * Lx;
        z.out_len = int(nz); z.out_shift = int(nz - 1); z.sign = -1;
        z.psf = kernel; z.psf_o = Lx; z.psf_s = Ly * Lx;
        z.psf_q = fused ? int(Lx / 4) : 0;
    z.tag = "toep_P3_z_psf";
        launch_col_pass<float2>(z, s);
        if (fused) {
            // P4 + P5 in one pass; bufA (the P1 output, consumed by P2) holds the partial sums
            ToepCropSumArgs f{};
            f.grid = bufB; f.grid_b = d.nB();
            f.out = output;
            f.partial = bufA; f.partial_elems = B * d.nA();
            f.m1 = out1; f.m2 = out2; f.io = io;
            f.scale = float(1.0 / double(d.nK()));
            f.nbatch = int(B); f.nz = int(nz);
            f.groups = toep_fused_groups(nz, B);
            f.accumulate = accumulate;
            f.n = int(nx);
            f.nframes = fr.n; f.grid_f = B * d.nB(); f.out_f = fr.output_f;
            f.partial_elems = fr.n * B * d.nA();
            f.partial_f = int64_t(f.groups) * d.nimg();
            launch_toep_crop_sum(f, s);
            return;
        }
        // P4
        ColArgs<float2> y = make_col_args<float2>(int(Ly));
        y.in = bufB; y.ni = int(Lx); y.no = int(nz); y.nbatch = int(B);
        y.in_s = Lx; y.in_o = Ly * Lx; y.in_b = nz * Ly * Lx; y.in_len = int(Ly); y.in_shift = 0;
        y.out = bufA; y.out_s = Lx; y.out_o = ny * Lx; y.out_b = nz * ny * Lx;
        y.out_len = int(ny); y.out_shift = int(ny - 1); y.sign = +1;
        y.tag = "toep_P4_y_ifft_crop";
        launch_col_pass<float2>(y, s);
    } else {
        // 2-D: the y pass carries the kernel product, in place on bufA
        ColArgs<float2> y = make_col_args<float2>(int(Ly));
        y.in = bufA; y.ni = int(Lx); y.no = 1; y.nbatch = int(B);
        y.in_s = Lx; y.in_o = 0; y.in_b = ny * Lx; y.in_len = int(ny); y.in_shift = 0;
        y.out = bufA; y.out_s = Lx; y.out_o = 0; y.out_b = ny * Lx;
        y.out_len = int(ny); y.out_shift = int(ny - 1); y.sign = -1;
        y.psf = kernel; y.psf_o = 0; y.psf_s = Lx;
        y.tag = "toep2d_y_psf";
        launch_col_pass<float2>(y, s);
    }
    // P5
    RowStoreArgs<float2> w{};
    w.in = bufA; w.out = output;
    w.nrows = nz * ny; w.nbatch = int(B);
    w.in_b = nz * ny * Lx; w.in_r = Lx;
    w.out_b = sum_batch ? 0 : d.nimg(); w.out_r = nx;
    w.L = int(Lx); w.LP = int(Lx) + 1; w.out_len = int(nx); w.out_shift = int(nx - 1);
    w.sign = +1; w.nl = rows_per_block<float2>(int(Lx), int(nx));
    w.scale = float(1.0 / double(d.nK()));
    w.m1 = out1; w.m2 = out2; w.io = io;
    w.dz = w.dy = w.dx = nullptr; w.ny = int(ny);
    w.accumulate = accumulate; w.sum_batch = sum_batch;
    w.scratch = bufB; w.scratch_elems = B * d.nB();   // bufB is free once P4 has read it
    w.plan = make_fft_plan(int(Lx));
    w.tw = reinterpret_cast<const float2*>(twiddles(int(Lx), false));
    w.tag = "toep_P5_x_ifft_crop";
    launch_row_store<float2>(w, s);
}

static int64_t scratch_elems(const ToeplitzDims& d, int64_t B) { return B * (d.nA() + d.nB()); }

// ---------------------------------------------------------------------------
// toeplitz_kernel: FP64 PSF via a type-1 NUFFT at 2n-1 modes, FP64 FFT.
// ---------------------------------------------------------------------------
__global__ void fill_ones_c128(double2* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(1.0, 0.0);
}
__global__ void c64_to_c128(const float2* src, double2* dst, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_double2(double(src[i].x), double(src[i].y));
}
__global__ void c128_to_c64_scaled(const double2* src, float2* dst, int64_t n, double scale) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_float2(float(src[i].x * scale), float(src[i].y * scale));
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (bytes && hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(HASTY_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <class T> T* as() { return static_cast<T*>(p); }
};

void nufft_setpts_typed(NufftPlan* p, int64_t M, const void* c0, const void* c1, const void* c2,
                        bool coords_f64);

static void toeplitz_kernel_impl(int dim, const int64_t* n, int64_t M, const void* c0,
                                 const void* c1, const void* c2, const void* weights,
                                 void* kernel, double tol, hipStream_t s) {
    ToeplitzDims d = toeplitz_dims(dim, n);
    HASTY_REQUIRE(kernel != nullptr, "kernel is NULL");
    HASTY_REQUIRE(M >= 0, "M must be >= 0");
    if (tol <= 0) tol = 1e-9;
    int64_t modes[3];
    for (int a = 0; a < dim; ++a) modes[a] = 2 * n[a] - 1;
    std::unique_ptr<NufftPlan> plan(nufft_makeplan(1, dim, modes, +1, 1, tol, true, nullptr, s));
    nufft_setpts_typed(plan.get(), M, c0, c1, c2, false);
    DevBuf wts(size_t(std::max<int64_t>(M, 1)) * 16);
    const unsigned blocks = unsigned(ceil_div(std::max<int64_t>(M, 1), 256));
    if (M > 0) {
        if (weights)
            hipLaunchKernelGGL(c64_to_c128, dim3(blocks), dim3(256), 0, s,
                               static_cast<const float2*>(weights), wts.as<double2>(), M);
        else
            hipLaunchKernelGGL(fill_ones_c128, dim3(blocks), dim3(256), 0, s, wts.as<double2>(), M);
        HASTY_CHECK_HIP(hipGetLastError());
    }
    int64_t nred = 1, nK = 1;
    for (int a = 0; a < dim; ++a) { nred *= modes[a]; nK *= 2 * n[a]; }
    DevBuf red(size_t(nred) * 16);
    nufft_execute(plan.get(), wts.p, red.p);
    HASTY_CHECK_HIP(hipStreamSynchronize(s));
    plan.reset();   // release the FP64 fine grid before the FFT buffers
    // embed lags at index m + n - 1 of the 2n grid and FFT (FP64)
    PadGeometry pg{};
    for (int sl = 0; sl < 3; ++sl) { pg.n[sl] = 1; pg.L[sl] = 1; pg.shift_in[sl] = 0; pg.shift_out[sl] = 0; }
    for (int a = 0; a < dim; ++a) { pg.n[3 - dim + a] = modes[a]; pg.L[3 - dim + a] = 2 * n[a]; }
    const int64_t nA = pg.n[0] * pg.n[1] * pg.L[2];
    const int64_t nB = pg.n[0] * pg.L[1] * pg.L[2];
    DevBuf K64(size_t(nK) * 16), bA(size_t(nA) * 16), bB(size_t(nB) * 16);
    DiagOp<double2> none{};
    pruned_forward<double2>(pg, 1, red.as<double2>(), K64.as<double2>(), bA.as<double2>(),
                            bB.as<double2>(), -1, none, none, nullptr, s);
    int64_t nimg = 1;
    for (int a = 0; a < dim; ++a) nimg *= n[a];
    hipLaunchKernelGGL(c128_to_c64_scaled, dim3(unsigned(ceil_div(nK, 256))), dim3(256), 0, s,
                       K64.as<double2>(), static_cast<float2*>(kernel), nK, 1.0 / double(nimg));
    HASTY_CHECK_HIP(hipGetLastError());
    HASTY_CHECK_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
}

// ---------------------------------------------------------------------------
// SENSE normal operators
// ---------------------------------------------------------------------------
// out[i][p] = sum_j Wm[i][j] in[j][p] with Wm = W ([C, C], row-major) or W^H
// (herm): a C x C mix of stacked images, one pixel per thread, 16 outputs per
// pass in registers.  W's index is wave-uniform, so its loads are scalar.
__global__ void __launch_bounds__(256) stack_mix_kernel(const float2* __restrict__ W,
                                                        const float2* __restrict__ in,
                                                        float2* __restrict__ out, int C,
                                                        int64_t nimg, int herm) {
    const int64_t px = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (px >= nimg) return;
    constexpr int U = 16;
    for (int i0 = 0; i0 < C; i0 += U) {
        float2 acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = make_float2(0.f, 0.f);
        for (int j = 0; j < C; ++j) {
            const float2 v = in[int64_t(j) * nimg + px];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (i0 + u < C) {
                    const float2 w = herm ? cconj(W[int64_t(j) * C + i0 + u]) : W[int64_t(i0 + u) * C + j];
                    acc[u].x += w.x * v.x - w.y * v.y;
                    acc[u].y += w.x * v.y + w.y * v.x;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i0 + u < C) out[int64_t(i0 + u) * nimg + px] = acc[u];
    }
}

static void stack_mix(const float2* W, const float2* in, float2* out, int C, int64_t nimg, int herm,
                      hipStream_t s) {
    hipLaunchKernelGGL(stack_mix_kernel, dim3(unsigned(ceil_div(nimg, 256))), dim3(256), 0, s, W, in,
                       out, C, nimg, herm);
    HASTY_CHECK_HIP(hipGetLastError());
}

// One handle for the normal-operator family of the reference
// (cpp/corelib/lib/mri/normal.cppm):
//   kind 1 (NORMAL_T_T1_OP, NORMAL_IDT_T1_OP, NORMAL_IDTW_T1_OP), x [n] -> [n]:
//     out = sum_c conj(S'_c) (.) sum_l conj(R_l) (.) T_l(R_l (.) S_c (.) x)
//     with S' = S, or S'_c = sum_i conj(W_ic) S_i when weights are given: the
//     reference's  sum_i conj(S_i) sum_j W_ij stack_j  regrouped as
//     sum_j conj(sum_i conj(W_ij) S_i) stack_j, so the coil sum stays fused
//     into the last FFT pass.  S' is formed once, at creation.
//   kind 2 (NORMAL_T_T2_OP, NORMAL_IDT_T2_OP, NORMAL_IDTW_T2_OP), Psi [B, n] -> [B, n]:
//     out_i = sum_j W_ij conj(x) (.) sum_l conj(R_l) (.) T_l(R_l (.) x (.) Psi_j)
//     (the map Psi_j -> stack_j is the same linear map for every j, so the
//     reference's mix of the outputs equals a mix of the inputs, done first).
// R_l (kerneldiags) may be absent (all ones).
// kind 3 (frames): F frames of NORMAL_T_T1_OP sharing the coil maps, frame
//   f with its own kernel: out[f] = sum_c conj(S_c) (.) T_f(S_c (.) x[f]),
//   x and out [F, n], kernels [F, 2n...]; every pass is one launch over the
//   frames and coils (toeplitz_core with frames).
struct NormalOp {
    ToeplitzDims d;
    int kind = 1;
    int nframes = 1;
    int nstack = 0, chunk = 0;          // coils / stacked images, and how many per launch
    int nkernels = 1;
    const float2* stack = nullptr;      // kind 1: coil maps S [C, n]
    const float2* out_maps = nullptr;   // kind 1: S' (== stack, or owned_maps)
    const float2* x = nullptr;          // kind 2: the image X [n]
    const float2* kernels = nullptr;    // [L, 2n...]
    const float2* kerneldiags = nullptr;  // [L, n] or NULL
    const float2* weights = nullptr;    // kind 2: [B, B] or NULL
    float2* owned_maps = nullptr;
    float2* mixbuf = nullptr;
    float2* bufA = nullptr;
    float2* bufB = nullptr;
    hipStream_t stream = nullptr;
    ~NormalOp() {
        for (float2* p : {owned_maps, mixbuf, bufA, bufB})
            if (p) (void)hipFree(p);
    }
};

static void normal_apply(NormalOp* op, const float2* in, float2* out) {
    const ToeplitzDims& d = op->d;
    const int64_t nimg = d.nimg(), nK = d.nK();
    auto bcast = [&](const float2* p, int type) {
        DiagOp<float2> o{};
        if (p) { o.ptr = p; o.sb = 0; o.sr = d.n[2]; o.type = type; }
        return o;
    };
    const DiagOp<float2> none{};
    if (op->kind == 3) {
        ToepFrames fr;
        fr.n = op->nframes;
        fr.input_f = 0;            // the coil maps are shared
        fr.kernel_f = nK;
        fr.output_f = nimg;
        DiagOp<float2> xin = bcast(in, HASTY_MULT);
        xin.sf = nimg;             // frame f's image
        for (int c0 = 0; c0 < op->nstack; c0 += op->chunk) {
            const int cb = std::min(op->chunk, op->nstack - c0);
            DiagOp<float2> io{};   // (.) conj(S_c), summed over the frame's coils
            io.ptr = op->out_maps + int64_t(c0) * nimg; io.sb = nimg; io.sr = d.n[2];
            io.type = HASTY_MULT_CONJ;
            toeplitz_core(d, cb, op->stack + int64_t(c0) * nimg, out, op->kernels, op->bufA, op->bufB, xin,
                          none, none, none, io, c0 > 0 ? 1 : 0, 1, op->stream, fr);
        }
        return;
    }
    if (op->kind == 1) {
        for (int l = 0; l < op->nkernels; ++l) {
            const float2* K = op->kernels + int64_t(l) * nK;
            const float2* R = op->kerneldiags ? op->kerneldiags + int64_t(l) * nimg : nullptr;
            for (int c0 = 0; c0 < op->nstack; c0 += op->chunk) {
                const int cb = std::min(op->chunk, op->nstack - c0);
                DiagOp<float2> io{};   // (.) conj(S'_c), summed over coils
                io.ptr = op->out_maps + int64_t(c0) * nimg; io.sb = nimg; io.sr = d.n[2];
                io.type = HASTY_MULT_CONJ;
                toeplitz_core(d, cb, op->stack + int64_t(c0) * nimg, out, K, op->bufA, op->bufB,
                              bcast(in, HASTY_MULT), bcast(R, HASTY_MULT), none,
                              bcast(R, HASTY_MULT_CONJ), io, (l > 0 || c0 > 0) ? 1 : 0, 1, op->stream);
            }
        }
    } else {
        const float2* src = in;
        if (op->weights) {
            stack_mix(op->weights, in, op->mixbuf, op->nstack, nimg, 0, op->stream);
            src = op->mixbuf;
        }
        for (int c0 = 0; c0 < op->nstack; c0 += op->chunk) {
            const int cb = std::min(op->chunk, op->nstack - c0);
            for (int l = 0; l < op->nkernels; ++l) {
                const float2* K = op->kernels + int64_t(l) * nK;
                const float2* R = op->kerneldiags ? op->kerneldiags + int64_t(l) * nimg : nullptr;
                toeplitz_core(d, cb, src + int64_t(c0) * nimg, out + int64_t(c0) * nimg, K, op->bufA,
                              op->bufB, bcast(op->x, HASTY_MULT), bcast(R, HASTY_MULT),
                              bcast(op->x, HASTY_MULT_CONJ), bcast(R, HASTY_MULT_CONJ), none,
                              l > 0 ? 1 : 0, 0, op->stream);
            }
        }
    }
}

static void* normal_alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes && hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(HASTY_ERR_ALLOC, "normal operator workspace allocation failed");
    }
    return p;
}

static NormalOp* normal_make(int kind, int dim, const int64_t* n, int nstack, const void* stack,
                             const void* x, int nkernels, const void* kernels,
                             const void* kerneldiags, const void* weights, int chunk,
                             hipStream_t stream, int nframes = 1) {
    HASTY_REQUIRE(nframes >= 1 && nframes <= 65535, "the number of frames must be in [1, 65535]");
    HASTY_REQUIRE(nstack >= 1, "the number of coils / stacked images must be >= 1");
    HASTY_REQUIRE(nkernels >= 1, "the number of kernels must be >= 1");
    HASTY_REQUIRE(kernels != nullptr, "kernels must be non-NULL");
    HASTY_REQUIRE(kind == 2 || stack != nullptr, "smaps must be non-NULL");
    HASTY_REQUIRE(kind != 2 || x != nullptr, "x must be non-NULL");
    auto op = std::make_unique<NormalOp>();
    op->d = toeplitz_dims(dim, n);
    op->kind = kind;
    op->nframes = nframes;
    op->nstack = nstack;
    op->chunk = chunk > 0 ? std::min(chunk, nstack) : nstack;
    op->nkernels = nkernels;
    op->stack = static_cast<const float2*>(stack);
    op->out_maps = op->stack;
    op->x = static_cast<const float2*>(x);
    op->kernels = static_cast<const float2*>(kernels);
    op->kerneldiags = static_cast<const float2*>(kerneldiags);
    op->stream = stream;
    const int64_t nimg = op->d.nimg();
    if (weights) {
        const size_t bytes = size_t(nstack) * size_t(nimg) * 8;
        if (kind == 1) {
            op->owned_maps = static_cast<float2*>(normal_alloc(bytes));
            stack_mix(static_cast<const float2*>(weights), op->stack, op->owned_maps, nstack, nimg, 1,
                      stream);
            op->out_maps = op->owned_maps;
        } else {
            op->weights = static_cast<const float2*>(weights);
            op->mixbuf = static_cast<float2*>(normal_alloc(bytes));
        }
    }
    op->bufA = static_cast<float2*>(normal_alloc(size_t(nframes) * op->chunk * op->d.nA() * 8));
    op->bufB = static_cast<float2*>(normal_alloc(size_t(nframes) * op->chunk * op->d.nB() * 8));
    return op.release();
}

}  // namespace hasty

// ===========================================================================
// C-ABI
// ===========================================================================
using namespace hasty;

extern "C" {

int hasty_toeplitz_kernel(int dim, const int64_t* n, int64_t M, const void* c0, const void* c1,
                          const void* c2, const void* weights, void* kernel, double tol,
                          void* stream) {
    return guarded([&] {
        toeplitz_kernel_impl(dim, n, M, c0, c1, c2, weights, kernel, tol,
                             static_cast<hipStream_t>(stream));
    });
}

int64_t hasty_toeplitz_scratch_bytes(int dim, const int64_t* n, int64_t batch) {
    int64_t r = -1;
    int st = guarded([&] { r = scratch_elems(toeplitz_dims(dim, n), batch) * 8; });
    return st == HASTY_OK ? r : -1;
}

int hasty_transform_toeplitz_kernel(int dim, const int64_t* n, void* kernel, void* stream) {
    return guarded([&] {
        (void)toeplitz_dims(dim, n);
        (void)stream;
        HASTY_REQUIRE(kernel != nullptr, "kernel is NULL");
    });
}

int hasty_toeplitz_multiplication(int dim, const int64_t* n, int64_t batch, const void* input,
                                  void* output, const void* kernel, void* scratch,
                                  int64_t scratch_bytes, const void* mult1, const void* mult2,
                                  const int32_t* types, int accumulate, void* stream) {
    return guarded([&] {
        ToeplitzDims d = toeplitz_dims(dim, n);
        HASTY_REQUIRE(batch >= 0, "batch must be >= 0");
        HASTY_REQUIRE(input && output && kernel, "input, output and kernel must be non-NULL");
        HASTY_REQUIRE(types != nullptr, "types is NULL");
        for (int i = 0; i < 5; ++i)
            HASTY_REQUIRE(types[i] >= 0 && types[i] <= 2, "invalid ToeplitzMultType");
        HASTY_REQUIRE(accumulate == 0 || accumulate == 1, "invalid ToeplitzAccumulateType");
        if (batch == 0) return;
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int64_t need = scratch_elems(d, batch) * 8;
        void* scr = scratch;
        bool owned = false;
        if (scr == nullptr) {
            HASTY_CHECK_HIP(hipMallocAsync(&scr, size_t(need), s));
            owned = true;
        } else {
            HASTY_REQUIRE(scratch_bytes >= need, "scratch buffer too small");
        }
        float2* bufA = static_cast<float2*>(scr);
        float2* bufB = bufA + batch * d.nA();
        auto diag = [&](const void* p, int type, bool batched) {
            DiagOp<float2> o{};
            if (p && type != HASTY_MULT_NONE) {
                o.ptr = static_cast<const float2*>(p);
                o.sb = batched ? d.nimg() : 0;
                o.sr = d.n[2];
                o.type = type;
            }
            return o;
        };
        toeplitz_core(d, batch, static_cast<const float2*>(input), static_cast<float2*>(output),
                      static_cast<const float2*>(kernel), bufA, bufB, diag(mult1, types[1], false),
                      diag(mult2, types[3], false), diag(mult1, types[2], false),
                      diag(mult2, types[4], false), diag(input, types[0], true), accumulate, 0, s);
        if (owned) HASTY_CHECK_HIP(hipFreeAsync(scr, s));
    });
}

int hasty_toeplitz_multiply(int dim, const int64_t* n, int64_t batch, void* inout,
                            const void* kernel, void* stream) {
    const int32_t types[5] = {0, 0, 0, 0, 0};
    return hasty_toeplitz_multiplication(dim, n, batch, inout, inout, kernel, nullptr, 0, nullptr,
                                         nullptr, types, 0, stream);
}

int hasty_normal_create(int dim, const int64_t* n, int ncoils, const void* smaps,
                        const void* kernel, int coil_batch, void* stream, hasty_normal_op* out) {
    return guarded([&] {
        HASTY_REQUIRE(out != nullptr, "output handle pointer is NULL");
        HASTY_REQUIRE(smaps && kernel, "smaps and kernel must be non-NULL");
        *out = reinterpret_cast<hasty_normal_op>(normal_make(1, dim, n, ncoils, smaps, nullptr, 1, kernel,
                                                             nullptr, nullptr, coil_batch,
                                                             static_cast<hipStream_t>(stream)));
    });
}

int hasty_normal_frames_create(int dim, const int64_t* n, int nframes, int ncoils, const void* smaps,
                               const void* kernels, int coil_batch, void* stream, hasty_normal_op* out) {
    return guarded([&] {
        HASTY_REQUIRE(out != nullptr, "output handle pointer is NULL");
        HASTY_REQUIRE(smaps && kernels, "smaps and kernels must be non-NULL");
        *out = reinterpret_cast<hasty_normal_op>(normal_make(3, dim, n, ncoils, smaps, nullptr, nframes, kernels,
                                                             nullptr, nullptr, coil_batch,
                                                             static_cast<hipStream_t>(stream), nframes));
    });
}

int hasty_normal_idt_create(int dim, const int64_t* n, int ncoils, const void* smaps, int nkernels,
                            const void* kernels, const void* kerneldiags, const void* weights,
                            int coil_batch, void* stream, hasty_normal_op* out) {
    return guarded([&] {
        HASTY_REQUIRE(out != nullptr, "output handle pointer is NULL");
        *out = reinterpret_cast<hasty_normal_op>(normal_make(1, dim, n, ncoils, smaps, nullptr, nkernels,
                                                             kernels, kerneldiags, weights, coil_batch,
                                                             static_cast<hipStream_t>(stream)));
    });
}

int hasty_normal_t2_create(int dim, const int64_t* n, int nstack, const void* x, int nkernels,
                           const void* kernels, const void* kerneldiags, const void* weights,
                           int stack_batch, void* stream, hasty_normal_op* out) {
    return guarded([&] {
        HASTY_REQUIRE(out != nullptr, "output handle pointer is NULL");
        *out = reinterpret_cast<hasty_normal_op>(normal_make(2, dim, n, nstack, nullptr, x, nkernels,
                                                             kernels, kerneldiags, weights, stack_batch,
                                                             static_cast<hipStream_t>(stream)));
    });
}

int hasty_normal_apply(hasty_normal_op h, const void* x, void* out) {
    return guarded([&] {
        HASTY_REQUIRE(h != nullptr, "normal operator handle is NULL");
        HASTY_REQUIRE(x && out, "x and out must be non-NULL");
        // out is written by the first launch while x is still read by later ones
        HASTY_REQUIRE(x != out, "x and out must not alias");
        normal_apply(reinterpret_cast<NormalOp*>(h), static_cast<const float2*>(x),
                     static_cast<float2*>(out));
    });
}

int hasty_normal_set_stream(hasty_normal_op h, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(h != nullptr, "normal operator handle is NULL");
        reinterpret_cast<NormalOp*>(h)->stream = static_cast<hipStream_t>(stream);
    });
}

int hasty_normal_destroy(hasty_normal_op h) {
    return guarded([&] { delete reinterpret_cast<NormalOp*>(h); });
}

}  // extern "C"
