// Conjugate-gradient vector operations with FP64 accumulation, scalars kept
// on the device.  Restates the step body of cg.cppm:137-150 and the driver
// of cg.cppm:157-180 (r = b - Ax, p = r, rz = Re<r,r>; per iteration
// alpha = rz / Re<p,Ap>, x += alpha p, r -= alpha Ap, rz' = Re<r,r>,
// p = r + (rz'/rz) p; stop once sqrt(rz) < tol) without a host sync per
// iteration: a device flag turns the remaining launches into no-ops.
#include <mutex>

#include "common.hpp"

namespace hasty {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;

// state layout (doubles)
enum : int {
    ST_RZ = 0, ST_PAP = 1, ST_ALPHA = 2, ST_RZNEW = 3, ST_BETA = 4, ST_DONE = 5, ST_TOL = 6,
    ST_ITER = 7, ST_PART = 16   // kRedBlocks partials follow
};
static_assert(ST_PART + 2 * kRedBlocks <= HASTY_CG_STATE_DOUBLES, "state too small");

__device__ inline double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// partial[blockIdx] = Re sum conj(a) b ; partial[kRedBlocks + blockIdx] = Im
__global__ void __launch_bounds__(kRedThreads) dot_partial_kernel(const float2* a, const float2* b,
                                                                  int64_t n, double* part,
                                                                  const double* done) {
    // batched systems (hasty_cg_*_batched): system blockIdx.y of [nsys, n]
    // vectors and [nsys, HASTY_CG_STATE_DOUBLES] states
    const int64_t sy = blockIdx.y;
    a += sy * n; b += sy * n; part += sy * HASTY_CG_STATE_DOUBLES;
    if (done) done += sy * HASTY_CG_STATE_DOUBLES;
    if (done && *done != 0.0) return;
    __shared__ double sh[kRedThreads];
    double re = 0.0, im = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float2 x = a[i], y = b[i];
        re += double(x.x) * double(y.x) + double(x.y) * double(y.y);
        im += double(x.x) * double(y.y) - double(x.y) * double(y.x);
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = re;
        part[kRedBlocks + blockIdx.x] = im;
    }
}

__global__ void __launch_bounds__(kRedThreads) sum_partials_kernel(const double* part, int nparts,
                                                                   double* out) {
    __shared__ double sh[kRedThreads];
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        re += part[i];
        im += part[kRedBlocks + i];
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) {
        out[0] = re;
        out[1] = im;
    }
}

// CG: alpha = rz / pAp
__global__ void cg_alpha_kernel(double* st) {
    __shared__ double sh[kRedThreads];
    st += int64_t(blockIdx.y) * HASTY_CG_STATE_DOUBLES;
    if (st[ST_DONE] != 0.0) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) v += st[ST_PART + i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        st[ST_PAP] = v;
        if (v == 0.0 || !(v == v)) {
            st[ST_DONE] = 1.0;
            st[ST_ALPHA] = 0.0;
        } else {
            st[ST_ALPHA] = st[ST_RZ] / v;
        }
    }
}

// x += alpha p ; r -= alpha Ap ; partial ||r||^2 (the preconditioned form
// takes Re<r, z> after the preconditioner instead: want_norm = 0)
__global__ void __launch_bounds__(kRedThreads) cg_update_kernel(float2* x, float2* r, const float2* p,
                                                                const float2* Ap, int64_t n, double* st,
                                                                int want_norm) {
    {
        const int64_t sy = blockIdx.y;
        x += sy * n; r += sy * n; p += sy * n; Ap += sy * n; st += sy * HASTY_CG_STATE_DOUBLES;
    }
    if (st[ST_DONE] != 0.0) return;
    __shared__ double sh[kRedThreads];
    const double alpha = st[ST_ALPHA];
    const float af = float(alpha);
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        float2 xi = x[i], ri = r[i];
        const float2 pi = p[i], api = Ap[i];
        xi.x += af * pi.x; xi.y += af * pi.y;
        ri.x -= af * api.x; ri.y -= af * api.y;
        x[i] = xi;
        r[i] = ri;
        acc += double(ri.x) * double(ri.x) + double(ri.y) * double(ri.y);
    }
    if (!want_norm) return;
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) st[ST_PART + kRedBlocks + blockIdx.x] = acc;
}

// Preconditioned direction (cg.cppm:45-56): rz' = Re<r, z> from the Re
// partials of dot_partial_kernel; beta = 0 on a restart (p = z), else rz'/rz.
__global__ void pcg_beta_kernel(double* st, int restart) {
    __shared__ double sh[kRedThreads];
    if (st[ST_DONE] != 0.0) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) v += st[ST_PART + i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        st[ST_RZNEW] = v;
        st[ST_BETA] = (restart || st[ST_RZ] == 0.0) ? 0.0 : v / st[ST_RZ];
        st[ST_RZ] = v;
        st[ST_ITER] += 1.0;
    }
}

// rz' = sum ; beta = rz'/rz ; rz = rz' ; done = sqrt(rz) < tol
__global__ void cg_beta_kernel(double* st) {
    __shared__ double sh[kRedThreads];
    st += int64_t(blockIdx.y) * HASTY_CG_STATE_DOUBLES;
    if (st[ST_DONE] != 0.0) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) v += st[ST_PART + kRedBlocks + i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        st[ST_RZNEW] = v;
        st[ST_BETA] = st[ST_RZ] != 0.0 ? v / st[ST_RZ] : 0.0;
        st[ST_RZ] = v;
        st[ST_ITER] += 1.0;
        // p still needs this iteration's update: signal convergence through
        // ST_DONE only after the p kernel has read ST_BETA (cg_finish_kernel)
    }
}

// p = r + beta p, then mark convergence
__global__ void __launch_bounds__(kRedThreads) cg_p_kernel(float2* p, const float2* r, int64_t n,
                                                           const double* st) {
    p += int64_t(blockIdx.y) * n; r += int64_t(blockIdx.y) * n; st += int64_t(blockIdx.y) * HASTY_CG_STATE_DOUBLES;
    if (st[ST_DONE] != 0.0) return;
    const float beta = float(st[ST_BETA]);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        float2 pi = p[i];
        const float2 ri = r[i];
        pi.x = ri.x + beta * pi.x;
        pi.y = ri.y + beta * pi.y;
        p[i] = pi;
    }
}

__global__ void cg_finish_kernel(double* st) {
    st += int64_t(blockIdx.y) * HASTY_CG_STATE_DOUBLES;
    if (st[ST_DONE] != 0.0) return;
    if (sqrt(st[ST_RZ]) < st[ST_TOL]) st[ST_DONE] = 1.0;
}

__global__ void __launch_bounds__(kRedThreads) copy_norm_kernel(const float2* r, float2* p, int64_t n,
                                                                double* st) {
    __shared__ double sh[kRedThreads];
    r += int64_t(blockIdx.y) * n; p += int64_t(blockIdx.y) * n; st += int64_t(blockIdx.y) * HASTY_CG_STATE_DOUBLES;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float2 ri = r[i];
        p[i] = ri;
        acc += double(ri.x) * double(ri.x) + double(ri.y) * double(ri.y);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) st[ST_PART + kRedBlocks + blockIdx.x] = acc;
}

// p <- z ; partial Re<r, z> (the preconditioned start, cg.cppm:81-85)
__global__ void __launch_bounds__(kRedThreads) copy_dot_kernel(const float2* r, const float2* z, float2* p,
                                                               int64_t n, double* st) {
    __shared__ double sh[kRedThreads];
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float2 ri = r[i], zi = z[i];
        p[i] = zi;
        acc += double(ri.x) * double(zi.x) + double(ri.y) * double(zi.y);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) st[ST_PART + kRedBlocks + blockIdx.x] = acc;
}

__global__ void cg_init_finish_kernel(double* st, double tol) {
    __shared__ double sh[kRedThreads];
    st += int64_t(blockIdx.y) * HASTY_CG_STATE_DOUBLES;
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) v += st[ST_PART + kRedBlocks + i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        st[ST_RZ] = v;
        st[ST_TOL] = tol;
        st[ST_ITER] = 0.0;
        st[ST_DONE] = sqrt(v) < tol ? 1.0 : 0.0;
    }
}

__global__ void __launch_bounds__(kRedThreads) norm_partial_kernel(const float2* x, int64_t n, double* part) {
    __shared__ double sh[kRedThreads];
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float2 v = x[i];
        acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void norm_finish_kernel(const double* part, double* out) {
    __shared__ double sh[kRedThreads];
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) v += part[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[0] = sqrt(v);
}

__global__ void scale_kernel(float2* x, int64_t n, const double* a, int invert) {
    const double s = invert ? 1.0 / a[0] : a[0];
    const float sf = float(s);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        float2 v = x[i];
        v.x *= sf; v.y *= sf;
        x[i] = v;
    }
}

// y <- y + a x (complex a in FP64, rounded to the c64 arithmetic once)
__global__ void __launch_bounds__(kRedThreads) axpy_kernel(float2* y, const float2* x, int64_t n,
                                                           float are, float aim) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const float2 v = x[i];
        float2 o = y[i];
        o.x += are * v.x - aim * v.y;
        o.y += are * v.y + aim * v.x;
        y[i] = o;
    }
}

__global__ void zero_doubles_kernel(double* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

static unsigned red_blocks(int64_t n) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, ceil_div(n, kRedThreads))));
}

// Scratch of hasty_dot_noalloc: the per-block FP64 partial sums (Re block,
// then Im block) of one dot product, 2 x kRedBlocks doubles = 8 KiB on every
// device the library runs on, held until the process ends.  Every partial a
// launch reads is written by the same launch, so nothing needs clearing.
// Host code never passes `scratch` itself (its host-side name is not a device
// address): it passes the per-device address from hipGetSymbolAddress.
constexpr int kDotScratchDoubles = 2 * kRedBlocks;
static __device__ __align__(256) double scratch[kDotScratchDoubles];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (on that device's first hasty_dot_noalloc call).
static double* dot_scratch_on_current_device() {
    int dev = -1;
    HASTY_CHECK_HIP(hipGetDevice(&dev));
    HASTY_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "device index out of range for the dot scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        HASTY_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        HASTY_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned NULL for the dot scratch");
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}

}  // namespace hasty

using namespace hasty;

extern "C" {

int hasty_dot(const void* a, const void* b, int64_t n, double* out, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(n >= 0 && out, "invalid arguments to hasty_dot");
        HASTY_REQUIRE(n == 0 || (a && b), "NULL vector");
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* part = nullptr;
        HASTY_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), 2 * kRedBlocks * sizeof(double), s));
        const unsigned nb = red_blocks(n);
        hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(kRedThreads), 0, s,
                           static_cast<const float2*>(a), static_cast<const float2*>(b), n, part,
                           static_cast<const double*>(nullptr));
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedThreads), 0, s, part, int(nb), out);
        HASTY_CHECK_HIP(hipGetLastError());
        HASTY_CHECK_HIP(hipFreeAsync(part, s));
    });
}

int hasty_dot_noalloc(const void* a, const void* b, int64_t n, double* out, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(n >= 0 && out, "invalid arguments to hasty_dot_noalloc");
        HASTY_REQUIRE(n == 0 || (a && b), "NULL vector");
        const unsigned nb = red_blocks(n);
        // the partials of nb blocks must fit the static scratch (nb <= kRedBlocks for any n)
        HASTY_REQUIRE(2 * int64_t(nb) <= kDotScratchDoubles,
                      "hasty_dot_noalloc: partial sums exceed the 8 KiB static scratch");
        double* part = dot_scratch_on_current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(kRedThreads), 0, s,
                           static_cast<const float2*>(a), static_cast<const float2*>(b), n, part,
                           static_cast<const double*>(nullptr));
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedThreads), 0, s, part, int(nb), out);
        HASTY_CHECK_HIP(hipGetLastError());
    });
}

int hasty_norm(const void* x, int64_t n, double* out, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(n >= 0 && out, "invalid arguments to hasty_norm");
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* part = nullptr;
        HASTY_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), kRedBlocks * sizeof(double), s));
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(2), dim3(256), 0, s, part, kRedBlocks);
        hipLaunchKernelGGL(norm_partial_kernel, dim3(red_blocks(n)), dim3(kRedThreads), 0, s,
                           static_cast<const float2*>(x), n, part);
        hipLaunchKernelGGL(norm_finish_kernel, dim3(1), dim3(kRedThreads), 0, s, part, out);
        HASTY_CHECK_HIP(hipGetLastError());
        HASTY_CHECK_HIP(hipFreeAsync(part, s));
    });
}

int hasty_scale_by_device(void* x, int64_t n, const double* a, int invert, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(n >= 0 && a, "invalid arguments to hasty_scale_by_device");
        hipLaunchKernelGGL(scale_kernel, dim3(red_blocks(n)), dim3(kRedThreads), 0,
                           static_cast<hipStream_t>(stream), static_cast<float2*>(x), n, a, invert);
        HASTY_CHECK_HIP(hipGetLastError());
    });
}

int hasty_axpy(double a_re, double a_im, const void* x, void* y, int64_t n, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(n >= 0 && (n == 0 || (x && y)), "invalid arguments to hasty_axpy");
        if (n == 0) return;
        hipLaunchKernelGGL(axpy_kernel, dim3(red_blocks(n)), dim3(kRedThreads), 0,
                           static_cast<hipStream_t>(stream), static_cast<float2*>(y),
                           static_cast<const float2*>(x), n, float(a_re), float(a_im));
        HASTY_CHECK_HIP(hipGetLastError());
    });
}

static void cg_init_launch(void* r, void* p, int64_t n, int nsys, double tol, double* state, hipStream_t s) {
    // unused partial slots must read as zero
    hipLaunchKernelGGL(zero_doubles_kernel, dim3(4), dim3(256), 0, s, state, nsys * HASTY_CG_STATE_DOUBLES);
    hipLaunchKernelGGL(copy_norm_kernel, dim3(red_blocks(n), unsigned(nsys)), dim3(kRedThreads), 0, s,
                       static_cast<const float2*>(r), static_cast<float2*>(p), n, state);
    hipLaunchKernelGGL(cg_init_finish_kernel, dim3(1, unsigned(nsys)), dim3(kRedThreads), 0, s, state, tol);
    HASTY_CHECK_HIP(hipGetLastError());
}

int hasty_cg_init(void* r, void* p, int64_t n, double tol, double* state, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(r && p && state && n >= 0, "invalid arguments to hasty_cg_init");
        cg_init_launch(r, p, n, 1, tol, state, static_cast<hipStream_t>(stream));
    });
}

int hasty_cg_init_batched(void* r, void* p, int64_t n, int nsys, double tol, double* state, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(r && p && state && n >= 0, "invalid arguments to hasty_cg_init_batched");
        HASTY_REQUIRE(nsys >= 1 && nsys <= 65535, "hasty_cg_init_batched: nsys must be in [1, 65535]");
        cg_init_launch(r, p, n, nsys, tol, state, static_cast<hipStream_t>(stream));
    });
}

// One CG update of nsys independent systems (blockIdx.y = system).
static void cg_step_launch(void* x, void* r, void* p, const void* Ap, int64_t n, int nsys, double* state,
                           hipStream_t s) {
    const unsigned nb = red_blocks(n);
    const unsigned ny = unsigned(nsys);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb, ny), dim3(kRedThreads), 0, s,
                       static_cast<const float2*>(p), static_cast<const float2*>(Ap), n,
                       state + ST_PART, static_cast<const double*>(state + ST_DONE));
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1, ny), dim3(kRedThreads), 0, s, state);
    hipLaunchKernelGGL(cg_update_kernel, dim3(nb, ny), dim3(kRedThreads), 0, s,
                       static_cast<float2*>(x), static_cast<float2*>(r),
                       static_cast<const float2*>(p), static_cast<const float2*>(Ap), n, state, 1);
    hipLaunchKernelGGL(cg_beta_kernel, dim3(1, ny), dim3(kRedThreads), 0, s, state);
    hipLaunchKernelGGL(cg_p_kernel, dim3(nb, ny), dim3(kRedThreads), 0, s, static_cast<float2*>(p),
                       static_cast<const float2*>(r), n, static_cast<const double*>(state));
    hipLaunchKernelGGL(cg_finish_kernel, dim3(1, ny), dim3(1), 0, s, state);
    HASTY_CHECK_HIP(hipGetLastError());
}

int hasty_cg_step(void* x, void* r, void* p, const void* Ap, int64_t n, double* state,
                  void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(x && r && p && Ap && state && n >= 0, "invalid arguments to hasty_cg_step");
        cg_step_launch(x, r, p, Ap, n, 1, state, static_cast<hipStream_t>(stream));
    });
}

int hasty_cg_step_batched(void* x, void* r, void* p, const void* Ap, int64_t n, int nsys, double* state,
                          void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(x && r && p && Ap && state && n >= 0, "invalid arguments to hasty_cg_step_batched");
        HASTY_REQUIRE(nsys >= 1 && nsys <= 65535, "hasty_cg_step_batched: nsys must be in [1, 65535]");
        cg_step_launch(x, r, p, Ap, n, nsys, state, static_cast<hipStream_t>(stream));
    });
}

int hasty_pcg_init(const void* r, const void* z, void* p, int64_t n, double tol, double* state,
                   void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(r && z && p && state && n >= 0, "invalid arguments to hasty_pcg_init");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(4), dim3(256), 0, s, state, HASTY_CG_STATE_DOUBLES);
        hipLaunchKernelGGL(copy_dot_kernel, dim3(red_blocks(n)), dim3(kRedThreads), 0, s,
                           static_cast<const float2*>(r), static_cast<const float2*>(z), static_cast<float2*>(p),
                           n, state);
        hipLaunchKernelGGL(cg_init_finish_kernel, dim3(1), dim3(kRedThreads), 0, s, state, tol);
        HASTY_CHECK_HIP(hipGetLastError());
    });
}

int hasty_pcg_update(void* x, void* r, const void* p, const void* Ap, int64_t n, double* state,
                     void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(x && r && p && Ap && state && n >= 0, "invalid arguments to hasty_pcg_update");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const unsigned nb = red_blocks(n);
        hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(kRedThreads), 0, s,
                           static_cast<const float2*>(p), static_cast<const float2*>(Ap), n,
                           state + ST_PART, static_cast<const double*>(state + ST_DONE));
        hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kRedThreads), 0, s, state);
        hipLaunchKernelGGL(cg_update_kernel, dim3(nb), dim3(kRedThreads), 0, s,
                           static_cast<float2*>(x), static_cast<float2*>(r),
                           static_cast<const float2*>(p), static_cast<const float2*>(Ap), n, state, 0);
        HASTY_CHECK_HIP(hipGetLastError());
    });
}

int hasty_pcg_direction(const void* r, const void* z, void* p, int64_t n, int restart, double* state,
                        void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(r && z && p && state && n >= 0, "invalid arguments to hasty_pcg_direction");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const unsigned nb = red_blocks(n);
        hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(kRedThreads), 0, s,
                           static_cast<const float2*>(r), static_cast<const float2*>(z), n,
                           state + ST_PART, static_cast<const double*>(state + ST_DONE));
        hipLaunchKernelGGL(pcg_beta_kernel, dim3(1), dim3(kRedThreads), 0, s, state, restart ? 1 : 0);
        hipLaunchKernelGGL(cg_p_kernel, dim3(nb), dim3(kRedThreads), 0, s, static_cast<float2*>(p),
                           static_cast<const float2*>(z), n, static_cast<const double*>(state));
        hipLaunchKernelGGL(cg_finish_kernel, dim3(1), dim3(1), 0, s, state);
        HASTY_CHECK_HIP(hipGetLastError());
    });
}

int hasty_cg_status(const double* state, double* resid, int* done, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(state != nullptr, "state is NULL");
        double h[8];
        hipStream_t s = static_cast<hipStream_t>(stream);
        HASTY_CHECK_HIP(hipMemcpyAsync(h, state, sizeof(h), hipMemcpyDeviceToHost, s));
        HASTY_CHECK_HIP(hipStreamSynchronize(s));
        if (resid) *resid = std::sqrt(h[ST_RZ]);
        if (done) *done = h[ST_DONE] != 0.0;
    });
}

}  // extern "C"
