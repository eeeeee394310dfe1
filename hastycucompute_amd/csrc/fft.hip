// Host side of the Stockham passes: radix plans, twiddle cache, launches.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft.hpp"
#include "fft_reg.hpp"

#include <cstdlib>

namespace hasty {

bool is_smooth(int64_t n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

int64_t next_smooth(int64_t n) {
    int64_t m = n < 1 ? 1 : n;
    while (!is_smooth(m)) ++m;
    return m;
}

FFTPlan1D make_fft_plan(int L) {
    HASTY_REQUIRE(L >= 1 && is_smooth(L), "FFT length " + std::to_string(L) + " is not 2^a 3^b 5^c");
    FFTPlan1D p{};
    p.L = L;
    p.nstages = 0;
    int m = L;
    // radix 8 first, then 4, 2 (at most one), 3, 5
    while (m % 8 == 0) { p.radix[p.nstages++] = 8; m /= 8; }
    while (m % 4 == 0) { p.radix[p.nstages++] = 4; m /= 4; }
    while (m % 2 == 0) { p.radix[p.nstages++] = 2; m /= 2; }
    while (m % 3 == 0) { p.radix[p.nstages++] = 3; m /= 3; }
    while (m % 5 == 0) { p.radix[p.nstages++] = 5; m /= 5; }
    HASTY_REQUIRE(p.nstages <= kMaxStages, "FFT length has too many factors");
    return p;
}

namespace {
std::mutex g_tw_mutex;
std::map<std::tuple<int, int, bool>, void*> g_tw;   // (device, L, f64) -> device table
}  // namespace

const void* twiddles(int L, bool f64) {
    const int dev = current_device();
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto key = std::make_tuple(dev, L, f64);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return it->second;
    const size_t esz = f64 ? 16 : 8;
    std::vector<unsigned char> host(esz * size_t(L));
    for (int k = 0; k < L; ++k) {
        // exact reduction of the angle to [0, pi/4] octants is not needed:
        // long double sin/cos of 2 pi k / L are accurate to < 1 ulp of double
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)L;
        const long double c = cosl(ang), s = sinl(ang);
        if (f64) {
            double v[2] = {(double)c, (double)s};
            std::memcpy(&host[esz * k], v, esz);
        } else {
            float v[2] = {(float)c, (float)s};
            std::memcpy(&host[esz * k], v, esz);
        }
    }
    void* d = nullptr;
    HASTY_CHECK_HIP(hipMalloc(&d, host.size()));
    HASTY_CHECK_HIP(hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice));
    g_tw[key] = d;
    return d;
}

// Dynamic LDS above 64 KiB must be enabled per kernel; done once per
// instantiation (the attribute is a property of the function).
template <class K>
void allow_lds(K kernel, size_t bytes) {
    static std::once_flag once;
    if (bytes <= 64 * 1024) return;
    std::call_once(once, [&] {
        HASTY_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            int(kPassLdsBudget)));
    });
}

// The register two-level kernels handle the common lengths; the LDS
// Stockham kernels cover every other 2^a 3^b 5^c length.  HASTY_FFT_STOCKHAM=1
// forces the Stockham kernels (used by the tests to cover both).
static bool force_stockham() {
    static const bool v = [] {
        const char* e = std::getenv("HASTY_FFT_STOCKHAM");
        return e && e[0] == '1';
    }();
    return v;
}
bool fft_force_stockham() { return force_stockham(); }

template <class V>
static PassArgs<V> blank_pass() {
    PassArgs<V> p{};
    p.scale = 1;
    p.ny = 1;
    return p;
}

template <class V>
void launch_row_load(RowLoadArgs<V> a, hipStream_t s) {
    if (!force_stockham()) {
        PassArgs<V> p = blank_pass<V>();
        p.in = a.in; p.out = a.out;
        p.in_b = a.in_b; p.in_o = 0; p.in_i = a.in_r; p.in_s = 1;
        p.out_b = a.out_b; p.out_o = 0; p.out_i = a.out_r; p.out_s = 1;
        p.ni = int(a.nrows); p.no = 1; p.nbatch = a.nbatch;
        p.in_len = a.in_len; p.in_shift = a.in_shift; p.out_len = a.L; p.out_shift = 0;
        p.lm1 = a.m1; p.lm2 = a.m2;
        p.dz = a.dz; p.dy = a.dy; p.dx = a.dx; p.ny = a.ny; p.dz_on_load = 1;
        p.tw = a.tw;
        p.perm4 = a.perm4;
        if (launch_fft2_pass<V>(p, a.L, a.sign, PM_ROW_LOAD, a.tag ? a.tag : "row_load", s)) return;
    }
    HASTY_REQUIRE(a.perm4 == 0, "the class-major row layout needs a register FFT pass");
    const int64_t lines = a.nrows * a.nbatch;
    if (lines == 0) return;
    const size_t bytes = size_t(2) * a.nl * a.LP * sizeof(V);
    const int64_t blocks = ceil_div(lines, a.nl);
    allow_lds(row_load_kernel<V>, bytes);
    HASTY_PROF(a.tag ? a.tag : "row_load", s);
    hipLaunchKernelGGL(row_load_kernel<V>, dim3(unsigned(blocks)), dim3(kFFTThreads), bytes, s, a);
    HASTY_CHECK_HIP(hipGetLastError());
}

template <class V>
void launch_row_store(RowStoreArgs<V> a, hipStream_t s) {
    if (!force_stockham()) {
        PassArgs<V> p = blank_pass<V>();
        p.in = a.in; p.out = a.out;
        p.in_b = a.in_b; p.in_o = 0; p.in_i = a.in_r; p.in_s = 1;
        p.out_b = a.out_b; p.out_o = 0; p.out_i = a.out_r; p.out_s = 1;
        p.ni = int(a.nrows); p.no = 1; p.nbatch = a.nbatch;
        p.in_len = a.L; p.in_shift = 0; p.out_len = a.out_len; p.out_shift = a.out_shift;
        p.scale = a.scale;
        p.sm1 = a.m1; p.sm2 = a.m2; p.sio = a.io;
        p.dz = a.dz; p.dy = a.dy; p.dx = a.dx; p.ny = a.ny; p.dz_on_load = 0;
        p.accumulate = a.accumulate; p.sum_batch = a.sum_batch;
        p.partial = a.scratch; p.partial_elems = a.scratch ? a.scratch_elems : 0;
        p.tw = a.tw;
        if (launch_fft2_pass<V>(p, a.L, a.sign, PM_ROW_STORE, a.tag ? a.tag : "row_store", s)) return;
    }
    const int64_t units = a.sum_batch ? a.nrows : a.nrows * a.nbatch;
    if (units == 0) return;
    HASTY_REQUIRE(int64_t(a.nl) * a.out_len <= int64_t(kRowAccMax) * kFFTThreads,
                  "row_store: too many outputs per block");
    const size_t bytes = size_t(2) * a.nl * a.LP * sizeof(V);
    const int64_t blocks = ceil_div(units, a.nl);
    allow_lds(row_store_kernel<V>, bytes);
    HASTY_PROF(a.tag ? a.tag : "row_store", s);
    hipLaunchKernelGGL(row_store_kernel<V>, dim3(unsigned(blocks)), dim3(kFFTThreads), bytes, s, a);
    HASTY_CHECK_HIP(hipGetLastError());
}

template <class V>
void launch_col_pass(ColArgs<V> a, hipStream_t s) {
    if (!force_stockham()) {
        PassArgs<V> p = blank_pass<V>();
        p.in = a.in; p.out = a.out;
        p.in_b = a.in_b; p.in_o = a.in_o; p.in_i = 1; p.in_s = a.in_s;
        p.out_b = a.out_b; p.out_o = a.out_o; p.out_i = 1; p.out_s = a.out_s;
        p.ni = a.ni; p.no = a.no; p.nbatch = a.nbatch;
        p.in_len = a.in_len; p.in_shift = a.in_shift; p.out_len = a.out_len; p.out_shift = a.out_shift;
        p.scale = a.scale;
        p.psf = a.psf; p.psf_o = a.psf_o; p.psf_i = 1; p.psf_s = a.psf_s;
        p.psf_f = a.psf_f; p.psf_div = a.psf ? a.psf_div : 0;
        p.tw = a.tw;
        p.psf_q = a.psf_q;
        if (launch_fft2_pass<V>(p, a.L, a.sign, a.psf ? PM_COL_PSF : PM_COL, a.tag ? a.tag : "col_pass", s)) return;
    }
    if (a.psf && a.psf_div > 0 && a.psf_div < a.nbatch) {
        // frames on the Stockham kernels: one launch per frame's kernel
        HASTY_REQUIRE(a.nbatch % a.psf_div == 0, "col pass: batch is not a whole number of frames");
        for (int f = 0; f < a.nbatch / a.psf_div; ++f) {
            ColArgs<V> g = a;
            g.in = a.in + int64_t(f) * a.psf_div * a.in_b;
            g.out = a.out + int64_t(f) * a.psf_div * a.out_b;
            g.psf = a.psf + int64_t(f) * a.psf_f;
            g.nbatch = a.psf_div;
            g.psf_div = 0;
            launch_col_pass<V>(g, s);
        }
        return;
    }
    const int64_t blocks = ceil_div(a.ni, a.tx) * int64_t(a.no) * a.nbatch;
    if (blocks == 0) return;
    const size_t bytes = size_t(2) * a.tx * a.LP * sizeof(V);
    allow_lds(col_pass_kernel<V>, bytes);
    HASTY_PROF(a.tag ? a.tag : "col_pass", s);
    hipLaunchKernelGGL(col_pass_kernel<V>, dim3(unsigned(blocks)), dim3(kFFTThreads), bytes, s, a);
    HASTY_CHECK_HIP(hipGetLastError());
}

template void launch_row_load<float2>(RowLoadArgs<float2>, hipStream_t);
template void launch_row_load<double2>(RowLoadArgs<double2>, hipStream_t);
template void launch_row_store<float2>(RowStoreArgs<float2>, hipStream_t);
template void launch_row_store<double2>(RowStoreArgs<double2>, hipStream_t);
template void launch_col_pass<float2>(ColArgs<float2>, hipStream_t);
template void launch_col_pass<double2>(ColArgs<double2>, hipStream_t);

}  // namespace hasty
