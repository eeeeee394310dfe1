// Shared helpers for the gfx950 NUFFT / Toeplitz / CG library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/hasty_hip.h"

namespace hasty {

// ---------------------------------------------------------------------------
// Errors: every C-ABI entry point catches hasty::Error and returns its code,
// with the message kept in a thread-local buffer (hasty_last_error()).
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define HASTY_CHECK_HIP(expr)                                                             \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            throw ::hasty::Error(HASTY_ERR_HIP, std::string(#expr) + ": " +               \
                                                    hipGetErrorString(_e));               \
    } while (0)

#define HASTY_REQUIRE(cond, msg)                                                          \
    do {                                                                                  \
        if (!(cond)) throw ::hasty::Error(HASTY_ERR_INVALID, (msg));                      \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return HASTY_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return HASTY_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return HASTY_ERR_INTERNAL;
    }
}

// ---------------------------------------------------------------------------
// Complex arithmetic on float2 / double2 (interleaved re, im as in c64/c128).
// ---------------------------------------------------------------------------
template <class R> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <class R> using cplx = typename vec2<R>::type;

template <class V> __host__ __device__ inline V cmake(decltype(V::x) re, decltype(V::x) im) {
    V v; v.x = re; v.y = im; return v;
}
template <class V> __host__ __device__ inline V cadd(V a, V b) { return cmake<V>(a.x + b.x, a.y + b.y); }
template <class V> __host__ __device__ inline V csub(V a, V b) { return cmake<V>(a.x - b.x, a.y - b.y); }
template <class V> __host__ __device__ inline V cmul(V a, V b) {
    return cmake<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <class V> __host__ __device__ inline V cmulc(V a, V b) {
    return cmake<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <class V> __host__ __device__ inline V cscale(V a, decltype(V::x) s) {
    return cmake<V>(a.x * s, a.y * s);
}
template <class V> __host__ __device__ inline V cconj(V a) { return cmake<V>(a.x, -a.y); }

// Diagonal-operand rule of the reference's Toeplitz load/store:
// 0 = skip, 1 = a*b, 2 = a*conj(b)   (toeplitz_load_3D.cu:2-15)
template <class V> __device__ inline V diag_mult(V a, V b, int type) {
    if (type == HASTY_MULT) return cmul(a, b);
    if (type == HASTY_MULT_CONJ) return cmulc(a, b);
    return a;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Opt-in launch timing (hasty_profile_enable / hasty_profile_read): when
// enabled on the calling thread, each named scope records a HIP event pair
// on the launch stream around the kernel(s) it encloses.
// ---------------------------------------------------------------------------
bool profiling_enabled();
void profile_record(const char* name, hipEvent_t start, hipEvent_t stop);

struct ProfScope {
    const char* name;
    hipStream_t stream;
    hipEvent_t start = nullptr, stop = nullptr;
    ProfScope(const char* n, hipStream_t s) : name(n), stream(s) {
        if (profiling_enabled()) {
            if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess ||
                hipEventRecord(start, stream) != hipSuccess) {
                start = stop = nullptr;
            }
        }
    }
    ~ProfScope() {
        if (start) {
            if (hipEventRecord(stop, stream) == hipSuccess) profile_record(name, start, stop);
        }
    }
};
#define HASTY_PROF_CAT2(a, b) a##b
#define HASTY_PROF_CAT(a, b) HASTY_PROF_CAT2(a, b)
#define HASTY_PROF(name, stream) ::hasty::ProfScope HASTY_PROF_CAT(_prof_, __LINE__)(name, stream)

// Current device of the calling thread.
inline int current_device() {
    int d = 0;
    HASTY_CHECK_HIP(hipGetDevice(&d));
    return d;
}

}  // namespace hasty
