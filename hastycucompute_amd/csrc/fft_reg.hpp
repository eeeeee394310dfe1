// Register-resident two-level FFT pass for gfx950.
//
// A line of length L = L1 * L2 is transformed by max(L1, L2) threads with
// one LDS exchange (Cooley-Tukey, p = L2 p1 + p2, k = k1 + L1 k2):
//   step 1  thread p2 holds x[L2 p1 + p2] (p1 < L1), does an L1-point DFT in
//           registers and multiplies by W_L^{p2 k1};
//   exchange through LDS;
//   step 2  thread k1 holds Y[p2][k1] (p2 < L2), does an L2-point DFT:
//           X[k1 + L1 k2] = sum_p2 W_L2^{p2 k2} Y[p2][k1].
// After step 2 thread k1 owns the stride-L1 positions k1 + L1 k2, which is
// exactly the step-1 layout of the factorisation L = L2 * L1, so a kernel
// product followed by the inverse transform (Toeplitz z-pass) continues from
// registers with a second exchange.
//
// The small DFTs are unrolled at compile time: radix 2/3/4/5/8 butterflies
// and composite sizes by an in-register Cooley-Tukey split with constant
// twiddles (cos / sin evaluated by constexpr Taylor series on [-pi, pi]).
#pragma once

#include "fft.hpp"

namespace hasty {

// ---------------------------------------------------------------------------
// constexpr trigonometry for compile-time twiddles
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_taylor(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 40; ++k) {
        term *= -x * x / double((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_taylor(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 40; ++k) {
        term *= -x * x / double((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
// cos / sin of 2 pi m / n for integers 0 <= m < n, reduced to [-pi, pi]
constexpr double cx_cos2pi(int m, int n) {
    m %= n;
    if (4 * m == n || 4 * m == 3 * n) return 0.0;
    if (2 * m == n) return -1.0;
    if (m == 0) return 1.0;
    double x = 2.0 * kPi * double(m) / double(n);
    if (x > kPi) x -= 2.0 * kPi;
    return cx_cos_taylor(x);
}
constexpr double cx_sin2pi(int m, int n) {
    m %= n;
    if (m == 0 || 2 * m == n) return 0.0;
    if (4 * m == n) return 1.0;
    if (4 * m == 3 * n) return -1.0;
    double x = 2.0 * kPi * double(m) / double(n);
    if (x > kPi) x -= 2.0 * kPi;
    return cx_sin_taylor(x);
}

// Twiddle table of an N-point DFT, built by the compiler: unrolled loops index
// it with constants, so every twiddle becomes an immediate.
template <int N> struct CxTwiddles {
    double c[N];
    double s[N];
    constexpr CxTwiddles() : c(), s() {
        for (int m = 0; m < N; ++m) {
            c[m] = cx_cos2pi(m, N);
            s[m] = cx_sin2pi(m, N);
        }
    }
};
template <int N> struct CxTw {
    static constexpr CxTwiddles<N> table{};
};

constexpr int cx_first_factor(int n) {
    // split n = a * b for the in-register Cooley-Tukey (a: the larger
    // balanced factor from {8, 5, 4, 3, 2})
    if (n % 4 == 0 && n / 4 >= 4 && n != 16 && n % 8 == 0) return 8;
    if (n % 4 == 0) return 4;
    if (n % 5 == 0) return 5;
    if (n % 3 == 0) return 3;
    return 2;
}

template <int N, int SIGN, class V> struct RegDFT;

template <int SIGN, class V> struct RegDFT<1, SIGN, V> {
    __device__ static inline void run(V*) {}
};
template <int SIGN, class V> struct RegDFT<2, SIGN, V> {
    __device__ static inline void run(V* v) { Butterfly<2, V>::run(v, SIGN); }
};
template <int SIGN, class V> struct RegDFT<3, SIGN, V> {
    __device__ static inline void run(V* v) { Butterfly<3, V>::run(v, SIGN); }
};
template <int SIGN, class V> struct RegDFT<4, SIGN, V> {
    __device__ static inline void run(V* v) { Butterfly<4, V>::run(v, SIGN); }
};
template <int SIGN, class V> struct RegDFT<5, SIGN, V> {
    __device__ static inline void run(V* v) { Butterfly<5, V>::run(v, SIGN); }
};
template <int SIGN, class V> struct RegDFT<8, SIGN, V> {
    __device__ static inline void run(V* v) { Butterfly<8, V>::run(v, SIGN); }
};

// Composite N = A * B: X[k1 + A k2] = sum_p2 W_B^{p2 k2} W_N^{p2 k1} sum_p1 x[B p1 + p2] W_A^{p1 k1}
template <int N, int SIGN, class V> struct RegDFT {
    static constexpr int A = cx_first_factor(N);
    static constexpr int B = N / A;
    static_assert(A * B == N && A > 1 && B > 1, "composite DFT size");
    __device__ static inline void run(V* v) {
        using R = decltype(V::x);
        V y[N];
#pragma unroll
        for (int p2 = 0; p2 < B; ++p2) {
            V t[A];
#pragma unroll
            for (int p1 = 0; p1 < A; ++p1) t[p1] = v[B * p1 + p2];
            RegDFT<A, SIGN, V>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                if (p2 * k1 == 0) {
                    y[p2 * A + k1] = t[k1];
                } else {
                    constexpr auto& tw = CxTw<N>::table;
                    const R c = R(tw.c[(p2 * k1) % N]);
                    const R s = R(SIGN * tw.s[(p2 * k1) % N]);
                    y[p2 * A + k1] = cmake<V>(t[k1].x * c - t[k1].y * s, t[k1].x * s + t[k1].y * c);
                }
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            V t[B];
#pragma unroll
            for (int p2 = 0; p2 < B; ++p2) t[p2] = y[p2 * A + k1];
            RegDFT<B, SIGN, V>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
        }
    }
};

// ---------------------------------------------------------------------------
// Unified pass descriptor
// ---------------------------------------------------------------------------
// Lines are indexed (b, o, i); element p of line (b, o, i) of the input is
// at in[b*in_b + o*in_o + i*in_i + j*in_s] with j = (p - in_shift) mod L
// (zero when j >= in_len).  Output element j = (q - out_shift) mod L of line
// position q is stored when j < out_len.  ROWS passes (lines contiguous,
// in_s = out_s = 1) may carry diagonal operands, deapodisation, a coil sum
// and accumulation; any pass may carry a kernel product followed by the
// inverse transform.
template <class V> struct PassArgs {
    using R = decltype(V::x);
    const V* in; V* out;
    int64_t in_b, in_o, in_i, in_s;
    int64_t out_b, out_o, out_i, out_s;
    int ni, no, nbatch, tx;
    int in_len, in_shift, out_len, out_shift;
    R scale;
    DiagOp<V> lm1, lm2;                 // load: element (b, i, j) at ptr[b*sb + i*sr + j]
    DiagOp<V> sm1, sm2, sio;            // store
    const R* dz; const R* dy; const R* dx;
    int ny;                             // line i -> (z = i / ny, y = i % ny)
    int dz_on_load;                     // deapodisation applied on load (1) or store (0)
    const V* psf; int64_t psf_o, psf_i, psf_s;
    int bchunk;                         // PSF passes: batch elements per block (kernel read once)
    int accumulate, sum_batch;
    const V* tw;                        // tw[k] = exp(-2 pi i k / L)
    int ldp;                            // LDS stride per line (complex elements)
};

template <int L1, int L2> struct RegShape {
    static constexpr int L = L1 * L2;
    static constexpr int T = L1 > L2 ? L1 : L2;
};

template <class V, int SIGN>
__device__ inline V twiddle_mul(V v, const V* tw, int idx) {
    V w = tw[idx];
    if (SIGN > 0) w.y = -w.y;
    return cmul(v, w);
}

// Pass modes (compile time)
enum PassMode : int {
    PM_COL = 0,        // strided lines, plain load / store
    PM_COL_PSF = 1,    // strided lines, forward FFT, kernel product, inverse FFT
    PM_ROW_LOAD = 2,   // contiguous lines, diagonals / deapodisation on load
    PM_ROW_STORE = 3,  // contiguous lines, diagonals / deapodisation / sums on store
};

template <class V, int L1, int L2, int SIGN, int MODE>
__global__ void __launch_bounds__(512) fft2_pass_kernel(PassArgs<V> a) {
    using R = decltype(V::x);
    constexpr int L = L1 * L2;
    constexpr int T = RegShape<L1, L2>::T;
    constexpr bool ROWS = MODE == PM_ROW_LOAD || MODE == PM_ROW_STORE;
    constexpr bool PSF = MODE == PM_COL_PSF;
    // output layout: stride L1 (L2 values per owner) after a forward pass,
    // stride L2 (L1 values per owner) after the PSF round trip
    constexpr int NOWN = PSF ? L1 : L2;
    constexpr int OSTRIDE = PSF ? L2 : L1;
    constexpr int NOWNERS = PSF ? L2 : L1;
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const int TX = a.tx;
    const int tid = threadIdx.x;
    const int t = ROWS ? (tid % T) : (tid / TX);
    const int c = ROWS ? (tid / T) : (tid % TX);
    const int ntile = (a.ni + TX - 1) / TX;
    int64_t bid = blockIdx.x;
    const int it = int(bid % ntile); bid /= ntile;
    const int o = int(bid % a.no); bid /= a.no;
    const int64_t bblk = bid;
    const int i = it * TX + c;
    const bool line_ok = i < a.ni;
    V* xl = lds + c * a.ldp;   // this line's exchange area
    // twiddle table W_L^k, k < L, staged once per block after the exchange areas
    V* twl = lds + int64_t(TX) * a.ldp;
    for (int k = tid; k < L; k += blockDim.x) twl[k] = a.tw[k];
    __syncthreads();

    const bool sum_b = MODE == PM_ROW_STORE && a.sum_batch;
    // a PSF block runs bchunk consecutive batch elements of its lines, so
    // each kernel value is read from HBM once per chunk, not once per element
    const int64_t b_first = PSF ? bblk * a.bchunk : bblk;
    const int nb_loop = sum_b ? a.nbatch
                              : (PSF ? int(min(int64_t(a.bchunk), int64_t(a.nbatch) - b_first)) : 1);
    // the kernel column of this thread's positions, the same for every
    // batch element of the chunk: loaded once, issued together
    V kv[PSF ? L2 : 1];
    if constexpr (PSF) {
        if (t < L1 && line_ok) {
            const V* k = a.psf + int64_t(o) * a.psf_o + int64_t(i) * a.psf_i;
#pragma unroll
            for (int k2 = 0; k2 < L2; ++k2) kv[k2] = k[int64_t(t + L1 * k2) * a.psf_s];
        }
    }
    R zy = R(1);
    if (ROWS && line_ok && a.dx) {
        const int z = i / a.ny;
        zy = a.dz[z] * a.dy[i - z * a.ny];
    }
    // coil-sum accumulators: outputs j = t + T m, m < NACC (the coil sum is
    // used with a crop to at most L/2 outputs; the host checks out_len)
    constexpr int NACC = (L / 2 + T - 1) / T;
    V acc[MODE == PM_ROW_STORE ? NACC : 1];
    if constexpr (MODE == PM_ROW_STORE) {
#pragma unroll
        for (int m = 0; m < NACC; ++m) acc[m] = cmake<V>(R(0), R(0));
    }
    for (int bb = 0; bb < nb_loop; ++bb) {
        const int64_t b = sum_b ? bb : b_first + bb;
        // per-iteration copies the compiler cannot hoist: otherwise the
        // per-position addresses and masks of every load and store stay live
        // in registers across the batch loop (and halve the occupancy)
        int64_t in_s = a.in_s, out_s = a.out_s;
        int in_shift = a.in_shift, in_len = a.in_len, out_shift = a.out_shift, out_len = a.out_len;
        if (nb_loop > 1)
            asm volatile("" : "+s"(in_s), "+s"(out_s), "+s"(in_shift), "+s"(in_len), "+s"(out_shift),
                         "+s"(out_len));
        V v[T];
        // ---- load (step-1 layout: thread t = p2 holds p = L2 p1 + t) ----
        if (t < L2) {
            const V* src = a.in + b * a.in_b + int64_t(o) * a.in_o + int64_t(i) * a.in_i;
#pragma unroll
            for (int p1 = 0; p1 < L1; ++p1) {
                const int p = L2 * p1 + t;
                int j = p - in_shift;
                j = j < 0 ? j + L : (j >= L ? j - L : j);
                V val = cmake<V>(R(0), R(0));
                if (line_ok && j < in_len) {
                    val = src[int64_t(j) * in_s];
                    if constexpr (MODE == PM_ROW_LOAD) {
                        if (a.lm1.ptr) val = diag_mult(val, a.lm1.ptr[b * a.lm1.sb + int64_t(i) * a.lm1.sr + j], a.lm1.type);
                        if (a.lm2.ptr) val = diag_mult(val, a.lm2.ptr[b * a.lm2.sb + int64_t(i) * a.lm2.sr + j], a.lm2.type);
                        if (a.dx) val = cscale(val, zy * a.dx[j]);
                    }
                }
                v[p1] = val;
            }
            RegDFT<L1, SIGN, V>::run(v);
#pragma unroll
            for (int k1 = 1; k1 < L1; ++k1) v[k1] = twiddle_mul<V, SIGN>(v[k1], twl, t * k1);
#pragma unroll
            for (int k1 = 0; k1 < L1; ++k1) xl[t * (L1 + 1) + k1] = v[k1];
        }
        __syncthreads();
        // ---- step 2: thread t = k1 -> positions t + L1 k2 ----
        if (t < L1) {
#pragma unroll
            for (int p2 = 0; p2 < L2; ++p2) v[p2] = xl[p2 * (L1 + 1) + t];
            RegDFT<L2, SIGN, V>::run(v);
        }
        if constexpr (PSF) {
            if (t < L1) {
                if (line_ok) {
#pragma unroll
                    for (int k2 = 0; k2 < L2; ++k2) v[k2] = cmul(v[k2], kv[k2]);
                }
                // inverse, factorisation L = L2 * L1: step 1 over the L2 held values
                RegDFT<L2, -SIGN, V>::run(v);
#pragma unroll
                for (int k1 = 1; k1 < L2; ++k1) v[k1] = twiddle_mul<V, -SIGN>(v[k1], twl, t * k1);
            }
            __syncthreads();   // step-2 reads of the first exchange are done
            if (t < L1) {
#pragma unroll
                for (int k1 = 0; k1 < L2; ++k1) xl[t * (L2 + 1) + k1] = v[k1];
            }
            __syncthreads();
            if (t < L2) {
#pragma unroll
                for (int p2 = 0; p2 < L1; ++p2) v[p2] = xl[p2 * (L2 + 1) + t];
                RegDFT<L1, -SIGN, V>::run(v);
            }
        }
        // ---- store ----
        if constexpr (MODE == PM_ROW_STORE) {
            // stage the line in natural order, then each thread stores the
            // consecutive outputs j = t + T m (coalesced, no register arrays)
            __syncthreads();
            if (t < L1) {
#pragma unroll
                for (int k2 = 0; k2 < L2; ++k2) xl[t + L1 * k2] = v[k2];
            }
            __syncthreads();
            if (line_ok) {
                V* dst = a.out + (sum_b ? 0 : b * a.out_b) + int64_t(o) * a.out_o + int64_t(i) * a.out_i;
                const V* m1 = a.sm1.ptr ? a.sm1.ptr + b * a.sm1.sb + int64_t(i) * a.sm1.sr : nullptr;
                const V* m2 = a.sm2.ptr ? a.sm2.ptr + b * a.sm2.sb + int64_t(i) * a.sm2.sr : nullptr;
                const V* io = a.sio.ptr ? a.sio.ptr + b * a.sio.sb + int64_t(i) * a.sio.sr : nullptr;
                if (sum_b) {
#pragma unroll
                    for (int m = 0; m < NACC; ++m) {
                        const int j = t + T * m;
                        if (j < out_len) {
                            int q = j + out_shift;
                            q = q >= L ? q - L : (q < 0 ? q + L : q);
                            V val = cscale(xl[q], a.scale);
                            if (m1) val = diag_mult(val, m1[j], a.sm1.type);
                            if (m2) val = diag_mult(val, m2[j], a.sm2.type);
                            if (io) val = diag_mult(val, io[j], a.sio.type);
                            if (a.dx) val = cscale(val, zy * a.dx[j]);
                            acc[m] = cadd(acc[m], val);
                        }
                    }
                } else {
                    for (int j = t; j < out_len; j += T) {
                        int q = j + out_shift;
                        q = q >= L ? q - L : (q < 0 ? q + L : q);
                        V val = cscale(xl[q], a.scale);
                        if (m1) val = diag_mult(val, m1[j], a.sm1.type);
                        if (m2) val = diag_mult(val, m2[j], a.sm2.type);
                        if (io) val = diag_mult(val, io[j], a.sio.type);
                        if (a.dx) val = cscale(val, zy * a.dx[j]);
                        V* d = dst + j;
                        if (a.accumulate) val = cadd(val, *d);
                        *d = val;
                    }
                }
            }
        } else {
            // owner t holds positions t + OSTRIDE k, k < NOWN
            if (t < NOWNERS && line_ok) {
                V* dst = a.out + b * a.out_b + int64_t(o) * a.out_o + int64_t(i) * a.out_i;
#pragma unroll
                for (int k = 0; k < NOWN; ++k) {
                    const int q = t + OSTRIDE * k;
                    int j = q - out_shift;
                    j = j < 0 ? j + L : (j >= L ? j - L : j);
                    if (j < out_len) dst[int64_t(j) * out_s] = cscale(v[k], a.scale);
                }
            }
        }
        if (nb_loop > 1) __syncthreads();   // exchange area reused by the next b
    }
    if constexpr (MODE == PM_ROW_STORE) {
        if (sum_b && line_ok) {
            V* dst = a.out + int64_t(o) * a.out_o + int64_t(i) * a.out_i;
#pragma unroll
            for (int m = 0; m < NACC; ++m) {
                const int j = t + T * m;
                if (j < a.out_len) {
                    V val = acc[m];
                    if (a.accumulate) val = cadd(val, dst[j]);
                    dst[j] = val;
                }
            }
        }
    }
}

// Host dispatch: returns false when L has no register factorisation here
// (the caller then uses the LDS Stockham kernels).
template <class V> bool launch_fft2_pass(PassArgs<V> a, int L, int sign, int mode,
                                         const char* tag, hipStream_t s);

}  // namespace hasty
