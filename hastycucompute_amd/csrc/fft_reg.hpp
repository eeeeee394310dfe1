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
    int64_t psf_f; int psf_div;         // frames: element b uses the kernel at psf + (b / psf_div) * psf_f
    int bchunk;                         // PSF passes: batch elements per block (kernel read once)
    int accumulate, sum_batch;
    V* partial; int64_t partial_elems;  // scratch for a batch sum in block groups:
    int groups;                         //   partial[g][row][j], g < groups
    const V* tw;                        // tw[k] = exp(-2 pi i k / L)
    int ldp;                            // LDS stride per line (complex elements)
    // byte strides as 32-bit values: within one (batch, o) slab every offset
    // a thread forms fits in 32 bits (checked on the host), so addresses are
    // a wave-uniform 64-bit base plus a 32-bit per-lane offset
    uint32_t in_ib, in_sb, out_ib, out_sb, psf_ib, psf_sb;
    uint32_t lm1_rb, lm2_rb, sm1_rb, sm2_rb, sio_rb;   // diagonal row strides in bytes
    // quarter turns folded into the step-1 twiddles (set by the host with
    // the half-line specs for a line shifted by -L/4, the centred modes of
    // the NUFFT): twist_out multiplies output k by i^(k mod 4) relative to
    // the unshifted transform, twist_in input p by i^(p mod 4), each taken
    // through the pass's sign convention (see the twiddle fill)
    int twist_out, twist_in;
    // class-major x layout of the fused Toeplitz path (both 0 elsewhere):
    // perm4 (row-load passes with out_len = L, L % 4 == 0) stores output k
    // at (k mod 4) (L/4) + k/4; psf_q = ni / 4 (PSF passes) reads, for line
    // i, kernel line 4 (i mod psf_q) + i / psf_q, the frequency stored there
    int perm4, psf_q;
};

template <int L1, int L2> struct RegShape {
    static constexpr int L = L1 * L2;
    static constexpr int T = L1 > L2 ? L1 : L2;
};

// v * w (SIGN < 0) or v * conj(w) (SIGN > 0) with w read from LDS
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

// Line shapes known at compile time: the input is the first half of the
// line (in_shift 0, in_len L/2), or the output keeps positions
// [L/2 - 1, L - 1) (out_shift L/2 - 1, out_len L/2; the Toeplitz crop) or
// [0, L/2) (HALF_OUT0).  Positions that are zero or dropped then cost
// nothing: no load, no arithmetic on the zeros (the passes are built with
// -fno-signed-zeros so x + 0 folds), no DFT output that is not stored.
// A line whose half sits at shift -L/4 (the NUFFT's centred modes, wrapped
// around 0) runs as HALF_IN / HALF_OUT0 with a quarter-turn twist: a shift
// by s multiplies the other side of the transform by W^(k s), and for
// s = -L/4 that is a power of i fixed by k mod 4.
enum PassSpec : int { PS_NONE = 0, PS_HALF_IN = 1, PS_HALF_OUT = 2, PS_HALF_OUT0 = 4 };

// v * i^r
template <class V>
__device__ inline V quarter_turn(V v, int r) {
    switch (r & 3) {
        case 1: return cmake<V>(-v.y, v.x);
        case 2: return cmake<V>(-v.x, -v.y);
        case 3: return cmake<V>(v.y, -v.x);
        default: return v;
    }
}

// a * b (sgn = 1) or a * conj(b) (sgn = -1); sgn is wave-uniform
template <class V>
__device__ inline V cmul_sgn(V a, V b, decltype(V::x) sgn) {
    b.y *= sgn;
    return cmul(a, b);
}

// NT: streaming access (nontemporal hint) for the line data of the column
// passes, which is read or written exactly once per pass
template <class V, bool NT = false>
__device__ inline V load_at(const char* base, uint32_t off) {
    if constexpr (NT) {
        const decltype(V::x)* p = reinterpret_cast<const decltype(V::x)*>(base + off);
        V v;
        v.x = __builtin_nontemporal_load(p);
        v.y = __builtin_nontemporal_load(p + 1);
        return v;
    } else {
        return *reinterpret_cast<const V*>(base + off);
    }
}
template <class V, bool NT = false>
__device__ inline void store_at(char* base, uint32_t off, V v) {
    if constexpr (NT) {
        decltype(V::x)* p = reinterpret_cast<decltype(V::x)*>(base + off);
        __builtin_nontemporal_store(v.x, p);
        __builtin_nontemporal_store(v.y, p + 1);
    } else {
        *reinterpret_cast<V*>(base + off) = v;
    }
}

// v[m] (.)= d[row, j[m]] for the positions with ok[m]: all loads of the
// diagonal are issued before the first product
template <class V, int NM>
__device__ inline void apply_diag(V* v, const DiagOp<V>& d, int64_t b, uint32_t row_off, const int* j,
                                  const bool* ok) {
    if (!d.ptr) return;
    const char* base = reinterpret_cast<const char*>(d.ptr + b * d.sb);
    const decltype(V::x) sgn = d.type == HASTY_MULT_CONJ ? -1 : 1;
    V w[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (ok[m]) w[m] = load_at<V>(base, row_off + uint32_t(j[m]) * uint32_t(sizeof(V)));
#pragma unroll
    for (int m = 0; m < NM; ++m)
        if (ok[m]) v[m] = cmul_sgn(v[m], w[m], sgn);
}

template <class V, int L1, int L2, int SIGN, int MODE, int SPEC>
__global__ void __launch_bounds__(512) fft2_pass_kernel(PassArgs<V> a) {
    using R = decltype(V::x);
    constexpr int L = L1 * L2;
    constexpr int T = RegShape<L1, L2>::T;
    constexpr bool ROWS = MODE == PM_ROW_LOAD || MODE == PM_ROW_STORE;
    constexpr bool PSF = MODE == PM_COL_PSF;
    constexpr bool HALF_IN = (SPEC & PS_HALF_IN) != 0;
    constexpr bool HALF_OUT = (SPEC & PS_HALF_OUT) != 0;
    constexpr bool HALF_OUT0 = (SPEC & PS_HALF_OUT0) != 0;
    constexpr bool HALF_KEEP = HALF_OUT || HALF_OUT0;   // L/2 outputs kept
    constexpr uint32_t VB = sizeof(V);
    // column passes stream their lines (nontemporal); the row load pass
    // fetches the next batch element's input while transforming the current
    // one (measured: the other passes gain nothing from it and lose occupancy)
    constexpr bool NT = MODE == PM_COL || MODE == PM_COL_PSF;
    constexpr bool PREFETCH = MODE == PM_ROW_LOAD && HALF_IN;
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
    // loads of a tail line read the last valid line; nothing of it is stored
    const int ic = line_ok ? i : a.ni - 1;
    V* xl = lds + c * a.ldp;   // this line's exchange area
    // twiddles after the exchange areas, one row per step-1 owner so a
    // thread's twiddles are consecutive (immediate LDS offsets):
    //   twa[p2 * L1 + k1] = W_L^{p2 k1}  (forward step 1, p2 < L2)
    //   twb[k1 * L2 + m]  = W_L^{k1 m}   (inverse step 1 of the PSF pass, k1 < L1)
    V* twa = lds + int64_t(TX) * a.ldp;
    V* twb = twa + L;
    // With a twist, entry (p2, k1) also carries i^r, r = k1 (output twist:
    // step 2 keeps k1 = k mod 4 when L1 % 4 == 0) plus p2 (input twist:
    // p2 = p mod 4 when L2 % 4 == 0).  The table is conjugated on use when
    // SIGN > 0, and the twist conjugates with it: both directions take i^r.
    // (only the half-line kernels take twists)
    for (int k = tid; k < L; k += blockDim.x) {
        int r = 0;
        if constexpr (HALF_IN && !PSF) r += a.twist_out ? k % L1 : 0;
        if constexpr (HALF_OUT0) r += a.twist_in ? k / L1 : 0;
        twa[k] = quarter_turn(a.tw[((k / L1) * (k % L1)) % L], r);
        if (PSF) twb[k] = a.tw[((k / L2) * (k % L2)) % L];
    }
    __syncthreads();

    const bool sum_b = MODE == PM_ROW_STORE && a.sum_batch;
    // a block runs bchunk consecutive batch elements of its lines (a batch
    // sum: those of its group): per-line data that is the same for every
    // element (the PSF kernel column, broadcast diagonals) is read once
    const int64_t b_first = bblk * a.bchunk;
    const int nb_loop = int(min(int64_t(a.bchunk), int64_t(a.nbatch) - b_first));
    V kv[PSF ? L2 : 1];
    if constexpr (PSF) {
        if (t < L1) {
            // (a block's batch chunk lies within one frame: bchunk divides psf_div)
            const int64_t kf = a.psf_div ? (b_first / a.psf_div) * a.psf_f : 0;
            const char* kb = reinterpret_cast<const char*>(a.psf + kf + int64_t(o) * a.psf_o);
            const int kcol = a.psf_q ? 4 * (ic % a.psf_q) + ic / a.psf_q : ic;
            const uint32_t ko = uint32_t(kcol) * a.psf_ib + uint32_t(t) * a.psf_sb;
#pragma unroll
            for (int k2 = 0; k2 < L2; ++k2) kv[k2] = load_at<V>(kb, ko + uint32_t(L1 * k2) * a.psf_sb);
        }
    }
    R zy = R(1);
    if (ROWS && a.dx) {
        const int z = ic / a.ny;
        zy = a.dz[z] * a.dy[ic - z * a.ny];
    }
    // step-1 input position p = L2 p1 + t of this thread -> line element j,
    // valid when ok (shift and length passed in, laundered per iteration)
    auto in_pos = [&](int p1, int shift, int len, int& j, bool& ok) {
        const int p = L2 * p1 + t;
        if constexpr (HALF_IN) {
            // whole rows of p1 are known valid / zero; one mixed row (L1 odd)
            // keeps a runtime test
            j = p;
            ok = t < L2 && ((L2 * p1 >= L / 2) ? false : (L2 * p1 + L2 <= L / 2 ? true : p < L / 2));
        } else {
            j = p - shift;
            j = j < 0 ? j + L : (j >= L ? j - L : j);
            ok = t < L2 && j < len;
        }
    };
    // input of batch element bx (zeros outside the data)
    auto fetch = [&](int64_t bx, V* dst) {
        int icl = ic;
        uint32_t in_sb = a.in_sb;
        int shift = a.in_shift, len = a.in_len;
        if (nb_loop > 1) asm volatile("" : "+v"(icl), "+s"(in_sb), "+s"(shift), "+s"(len));
        const char* src = reinterpret_cast<const char*>(a.in + bx * a.in_b + int64_t(o) * a.in_o);
        const uint32_t lo = uint32_t(icl) * a.in_ib;
#pragma unroll
        for (int p1 = 0; p1 < L1; ++p1) {
            int j;
            bool ok;
            in_pos(p1, shift, len, j, ok);
            dst[p1] = cmake<V>(R(0), R(0));
            if (ok) dst[p1] = load_at<V, NT>(src, lo + uint32_t(j) * in_sb);
        }
    };
    // broadcast load diagonals (batch stride 0) of the prefetching row pass
    // are read once per block
    V lw1[PREFETCH ? L1 : 1], lw2[PREFETCH ? L1 : 1];
    if constexpr (PREFETCH) {
        auto pre = [&](const DiagOp<V>& d, uint32_t rb, V* w) {
            if (d.ptr && d.sb == 0) {
                const char* base = reinterpret_cast<const char*>(d.ptr);
#pragma unroll
                for (int p1 = 0; p1 < L1; ++p1) {
                    int j;
                    bool ok;
                    in_pos(p1, 0, L / 2, j, ok);
                    if (ok) w[p1] = load_at<V>(base, uint32_t(ic) * rb + uint32_t(j) * VB);
                }
            }
        };
        pre(a.lm1, a.lm1_rb, lw1);
        pre(a.lm2, a.lm2_rb, lw2);
    }
    // deapodisation of the prefetching row pass, zy * dx[j] per held input,
    // read once per block (the same for every batch element)
    R dxw[PREFETCH ? L1 : 1];
    if constexpr (PREFETCH) {
        if (a.dx) {
#pragma unroll
            for (int p1 = 0; p1 < L1; ++p1) {
                int j;
                bool ok;
                in_pos(p1, 0, L / 2, j, ok);
                dxw[p1] = ok ? zy * a.dx[j] : R(0);
            }
        }
    }
    // coil-sum accumulators: outputs j = t + T m, m < NACC (the coil sum is
    // used with a crop to at most L/2 outputs; the host checks out_len)
    constexpr int NACC = (L / 2 + T - 1) / T;
    V acc[MODE == PM_ROW_STORE ? NACC : 1];
    if constexpr (MODE == PM_ROW_STORE) {
#pragma unroll
        for (int m = 0; m < NACC; ++m) acc[m] = cmake<V>(R(0), R(0));
    }
    V nxt[PREFETCH ? L1 : 1];
    if constexpr (PREFETCH) fetch(b_first, nxt);
    for (int bb = 0; bb < nb_loop; ++bb) {
        const int64_t b = b_first + bb;
        // per-iteration copies the compiler cannot hoist: otherwise the
        // per-position offsets and masks of every load and store stay live in
        // registers across the batch loop (and cut the occupancy)
        uint32_t out_sb = a.out_sb;
        int out_shift = a.out_shift, out_len = a.out_len, in_shift = a.in_shift, in_len = a.in_len;
        int icl = ic;
        if (nb_loop > 1)
            asm volatile("" : "+s"(out_sb), "+s"(out_shift), "+s"(out_len), "+s"(in_shift), "+s"(in_len),
                         "+v"(icl));
        V v[T];
        if constexpr (PREFETCH) {
#pragma unroll
            for (int p1 = 0; p1 < L1; ++p1) v[p1] = nxt[p1];
            if (bb + 1 < nb_loop) fetch(b + 1, nxt);
        } else {
            fetch(b, v);
        }
        // ---- step 1: thread t = p2 holds p = L2 p1 + t ----
        if (t < L2) {
            if constexpr (MODE == PM_ROW_LOAD) {
                int jin[L1];
                bool okin[L1];
#pragma unroll
                for (int p1 = 0; p1 < L1; ++p1) in_pos(p1, in_shift, in_len, jin[p1], okin[p1]);
                auto use = [&](const DiagOp<V>& d, uint32_t rb, const V* w) {
                    if (!d.ptr) return;
                    if constexpr (PREFETCH) {
                        if (d.sb == 0) {
                            const decltype(V::x) sgn = d.type == HASTY_MULT_CONJ ? -1 : 1;
#pragma unroll
                            for (int p1 = 0; p1 < L1; ++p1)
                                if (okin[p1]) v[p1] = cmul_sgn(v[p1], w[p1], sgn);
                            return;
                        }
                    }
                    apply_diag<V, L1>(v, d, b, uint32_t(icl) * rb, jin, okin);
                };
                use(a.lm1, a.lm1_rb, lw1);
                use(a.lm2, a.lm2_rb, lw2);
                if (a.dx) {
                    if constexpr (PREFETCH) {
#pragma unroll
                        for (int p1 = 0; p1 < L1; ++p1)
                            if (okin[p1]) v[p1] = cscale(v[p1], dxw[p1]);
                    } else {
#pragma unroll
                        for (int p1 = 0; p1 < L1; ++p1)
                            if (okin[p1]) v[p1] = cscale(v[p1], zy * a.dx[jin[p1]]);
                    }
                }
            }
            RegDFT<L1, SIGN, V>::run(v);
            const V* tw = twa + t * L1;
            // an input twist is the same for every p1 of this thread, so it
            // applies after the L1-point DFT, to k1 = 0 as well
            if constexpr (HALF_OUT0) {
                if (a.twist_in) v[0] = twiddle_mul<V, SIGN>(v[0], tw, 0);
            }
#pragma unroll
            for (int k1 = 1; k1 < L1; ++k1) v[k1] = twiddle_mul<V, SIGN>(v[k1], tw, k1);
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
#pragma unroll
                for (int k2 = 0; k2 < L2; ++k2) v[k2] = cmul(v[k2], kv[k2]);
                // inverse, factorisation L = L2 * L1: step 1 over the L2 held values
                RegDFT<L2, -SIGN, V>::run(v);
                const V* tw = twb + t * L2;
#pragma unroll
                for (int k1 = 1; k1 < L2; ++k1) v[k1] = twiddle_mul<V, -SIGN>(v[k1], tw, k1);
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
                // outputs j = t + T m: HALF_OUT keeps j < L/2 at line positions
                // q = j + L/2 - 1, HALF_OUT0 at q = j; otherwise j < out_len at
                // q = j + out_shift mod L, taken in rounds of NACC outputs per
                // thread (the coil sum has one round: the host keeps
                // out_len <= NACC T)
                constexpr int NR = HALF_KEEP ? 1 : (L + T * NACC - 1) / (T * NACC);
                char* dst = reinterpret_cast<char*>(a.out + (sum_b ? 0 : b * a.out_b) + int64_t(o) * a.out_o);
                const uint32_t lo = uint32_t(icl) * a.out_ib;
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    if (!HALF_KEEP && T * NACC * r >= out_len) break;
                    int jj[NACC];
                    bool ok[NACC];
                    V val[NACC];
#pragma unroll
                    for (int m = 0; m < NACC; ++m) {
                        const int j = t + T * (NACC * r + m);
                        int q;
                        if constexpr (HALF_KEEP) {
                            ok[m] = T * (m + 1) <= L / 2 || j < L / 2;
                            q = HALF_OUT ? j + L / 2 - 1 : j;
                        } else {
                            ok[m] = j < out_len;
                            q = j + out_shift;
                            q = q >= L ? q - L : (q < 0 ? q + L : q);
                        }
                        jj[m] = j;
                        val[m] = cmake<V>(R(0), R(0));
                        if (ok[m]) val[m] = cscale(xl[q], a.scale);
                    }
                    apply_diag<V, NACC>(val, a.sm1, b, uint32_t(icl) * a.sm1_rb, jj, ok);
                    apply_diag<V, NACC>(val, a.sm2, b, uint32_t(icl) * a.sm2_rb, jj, ok);
                    apply_diag<V, NACC>(val, a.sio, b, uint32_t(icl) * a.sio_rb, jj, ok);
                    if (a.dx) {
#pragma unroll
                        for (int m = 0; m < NACC; ++m)
                            if (ok[m]) val[m] = cscale(val[m], zy * a.dx[jj[m]]);
                    }
                    if (sum_b) {
#pragma unroll
                        for (int m = 0; m < NACC; ++m)
                            if (ok[m]) acc[m] = cadd(acc[m], val[m]);
                    } else {
#pragma unroll
                        for (int m = 0; m < NACC; ++m) {
                            if (ok[m]) {
                                const uint32_t off = lo + uint32_t(jj[m]) * VB;
                                store_at<V>(dst, off, a.accumulate ? cadd(val[m], load_at<V>(dst, off)) : val[m]);
                            }
                        }
                    }
                }
            }
        } else {
            bool staged = false;
            if constexpr (MODE == PM_ROW_LOAD && L % 4 == 0) {
                if (a.perm4) {
                    // class-major store (full lines, no shift): stage the line
                    // in natural order, then each thread stores consecutive
                    // positions q = r L/4 + m, which hold frequency 4 m + r
                    staged = true;
                    __syncthreads();
                    if (t < NOWNERS) {
#pragma unroll
                        for (int k = 0; k < NOWN; ++k) xl[t + OSTRIDE * k] = v[k];
                    }
                    __syncthreads();
                    if (line_ok) {
                        char* dst = reinterpret_cast<char*>(a.out + b * a.out_b + int64_t(o) * a.out_o);
                        const uint32_t lo = uint32_t(icl) * a.out_ib;
#pragma unroll
                        for (int u = 0; u < L / T; ++u) {
                            const int q = t + T * u;
                            const int f = 4 * (q % (L / 4)) + q / (L / 4);
                            store_at<V>(dst, lo + uint32_t(q) * out_sb, cscale(xl[f], a.scale));
                        }
                    }
                }
            }
            // owner t holds positions t + OSTRIDE k, k < NOWN
            if (!staged && t < NOWNERS && line_ok) {
                char* dst = reinterpret_cast<char*>(a.out + b * a.out_b + int64_t(o) * a.out_o);
                const uint32_t lo = uint32_t(icl) * a.out_ib;
#pragma unroll
                for (int k = 0; k < NOWN; ++k) {
                    const int q = t + OSTRIDE * k;
                    int j;
                    bool ok;
                    if constexpr (HALF_KEEP) {
                        // kept: lo_q <= q < hi_q, q in [OSTRIDE k, OSTRIDE (k + 1))
                        constexpr int lo_q = HALF_OUT ? L / 2 - 1 : 0, hi_q = HALF_OUT ? L - 1 : L / 2;
                        j = q - lo_q;
                        const bool none = OSTRIDE * (k + 1) <= lo_q || OSTRIDE * k >= hi_q;
                        const bool all = OSTRIDE * k >= lo_q && OSTRIDE * (k + 1) <= hi_q;
                        ok = none ? false : (all ? true : (q >= lo_q && q < hi_q));
                    } else {
                        j = q - out_shift;
                        j = j < 0 ? j + L : (j >= L ? j - L : j);
                        ok = j < out_len;
                    }
                    if (ok) store_at<V, NT>(dst, lo + uint32_t(j) * out_sb, cscale(v[k], a.scale));
                }
            }
        }
        if (nb_loop > 1) __syncthreads();   // exchange area reused by the next b
    }
    if constexpr (MODE == PM_ROW_STORE) {
        if (sum_b && line_ok) {
            const int out_len = HALF_KEEP ? L / 2 : a.out_len;
            if (a.groups > 1) {
                // this group's partial sum, rows of out_len values
                V* part = a.partial + (bblk * a.ni + i) * int64_t(out_len);
#pragma unroll
                for (int m = 0; m < NACC; ++m) {
                    const int j = t + T * m;
                    if (j < out_len) part[j] = acc[m];
                }
            } else {
                char* dst = reinterpret_cast<char*>(a.out + int64_t(o) * a.out_o);
                const uint32_t lo = uint32_t(i) * a.out_ib;
#pragma unroll
                for (int m = 0; m < NACC; ++m) {
                    const int j = t + T * m;
                    if (j < out_len) {
                        const uint32_t off = lo + uint32_t(j) * VB;
                        store_at<V>(dst, off, a.accumulate ? cadd(acc[m], load_at<V>(dst, off)) : acc[m]);
                    }
                }
            }
        }
    }
}

// out[row][j] (+)= sum_g partial[g][row][j], groups summed in order
template <class V>
__global__ void __launch_bounds__(256) sum_partials_kernel(const V* __restrict__ partial, V* out, int64_t nrows,
                                                           int out_len, int64_t out_r, int groups,
                                                           int accumulate, int64_t out_f = 0) {
    const int64_t n = nrows * out_len;
    // frames (gridDim.y): frame f's partials follow frame f-1's groups
    partial += int64_t(blockIdx.y) * groups * n;
    out += int64_t(blockIdx.y) * out_f;
    for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n;
         idx += int64_t(gridDim.x) * blockDim.x) {
        const int64_t r = idx / out_len;
        const int j = int(idx - r * out_len);
        V s = partial[idx];
        for (int g = 1; g < groups; ++g) s = cadd(s, partial[g * n + idx]);
        V* d = out + r * out_r + j;
        *d = accumulate ? cadd(s, *d) : s;
    }
}

// Host dispatch: returns false when L has no register factorisation here
// (the caller then uses the LDS Stockham kernels).
template <class V> bool launch_fft2_pass(PassArgs<V> a, int L, int sign, int mode,
                                         const char* tag, hipStream_t s);

}  // namespace hasty
