// Stockham mixed-radix (2,3,4,5,8) FFT passes for gfx950.
//
// Every multi-dimensional transform in this library (NUFFT fine grid,
// Toeplitz 2N grid, PSF kernel) is built from two pass kernels:
//   * row_load / row_store : lines along the contiguous (x) axis,
//   * col_pass             : lines along a strided axis, a tile of `tx`
//                            neighbouring columns per block so that every
//                            global access is a run of tx complex values.
// A line is loaded into LDS with zero padding / index rotation folded into
// the load ("pad"), transformed with the Stockham autosort recurrence
// (Govindaraju et al., SC'08: out[(j/Ns)*Ns*R + j%Ns + r*Ns] =
//  DFT_R(in[j + r*L/R] * w^{r*(j%Ns)})), and written back with only the kept
// outputs ("crop"), optional diagonal products, deapodisation and coil sums.
// Pruning falls out of the pass structure: a pass only runs over lines that
// hold nonzero input and only stores outputs that a later pass reads.
#pragma once

#include "common.hpp"

namespace hasty {

constexpr int kFFTThreads = 256;
constexpr int kMaxStages = 12;

struct FFTPlan1D {
    int L;
    int nstages;
    int radix[kMaxStages];
};

FFTPlan1D make_fft_plan(int L);               // throws if L is not 2^a 3^b 5^c
bool is_smooth(int64_t n);
int64_t next_smooth(int64_t n);
// Device twiddle table tw[k] = exp(-2 pi i k / L), k in [0, L), cached per
// (device, L, precision); lives until hasty_release_caches / process end.
const void* twiddles(int L, bool f64);
// HASTY_FFT_STOCKHAM set: the LDS Stockham kernels replace the register passes
bool fft_force_stockham();

// --------------------------------------------------------------------------
// Radix butterflies: X[q] = sum_r v[r] exp(sign * 2 pi i r q / R)
// --------------------------------------------------------------------------
template <class V> __device__ inline V mul_i(V a, int sign) {  // (sign * i) * a
    return sign < 0 ? cmake<V>(a.y, -a.x) : cmake<V>(-a.y, a.x);
}

template <int R, class V> struct Butterfly;

template <class V> struct Butterfly<2, V> {
    __device__ static void run(V* v, int) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <class V> struct Butterfly<4, V> {
    __device__ static void run(V* v, int sign) {
        V s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        V s13 = cadd(v[1], v[3]), d13 = mul_i(csub(v[1], v[3]), sign);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};
template <class V> struct Butterfly<3, V> {
    __device__ static void run(V* v, int sign) {
        using R = decltype(V::x);
        const R h = R(0.86602540378443864676372317075294);  // sqrt(3)/2
        V t = cadd(v[1], v[2]);
        V d = mul_i(cscale(csub(v[1], v[2]), h), sign);
        V m = csub(v[0], cscale(t, R(0.5)));
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, d);
        v[2] = csub(m, d);
    }
};
template <class V> struct Butterfly<5, V> {
    __device__ static void run(V* v, int sign) {
        using R = decltype(V::x);
        const R c1 = R(0.30901699437494742410229341718282);   // cos(2pi/5)
        const R c2 = R(-0.80901699437494742410229341718282);  // cos(4pi/5)
        const R s1 = R(0.95105651629515357211643933337938);   // sin(2pi/5)
        const R s2 = R(0.58778525229247312916870595463907);   // sin(4pi/5)
        V a14 = cadd(v[1], v[4]), b14 = csub(v[1], v[4]);
        V a23 = cadd(v[2], v[3]), b23 = csub(v[2], v[3]);
        V m1 = cadd(v[0], cadd(cscale(a14, c1), cscale(a23, c2)));
        V m2 = cadd(v[0], cadd(cscale(a14, c2), cscale(a23, c1)));
        V n1 = mul_i(cadd(cscale(b14, s1), cscale(b23, s2)), sign);
        V n2 = mul_i(csub(cscale(b14, s2), cscale(b23, s1)), sign);
        v[0] = cadd(v[0], cadd(a14, a23));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};
template <class V> struct Butterfly<8, V> {
    __device__ static void run(V* v, int sign) {
        using R = decltype(V::x);
        const R r2 = R(0.70710678118654752440084436210485);
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        Butterfly<4, V>::run(e, sign);
        Butterfly<4, V>::run(o, sign);
        // w8^1 = r2 (1 + s i), w8^2 = s i, w8^3 = r2 (-1 + s i)
        V o1 = cscale(cadd(o[1], mul_i(o[1], sign)), r2);
        V o2 = mul_i(o[2], sign);
        V o3 = cscale(csub(mul_i(o[3], sign), o[3]), r2);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// One Stockham stage over `nlines` lines (line stride LP) from src to dst.
template <int R, class V>
__device__ inline void stockham_stage(const V* __restrict__ src, V* __restrict__ dst, int nlines,
                                      int L, int LP, int Ns, int sign, const V* __restrict__ tw) {
    const int nb = L / R;
    const int tstride = L / (Ns * R);
    const int total = nlines * nb;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int line = t / nb;
        const int j = t - line * nb;
        const V* in = src + line * LP;
        V* out = dst + line * LP;
        const int k = j % Ns;
        V v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[j + r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                V w = tw[r * k * tstride];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        Butterfly<R, V>::run(v, sign);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
    }
}

// Transforms nlines lines held in buf0 (stride LP); buf1 is scratch of the
// same size.  Returns the buffer holding the result.  Block-synchronous.
template <class V>
__device__ V* lds_fft(V* buf0, V* buf1, int nlines, int LP, const FFTPlan1D& plan, int sign,
                      const V* __restrict__ tw) {
    V* src = buf0;
    V* dst = buf1;
    int Ns = 1;
    for (int s = 0; s < plan.nstages; ++s) {
        const int R = plan.radix[s];
        switch (R) {
            case 8: stockham_stage<8, V>(src, dst, nlines, plan.L, LP, Ns, sign, tw); break;
            case 4: stockham_stage<4, V>(src, dst, nlines, plan.L, LP, Ns, sign, tw); break;
            case 2: stockham_stage<2, V>(src, dst, nlines, plan.L, LP, Ns, sign, tw); break;
            case 3: stockham_stage<3, V>(src, dst, nlines, plan.L, LP, Ns, sign, tw); break;
            default: stockham_stage<5, V>(src, dst, nlines, plan.L, LP, Ns, sign, tw); break;
        }
        __syncthreads();
        V* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

__device__ inline int wrap_mod(int a, int L) {
    a %= L;
    return a < 0 ? a + L : a;
}

// --------------------------------------------------------------------------
// Pass descriptors (passed by value as kernel arguments)
// --------------------------------------------------------------------------
template <class V> struct DiagOp {
    const V* ptr = nullptr;   // element (b, r, j) at ptr[b*sb + r*sr + j]
    int64_t sb = 0, sr = 0;
    int type = HASTY_MULT_NONE;
    int64_t sf = 0;           // frame stride: passes batched over frames read frame f at ptr + f*sf
};

template <class V> struct RowLoadArgs {
    const V* in; V* out;
    int64_t nrows; int nbatch;
    int64_t in_b, in_r, out_b, out_r;
    int L, LP, in_len, in_shift, sign, nl;
    DiagOp<V> m1, m2;
    const decltype(V::x)* dz; const decltype(V::x)* dy; const decltype(V::x)* dx;
    int ny;                                  // row r -> (z = r / ny, y = r % ny)
    FFTPlan1D plan; const V* tw;
    const char* tag;                         // launch name for profiling (host only)
    int perm4 = 0;                           // class-major store (register passes only)
};

template <class V> struct RowStoreArgs {
    const V* in; V* out;
    int64_t nrows; int nbatch;
    int64_t in_b, in_r, out_b, out_r;
    int L, LP, out_len, out_shift, sign, nl;
    decltype(V::x) scale;
    DiagOp<V> m1, m2, io;
    const decltype(V::x)* dz; const decltype(V::x)* dy; const decltype(V::x)* dx;
    int ny;
    int accumulate;   // out += result
    int sum_batch;    // out has no batch axis; results summed over b
    // optional device scratch (complex elements) free during this pass: lets
    // a batch sum run in several block groups whose partial sums are added
    // by a second kernel (deterministic order)
    V* scratch; int64_t scratch_elems;
    FFTPlan1D plan; const V* tw;
    const char* tag;
};

template <class V> struct ColArgs {
    const V* in; V* out;
    int64_t in_b, in_o, in_s, out_b, out_o, out_s;
    int ni, no, nbatch, tx;
    int L, LP, in_len, in_shift, out_len, out_shift, sign;
    const V* psf; int64_t psf_o, psf_s;      // optional: multiply then inverse FFT
    decltype(V::x) scale;
    FFTPlan1D plan; const V* tw;
    const char* tag;
    int psf_q = 0;                           // class-major kernel columns: line i reads 4 (i % psf_q) + i / psf_q
    // frames: batch element b belongs to frame b / psf_div, whose kernel is
    // psf + (b / psf_div) * psf_f (psf_div = 0: one kernel for the batch)
    int64_t psf_f = 0; int psf_div = 0;
};

constexpr int kRowAccMax = 16;

// --------------------------------------------------------------------------
// Kernels
// --------------------------------------------------------------------------
template <class V>
__global__ void __launch_bounds__(kFFTThreads) row_load_kernel(RowLoadArgs<V> a) {
    using R = decltype(V::x);
    extern __shared__ unsigned char smem_raw[];
    V* buf0 = reinterpret_cast<V*>(smem_raw);
    V* buf1 = buf0 + a.nl * a.LP;
    const int64_t nlines_total = a.nrows * a.nbatch;
    const int64_t g0 = int64_t(blockIdx.x) * a.nl;
    const int nl = int(min<int64_t>(a.nl, nlines_total - g0));
    for (int t = threadIdx.x; t < nl * a.L; t += blockDim.x) {
        const int li = t / a.L;
        const int p = t - li * a.L;
        const int64_t g = g0 + li;
        const int64_t b = g / a.nrows;
        const int64_t r = g - b * a.nrows;
        const int j = wrap_mod(p - a.in_shift, a.L);
        V v = cmake<V>(R(0), R(0));
        if (j < a.in_len) {
            v = a.in[b * a.in_b + r * a.in_r + j];
            if (a.m1.ptr) v = diag_mult(v, a.m1.ptr[b * a.m1.sb + r * a.m1.sr + j], a.m1.type);
            if (a.m2.ptr) v = diag_mult(v, a.m2.ptr[b * a.m2.sb + r * a.m2.sr + j], a.m2.type);
            if (a.dx) {
                const int64_t z = r / a.ny, y = r - z * a.ny;
                v = cscale(v, a.dz[z] * a.dy[y] * a.dx[j]);
            }
        }
        buf0[li * a.LP + p] = v;
    }
    __syncthreads();
    V* res = lds_fft(buf0, buf1, nl, a.LP, a.plan, a.sign, a.tw);
    for (int t = threadIdx.x; t < nl * a.L; t += blockDim.x) {
        const int li = t / a.L;
        const int p = t - li * a.L;
        const int64_t g = g0 + li;
        const int64_t b = g / a.nrows;
        const int64_t r = g - b * a.nrows;
        a.out[b * a.out_b + r * a.out_r + p] = res[li * a.LP + p];
    }
}

template <class V>
__global__ void __launch_bounds__(kFFTThreads) row_store_kernel(RowStoreArgs<V> a) {
    using R = decltype(V::x);
    extern __shared__ unsigned char smem_raw[];
    V* buf0 = reinterpret_cast<V*>(smem_raw);
    V* buf1 = buf0 + a.nl * a.LP;
    // With sum_batch the block owns nl rows and loops over b; otherwise it
    // owns nl (b, row) lines.
    const int64_t nunits = a.sum_batch ? a.nrows : a.nrows * a.nbatch;
    const int64_t g0 = int64_t(blockIdx.x) * a.nl;
    const int nl = int(min<int64_t>(a.nl, nunits - g0));
    const int nb_loop = a.sum_batch ? a.nbatch : 1;
    V acc[kRowAccMax];
#pragma unroll
    for (int k = 0; k < kRowAccMax; ++k) acc[k] = cmake<V>(R(0), R(0));
    for (int bb = 0; bb < nb_loop; ++bb) {
        for (int t = threadIdx.x; t < nl * a.L; t += blockDim.x) {
            const int li = t / a.L;
            const int p = t - li * a.L;
            const int64_t g = g0 + li;
            int64_t b, r;
            if (a.sum_batch) { b = bb; r = g; } else { b = g / a.nrows; r = g - b * a.nrows; }
            buf0[li * a.LP + p] = a.in[b * a.in_b + r * a.in_r + p];
        }
        __syncthreads();
        V* res = lds_fft(buf0, buf1, nl, a.LP, a.plan, a.sign, a.tw);
#pragma unroll
        for (int k = 0; k < kRowAccMax; ++k) {
            const int t = threadIdx.x + k * blockDim.x;
            if (t < nl * a.out_len) {
                const int li = t / a.out_len;
                const int j = t - li * a.out_len;
                const int64_t g = g0 + li;
                int64_t b, r;
                if (a.sum_batch) { b = bb; r = g; } else { b = g / a.nrows; r = g - b * a.nrows; }
                const int p = wrap_mod(j + a.out_shift, a.L);
                V v = cscale(res[li * a.LP + p], a.scale);
                if (a.m1.ptr) v = diag_mult(v, a.m1.ptr[b * a.m1.sb + r * a.m1.sr + j], a.m1.type);
                if (a.m2.ptr) v = diag_mult(v, a.m2.ptr[b * a.m2.sb + r * a.m2.sr + j], a.m2.type);
                if (a.io.ptr) v = diag_mult(v, a.io.ptr[b * a.io.sb + r * a.io.sr + j], a.io.type);
                if (a.dx) {
                    const int64_t z = r / a.ny, y = r - z * a.ny;
                    v = cscale(v, a.dz[z] * a.dy[y] * a.dx[j]);
                }
                acc[k] = cadd(acc[k], v);
            }
        }
        __syncthreads();   // buffers are reused by the next b
    }
#pragma unroll
    for (int k = 0; k < kRowAccMax; ++k) {
        const int t = threadIdx.x + k * blockDim.x;
        if (t < nl * a.out_len) {
            const int li = t / a.out_len;
            const int j = t - li * a.out_len;
            const int64_t g = g0 + li;
            int64_t b, r;
            if (a.sum_batch) { b = 0; r = g; } else { b = g / a.nrows; r = g - b * a.nrows; }
            V* o = a.out + b * a.out_b + r * a.out_r + j;
            V v = acc[k];
            if (a.accumulate) v = cadd(v, *o);
            *o = v;
        }
    }
}

template <class V>
__global__ void __launch_bounds__(kFFTThreads) col_pass_kernel(ColArgs<V> a) {
    extern __shared__ unsigned char smem_raw[];
    using R = decltype(V::x);
    V* buf0 = reinterpret_cast<V*>(smem_raw);
    V* buf1 = buf0 + a.tx * a.LP;
    const int ntile = (a.ni + a.tx - 1) / a.tx;
    int64_t bid = blockIdx.x;
    const int it = int(bid % ntile); bid /= ntile;
    const int o = int(bid % a.no); bid /= a.no;
    const int64_t b = bid;
    const int i0 = it * a.tx;
    const int tx = min(a.tx, a.ni - i0);
    const V* in = a.in + b * a.in_b + int64_t(o) * a.in_o + i0;
    for (int t = threadIdx.x; t < tx * a.L; t += blockDim.x) {
        const int c = t % tx;
        const int p = t / tx;
        const int j = wrap_mod(p - a.in_shift, a.L);
        V v = cmake<V>(R(0), R(0));
        if (j < a.in_len) v = in[int64_t(j) * a.in_s + c];
        buf0[c * a.LP + p] = v;
    }
    __syncthreads();
    V* res = lds_fft(buf0, buf1, tx, a.LP, a.plan, a.sign, a.tw);
    int sign_out = a.sign;
    if (a.psf) {
        const V* k = a.psf + int64_t(o) * a.psf_o;
        for (int t = threadIdx.x; t < tx * a.L; t += blockDim.x) {
            const int c = t % tx;
            const int p = t / tx;
            const int i = i0 + c;
            const int kc = a.psf_q ? 4 * (i % a.psf_q) + i / a.psf_q : i;   // class-major columns
            res[c * a.LP + p] = cmul(res[c * a.LP + p], k[int64_t(p) * a.psf_s + kc]);
        }
        __syncthreads();
        V* other = (res == buf0) ? buf1 : buf0;
        res = lds_fft(res, other, tx, a.LP, a.plan, -a.sign, a.tw);
        sign_out = -a.sign;
    }
    (void)sign_out;
    V* out = a.out + b * a.out_b + int64_t(o) * a.out_o + i0;
    for (int t = threadIdx.x; t < tx * a.out_len; t += blockDim.x) {
        const int c = t % tx;
        const int j = t / tx;
        const int p = wrap_mod(j + a.out_shift, a.L);
        out[int64_t(j) * a.out_s + c] = cscale(res[c * a.LP + p], a.scale);
    }
}

// --------------------------------------------------------------------------
// Host launch helpers (defined in fft.hip for float2 and double2)
// --------------------------------------------------------------------------
template <class V> void launch_row_load(RowLoadArgs<V> a, hipStream_t s);
template <class V> void launch_row_store(RowStoreArgs<V> a, hipStream_t s);
template <class V> void launch_col_pass(ColArgs<V> a, hipStream_t s);

// LDS bytes a pass may use per block.
constexpr int64_t kPassLdsBudget = 80 * 1024;

}  // namespace hasty
