// Spread (type 1) and interpolation (type 2) kernels for gfx950.
//
// Points are bin-sorted by fine-grid tile at setpts; for every sorted point s
// the plan holds its first support cell per slot (sl0), the fractional offset
// (sfrac) and the w window weights per slot (sw, evaluated once in FP64).
#pragma once

#include "common.hpp"
#include "kb.hpp"

namespace hasty {

constexpr int kNufftThreads = 512;
constexpr int kMaxWidthF32 = 9;
constexpr int kMaxWidthF64 = 14;

struct Geo {
    int64_t nf[3];
    int T[3];
    int ntile[3];
    int wa[3];        // support width per slot (1 on a degenerate slot)
    double half_w;
};

__device__ inline void tile_origin(const Geo& g, int tile, int t0[3]) {
    const int tx = tile % g.ntile[2];
    const int ty = (tile / g.ntile[2]) % g.ntile[1];
    const int tz = tile / (g.ntile[2] * g.ntile[1]);
    t0[0] = tz * g.T[0];
    t0[1] = ty * g.T[1];
    t0[2] = tx * g.T[2];
}

template <class V>
struct GridArgs {
    using R = decltype(V::x);
    V* fine;                // [B, nf]
    V* cdata;               // [B, M]   (original point order)
    const V* csort;         // [M, B]   (sorted point order, transform fastest)
    const int* sidx;
    const int* sl0;         // [3][M]
    const R* sw;            // [M][ws]: w weights per slot, slot-major, padded to ws
    const int4* subs;       // (tile, first, count, split)
    const int4* visits;     // (8 s + tile-neighbour bits, l0z, l0y, l0x)
    int64_t M;
    int nbatch, group, w, ws;
    int cstride;            // LDS words between transforms (spread)
    V* partial;             // [slot][nbatch][tile cells]: partial tiles of split tiles
    const unsigned* cmax;   // bits of max(|Re c|, |Im c|) over the batch (spread)
    double wmax3;           // (window peak)^d, bound of the weight product
    Geo g;
};

// ---------------------------------------------------------------------------
// Interpolation: one block per (tile subproblem, transform group); the padded
// tile footprint [t0, t0 + T + w - 1) of each transform is staged in LDS and
// each thread gathers the w^d cells of its points.
// ---------------------------------------------------------------------------
template <class V, int W>
__global__ void __launch_bounds__(kNufftThreads) interp_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const int4 sub = a.subs[blockIdx.x];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    const int Pz = a.g.T[0] + a.g.wa[0] - 1;
    const int Py = a.g.T[1] + a.g.wa[1] - 1;
    const int Px = a.g.T[2] + a.g.wa[2] - 1;
    const int cells = Pz * Py * Px;
    const int g0 = blockIdx.y * a.group;
    const int ng = min(a.group, a.nbatch - g0);
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    // stage, 4 loads in flight per thread
    const int total = ng * cells;
    for (int base = threadIdx.x; base < total; base += 4 * blockDim.x) {
        V tmp[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = base + u * blockDim.x;
            if (idx < total) {
                const int gi = idx / cells;
                const int c = idx - gi * cells;
                const int px = c % Px;
                const int py = (c / Px) % Py;
                const int pz = c / (Px * Py);
                int gz = t0[0] + pz; gz -= gz >= a.g.nf[0] ? int(a.g.nf[0]) : 0;
                int gy = t0[1] + py; gy -= gy >= a.g.nf[1] ? int(a.g.nf[1]) : 0;
                int gx = t0[2] + px; gx -= gx >= a.g.nf[2] ? int(a.g.nf[2]) : 0;
                tmp[u] = a.fine[int64_t(g0 + gi) * G + (int64_t(gz) * a.g.nf[1] + gy) * a.g.nf[2] + gx];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = base + u * blockDim.x;
            if (idx < total) lds[idx] = tmp[u];
        }
    }
    __syncthreads();
    const int wz = a.g.wa[0];
    for (int s = sub.y + threadIdx.x; s < sub.y + sub.z; s += blockDim.x) {
        const int lz = a.sl0[s] - t0[0];
        const int ly = a.sl0[a.M + s] - t0[1];
        const int lx = a.sl0[2 * a.M + s] - t0[2];
        const R* w = a.sw + int64_t(s) * a.ws;
        R wz_[W], wy[W], wx[W];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            wz_[q] = w[q];
            wy[q] = w[W + q];
            wx[q] = w[2 * W + q];
        }
        const int orig = a.sidx[s];
        for (int gi = 0; gi < ng; ++gi) {
            V acc = cmake<V>(R(0), R(0));
#pragma unroll
            for (int qz = 0; qz < W; ++qz) {
                if (qz < wz) {
                    const V* base = lds + gi * cells + ((lz + qz) * Py + ly) * Px + lx;
                    V accy = cmake<V>(R(0), R(0));
#pragma unroll
                    for (int qy = 0; qy < W; ++qy) {
                        V accx = cmake<V>(R(0), R(0));
#pragma unroll
                        for (int qx = 0; qx < W; ++qx) {
                            const V v = base[qy * Px + qx];
                            accx.x += wx[qx] * v.x;
                            accx.y += wx[qx] * v.y;
                        }
                        accy.x += wy[qy] * accx.x;
                        accy.y += wy[qy] * accx.y;
                    }
                    acc.x += wz_[qz] * accy.x;
                    acc.y += wz_[qz] * accy.y;
                }
            }
            a.cdata[int64_t(g0 + gi) * a.M + orig] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Spreading: one block per (tile subproblem, transform group of <= 32); the
// block owns the tile's T_z x T_y x T_x cells of every transform in its group.
//
// Accumulation is exact 64-bit fixed point with integer LDS atomics (float
// atomics on LDS run ~30x slower than integer ones on gfx950).  Every
// contribution v = c w_z w_y w_x is scaled by 2^k, k chosen per batch so that
// |v| 2^k < 2^51 (max |c| of the batch times the window peak cubed), and
// rounded to an integer with the 1.5 * 2^52 magic-number addition in FP64;
// a tile sums at most max_subproblem such terms, far below 2^63.  The sum is
// therefore independent of the order of the atomics (deterministic), and its
// absolute error is below (terms) * 2^-k-1.
//
// Each wave takes one visit at a time; its lanes are the (transform, re/im)
// pairs, so the visit record, weights and clip ranges are wave-uniform
// (scalar loads / branches) and the 64 atomics of one instruction hit 64
// distinct words.  The next visit's record, weights and value are loaded
// before the current visit's atomics are issued.  Split tiles write their
// share to a partial slot.
// ---------------------------------------------------------------------------
// Broadcast lane 0's value to the wave as a scalar (the value is known to be
// wave-uniform; this lets the compiler keep it in SGPRs and branch on it).
__device__ inline int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline float wave_uniform(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ inline double wave_uniform(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane(int(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane(int(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// round(x) for |x| < 2^51 as a two's-complement 64-bit integer
__device__ inline unsigned long long fixed_round(double x) {
    const double magic = 6755399441055744.0;   // 1.5 * 2^52
    return static_cast<unsigned long long>(__double_as_longlong(x + magic) -
                                           __double_as_longlong(magic));
}

// Fixed-point scale 2^k with (peak contribution) * 2^k <= 2^51.
__device__ inline double spread_scale(const unsigned* cmax_bits, double wmax3) {
    if (cmax_bits == nullptr) return 1.0;
    const double bound = double(__uint_as_float(*cmax_bits)) * wmax3;
    if (!(bound > 0.0) || !(bound < 1e300)) return 1.0;
    int e = 0;
    (void)frexp(bound, &e);          // bound < 2^e
    return ldexp(1.0, 51 - e);
}

template <class V, int W>
__global__ void __launch_bounds__(kNufftThreads) spread_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    using Acc = unsigned long long;
    extern __shared__ unsigned char smem_raw[];
    Acc* lds = reinterpret_cast<Acc*>(smem_raw);
    const int4 sub = a.subs[blockIdx.x];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    const int Tz = a.g.T[0], Ty = a.g.T[1], Tx = a.g.T[2];
    const int cells = Tz * Ty * Tx;
    const int g0 = blockIdx.y * a.group;
    const int ng = min(a.group, a.nbatch - g0);
    const int CS = a.cstride;
    for (int idx = threadIdx.x; idx < a.group * CS; idx += blockDim.x) lds[idx] = 0ull;
    const double scale = spread_scale(a.cmax, a.wmax3);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = wave_uniform(int(threadIdx.x >> 6));
    const int nwaves = blockDim.x >> 6;
    const int gi = lane >> 1;
    const int comp = lane & 1;
    const bool active = gi < ng;
    Acc* my = lds + (active ? gi : 0) * CS + comp;
    const int end = sub.y + sub.z;
    const int WS = a.ws;
    int v = sub.y + wave;
    int4 rec_n = make_int4(0, 0, 0, 0);
    R w_n[3 * W];
    R val_n = R(0);
    auto fetch = [&](int vv) {
        rec_n = a.visits[vv];
        const int sn = rec_n.x >> 3;
        const R* w = a.sw + int64_t(sn) * WS;
#pragma unroll
        for (int q = 0; q < 3 * W; ++q) w_n[q] = w[q];
        val_n = R(0);
        if (active) {
            const V cv = a.csort[int64_t(sn) * a.nbatch + g0 + gi];
            val_n = comp ? cv.y : cv.x;
        }
    };
    if (v < end) fetch(v);
    for (; v < end; v += nwaves) {
        const int4 rec = make_int4(wave_uniform(rec_n.x), wave_uniform(rec_n.y), wave_uniform(rec_n.z),
                                   wave_uniform(rec_n.w));
        R wt[3 * W];
#pragma unroll
        for (int q = 0; q < 3 * W; ++q) wt[q] = wave_uniform(w_n[q]);
        const double val = double(val_n) * scale;
        if (v + nwaves < end) fetch(v + nwaves);
        const int bits = rec.x & 7;
        const int l0[3] = {rec.y, rec.z, rec.w};
        int qlo[3], qhi[3], off[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int wa = a.g.wa[d];
            if (a.g.ntile[d] == 1) {
                qlo[d] = 0; qhi[d] = wa; off[d] = l0[d];
            } else {
                const int r = l0[d] % a.g.T[d];
                const int split = min(wa, a.g.T[d] - r);
                if ((bits >> (2 - d)) & 1) { qlo[d] = split; qhi[d] = wa; off[d] = -split; }
                else { qlo[d] = 0; qhi[d] = split; off[d] = r; }
            }
        }
        if (!active) continue;
        double wxd[W];
#pragma unroll
        for (int q = 0; q < W; ++q) wxd[q] = double(wt[2 * W + q]);
        const bool xwrap = a.g.ntile[2] == 1;   // single tile along x: indices wrap
#pragma unroll
        for (int qz = 0; qz < W; ++qz) {
            if (qz >= qlo[0] && qz < qhi[0]) {
                const double vz = val * double(wt[qz]);
                int lz = off[0] + qz; lz -= lz >= Tz ? Tz : 0;
#pragma unroll
                for (int qy = 0; qy < W; ++qy) {
                    if (qy >= qlo[1] && qy < qhi[1]) {
                        int ly = off[1] + qy; ly -= ly >= Ty ? Ty : 0;
                        const double vzy = vz * double(wt[W + qy]);
                        Acc* row = my + 2 * ((lz * Ty + ly) * Tx);
                        if (!xwrap) {
                            // off + qx < Tx for every qx in range: constant offsets
                            Acc* rq = row + 2 * off[2];
#pragma unroll
                            for (int qx = 0; qx < W; ++qx)
                                if (qx >= qlo[2] && qx < qhi[2])
                                    atomicAdd(rq + 2 * qx, fixed_round(vzy * wxd[qx]));
                        } else {
#pragma unroll
                            for (int qx = 0; qx < W; ++qx) {
                                if (qx >= qlo[2] && qx < qhi[2]) {
                                    int lx = off[2] + qx; lx -= lx >= Tx ? Tx : 0;
                                    atomicAdd(row + 2 * lx, fixed_round(vzy * wxd[qx]));
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    const double inv = 1.0 / scale;
    auto value = [&](int g, int c) {
        return cmake<V>(R(double(static_cast<long long>(lds[g * CS + 2 * c])) * inv),
                        R(double(static_cast<long long>(lds[g * CS + 2 * c + 1])) * inv));
    };
    if (sub.w != 0) {
        // split tile: this block's share goes to its partial slot (summed by
        // reduce_split_tiles_kernel), contiguous per transform
        V* dst = a.partial + (int64_t(sub.w - 1) * a.nbatch + g0) * cells;
        for (int idx = threadIdx.x; idx < ng * cells; idx += blockDim.x) {
            const int g = idx / cells;
            dst[idx] = value(g, idx - g * cells);
        }
        return;
    }
    for (int idx = threadIdx.x; idx < ng * cells; idx += blockDim.x) {
        const int g = idx / cells;
        const int c = idx - g * cells;
        const int lx = c % Tx;
        const int ly = (c / Tx) % Ty;
        const int lz = c / (Tx * Ty);
        V* dst = a.fine + int64_t(g0 + g) * G +
                 (int64_t(t0[0] + lz) * a.g.nf[1] + (t0[1] + ly)) * a.g.nf[2] + (t0[2] + lx);
        *dst = value(g, c);
    }
}

// Sum the partial tiles of every split tile into the fine grid.
// info[k] = (tile, first slot, slot count, -).  Deterministic order.
template <class V>
__global__ void reduce_split_tiles_kernel(V* fine, const V* partial, const int4* info, int nbatch,
                                          Geo g) {
    using R = decltype(V::x);
    const int4 t = info[blockIdx.y];
    const int cells = g.T[0] * g.T[1] * g.T[2];
    const int64_t G = g.nf[0] * g.nf[1] * g.nf[2];
    int t0[3];
    tile_origin(g, t.x, t0);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nbatch * cells;
         idx += gridDim.x * blockDim.x) {
        const int b = idx / cells;
        const int c = idx - b * cells;
        V acc = cmake<V>(R(0), R(0));
        for (int k = 0; k < t.z; ++k) acc = cadd(acc, partial[(int64_t(t.y + k) * nbatch + b) * cells + c]);
        const int lx = c % g.T[2];
        const int ly = (c / g.T[2]) % g.T[1];
        const int lz = c / (g.T[2] * g.T[1]);
        fine[int64_t(b) * G + (int64_t(t0[0] + lz) * g.nf[1] + (t0[1] + ly)) * g.nf[2] + (t0[2] + lx)] = acc;
    }
}

// ---------------------------------------------------------------------------
// Global-memory variants (method GM, or when a tile does not fit in LDS):
// one thread per bin-sorted point reading / atomically adding the fine grid.
// ---------------------------------------------------------------------------
template <class V, int W>
__global__ void __launch_bounds__(256) interp_gm_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.M) return;
    int iy[W], ix[W];
    R wy[W], wx[W];
    const R* w = a.sw + s * a.ws;
    const int l0z = a.sl0[s], l0y = a.sl0[a.M + s], l0x = a.sl0[2 * a.M + s];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        int l = l0y + q; l -= l >= int(a.g.nf[1]) ? int(a.g.nf[1]) : 0; iy[q] = l;
        l = l0x + q; l -= l >= int(a.g.nf[2]) ? int(a.g.nf[2]) : 0; ix[q] = l;
        wy[q] = w[W + q];
        wx[q] = w[2 * W + q];
    }
    const int orig = a.sidx[s];
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    for (int t = 0; t < a.nbatch; ++t) {
        const V* g = a.fine + int64_t(t) * G;
        V acc = cmake<V>(R(0), R(0));
        for (int qz = 0; qz < a.g.wa[0]; ++qz) {
            int z = l0z + qz; z -= z >= int(a.g.nf[0]) ? int(a.g.nf[0]) : 0;
            V accy = cmake<V>(R(0), R(0));
#pragma unroll
            for (int qy = 0; qy < W; ++qy) {
                const V* row = g + (int64_t(z) * a.g.nf[1] + iy[qy]) * a.g.nf[2];
                V accx = cmake<V>(R(0), R(0));
#pragma unroll
                for (int qx = 0; qx < W; ++qx) {
                    const V v = row[ix[qx]];
                    accx.x += wx[qx] * v.x;
                    accx.y += wx[qx] * v.y;
                }
                accy.x += wy[qy] * accx.x;
                accy.y += wy[qy] * accx.y;
            }
            acc.x += w[qz] * accy.x;
            acc.y += w[qz] * accy.y;
        }
        a.cdata[int64_t(t) * a.M + orig] = acc;
    }
}

template <class V, int W>
__global__ void __launch_bounds__(256) spread_gm_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.M) return;
    int iy[W], ix[W];
    R wy[W], wx[W];
    const R* w = a.sw + s * a.ws;
    const int l0z = a.sl0[s], l0y = a.sl0[a.M + s], l0x = a.sl0[2 * a.M + s];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        int l = l0y + q; l -= l >= int(a.g.nf[1]) ? int(a.g.nf[1]) : 0; iy[q] = l;
        l = l0x + q; l -= l >= int(a.g.nf[2]) ? int(a.g.nf[2]) : 0; ix[q] = l;
        wy[q] = w[W + q];
        wx[q] = w[2 * W + q];
    }
    const int orig = a.sidx[s];
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    for (int t = 0; t < a.nbatch; ++t) {
        V* g = a.fine + int64_t(t) * G;
        const V c = a.cdata[int64_t(t) * a.M + orig];
        for (int qz = 0; qz < a.g.wa[0]; ++qz) {
            int z = l0z + qz; z -= z >= int(a.g.nf[0]) ? int(a.g.nf[0]) : 0;
            const R wzq = w[qz];
#pragma unroll
            for (int qy = 0; qy < W; ++qy) {
                V* row = g + (int64_t(z) * a.g.nf[1] + iy[qy]) * a.g.nf[2];
                const R wzy = wzq * wy[qy];
#pragma unroll
                for (int qx = 0; qx < W; ++qx) {
                    const R ww = wzy * wx[qx];
                    unsafeAtomicAdd(&row[ix[qx]].x, ww * c.x);
                    unsafeAtomicAdd(&row[ix[qx]].y, ww * c.y);
                }
            }
        }
    }
}

// csort[s * B + t] = c[t * M + sidx[s]]: the nonuniform data in sorted,
// transform-fastest order for the spreading lanes.
// Also accumulates max(|Re|, |Im|) into *cmax_bits (non-negative floats
// order like their bit patterns).
template <class V>
__global__ void __launch_bounds__(256) gather_sorted_kernel(const V* __restrict__ c, V* __restrict__ csort,
                                                            const int* __restrict__ sidx, int64_t M, int B,
                                                            unsigned* cmax_bits) {
    using R = decltype(V::x);
    __shared__ float red[256];
    const int64_t n = M * B;
    float m = 0.f;
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += int64_t(gridDim.x) * blockDim.x) {
        const int64_t s = k / B;
        const int t = int(k - s * B);
        const V v = c[int64_t(t) * M + sidx[s]];
        csort[k] = v;
        m = fmaxf(m, float(fmax(fabs(v.x), fabs(v.y))));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // round up so the bound stays a bound after the float conversion
        atomicMax(cmax_bits, __float_as_uint(red[0]) + (red[0] > 0.f ? 1u : 0u));
    }
    (void)sizeof(R);
}

template <class V>
__global__ void fill_zero_kernel(V* p, int64_t n) {
    using R = decltype(V::x);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = cmake<V>(R(0), R(0));
}

// Zero the cells of split tiles before their blocks accumulate atomically.
template <class V>
__global__ void zero_tiles_kernel(V* fine, const int* tiles, int nbatch, Geo g) {
    using R = decltype(V::x);
    const int cells = g.T[0] * g.T[1] * g.T[2];
    const int64_t G = g.nf[0] * g.nf[1] * g.nf[2];
    const int tile = tiles[blockIdx.x];
    int t0[3];
    tile_origin(g, tile, t0);
    for (int idx = threadIdx.x; idx < nbatch * cells; idx += blockDim.x) {
        const int b = idx / cells;
        const int c = idx - b * cells;
        const int lx = c % g.T[2];
        const int ly = (c / g.T[2]) % g.T[1];
        const int lz = c / (g.T[2] * g.T[1]);
        fine[int64_t(b) * G + (int64_t(t0[0] + lz) * g.nf[1] + (t0[1] + ly)) * g.nf[2] + (t0[2] + lx)] =
            cmake<V>(R(0), R(0));
    }
}

// which: 0 interp (LDS), 1 spread (LDS), 2 interp (GM), 3 spread (GM)
template <class V> void* nufft_kernel_ptr(int which, int w);

}  // namespace hasty
