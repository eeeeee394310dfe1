// Type-1 / type-2 NUFFT for gfx950: bin sort, LDS-tiled spread and interp,
// pruned Stockham FFT passes with fused deapodisation.
#include <algorithm>
#include <memory>
#include <type_traits>
#include <cmath>
#include <numeric>
#include <vector>

#include "kb.hpp"
#include "nufft.hpp"

namespace hasty {

// ===========================================================================
// Pruned FFT pass sequences
// ===========================================================================
template <class V> ColArgs<V> make_col_args(int L) {
    ColArgs<V> a{};
    a.L = L;
    a.LP = L + 1;
    a.plan = make_fft_plan(L);
    a.tw = reinterpret_cast<const V*>(twiddles(L, sizeof(V) == 16));
    int64_t tx = kPassLdsBudget / (2 * int64_t(a.LP) * int64_t(sizeof(V)));
    a.tx = int(std::max<int64_t>(1, std::min<int64_t>(16, tx)));
    a.psf = nullptr;
    a.scale = 1;
    return a;
}

template <class V> int rows_per_block(int L, int out_len) {
    int64_t nl = kPassLdsBudget / (2 * int64_t(L + 1) * int64_t(sizeof(V)));
    nl = std::min<int64_t>(nl, 16);
    nl = std::min<int64_t>(nl, int64_t(kRowAccMax) * kFFTThreads / std::max(out_len, 1));
    return int(std::max<int64_t>(1, nl));
}

template <class V>
void pruned_forward(const PadGeometry& g, int64_t B, const V* img, V* grid, V* bufA, V* bufB,
                    int sign, const DiagOp<V>& m1, const DiagOp<V>& m2,
                    const decltype(V::x)* const* deapod, hipStream_t s) {
    const int64_t nz = g.n[0], ny = g.n[1], nx = g.n[2];
    const int64_t Lz = g.L[0], Ly = g.L[1], Lx = g.L[2];
    // x: image rows -> bufA [B, nz, ny, Lx]
    RowLoadArgs<V> r{};
    r.in = img; r.out = bufA;
    r.nrows = nz * ny; r.nbatch = int(B);
    r.in_b = nz * ny * nx; r.in_r = nx;
    r.out_b = nz * ny * Lx; r.out_r = Lx;
    r.L = int(Lx); r.LP = int(Lx) + 1; r.in_len = int(nx); r.in_shift = g.shift_in[2];
    r.sign = sign; r.nl = rows_per_block<V>(int(Lx), int(Lx));
    r.m1 = m1; r.m2 = m2;
    r.dz = deapod ? deapod[0] : nullptr; r.dy = deapod ? deapod[1] : nullptr;
    r.dx = deapod ? deapod[2] : nullptr; r.ny = int(ny);
    r.plan = make_fft_plan(int(Lx));
    r.tw = reinterpret_cast<const V*>(twiddles(int(Lx), sizeof(V) == 16));
    r.tag = "fwd_x_pad_fft";
    launch_row_load<V>(r, s);
    // y: bufA -> (3-D) bufB [B, nz, Ly, Lx] or (2-D) grid [B, Ly, Lx]
    ColArgs<V> c = make_col_args<V>(int(Ly));
    c.in = bufA; c.ni = int(Lx); c.no = int(nz); c.nbatch = int(B);
    c.in_s = Lx; c.in_o = ny * Lx; c.in_b = nz * ny * Lx;
    c.in_len = int(ny); c.in_shift = g.shift_in[1];
    c.out = (Lz == 1) ? grid : bufB;
    c.out_s = Lx; c.out_o = Ly * Lx; c.out_b = nz * Ly * Lx;
    c.out_len = int(Ly); c.out_shift = 0; c.sign = sign;
    c.tag = "fwd_y_pad_fft";
    launch_col_pass<V>(c, s);
    if (Lz == 1) return;
    // z: bufB -> grid [B, Lz, Ly, Lx]
    ColArgs<V> z = make_col_args<V>(int(Lz));
    z.in = bufB; z.ni = int(Lx); z.no = int(Ly); z.nbatch = int(B);
    z.in_s = Ly * Lx; z.in_o = Lx; z.in_b = nz * Ly * Lx;
    z.in_len = int(nz); z.in_shift = g.shift_in[0];
    z.out = grid; z.out_s = Ly * Lx; z.out_o = Lx; z.out_b = Lz * Ly * Lx;
    z.out_len = int(Lz); z.out_shift = 0; z.sign = sign;
    z.tag = "fwd_z_pad_fft";
    launch_col_pass<V>(z, s);
}

template <class V>
void pruned_inverse(const PadGeometry& g, int64_t B, const V* grid, V* img, V* bufA, V* bufB,
                    int sign, decltype(V::x) scale, const DiagOp<V>& m1, const DiagOp<V>& m2,
                    const DiagOp<V>& io, const decltype(V::x)* const* deapod, int accumulate,
                    int sum_batch, hipStream_t s) {
    const int64_t nz = g.n[0], ny = g.n[1], nx = g.n[2];
    const int64_t Lz = g.L[0], Ly = g.L[1], Lx = g.L[2];
    const V* ysrc = grid;
    if (Lz > 1) {
        ColArgs<V> z = make_col_args<V>(int(Lz));
        z.in = grid; z.ni = int(Lx); z.no = int(Ly); z.nbatch = int(B);
        z.in_s = Ly * Lx; z.in_o = Lx; z.in_b = Lz * Ly * Lx;
        z.in_len = int(Lz); z.in_shift = 0;
        z.out = bufB; z.out_s = Ly * Lx; z.out_o = Lx; z.out_b = nz * Ly * Lx;
        z.out_len = int(nz); z.out_shift = g.shift_out[0]; z.sign = sign;
        z.tag = "inv_z_fft_crop";
        launch_col_pass<V>(z, s);
        ysrc = bufB;
    }
    ColArgs<V> c = make_col_args<V>(int(Ly));
    c.in = ysrc; c.ni = int(Lx); c.no = int(nz); c.nbatch = int(B);
    c.in_s = Lx; c.in_o = Ly * Lx; c.in_b = nz * Ly * Lx;
    c.in_len = int(Ly); c.in_shift = 0;
    c.out = bufA; c.out_s = Lx; c.out_o = ny * Lx; c.out_b = nz * ny * Lx;
    c.out_len = int(ny); c.out_shift = g.shift_out[1]; c.sign = sign;
    c.tag = "inv_y_fft_crop";
    launch_col_pass<V>(c, s);
    RowStoreArgs<V> r{};
    r.in = bufA; r.out = img;
    r.nrows = nz * ny; r.nbatch = int(B);
    r.in_b = nz * ny * Lx; r.in_r = Lx;
    r.out_b = sum_batch ? 0 : nz * ny * nx; r.out_r = nx;
    r.L = int(Lx); r.LP = int(Lx) + 1; r.out_len = int(nx); r.out_shift = g.shift_out[2];
    r.sign = sign; r.nl = rows_per_block<V>(int(Lx), int(nx));
    r.scale = scale;
    r.m1 = m1; r.m2 = m2; r.io = io;
    r.dz = deapod ? deapod[0] : nullptr; r.dy = deapod ? deapod[1] : nullptr;
    r.dx = deapod ? deapod[2] : nullptr; r.ny = int(ny);
    r.accumulate = accumulate; r.sum_batch = sum_batch;
    r.plan = make_fft_plan(int(Lx));
    r.tw = reinterpret_cast<const V*>(twiddles(int(Lx), sizeof(V) == 16));
    r.tag = "inv_x_fft_crop";
    launch_row_store<V>(r, s);
}

template void pruned_forward<float2>(const PadGeometry&, int64_t, const float2*, float2*, float2*,
                                     float2*, int, const DiagOp<float2>&, const DiagOp<float2>&,
                                     const float* const*, hipStream_t);
template void pruned_forward<double2>(const PadGeometry&, int64_t, const double2*, double2*,
                                      double2*, double2*, int, const DiagOp<double2>&,
                                      const DiagOp<double2>&, const double* const*, hipStream_t);
template void pruned_inverse<float2>(const PadGeometry&, int64_t, const float2*, float2*, float2*,
                                     float2*, int, float, const DiagOp<float2>&,
                                     const DiagOp<float2>&, const DiagOp<float2>&,
                                     const float* const*, int, int, hipStream_t);
template void pruned_inverse<double2>(const PadGeometry&, int64_t, const double2*, double2*,
                                      double2*, double2*, int, double, const DiagOp<double2>&,
                                      const DiagOp<double2>&, const DiagOp<double2>&,
                                      const double* const*, int, int, hipStream_t);
template ColArgs<float2> make_col_args<float2>(int);
template ColArgs<double2> make_col_args<double2>(int);
template int rows_per_block<float2>(int, int);
template int rows_per_block<double2>(int, int);

// ===========================================================================
// Point folding and bin sort
// ===========================================================================
struct Geo {
    int64_t nf[3];
    int T[3];
    int ntile[3];
    int wa[3];        // support width per slot (1 on a degenerate slot)
    double half_w;
};

template <class CT>
struct FoldArgs {
    const CT* coord[3];   // per slot; nullptr on a degenerate slot
    int64_t M;
    Geo g;
};

// u = frac(k / 2pi) * nf in FP64; l0 = ceil(u - w/2); frac = u - l0.
template <class CT>
__device__ inline void fold_point(const FoldArgs<CT>& a, int64_t j, int l0w[3], double fr[3],
                                  int tile3[3]) {
    for (int s = 0; s < 3; ++s) {
        if (a.coord[s] == nullptr) {
            l0w[s] = 0; fr[s] = 0.0; tile3[s] = 0;
            continue;
        }
        const double nf = double(a.g.nf[s]);
        double t = double(a.coord[s][j]) * 0.15915494309189533576888376337251;  // 1/(2 pi)
        t -= floor(t);
        double u = t * nf;
        if (u >= nf) u -= nf;
        const double l0 = ceil(u - a.g.half_w);
        fr[s] = u - l0;
        int64_t lw = int64_t(l0) % a.g.nf[s];
        if (lw < 0) lw += a.g.nf[s];
        l0w[s] = int(lw);
        tile3[s] = int(lw / a.g.T[s]);
    }
}

__device__ inline int tile_id(const Geo& g, int tz, int ty, int tx) {
    return (tz * g.ntile[1] + ty) * g.ntile[2] + tx;
}

// Does the support along slot s reach into the next tile?
__device__ inline int second_tile(const Geo& g, int s, int l0w) {
    if (g.ntile[s] <= 1) return 0;
    return (l0w % g.T[s]) + g.wa[s] > g.T[s] ? 1 : 0;
}

template <class CT>
__global__ void fold_count_kernel(FoldArgs<CT> a, int* pcount, int* vcount) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= a.M) return;
    int l0w[3], t3[3];
    double fr[3];
    fold_point(a, j, l0w, fr, t3);
    atomicAdd(&pcount[tile_id(a.g, t3[0], t3[1], t3[2])], 1);
    const int e0 = second_tile(a.g, 0, l0w[0]), e1 = second_tile(a.g, 1, l0w[1]),
              e2 = second_tile(a.g, 2, l0w[2]);
    for (int bz = 0; bz <= e0; ++bz)
        for (int by = 0; by <= e1; ++by)
            for (int bx = 0; bx <= e2; ++bx)
                atomicAdd(&vcount[tile_id(a.g, (t3[0] + bz) % a.g.ntile[0],
                                          (t3[1] + by) % a.g.ntile[1],
                                          (t3[2] + bx) % a.g.ntile[2])],
                          1);
}

template <class CT, class R>
__global__ void fold_scatter_kernel(FoldArgs<CT> a, int* pcursor, int* vcursor, int* sidx,
                                    int* sl0, R* sfrac, int* visits) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= a.M) return;
    int l0w[3], t3[3];
    double fr[3];
    fold_point(a, j, l0w, fr, t3);
    const int pos = atomicAdd(&pcursor[tile_id(a.g, t3[0], t3[1], t3[2])], 1);
    sidx[pos] = int(j);
    for (int s = 0; s < 3; ++s) {
        sl0[int64_t(s) * a.M + pos] = l0w[s];
        sfrac[int64_t(s) * a.M + pos] = R(fr[s]);
    }
    const int e0 = second_tile(a.g, 0, l0w[0]), e1 = second_tile(a.g, 1, l0w[1]),
              e2 = second_tile(a.g, 2, l0w[2]);
    for (int bz = 0; bz <= e0; ++bz)
        for (int by = 0; by <= e1; ++by)
            for (int bx = 0; bx <= e2; ++bx) {
                const int t = tile_id(a.g, (t3[0] + bz) % a.g.ntile[0],
                                      (t3[1] + by) % a.g.ntile[1], (t3[2] + bx) % a.g.ntile[2]);
                const int v = atomicAdd(&vcursor[t], 1);
                visits[v] = pos * 8 + (bz << 2 | by << 1 | bx);
            }
}

// ===========================================================================
// Interpolation (type 2): one block per (tile subproblem, transform group).
// The block stages the tile's padded footprint [t0, t0 + T + w - 1) of every
// transform in its group into LDS, then each thread gathers w^d cells per
// point from LDS.
// ===========================================================================
template <class V>
struct GridArgs {
    V* fine;                // [B, nf]
    V* cdata;               // [B, M]
    const int* sidx;
    const int* sl0;
    const decltype(V::x)* sfrac;
    const int4* subs;
    int64_t M;
    int nbatch, group;
    Geo g;
    decltype(V::x) beta, inv_half_w;
};

__device__ inline void tile_origin(const Geo& g, int tile, int t0[3]) {
    const int tx = tile % g.ntile[2];
    const int ty = (tile / g.ntile[2]) % g.ntile[1];
    const int tz = tile / (g.ntile[2] * g.ntile[1]);
    t0[0] = tz * g.T[0];
    t0[1] = ty * g.T[1];
    t0[2] = tx * g.T[2];
}

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
    for (int idx = threadIdx.x; idx < ng * cells; idx += blockDim.x) {
        const int gi = idx / cells;
        const int c = idx - gi * cells;
        const int px = c % Px;
        const int py = (c / Px) % Py;
        const int pz = c / (Px * Py);
        int gz = t0[0] + pz; gz -= gz >= a.g.nf[0] ? int(a.g.nf[0]) : 0;
        int gy = t0[1] + py; gy -= gy >= a.g.nf[1] ? int(a.g.nf[1]) : 0;
        int gx = t0[2] + px; gx -= gx >= a.g.nf[2] ? int(a.g.nf[2]) : 0;
        lds[idx] = a.fine[int64_t(g0 + gi) * G + (int64_t(gz) * a.g.nf[1] + gy) * a.g.nf[2] + gx];
    }
    __syncthreads();
    const int wz = a.g.wa[0];
    for (int s = sub.y + threadIdx.x; s < sub.y + sub.z; s += blockDim.x) {
        const int lz = a.sl0[s] - t0[0];
        const int ly = a.sl0[a.M + s] - t0[1];
        const int lx = a.sl0[2 * a.M + s] - t0[2];
        const R fz = a.sfrac[s], fy = a.sfrac[a.M + s], fx = a.sfrac[2 * a.M + s];
        R wy[W], wx[W];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            wy[q] = (a.g.wa[1] == 1) ? R(q == 0) : kb_phi(fy - R(q), a.inv_half_w, a.beta);
            wx[q] = kb_phi(fx - R(q), a.inv_half_w, a.beta);
        }
        V acc[kMaxGroup];
#pragma unroll
        for (int gi = 0; gi < kMaxGroup; ++gi) acc[gi] = cmake<V>(R(0), R(0));
        for (int qz = 0; qz < wz; ++qz) {
            const R wzq = (wz == 1) ? R(1) : kb_phi(fz - R(qz), a.inv_half_w, a.beta);
#pragma unroll
            for (int gi = 0; gi < kMaxGroup; ++gi) {
                if (gi < ng) {
                    const V* base = lds + gi * cells + ((lz + qz) * Py + ly) * Px + lx;
                    V accy = cmake<V>(R(0), R(0));
#pragma unroll
                    for (int qy = 0; qy < W; ++qy) {
                        if (qy < a.g.wa[1]) {
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
                    }
                    acc[gi].x += wzq * accy.x;
                    acc[gi].y += wzq * accy.y;
                }
            }
        }
        const int orig = a.sidx[s];
#pragma unroll
        for (int gi = 0; gi < kMaxGroup; ++gi)
            if (gi < ng) a.cdata[int64_t(g0 + gi) * a.M + orig] = acc[gi];
    }
}

// ===========================================================================
// Spreading (type 1): one block per (tile subproblem, transform group), the
// block owns the tile's T_z x T_y x T_x cells.  Every point whose support
// touches the tile (the tile's visit list) adds its clipped footprint with
// LDS atomics; the tile is then written with plain stores, or with global
// atomics when the tile's visits are split over several blocks (such tiles
// are zeroed first).  No cell outside the tile is touched, so neighbouring
// tiles never race.
// ===========================================================================
template <class V, int W>
__global__ void __launch_bounds__(kNufftThreads) spread_kernel(GridArgs<V> a, const int* visits) {
    using R = decltype(V::x);
    extern __shared__ unsigned char smem_raw[];
    R* lds = reinterpret_cast<R*>(smem_raw);   // [group][Tz][Ty][Tx][2]
    const int4 sub = a.subs[blockIdx.x];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    const int Tz = a.g.T[0], Ty = a.g.T[1], Tx = a.g.T[2];
    const int cells = Tz * Ty * Tx;
    const int g0 = blockIdx.y * a.group;
    const int ng = min(a.group, a.nbatch - g0);
    for (int idx = threadIdx.x; idx < 2 * ng * cells; idx += blockDim.x) lds[idx] = R(0);
    __syncthreads();
    for (int v = sub.y + threadIdx.x; v < sub.y + sub.z; v += blockDim.x) {
        const int rec = visits[v];
        const int s = rec >> 3;
        const int bits = rec & 7;
        // per slot: q range [qlo, qhi) and local offset (local = off + q, wrapped by T)
        int qlo[3], qhi[3], off[3];
        for (int d = 0; d < 3; ++d) {
            const int l0w = a.sl0[int64_t(d) * a.M + s];
            const int wa = a.g.wa[d];
            if (a.g.ntile[d] == 1) {
                qlo[d] = 0; qhi[d] = wa; off[d] = l0w;
            } else {
                const int r = l0w % a.g.T[d];
                const int split = min(wa, a.g.T[d] - r);
                const int second = (bits >> (2 - d)) & 1;
                if (second) { qlo[d] = split; qhi[d] = wa; off[d] = -split; }
                else { qlo[d] = 0; qhi[d] = split; off[d] = r; }
            }
        }
        const R fz = a.sfrac[s], fy = a.sfrac[a.M + s], fx = a.sfrac[2 * a.M + s];
        R wy[W], wx[W];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            wy[q] = (a.g.wa[1] == 1) ? R(q == 0) : kb_phi(fy - R(q), a.inv_half_w, a.beta);
            wx[q] = kb_phi(fx - R(q), a.inv_half_w, a.beta);
        }
        const int orig = a.sidx[s];
        V cv[kMaxGroup];
#pragma unroll
        for (int gi = 0; gi < kMaxGroup; ++gi)
            cv[gi] = gi < ng ? a.cdata[int64_t(g0 + gi) * a.M + orig] : cmake<V>(R(0), R(0));
        for (int qz = qlo[0]; qz < qhi[0]; ++qz) {
            const R wzq = (a.g.wa[0] == 1) ? R(1) : kb_phi(fz - R(qz), a.inv_half_w, a.beta);
            int lz = off[0] + qz; lz -= lz >= Tz ? Tz : 0;
#pragma unroll
            for (int qy = 0; qy < W; ++qy) {
                if (qy >= qlo[1] && qy < qhi[1]) {
                    int ly = off[1] + qy; ly -= ly >= Ty ? Ty : 0;
                    const R wzy = wzq * wy[qy];
#pragma unroll
                    for (int qx = 0; qx < W; ++qx) {
                        if (qx >= qlo[2] && qx < qhi[2]) {
                            int lx = off[2] + qx; lx -= lx >= Tx ? Tx : 0;
                            const R ww = wzy * wx[qx];
                            const int cell = (lz * Ty + ly) * Tx + lx;
#pragma unroll
                            for (int gi = 0; gi < kMaxGroup; ++gi) {
                                if (gi < ng) {
                                    atomicAdd(&lds[2 * (gi * cells + cell)], ww * cv[gi].x);
                                    atomicAdd(&lds[2 * (gi * cells + cell) + 1], ww * cv[gi].y);
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
    const bool split = sub.w != 0;
    for (int idx = threadIdx.x; idx < ng * cells; idx += blockDim.x) {
        const int gi = idx / cells;
        const int c = idx - gi * cells;
        const int lx = c % Tx;
        const int ly = (c / Tx) % Ty;
        const int lz = c / (Tx * Ty);
        V* dst = a.fine + int64_t(g0 + gi) * G +
                 (int64_t(t0[0] + lz) * a.g.nf[1] + (t0[1] + ly)) * a.g.nf[2] + (t0[2] + lx);
        const R re = lds[2 * idx], im = lds[2 * idx + 1];
        if (split) {
            atomicAdd(&dst->x, re);
            atomicAdd(&dst->y, im);
        } else {
            *dst = cmake<V>(re, im);
        }
    }
}

// ===========================================================================
// Global-memory variants (method GM, or when a tile does not fit in LDS):
// one thread per bin-sorted point, reading / atomically adding the fine grid
// directly with wrapped indices.
// ===========================================================================
template <class V, int W>
__global__ void __launch_bounds__(kNufftThreads) interp_gm_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.M) return;
    int iz[W], iy[W], ix[W];
    R wz[W], wy[W], wx[W];
    const R* fr[3] = {a.sfrac, a.sfrac + a.M, a.sfrac + 2 * a.M};
    int* idx[3] = {iz, iy, ix};
    R* wt[3] = {wz, wy, wx};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int l0 = a.sl0[int64_t(d) * a.M + s];
        const R f = fr[d][s];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            int l = l0 + q;
            l -= l >= int(a.g.nf[d]) ? int(a.g.nf[d]) : 0;
            idx[d][q] = l;
            wt[d][q] = a.g.wa[d] == 1 ? R(q == 0) : kb_phi(f - R(q), a.inv_half_w, a.beta);
        }
    }
    const int orig = a.sidx[s];
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    for (int t = 0; t < a.nbatch; ++t) {
        const V* g = a.fine + int64_t(t) * G;
        V acc = cmake<V>(R(0), R(0));
        for (int qz = 0; qz < a.g.wa[0]; ++qz) {
            V accy = cmake<V>(R(0), R(0));
            for (int qy = 0; qy < a.g.wa[1]; ++qy) {
                const V* row = g + (int64_t(iz[qz]) * a.g.nf[1] + iy[qy]) * a.g.nf[2];
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
            acc.x += wz[qz] * accy.x;
            acc.y += wz[qz] * accy.y;
        }
        a.cdata[int64_t(t) * a.M + orig] = acc;
    }
}

template <class V, int W>
__global__ void __launch_bounds__(kNufftThreads) spread_gm_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.M) return;
    int iz[W], iy[W], ix[W];
    R wz[W], wy[W], wx[W];
    const R* fr[3] = {a.sfrac, a.sfrac + a.M, a.sfrac + 2 * a.M};
    int* idx[3] = {iz, iy, ix};
    R* wt[3] = {wz, wy, wx};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int l0 = a.sl0[int64_t(d) * a.M + s];
        const R f = fr[d][s];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            int l = l0 + q;
            l -= l >= int(a.g.nf[d]) ? int(a.g.nf[d]) : 0;
            idx[d][q] = l;
            wt[d][q] = a.g.wa[d] == 1 ? R(q == 0) : kb_phi(f - R(q), a.inv_half_w, a.beta);
        }
    }
    const int orig = a.sidx[s];
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    for (int t = 0; t < a.nbatch; ++t) {
        V* g = a.fine + int64_t(t) * G;
        const V c = a.cdata[int64_t(t) * a.M + orig];
        for (int qz = 0; qz < a.g.wa[0]; ++qz) {
            for (int qy = 0; qy < a.g.wa[1]; ++qy) {
                V* row = g + (int64_t(iz[qz]) * a.g.nf[1] + iy[qy]) * a.g.nf[2];
                const R wzy = wz[qz] * wy[qy];
#pragma unroll
                for (int qx = 0; qx < W; ++qx) {
                    const R ww = wzy * wx[qx];
                    atomicAdd(&row[ix[qx]].x, ww * c.x);
                    atomicAdd(&row[ix[qx]].y, ww * c.y);
                }
            }
        }
    }
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
__global__ void zero_tiles_kernel(V* fine, const int* tiles, int ntiles_split, int nbatch, Geo g) {
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

// ===========================================================================
// Plan
// ===========================================================================
NufftPlan::~NufftPlan() {
    void* ptrs[] = {deapod[0], deapod[1], deapod[2], fine, bufA, bufB, sidx, sl0, sfrac,
                    visits, sub_interp, sub_spread, split_tiles};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

namespace {

template <class T> T* dev_alloc(int64_t count, int64_t& tally) {
    if (count <= 0) return nullptr;
    void* p = nullptr;
    const size_t bytes = size_t(count) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(HASTY_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    tally += int64_t(bytes);
    return static_cast<T*>(p);
}

template <class T> void dev_free(T*& p, int64_t bytes, int64_t& tally) {
    if (p) {
        HASTY_CHECK_HIP(hipFree(p));
        tally -= bytes;
        p = nullptr;
    }
}

int pick_tile(int64_t nf, int target, int w) {
    if (nf <= 1) return 1;
    int best = 0;
    for (int t = std::min<int64_t>(target, nf); t >= 1; --t)
        if (nf % t == 0 && t >= w) { best = t; break; }
    return best ? best : int(nf);
}

Geo geo_of(const NufftPlan* p) {
    Geo g{};
    for (int s = 0; s < 3; ++s) {
        g.nf[s] = p->nf[s];
        g.T[s] = p->T[s];
        g.ntile[s] = p->ntile[s];
        g.wa[s] = p->nf[s] == 1 ? 1 : p->w;
    }
    g.half_w = 0.5 * p->w;
    return g;
}

template <class V> using InterpFn = void (*)(GridArgs<V>);
template <class V> using SpreadFn = void (*)(GridArgs<V>, const int*);

#define HASTY_WIDTH_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)

template <class V, int W> struct KernelTable {
    static void* interp() { return reinterpret_cast<void*>(&interp_kernel<V, W>); }
    static void* spread() { return reinterpret_cast<void*>(&spread_kernel<V, W>); }
    static void* interp_gm() { return reinterpret_cast<void*>(&interp_gm_kernel<V, W>); }
    static void* spread_gm() { return reinterpret_cast<void*>(&spread_gm_kernel<V, W>); }
};

template <class V> void* kernel_for_width(int w, int which) {
    switch (w) {
#define C(Wv)                                                        \
    case Wv:                                                         \
        return which == 0 ? KernelTable<V, Wv>::interp()             \
             : which == 1 ? KernelTable<V, Wv>::spread()             \
             : which == 2 ? KernelTable<V, Wv>::interp_gm()          \
                          : KernelTable<V, Wv>::spread_gm();
        HASTY_WIDTH_CASES(C)
#undef C
    }
    throw Error(HASTY_ERR_INVALID, "kernel width " + std::to_string(w) + " unsupported");
}
template <class V> void* interp_for_width(int w) { return kernel_for_width<V>(w, 0); }
template <class V> void* spread_for_width(int w) { return kernel_for_width<V>(w, 1); }
template <class V> void* interp_gm_for_width(int w) { return kernel_for_width<V>(w, 2); }
template <class V> void* spread_gm_for_width(int w) { return kernel_for_width<V>(w, 3); }

constexpr int kMaxWidth = 14;

size_t interp_lds_bytes(const NufftPlan* p, int group) {
    size_t cells = 1;
    for (int s = 0; s < 3; ++s) cells *= size_t(p->T[s] + (p->nf[s] == 1 ? 1 : p->w) - 1);
    return cells * size_t(group) * (p->f64 ? 16 : 8);
}
size_t spread_lds_bytes(const NufftPlan* p, int group) {
    return size_t(p->T[0]) * p->T[1] * p->T[2] * size_t(group) * (p->f64 ? 16 : 8);
}

void set_lds_limit(void* fn, size_t bytes) {
    HASTY_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}

constexpr size_t kLdsMax = 160 * 1024;

}  // namespace

NufftPlan* nufft_makeplan(int type, int dim, const int64_t* nmodes, int sign, int ntransf,
                          double tol, bool f64, const hasty_nufft_opts* opts, hipStream_t stream) {
    HASTY_REQUIRE(type == 1 || type == 2, "nufft type must be 1 or 2 (type 3 is not supported)");
    HASTY_REQUIRE(dim == 2 || dim == 3, "nufft dim must be 2 or 3");
    HASTY_REQUIRE(nmodes != nullptr, "nmodes is NULL");
    HASTY_REQUIRE(ntransf >= 1, "ntransf must be >= 1");
    HASTY_REQUIRE(sign == 1 || sign == -1, "sign must be +1 or -1");
    HASTY_REQUIRE(tol > 0 && tol < 1, "tol must be in (0, 1)");
    auto p = std::make_unique<NufftPlan>();
    p->type = type; p->dim = dim; p->sign = sign; p->ntransf = ntransf; p->f64 = f64;
    p->tol = tol;
    p->sigma = (opts && opts->upsamp > 0) ? opts->upsamp : 2.0;
    HASTY_REQUIRE(p->sigma >= 1.25, "upsampling factor must be >= 1.25");
    p->w = kb_width_for_tol(tol);
    HASTY_REQUIRE(p->w <= kMaxWidth, "requested tolerance needs a kernel wider than 14");
    p->beta = kb_beta(p->w, p->sigma);
    p->stream = stream;
    p->device = current_device();
    for (int a = 0; a < dim; ++a) {
        HASTY_REQUIRE(nmodes[a] >= 1, "nmodes must be >= 1");
        p->n[3 - dim + a] = nmodes[a];
    }
    for (int s = 3 - dim; s < 3; ++s) {
        p->nf[s] = next_smooth(std::max<int64_t>(int64_t(std::ceil(p->sigma * p->n[s])), 2 * p->w));
        HASTY_REQUIRE(p->nf[s] <= (1 << 20), "fine grid too large");
    }
    // tiles: target 16 cells per axis in 3-D, 32 in 2-D (divisors of nf, >= w)
    const int target = dim == 3 ? 16 : 32;
    for (int s = 0; s < 3; ++s) {
        int want = (opts && opts->tile[s] > 0) ? opts->tile[s] : target;
        p->T[s] = p->nf[s] == 1 ? 1 : pick_tile(p->nf[s], want, p->w);
        p->ntile[s] = int(p->nf[s] / p->T[s]);
    }
    p->ntiles = int64_t(p->ntile[0]) * p->ntile[1] * p->ntile[2];
    int group = (opts && opts->coil_group > 0) ? opts->coil_group : (f64 ? 1 : 2);
    group = std::min(std::min(group, kMaxGroup), ntransf);
    // LDS-tiled kernels when the tile fits (method SM / DEFAULT); the
    // global-memory kernels for method GM_* or when it does not fit.
    const int method = opts ? opts->method : 3;
    auto lds_of = [&](int g) {
        return type == 2 ? interp_lds_bytes(p.get(), g) : spread_lds_bytes(p.get(), g);
    };
    while (group > 1 && lds_of(group) > kLdsMax) --group;
    p->gm = (method == 0 || method == 1) || lds_of(1) > kLdsMax;
    p->group = group;
    p->max_batch = (opts && opts->max_batch > 0) ? std::min(opts->max_batch, ntransf) : ntransf;
    p->max_sub = (opts && opts->max_subproblem > 0) ? opts->max_subproblem : 8192;
    if (!p->gm) {
        void* fn = type == 2 ? (f64 ? interp_for_width<double2>(p->w) : interp_for_width<float2>(p->w))
                             : (f64 ? spread_for_width<double2>(p->w) : spread_for_width<float2>(p->w));
        set_lds_limit(fn, lds_of(group));
    }
    // deapodisation factors 1/Phi(2 pi m / nf) per slot, centred modes
    for (int s = 0; s < 3; ++s) {
        const int64_t n = p->n[s];
        std::vector<double> h(size_t(n), 1.0);
        if (p->nf[s] > 1)
            for (int64_t j = 0; j < n; ++j) {
                const double m = double(j - n / 2);
                h[j] = 1.0 / kb_hat_host(2.0 * M_PI * m / double(p->nf[s]), p->w, p->beta);
            }
        if (f64) {
            double* d = dev_alloc<double>(n, p->device_bytes);
            HASTY_CHECK_HIP(hipMemcpy(d, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
            p->deapod[s] = d;
        } else {
            std::vector<float> hf(h.begin(), h.end());
            float* d = dev_alloc<float>(n, p->device_bytes);
            HASTY_CHECK_HIP(hipMemcpy(d, hf.data(), n * sizeof(float), hipMemcpyHostToDevice));
            p->deapod[s] = d;
        }
    }
    // work buffers for one internal batch
    const int64_t B = p->max_batch;
    const int64_t esz = f64 ? 16 : 8;
    const int64_t nfine = p->nf[0] * p->nf[1] * p->nf[2];
    const int64_t nA = p->n[0] * p->n[1] * p->nf[2];
    const int64_t nB = p->n[0] * p->nf[1] * p->nf[2];
    p->fine = dev_alloc<unsigned char>(B * nfine * esz, p->device_bytes);
    p->bufA = dev_alloc<unsigned char>(B * nA * esz, p->device_bytes);
    p->bufB = dev_alloc<unsigned char>(B * nB * esz, p->device_bytes);
    return p.release();
}

void nufft_setpts_typed(NufftPlan* p, int64_t M, const void* c0, const void* c1, const void* c2,
                        bool coords_f64) {
    HASTY_REQUIRE(M >= 0, "number of points must be >= 0");
    HASTY_REQUIRE(M * 8 < (int64_t(1) << 31), "too many points for one plan (max 2^28)");
    const void* cs[3] = {c0, c1, c2};
    for (int a = 0; a < p->dim; ++a)
        HASTY_REQUIRE(M == 0 || cs[a] != nullptr, "coordinate pointer is NULL");
    p->have_pts = false;
    const hipStream_t s = p->stream;
    const Geo g = geo_of(p);
    p->M = M;
    // grow point storage
    if (M > p->pts_cap) {
        dev_free(p->sidx, p->pts_cap * 4, p->device_bytes);
        dev_free(p->sl0, p->pts_cap * 12, p->device_bytes);
        unsigned char* sf = static_cast<unsigned char*>(p->sfrac);
        dev_free(sf, p->pts_cap * 3 * (p->f64 ? 8 : 4), p->device_bytes);
        p->sfrac = nullptr;
        p->sidx = dev_alloc<int>(M, p->device_bytes);
        p->sl0 = dev_alloc<int>(3 * M, p->device_bytes);
        p->sfrac = dev_alloc<unsigned char>(3 * M * (p->f64 ? 8 : 4), p->device_bytes);
        p->pts_cap = M;
    }
    const int64_t nt = p->ntiles;
    int64_t tally_tmp = 0;
    int* counts = dev_alloc<int>(4 * nt, tally_tmp);   // pcount, vcount, pcursor, vcursor
    HASTY_CHECK_HIP(hipMemsetAsync(counts, 0, size_t(4 * nt) * sizeof(int), s));
    auto run_fold = [&](auto* tag) {
        using CT = std::remove_pointer_t<decltype(tag)>;
        FoldArgs<CT> fa{};
        for (int a = 0; a < 3; ++a) fa.coord[a] = nullptr;
        for (int a = 0; a < p->dim; ++a) fa.coord[3 - p->dim + a] = static_cast<const CT*>(cs[a]);
        fa.M = M;
        fa.g = g;
        return fa;
    };
    const unsigned blocks = unsigned(ceil_div(std::max<int64_t>(M, 1), 256));
    if (M > 0) {
        if (coords_f64) {
            auto fa = run_fold((double*)nullptr);
            hipLaunchKernelGGL(fold_count_kernel<double>, dim3(blocks), dim3(256), 0, s, fa, counts, counts + nt);
        } else {
            auto fa = run_fold((float*)nullptr);
            hipLaunchKernelGGL(fold_count_kernel<float>, dim3(blocks), dim3(256), 0, s, fa, counts, counts + nt);
        }
        HASTY_CHECK_HIP(hipGetLastError());
    }
    std::vector<int> hcount(size_t(2 * nt));
    HASTY_CHECK_HIP(hipMemcpyAsync(hcount.data(), counts, size_t(2 * nt) * sizeof(int),
                                   hipMemcpyDeviceToHost, s));
    HASTY_CHECK_HIP(hipStreamSynchronize(s));
    // offsets and subproblems
    std::vector<int> cursors(size_t(2 * nt));
    std::vector<int4> isubs, ssubs;
    std::vector<int> splits;
    int64_t poff = 0, voff = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const int pc = hcount[t], vc = hcount[nt + t];
        cursors[t] = int(poff);
        cursors[nt + t] = int(voff);
        for (int st = 0; st < pc; st += p->max_sub)
            isubs.push_back(make_int4(int(t), int(poff + st), std::min(p->max_sub, pc - st), 0));
        if (vc <= p->max_sub) {
            ssubs.push_back(make_int4(int(t), int(voff), vc, 0));
        } else {
            splits.push_back(int(t));
            for (int st = 0; st < vc; st += p->max_sub)
                ssubs.push_back(make_int4(int(t), int(voff + st), std::min(p->max_sub, vc - st), 1));
        }
        poff += pc;
        voff += vc;
    }
    HASTY_REQUIRE(poff == M, "internal: bin counts do not sum to M");
    p->nvisits = voff;
    if (voff > p->visits_cap) {
        dev_free(p->visits, p->visits_cap * 4, p->device_bytes);
        p->visits = dev_alloc<int>(voff, p->device_bytes);
        p->visits_cap = voff;
    }
    auto upload_int4 = [&](std::vector<int4>& v, int4*& ptr, int64_t& cap, int& n) {
        n = int(v.size());
        if (int64_t(v.size()) > cap) {
            dev_free(ptr, cap * 16, p->device_bytes);
            ptr = dev_alloc<int4>(int64_t(v.size()), p->device_bytes);
            cap = int64_t(v.size());
        }
        if (!v.empty())
            HASTY_CHECK_HIP(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    };
    upload_int4(isubs, p->sub_interp, p->sub_interp_cap, p->nsub_interp);
    upload_int4(ssubs, p->sub_spread, p->sub_spread_cap, p->nsub_spread);
    p->nsplit = int(splits.size());
    if (int64_t(splits.size()) > p->split_cap) {
        dev_free(p->split_tiles, p->split_cap * 4, p->device_bytes);
        p->split_tiles = dev_alloc<int>(int64_t(splits.size()), p->device_bytes);
        p->split_cap = int64_t(splits.size());
    }
    if (!splits.empty())
        HASTY_CHECK_HIP(hipMemcpyAsync(p->split_tiles, splits.data(), splits.size() * sizeof(int),
                                       hipMemcpyHostToDevice, s));
    HASTY_CHECK_HIP(hipMemcpyAsync(counts + 2 * nt, cursors.data(), size_t(2 * nt) * sizeof(int),
                                   hipMemcpyHostToDevice, s));
    if (M > 0) {
        auto scatter = [&](auto* ctag, auto* rtag) {
            using CT = std::remove_pointer_t<decltype(ctag)>;
            using RT = std::remove_pointer_t<decltype(rtag)>;
            auto fa = run_fold((CT*)nullptr);
            hipLaunchKernelGGL((fold_scatter_kernel<CT, RT>), dim3(blocks), dim3(256), 0, s, fa,
                               counts + 2 * nt, counts + 3 * nt, p->sidx, p->sl0,
                               static_cast<RT*>(p->sfrac), p->visits);
        };
        if (coords_f64 && p->f64) scatter((double*)nullptr, (double*)nullptr);
        else if (coords_f64) scatter((double*)nullptr, (float*)nullptr);
        else if (p->f64) scatter((float*)nullptr, (double*)nullptr);
        else scatter((float*)nullptr, (float*)nullptr);
        HASTY_CHECK_HIP(hipGetLastError());
    }
    // the host vectors above are copied asynchronously: wait before they go
    HASTY_CHECK_HIP(hipStreamSynchronize(s));
    HASTY_CHECK_HIP(hipFree(counts));
    p->have_pts = true;
}

void nufft_setpts(NufftPlan* p, int64_t M, const void* c0, const void* c1, const void* c2) {
    nufft_setpts_typed(p, M, c0, c1, c2, p->f64);
}

namespace {

template <class V>
void execute_impl(NufftPlan* p, V* c, V* f) {
    using R = decltype(V::x);
    const hipStream_t s = p->stream;
    const Geo g = geo_of(p);
    const int64_t nimg = p->n[0] * p->n[1] * p->n[2];
    const int64_t nfine = p->nf[0] * p->nf[1] * p->nf[2];
    PadGeometry pg{};
    for (int a = 0; a < 3; ++a) {
        pg.n[a] = p->n[a];
        pg.L[a] = p->nf[a];
        pg.shift_in[a] = int(-(p->n[a] / 2));
        pg.shift_out[a] = int(-(p->n[a] / 2));
    }
    const R* deapod[3] = {static_cast<const R*>(p->deapod[0]), static_cast<const R*>(p->deapod[1]),
                          static_cast<const R*>(p->deapod[2])};
    V* fine = static_cast<V*>(p->fine);
    V* bufA = static_cast<V*>(p->bufA);
    V* bufB = static_cast<V*>(p->bufB);
    for (int b0 = 0; b0 < p->ntransf; b0 += p->max_batch) {
        const int nb = std::min(p->max_batch, p->ntransf - b0);
        V* cb = c + int64_t(b0) * p->M;
        V* fb = f + int64_t(b0) * nimg;
        GridArgs<V> ga{};
        ga.fine = fine; ga.cdata = cb; ga.sidx = p->sidx; ga.sl0 = p->sl0;
        ga.sfrac = static_cast<const R*>(p->sfrac);
        ga.M = p->M; ga.nbatch = nb; ga.group = p->group; ga.g = g;
        ga.beta = R(p->beta); ga.inv_half_w = R(2.0 / p->w);
        const unsigned ngroups = unsigned(ceil_div(nb, p->group));
        if (p->type == 2) {
            DiagOp<V> none{};
            pruned_forward<V>(pg, nb, fb, fine, bufA, bufB, p->sign, none, none, deapod, s);
            if (p->gm && p->M > 0) {
                void* fn = interp_gm_for_width<V>(p->w);
                HASTY_PROF("nufft_interp_gm", s);
                hipLaunchKernelGGL(reinterpret_cast<InterpFn<V>>(fn), dim3(unsigned(ceil_div(p->M, kNufftThreads))),
                                   dim3(kNufftThreads), 0, s, ga);
                HASTY_CHECK_HIP(hipGetLastError());
            } else if (p->nsub_interp > 0) {
                ga.subs = p->sub_interp;
                void* fn = interp_for_width<V>(p->w);
                HASTY_PROF("nufft_interp", s);
                hipLaunchKernelGGL(reinterpret_cast<InterpFn<V>>(fn), dim3(unsigned(p->nsub_interp), ngroups),
                                   dim3(kNufftThreads), interp_lds_bytes(p, p->group), s, ga);
                HASTY_CHECK_HIP(hipGetLastError());
            }
        } else {
            if (p->gm) {
                HASTY_PROF("nufft_spread_gm", s);
                hipLaunchKernelGGL(fill_zero_kernel<V>, dim3(unsigned(std::min<int64_t>(8192, ceil_div(nb * nfine, 256)))),
                                   dim3(256), 0, s, fine, int64_t(nb) * nfine);
                if (p->M > 0) {
                    void* fn = spread_gm_for_width<V>(p->w);
                    hipLaunchKernelGGL(reinterpret_cast<InterpFn<V>>(fn), dim3(unsigned(ceil_div(p->M, kNufftThreads))),
                                       dim3(kNufftThreads), 0, s, ga);
                }
                HASTY_CHECK_HIP(hipGetLastError());
            } else {
            if (p->nsplit > 0) {
                hipLaunchKernelGGL(zero_tiles_kernel<V>, dim3(unsigned(p->nsplit)), dim3(256), 0, s, fine,
                                   p->split_tiles, p->nsplit, nb, g);
                HASTY_CHECK_HIP(hipGetLastError());
            }
            ga.subs = p->sub_spread;
            void* fn = spread_for_width<V>(p->w);
            {
            HASTY_PROF("nufft_spread", s);
            hipLaunchKernelGGL(reinterpret_cast<SpreadFn<V>>(fn), dim3(unsigned(p->nsub_spread), ngroups),
                               dim3(kNufftThreads), spread_lds_bytes(p, p->group), s, ga,
                               static_cast<const int*>(p->visits));
            HASTY_CHECK_HIP(hipGetLastError());
            }
            }
            DiagOp<V> none{};
            pruned_inverse<V>(pg, nb, fine, fb, bufA, bufB, p->sign, R(1), none, none, none, deapod,
                              0, 0, s);
        }
    }
}

}  // namespace

void nufft_execute(NufftPlan* p, void* c, void* f) {
    if (!p->have_pts) throw Error(HASTY_ERR_STATE, "nufft execute called before setpts");
    HASTY_REQUIRE(f != nullptr, "mode array is NULL");
    HASTY_REQUIRE(c != nullptr || p->M == 0, "nonuniform array is NULL");
    if (p->f64) execute_impl<double2>(p, static_cast<double2*>(c), static_cast<double2*>(f));
    else execute_impl<float2>(p, static_cast<float2*>(c), static_cast<float2*>(f));
}

}  // namespace hasty
