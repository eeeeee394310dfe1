// Type-1 / type-2 NUFFT for gfx950: bin sort, LDS-tiled spread and interp,
// pruned Stockham FFT passes with fused deapodisation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <type_traits>
#include <cmath>
#include <numeric>
#include <vector>

#include "kb.hpp"
#include "nufft.hpp"
#include "toeplitz_fused.hpp"
#include "nufft_kernels.hpp"

#include <hipcub/hipcub.hpp>

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

// The x and y passes of 3-D c64 plans with 128 x 128 planes on a 256 x 256
// fine grid run as one fused kernel (nufft_fused.hip), forward and inverse
// (the inverse without a batch sum); HASTY_NUFFT_FUSED=0 keeps the separate
// passes.
static bool nufft_fused_enabled() {
    static const bool v = [] {
        const char* e = std::getenv("HASTY_NUFFT_FUSED");
        return !(e && e[0] == '0');
    }();
    return v;
}

template <class V>
void pruned_forward(const PadGeometry& g, int64_t B, const V* img, V* grid, V* bufA, V* bufB,
                    int sign, const DiagOp<V>& m1, const DiagOp<V>& m2,
                    const decltype(V::x)* const* deapod, hipStream_t s) {
    const int64_t nz = g.n[0], ny = g.n[1], nx = g.n[2];
    const int64_t Lz = g.L[0], Ly = g.L[1], Lx = g.L[2];
    bool fused_xy = false;
    if constexpr (std::is_same_v<V, float2>) {
        fused_xy = Lz > 1 && B >= 1 && nufft_fused_enabled() && !fft_force_stockham() &&
                   nufft_fwd_xy_supported(ny, nx, Ly, Lx, g.shift_in[1], g.shift_in[2]) &&
                   (!deapod || (deapod[0] && deapod[1] && deapod[2]));
        if (fused_xy) {
            NufftFwdXyArgs f{};
            f.in = img; f.in_b = nz * ny * nx;
            f.m1 = m1; f.m2 = m2;
            f.dz = deapod ? deapod[0] : nullptr; f.dy = deapod ? deapod[1] : nullptr;
            f.dx = deapod ? deapod[2] : nullptr;
            f.grid = bufB; f.grid_b = nz * Ly * Lx;
            f.nbatch = int(B); f.nz = int(nz); f.sign = sign;
            launch_nufft_fwd_xy(f, s);
        }
    }
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
    if (!fused_xy) launch_row_load<V>(r, s);
    // y: bufA -> (3-D) bufB [B, nz, Ly, Lx] or (2-D) grid [B, Ly, Lx]
    ColArgs<V> c = make_col_args<V>(int(Ly));
    c.in = bufA; c.ni = int(Lx); c.no = int(nz); c.nbatch = int(B);
    c.in_s = Lx; c.in_o = ny * Lx; c.in_b = nz * ny * Lx;
    c.in_len = int(ny); c.in_shift = g.shift_in[1];
    c.out = (Lz == 1) ? grid : bufB;
    c.out_s = Lx; c.out_o = Ly * Lx; c.out_b = nz * Ly * Lx;
    c.out_len = int(Ly); c.out_shift = 0; c.sign = sign;
    c.tag = "fwd_y_pad_fft";
    if (!fused_xy) launch_col_pass<V>(c, s);
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
        // the y and x passes as one fused kernel (nufft_fused.hip) for 3-D
        // c64 plans with 128 x 128 planes on a 256 x 256 fine grid, without
        // a batch sum
        if constexpr (std::is_same_v<V, float2>) {
            if (!sum_batch && nufft_fused_enabled() && !fft_force_stockham() &&
                nufft_inv_xy_supported(ny, nx, Ly, Lx, g.shift_out[1], g.shift_out[2]) &&
                (!deapod || (deapod[0] && deapod[1] && deapod[2]))) {
                NufftInvXyArgs f{};
                f.in = bufB; f.in_b = nz * Ly * Lx;
                f.out = img; f.out_b = nz * ny * nx;
                f.scale = scale;
                f.m1 = m1; f.m2 = m2; f.io = io;
                f.dz = deapod ? deapod[0] : nullptr; f.dy = deapod ? deapod[1] : nullptr;
                f.dx = deapod ? deapod[2] : nullptr;
                f.accumulate = accumulate;
                f.nbatch = int(B); f.nz = int(nz); f.sign = sign;
                launch_nufft_inv_xy(f, s);
                return;
            }
        }
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
    // in 3-D bufB is free once the y pass has read it (scratch for a batch sum)
    r.scratch = Lz > 1 ? bufB : nullptr; r.scratch_elems = B * nz * Ly * Lx;
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

// Sort key of a point: its tile, then its first support cell inside the
// tile (so the points of a tile are grouped by cell).
__device__ inline int64_t cell_key(const Geo& g, const int l0w[3], const int t3[3]) {
    const int64_t cpt = int64_t(g.T[0]) * g.T[1] * g.T[2];
    const int lz = l0w[0] - t3[0] * g.T[0], ly = l0w[1] - t3[1] * g.T[1], lx = l0w[2] - t3[2] * g.T[2];
    return int64_t(tile_id(g, t3[0], t3[1], t3[2])) * cpt + (int64_t(lz) * g.T[1] + ly) * g.T[2] + lx;
}

template <class CT>
__global__ void fold_count_kernel(FoldArgs<CT> a, int* ccount, int* vcount) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= a.M) return;
    int l0w[3], t3[3];
    double fr[3];
    fold_point(a, j, l0w, fr, t3);
    atomicAdd(&ccount[cell_key(a.g, l0w, t3)], 1);
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

// Sort keys of the points in caller order (cell key, then the point index as
// the value of a stable radix sort: the sorted order is the same on every
// call for the same trajectory).
template <class CT>
__global__ void fold_key_kernel(FoldArgs<CT> a, unsigned* keys, int* idx) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= a.M) return;
    int l0w[3], t3[3];
    double fr[3];
    fold_point(a, j, l0w, fr, t3);
    keys[j] = unsigned(cell_key(a.g, l0w, t3));
    idx[j] = int(j);
}

// Per sorted point s: support start, fractional offset, window weights
// (FP64 evaluation); per original point j the number of tiles its support
// touches.
template <class CT, class R>
__global__ void fold_sorted_kernel(FoldArgs<CT> a, const int* sidx, int* sinv, int* sl0, R* sfrac, R* sw,
                                   int w, int ws, double beta, int* vcnt) {
    const int64_t pos = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pos >= a.M) return;
    const int64_t j = sidx[pos];
    sinv[j] = int(pos);
    int l0w[3], t3[3];
    double fr[3];
    fold_point(a, j, l0w, fr, t3);
    const double ihw = 2.0 / double(w);
    for (int s = 0; s < 3; ++s) {
        sl0[int64_t(s) * a.M + pos] = l0w[s];
        sfrac[int64_t(s) * a.M + pos] = R(fr[s]);
        R* wp = sw + pos * ws + s * w;
        for (int q = 0; q < w; ++q)
            wp[q] = a.g.wa[s] == 1 ? R(q == 0 ? 1.0 : 0.0) : R(kb_phi<double>(fr[s] - double(q), ihw, beta));
    }
    for (int q = 3 * w; q < ws; ++q) sw[pos * ws + q] = R(0);
    vcnt[j] = (1 + second_tile(a.g, 0, l0w[0])) * (1 + second_tile(a.g, 1, l0w[1])) *
              (1 + second_tile(a.g, 2, l0w[2]));
}

// The (tile, 8 s + neighbour bits) pair of every visit, in emission order:
// thread i emits point j = (i * stride) mod M at vpos[i] (vpos = exclusive
// scan of the counts in that order, permute_counts_kernel); a stable sort by
// tile then groups them, keeping that order inside each tile.
__global__ void visit_emit_kernel(const int* sl0, const int* sinv, int64_t M, Geo g, const int* vpos,
                                  unsigned* vkey, int* vval, int64_t stride) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t j = (i * stride) % M;   // emission order (see visit_stride)
    const int64_t s = sinv[j];
    int l0w[3], t3[3];
    for (int d = 0; d < 3; ++d) {
        l0w[d] = sl0[int64_t(d) * M + s];
        t3[d] = l0w[d] / g.T[d];
    }
    const int e0 = second_tile(g, 0, l0w[0]), e1 = second_tile(g, 1, l0w[1]), e2 = second_tile(g, 2, l0w[2]);
    int v = vpos[i];
    for (int bz = 0; bz <= e0; ++bz)
        for (int by = 0; by <= e1; ++by)
            for (int bx = 0; bx <= e2; ++bx) {
                vkey[v] = unsigned(tile_id(g, (t3[0] + bz) % g.ntile[0], (t3[1] + by) % g.ntile[1],
                                           (t3[2] + bx) % g.ntile[2]));
                vval[v] = int(s) * 8 + (bz << 2 | by << 1 | bx);
                ++v;
            }
}

// counts in emission order: out[i] = vcnt[(i * stride) mod M]
__global__ void permute_counts_kernel(const int* vcnt, int* out, int64_t M, int64_t stride) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < M) out[i] = vcnt[(i * stride) % M];
}

// Emission order of the visits: point j = (i * stride) mod M is emitted i-th,
// with stride the integer nearest 0.618 M that is coprime to M (a golden-ratio
// low-discrepancy permutation).  After the stable sort by tile, the visit list
// of a tile then interleaves points from all over the acquisition instead of
// runs of neighbouring samples along one spoke, so each chunk of visits is
// spread evenly over the tile's sub-blocks.  The order is fixed for a given M.
static int64_t visit_stride(int64_t M) {
    if (M <= 1) return 1;
    int64_t P = std::max<int64_t>(1, int64_t(0.6180339887498949 * double(M)));
    while (std::gcd(P, M) != 1) ++P;
    return P;
}

__global__ void visit_records_kernel(const int* vval, const int* sl0, int64_t M, int64_t nvis, int4* visits) {
    const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nvis) return;
    const int val = vval[v];
    const int64_t s = val >> 3;
    visits[v] = make_int4(val, sl0[s], sl0[M + s], sl0[2 * M + s]);
}

// Significant bits of keys below n.
inline int key_bits(int64_t n) {
    int b = 1;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

// Interpolation reads a point's window as 16-B LDS values at fixed offsets
// from its first cell, so the bank quad of every read is fixed by
// class = (padded-tile index of the first cell) mod 16, and 16 lanes whose
// classes differ read conflict-free (measured: 214 against 89 B/clk per CU
// for random classes).  Within each interpolation subproblem this puts the
// sorted points in (rank within class, class) order: the k-th point of every
// class, then the (k+1)-th, ... so runs of up to 16 consecutive points (the
// lanes of a quarter wave) have distinct classes.  Ranks follow the sorted
// order, so the result is the same on every call.  One block per subproblem;
// sidx_in / keys are the sorted order, sidx_out receives the new one.
constexpr int kClassThreads = 1024;
__global__ void __launch_bounds__(kClassThreads)
    interleave_classes_kernel(const int4* subs, const unsigned* keys, const int* sidx_in, int* sidx_out,
                              int* rank, int64_t cpt, int T1, int T2, int Py, int Px) {
    __shared__ int cnt[16];
    __shared__ int wave_cnt[kClassThreads / 64][16];
    const int4 sub = subs[blockIdx.x];
    const int first = sub.y, count = sub.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto class_of = [&](int pos) {
        const int64_t c = int64_t(keys[pos]) % cpt;
        const int lx = int(c % T2), ly = int((c / T2) % T1), lz = int(c / (int64_t(T1) * T2));
        return ((lz * Py + ly) * Px + lx) & 15;
    };
    if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
    __syncthreads();
    // stable rank of each point within its class, chunk by chunk
    for (int base = 0; base < count; base += kClassThreads) {
        const int i = base + int(threadIdx.x);
        const int cl = i < count ? class_of(first + i) : -1;
        int below = 0;
        for (int c = 0; c < 16; ++c) {
            const unsigned long long m = __ballot(cl == c);
            if (cl == c) below = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wave_cnt[wave][c] = __popcll(m);
        }
        __syncthreads();
        if (cl >= 0) {
            int r = cnt[cl] + below;
            for (int w = 0; w < wave; ++w) r += wave_cnt[w][cl];
            rank[first + i] = r;
        }
        __syncthreads();
        if (threadIdx.x < 16) {
            int t = 0;
            for (int w = 0; w < kClassThreads / 64; ++w) t += wave_cnt[w][threadIdx.x];
            cnt[threadIdx.x] += t;
        }
        __syncthreads();
    }
    // new position of the rank-k point of class c: all points of lower rank,
    // sum_c' min(cnt[c'], k), then the classes below c that have a rank-k point
    for (int i = int(threadIdx.x); i < count; i += kClassThreads) {
        const int cl = class_of(first + i), k = rank[first + i];
        int pos = 0;
        for (int c = 0; c < 16; ++c) pos += min(cnt[c], k) + (c < cl && cnt[c] > k ? 1 : 0);
        sidx_out[first + pos] = sidx_in[first + i];
    }
}

// toff[t] = first sorted point of tile t (t <= ntiles) from the cell offsets.
__global__ void tile_offsets_kernel(const int* coff, int64_t cpt, int64_t ntiles, int M, int* toff) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < ntiles) toff[t] = coff[t * cpt];
    if (t == ntiles) toff[t] = M;
}

// ===========================================================================
// Plan
// ===========================================================================
NufftPlan::~NufftPlan() {
    void* ptrs[] = {deapod[0], deapod[1], deapod[2], fine, bufA, bufB, sidx, sinv, sl0, sfrac, sw,
                    csort, visits, sub_interp, sub_spread, split_info, partial};
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

void dev_free_void(void*& p, int64_t bytes, int64_t& tally) {
    if (p) {
        HASTY_CHECK_HIP(hipFree(p));
        tally -= bytes;
        p = nullptr;
    }
}

// floats per point in the weight table: 3 slots x w, padded to 16 bytes
int weight_stride(int w) { return (3 * w + 3) / 4 * 4; }

// Tile edge along one axis: the largest divisor of nf that is <= target and
// >= w (a support then touches at most two tiles), else the smallest divisor
// >= w.
int pick_tile(int64_t nf, int target, int w) {
    if (nf <= 1) return 1;
    for (int t = int(std::min<int64_t>(target, nf)); t >= w; --t)
        if (nf % t == 0) return t;
    for (int64_t t = std::max(w, 1); t < nf; ++t)
        if (nf % t == 0) return int(t);
    return int(nf);
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

size_t interp_lds_bytes(const NufftPlan* p, int group) {
    size_t cells = 1;
    for (int s = 0; s < 3; ++s) cells *= size_t(p->T[s] + (p->nf[s] == 1 ? 1 : p->w) - 1);
    return cells * size_t(group) * (p->f64 ? 16 : 8);
}

void set_lds_limit(void* fn, size_t bytes) {
    if (bytes > 64 * 1024)
        HASTY_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}

constexpr size_t kLdsMax = 160 * 1024;

void* kernel_ptr(const NufftPlan* p, int which) {
    return p->f64 ? nufft_kernel_ptr<double2>(which, p->w) : nufft_kernel_ptr<float2>(which, p->w);
}

// The interpolation kernel for this plan: the compile-time-tile variant when
// the plan uses the standard tile (16^3 in 3-D, 32^2 in 2-D).
void* interp_ptr(const NufftPlan* p) {
    const bool std3 = p->dim == 3 && p->T[0] == 16 && p->T[1] == 16 && p->T[2] == 16;
    const bool std2 = p->dim == 2 && p->T[0] == 1 && p->T[1] == 32 && p->T[2] == 32;
    if (std3 || std2) {
        void* fn = p->f64 ? nufft_interp_fixed_ptr<double2>(p->dim, p->w)
                          : nufft_interp_fixed_ptr<float2>(p->dim, p->w);
        if (fn) return fn;
    }
    return kernel_ptr(p, 0);
}

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
    p->w = kb_width_for(tol, p->sigma, dim, f64);
    if (p->w > (f64 ? kMaxWidthF64 : kMaxWidthF32)) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "tolerance %.3g at upsampling %.3g needs a kernel wider than the %d cells "
                      "%s plans support",
                      tol, p->sigma, f64 ? kMaxWidthF64 : kMaxWidthF32, f64 ? "complex128" : "complex64");
        throw Error(HASTY_ERR_INVALID, msg);
    }
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
    // tiles: interpolation stages a padded 16^3 / 32^2 tile per transform;
    // spreading owns an unpadded 8^3 / 16^2 tile of 64-bit accumulators for a
    // group of transforms (16 / 32 of them fit in 128 KiB of LDS)
    const int target3[3] = {type == 2 ? 16 : 8, type == 2 ? 16 : 8, type == 2 ? 16 : 8};
    const int target2 = type == 2 ? 32 : 16;
    for (int s = 0; s < 3; ++s) {
        const int target = dim == 3 ? target3[s] : target2;
        int want = (opts && opts->tile[s] > 0) ? opts->tile[s] : target;
        p->T[s] = p->nf[s] == 1 ? 1 : pick_tile(p->nf[s], want, p->w);
        p->ntile[s] = int(p->nf[s] / p->T[s]);
    }
    p->ntiles = int64_t(p->ntile[0]) * p->ntile[1] * p->ntile[2];
    const int method = opts ? opts->method : HASTY_METHOD_DEFAULT;
    HASTY_REQUIRE(method >= HASTY_METHOD_GM_NO_SORT && method <= HASTY_METHOD_DEFAULT,
                  "Invalid nufft_cuda_method");
    const bool gm_method = method == HASTY_METHOD_GM_NO_SORT || method == HASTY_METHOD_GM_SORT;
    if (type == 2) {
        // one transform per block by default: a second staged tile halves
        // the occupancy, which costs more than re-reading the point data
        int group = (opts && opts->coil_group > 0) ? opts->coil_group : 1;
        group = std::min(std::min(group, kMaxGroupInterp), ntransf);
        while (group > 1 && interp_lds_bytes(p.get(), group) > kLdsMax) --group;
        p->group = group;
        p->gm = gm_method || interp_lds_bytes(p.get(), 1) > kLdsMax;
    } else {
        // transforms per spreading block: a power of two covering the batch
        // (register accumulators per lane), capped at 32 (16 for complex128)
        const int want = (opts && opts->coil_group > 0) ? opts->coil_group : ntransf;
        const int cap = f64 ? kMaxGroupSpread / 2 : kMaxGroupSpread;
        int group = 1;
        while (group < std::min(want, cap)) group *= 2;
        p->group = group;
        p->gm = gm_method;
    }
    p->max_batch = (opts && opts->max_batch > 0) ? std::min(opts->max_batch, ntransf) : ntransf;
    p->max_sub = (opts && opts->max_subproblem > 0) ? opts->max_subproblem : (type == 2 ? 8192 : 2048);
    if (!p->gm && type == 2) set_lds_limit(interp_ptr(p.get()), interp_lds_bytes(p.get(), p->group));
    // complex64 plans with a batch and the standard tile interpolate two
    // transforms per block (16-B LDS reads), when the pair's tile fits in LDS
    // and no transform group was asked for (HASTY_INTERP_PAIR=0 turns it off)
    if (!p->gm && type == 2 && !f64 && ntransf >= 2 && !(opts && opts->coil_group > 0)) {
        static const char* env = std::getenv("HASTY_INTERP_PAIR");
        const bool std3 = dim == 3 && p->T[0] == 16 && p->T[1] == 16 && p->T[2] == 16;
        void* fn = std3 ? nufft_interp_pair_ptr(dim, p->w) : nullptr;
        // (the kernel addresses one transform's fine grid with 32-bit byte offsets)
        const bool small_grid = p->nf[0] * p->nf[1] * p->nf[2] * 8 <= int64_t(UINT32_MAX);
        if (fn && small_grid && interp_lds_bytes(p.get(), 2) <= kLdsMax && !(env && env[0] == '0')) {
            p->pair = true;
            set_lds_limit(fn, interp_lds_bytes(p.get(), 2));
        }
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
    const int64_t rsz = p->f64 ? 8 : 4;
    p->M = M;
    if (M > p->pts_cap) {
        dev_free(p->sidx, p->pts_cap * 4, p->device_bytes);
        dev_free(p->sinv, p->pts_cap * 4, p->device_bytes);
        dev_free(p->sl0, p->pts_cap * 12, p->device_bytes);
        dev_free_void(p->sfrac, p->pts_cap * 3 * rsz, p->device_bytes);
        dev_free_void(p->sw, p->pts_cap * weight_stride(p->w) * rsz, p->device_bytes);
        p->sidx = dev_alloc<int>(M, p->device_bytes);
        p->sinv = dev_alloc<int>(M, p->device_bytes);
        p->sl0 = dev_alloc<int>(3 * M, p->device_bytes);
        p->sfrac = dev_alloc<unsigned char>(3 * M * rsz, p->device_bytes);
        p->sw = dev_alloc<unsigned char>(M * weight_stride(p->w) * rsz, p->device_bytes);
        p->pts_cap = M;
    }
    const int64_t nt = p->ntiles;
    const int64_t cpt = int64_t(p->T[0]) * p->T[1] * p->T[2];
    const int64_t nkeys = nt * cpt;                      // = fine-grid cells
    HASTY_REQUIRE(nkeys <= (int64_t(1) << 32), "fine grid too large for one plan (2^32 cells)");
    int64_t tally_tmp = 0;
    // ccount[nkeys] -> (scan) coff[nkeys] -> tile offsets toff [nt + 1];
    // vcount [nt] (visits per tile)
    int* ccount = dev_alloc<int>(nkeys, tally_tmp);
    int* coff = dev_alloc<int>(nkeys, tally_tmp);
    int* vbuf = dev_alloc<int>(2 * nt, tally_tmp);
    int* toff = dev_alloc<int>(nt + 1, tally_tmp);
    HASTY_CHECK_HIP(hipMemsetAsync(ccount, 0, size_t(nkeys) * sizeof(int), s));
    HASTY_CHECK_HIP(hipMemsetAsync(vbuf, 0, size_t(2 * nt) * sizeof(int), s));
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
            hipLaunchKernelGGL(fold_count_kernel<double>, dim3(blocks), dim3(256), 0, s, fa, ccount, vbuf);
        } else {
            auto fa = run_fold((float*)nullptr);
            hipLaunchKernelGGL(fold_count_kernel<float>, dim3(blocks), dim3(256), 0, s, fa, ccount, vbuf);
        }
        HASTY_CHECK_HIP(hipGetLastError());
    }
    {
        size_t tmp_bytes = 0;
        HASTY_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, ccount, coff, nkeys, s));
        void* tmp = dev_alloc<unsigned char>(int64_t(std::max<size_t>(tmp_bytes, 1)), tally_tmp);
        HASTY_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ccount, coff, nkeys, s));
        hipLaunchKernelGGL(tile_offsets_kernel, dim3(unsigned(ceil_div(nt + 1, 256))), dim3(256), 0, s,
                           static_cast<const int*>(coff), cpt, nt, int(M), toff);
        HASTY_CHECK_HIP(hipGetLastError());
        HASTY_CHECK_HIP(hipStreamSynchronize(s));
        HASTY_CHECK_HIP(hipFree(tmp));
    }
    std::vector<int> htoff(static_cast<size_t>(nt + 1)), hvcount(static_cast<size_t>(nt));
    HASTY_CHECK_HIP(hipMemcpyAsync(htoff.data(), toff, size_t(nt + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    HASTY_CHECK_HIP(hipMemcpyAsync(hvcount.data(), vbuf, size_t(nt) * sizeof(int), hipMemcpyDeviceToHost, s));
    HASTY_CHECK_HIP(hipStreamSynchronize(s));
    // subproblems; split tiles get one partial slot per subproblem
    std::vector<int4> isubs, ssubs, splits;
    int64_t voff = 0;
    int nslots = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const int poff = htoff[t], pc = htoff[t + 1] - htoff[t], vc = hvcount[t];
        for (int st = 0; st < pc; st += p->max_sub)
            isubs.push_back(make_int4(int(t), poff + st, std::min(p->max_sub, pc - st), 0));
        if (vc <= p->max_sub) {
            ssubs.push_back(make_int4(int(t), int(voff), vc, 0));
        } else {
            const int first = nslots;
            for (int st = 0; st < vc; st += p->max_sub)
                ssubs.push_back(make_int4(int(t), int(voff + st), std::min(p->max_sub, vc - st), 1 + nslots++));
            splits.push_back(make_int4(int(t), first, nslots - first, 0));
        }
        voff += vc;
    }
    HASTY_REQUIRE(htoff[nt] == M, "internal: bin counts do not sum to M");
    p->nvisits = voff;
    p->nslots = nslots;
    if (voff > p->visits_cap) {
        dev_free(p->visits, p->visits_cap * 16, p->device_bytes);
        p->visits = dev_alloc<int4>(voff, p->device_bytes);
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
    upload_int4(splits, p->split_info, p->split_cap, p->nsplit);
    if (M > 0) {
        // 1. stable radix sort of (cell key, point index): the sorted order
        unsigned* keys_in = dev_alloc<unsigned>(M, tally_tmp);
        unsigned* keys_out = dev_alloc<unsigned>(M, tally_tmp);
        int* idx_in = dev_alloc<int>(M, tally_tmp);
        int* vcnt = dev_alloc<int>(M, tally_tmp);
        int* vpos = dev_alloc<int>(M, tally_tmp);
        if (coords_f64) {
            auto fa = run_fold((double*)nullptr);
            hipLaunchKernelGGL(fold_key_kernel<double>, dim3(blocks), dim3(256), 0, s, fa, keys_in, idx_in);
        } else {
            auto fa = run_fold((float*)nullptr);
            hipLaunchKernelGGL(fold_key_kernel<float>, dim3(blocks), dim3(256), 0, s, fa, keys_in, idx_in);
        }
        HASTY_CHECK_HIP(hipGetLastError());
        auto sort_pairs = [&](const unsigned* kin, unsigned* kout, const int* vin, int* vout, int64_t n, int bits) {
            size_t tb = 0;
            HASTY_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, int(n), 0, bits, s));
            void* tmp = dev_alloc<unsigned char>(int64_t(std::max<size_t>(tb, 1)), tally_tmp);
            HASTY_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin, vout, int(n), 0, bits, s));
            HASTY_CHECK_HIP(hipStreamSynchronize(s));
            HASTY_CHECK_HIP(hipFree(tmp));
        };
        sort_pairs(keys_in, keys_out, idx_in, p->sidx, M, key_bits(nkeys));
        // 1b. interpolation with 16-B cells (paired complex64, complex128) on
        // the standard tile: bank-class interleaving inside each subproblem
        // (HASTY_INTERP_INTERLEAVE=0 keeps the cell order)
        static const char* env_il = std::getenv("HASTY_INTERP_INTERLEAVE");
        const bool std3 = p->dim == 3 && p->T[0] == 16 && p->T[1] == 16 && p->T[2] == 16;
        if (p->type == 2 && !p->gm && std3 && (p->pair || p->f64) && p->nsub_interp > 0 &&
            !(env_il && env_il[0] == '0')) {
            int* sidx_cell = dev_alloc<int>(M, tally_tmp);
            int* rank = dev_alloc<int>(M, tally_tmp);
            HASTY_CHECK_HIP(hipMemcpyAsync(sidx_cell, p->sidx, size_t(M) * sizeof(int), hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(interleave_classes_kernel, dim3(unsigned(p->nsub_interp)), dim3(kClassThreads), 0, s,
                               static_cast<const int4*>(p->sub_interp), static_cast<const unsigned*>(keys_out),
                               static_cast<const int*>(sidx_cell), p->sidx, rank, cpt, p->T[1], p->T[2],
                               p->T[1] + p->w - 1, p->T[2] + p->w - 1);
            HASTY_CHECK_HIP(hipGetLastError());
            HASTY_CHECK_HIP(hipStreamSynchronize(s));
            HASTY_CHECK_HIP(hipFree(sidx_cell));
            HASTY_CHECK_HIP(hipFree(rank));
        }
        // 2. per sorted point: support, weights, visit count
        auto fold_sorted = [&](auto* ctag, auto* rtag) {
            using CT = std::remove_pointer_t<decltype(ctag)>;
            using RT = std::remove_pointer_t<decltype(rtag)>;
            auto fa = run_fold((CT*)nullptr);
            hipLaunchKernelGGL((fold_sorted_kernel<CT, RT>), dim3(blocks), dim3(256), 0, s, fa,
                               static_cast<const int*>(p->sidx), p->sinv, p->sl0,
                               static_cast<RT*>(p->sfrac), static_cast<RT*>(p->sw), p->w,
                               weight_stride(p->w), p->beta, vcnt);
        };
        if (coords_f64 && p->f64) fold_sorted((double*)nullptr, (double*)nullptr);
        else if (coords_f64) fold_sorted((double*)nullptr, (float*)nullptr);
        else if (p->f64) fold_sorted((float*)nullptr, (double*)nullptr);
        else fold_sorted((float*)nullptr, (float*)nullptr);
        HASTY_CHECK_HIP(hipGetLastError());
        // 3. visits in emission order (visit_stride), then a stable sort by tile
        const int64_t stride = visit_stride(M);
        int* vcnt_e = dev_alloc<int>(M, tally_tmp);
        hipLaunchKernelGGL(permute_counts_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const int*>(vcnt),
                           vcnt_e, M, stride);
        HASTY_CHECK_HIP(hipGetLastError());
        {
            size_t tb = 0;
            HASTY_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, vcnt_e, vpos, M, s));
            void* tmp = dev_alloc<unsigned char>(int64_t(std::max<size_t>(tb, 1)), tally_tmp);
            HASTY_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, vcnt_e, vpos, M, s));
            HASTY_CHECK_HIP(hipStreamSynchronize(s));
            HASTY_CHECK_HIP(hipFree(tmp));
        }
        unsigned* vkey_in = dev_alloc<unsigned>(voff, tally_tmp);
        unsigned* vkey_out = dev_alloc<unsigned>(voff, tally_tmp);
        int* vval_in = dev_alloc<int>(voff, tally_tmp);
        int* vval_out = dev_alloc<int>(voff, tally_tmp);
        hipLaunchKernelGGL(visit_emit_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const int*>(p->sl0),
                           static_cast<const int*>(p->sinv), M, g, static_cast<const int*>(vpos), vkey_in,
                           vval_in, stride);
        HASTY_CHECK_HIP(hipGetLastError());
        sort_pairs(vkey_in, vkey_out, vval_in, vval_out, voff, key_bits(nt));
        hipLaunchKernelGGL(visit_records_kernel, dim3(unsigned(ceil_div(voff, 256))), dim3(256), 0, s,
                           static_cast<const int*>(vval_out), static_cast<const int*>(p->sl0), M, voff,
                           p->visits);
        HASTY_CHECK_HIP(hipGetLastError());
        HASTY_CHECK_HIP(hipStreamSynchronize(s));
        for (void* q : {static_cast<void*>(keys_in), static_cast<void*>(keys_out), static_cast<void*>(idx_in),
                        static_cast<void*>(vcnt), static_cast<void*>(vcnt_e), static_cast<void*>(vpos),
                        static_cast<void*>(vkey_in),
                        static_cast<void*>(vkey_out), static_cast<void*>(vval_in), static_cast<void*>(vval_out)})
            HASTY_CHECK_HIP(hipFree(q));
    }
    // the host vectors above are copied asynchronously: wait before they go
    HASTY_CHECK_HIP(hipStreamSynchronize(s));
    HASTY_CHECK_HIP(hipFree(ccount));
    HASTY_CHECK_HIP(hipFree(coff));
    HASTY_CHECK_HIP(hipFree(vbuf));
    HASTY_CHECK_HIP(hipFree(toff));
    p->have_pts = true;
}

void nufft_setpts(NufftPlan* p, int64_t M, const void* c0, const void* c1, const void* c2) {
    nufft_setpts_typed(p, M, c0, c1, c2, p->f64);
}

namespace {

template <class V> using GridFn = void (*)(GridArgs<V>);
template <class V>
using SpreadFn = void (*)(GridArgs<V>, const int4*, const V*, const decltype(V::x)*);

// Consecutive subproblems per XCD run of the spread / interpolation grids.
// Measured at 128^3 x 32 coils x 8.4M points: spreading gains from runs of
// 32 tiles (a row of tiles along x shares its straddling points in one L2:
// 28.0 -> 24.7 ms); interpolation is fastest with single subproblems
// interleaved over the XCDs.  HASTY_XCD_CHUNK overrides both (tuning).
// The MFMA spreading kernel, when it applies: complex64 3-D plans spreading
// 16 or 32 transforms per block, an even batch (16-B aligned value rows),
// every fine-grid axis >= w + 4 (a support meets a 4-cell sub-block in one
// piece).  HASTY_SPREAD_LANES=0 keeps the cell-per-lane kernel.
static void* spread_lanes_fn(const NufftPlan* p, int nb) {
    static const bool off = [] {
        const char* e = std::getenv("HASTY_SPREAD_LANES");
        return e && e[0] == '0';
    }();
    if (off || p->f64 || p->dim != 3 || (p->group != 32 && p->group != 16) || nb % 2 != 0) return nullptr;
    for (int a = 0; a < 3; ++a)
        if (p->nf[a] < p->w + 4) return nullptr;
    return nufft_spread_lanes_ptr(p->w, p->group);
}

// Transform pairs per paired-interpolation block: half of the next pair's
// tile is fetched while the current one is gathered.  Measured at 128^3 x 32
// coils x 8.4M points: 1 pair 6.78 ms, 2 pairs 6.59, 4 pairs 6.65, 8 pairs
// 7.22 (fewer, longer blocks balance worse).  HASTY_INTERP_PAIR_RUN
// overrides (tuning).
static int interp_pair_run() {
    static const int env = [] {
        const char* e = std::getenv("HASTY_INTERP_PAIR_RUN");
        return e ? std::max(1, std::atoi(e)) : 0;
    }();
    return env > 0 ? env : 2;
}

static int xcd_chunk(bool spread) {
    static const int env = [] {
        const char* e = std::getenv("HASTY_XCD_CHUNK");
        return e ? std::max(1, std::atoi(e)) : 0;
    }();
    return env > 0 ? env : (spread ? 32 : 1);
}

template <class V>
void execute_impl(NufftPlan* p, V* c, V* f, const NufftHooks& hk, int ntransf) {
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
    if (p->type == 1 && !p->gm && p->M > 0) {
        // + padding: the spreading kernel reads a full group past the last point
        const int64_t need = p->M * int64_t(p->max_batch) + kMaxGroupSpread;
        if (need > p->csort_cap) {
            dev_free_void(p->csort, p->csort_cap * int64_t(sizeof(V)), p->device_bytes);
            p->csort = dev_alloc<V>(need, p->device_bytes);
            p->csort_cap = need;
        }
        const int64_t pneed = int64_t(p->nslots) * p->max_batch * p->T[0] * p->T[1] * p->T[2];
        if (pneed > p->partial_cap) {
            dev_free_void(p->partial, p->partial_cap * int64_t(sizeof(V)), p->device_bytes);
            p->partial = dev_alloc<V>(pneed, p->device_bytes);
            p->partial_cap = pneed;
        }
    }
    // hook diagonals for the internal batch starting at transform b0
    auto diag = [&](const void* ptr, int64_t sb, int type, int b0) {
        DiagOp<V> d{};
        if (ptr && type != HASTY_MULT_NONE) {
            d.ptr = static_cast<const V*>(ptr) + int64_t(b0) * sb;
            d.sb = sb;
            d.sr = p->n[2];
            d.type = type;
        }
        return d;
    };
    const V* pw = static_cast<const V*>(hk.pw);
    for (int b0 = 0; b0 < ntransf; b0 += p->max_batch) {
        const int nb = std::min(p->max_batch, ntransf - b0);
        V* cb = c + int64_t(b0) * p->M;
        V* fb = f + int64_t(b0) * nimg;
        const DiagOp<V> m1 = diag(hk.m1, hk.m1_sb, hk.m1_type, b0);
        const DiagOp<V> m2 = diag(hk.m2, hk.m2_sb, hk.m2_type, b0);
        GridArgs<V> ga{};
        ga.pw = pw;
        ga.accumulate = hk.accumulate;
        ga.fine = fine; ga.cdata = cb; ga.csort = static_cast<const V*>(p->csort);
        ga.sidx = p->sidx; ga.sl0 = p->sl0; ga.sw = static_cast<const R*>(p->sw);
        ga.visits = p->visits;
        ga.M = p->M; ga.nbatch = nb; ga.group = std::min(p->group, nb); ga.w = p->w;
        ga.ws = weight_stride(p->w);
        ga.g = g;
        ga.partial = static_cast<V*>(p->partial);
        const unsigned ngroups = unsigned(ceil_div(nb, ga.group));
        if (p->type == 2) {
            pruned_forward<V>(pg, nb, fb, fine, bufA, bufB, p->sign, m1, m2, deapod, s);
            if (p->M == 0) continue;
            if (p->gm) {
                HASTY_PROF("nufft_interp_gm", s);
                hipLaunchKernelGGL(reinterpret_cast<GridFn<V>>(kernel_ptr(p, 2)),
                                   dim3(unsigned(ceil_div(p->M, 256))), dim3(256), 0, s, ga);
            } else {
                ga.subs = p->sub_interp;
                ga.nsub = p->nsub_interp;
                ga.ngroups = int(ngroups);
                ga.xchunk = xcd_chunk(false);
                HASTY_PROF("nufft_interp", s);
                // XCD-aware numbering (xcd_block) over a 1-D grid
                if (p->pair && nb >= 2) {
                    ga.group = 2;
                    ga.ngroups = int(ceil_div(nb, 2));
                    ga.pair_run = interp_pair_run();
                    const int64_t nblocks =
                        xcd_grid(p->nsub_interp, ceil_div(ga.ngroups, ga.pair_run), ga.xchunk);
                    hipLaunchKernelGGL(reinterpret_cast<GridFn<V>>(nufft_interp_pair_ptr(p->dim, p->w)),
                                       dim3(unsigned(nblocks)), dim3(kInterpPairThreads),
                                       interp_lds_bytes(p, 2), s, ga);
                } else {
                    const int64_t nblocks = xcd_grid(p->nsub_interp, ngroups, ga.xchunk);
                    hipLaunchKernelGGL(reinterpret_cast<GridFn<V>>(interp_ptr(p)), dim3(unsigned(nblocks)),
                                       dim3(kNufftThreads), interp_lds_bytes(p, ga.group), s, ga);
                }
            }
            HASTY_CHECK_HIP(hipGetLastError());
        } else {
            if (p->gm) {
                HASTY_PROF("nufft_spread_gm", s);
                hipLaunchKernelGGL(fill_zero_kernel<V>,
                                   dim3(unsigned(std::min<int64_t>(8192, ceil_div(nb * nfine, 256)))),
                                   dim3(256), 0, s, fine, int64_t(nb) * nfine);
                if (p->M > 0)
                    hipLaunchKernelGGL(reinterpret_cast<GridFn<V>>(kernel_ptr(p, 3)),
                                       dim3(unsigned(ceil_div(p->M, 256))), dim3(256), 0, s, ga);
                HASTY_CHECK_HIP(hipGetLastError());
            } else {
                if (p->M > 0) {
                    HASTY_PROF("nufft_sort_values", s);
                    hipLaunchKernelGGL(sort_values_kernel<V>, dim3(unsigned(ceil_div(p->M, 64))), dim3(256),
                                       0, s, static_cast<const V*>(cb), static_cast<V*>(p->csort),
                                       static_cast<const int*>(p->sinv), pw, p->M, nb);
                    HASTY_CHECK_HIP(hipGetLastError());
                }
                ga.subs = p->sub_spread;
                if (void* lanes_fn = spread_lanes_fn(p, nb)) {
                    HASTY_PROF("nufft_spread", s);
                    const int ng = int(ceil_div(nb, p->group));
                    ga.nsub = p->nsub_spread;
                    ga.ngroups = ng;
                    ga.xchunk = xcd_chunk(true);
                    hipLaunchKernelGGL(reinterpret_cast<SpreadFn<V>>(lanes_fn),
                                       dim3(unsigned(xcd_grid(p->nsub_spread, ng, ga.xchunk))),
                                       dim3(kLaneSpreadThreads), 0, s, ga, static_cast<const int4*>(p->visits),
                                       static_cast<const V*>(p->csort), static_cast<const R*>(p->sw));
                    HASTY_CHECK_HIP(hipGetLastError());
                } else {
                    HASTY_PROF("nufft_spread", s);
                    // one wave per 64-cell sub-block of the tile, up to 8 per block
                    const int b0 = p->T[0] == 1 ? 1 : 4, b1 = p->T[0] == 1 ? 8 : 4, b2 = b1;
                    const int64_t nsb = ceil_div(p->T[0], b0) * ceil_div(p->T[1], b1) * ceil_div(p->T[2], b2);
                    // waves per block: one per 64-cell sub-block, up to 8, and as many as
                    // keep each wave's staged weight rows (64 x ws) within 64 KB of LDS
                    const int64_t row_bytes = int64_t(64) * weight_stride(p->w) * int64_t(sizeof(R));
                    const int64_t nw = std::max<int64_t>(1, std::min<int64_t>({8, nsb, (64 * 1024) / row_bytes}));
                    const unsigned threads = unsigned(64 * nw);
                    const size_t lds = size_t(nw * row_bytes);
                    const int ng = int(ceil_div(nb, p->group));
                    ga.nsub = p->nsub_spread;
                    ga.ngroups = ng;
                    ga.xchunk = xcd_chunk(true);
                    void* fn = p->f64 ? nufft_spread_ptr<double2>(p->group) : nufft_spread_ptr<float2>(p->group);
                    hipLaunchKernelGGL(reinterpret_cast<SpreadFn<V>>(fn),
                                       dim3(unsigned(xcd_grid(p->nsub_spread, ng, ga.xchunk))),
                                       dim3(threads), lds, s, ga, static_cast<const int4*>(p->visits),
                                       static_cast<const V*>(p->csort), static_cast<const R*>(p->sw));
                    HASTY_CHECK_HIP(hipGetLastError());
                }
                if (p->nsplit > 0) {
                    HASTY_PROF("nufft_spread_reduce", s);
                    const int cells = p->T[0] * p->T[1] * p->T[2];
                    const unsigned bx = unsigned(std::min<int64_t>(64, ceil_div(int64_t(nb) * cells, 256)));
                    hipLaunchKernelGGL(reduce_split_tiles_kernel<V>, dim3(bx, unsigned(p->nsplit)), dim3(256),
                                       0, s, fine, static_cast<const V*>(p->partial),
                                       static_cast<const int4*>(p->split_info), nb, g);
                    HASTY_CHECK_HIP(hipGetLastError());
                }
            }
            const DiagOp<V> io = diag(hk.io, hk.io_sb, hk.io_type, b0);
            pruned_inverse<V>(pg, nb, fine, hk.sum_batch ? f : fb, bufA, bufB, p->sign, R(1), m1, m2,
                              io, deapod, hk.accumulate || (hk.sum_batch && b0 > 0), hk.sum_batch, s);
        }
    }
}

}  // namespace

namespace {

template <class V>
void coord_grad_impl(NufftPlan* p, const V* h, const V* w, decltype(V::x)* out) {
    using R = decltype(V::x);
    const hipStream_t s = p->stream;
    const int64_t nimg = p->n[0] * p->n[1] * p->n[2];
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
    for (int b0 = 0; b0 < p->ntransf; b0 += p->max_batch) {
        const int nb = std::min(p->max_batch, p->ntransf - b0);
        DiagOp<V> none{};
        pruned_forward<V>(pg, nb, h + int64_t(b0) * nimg, fine, static_cast<V*>(p->bufA),
                          static_cast<V*>(p->bufB), p->sign, none, none, deapod, s);
        if (p->M == 0) continue;
        CoordGradArgs<V> a{};
        a.fine = fine;
        a.w = w + int64_t(b0) * p->M;
        a.out = out;
        a.sidx = p->sidx;
        a.sl0 = p->sl0;
        a.sfrac = static_cast<const R*>(p->sfrac);
        a.sw = static_cast<const R*>(p->sw);
        a.M = p->M;
        a.nbatch = nb;
        a.ws = weight_stride(p->w);
        a.dim = p->dim;
        a.accumulate = b0 > 0;
        a.beta = p->beta;
        a.g = geo_of(p);
        HASTY_PROF("nufft_coord_grad", s);
        hipLaunchKernelGGL(reinterpret_cast<void (*)(CoordGradArgs<V>)>(kernel_ptr(p, 4)),
                           dim3(unsigned(ceil_div(p->M, 256))), dim3(256), 0, s, a);
        HASTY_CHECK_HIP(hipGetLastError());
    }
}

}  // namespace

void nufft_coord_grad(NufftPlan* p, const void* h, const void* w, void* out) {
    if (!p->have_pts) throw Error(HASTY_ERR_STATE, "nufft coordinate gradient called before setpts");
    HASTY_REQUIRE(p->type == 2, "the coordinate gradient runs on a type-2 plan");
    HASTY_REQUIRE(h != nullptr && out != nullptr, "mode or output array is NULL");
    HASTY_REQUIRE(w != nullptr || p->M == 0, "nonuniform array is NULL");
    if (p->f64)
        coord_grad_impl<double2>(p, static_cast<const double2*>(h), static_cast<const double2*>(w),
                                 static_cast<double*>(out));
    else
        coord_grad_impl<float2>(p, static_cast<const float2*>(h), static_cast<const float2*>(w),
                                static_cast<float*>(out));
}

void nufft_execute(NufftPlan* p, void* c, void* f) {
    if (!p->have_pts) throw Error(HASTY_ERR_STATE, "nufft execute called before setpts");
    HASTY_REQUIRE(f != nullptr, "mode array is NULL");
    HASTY_REQUIRE(c != nullptr || p->M == 0, "nonuniform array is NULL");
    const NufftHooks none{};
    if (p->f64) execute_impl<double2>(p, static_cast<double2*>(c), static_cast<double2*>(f), none, p->ntransf);
    else execute_impl<float2>(p, static_cast<float2*>(c), static_cast<float2*>(f), none, p->ntransf);
}

void nufft_execute_hooked(NufftPlan* p, void* c, void* f, const NufftHooks& hk, int ntransf) {
    if (!p->have_pts) throw Error(HASTY_ERR_STATE, "nufft execute called before setpts");
    HASTY_REQUIRE(f != nullptr, "mode array is NULL");
    HASTY_REQUIRE(c != nullptr || p->M == 0, "nonuniform array is NULL");
    HASTY_REQUIRE(ntransf >= 1 && ntransf <= p->ntransf, "transform count out of range");
    HASTY_REQUIRE(p->type == 1 || (hk.io == nullptr && !hk.sum_batch),
                  "io diagonal and batch sum apply to type-1 plans");
    if (p->f64) execute_impl<double2>(p, static_cast<double2*>(c), static_cast<double2*>(f), hk, ntransf);
    else execute_impl<float2>(p, static_cast<float2*>(c), static_cast<float2*>(f), hk, ntransf);
}

void sense_offres_forward(NufftPlan* p, int ncoils, int L, const void* image, const void* smaps,
                          const void* bil, const void* cjl, void* kdata) {
    HASTY_REQUIRE(p->type == 2, "the SENSE forward runs on a type-2 plan");
    HASTY_REQUIRE(ncoils >= 1 && L >= 1, "ncoils and L must be >= 1");
    HASTY_REQUIRE(image && smaps && (kdata || p->M == 0), "image, smaps and kdata must be non-NULL");
    HASTY_REQUIRE(L == 1 || (bil && cjl), "bil and cjl are required for L > 1");
    const int64_t es = p->f64 ? 16 : 8;
    const int64_t nimg = p->n[0] * p->n[1] * p->n[2];
    auto at = [&](const void* base, int64_t off) {
        return base ? static_cast<const char*>(base) + off * es : nullptr;
    };
    for (int c0 = 0; c0 < ncoils; c0 += p->ntransf) {
        const int cb = std::min(p->ntransf, ncoils - c0);
        for (int l = 0; l < L; ++l) {
            NufftHooks hk{};
            hk.m1 = image; hk.m1_sb = 0; hk.m1_type = HASTY_MULT;
            hk.m2 = at(cjl, int64_t(l) * nimg); hk.m2_sb = 0; hk.m2_type = HASTY_MULT;
            hk.pw = at(bil, int64_t(l) * p->M);
            hk.accumulate = l > 0;
            nufft_execute_hooked(p, const_cast<char*>(at(kdata, int64_t(c0) * p->M)),
                                 const_cast<char*>(at(smaps, int64_t(c0) * nimg)), hk, cb);
        }
    }
}

void sense_offres_adjoint(NufftPlan* p, int ncoils, int L, const void* kdata, const void* smaps,
                          const void* bil, const void* cjl, void* images, int combine) {
    HASTY_REQUIRE(p->type == 1, "the SENSE adjoint runs on a type-1 plan");
    HASTY_REQUIRE(ncoils >= 1 && L >= 1, "ncoils and L must be >= 1");
    HASTY_REQUIRE(images && smaps && (kdata || p->M == 0), "kdata, smaps and images must be non-NULL");
    HASTY_REQUIRE(L == 1 || (bil && cjl), "bil and cjl are required for L > 1");
    const int64_t es = p->f64 ? 16 : 8;
    const int64_t nimg = p->n[0] * p->n[1] * p->n[2];
    auto at = [&](const void* base, int64_t off) {
        return base ? static_cast<const char*>(base) + off * es : nullptr;
    };
    for (int c0 = 0; c0 < ncoils; c0 += p->ntransf) {
        const int cb = std::min(p->ntransf, ncoils - c0);
        for (int l = 0; l < L; ++l) {
            NufftHooks hk{};
            hk.m1 = at(cjl, int64_t(l) * nimg); hk.m1_sb = 0; hk.m1_type = HASTY_MULT_CONJ;
            hk.m2 = at(smaps, int64_t(c0) * nimg); hk.m2_sb = nimg; hk.m2_type = HASTY_MULT_CONJ;
            hk.pw = at(bil, int64_t(l) * p->M);
            hk.accumulate = l > 0 || (combine && c0 > 0);
            hk.sum_batch = combine;
            void* out = combine ? images : static_cast<char*>(images) + int64_t(c0) * nimg * es;
            nufft_execute_hooked(p, const_cast<char*>(at(kdata, int64_t(c0) * p->M)), out, hk, cb);
        }
    }
}

}  // namespace hasty
