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
    const int4* visits;     // (8 s + tile-neighbour bits, l0z, l0y, l0x) per tile visit
    int64_t M;
    int nbatch, group, w, ws;
    V* partial;             // [slot][nbatch][tile cells]: partial tiles of split tiles
    int nsub, ngroups;      // subproblems and transform groups (xcd_block)
    int xchunk;             // consecutive subproblems per XCD run (xcd_block)
    int pair_run;           // paired interpolation: transform pairs per block, run one after another
    const V* pw;            // optional per-point factor [M], caller order (SENSE time segments)
    int accumulate;         // interpolation: add into cdata
    Geo g;
};

// Interpolation output of point `orig` of transform row `row`: pw[orig] * v,
// plus the value already there when accumulating.
template <class V>
__device__ inline void store_point(const GridArgs<V>& a, int64_t row, int orig, V v) {
    V* dst = a.cdata + row * a.M + orig;
    if (a.pw) v = cmul(v, a.pw[orig]);
    if (a.accumulate) v = cadd(v, *dst);
    *dst = v;
}

// Spread / interpolation blocks are numbered with the 8 XCDs in mind:
// workgroups go to the XCDs round-robin by id, so XCD x runs ids x, x + 8,
// x + 16, ...  Here its k-th block takes transform group k % ngroups of
// subproblem ((r * 8 + x) * chunk + j), with r, j = divmod(k / ngroups, chunk):
// each XCD works through runs of `chunk` consecutive subproblems (tiles
// that are neighbours along x, whose halos and straddling points it then
// reads from its own L2), all transform groups of a subproblem one after the
// other, and the runs alternate between XCDs so the work stays balanced.
// Returns false for the padding ids past the last subproblem.
constexpr int kXcds = 8;
__device__ inline bool xcd_block(int nsub, int ngroups, int chunk, int& sub, int& group) {
    const int id = int(blockIdx.x);
    const int x = id % kXcds, k = id / kXcds;
    group = k % ngroups;
    const int q = k / ngroups;
    sub = ((q / chunk) * kXcds + x) * chunk + q % chunk;
    return sub < nsub;
}
// number of ids the mapping above needs for nsub subproblems
__host__ inline int64_t xcd_grid(int64_t nsub, int64_t ngroups, int64_t chunk) {
    const int64_t per_round = kXcds * chunk;
    return ((nsub + per_round - 1) / per_round) * per_round * ngroups;
}

// ---------------------------------------------------------------------------
// Interpolation: one block per (tile subproblem, transform group); the padded
// tile footprint [t0, t0 + T + w - 1) of each transform is staged in LDS and
// each thread gathers the w^d cells of its points.  TZ/TY/TX > 0 fix the tile
// edge at compile time (the standard 16^3 / 32^2 tiles): the padded pitch is
// then a constant and every cell read is an immediate offset from one
// per-point base address; 0 reads the tile from the plan.
// ---------------------------------------------------------------------------
template <class V, int W, int TZ = 0, int TY = 0, int TX = 0>
__global__ void __launch_bounds__(kNufftThreads) interp_kernel(GridArgs<V> a) {
    using R = decltype(V::x);
    extern __shared__ unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    int isub, igroup;
    if (!xcd_block(a.nsub, a.ngroups, a.xchunk, isub, igroup)) return;
    const int4 sub = a.subs[isub];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    constexpr bool FIXED = TX > 0;
    const int wz = FIXED ? (TZ == 1 ? 1 : W) : a.g.wa[0];
    const int Pz = FIXED ? (TZ == 1 ? 1 : TZ + W - 1) : a.g.T[0] + a.g.wa[0] - 1;
    const int Py = FIXED ? TY + W - 1 : a.g.T[1] + a.g.wa[1] - 1;
    const int Px = FIXED ? TX + W - 1 : a.g.T[2] + a.g.wa[2] - 1;
    const int cells = Pz * Py * Px;
    const int g0 = igroup * a.group;
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
        const V* pbase = lds + (lz * Py + ly) * Px + lx;
        for (int gi = 0; gi < ng; ++gi) {
            const V* tb = pbase + gi * cells;
            V acc = cmake<V>(R(0), R(0));
#pragma unroll
            for (int qz = 0; qz < W; ++qz) {
                if (qz < wz) {
                    V accy = cmake<V>(R(0), R(0));
#pragma unroll
                    for (int qy = 0; qy < W; ++qy) {
                        // the whole row first: its LDS reads are in flight together
                        V row[W];
#pragma unroll
                        for (int qx = 0; qx < W; ++qx) row[qx] = tb[(qz * Py + qy) * Px + qx];
                        V accx = cmake<V>(R(0), R(0));
#pragma unroll
                        for (int qx = 0; qx < W; ++qx) {
                            accx.x += wx[qx] * row[qx].x;
                            accx.y += wx[qx] * row[qx].y;
                        }
                        accy.x += wy[qy] * accx.x;
                        accy.y += wy[qy] * accx.y;
                    }
                    acc.x += wz_[qz] * accy.x;
                    acc.y += wz_[qz] * accy.y;
                }
            }
            store_point(a, int64_t(g0 + gi), orig, acc);
        }
    }
}

// ---------------------------------------------------------------------------
// Paired interpolation (complex64, compile-time tile): one block per (tile
// subproblem, pair of transforms).  The pair is staged interleaved, one 16-B
// value (transform 2k, transform 2k + 1) per padded cell, so every cell of a
// point's window is a single aligned 16-B LDS read serving both transforms:
// on gfx950 16-B LDS reads move about twice the bytes per clock of the 8-B
// pairs the single-transform kernel issues.  The tile of a pair takes twice
// the LDS (one block per CU at 16^3, w = 6), so the block has 1024 threads to
// keep 16 waves per CU.  An odd last transform stages zeros in its partner
// slot and stores nothing for it.
// ---------------------------------------------------------------------------
constexpr int kInterpPairThreads = 1024;

template <int W, int TZ, int TY, int TX>
__global__ void __launch_bounds__(kInterpPairThreads) interp_pair_kernel(GridArgs<float2> a) {
    extern __shared__ unsigned char smem_raw[];
    float4* lds = reinterpret_cast<float4*>(smem_raw);
    // a block runs transform pairs [p0, p1) of one subproblem, one after the
    // other; the XCD mapping numbers its runs like transform groups
    const int nruns = (a.ngroups + a.pair_run - 1) / a.pair_run;
    int isub, irun;
    if (!xcd_block(a.nsub, nruns, a.xchunk, isub, irun)) return;
    const int p0 = irun * a.pair_run, p1 = min(a.ngroups, p0 + a.pair_run);
    const int4 sub = a.subs[isub];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    constexpr int wz = TZ == 1 ? 1 : W;
    constexpr int Pz = TZ == 1 ? 1 : TZ + W - 1, Py = TY + W - 1, Px = TX + W - 1;
    constexpr int cells = Pz * Py * Px;
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    // Staging: thread t stages padded cells t + 1024 u of a pair's tile: its
    // first cell's (pz, py, px) once, then the compile-time steps of 1024
    // cells with their carries; the fine grid wraps with one unsigned min per
    // axis (t0 + p < 2 nf) and the offsets are 32-bit byte offsets (the plan
    // pairs transforms only when one transform's fine grid is under 4 GiB).
    // The loads of the first half of pair p + 1's cells are issued right after
    // the point data of pair p's first round, so they stream in while pair p
    // is gathered; the second half is loaded when pair p is done, then the
    // whole tile goes to LDS.
    constexpr int SU = (cells + kInterpPairThreads - 1) / kInterpPairThreads;
    // cells held in registers across a gather: the first SP steps (all of
    // them would spill at the 128-VGPR cap of 1024-thread blocks)
    constexpr int SP = SU / 2;
    float2 v0[SU], v1[SU];
    auto fetch = [&](int pair, int ulo, int uhi) {
        const int g0 = 2 * pair;
        const bool two = g0 + 1 < a.nbatch;
        const float2* f0 = a.fine + int64_t(g0) * G;
        const float2* f1 = two ? f0 + G : f0;
        const int base = int(threadIdx.x);
        const unsigned nf0 = unsigned(a.g.nf[0]), nf1 = unsigned(a.g.nf[1]), nf2 = unsigned(a.g.nf[2]);
        const int px0 = base % Px, py0 = (base / Px) % Py, pz0 = base / (Px * Py);
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            constexpr int kStep = kInterpPairThreads;
            const int dx = (u * kStep) % Px, dy = ((u * kStep) / Px) % Py, dz = (u * kStep) / (Px * Py);
            int px = px0 + dx, py = py0 + dy, pz = pz0 + dz;
            if (px >= Px) { px -= Px; ++py; }
            if (py >= Py) { py -= Py; ++pz; }
            if (u >= ulo && u < uhi && base + u * kStep < cells) {
                unsigned gz = unsigned(t0[0] + pz), gy = unsigned(t0[1] + py), gx = unsigned(t0[2] + px);
                gz = min(gz, gz - nf0);
                gy = min(gy, gy - nf1);
                gx = min(gx, gx - nf2);
                // byte offset from the (uniform) grid base: a 32-bit VGPR offset
                const unsigned boff = ((gz * nf1 + gy) * nf2 + gx) * unsigned(sizeof(float2));
                v0[u] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(f0) + boff);
                v1[u] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(f1) + boff);
                if (!two) v1[u] = make_float2(0.f, 0.f);
            }
        }
    };
    fetch(p0, 0, SP);
    for (int pair = p0; pair < p1; ++pair) {
        const int g0 = 2 * pair;
        const bool two = g0 + 1 < a.nbatch;
        fetch(pair, SP, SU);
        if (pair > p0) __syncthreads();   // the previous pair's gather is done with the tile
        {
            const int base = int(threadIdx.x);
#pragma unroll
            for (int u = 0; u < SU; ++u)
                if (base + u * kInterpPairThreads < cells)
                    lds[base + u * kInterpPairThreads] = make_float4(v0[u].x, v0[u].y, v1[u].x, v1[u].y);
        }
        __syncthreads();
        bool fetched = pair + 1 >= p1;
        for (int s = sub.y + threadIdx.x; s < sub.y + sub.z; s += blockDim.x) {
            const int lz = a.sl0[s] - t0[0];
            const int ly = a.sl0[a.M + s] - t0[1];
            const int lx = a.sl0[2 * a.M + s] - t0[2];
            const float* w = a.sw + int64_t(s) * a.ws;
            float wz_[W], wy[W], wx[W];
#pragma unroll
            for (int q = 0; q < W; ++q) {
                wz_[q] = w[q];
                wy[q] = w[W + q];
                wx[q] = w[2 * W + q];
            }
            const int orig = a.sidx[s];
            if (!fetched) {
                fetch(pair + 1, 0, SP);
                fetched = true;
            }
            const float4* tb = lds + (lz * Py + ly) * Px + lx;
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            // one z plane of the window per iteration: unrolled whole, its 16-B
            // reads (4 VGPRs each) would be hoisted together and spill
#pragma unroll 1
            for (int qz = 0; qz < wz; ++qz) {
                float4 accy = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int qy = 0; qy < W; ++qy) {
                    float4 row[W];
#pragma unroll
                    for (int qx = 0; qx < W; ++qx) row[qx] = tb[(qz * Py + qy) * Px + qx];
                    float4 accx = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                    for (int qx = 0; qx < W; ++qx) {
                        accx.x += wx[qx] * row[qx].x;
                        accx.y += wx[qx] * row[qx].y;
                        accx.z += wx[qx] * row[qx].z;
                        accx.w += wx[qx] * row[qx].w;
                    }
                    accy.x += wy[qy] * accx.x;
                    accy.y += wy[qy] * accx.y;
                    accy.z += wy[qy] * accx.z;
                    accy.w += wy[qy] * accx.w;
                }
                acc.x += wz_[qz] * accy.x;
                acc.y += wz_[qz] * accy.y;
                acc.z += wz_[qz] * accy.z;
                acc.w += wz_[qz] * accy.w;
            }
            store_point(a, int64_t(g0), orig, make_float2(acc.x, acc.y));
            if (two) store_point(a, int64_t(g0 + 1), orig, make_float2(acc.z, acc.w));
        }
        if (!fetched) fetch(pair + 1, 0, SP);   // (no points for this thread)
    }
}

// ---------------------------------------------------------------------------
// Spreading, formulated as a gather: one block per (tile subproblem, group of
// CG transforms).  Each wave owns a sub-block of 64 cells of the tile
// (4 x 4 x 4 in 3-D, 1 x 8 x 8 in 2-D) and each lane one cell of it, with
// the CG complex accumulators of its cell in registers (FP32 for complex64).
// The wave walks the tile's visit list (points whose support touches the
// tile) 64 records at a time, one per lane; the lanes test their records
// against the sub-block and the wave takes the hits in visit order (ballot),
// every lane adding w_z w_y w_x c over the CG transforms of the point.  No atomics and no LDS; each cell's sum is taken in visit order, so
// the result is the same on every run of a plan.  A split tile writes its
// share to a partial slot (summed by reduce_split_tiles_kernel).
// ---------------------------------------------------------------------------
// Broadcast lane 0's value to the wave as a scalar (the value is known to be
// wave-uniform; this lets the compiler keep it in SGPRs and branch on it).
__device__ inline int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// x mod n for x in (-n, 2n)
__device__ inline int wrap_once(int x, int n) {
    x += x < 0 ? n : 0;
    return x - (x >= n ? n : 0);
}

// Do the arcs [a, a + la) and [b, b + lb) of the circle Z_n intersect?
__device__ inline bool arcs_meet(int a, int la, int b, int lb, int n) {
    return wrap_once(b - a, n) < la || wrap_once(a - b, n) < lb;
}

// The read-only arrays come as separate restrict-qualified arguments: with no
// aliasing stores, the wave-uniform loads (the point's CG values) become
// scalar loads.  csort holds CG - 1 padding elements past the
// last point, so the CG values of a point are read without bounds tests.
template <class V, int CG>
__device__ __forceinline__ void spread_body(const GridArgs<V>& a, const int4* __restrict__ visits,
                                            const V* __restrict__ csort,
                                            const decltype(V::x)* __restrict__ sw) {
    using R = decltype(V::x);
    int isub, igroup;
    if (!xcd_block(a.nsub, a.ngroups, a.xchunk, isub, igroup)) return;
    const int4 sub = a.subs[isub];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    const int T0 = a.g.T[0], T1 = a.g.T[1], T2 = a.g.T[2];
    const int cells = T0 * T1 * T2;
    const int n0 = int(a.g.nf[0]), n1 = int(a.g.nf[1]), n2 = int(a.g.nf[2]);
    const int wa0 = a.g.wa[0], wa1 = a.g.wa[1], wa2 = a.g.wa[2];
    const int B0 = T0 == 1 ? 1 : 4, B1 = T0 == 1 ? 8 : 4, B2 = T0 == 1 ? 8 : 4;
    const int ns1 = (T1 + B1 - 1) / B1, ns2 = (T2 + B2 - 1) / B2;
    const int nsb = ((T0 + B0 - 1) / B0) * ns1 * ns2;
    const int g0 = igroup * CG;
    const int ng = min(CG, a.nbatch - g0);
    const int lane = threadIdx.x & 63;
    const int c2 = lane % B2, c1 = (lane / B2) % B1, c0 = lane / (B2 * B1);
    const int wave = wave_uniform(int(threadIdx.x >> 6));
    const int nwaves = blockDim.x >> 6;
    const int W = a.w;
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    // per wave, the window weights of the 64 records in flight (one row per
    // lane, staged by the lane that loaded the record): a hit's weights are
    // then read from LDS, not gathered from global memory
    extern __shared__ unsigned char smem_raw[];
    R* wl = reinterpret_cast<R*>(smem_raw) + int64_t(wave) * 64 * a.ws;
    constexpr int RV = 16 / int(sizeof(R));   // elements per 16-B copy (rows are 16-B multiples)
    for (int sb = wave; sb < nsb; sb += nwaves) {
        const int o2 = (sb % ns2) * B2, o1 = ((sb / ns2) % ns1) * B1, o0 = (sb / (ns2 * ns1)) * B0;
        const int l0c = o0 + c0, l1c = o1 + c1, l2c = o2 + c2;     // lane's tile-local cell
        const bool mine = l0c < T0 && l1c < T1 && l2c < T2;
        const int G0 = t0[0] + o0, G1 = t0[1] + o1, G2 = t0[2] + o2;  // sub-block origin
        const int x0 = G0 + c0, x1 = G1 + c1, x2 = G2 + c2;           // lane's cell (< nf + B)
        R acc[2 * CG];
#pragma unroll
        for (int k = 0; k < 2 * CG; ++k) acc[k] = R(0);
        // 64 visit records at a time, one per lane: every lane tests its
        // record against the sub-block, and the wave then takes the
        // overlapping ones in visit order (ballot bits, low to high).
        const int vend = sub.y + sub.z;
        for (int vb = sub.y; vb < vend; vb += 64) {
            const int v = vb + lane;
            int4 rec = make_int4(0, 0, 0, 0);
            bool hit = false;
            if (v < vend) {
                rec = visits[v];
                hit = arcs_meet(rec.y, wa0, G0, B0, n0) && arcs_meet(rec.z, wa1, G1, B1, n1) &&
                      arcs_meet(rec.w, wa2, G2, B2, n2);
                if (hit) {
                    const uint4* src = reinterpret_cast<const uint4*>(sw + int64_t(rec.x >> 3) * a.ws);
                    uint4* dst = reinterpret_cast<uint4*>(wl + lane * a.ws);
                    for (int k = 0; k < a.ws / RV; ++k) dst[k] = src[k];
                }
            }
            unsigned long long mask = __ballot(hit);
            while (mask) {
                const int i = __ffsll(mask) - 1;
                mask &= mask - 1;
                const int s = __builtin_amdgcn_readlane(rec.x, i) >> 3;
                const int p0 = __builtin_amdgcn_readlane(rec.y, i);
                const int p1 = __builtin_amdgcn_readlane(rec.z, i);
                const int p2 = __builtin_amdgcn_readlane(rec.w, i);
                // q = (x - p) mod n for x, p in [0, n): u = x - p + n in [1, 2n),
                // q = min(u, u - n) unsigned (lanes past the grid get q >= n,
                // outside the support)
                const unsigned u0 = unsigned(x0 + (n0 - p0)), u1 = unsigned(x1 + (n1 - p1)),
                               u2 = unsigned(x2 + (n2 - p2));
                const unsigned q0 = min(u0, u0 - unsigned(n0)), q1 = min(u1, u1 - unsigned(n1)),
                               q2 = min(u2, u2 - unsigned(n2));
                const bool in = q0 < unsigned(wa0) && q1 < unsigned(wa1) && q2 < unsigned(wa2);
                // out-of-support lanes read weight 0 of the point (in bounds)
                const R* wp = wl + i * a.ws;
                const R w0 = wp[in ? q0 : 0], w1 = wp[W + (in ? q1 : 0)], w2 = wp[2 * W + (in ? q2 : 0)];
                const V* cv = csort + int64_t(s) * a.nbatch + g0;
                V c[CG];
#pragma unroll
                for (int k = 0; k < CG; ++k) c[k] = cv[k];
                const R wt = in ? w0 * w1 * w2 : R(0);
#pragma unroll
                for (int k = 0; k < CG; ++k) {
                    acc[2 * k] += wt * c[k].x;
                    acc[2 * k + 1] += wt * c[k].y;
                }
            }
        }
        if (!mine) continue;
        const int cell = (l0c * T1 + l1c) * T2 + l2c;
        if (sub.w != 0) {
            V* dst = a.partial + (int64_t(sub.w - 1) * a.nbatch + g0) * cells + cell;
#pragma unroll
            for (int k = 0; k < CG; ++k)
                if (k < ng) dst[int64_t(k) * cells] = cmake<V>(acc[2 * k], acc[2 * k + 1]);
        } else {
            V* dst = a.fine + int64_t(g0) * G + (int64_t(x0) * n1 + x1) * n2 + x2;
#pragma unroll
            for (int k = 0; k < CG; ++k)
                if (k < ng) dst[int64_t(k) * G] = cmake<V>(acc[2 * k], acc[2 * k + 1]);
        }
    }
}

template <class V, int CG>
__global__ void __launch_bounds__(kNufftThreads)
    spread_kernel(GridArgs<V> a, const int4* __restrict__ visits, const V* __restrict__ csort,
                  const decltype(V::x)* __restrict__ sw) {
    spread_body<V, CG>(a, visits, csort, sw);
}

// 32 complex64 transforms per block: the 64 accumulator VGPRs leave room
// for 6 waves per SIMD only under an 80-VGPR budget (the compiler spills a
// little outside the hit loop).  Measured at 128^3 x 32 coils: 24.4 ms at
// 88 VGPRs / 5 waves, 23.4 ms at 80 / 6; at 7 waves the accumulators spill
// inside the hit loop and it takes 46 ms.
template <class V, int CG>
__global__ void __launch_bounds__(kNufftThreads) __attribute__((amdgpu_waves_per_eu(6, 8)))
    spread_kernel_w6(GridArgs<V> a, const int4* __restrict__ visits, const V* __restrict__ csort,
                     const decltype(V::x)* __restrict__ sw) {
    spread_body<V, CG>(a, visits, csort, sw);
}

// ---------------------------------------------------------------------------
// Spreading on the matrix cores (complex64, 16 or 32 transforms per block,
// 3-D).  For one wave's 4 x 4 x 4 sub-block of a tile, the sum over the
// points that touch it is a rank-K update
//   acc[tp][cell] += sum_hits value[hit][tp] * weight[hit][cell],
// tp = 2 transform + part (real, imaginary), cell = 16 z + 4 y + x, so it runs
// as FP32 MFMAs: v_mfma_f32_32x32x2_f32 takes the hits two at a time (lanes
// 0-31 carry hit k = 0, lanes 32-63 hit k = 1 of the pair; an odd last hit
// pairs with zero weights), A = the hits' values (32 transform-parts per
// tile), B = their window weights w_z w_y w_x at the 64 cells (two 32-cell
// tiles).  Each lane keeps 16 FP32 accumulators per tile, 32 or 64 in all.
// Cells outside a support take zero weights; that costs matrix work but no
// branches.  MFMA products and sums are FP32.
//
// The block (8 waves, one 4^3 sub-block each; larger tiles take several
// passes) walks the tile's visit list in chunks of 128 records.  Every chunk
// is staged once in LDS for all 8 waves: the records, the points' values
// (2 NTG floats each) and their window weights in zero-padded rows (index
// q + 3 for q = cell - first support cell, so any cell of a sub-block reads
// a weight or a zero).  Each wave tests the records against its sub-block
// and takes the hits in visit order, so every cell's sum has a fixed order
// and the result is the same on every run.  No atomics.  Requires every axis
// of the fine grid to be >= w + 4 (a support then meets a sub-block in one
// piece).
// ---------------------------------------------------------------------------
constexpr int kLaneSpreadThreads = 512;
constexpr int kLaneSpreadChunk = 128;   // records staged per chunk (a multiple of 64)
constexpr int kLaneSpreadPad = 16;   // padded weight row (w + 6 <= 16)

__device__ __forceinline__ int lane_hit_pop(unsigned long long& mask) {
    const int i = __ffsll(mask) - 1;
    mask &= mask - 1;
    return i;
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

template <int W, int NTG>
__global__ void __launch_bounds__(kLaneSpreadThreads)
    spread_lanes_kernel(GridArgs<float2> a, const int4* __restrict__ visits,
                          const float2* __restrict__ csort, const float* __restrict__ sw) {
    static_assert(W + 6 <= kLaneSpreadPad, "padded weight row too short");
    static_assert(kLaneSpreadChunk % 64 == 0 && kLaneSpreadThreads == 512, "staging: 8 threads per record");
    constexpr int K = kLaneSpreadChunk, PW = kLaneSpreadPad;
    constexpr int ZROW = K * 3 * PW;   // 16 zeros after the weight rows (padding hits)
    __shared__ int4 lrec[K];
    __shared__ float lval[K * 64];
    __shared__ float lw[K * 3 * PW + 16];
    // per wave: the hits of the current 64 records, compacted in visit order
    // into three runs (both N-tiles, N-tile 0 only, N-tile 1 only), each
    // padded to an even length, as (value row, z / y / x weight-row starts)
    __shared__ int4 lhit[8][67];
    int isub, igroup;
    if (!xcd_block(a.nsub, a.ngroups, a.xchunk, isub, igroup)) return;
    const int4 sub = a.subs[isub];
    int t0[3];
    tile_origin(a.g, sub.x, t0);
    const int T0 = a.g.T[0], T1 = a.g.T[1], T2 = a.g.T[2];
    const int cells = T0 * T1 * T2;
    const int n0 = int(a.g.nf[0]), n1 = int(a.g.nf[1]), n2 = int(a.g.nf[2]);
    const int ns1 = (T1 + 3) / 4, ns2 = (T2 + 3) / 4;
    const int nsb = ((T0 + 3) / 4) * ns1 * ns2;
    static_assert(NTG == 16 || NTG == 32, "16 or 32 transforms per block");
    constexpr int MT = NTG / 16;                    // 32-row MFMA tiles of transform-parts
    const int g0 = igroup * NTG;
    const int ng = min(NTG, a.nbatch - g0);
    const int lane = int(threadIdx.x) & 63;
    const int wave = wave_uniform(int(threadIdx.x >> 6));
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    const int vend = sub.y + sub.z;
    const int npass = (nsb + 7) / 8;
    // padded weight rows: entries q + 3 for taps q < W are written per chunk,
    // the rest stay zero (the barrier orders these zeros before the first
    // chunk's weights, which land on some of the same entries)
    for (int e = int(threadIdx.x); e < K * 3 * PW + 16; e += kLaneSpreadThreads) lw[e] = 0.f;
    __syncthreads();
    for (int pass = 0; pass < npass; ++pass) {
        const int sb = pass * 8 + wave;
        const bool active = sb < nsb;
        const int o2 = (sb % ns2) * 4, o1 = ((sb / ns2) % ns1) * 4, o0 = (sb / (ns2 * ns1)) * 4;
        const int G0 = t0[0] + o0, G1 = t0[1] + o1, G2 = t0[2] + o2;   // sub-block origin
        // acc[m][n]: the 32 x 32 MFMA tile of transform-parts 32 m .. 32 m + 31
        // by sub-block cells 32 n .. 32 n + 31 (cell = 16 z + 4 y + x)
        floatx16 acc[MT][2];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < 2; ++n)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[m][n][v] = 0.f;
        // the next chunk's records, loaded while the current chunk is added up
        int4 nrec_pf[K / 64];
#pragma unroll
        for (int rb = 0; rb < K; rb += 64) {
            const int v = sub.y + rb + (int(threadIdx.x) >> 3);
            nrec_pf[rb / 64] = v < vend ? visits[v] : make_int4(0, 0, 0, 0);
        }
        for (int vb = sub.y; vb < vend; vb += K) {
            // ---- stage the chunk: records, values, weights ----
#pragma unroll
            for (int rb = 0; rb < K; rb += 64) {
                const int t = int(threadIdx.x);
                const int r = rb + (t >> 3), j = t & 7;
                if (vb + r < vend) {
                    const int4 rec = nrec_pf[rb / 64];
                    const int s = rec.x >> 3;
                    if (j == 0) lrec[r] = rec;
                    // the point's 2 NTG floats: 8 per thread, threads j < NTG / 4
                    const float4* src = reinterpret_cast<const float4*>(csort + int64_t(s) * a.nbatch + g0) + 2 * j;
                    float4 u0 = make_float4(0.f, 0.f, 0.f, 0.f), u1 = u0;
                    if (j < NTG / 4) {
                        u0 = src[0];
                        u1 = src[1];
                    }
                    // the 3W weights (rows padded to 16 B in sw): float4 j < ceil(3W/4)
                    float4 wq = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (4 * j < 3 * W) wq = reinterpret_cast<const float4*>(sw + int64_t(s) * a.ws)[j];
                    float4* dst = reinterpret_cast<float4*>(lval + r * 64) + 2 * j;
                    dst[0] = u0;
                    dst[1] = u1;
                    const float wv[4] = {wq.x, wq.y, wq.z, wq.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int k = 4 * j + e;     // weight k = axis k / W, tap k % W
                        if (k < 3 * W) lw[r * 3 * PW + (k / W) * PW + (k % W) + 3] = wv[e];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int rb = 0; rb < K; rb += 64) {
                const int v = vb + K + rb + (int(threadIdx.x) >> 3);
                if (v < vend) nrec_pf[rb / 64] = visits[v];
            }
            const int nrec = min(K, vend - vb);
            for (int rb = 0; active && rb < nrec; rb += 64) {
                // each lane takes one record: the offsets of the point's first
                // support cell from the sub-block origin, wrapped once per axis
                // (each fine-grid axis is >= W + 4); the support meets the
                // sub-block exactly when all three lie in [-(W - 1), 3].  The
                // test is branch-free; lanes past the chunk's records (stale
                // LDS entries, rb + lane < K) are masked out.
                const int4 rec = lrec[rb + lane];
                int oz = rec.y - G0, oy = rec.z - G1, ox = rec.w - G2;
                oz += oz < -(W - 1) ? n0 : 0; oz -= oz > 3 ? n0 : 0;
                oy += oy < -(W - 1) ? n1 : 0; oy -= oy > 3 ? n1 : 0;
                ox += ox < -(W - 1) ? n2 : 0; ox -= ox > 3 ? n2 : 0;
                const bool hit = (rb + lane < nrec) & (unsigned(oz + (W - 1)) <= unsigned(W + 2)) &
                                 (unsigned(oy + (W - 1)) <= unsigned(W + 2)) &
                                 (unsigned(ox + (W - 1)) <= unsigned(W + 2));
                // as padded-row starts 3 - o: the LDS indices of the point's
                // value row and of its z, y, x weights for the sub-block's
                // first cell along each axis
                const int k = lane >> 5, j = lane & 31;
                const int cz = j >> 4, cy = (j >> 2) & 3, cx = j & 3;   // cell j of N-tile 0 (N-tile 1: z + 2)
                // The support meets N-tile 0 (z = 0, 1 of the sub-block) when
                // oz <= 1 and N-tile 1 (z = 2, 3) when oz >= 3 - W; a hit that
                // misses one of them skips its two MFMAs.
                const bool t0 = hit && oz <= 1, t1 = hit && oz >= 3 - W;
                const unsigned long long mb = __ballot(t0 && t1), m0 = __ballot(t0 && !t1),
                                         m1 = __ballot(t1 && !t0);
                const int nb = __popcll(mb), nt0 = __popcll(m0), nt1 = __popcll(m1);
                const int base0 = (nb + 1) & ~1, base1 = base0 + ((nt0 + 1) & ~1);
                int4* hl = lhit[wave];
                if (hit) {
                    const unsigned long long mm = t0 && t1 ? mb : (t0 ? m0 : m1);
                    const int base = t0 && t1 ? 0 : (t0 ? base0 : base1);
                    const int wrow = (rb + lane) * 3 * PW;
                    // rank within its run = its run's hits in lower lanes: visit order
                    const int rank = __builtin_amdgcn_mbcnt_hi(unsigned(mm >> 32),
                                                              __builtin_amdgcn_mbcnt_lo(unsigned(mm), 0u));
                    hl[base + rank] = make_int4((rb + lane) * 64, wrow + 3 - oz, wrow + PW + 3 - oy,
                                                wrow + 2 * PW + 3 - ox);
                }
                // an odd run ends with a zero-weight entry (any staged value row)
                const int4 pad = make_int4(rb * 64, ZROW, ZROW, ZROW);
                if (lane == 0 && (nb & 1)) hl[nb] = pad;
                if (lane == 1 && (nt0 & 1)) hl[base0 + nt0] = pad;
                if (lane == 2 && (nt1 & 1)) hl[base1 + nt1] = pad;
                __builtin_amdgcn_wave_barrier();
                // each run's hits in visit order, two per MFMA step: lanes
                // 0-31 carry hit 2q and lanes 32-63 hit 2q + 1
                auto load = [&](int e, float& a0, float& a1, float& wz0, float& wz1, float& wyx) {
                    const int4 h = hl[e + k];
                    a0 = lval[h.x + j];
                    a1 = MT == 2 ? lval[h.x + 32 + j] : 0.f;
                    wz0 = lw[h.y + cz];
                    wz1 = lw[h.y + cz + 2];
                    wyx = lw[h.z + cy] * lw[h.w + cx];
                };
                for (int q = 0; q < nb; q += 2) {
                    float a0, a1, wz0, wz1, wyx;
                    load(q, a0, a1, wz0, wz1, wyx);
                    const float b0 = wz0 * wyx, b1 = wz1 * wyx;
                    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
                    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
                    if constexpr (MT == 2) {
                        acc[MT - 1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[MT - 1][0], 0, 0, 0);
                        acc[MT - 1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[MT - 1][1], 0, 0, 0);
                    }
                }
                for (int q = 0; q < nt0; q += 2) {
                    float a0, a1, wz0, wz1, wyx;
                    load(base0 + q, a0, a1, wz0, wz1, wyx);
                    const float b0 = wz0 * wyx;
                    acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
                    if constexpr (MT == 2)
                        acc[MT - 1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[MT - 1][0], 0, 0, 0);
                }
                for (int q = 0; q < nt1; q += 2) {
                    float a0, a1, wz0, wz1, wyx;
                    load(base1 + q, a0, a1, wz0, wz1, wyx);
                    const float b1 = wz1 * wyx;
                    acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
                    if constexpr (MT == 2)
                        acc[MT - 1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[MT - 1][1], 0, 0, 0);
                }
                // the next group's entries overwrite this list
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
        }
        if (!active) continue;
        // ---- write the sub-block's cells of this tile ----
        // 32 x 32 MFMA output: register v of lane l holds row
        // 8 (v / 4) + 4 (l / 32) + v % 4 and column l % 32 of its tile, so
        // registers v and v + 1 (v even) hold the real and imaginary parts of
        // one transform at one cell: each lane stores whole complex values.
        // (An opaque copy of the lane index keeps the compiler from hoisting
        // these 64 addresses out of the loops into registers.)
        int lw_ = lane;
        asm volatile("" : "+v"(lw_));
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < 2; ++n)
#pragma unroll
                for (int v = 0; v < 16; v += 2) {
                    const int tr = 16 * m + 4 * (v >> 2) + 2 * (lw_ >> 5) + ((v & 3) >> 1);   // transform
                    const int cell_sb = 32 * n + (lw_ & 31);
                    const int lz = o0 + (cell_sb >> 4), ly = o1 + ((cell_sb >> 2) & 3), lx = o2 + (cell_sb & 3);
                    if (tr >= ng || lz >= T0 || ly >= T1 || lx >= T2) continue;
                    const int cell = (lz * T1 + ly) * T2 + lx;
                    float2* dst;
                    if (sub.w != 0)
                        dst = a.partial + (int64_t(sub.w - 1) * a.nbatch + g0 + tr) * cells + cell;
                    else
                        dst = a.fine + int64_t(g0 + tr) * G +
                              (int64_t(t0[0] + lz) * n1 + (t0[1] + ly)) * n2 + (t0[2] + lx);
                    *dst = make_float2(acc[m][n][v], acc[m][n][v + 1]);
                }
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
        // slots in order; 8 loads in flight at a time (the kooshball centre
        // splits into ~200 slots, which one dependent load at a time made
        // latency-bound)
        const V* src = partial + (int64_t(t.y) * nbatch + b) * cells + c;
        const int64_t slot = int64_t(nbatch) * cells;
        V acc = cmake<V>(R(0), R(0));
        int k = 0;
        for (; k + 8 <= t.z; k += 8) {
            V v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = src[(k + u) * slot];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = cadd(acc, v[u]);
        }
        for (; k < t.z; ++k) acc = cadd(acc, src[k * slot]);
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
        store_point(a, int64_t(t), orig, acc);
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
        V c = a.cdata[int64_t(t) * a.M + orig];
        if (a.pw) c = cmulc(c, a.pw[orig]);
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

// csort[sinv[j]][b] = c[b][j] (times conj(pw[j]) when pw is given): the
// nonuniform values in sorted point order, transforms fastest.  A block reads 64 consecutive points of up to 32
// transforms (coalesced rows), transposes them in LDS and writes each point's
// transforms as one contiguous run.
template <class V>
__global__ void __launch_bounds__(256) sort_values_kernel(const V* __restrict__ c, V* __restrict__ csort,
                                                          const int* __restrict__ sinv,
                                                          const V* __restrict__ pw, int64_t M, int B) {
    __shared__ V tile[64][33];
    const int64_t j0 = int64_t(blockIdx.x) * 64;
    const int lane = threadIdx.x & 63, k0 = threadIdx.x >> 6;
    const int64_t j = j0 + lane;
    for (int b0 = 0; b0 < B; b0 += 32) {
        const int nbc = min(32, B - b0);
        if (j < M) {
            const V f = pw ? cconj(pw[j]) : cmake<V>(1, 0);
            for (int k = k0; k < nbc; k += 4) {
                const V v = c[int64_t(b0 + k) * M + j];
                tile[lane][k] = pw ? cmul(v, f) : v;
            }
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < 64 * nbc; idx += 256) {
            const int pnt = idx / nbc, k = idx - pnt * nbc;
            if (j0 + pnt < M) csort[int64_t(sinv[j0 + pnt]) * B + b0 + k] = tile[pnt][k];
        }
        __syncthreads();
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
// ---------------------------------------------------------------------------
// Trajectory gradient by kernel-derivative interpolation.  With the fine grid
// F_t = FFT(deapodised, padded h_t) of a type-2 plan and u_a = k_a nf_a / 2 pi,
//   d/dk_a (type2(h)_t)_j = nf_a / (2 pi) sum_l d_a phi(u_j - l) F_t[l],
// where d_a phi differentiates the window along slot a only.  Each thread
// takes one sorted point, forms the three derivatives separably (x-row sums
// of phi and phi', then y, then z) for every transform t and accumulates
//   out[a][j] (+)= Re sum_t conj(w_t[j]) d/dk_a (type2(h)_t)_j.
// ---------------------------------------------------------------------------
template <class V>
struct CoordGradArgs {
    using R = decltype(V::x);
    const V* fine;          // [nbatch, nf]
    const V* w;             // [nbatch, M] original point order
    R* out;                 // [dim, M] original point order
    const int* sidx;
    const int* sl0;         // [3][M]
    const R* sfrac;         // [3][M]
    const R* sw;            // [M][ws]
    int64_t M;
    int nbatch, ws, dim, accumulate;
    double beta;
    Geo g;
};

template <class V, int W>
__global__ void __launch_bounds__(256) coord_grad_kernel(CoordGradArgs<V> a) {
    using R = decltype(V::x);
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.M) return;
    const R ihw = R(2.0 / double(W));
    const R beta = R(a.beta);
    R wv[3][W], dw[3][W];
    int idx[3][W];
    for (int d = 0; d < 3; ++d) {
        const int l0 = a.sl0[int64_t(d) * a.M + s];
        const R fr = a.sfrac[int64_t(d) * a.M + s];
        const int n = int(a.g.nf[d]);
#pragma unroll
        for (int q = 0; q < W; ++q) {
            int l = l0 + q;
            l -= l >= n ? n : 0;
            idx[d][q] = l;
            wv[d][q] = a.sw[s * a.ws + d * W + q];
            dw[d][q] = a.g.wa[d] == 1 ? R(0) : kb_dphi<R>(fr - R(q), ihw, beta);
        }
    }
    const int orig = a.sidx[s];
    const int64_t G = a.g.nf[0] * a.g.nf[1] * a.g.nf[2];
    R acc[3] = {R(0), R(0), R(0)};
    for (int t = 0; t < a.nbatch; ++t) {
        const V* g = a.fine + int64_t(t) * G;
        V Dz = cmake<V>(R(0), R(0)), Dy = Dz, Dx = Dz;
        for (int qz = 0; qz < a.g.wa[0]; ++qz) {
            V S = cmake<V>(R(0), R(0)), Sdy = S, Sdx = S;
#pragma unroll
            for (int qy = 0; qy < W; ++qy) {
                if (qy < a.g.wa[1]) {
                    const V* row = g + (int64_t(idx[0][qz]) * a.g.nf[1] + idx[1][qy]) * a.g.nf[2];
                    V sx = cmake<V>(R(0), R(0)), sdx = sx;
#pragma unroll
                    for (int qx = 0; qx < W; ++qx) {
                        const V f = row[idx[2][qx]];
                        sx.x += wv[2][qx] * f.x;
                        sx.y += wv[2][qx] * f.y;
                        sdx.x += dw[2][qx] * f.x;
                        sdx.y += dw[2][qx] * f.y;
                    }
                    S.x += wv[1][qy] * sx.x;
                    S.y += wv[1][qy] * sx.y;
                    Sdy.x += dw[1][qy] * sx.x;
                    Sdy.y += dw[1][qy] * sx.y;
                    Sdx.x += wv[1][qy] * sdx.x;
                    Sdx.y += wv[1][qy] * sdx.y;
                }
            }
            Dz.x += dw[0][qz] * S.x;
            Dz.y += dw[0][qz] * S.y;
            Dy.x += wv[0][qz] * Sdy.x;
            Dy.y += wv[0][qz] * Sdy.y;
            Dx.x += wv[0][qz] * Sdx.x;
            Dx.y += wv[0][qz] * Sdx.y;
        }
        const V c = a.w[int64_t(t) * a.M + orig];
        // Re(conj(c) D) = c.x D.x + c.y D.y
        acc[0] += c.x * Dz.x + c.y * Dz.y;
        acc[1] += c.x * Dy.x + c.y * Dy.y;
        acc[2] += c.x * Dx.x + c.y * Dx.y;
    }
    for (int d = 3 - a.dim; d < 3; ++d) {
        const R v = acc[d] * R(double(a.g.nf[d]) * 0.15915494309189533576888376337251);
        R* o = a.out + int64_t(d - (3 - a.dim)) * a.M + orig;
        *o = a.accumulate ? *o + v : v;
    }
}

template <class V> void* nufft_kernel_ptr(int which, int w);
// interpolation kernel with the tile fixed at compile time (16^3 in 3-D,
// 32^2 in 2-D), or nullptr when this width has none
template <class V> void* nufft_interp_fixed_ptr(int dim, int w);
// paired interpolation kernel (complex64, standard 3-D tile, w <= 6), or NULL
void* nufft_interp_pair_ptr(int dim, int w);
// spreading kernel for CG transforms per block (1, 2, 4, 8, 16, 32 / 16 for complex128)
template <class V> void* nufft_spread_ptr(int cg);
// complex64 spreading with the 16 or 32 transforms of a block on the MFMA
// rows (spread_lanes_kernel), or nullptr for widths without an instance
void* nufft_spread_lanes_ptr(int w, int group);

}  // namespace hasty
