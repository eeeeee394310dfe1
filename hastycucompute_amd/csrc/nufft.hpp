// NUFFT plan (type 1 spread / type 2 interp) for gfx950.
//
// Replaces nufft_plan<cuda_t,TT,DIM,NT> (nufft.cppm:103-310), which wraps
// cufinufft.  Internally every transform is 3-D with slots (z, y, x); a 2-D
// transform has n_z = nf_z = 1.
//
// Data layout on the device (per internal batch of transforms):
//   fine  [B, nf_z, nf_y, nf_x]   oversampled grid (x fastest)
//   bufB  [B, n_z,  nf_y, nf_x]   after the z-axis pass (pruned)
//   bufA  [B, n_z,  n_y,  nf_x]   after the y-axis pass (pruned)
// Points are kept bin-sorted by fine-grid tile (T_z x T_y x T_x cells):
//   sidx[s]          original index of sorted point s
//   sinv[j]          sorted position of original point j
//   sl0[a][s]        first support cell along slot a (wrapped into [0, nf_a))
//   sfrac[a][s]      u_a - l0_a in (w/2 - 1, w/2]  (fractional offset, R)
//   sw[s][a][q]      phi(sfrac[a][s] - q), q < w   (window weights, from FP64)
//   visits[v]        (8 s + tile-neighbour bits, l0) for spreading: tile-major
//                    list of the points whose support touches each tile.
#pragma once

#include <vector>

#include "common.hpp"
#include "fft.hpp"

namespace hasty {

constexpr int kMaxGroupSpread = 32;   // transforms per spreading block (registers per lane)
constexpr int kMaxGroupInterp = 4;    // transforms staged per interpolation block

struct NufftPlan {
    int type = 1, dim = 3, sign = 1, ntransf = 1;
    bool f64 = false;
    double tol = 1e-5, sigma = 2.0, beta = 0;
    int w = 6;
    int64_t n[3] = {1, 1, 1};
    int64_t nf[3] = {1, 1, 1};
    int T[3] = {1, 1, 1};
    int ntile[3] = {1, 1, 1};
    int64_t ntiles = 1;
    int group = 1;
    int max_batch = 1;
    int max_sub = 4096;
    hipStream_t stream = nullptr;
    int device = 0;

    void* deapod[3] = {nullptr, nullptr, nullptr};
    void* fine = nullptr;
    void* bufA = nullptr;
    void* bufB = nullptr;

    int64_t M = 0;
    bool have_pts = false;
    bool gm = false;             // global-memory spread / interp kernels
    bool pair = false;           // type 2: paired interpolation (two transforms per block)
    int* sidx = nullptr;
    int* sinv = nullptr;
    int* sl0 = nullptr;          // [3][M]
    void* sfrac = nullptr;       // [3][M] of R
    void* sw = nullptr;          // [M][3][w] of R: window weights per slot
    void* csort = nullptr;       // [M][batch] sorted nonuniform data (spreading)
    int64_t csort_cap = 0;
    int4* visits = nullptr;       // (8 s + neighbour bits, l0z, l0y, l0x) per tile visit
    int64_t nvisits = 0;
    int4* sub_interp = nullptr;
    int nsub_interp = 0;
    int4* sub_spread = nullptr;
    int nsub_spread = 0;
    int4* split_info = nullptr;  // (tile, first partial slot, slot count, -) per split tile
    int nsplit = 0;
    int nslots = 0;              // partial slots (one per subproblem of a split tile)
    void* partial = nullptr;     // [nslots][batch][tile cells] partial spreading tiles
    int64_t partial_cap = 0;
    int64_t pts_cap = 0, visits_cap = 0;
    int64_t sub_interp_cap = 0, sub_spread_cap = 0, split_cap = 0;

    int64_t device_bytes = 0;

    ~NufftPlan();
};

NufftPlan* nufft_makeplan(int type, int dim, const int64_t* nmodes, int sign, int ntransf,
                          double tol, bool f64, const hasty_nufft_opts* opts, hipStream_t stream);
void nufft_setpts(NufftPlan* p, int64_t M, const void* c0, const void* c1, const void* c2);
void nufft_execute(NufftPlan* p, void* c, void* f);

// Diagonal factors folded into one execution of a plan (the SENSE operators).
// Arrays have the plan's complex type; image-shaped ones have row stride n_x
// and batch stride `*_sb` (0 = the same array for every transform).
//   type 2: the mode array read is  f_b (.) m1 (.) m2   (types MULT / MULT_CONJ)
//           and  c_b[j] = pw[j] * interp_b(j)  (+ c_b[j] if accumulate);
//   type 1: the spread input is  conj(pw[j]) * c_b[j]  and the stored image is
//           f_b = result (.) m1 (.) m2 (.) io  (+ f_b if accumulate), summed
//           over the transforms into one image when sum_batch.
struct NufftHooks {
    const void* m1 = nullptr; int64_t m1_sb = 0; int m1_type = HASTY_MULT_NONE;
    const void* m2 = nullptr; int64_t m2_sb = 0; int m2_type = HASTY_MULT_NONE;
    const void* io = nullptr; int64_t io_sb = 0; int io_type = HASTY_MULT_NONE;
    const void* pw = nullptr;
    int accumulate = 0;
    int sum_batch = 0;
};
// Execute the first `ntransf` (<= plan ntransf) transforms with hooks.
void nufft_execute_hooked(NufftPlan* p, void* c, void* f, const NufftHooks& hk, int ntransf);

// Time-segmented SENSE (off-resonance) operators over a plan whose transforms
// are a chunk of coils, L segments:
//   forward: y_c = sum_l bil_l (.) type2(cjl_l (.) S_c (.) x)          [C, M]
//   adjoint: x_c = conj(S_c) (.) sum_l conj(cjl_l) (.) type1(conj(bil_l) (.) y_c)
//            [C, n...], or summed over coils into one image when `combine`.
// bil [L, M] and cjl [L, n...] may be NULL (all ones).
void sense_offres_forward(NufftPlan* p, int ncoils, int L, const void* image, const void* smaps,
                          const void* bil, const void* cjl, void* kdata);
void sense_offres_adjoint(NufftPlan* p, int ncoils, int L, const void* kdata, const void* smaps,
                          const void* bil, const void* cjl, void* images, int combine);
// out[a][j] = Re sum_t conj(w[t][j]) d/dk_a (type2(h)_t)_j on a type-2 plan
void nufft_coord_grad(NufftPlan* p, const void* h, const void* w, void* out);

// Pruned multi-dimensional FFT helpers shared with the Toeplitz code.
// Geometry in slots (z, y, x): image extent n, FFT length L, shifts.
struct PadGeometry {
    int64_t n[3];
    int64_t L[3];
    int shift_in[3];    // image index j -> line position (j + shift) mod L
    int shift_out[3];   // output index j <- line position (j + shift) mod L
};

// img [B, n] (row loads with optional diagonals and deapodisation) ->
// grid [B, L] = FFT_sign(pad(img)), using bufA [B, n_z, n_y, L_x] and
// bufB [B, n_z, L_y, L_x].
template <class V>
void pruned_forward(const PadGeometry& g, int64_t B, const V* img, V* grid, V* bufA, V* bufB,
                    int sign, const DiagOp<V>& m1, const DiagOp<V>& m2,
                    const decltype(V::x)* const* deapod, hipStream_t s);

// grid [B, L] -> img [B, n] = crop(FFT_sign(grid)) (.) diagonals, deapod,
// scale, optional accumulate / sum over B.
template <class V>
void pruned_inverse(const PadGeometry& g, int64_t B, const V* grid, V* img, V* bufA, V* bufB,
                    int sign, decltype(V::x) scale, const DiagOp<V>& m1, const DiagOp<V>& m2,
                    const DiagOp<V>& io, const decltype(V::x)* const* deapod, int accumulate,
                    int sum_batch, hipStream_t s);

template <class V> ColArgs<V> make_col_args(int L);
template <class V> int rows_per_block(int L, int out_len);

}  // namespace hasty
