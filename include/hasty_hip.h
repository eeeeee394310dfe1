/*
 * hasty_hip.h -- C-ABI of the MI355X (gfx950) non-Cartesian reconstruction path.
 *
 * Plain pointers and sizes only.  Every device pointer is caller-owned memory
 * on the current HIP device; every call enqueues work on the given stream
 * (NULL = the default stream) and returns without a device-wide sync unless
 * stated.  Status 0 = OK; on failure hasty_last_error() returns a message
 * (thread-local).
 *
 * Layouts (C-order, last axis fastest), complex data interleaved (re, im):
 *   image / modes   [ntransf, n0, n1, n2]     (n0 slowest)
 *   nonuniform data [ntransf, M]
 *   coordinates     dim separate real arrays [M], radians, coords[a] <-> axis a
 *   Toeplitz kernel [2 n0, 2 n1, 2 n2]
 *   coil maps       [C, n0, n1, n2]
 * Complex dtypes: c64 = 2 x float, c128 = 2 x double.
 */
#ifndef HASTY_HIP_H
#define HASTY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HASTY_OK 0
#define HASTY_ERR_INVALID 1   /* bad argument / shape / enum            */
#define HASTY_ERR_ALLOC 2     /* device or host allocation failed       */
#define HASTY_ERR_HIP 3       /* a HIP runtime call failed              */
#define HASTY_ERR_STATE 4     /* plan misuse (execute before setpts ...) */
#define HASTY_ERR_INTERNAL 5

/* ---- enums mirrored from the reference --------------------------------- */
/* nufft_type (nufft.cppm:42-48) */
#define HASTY_TYPE_1 1        /* BACKWARD: nonuniform -> modes (spread)  */
#define HASTY_TYPE_2 2        /* FORWARD:  modes -> nonuniform (interp)  */
/* nufft_method_cuda (nufft.cppm:65-70; the reference maps it onto cufinufft's
 * gpu_method / gpu_sort at nufft.cppm:118-139).  The values are the
 * reference's, so 0 is GM_NO_SORT, not "auto": a caller that zero-fills
 * hasty_nufft_opts must set `method` (or start from hasty_nufft_default_opts).
 *   GM_NO_SORT, GM_SORT: global-memory kernels, one thread per point over the
 *     bin-sorted points (the sort also holds the precomputed window weights, so
 *     GM_NO_SORT runs the same kernels as GM_SORT).  Type 1 adds into the fine
 *     grid with FP32 / FP64 global atomics, so its sums are not in a fixed
 *     order and its results are not bitwise reproducible.
 *   SM, DEFAULT: the LDS-tiled kernels (type 2 stages a padded tile per block;
 *     type 1 gathers each tile's visit list with register accumulators, no
 *     atomics, bitwise reproducible).                                       */
#define HASTY_METHOD_GM_NO_SORT 0
#define HASTY_METHOD_GM_SORT 1
#define HASTY_METHOD_SM 2
#define HASTY_METHOD_DEFAULT 3
/* ToeplitzMultType / ToeplitzAccumulateType (toeplitz.cppm:47-56) */
#define HASTY_MULT_NONE 0
#define HASTY_MULT 1
#define HASTY_MULT_CONJ 2
#define HASTY_ACCUMULATE_NONE 0
#define HASTY_ACCUMULATE 1

/* precision of a plan: real type of coordinates and data */
#define HASTY_F32 0           /* coords float,  data c64  */
#define HASTY_F64 1           /* coords double, data c128 */

typedef struct hasty_nufft_plan_s* hasty_nufft_plan;
typedef struct hasty_normal_op_s* hasty_normal_op;

/* Optional plan tuning (NULL = defaults).  Mirrors the option block of
 * nufft_opts (nufft.cppm:78-95) plus this library's tiling knobs. */
typedef struct hasty_nufft_opts {
    double upsamp;        /* oversampling sigma; <= 0 -> 2.0 (nufft_upsamp_cuda)  */
    int32_t method;       /* HASTY_METHOD_* (nufft_method_cuda); 0 = GM_NO_SORT   */
    int32_t tile[3];      /* fine-grid tile per axis (slots z,y,x); 0 -> auto     */
    int32_t coil_group;   /* transforms per spread/interp block; 0 -> auto        */
    int32_t max_batch;    /* transforms per internal batch; 0 -> all              */
    int32_t max_subproblem; /* points per spread/interp block; 0 -> auto          */
} hasty_nufft_opts;

const char* hasty_last_error(void);
int hasty_version(void);
int hasty_device_synchronize(void);
/* Launch timing for the calling thread: when enabled, every kernel launch of
 * the library is bracketed by a HIP event pair on its stream.  read()
 * synchronises on the recorded events, aggregates elapsed milliseconds and
 * launch counts per kernel name ('\n'-separated in names), and clears. */
int hasty_profile_enable(int on);
int hasty_profile_read(char* names, int64_t names_len, double* ms, int32_t* counts,
                       int32_t max_entries, int32_t* n_entries);

/* ---- NUFFT plan: replaces nufft_plan<cuda_t,TT,DIM,NT> (nufft.cppm:103-310),
 *      the cufinufft(f)_makeplan / setpts / execute / destroy call sites
 *      (nufft.cppm:160-164, 191-198, 214-245, 282-285). ------------------- */
int hasty_nufft_makeplan(int type, int dim, const int64_t* nmodes, int sign, int ntransf,
                         double tol, int precision, const hasty_nufft_opts* opts,
                         void* stream, hasty_nufft_plan* plan_out);
/* coordinates are copied into the plan's bin-sorted layout; the caller's
 * arrays may be freed after the call returns and the stream has run it.
 * c0 pairs with axis 0 (slowest), c_{dim-1} with the fastest axis; unused
 * pointers may be NULL. M == 0 is allowed (execute then writes zeros / nothing). */
int hasty_nufft_setpts(hasty_nufft_plan plan, int64_t M, const void* c0, const void* c1,
                       const void* c2);
/* FINUFFT argument convention: c = nonuniform data [ntransf, M],
 * f = modes [ntransf, n...].  Type 1 reads c and writes f; type 2 reads f
 * and writes c. */
int hasty_nufft_execute(hasty_nufft_plan plan, void* c, void* f);
int hasty_nufft_set_stream(hasty_nufft_plan plan, void* stream);
/* Trajectory gradient (no counterpart in the reference, whose NUFFT is not
 * differentiable; this backs the d/dtrajectory autograd of the Python layer).
 * On a type-2 plan with points set:
 *   out[a][j] = Re sum_t conj(w[t][j]) d/dk_{j,a} (type2(h)_t)_j,
 * h = modes [ntransf, n...], w = nonuniform [ntransf, M], out = real [dim, M]
 * (float for HASTY_F32 plans, double for HASTY_F64) in the caller's point
 * order, by interpolation with the derivative of the window on the fine grid
 * of h.  For y = type2(f): grad_k = out(h = f, w = dL/dy); for f = type1(c)
 * with the opposite sign: grad_k = out(h = dL/df, w = c). */
int hasty_nufft_coord_grad(hasty_nufft_plan plan, const void* h, const void* w, void* out);
int hasty_nufft_destroy(hasty_nufft_plan plan);
/* cufinufft_default_opts (called at nufft.cppm:117): upsamp 2, method
 * DEFAULT, every tiling knob 0 (auto). */
int hasty_nufft_default_opts(hasty_nufft_opts* opts);
/* Kernel width a plan of this precision takes for (tol, upsamp, dim), without
 * making one.  upsamp 2 (or <= 0): w = ceil(-log10(tol / 10)).  Other
 * upsampling: the smallest w whose estimated relative L2 error (aliasing of
 * the window plus the rounding of the plan's arithmetic, which the wide
 * sigma-1.25 windows amplify) is <= tol: upsamp 1.25 gives w = 9 at 1e-5 for
 * HASTY_F64 and w = 6 at 1e-3; HASTY_F32 bottoms out near 6e-5 there.
 * 17 = no width up to 16 reaches tol.  makeplan rejects widths above 9
 * (complex64) or 14 (complex128). */
int hasty_nufft_kernel_width(double tol, double upsamp, int dim, int precision, int* width);
/* introspection: width, beta, fine grid, tiles, device bytes held */
int hasty_nufft_info(hasty_nufft_plan plan, int* width, double* beta, int64_t* nf3,
                     int32_t* tile3, int64_t* device_bytes);

/* ---- Toeplitz ----------------------------------------------------------- */
/* toeplitz_kernel (nufft.cppm:386-441): kernel[2n...] = fftn(embed(psf)) / prod(n),
 * psf[m] = sum_j w_j e^{+i k_j.m}, m in [-(n-1), n-1], lag m stored at index
 * m + n - 1 (index 2n-1 stays zero).  weights = c64 [M] or NULL (all ones).
 * The PSF is computed with an FP64 type-1 NUFFT at tolerance tol (<= 0 ->
 * 1e-9) and an FP64 FFT, then rounded to c64. coords are float [M]. */
int hasty_toeplitz_kernel(int dim, const int64_t* n, int64_t M, const void* c0,
                          const void* c1, const void* c2, const void* weights, void* kernel,
                          double tol, void* stream);
/* toeplitz_multiply (nufft.cppm:448-479): inout[b] <- crop_{n-1}(ifftn(fftn_{2n}(inout[b]) * kernel)),
 * batch >= 1, c64. */
int hasty_toeplitz_multiply(int dim, const int64_t* n, int64_t batch, void* inout,
                            const void* kernel, void* stream);
/* transform_toeplitz_kernel (toeplitz.cppm:510-609): the kernel is kept in
 * natural frequency order here, so this is the identity (validates shape). */
int hasty_transform_toeplitz_kernel(int dim, const int64_t* n, void* kernel, void* stream);
/* toeplitz_multiplication (toeplitz.cppm:611-712, load/store rules of
 * toeplitz_load_3D.cu:18-79): for each b in [0,batch)
 *   s      = input[b] (.)in_m1 mult1 (.)in_m2 mult2
 *   out[b] = crop(T s) (.)out_m1 mult1 (.)out_m2 mult2 (.)io input[b]  (+ out[b])
 * types[5] = {input_output, input_mult1, output_mult1, input_mult2, output_mult2}
 * mult1 / mult2 may be NULL (then their types are ignored); scratch may be
 * NULL (the library allocates stream-ordered) or caller memory of
 * scratch_bytes >= hasty_toeplitz_scratch_bytes(dim, n, batch). */
int hasty_toeplitz_multiplication(int dim, const int64_t* n, int64_t batch, const void* input,
                                  void* output, const void* kernel, void* scratch,
                                  int64_t scratch_bytes, const void* mult1, const void* mult2,
                                  const int32_t* types, int accumulate, void* stream);
int64_t hasty_toeplitz_scratch_bytes(int dim, const int64_t* n, int64_t batch);

/* ---- SENSE normal operator: replaces NORMAL_T_T1_OP (normal.cppm:43-93):
 *      out = sum_c conj(S_c) . T(S_c . x), all coils batched per launch,
 *      coil sum fused into the last pass (no [C, n...] intermediate). ------ */
int hasty_normal_create(int dim, const int64_t* n, int ncoils, const void* smaps,
                        const void* kernel, int coil_batch, void* stream, hasty_normal_op* op);
/* ---- multi-kernel / whitened normal operators, kind 1 (image -> image):
 *      replaces NORMAL_IDT_T1_OP (normal.cppm:185-272) and NORMAL_IDTW_T1_OP
 *      (normal.cppm:405-564):
 *        out = sum_c conj(S'_c) . sum_l conj(R_l) . T_l(R_l . S_c . x)
 *      kernels [L, 2n...], kerneldiags R [L, n...] (NULL = all ones),
 *      weights W [C, C] (NULL = identity; else S'_c = sum_i conj(W_ic) S_i,
 *      formed once here, which equals the reference's sum_i conj(S_i) sum_j W_ij stack_j).
 *      smaps, kernels, kerneldiags and weights are caller memory, read by the
 *      applies (weights and smaps also here); all c64.  Apply with
 *      hasty_normal_apply(op, x [n...], out [n...]). */
int hasty_normal_idt_create(int dim, const int64_t* n, int ncoils, const void* smaps, int nkernels,
                            const void* kernels, const void* kerneldiags, const void* weights,
                            int coil_batch, void* stream, hasty_normal_op* op);
/* ---- kind 2 (stacked images -> stacked images): replaces NORMAL_T_T2_OP
 *      (normal.cppm:113-160), NORMAL_IDT_T2_OP (normal.cppm:292-375) and
 *      NORMAL_IDTW_T2_OP (normal.cppm:586-735):
 *        out_i = sum_j W_ij conj(X) . sum_l conj(R_l) . T_l(R_l . X . Psi_j)
 *      x = X [n...]; weights [B, B] or NULL; apply with
 *      hasty_normal_apply(op, Psi [B, n...], out [B, n...]); stack_batch
 *      caps the stacked images per launch (0 = all). */
int hasty_normal_t2_create(int dim, const int64_t* n, int nstack, const void* x, int nkernels,
                           const void* kernels, const void* kerneldiags, const void* weights,
                           int stack_batch, void* stream, hasty_normal_op* op);
/* ---- batched dynamic frames (no counterpart in the reference, which builds
 *      one NORMAL_T_T1_OP per frame, normal.cppm:43-93): F frames that share
 *      the coil maps, frame f with its own kernel,
 *        out[f] = sum_c conj(S_c) . T_f(S_c . x[f]),
 *      kernels [F, 2n...]; apply with hasty_normal_apply(op, x [F, n...],
 *      out [F, n...]).  On the fused 3-D passes (n x n planes, n = 64, 96,
 *      128, >= 2 coils per launch) every pass is one launch over all frames
 *      and coils; other shapes run the frames one after another.  Frame f's
 *      result is bitwise that of hasty_normal_create on kernel f with the
 *      same coil_batch. */
int hasty_normal_frames_create(int dim, const int64_t* n, int nframes, int ncoils, const void* smaps,
                               const void* kernels, int coil_batch, void* stream, hasty_normal_op* op);
/* x and out must be distinct buffers. */
int hasty_normal_apply(hasty_normal_op op, const void* x, void* out);
int hasty_normal_set_stream(hasty_normal_op op, void* stream);
int hasty_normal_destroy(hasty_normal_op op);

/* ---- time-segmented (off-resonance) SENSE over a NUFFT plan whose
 *      transforms are a chunk of coils (coils run in chunks of the plan's
 *      ntransf; data in the plan's precision). bil [L, M] and cjl [L, n...]
 *      may both be NULL when L == 1 (plain SENSE).
 * Replaces sense_forward_offresonance (forward.cppm:36-165), on a type-2 plan:
 *   kdata[c] = sum_l bil[l] . type2(cjl[l] . smaps[c] . image)      [C, M]
 * with the products folded into the plan's load and interpolation passes. */
int hasty_sense_forward(hasty_nufft_plan plan, int ncoils, int L, const void* image,
                        const void* smaps, const void* bil, const void* cjl, void* kdata);
/* Replaces sense_adjoint_offresonance (adjoint.cppm:33-171), on a type-1 plan:
 *   images[c] = conj(smaps[c]) . sum_l conj(cjl[l]) . type1(conj(bil[l]) . kdata[c])
 * [C, n...]; with combine != 0 the coils are summed into one image [n...]
 * inside the last FFT pass.  (The reference applies conj(smaps) twice when
 * the NUFFT and data precisions match, adjoint.cppm:116-117; once here.) */
int hasty_sense_adjoint(hasty_nufft_plan plan, int ncoils, int L, const void* kdata,
                        const void* smaps, const void* bil, const void* cjl, void* images,
                        int combine);

/* ---- CG / vector ops (cg.cppm:122-184, alg.cppm:13-30) with FP64
 *      accumulation.  state = caller device buffer of HASTY_CG_STATE_DOUBLES
 *      doubles; all scalars stay on the device. ------------------------------ */
#define HASTY_CG_STATE_DOUBLES 1040
/* out[0] = Re sum conj(a) b, out[1] = Im (device doubles), n complex elements */
int hasty_dot(const void* a, const void* b, int64_t n, double* out, void* stream);
/* hasty_dot without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * hasty_dot, bit for bit (same kernels and summation order); its partial sums
 * live in a static 8 KiB device array per GPU (2 x 512 doubles, enough for
 * any n) instead of a hipMallocAsync buffer, and the call allocates nothing.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams, or host threads without ordering) race. */
int hasty_dot_noalloc(const void* a, const void* b, int64_t n, double* out, void* stream);
/* state <- {rz = ||r||^2, done = sqrt(rz) < tol}; p <- r */
int hasty_cg_init(void* r, void* p, int64_t n, double tol, double* state, void* stream);
/* one CG update given Ap: pAp, alpha, x += alpha p, r -= alpha Ap, rz', p = r + (rz'/rz) p.
 * No-op once state.done is set.  */
int hasty_cg_step(void* x, void* r, void* p, const void* Ap, int64_t n, double* state,
                  void* stream);
/* hasty_cg_init / hasty_cg_step on nsys independent systems in one launch
 * per kernel (the batched frames; the reference runs cg.cppm:122-184 once
 * per system): vectors [nsys, n] (system s at offset s*n), states
 * [nsys, HASTY_CG_STATE_DOUBLES].  System s gets, bit for bit, what the
 * single-system call gives on its own vectors; each system stops on its own
 * done flag. */
int hasty_cg_init_batched(void* r, void* p, int64_t n, int nsys, double tol, double* state, void* stream);
int hasty_cg_step_batched(void* x, void* r, void* p, const void* Ap, int64_t n, int nsys, double* state,
                          void* stream);
/* Preconditioned CG with restarts (cg.cppm:16-119), as three calls around the
 * caller's A and P applies: z = P(r) comes from the caller between calls.
 *   pcg_init:      rz = Re<r, z>, p <- z, done = sqrt(rz) < tol
 *   pcg_update:    pAp = Re<p, Ap>, alpha = rz / pAp, x += alpha p, r -= alpha Ap
 *   pcg_direction: rz' = Re<r, z>; p <- z if restart, else z + (rz'/rz) p;
 *                  rz = rz', done = sqrt(rz) < tol
 * FP64 reductions and scalars on the device; update / direction are no-ops
 * once done is set.  state as for hasty_cg_*; hasty_cg_status reads it. */
int hasty_pcg_init(const void* r, const void* z, void* p, int64_t n, double tol, double* state,
                   void* stream);
int hasty_pcg_update(void* x, void* r, const void* p, const void* Ap, int64_t n, double* state,
                     void* stream);
int hasty_pcg_direction(const void* r, const void* z, void* p, int64_t n, int restart, double* state,
                        void* stream);
/* host read of residual and convergence flag (synchronises the stream) */
int hasty_cg_status(const double* state, double* resid, int* done, void* stream);
/* y <- a * x (a real FP64 scalar from device memory, scaled by 1/a_ptr[0] if invert) */
int hasty_scale_by_device(void* x, int64_t n, const double* a, int invert, void* stream);
/* out[0] = ||x||_2 (FP64, device) */
int hasty_norm(const void* x, int64_t n, double* out, void* stream);
/* y <- y + a x over n c64 elements (the b - A x and x + alpha p updates that
 * cg.cppm:154-157 writes as tensor expressions) */
int hasty_axpy(double a_re, double a_im, const void* x, void* y, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HASTY_HIP_H */
