/*
 * OpenMP Kaiser-Bessel spread / interpolate on a periodic fine grid.
 * TEST INFRASTRUCTURE ONLY: the CPU oracle's gridding step and the timed CPU
 * baseline.  Nothing in the product links this file.
 *
 * Restates the gridding step of the NUFFT that the reference delegates to
 * cufinufft (call sites nufft.cppm:160-164,191-198,214-245): a type-1
 * transform spreads every nonuniform value onto the w^d fine-grid cells
 * around its point with a separable window, a type-2 transform interpolates
 * with the same window.  The window here is Kaiser-Bessel,
 *     phi(t) = I0(beta * sqrt(1 - (2t/w)^2)),  |t| <= w/2,
 * with I0 summed from its power series sum_k (x^2/4)^k / (k!)^2 (Abramowitz &
 * Stegun 9.6.12) until the term is below 1e-17 of the sum.
 *
 * Grid: C-order [nf0][nf1][nf2] (axis 0 slowest), complex interleaved.
 * Points: fine-grid coordinates u_a in [0, nf_a) (already folded).
 * Support of a point along axis a: cells l0 .. l0+w-1 with l0 = ceil(u - w/2),
 * taken modulo nf_a.
 *
 * Spreading is parallel over z-slabs of height >= w processed in two colours
 * (even / odd slabs, plus the wrapped last slab alone when the slab count is
 * odd), so no two threads ever add into the same cell and the result is
 * deterministic.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double bessel_i0(double x)
{
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 200; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

static double kb_window(double t, int w, double beta)
{
    double z = 2.0 * t / (double)w;
    z = 1.0 - z * z;
    if (z < 0.0) return 0.0;
    return bessel_i0(beta * sqrt(z));
}

static inline int64_t wrap(int64_t l, int64_t n)
{
    l %= n;
    return l < 0 ? l + n : l;
}

/* Per-point support: first cell (wrapped per cell later) and w weights per axis. */
static void point_weights(int d, const int64_t *nf, const double *const *u, int64_t j,
                          int w, double beta, int64_t l0[3], double wt[3][32])
{
    for (int a = 0; a < d; ++a) {
        double uj = u[a][j];
        int64_t first = (int64_t)ceil(uj - 0.5 * w);
        l0[a] = first;
        for (int q = 0; q < w; ++q) wt[a][q] = kb_window(uj - (double)(first + q), w, beta);
    }
    for (int a = d; a < 3; ++a) {
        l0[a] = 0;
        wt[a][0] = 1.0;
    }
    (void)nf;
}

#define DEFINE_KB(REAL, SUFFIX)                                                              \
void kb_interp_##SUFFIX(int d, const int64_t *nf, int ntransf, const REAL *grid,          \
                        int64_t M, const double *u0, const double *u1, const double *u2,   \
                        int w, double beta, REAL *out)                                     \
{                                                                                          \
    const double *u[3] = {u0, u1, u2};                                                     \
    int64_t n[3] = {1, 1, 1};                                                              \
    for (int a = 0; a < d; ++a) n[3 - d + a] = nf[a];                                      \
    int64_t G = n[0] * n[1] * n[2];                                                        \
    int wz = d == 3 ? w : 1, wy = d >= 2 ? w : 1;                                          \
    _Pragma("omp parallel for schedule(dynamic, 256)")                                     \
    for (int64_t j = 0; j < M; ++j) {                                                      \
        int64_t l0[3];                                                                     \
        double wt[3][32];                                                                  \
        point_weights(d, nf, u, j, w, beta, l0, wt);                                       \
        /* re-index so that axis slots are (z, y, x) with unused leading slots */          \
        int64_t L[3] = {0, 0, 0};                                                          \
        double *W[3];                                                                      \
        double one = 1.0;                                                                  \
        for (int s = 0; s < 3; ++s) W[s] = &one;                                           \
        for (int a = 0; a < d; ++a) { L[3 - d + a] = l0[a]; W[3 - d + a] = wt[a]; }        \
        for (int t = 0; t < ntransf; ++t) {                                                \
            const REAL *g = grid + 2 * G * t;                                              \
            double re = 0.0, im = 0.0;                                                     \
            for (int qz = 0; qz < wz; ++qz) {                                              \
                int64_t z = wrap(L[0] + qz, n[0]);                                         \
                for (int qy = 0; qy < wy; ++qy) {                                          \
                    int64_t y = wrap(L[1] + qy, n[1]);                                     \
                    double wzy = W[0][qz] * W[1][qy];                                      \
                    for (int qx = 0; qx < w; ++qx) {                                       \
                        int64_t x = wrap(L[2] + qx, n[2]);                                 \
                        int64_t c = (z * n[1] + y) * n[2] + x;                             \
                        double ww = wzy * W[2][qx];                                        \
                        re += ww * (double)g[2 * c];                                       \
                        im += ww * (double)g[2 * c + 1];                                   \
                    }                                                                      \
                }                                                                          \
            }                                                                              \
            out[2 * (t * M + j)] = (REAL)re;                                               \
            out[2 * (t * M + j) + 1] = (REAL)im;                                           \
        }                                                                                  \
    }                                                                                      \
}                                                                                          \
                                                                                           \
static void spread_point_##SUFFIX(int d, const int64_t *n, int ntransf, REAL *grid,       \
                                  int64_t G, const REAL *c, int64_t M, const int64_t *nf,  \
                                  const double *const *u, int64_t j, int w, double beta)   \
{                                                                                          \
    int64_t l0[3];                                                                         \
    double wt[3][32];                                                                      \
    point_weights(d, nf, u, j, w, beta, l0, wt);                                           \
    int64_t L[3] = {0, 0, 0};                                                              \
    double one = 1.0;                                                                      \
    double *W[3] = {&one, &one, &one};                                                     \
    for (int a = 0; a < d; ++a) { L[3 - d + a] = l0[a]; W[3 - d + a] = wt[a]; }            \
    int wz = d == 3 ? w : 1, wy = d >= 2 ? w : 1;                                          \
    for (int t = 0; t < ntransf; ++t) {                                                    \
        double cre = (double)c[2 * (t * M + j)], cim = (double)c[2 * (t * M + j) + 1];     \
        REAL *g = grid + 2 * G * t;                                                        \
        for (int qz = 0; qz < wz; ++qz) {                                                  \
            int64_t z = wrap(L[0] + qz, n[0]);                                             \
            for (int qy = 0; qy < wy; ++qy) {                                              \
                int64_t y = wrap(L[1] + qy, n[1]);                                         \
                double wzy = W[0][qz] * W[1][qy];                                          \
                for (int qx = 0; qx < w; ++qx) {                                           \
                    int64_t x = wrap(L[2] + qx, n[2]);                                     \
                    int64_t cc = (z * n[1] + y) * n[2] + x;                                \
                    double ww = wzy * W[2][qx];                                            \
                    g[2 * cc] += (REAL)(ww * cre);                                         \
                    g[2 * cc + 1] += (REAL)(ww * cim);                                     \
                }                                                                          \
            }                                                                              \
        }                                                                                  \
    }                                                                                      \
}                                                                                          \
                                                                                           \
void kb_spread_##SUFFIX(int d, const int64_t *nf, int ntransf, const REAL *c, int64_t M,  \
                        const double *u0, const double *u1, const double *u2, int w,       \
                        double beta, REAL *grid)                                           \
{                                                                                          \
    const double *u[3] = {u0, u1, u2};                                                     \
    int64_t n[3] = {1, 1, 1};                                                              \
    for (int a = 0; a < d; ++a) n[3 - d + a] = nf[a];                                      \
    int64_t G = n[0] * n[1] * n[2];                                                        \
    memset(grid, 0, sizeof(REAL) * 2 * G * ntransf);                                       \
    /* slab colouring along axis 0 of the grid (the slowest axis) */                       \
    int64_t n0 = nf[0];                                                                    \
    int64_t H = w;                                                                         \
    int64_t nslab = n0 / H;                                                                \
    if (nslab < 3) { /* too small to colour: serial */                                     \
        for (int64_t j = 0; j < M; ++j)                                                    \
            spread_point_##SUFFIX(d, n, ntransf, grid, G, c, M, nf, u, j, w, beta);        \
        return;                                                                            \
    }                                                                                      \
    int64_t *cnt = (int64_t *)calloc((size_t)nslab + 1, sizeof(int64_t));                  \
    int64_t *slab = (int64_t *)malloc(sizeof(int64_t) * (size_t)M);                        \
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)M);                        \
    for (int64_t j = 0; j < M; ++j) {                                                      \
        int64_t s = wrap((int64_t)ceil(u[0][j] - 0.5 * w), n0) / H;                        \
        if (s >= nslab) s = nslab - 1; /* remainder rows join the last slab */             \
        slab[j] = s;                                                                       \
        cnt[s + 1]++;                                                                      \
    }                                                                                      \
    for (int64_t s = 0; s < nslab; ++s) cnt[s + 1] += cnt[s];                              \
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)nslab);                     \
    memcpy(pos, cnt, sizeof(int64_t) * (size_t)nslab);                                     \
    for (int64_t j = 0; j < M; ++j) perm[pos[slab[j]]++] = j;                              \
    for (int colour = 0; colour < 3; ++colour) {                                           \
        _Pragma("omp parallel for schedule(dynamic, 1)")                                   \
        for (int64_t s = 0; s < nslab; ++s) {                                              \
            int odd_last = (nslab % 2 == 1) && (s == nslab - 1);                           \
            int mine = odd_last ? (colour == 2) : ((int)(s % 2) == colour);                \
            if (!mine) continue;                                                           \
            for (int64_t k = cnt[s]; k < cnt[s + 1]; ++k)                                  \
                spread_point_##SUFFIX(d, n, ntransf, grid, G, c, M, nf, u, perm[k], w,     \
                                      beta);                                               \
        }                                                                                  \
    }                                                                                      \
    free(pos); free(perm); free(slab); free(cnt);                                          \
}

DEFINE_KB(double, f64)
DEFINE_KB(float, f32)

/* Kaiser-Bessel window at one offset (exported for the tests' closed-form checks). */
double kb_eval(double t, int w, double beta) { return kb_window(t, w, beta); }
