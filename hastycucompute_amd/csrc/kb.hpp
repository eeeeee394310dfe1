// Kaiser-Bessel window: device evaluation and host deapodisation factors.
//
//   phi(t) = I0(beta sqrt(1 - (2t/w)^2)),  |t| <= w/2     (t in fine-grid cells)
//   Phi(a) = w sinh(sqrt(beta^2 - (a w/2)^2)) / sqrt(beta^2 - (a w/2)^2)
//
// I0 is summed from its power series sum_k (x^2/4)^k / (k!)^2
// (Abramowitz & Stegun 9.6.12); beta follows Beatty et al. (IEEE TMI 2005):
//   beta = pi sqrt(w^2/sigma^2 (sigma - 1/2)^2 - 0.8).
#pragma once

#include <cmath>

#include "common.hpp"

namespace hasty {

// I0 and the window on the host (same series as the device kb_i0 below).
inline double kb_i0_host(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 200; ++k) {
        term *= q / double(k * k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

inline double kb_phi_host(double t, double inv_half_w, double beta) {
    double z = t * inv_half_w;
    z = 1.0 - z * z;
    return z > 0.0 ? kb_i0_host(beta * std::sqrt(z)) : (z == 0.0 ? 1.0 : 0.0);
}

inline double kb_beta(int w, double sigma) {
    return M_PI * std::sqrt(double(w) * w / (sigma * sigma) * (sigma - 0.5) * (sigma - 0.5) - 0.8);
}

inline double kb_hat_host(double alpha, int w, double beta) {
    const double a = 0.5 * w;
    const double s = beta * beta - (a * alpha) * (a * alpha);
    const double r = std::sqrt(std::fabs(s));
    if (r == 0.0) return 2.0 * a;
    return s > 0 ? 2.0 * a * std::sinh(r) / r : 2.0 * a * std::sin(r) / r;
}

// Estimated relative L2 error of gridding with a width-w window at
// oversampling sigma in dim dimensions.  The window is compactly supported,
// so the only error is aliasing: by Poisson summation a mode at angular
// frequency alpha (|alpha| <= pi / sigma on the fine grid) picks up the
// replicas Phi(alpha + 2 pi p), p != 0, relative to Phi(alpha):
//   e(alpha)^2 = sum_{p != 0} Phi(alpha + 2 pi p)^2 / Phi(alpha)^2.
// The estimate is the RMS of e over alpha (64 midpoints of [0, pi / sigma],
// |p| <= 400), times sqrt(dim) for the independent per-axis errors.
inline double kb_error_estimate(int w, double sigma, int dim) {
    const double beta = kb_beta(w, sigma);
    constexpr int kAlpha = 64, kReplicas = 400;
    double acc = 0.0;
    for (int i = 0; i < kAlpha; ++i) {
        const double al = M_PI / sigma * (i + 0.5) / kAlpha;
        const double h0 = kb_hat_host(al, w, beta);
        double s = 0.0;
        for (int p = 1; p <= kReplicas; ++p) {
            const double hp = kb_hat_host(al + 2.0 * M_PI * p, w, beta);
            const double hm = kb_hat_host(al - 2.0 * M_PI * p, w, beta);
            s += hp * hp + hm * hm;
        }
        acc += s / (h0 * h0);
    }
    return std::sqrt(dim * acc / kAlpha);
}

// How much the complex-arithmetic rounding of the fine-grid FFT grows in
// the gridding steps, per axis: white rounding noise on the fine grid is
// scaled by the RMS deapodisation factor sqrt(mean_alpha Phi(alpha)^-2)
// (alpha in [0, pi / sigma]) and summed by the window, sqrt(int phi^2).
// At sigma = 2 this is about 0.8-1.1; at sigma = 1.25 it is 2.6 (w = 6) to
// 7.4 (w = 9), so complex64 plans there have a rounding floor of ~1e-5.
inline double kb_rounding_gain(int w, double sigma) {
    const double beta = kb_beta(w, sigma);
    constexpr int kAlpha = 64, kT = 4096;
    double inv2 = 0.0;
    for (int i = 0; i < kAlpha; ++i) {
        const double h = kb_hat_host(M_PI / sigma * (i + 0.5) / kAlpha, w, beta);
        inv2 += 1.0 / (h * h);
    }
    double phi2 = 0.0;   // midpoint rule over [-w/2, w/2]
    for (int i = 0; i < kT; ++i) {
        const double t = -0.5 * w + w * (i + 0.5) / kT;
        const double v = kb_phi_host(t, 2.0 / w, beta);
        phi2 += v * v;
    }
    phi2 *= double(w) / kT;
    return std::sqrt(inv2 / kAlpha) * std::sqrt(phi2);
}

// Window width for a tolerance at oversampling sigma, for plans computing in
// FP32 (f64 = false) or FP64.  sigma = 2 keeps FINUFFT's published rule
// w = ceil(-log10(tol / 10)) (Barnett, Magland and af Klinteberg, SIAM J.
// Sci. Comput. 2019), whose error the aliasing estimate puts at 0.8-1.6 tol
// for this window.  Any other sigma takes the smallest width whose aliasing
// estimate plus rounding floor, 4 eps gain^dim (eps = 2^-24 or 2^-53; the
// factor 4 covers type 1, whose floor measures ~2.7x type 2's), is <= tol:
// sigma = 1.25 gives w = 9 at 1e-5 in FP64 and w = 6 at 1e-3; in FP32 its
// floor is ~6e-5 (w = 8), so 1e-5 is out of reach.  Returns 17 when no width
// up to 16 reaches tol.
inline int kb_width_for(double tol, double sigma, int dim, bool f64) {
    if (sigma == 2.0) {
        const int w = int(std::ceil(-std::log10(tol / 10.0)));
        return w < 2 ? 2 : (w > 16 ? 17 : w);
    }
    const double eps = f64 ? 0x1p-53 : 0x1p-24;
    for (int w = 2; w <= 16; ++w) {
        const double err = kb_error_estimate(w, sigma, dim) + 4.0 * eps * std::pow(kb_rounding_gain(w, sigma), dim);
        if (err <= tol) return w;
    }
    return 17;
}

// I0 by its power series.  Terms peak near k ~ x/2 and are summed until they
// fall below 2^-(mantissa bits) of the partial sum.
template <class R> __device__ inline R kb_i0(R x) {
    const R q = R(0.25) * x * x;
    R term = R(1), sum = R(1);
    const R stop = sizeof(R) == 4 ? R(6e-8) : R(1e-16);
    for (int k = 1; k < 80; ++k) {
        term *= q / R(k * k);
        sum += term;
        if (term < stop * sum) break;
    }
    return sum;
}

template <class R> __device__ inline R kb_phi(R t, R inv_half_w, R beta) {
    R z = t * inv_half_w;
    z = R(1) - z * z;
    return z > R(0) ? kb_i0(beta * sqrt(z)) : (z == R(0) ? R(1) : R(0));
}

// d phi / dt inside the support.  With c = 2/w and s = sqrt(1 - (c t)^2),
// d/dt I0(beta s) = I1(beta s) beta ds/dt = -beta^2 c^2 t [I1(z) / z]_{z = beta s},
// and I1(z)/z = 1/2 sum_k (z^2/4)^k / (k! (k+1)!)  (A&S 9.6.10), finite at s = 0.
template <class R> __device__ inline R kb_dphi(R t, R inv_half_w, R beta) {
    const R ct = t * inv_half_w;
    const R z2 = R(1) - ct * ct;
    if (!(z2 >= R(0))) return R(0);
    const R q = R(0.25) * beta * beta * z2;
    R term = R(0.5), sum = R(0.5);
    const R stop = sizeof(R) == 4 ? R(6e-8) : R(1e-16);
    for (int k = 1; k < 80; ++k) {
        term *= q / R(k * (k + 1));
        sum += term;
        if (term < stop * sum) break;
    }
    return -beta * beta * inv_half_w * inv_half_w * t * sum;
}

}  // namespace hasty
