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

inline int kb_width_for_tol(double tol) {
    int w = int(std::ceil(-std::log10(tol / 10.0)));
    return w < 2 ? 2 : (w > 16 ? 16 : w);
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
