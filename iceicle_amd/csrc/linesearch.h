// Linesearches of the Newton solve (nonlinear_solver_utils.hpp:74-199),
// host-only: they see the merit function phi(alpha) = ||R(u - alpha du)||_2
// through a callable, so they are tested on the CPU with closed-form phi.
// The control flow decides the Newton iterates, so each routine follows the
// reference's steps in order.
#pragma once
#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>

#include "../../include/iceicle_amd.h"

namespace ice {

using Phi = std::function<double(double)>;

// A non-finite merit value enters the linesearch as 1e9 (BIG_RESIDUAL,
// petsc_newton.hpp:258-282), so a step into NaN or Inf is rejected like a
// step that raises the residual.
inline double ls_merit(double v) { return std::isfinite(v) ? v : 1e9; }

// finite_difference, nonlinear_solver_utils.hpp:74-79
inline double fd_slope(const Phi& f, double eps, double x) {
    const double p1 = f(x);
    const double p2 = f(x + eps);
    return (p2 - p1) / eps;
}

// zoom, nonlinear_solver_utils.hpp:82-124 (cubic interpolation between
// alpha_lo and alpha_hi; the midpoint if it does not finish)
inline double ls_zoom(const Phi& f, double lo, double hi, int kmax, double c1, double c2) {
    const double eps = std::sqrt(std::numeric_limits<double>::epsilon());
    for (int k = 0; k < kmax; ++k) {
        const double phi_lo = f(lo), dphi_lo = fd_slope(f, eps, lo);
        const double phi_hi = f(hi), dphi_hi = fd_slope(f, eps, hi);
        const double d1 = dphi_lo + dphi_hi - 3 * (phi_lo - phi_hi) / (lo - hi);
        double d2 = std::sqrt(std::max(d1 * d1 - dphi_lo * dphi_hi, 1e-8));
        d2 = std::copysign(d2, hi - lo);
        const double aj = hi - (hi - lo) * ((dphi_hi + d2 - d1) / (dphi_hi - dphi_lo + 2 * d2));
        const double phi_aj = f(aj);
        const double dphi_0 = fd_slope(f, eps, 0.0);
        if (phi_aj > f(0.0) + c1 * aj * dphi_0 || phi_aj >= phi_lo) {
            hi = aj;
        } else {
            const double dphi_aj = fd_slope(f, eps, aj);
            if (std::abs(dphi_aj) <= -c2 * dphi_0) return aj;
            if (dphi_aj * (hi - lo) >= 0) hi = lo;
            lo = aj;
        }
    }
    return 0.5 * (lo + hi);
}

// wolfe_ls, nonlinear_solver_utils.hpp:139-170
inline double wolfe_linesearch(const Phi& f, const ice_newton_opts& o) {
    const double eps = std::sqrt(std::numeric_limits<double>::epsilon());
    double a_im1 = 0.0, a_i = o.ls_alpha_initial;
    const double phi_0 = f(0.0);
    const double dphi_0 = fd_slope(f, eps, 0.0);
    for (int k = 0; k < o.ls_kmax; ++k) {
        const double phi_i = f(a_i);
        const double phi_im1 = f(a_im1);
        if (phi_i > phi_0 + o.ls_c1 * a_i * dphi_0 || (phi_i >= phi_im1 && k > 0))
            return ls_zoom(f, a_im1, a_i, o.ls_kmax, o.ls_c1, o.ls_c2);
        const double dphi_i = fd_slope(f, eps, a_i);
        if (std::abs(dphi_i) <= -o.ls_c2 * dphi_0) return a_i;
        if (dphi_i >= 0) return ls_zoom(f, a_i, a_im1, o.ls_kmax, o.ls_c1, o.ls_c2);
        const double next = 0.5 * (a_i + o.ls_alpha_max);
        a_im1 = a_i;
        a_i = next;
    }
    return a_i;
}

// corrigan_ls, nonlinear_solver_utils.hpp:172-191: bisection between
// alpha_min and alpha_max; returns the last alpha tried
inline double corrigan_linesearch(const Phi& f, const ice_newton_opts& o) {
    double a = o.ls_alpha_min, b = o.ls_alpha_max;
    double alpha = o.ls_alpha_initial;
    for (int k = 0; k < o.ls_kmax; ++k) {
        const double phi_0 = f(0.0);
        if (k > 0) alpha = 0.5 * (a + b);
        const double phi_star = f(alpha);
        if (std::abs(phi_star) < std::abs(phi_0)) {
            if (alpha == o.ls_alpha_initial) break;
            a = alpha;
        } else {
            b = alpha;
        }
    }
    return alpha;
}

}  // namespace ice
