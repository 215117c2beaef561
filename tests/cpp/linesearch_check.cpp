// CPU check of the Newton linesearches (iceicle_amd/csrc/linesearch.h) with
// closed-form merit functions, including a step that lands on a non-finite
// residual (petsc_newton.hpp:258-293 maps it to 1e9 before the linesearch
// sees it).  Prints "OK" and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "../../iceicle_amd/csrc/linesearch.h"

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

static ice_newton_opts defaults() {
    // the documented defaults of ice_newton_opts_default (wolfe_ls /
    // corrigan_ls parameters, nonlinear_solver_utils.hpp:130-191)
    ice_newton_opts o;
    std::memset(&o, 0, sizeof(o));
    o.ls_kmax = 20;
    o.ls_alpha_initial = 1.0;
    o.ls_alpha_max = 10.0;
    o.ls_c1 = 1e-4;
    o.ls_c2 = 0.9;
    o.ls_alpha_min = 0.0;
    return o;
}

int main() {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const double inf = std::numeric_limits<double>::infinity();
    CHECK(ice::ls_merit(nan) == 1e9);
    CHECK(ice::ls_merit(inf) == 1e9);
    CHECK(ice::ls_merit(-inf) == 1e9);
    CHECK(ice::ls_merit(2.5) == 2.5);

    const ice_newton_opts o = defaults();
    // merit of a residual norm that is finite and decreasing up to alpha = 0.6
    // with its minimum at 0.5, and not a number past 0.9 (a step into
    // negative pressure, say)
    int calls = 0;
    auto raw = [&](double a) -> double {
        ++calls;
        if (a >= 0.9) return nan;
        return 1.0 + (a - 0.5) * (a - 0.5);
    };
    ice::Phi phi = [&](double a) { return ice::ls_merit(raw(a)); };
    const double phi0 = phi(0.0);

    const double aw = ice::wolfe_linesearch(phi, o);
    std::printf("wolfe: alpha = %.6f phi = %.6f (phi0 %.6f)\n", aw, phi(aw), phi0);
    CHECK(std::isfinite(aw) && aw > 0.0 && aw < 0.9);
    CHECK(phi(aw) < phi0);

    const double ac = ice::corrigan_linesearch(phi, o);
    std::printf("corrigan: alpha = %.6f phi = %.6f\n", ac, phi(ac));
    // bisection: alpha = 1 (NaN -> 1e9) is rejected, every later trial
    // below 0.9 is accepted and raises the lower end, so the search closes
    // in on 0.9 from below and returns the last (finite) trial
    CHECK(ac > 0.85 && ac < 0.9);
    CHECK(phi(ac) < phi0);

    // an accepted full step returns at once on both searches
    ice::Phi bowl = [](double a) { return (a - 1.0) * (a - 1.0) + 0.1; };
    CHECK(ice::corrigan_linesearch(bowl, o) == 1.0);
    const double aw2 = ice::wolfe_linesearch(bowl, o);
    CHECK(std::abs(aw2 - 1.0) < 1e-6);

    // the zoom's cubic step lands inside the bracket
    const double az = ice::ls_zoom(phi, 0.0, 1.0, o.ls_kmax, o.ls_c1, o.ls_c2);
    CHECK(az > 0.0 && az < 1.0);
    if (failures == 0) std::printf("OK\n");
    return failures == 0 ? 0 : 1;
}
