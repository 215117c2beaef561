/* Analytic fields used as Dirichlet/Neumann values, sources and initial
 * states.  The reference passes these as std::function callbacks taking a
 * physical point and writing neq values (conservation_law.hpp:146-170); here
 * they are closed-form functions identified by ICE_FIELD_* kinds so that the
 * same definition is evaluated on the host and inside device kernels. */
#pragma once
#include <math.h>

#ifndef ICE_HD
#if defined(__HIPCC__)
#define ICE_HD __host__ __device__
#else
#define ICE_HD
#endif
#endif

#include "../../include/iceicle_amd.h"
#include "ns_mms_generated.h"

/* Isentropic vortex (examples/conservation_law/isentropic_vortex.lua:37-57
 * form, extruded in z): p = (x0, y0, beta, rho_inf(unused, 1), u_inf, v_inf),
 * gamma = 1.4, unit reference parameters. */
ICE_HD inline void ice_isentropic_vortex(const double* p, int ndim, const double* x, double* out) {
    const double gamma = 1.4;
    const double x0 = p[0], y0 = p[1], beta = p[2], uinf = p[4], vinf = p[5];
    const double dx = x[0] - x0, dy = x[1] - y0;
    const double r2 = dx * dx + dy * dy;
    const double pi = 3.14159265358979323846;
    const double f = exp(0.5 * (1.0 - r2));
    const double u = uinf - beta / (2.0 * pi) * f * dy;
    const double v = vinf + beta / (2.0 * pi) * f * dx;
    /* T / T_inf with c_inf^2 = gamma p_inf / rho_inf = 1 (the deck's T_hat) */
    const double T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * pi * pi) * exp(1.0 - r2);
    const double rho = pow(T, 1.0 / (gamma - 1.0));
    const double pr = pow(rho, gamma) / gamma;  /* p = rho^gamma / gamma  (c_inf = 1) */
    out[0] = rho;
    out[1] = rho * u;
    out[2] = rho * v;
    double ke = 0.5 * rho * (u * u + v * v);
    if (ndim == 3) {
        out[3] = 0.0;
        out[4] = pr / (gamma - 1.0) + ke;
    } else {
        out[3] = pr / (gamma - 1.0) + ke;
    }
}

/* Evaluate field `f` at physical point x (ndim coordinates), writing neq values. */
ICE_HD inline void ice_field_eval(const ice_field* f, int ndim, int neq, const double* x, double* out) {
    const double* p = f->p;
    switch (f->kind) {
    case ICE_FIELD_CONSTANT:
        for (int v = 0; v < neq; ++v) out[v] = p[v];
        break;
    case ICE_FIELD_SINE:
        out[0] = p[0] * sin(p[1] * x[0] + p[2]);
        for (int v = 1; v < neq; ++v) out[v] = 0.0;
        break;
    case ICE_FIELD_SINUSOIDS: {
        const double twopi = 6.28318530717958647692;
        int K = (int)p[0];
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            const double A = p[1 + 4 * k], m = p[2 + 4 * k], n = p[3 + 4 * k], l = p[4 + 4 * k];
            double arg = m * x[0];
            if (ndim > 1) arg += n * x[1];
            if (ndim > 2) arg += l * x[2];
            s += A * sin(twopi * arg);
        }
        out[0] = s;
        for (int v = 1; v < neq; ++v) out[v] = 0.0;
        break;
    }
    case ICE_FIELD_NS_MMS_STATE:
        ice_ns_mms_state(x, out);
        break;
    case ICE_FIELD_NS_MMS_SOURCE:
        ice_ns_mms_source(x, out);
        break;
    case ICE_FIELD_ISENTROPIC_VORTEX:
        ice_isentropic_vortex(p, ndim, x, out);
        break;
    case ICE_FIELD_EXP_DECAY_SINE:
        out[0] = exp(-p[0] * p[1]) * sin(x[0]);
        for (int v = 1; v < neq; ++v) out[v] = 0.0;
        break;
    default:
        for (int v = 0; v < neq; ++v) out[v] = 0.0;
        break;
    }
}
