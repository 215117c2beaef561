// Pointwise flux functors for the device kernels.
//
// Each family implements the reference's flux-functor concepts
// (conservation_law.hpp:17-90) as static device functions over plain arrays:
//   flux(u, gradu[M][D]) -> F[M][D]           physical_flux
//   conv(uL, uR, n)      -> [M]               convective_numerical_flux
//   diff(u, gradu, n)    -> [M]               diffusion_flux (+ neumann)
//   apply_bc(...)                               implements_bcs (NS only)
//   homogeneity(u) -> G[M][D][M][D]             computes_homogeneity_tensor (Burgers)
// Formulas follow disc/burgers.hpp:41-255 and disc/navier_stokes.hpp:321-1301.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/iceicle_amd.h"

namespace ice {

// Parameters of every flux family, flattened for a kernel argument, in the
// arithmetic type R of the kernels that use them.
template <class R>
struct PhysT {
    // Burgers (BurgersCoefficients, burgers.hpp:14-40)
    R mu;
    R a[3];
    R b[3];
    // Navier-Stokes (Physics + CaloricallyPerfectEoS + Nondimensionalization)
    R gamma;
    R Pr;
    R Re, Eu, e_coeff;
    R T_coeff;           // cp * ref.T * ref.rho / ref.p
    int visc_kind;
    R visc_mu;
    R mu_ratio, T0r, Tsr;  // DimensionlessSutherlands ratios
    R isoT[ICE_MAX_BCFLAG];
    R fs_rho, fs_u, fs_dir[3], fs_T;
    int textbook_energy;
    // ConservationLawDDG
    R sigma_ic;
    R sigma_face, sigma_bdy;  // DDGIC factors per interior side and on Dirichlet faces
    int ic_mirror;            // DDGIC test gradients averaged with the mirror point's
    R beta0, beta1, beta0_bdy;
    // parameter-only factors of the flux formulas, computed once on the host
    // (set_derived) so the pointwise fluxes multiply instead of divide
    R p_fac;        // (gamma - 1) / (e_coeff Eu): p = p_fac rho e
    R T_fac;        // gamma / ((gamma - 1) T_coeff): T = T_fac p / rho
    R inv_e_coeff;  // 1 / e_coeff
    R g_over_Pr;    // gamma / Pr
    R inv_Re;       // 1 / Re
    R inv_gamma;    // 1 / gamma
    R vl_fac;       // 1 / (2 (gamma^2 - 1)), VanLeer energy split
};
using PhysParams = PhysT<double>;

// fill the derived factors from the primary parameters
inline void set_derived(PhysParams& p) {
    p.p_fac = (p.gamma - 1) / (p.e_coeff * p.Eu);
    p.T_fac = p.gamma / ((p.gamma - 1) * p.T_coeff);
    p.inv_e_coeff = 1.0 / p.e_coeff;
    p.g_over_Pr = p.gamma / p.Pr;
    p.inv_Re = 1.0 / p.Re;
    p.inv_gamma = 1.0 / p.gamma;
    p.vl_fac = 1.0 / (2 * (p.gamma * p.gamma - 1));
}

// the same parameters rounded to another arithmetic type
template <class R>
inline PhysT<R> phys_cast(const PhysParams& p) {
    PhysT<R> q;
    q.mu = (R)p.mu;
    for (int k = 0; k < 3; ++k) {
        q.a[k] = (R)p.a[k];
        q.b[k] = (R)p.b[k];
        q.fs_dir[k] = (R)p.fs_dir[k];
    }
    q.gamma = (R)p.gamma;
    q.Pr = (R)p.Pr;
    q.Re = (R)p.Re;
    q.Eu = (R)p.Eu;
    q.e_coeff = (R)p.e_coeff;
    q.T_coeff = (R)p.T_coeff;
    q.visc_kind = p.visc_kind;
    q.visc_mu = (R)p.visc_mu;
    q.mu_ratio = (R)p.mu_ratio;
    q.T0r = (R)p.T0r;
    q.Tsr = (R)p.Tsr;
    for (int k = 0; k < ICE_MAX_BCFLAG; ++k) q.isoT[k] = (R)p.isoT[k];
    q.fs_rho = (R)p.fs_rho;
    q.fs_u = (R)p.fs_u;
    q.fs_T = (R)p.fs_T;
    q.textbook_energy = p.textbook_energy;
    q.sigma_ic = (R)p.sigma_ic;
    q.sigma_face = (R)p.sigma_face;
    q.sigma_bdy = (R)p.sigma_bdy;
    q.ic_mirror = p.ic_mirror;
    q.beta0 = (R)p.beta0;
    q.beta1 = (R)p.beta1;
    q.beta0_bdy = (R)p.beta0_bdy;
    q.p_fac = (R)p.p_fac;
    q.T_fac = (R)p.T_fac;
    q.inv_e_coeff = (R)p.inv_e_coeff;
    q.g_over_Pr = (R)p.g_over_Pr;
    q.inv_Re = (R)p.inv_Re;
    q.inv_gamma = (R)p.inv_gamma;
    q.vl_fac = (R)p.vl_fac;
    return q;
}

template <int D, class R = double>
struct Burgers {
    using real = R;
    using Params = PhysT<R>;
    static constexpr int M = 1;
    static constexpr bool IMPLEMENTS_BCS = false;
    static constexpr bool HAS_HOMOGENEITY = true;

    __device__ __forceinline__ static void flux(const Params& p, const R* u, const R* g, R* F) {
#pragma unroll
        for (int d = 0; d < D; ++d) F[d] = (p.a[d] + R(0.5) * p.b[d] * u[0]) * u[0] - p.mu * g[d];
    }
    // |a + b u / 2|, the wave speed the reference tracks in lambda_max (burgers.hpp:79-90)
    __device__ __forceinline__ static R wave_speed(const Params& p, const R* u) {
        R s = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const R l = p.a[d] + R(0.5) * p.b[d] * u[0];
            s += l * l;
        }
        return sqrt(s);
    }
    __device__ __forceinline__ static R visc_at(const Params& p, const R*) { return p.mu; }
    __device__ __forceinline__ static void conv(const Params& p, const R* uL, const R* uR,
                                                const R* n, R* out) {
        const R ua = R(0.5) * (uL[0] + uR[0]);
        R P = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) P += n[d] * (p.a[d] + R(0.5) * p.b[d] * ua);
        const R uu = (P > 0) ? uL[0] : uR[0];
        R f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) f += n[d] * (p.a[d] + R(0.5) * p.b[d] * uu);
        out[0] = f * uu;
    }
    __device__ __forceinline__ static void diff(const Params& p, const R*, const R* g, const R* n,
                                                R* out) {
        R f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) f += p.mu * g[d] * n[d];
        out[0] = f;
    }
    __device__ __forceinline__ static void neumann(const Params& p, const R* g, R* out) { out[0] = p.mu * g[0]; }
    // G[0][k][0][s] = mu delta_ks
    __device__ __forceinline__ static void homogeneity(const Params& p, const R*, R* G) {
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
            for (int s = 0; s < D; ++s) G[k * D + s] = (k == s) ? p.mu : R(0.0);
    }
    __device__ __forceinline__ static void apply_bc(const Params&, const R* uL, const R*, const R*, int, int,
                                                    R* uR) { uR[0] = uL[0]; }
};

// Compressible Navier-Stokes / Euler in conservative variables.
template <int D, bool FULL, class R = double>
struct NavierStokes {
    using real = R;
    using Params = PhysT<R>;
    static constexpr int M = D + 2;
    static constexpr bool IMPLEMENTS_BCS = true;
    static constexpr bool HAS_HOMOGENEITY = false;

    struct State {
        R rho, irho, rhoE, E, e, p, T, c, vv;
        R m[D], v[D];
    };

    // calc_thermo_state<CONSERVATIVE> (navier_stokes.hpp:405-449)
    __device__ __forceinline__ static State state(const Params& p, const R* u) {
        State s;
        s.rho = fmax(R(1e-14), u[0]);
        s.irho = R(1.0) / s.rho;
        s.vv = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            s.m[d] = u[1 + d];
            s.v[d] = u[1 + d] * s.irho;
            s.vv += s.v[d] * s.v[d];
        }
        s.rhoE = u[D + 1];
        s.E = s.rhoE * s.irho;
        s.e = s.E - R(0.5) * p.e_coeff * s.vv;
        s.p = fmax(R(1e-12), p.p_fac * s.rho * s.e);
        s.T = s.p * s.irho * p.T_fac;
        s.c = sqrt((p.gamma * p.Eu) * s.p * s.irho);
        return s;
    }
    __device__ __forceinline__ static R viscosity(const Params& p, R T) {
        if (p.visc_kind == ICE_VISC_CONSTANT) return p.visc_mu;
        return p.mu_ratio * pow(T / p.T0r, R(1.5)) * (p.T0r + p.Tsr) / (T + p.Tsr);
    }
    __device__ __forceinline__ static R wave_speed(const Params& p, const R* u) {
        const State s = state(p, u);
        return s.c + sqrt(s.vv);
    }
    __device__ __forceinline__ static R visc_at(const Params& p, const R* u) {
        return FULL ? viscosity(p, state(p, u).T) : R(0.0);
    }

    // velocity and total-energy gradients (calc_state_gradients<CONSERVATIVE>, :541-575),
    // shear stress (:700-723) and heat flux (:730-746)
    __device__ __forceinline__ static void viscous(const Params& p, const State& s, const R* g, R mu,
                                                   R (&tau)[D][D], R (&gE)[D], R (&q)[D]) {
        R gv[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) gv[i][j] = (g[(1 + i) * D + j] - s.v[i] * g[j]) * s.irho;
#pragma unroll
        for (int j = 0; j < D; ++j) gE[j] = (g[(D + 1) * D + j] - s.E * g[j]) * s.irho;
        R div = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) div += gv[k][k];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) tau[i][j] = mu * (gv[i][j] + gv[j][i]);
#pragma unroll
        for (int i = 0; i < D; ++i) tau[i][i] -= R(2.0) / R(3.0) * mu * div;
        const R iec = p.inv_e_coeff, kq = mu * p.g_over_Pr;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            R qj = gE[j] * iec;
#pragma unroll
            for (int k = 0; k < D; ++k) qj -= s.v[k] * gv[k][j];
            q[j] = qj * kq;
        }
    }

    // Flux::operator() (:915-973)
    __device__ __forceinline__ static void flux(const Params& p, const R* u, const R* g, R* F) {
        const State s = state(p, u);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            F[j] = s.m[j];
#pragma unroll
            for (int i = 0; i < D; ++i) F[(1 + i) * D + j] = s.m[i] * s.v[j];
            F[(1 + j) * D + j] += p.Eu * s.p;
            F[(D + 1) * D + j] = s.v[j] * (s.rhoE + p.Eu * p.e_coeff * s.p);
        }
        if constexpr (FULL) {
            R tau[D][D], gE[D], q[D];
            viscous(p, s, g, viscosity(p, s.T), tau, gE, q);
            const R iRe = p.inv_Re, ecRe = p.e_coeff * iRe;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                R ef = q[j];
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    F[(1 + i) * D + j] -= tau[i][j] * iRe;
                    ef += s.v[i] * tau[i][j];
                }
                F[(D + 1) * D + j] -= ef * ecRe;
            }
        }
    }

    // VanLeer (:777-864)
    __device__ __forceinline__ static void conv(const Params& p, const R* uL, const R* uR, const R* n,
                                                R* f) {
        const State SL = state(p, uL), SR = state(p, uR);
        const R Eu = p.Eu, ec = p.e_coeff, g = p.gamma;
        R vnL = 0, vnR = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) { vnL += SL.v[d] * n[d]; vnR += SR.v[d] * n[d]; }
        const R mL = vnL / SL.c, mR = vnR / SR.c;
        if (mL > 1) {
            f[0] = SL.rho * vnL;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] = SL.m[d] * vnL + Eu * SL.p * n[d];
            f[D + 1] = vnL * (SL.rhoE + Eu * ec * SL.p);
        } else if (mL < -1) {
#pragma unroll
            for (int k = 0; k < M; ++k) f[k] = 0;
        } else {
            const R fm = SL.rho * SL.c * (mL + 1) * (mL + 1) / R(4.0);
            f[0] = fm;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] = fm * (SL.v[d] + n[d] * (-vnL + 2 * SL.c) * p.inv_gamma);
            const R t = (g - 1) * vnL + 2 * SL.c;
            f[D + 1] = fm * ((SL.vv - vnL * vnL) / 2 + t * t * p.vl_fac);
        }
        if (mR < -1) {
            f[0] += SR.rho * vnR;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] += SR.m[d] * vnR + Eu * SR.p * n[d];
            f[D + 1] += vnR * (SR.rhoE + Eu * ec * SR.p);
        } else if (mR <= 1) {
            const R fm = -SR.rho * SR.c * (mR - 1) * (mR - 1) / R(4.0);
            f[0] += fm;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] += fm * (SR.v[d] + n[d] * (-vnR - 2 * SR.c) * p.inv_gamma);
            const R t = (g - 1) * vnR - 2 * SR.c;
            f[D + 1] += fm * ((SR.vv - vnR * vnR) / 2 + t * t * p.vl_fac);
        }
    }

    // DiffusionFlux (:1241-1289).  Energy: mu gamma / Pr * dE/dx . n as in
    // :1277-1279, or the volume heat flux q . n when textbook_energy is set.
    __device__ __forceinline__ static void diff(const Params& p, const R* u, const R* g, const R* n,
                                                R* f) {
#pragma unroll
        for (int k = 0; k < M; ++k) f[k] = 0;
        if constexpr (FULL) {
            const State s = state(p, u);
            const R mu = viscosity(p, s.T);
            R tau[D][D], gE[D], q[D];
            viscous(p, s, g, mu, tau, gE, q);
            const R iRe = p.inv_Re, ecRe = p.e_coeff * iRe, kE = mu * p.g_over_Pr;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                R ef = p.textbook_energy ? q[j] * n[j] : kE * gE[j] * n[j];
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    f[1 + i] += tau[i][j] * iRe * n[j];
                    ef += s.v[i] * tau[i][j] * n[j];
                }
                f[D + 1] += ef * ecRe;
            }
        }
    }
    __device__ __forceinline__ static void neumann(const Params&, const R*, R* out) {
#pragma unroll
        for (int k = 0; k < M; ++k) out[k] = 0;
    }
    __device__ __forceinline__ static void homogeneity(const Params&, const R*, R*) {}

    // conservative state from (rho, velocity, T) (calc_thermo_state<RHO_U_T>, :450-488)
    __device__ __forceinline__ static void cons_from_rho_u_T(const Params& p, const R* prim, R* out) {
        const R rho = fmax(R(1e-14), prim[0]);
        R vv = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) vv += prim[1 + d] * prim[1 + d];
        const R pr = fmax(R(1e-12), rho * (p.gamma - 1) / p.gamma * p.T_coeff * prim[D + 1]);
        const R e = pr * p.e_coeff * p.Eu / (p.gamma - 1) / rho;
        const R E = e + R(0.5) * p.e_coeff * vv;
        out[0] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) out[1 + d] = prim[1 + d] * rho;
        out[D + 1] = rho * E;
    }
    // conservative state from (rho, velocity, p) (calc_thermo_state<RHO_U_P>, :489-525)
    __device__ __forceinline__ static void cons_from_rho_u_p(const Params& p, const R* prim, R* out) {
        const R rho = fmax(R(1e-14), prim[0]);
        const R pr = fmax(R(1e-12), prim[D + 1]);
        R vv = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) vv += prim[1 + d] * prim[1 + d];
        const R e = pr * p.e_coeff * p.Eu / (p.gamma - 1) / rho;
        const R E = e + R(0.5) * p.e_coeff * vv;
        out[0] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) out[1 + d] = prim[1 + d] * rho;
        out[D + 1] = rho * E;
    }

    // Flux::apply_bc (:975-1189); exterior gradient = interior gradient
    __device__ __forceinline__ static void apply_bc(const Params& p, const R* uL, const R*, const R* n,
                                                    int bctype, int bcflag, R* uR) {
        if (bctype == ICE_BC_WALL_GENERAL) bctype = FULL ? ICE_BC_NO_SLIP_ISOTHERMAL : ICE_BC_SLIP_WALL;
        const State sL = state(p, uL);
        if (bctype == ICE_BC_SLIP_WALL) {
            uR[0] = uL[0];
            uR[D + 1] = uL[D + 1];
            R mn = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) mn += sL.m[d] * n[d];
#pragma unroll
            for (int d = 0; d < D; ++d) uR[1 + d] = sL.m[d] + (-2 * mn) * n[d];
        } else if (bctype == ICE_BC_NO_SLIP_ISOTHERMAL) {
            R prim[M];
            prim[0] = sL.rho;
#pragma unroll
            for (int d = 0; d < D; ++d) prim[1 + d] = 0;
            // wall temperature by an unrolled select: a runtime index into
            // the parameter copy would move the whole copy to scratch memory
            const int wall = bcflag & (ICE_MAX_BCFLAG - 1);
            R Tw = p.isoT[0];
#pragma unroll
            for (int k = 1; k < ICE_MAX_BCFLAG; ++k) Tw = (wall == k) ? p.isoT[k] : Tw;
            prim[D + 1] = Tw;
            cons_from_rho_u_T(p, prim, uR);
        } else {  // RIEMANN (Carlson, NASA/TM-2011-217181 as in :1014-1125)
            const R g = p.gamma;
            // free-stream state from (rho, u dir, T) as calc_thermo_state<RHO_U_T>
            const R rho_f = fmax(R(1e-14), p.fs_rho);
            const R p_f = fmax(R(1e-12), rho_f * (g - 1) / g * p.T_coeff * p.fs_T);
            const R c_f = sqrt((g * p.Eu * p_f) / rho_f);
            R vni = 0, vno = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) { vni += sL.v[d] * n[d]; vno += p.fs_dir[d] * p.fs_u * n[d]; }
            const R mach = vni / sL.c;
            R Rp = vni + 2 * sL.c / (g - 1);
            const R Rm = vno - 2 * c_f / (g - 1);
            if (mach < 0) Rp = vno + 2 * c_f / (g - 1);  // the outflow branch leaves both unchanged
            const R Ub = R(0.5) * (Rp + Rm);
            const R cb = R(0.25) * (g - 1) * (Rp - Rm);
            const R sb = sL.c * sL.c / (g * pow(sL.rho, g - 1));
            const R rhoR = pow(cb * cb / (g * sb), R(1.0) / (g - 1));
            const R pR = rhoR * cb * cb / g;
            R prim[M];
            prim[0] = rhoR;
#pragma unroll
            for (int d = 0; d < D; ++d) prim[1 + d] = sL.v[d] + (Ub - vni) * n[d];
            prim[D + 1] = pR;
            cons_from_rho_u_p(p, prim, uR);
        }
    }
};

}  // namespace ice
