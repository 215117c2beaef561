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

// Parameters of every flux family, flattened for a kernel argument.
struct PhysParams {
    // Burgers (BurgersCoefficients, burgers.hpp:14-40)
    double mu;
    double a[3];
    double b[3];
    // Navier-Stokes (Physics + CaloricallyPerfectEoS + Nondimensionalization)
    double gamma;
    double Pr;
    double Re, Eu, e_coeff;
    double T_coeff;           // cp * ref.T * ref.rho / ref.p
    int visc_kind;
    double visc_mu;
    double mu_ratio, T0r, Tsr;  // DimensionlessSutherlands ratios
    double isoT[ICE_MAX_BCFLAG];
    double fs_rho, fs_u, fs_dir[3], fs_T;
    int textbook_energy;
    // ConservationLawDDG
    double sigma_ic;
    double beta0, beta1, beta0_bdy;
};

template <int D>
struct Burgers {
    static constexpr int M = 1;
    static constexpr bool IMPLEMENTS_BCS = false;
    static constexpr bool HAS_HOMOGENEITY = true;

    __device__ __forceinline__ static void flux(const PhysParams& p, const double* u, const double* g, double* F) {
#pragma unroll
        for (int d = 0; d < D; ++d) F[d] = (p.a[d] + 0.5 * p.b[d] * u[0]) * u[0] - p.mu * g[d];
    }
    // |a + b u / 2|, the wave speed the reference tracks in lambda_max (burgers.hpp:79-90)
    __device__ __forceinline__ static double wave_speed(const PhysParams& p, const double* u) {
        double s = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double l = p.a[d] + 0.5 * p.b[d] * u[0];
            s += l * l;
        }
        return sqrt(s);
    }
    __device__ __forceinline__ static double visc_at(const PhysParams& p, const double*) { return p.mu; }
    __device__ __forceinline__ static void conv(const PhysParams& p, const double* uL, const double* uR,
                                                const double* n, double* out) {
        const double ua = 0.5 * (uL[0] + uR[0]);
        double P = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) P += n[d] * (p.a[d] + 0.5 * p.b[d] * ua);
        const double uu = (P > 0) ? uL[0] : uR[0];
        double f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) f += n[d] * (p.a[d] + 0.5 * p.b[d] * uu);
        out[0] = f * uu;
    }
    __device__ __forceinline__ static void diff(const PhysParams& p, const double*, const double* g, const double* n,
                                                double* out) {
        double f = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) f += p.mu * g[d] * n[d];
        out[0] = f;
    }
    __device__ __forceinline__ static void neumann(const PhysParams& p, const double* g, double* out) { out[0] = p.mu * g[0]; }
    // G[0][k][0][s] = mu delta_ks
    __device__ __forceinline__ static void homogeneity(const PhysParams& p, const double*, double* G) {
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
            for (int s = 0; s < D; ++s) G[k * D + s] = (k == s) ? p.mu : 0.0;
    }
    __device__ __forceinline__ static void apply_bc(const PhysParams&, const double* uL, const double*, const double*, int, int,
                                                    double* uR) { uR[0] = uL[0]; }
};

// Compressible Navier-Stokes / Euler in conservative variables.
template <int D, bool FULL>
struct NavierStokes {
    static constexpr int M = D + 2;
    static constexpr bool IMPLEMENTS_BCS = true;
    static constexpr bool HAS_HOMOGENEITY = false;

    struct State {
        double rho, irho, rhoE, E, e, p, T, c, vv;
        double m[D], v[D];
    };

    // calc_thermo_state<CONSERVATIVE> (navier_stokes.hpp:405-449)
    __device__ __forceinline__ static State state(const PhysParams& p, const double* u) {
        State s;
        s.rho = fmax(1e-14, u[0]);
        s.irho = 1.0 / s.rho;
        s.vv = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            s.m[d] = u[1 + d];
            s.v[d] = u[1 + d] * s.irho;
            s.vv += s.v[d] * s.v[d];
        }
        s.rhoE = u[D + 1];
        s.E = s.rhoE * s.irho;
        s.e = s.E - 0.5 * p.e_coeff * s.vv;
        s.p = fmax(1e-12, (p.gamma - 1) / (p.e_coeff * p.Eu) * s.rho * s.e);
        s.T = s.p * s.irho * (p.gamma / ((p.gamma - 1) * p.T_coeff));
        s.c = sqrt((p.gamma * p.Eu) * s.p * s.irho);
        return s;
    }
    __device__ __forceinline__ static double viscosity(const PhysParams& p, double T) {
        if (p.visc_kind == ICE_VISC_CONSTANT) return p.visc_mu;
        return p.mu_ratio * pow(T / p.T0r, 1.5) * (p.T0r + p.Tsr) / (T + p.Tsr);
    }
    __device__ __forceinline__ static double wave_speed(const PhysParams& p, const double* u) {
        const State s = state(p, u);
        return s.c + sqrt(s.vv);
    }
    __device__ __forceinline__ static double visc_at(const PhysParams& p, const double* u) {
        return FULL ? viscosity(p, state(p, u).T) : 0.0;
    }

    // velocity and total-energy gradients (calc_state_gradients<CONSERVATIVE>, :541-575),
    // shear stress (:700-723) and heat flux (:730-746)
    __device__ __forceinline__ static void viscous(const PhysParams& p, const State& s, const double* g, double mu,
                                                   double (&tau)[D][D], double (&gE)[D], double (&q)[D]) {
        double gv[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) gv[i][j] = (g[(1 + i) * D + j] - s.v[i] * g[j]) * s.irho;
#pragma unroll
        for (int j = 0; j < D; ++j) gE[j] = (g[(D + 1) * D + j] - s.E * g[j]) * s.irho;
        double div = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) div += gv[k][k];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) tau[i][j] = mu * (gv[i][j] + gv[j][i]);
#pragma unroll
        for (int i = 0; i < D; ++i) tau[i][i] -= 2.0 / 3.0 * mu * div;
        const double iec = 1.0 / p.e_coeff, kq = mu * p.gamma / p.Pr;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double qj = gE[j] * iec;
#pragma unroll
            for (int k = 0; k < D; ++k) qj -= s.v[k] * gv[k][j];
            q[j] = qj * kq;
        }
    }

    // Flux::operator() (:915-973)
    __device__ __forceinline__ static void flux(const PhysParams& p, const double* u, const double* g, double* F) {
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
            double tau[D][D], gE[D], q[D];
            viscous(p, s, g, viscosity(p, s.T), tau, gE, q);
            const double iRe = 1.0 / p.Re, ecRe = p.e_coeff * iRe;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double ef = q[j];
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
    __device__ __forceinline__ static void conv(const PhysParams& p, const double* uL, const double* uR, const double* n,
                                                double* f) {
        const State L = state(p, uL), R = state(p, uR);
        const double Eu = p.Eu, ec = p.e_coeff, g = p.gamma;
        double vnL = 0, vnR = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) { vnL += L.v[d] * n[d]; vnR += R.v[d] * n[d]; }
        const double mL = vnL / L.c, mR = vnR / R.c;
        if (mL > 1) {
            f[0] = L.rho * vnL;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] = L.m[d] * vnL + Eu * L.p * n[d];
            f[D + 1] = vnL * (L.rhoE + Eu * ec * L.p);
        } else if (mL < -1) {
#pragma unroll
            for (int k = 0; k < M; ++k) f[k] = 0;
        } else {
            const double fm = L.rho * L.c * (mL + 1) * (mL + 1) / 4.0;
            f[0] = fm;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] = fm * (L.v[d] + n[d] * (-vnL + 2 * L.c) / g);
            const double t = (g - 1) * vnL + 2 * L.c;
            f[D + 1] = fm * ((L.vv - vnL * vnL) / 2 + t * t / (2 * (g * g - 1)));
        }
        if (mR < -1) {
            f[0] += R.rho * vnR;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] += R.m[d] * vnR + Eu * R.p * n[d];
            f[D + 1] += vnR * (R.rhoE + Eu * ec * R.p);
        } else if (mR <= 1) {
            const double fm = -R.rho * R.c * (mR - 1) * (mR - 1) / 4.0;
            f[0] += fm;
#pragma unroll
            for (int d = 0; d < D; ++d) f[1 + d] += fm * (R.v[d] + n[d] * (-vnR - 2 * R.c) / g);
            const double t = (g - 1) * vnR - 2 * R.c;
            f[D + 1] += fm * ((R.vv - vnR * vnR) / 2 + t * t / (2 * (g * g - 1)));
        }
    }

    // DiffusionFlux (:1241-1289).  Energy: mu gamma / Pr * dE/dx . n as in
    // :1277-1279, or the volume heat flux q . n when textbook_energy is set.
    __device__ __forceinline__ static void diff(const PhysParams& p, const double* u, const double* g, const double* n,
                                                double* f) {
#pragma unroll
        for (int k = 0; k < M; ++k) f[k] = 0;
        if constexpr (FULL) {
            const State s = state(p, u);
            const double mu = viscosity(p, s.T);
            double tau[D][D], gE[D], q[D];
            viscous(p, s, g, mu, tau, gE, q);
            const double iRe = 1.0 / p.Re, ecRe = p.e_coeff * iRe, kE = mu * p.gamma / p.Pr;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double ef = p.textbook_energy ? q[j] * n[j] : kE * gE[j] * n[j];
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    f[1 + i] += tau[i][j] * iRe * n[j];
                    ef += s.v[i] * tau[i][j] * n[j];
                }
                f[D + 1] += ef * ecRe;
            }
        }
    }
    __device__ __forceinline__ static void neumann(const PhysParams&, const double*, double* out) {
#pragma unroll
        for (int k = 0; k < M; ++k) out[k] = 0;
    }
    __device__ __forceinline__ static void homogeneity(const PhysParams&, const double*, double*) {}

    // conservative state from (rho, velocity, T) (calc_thermo_state<RHO_U_T>, :450-488)
    __device__ __forceinline__ static void cons_from_rho_u_T(const PhysParams& p, const double* prim, double* out) {
        const double rho = fmax(1e-14, prim[0]);
        double vv = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) vv += prim[1 + d] * prim[1 + d];
        const double pr = fmax(1e-12, rho * (p.gamma - 1) / p.gamma * p.T_coeff * prim[D + 1]);
        const double e = pr * p.e_coeff * p.Eu / (p.gamma - 1) / rho;
        const double E = e + 0.5 * p.e_coeff * vv;
        out[0] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) out[1 + d] = prim[1 + d] * rho;
        out[D + 1] = rho * E;
    }
    // conservative state from (rho, velocity, p) (calc_thermo_state<RHO_U_P>, :489-525)
    __device__ __forceinline__ static void cons_from_rho_u_p(const PhysParams& p, const double* prim, double* out) {
        const double rho = fmax(1e-14, prim[0]);
        const double pr = fmax(1e-12, prim[D + 1]);
        double vv = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) vv += prim[1 + d] * prim[1 + d];
        const double e = pr * p.e_coeff * p.Eu / (p.gamma - 1) / rho;
        const double E = e + 0.5 * p.e_coeff * vv;
        out[0] = rho;
#pragma unroll
        for (int d = 0; d < D; ++d) out[1 + d] = prim[1 + d] * rho;
        out[D + 1] = rho * E;
    }

    // Flux::apply_bc (:975-1189); exterior gradient = interior gradient
    __device__ __forceinline__ static void apply_bc(const PhysParams& p, const double* uL, const double*, const double* n,
                                                    int bctype, int bcflag, double* uR) {
        if (bctype == ICE_BC_WALL_GENERAL) bctype = FULL ? ICE_BC_NO_SLIP_ISOTHERMAL : ICE_BC_SLIP_WALL;
        const State sL = state(p, uL);
        if (bctype == ICE_BC_SLIP_WALL) {
            uR[0] = uL[0];
            uR[D + 1] = uL[D + 1];
            double mn = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) mn += sL.m[d] * n[d];
#pragma unroll
            for (int d = 0; d < D; ++d) uR[1 + d] = sL.m[d] + (-2 * mn) * n[d];
        } else if (bctype == ICE_BC_NO_SLIP_ISOTHERMAL) {
            double prim[M];
            prim[0] = sL.rho;
#pragma unroll
            for (int d = 0; d < D; ++d) prim[1 + d] = 0;
            prim[D + 1] = p.isoT[bcflag & (ICE_MAX_BCFLAG - 1)];
            cons_from_rho_u_T(p, prim, uR);
        } else {  // RIEMANN (Carlson, NASA/TM-2011-217181 as in :1014-1125)
            const double g = p.gamma;
            // free-stream state from (rho, u dir, T) as calc_thermo_state<RHO_U_T>
            const double rho_f = fmax(1e-14, p.fs_rho);
            const double p_f = fmax(1e-12, rho_f * (g - 1) / g * p.T_coeff * p.fs_T);
            const double c_f = sqrt((g * p.Eu * p_f) / rho_f);
            double vni = 0, vno = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) { vni += sL.v[d] * n[d]; vno += p.fs_dir[d] * p.fs_u * n[d]; }
            const double mach = vni / sL.c;
            double Rp = vni + 2 * sL.c / (g - 1);
            const double Rm = vno - 2 * c_f / (g - 1);
            if (mach < 0) Rp = vno + 2 * c_f / (g - 1);  // the outflow branch leaves both unchanged
            const double Ub = 0.5 * (Rp + Rm);
            const double cb = 0.25 * (g - 1) * (Rp - Rm);
            const double sb = sL.c * sL.c / (g * pow(sL.rho, g - 1));
            const double rhoR = pow(cb * cb / (g * sb), 1.0 / (g - 1));
            const double pR = rhoR * cb * cb / g;
            double prim[M];
            prim[0] = rhoR;
#pragma unroll
            for (int d = 0; d < D; ++d) prim[1 + d] = sL.v[d] + (Ub - vni) * n[d];
            prim[D + 1] = pR;
            cons_from_rho_u_p(p, prim, uR);
        }
    }
};

}  // namespace ice
