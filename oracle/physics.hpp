// CPU ORACLE (test infrastructure only; never linked into the product).
//
// Flux functors restated from the reference:
//  Burgers: disc/burgers.hpp:14-255
//  Navier-Stokes / Euler (CONSERVATIVE variable set): disc/navier_stokes.hpp:12-1307
// Each functor keeps the reference's call shape:
//  physical flux  F(u, gradu) -> [neq][D]       (conservation_law.hpp:20-28)
//  convective     Fc(uL, uR, n) -> [neq]        (:34-42)
//  diffusive      Fd(u, gradu, n) -> [neq]      (:50-59), neumann_flux
//  optional apply_bc (:63-80), optional homogeneity_tensor (:84-90)
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

#include "mesh.hpp"

namespace oracle {

// ======================= Burgers ==========================================
struct BurgersCoefficients {
    double mu = 0.001;
    double a[3] = {0, 0, 0};
    double b[3] = {0, 0, 0};
};

template <int D>
struct BurgersFlux {
    static constexpr int neq = 1;
    static constexpr bool implements_bcs = false;
    const BurgersCoefficients* c;
    mutable double lambda_max = 0.0;
    // burgers.hpp:72-91
    void operator()(const double* u, const double* gradu, double* F) const {
        double lnorm = 0;
        for (int d = 0; d < D; ++d) {
            const double lambda = c->a[d] + 0.5 * c->b[d] * u[0];
            lnorm += lambda * lambda;
            F[d] = lambda * u[0] - c->mu * gradu[d];
        }
        lnorm = std::sqrt(lnorm);
        lambda_max = std::max(lambda_max, lnorm);
    }
    // burgers.hpp:103-108
    double dt_from_cfl(double cfl, double reflen) const {
        return (reflen * cfl) / (c->mu / reflen + lambda_max);
    }
    void apply_bc(double*, const double*, const double*, int, int, double*, double*) const {}
};

template <int D>
struct BurgersUpwind {
    static constexpr int neq = 1;
    const BurgersCoefficients* c;
    // burgers.hpp:136-169
    void operator()(const double* uL, const double* uR, const double* n, double* out) const {
        const double uavg = 0.5 * (uL[0] + uR[0]);
        double P = 0;
        for (int d = 0; d < D; ++d) P += n[d] * (c->a[d] + 0.5 * c->b[d] * uavg);
        const double uu = (P > 0) ? uL[0] : uR[0];
        double f = 0;
        for (int d = 0; d < D; ++d) f += n[d] * (c->a[d] + 0.5 * c->b[d] * uu);
        out[0] = f * uu;
    }
};

template <int D>
struct BurgersDiffusionFlux {
    static constexpr int neq = 1;
    static constexpr bool has_homogeneity = true;
    const BurgersCoefficients* c;
    // burgers.hpp:206-220
    void operator()(const double*, const double* gradu, const double* n, double* out) const {
        double f = 0;
        for (int d = 0; d < D; ++d) f += c->mu * gradu[d] * n[d];
        out[0] = f;
    }
    // burgers.hpp:224-229
    void neumann_flux(const double* g, double* out) const { out[0] = c->mu * g[0]; }
    // burgers.hpp:233-249: G[0][k][0][s] = mu delta_ks
    void homogeneity_tensor(const double*, double* G) const {
        for (int k = 0; k < D; ++k)
            for (int s = 0; s < D; ++s) G[k * D + s] = (k == s) ? c->mu : 0.0;
    }
};

// ======================= Navier-Stokes ====================================
struct ReferenceParameters { double t = 1, l = 1, rho = 1, u = 1, e = 1, p = 1, T = 1, mu = 1; };
struct Nondim { double Re, Eu, Sr, e_coeff; };

// create_nondim (navier_stokes.hpp:207-213)
inline Nondim create_nondim(const ReferenceParameters& r) {
    return Nondim{r.rho * r.u * r.l / r.mu, r.p / (r.rho * r.u * r.u), r.l / (r.u * r.t), r.u * r.u / r.e};
}

template <int D>
struct ThermoState {
    double rho, rhoE, gamma, cp, T, vv, p, csound, e, E, H;
    double momentum[D], velocity[D];
};

template <int D>
struct NSPhysics {
    static constexpr double MIN_PRESSURE = 1e-12;  // navier_stokes.hpp:332
    static constexpr double MIN_DENSITY = 1e-14;   // :334
    double gamma = 1.4;
    double Rgas = 287.052874;
    double Pr = 0.72;
    ReferenceParameters ref{};
    Nondim nondim{1, 1, 1, 1};
    int visc_kind = 0;        // 0 constant, 1 dimensionless Sutherland
    double visc_mu = 1.0;
    double isothermal_T[4] = {1, 1, 1, 1};
    double fs_rho = 1, fs_u = 0, fs_dir[3] = {1, 0, 0}, fs_T = 1;

    // constant_viscosity (:231-242) / DimensionlessSutherlands (:271-301)
    double viscosity(double That) const {
        if (visc_kind == 0) return visc_mu;
        const double mu_0 = 1.716e-5, T_0 = 273.1, T_s = 110.5;
        const double mu_ratio = mu_0 / ref.mu, T0r = T_0 / ref.T, Tsr = T_s / ref.T;
        return mu_ratio * std::pow(That / T0r, 1.5) * (T0r + Tsr) / (That + Tsr);
    }

    // calc_thermo_state<CONSERVATIVE> (:405-449)
    ThermoState<D> state_cons(const double* u) const {
        ThermoState<D> s;
        s.rho = std::max(MIN_DENSITY, u[0]);
        s.vv = 0;
        for (int d = 0; d < D; ++d) {
            s.momentum[d] = u[1 + d];
            s.velocity[d] = u[1 + d] / s.rho;
            s.vv += s.velocity[d] * s.velocity[d];
        }
        s.rhoE = u[D + 1];
        s.gamma = gamma;
        s.cp = gamma / (gamma - 1) * Rgas;
        const double T_coeff = s.cp * ref.T * ref.rho / ref.p;
        s.E = s.rhoE / s.rho;
        s.e = s.E - 0.5 * nondim.e_coeff * s.vv;
        s.p = std::max(MIN_PRESSURE, (gamma - 1) / (nondim.e_coeff * nondim.Eu) * s.rho * s.e);
        s.T = s.p / s.rho * gamma / (gamma - 1) / T_coeff;
        s.csound = std::sqrt((gamma * nondim.Eu * s.p) / s.rho);
        s.H = s.E + s.p / s.rho;
        return s;
    }
    // calc_thermo_state<RHO_U_T> (:450-488)
    ThermoState<D> state_rho_u_T(const double* u) const {
        ThermoState<D> s;
        s.rho = std::max(MIN_DENSITY, u[0]);
        s.vv = 0;
        for (int d = 0; d < D; ++d) {
            s.velocity[d] = u[1 + d];
            s.momentum[d] = u[1 + d] * s.rho;
            s.vv += s.velocity[d] * s.velocity[d];
        }
        s.T = u[D + 1];
        s.gamma = gamma;
        s.cp = gamma / (gamma - 1) * Rgas;
        const double T_coeff = s.cp * ref.T * ref.rho / ref.p;
        s.p = std::max(MIN_PRESSURE, s.rho * (gamma - 1) / gamma * T_coeff * s.T);
        s.e = s.p * nondim.e_coeff * nondim.Eu / (gamma - 1) / s.rho;
        s.E = s.e + 0.5 * nondim.e_coeff * s.vv;
        s.rhoE = s.rho * s.E;
        s.csound = std::sqrt((gamma * nondim.Eu * s.p) / s.rho);
        s.H = s.E + s.p / s.rho;
        return s;
    }
    // calc_thermo_state<RHO_U_P> (:489-525)
    ThermoState<D> state_rho_u_p(const double* u) const {
        ThermoState<D> s;
        s.rho = std::max(MIN_DENSITY, u[0]);
        s.p = std::max(MIN_PRESSURE, u[D + 1]);
        s.gamma = gamma;
        s.cp = gamma / (gamma - 1) * Rgas;
        const double T_coeff = s.cp * ref.T * ref.rho / ref.p;
        s.T = s.p / s.rho * gamma / (gamma - 1) / T_coeff;
        s.vv = 0;
        for (int d = 0; d < D; ++d) {
            s.velocity[d] = u[1 + d];
            s.momentum[d] = u[1 + d] * s.rho;
            s.vv += s.velocity[d] * s.velocity[d];
        }
        s.e = s.p * nondim.e_coeff * nondim.Eu / (gamma - 1) / s.rho;
        s.E = s.e + 0.5 * nondim.e_coeff * s.vv;
        s.rhoE = s.rho * s.E;
        s.csound = std::sqrt((gamma * nondim.Eu * s.p) / s.rho);
        s.H = s.E + s.p / s.rho;
        return s;
    }
    static void to_cons(const ThermoState<D>& s, double* out) {
        out[0] = s.rho;
        for (int d = 0; d < D; ++d) out[1 + d] = s.momentum[d];
        out[D + 1] = s.rhoE;
    }
    // calc_state_gradients<CONSERVATIVE> (:541-575): velocity and E gradients
    void state_grads(const ThermoState<D>& s, const double* gradu, double (&gv)[D][D], double (&gE)[D]) const {
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j)
                gv[i][j] = gradu[(1 + i) * D + j] / s.rho - s.velocity[i] / s.rho * gradu[0 * D + j];
        for (int j = 0; j < D; ++j) gE[j] = (gradu[(D + 1) * D + j] - s.E * gradu[0 * D + j]) / s.rho;
    }
    // calc_shear_stress (:700-723)
    void shear(const ThermoState<D>& s, const double (&gv)[D][D], double (&tau)[D][D]) const {
        const double mu = viscosity(s.T);
        double div = 0;
        for (int k = 0; k < D; ++k) div += gv[k][k];
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) tau[i][j] = mu * (gv[i][j] + gv[j][i]);
        for (int i = 0; i < D; ++i) tau[i][i] -= 2.0 / 3.0 * mu * div;
    }
    // calc_heat_flux (:730-746)
    void heat(const ThermoState<D>& s, const double (&gv)[D][D], const double (&gE)[D], double (&q)[D]) const {
        const double mu = viscosity(s.T);
        for (int j = 0; j < D; ++j) {
            q[j] = gE[j] / nondim.e_coeff;
            for (int k = 0; k < D; ++k) q[j] -= s.velocity[k] * gv[k][j];
            q[j] *= mu * s.gamma / Pr;
        }
    }
};

// Flux<real, ndim, EoS, CONSERVATIVE, full_ns> (navier_stokes.hpp:873-1204)
template <int D, bool FULL_NS>
struct NSFlux {
    static constexpr int neq = D + 2;
    static constexpr bool implements_bcs = true;
    const NSPhysics<D>* ph;
    mutable double lambda_max = 0.0, gamma_max = 0.0, visc_max = 0.0;

    // :915-973
    void operator()(const double* u, const double* gradu, double* F) const {
        const auto s = ph->state_cons(u);
        const double Re = ph->nondim.Re, ec = ph->nondim.e_coeff, Eu = ph->nondim.Eu;
        lambda_max = s.csound + std::sqrt(s.vv);
        for (int j = 0; j < D; ++j) {
            F[0 * D + j] = s.momentum[j];
            for (int i = 0; i < D; ++i) F[(1 + i) * D + j] = s.momentum[i] * s.velocity[j];
            F[(1 + j) * D + j] += Eu * s.p;
            F[(D + 1) * D + j] = s.velocity[j] * (s.rhoE + Eu * ec * s.p);
        }
        if (FULL_NS) {
            double gv[D][D], gE[D], tau[D][D], q[D];
            ph->state_grads(s, gradu, gv, gE);
            ph->shear(s, gv, tau);
            ph->heat(s, gv, gE, q);
            for (int j = 0; j < D; ++j) {
                double ef = q[j];
                for (int i = 0; i < D; ++i) {
                    F[(1 + i) * D + j] -= tau[i][j] / Re;
                    ef += s.velocity[i] * tau[i][j];
                }
                ef *= ec / Re;
                F[(D + 1) * D + j] -= ef;
            }
            const double mu = ph->viscosity(s.T);
            visc_max = std::max(visc_max, mu);
            gamma_max = std::max(gamma_max, s.gamma);
        }
    }

    // apply_bc (:975-1189): SLIP_WALL, NO_SLIP_ISOTHERMAL (+WALL_GENERAL), RIEMANN
    void apply_bc(double* uL, const double* graduL, const double* n, int bctype, int bcflag,
                  double* uR, double* graduR) const {
        for (int k = 0; k < neq * D; ++k) graduR[k] = graduL[k];
        if (bctype == BC_WALL_GENERAL) bctype = FULL_NS ? BC_NO_SLIP_ISOTHERMAL : BC_SLIP_WALL;
        if (bctype == BC_SLIP_WALL) {
            const auto sL = ph->state_cons(uL);
            uR[0] = uL[0];
            uR[D + 1] = uL[D + 1];
            double mn = 0;
            for (int d = 0; d < D; ++d) mn += sL.momentum[d] * n[d];
            for (int d = 0; d < D; ++d) uR[1 + d] = sL.momentum[d] + (-2 * mn) * n[d];
        } else if (bctype == BC_NO_SLIP_ISOTHERMAL) {
            const auto sL = ph->state_cons(uL);
            double prim[D + 2];
            prim[0] = sL.rho;
            for (int d = 0; d < D; ++d) prim[1 + d] = 0;
            prim[D + 1] = ph->isothermal_T[bcflag];
            NSPhysics<D>::to_cons(ph->state_rho_u_T(prim), uR);
        } else if (bctype == BC_RIEMANN) {
            const auto sL = ph->state_cons(uL);
            const double g = sL.gamma;
            double fs[D + 2];
            fs[0] = ph->fs_rho;
            for (int d = 0; d < D; ++d) fs[1 + d] = ph->fs_dir[d] * ph->fs_u;
            fs[D + 1] = ph->fs_T;
            const auto sfs = ph->state_rho_u_T(fs);
            double vni = 0, vno = 0;
            for (int d = 0; d < D; ++d) { vni += sL.velocity[d] * n[d]; vno += sfs.velocity[d] * n[d]; }
            const double mach = vni / sL.csound;
            double Rplus = vni + 2 * sL.csound / (g - 1);
            double Rminus = vno - 2 * sfs.csound / (g - 1);
            if (mach > 0) {
                // the reference declares a new local Rminus here (:1073-1076): no effect
            } else if (mach < 0) {
                Rplus = vno + 2 * sfs.csound / (g - 1);
            }
            const double Ub = 0.5 * (Rplus + Rminus);
            const double cb = 0.25 * (g - 1) * (Rplus - Rminus);
            double uadv[D];
            for (int d = 0; d < D; ++d) uadv[d] = sL.velocity[d] + (Ub - vni) * n[d];
            const double sb = sL.csound * sL.csound / (g * std::pow(sL.rho, g - 1));
            const double rhoR = std::pow(cb * cb / (g * sb), 1.0 / (g - 1));
            const double pR = rhoR * cb * cb / g;
            double prim[D + 2];
            prim[0] = rhoR;
            for (int d = 0; d < D; ++d) prim[1 + d] = uadv[d];
            prim[D + 1] = pR;
            NSPhysics<D>::to_cons(ph->state_rho_u_p(prim), uR);
        } else {
            throw std::runtime_error("NS apply_bc: unsupported boundary condition (reference logs an anomaly and returns an uninitialised state)");
        }
    }

    // dt_from_cfl (:1191-1203)
    double dt_from_cfl(double cfl, double reflen) const {
        double dt = (reflen * cfl) / lambda_max;
        if (FULL_NS) dt = std::min(dt, reflen * reflen * cfl / visc_max);
        lambda_max = 0;
        visc_max = 0;
        return dt;
    }
};

// VanLeer (navier_stokes.hpp:777-864)
template <int D>
struct VanLeer {
    static constexpr int neq = D + 2;
    const NSPhysics<D>* ph;
    void operator()(const double* uL, const double* uR, const double* n, double* f) const {
        const auto sL = ph->state_cons(uL);
        const auto sR = ph->state_cons(uR);
        const double Eu = ph->nondim.Eu, ec = ph->nondim.e_coeff;
        double vnL = 0, vnR = 0;
        for (int d = 0; d < D; ++d) { vnL += sL.velocity[d] * n[d]; vnR += sR.velocity[d] * n[d]; }
        const double mL = vnL / sL.csound, mR = vnR / sR.csound;
        if (mL > 1) {
            f[0] = sL.rho * vnL;
            for (int d = 0; d < D; ++d) f[1 + d] = sL.momentum[d] * vnL + Eu * sL.p * n[d];
            f[D + 1] = vnL * (sL.rhoE + Eu * ec * sL.p);
        } else if (mL < -1) {
            for (int k = 0; k < neq; ++k) f[k] = 0;
        } else {
            const double fm = sL.rho * sL.csound * (mL + 1) * (mL + 1) / 4.0;
            f[0] = fm;
            for (int d = 0; d < D; ++d)
                f[1 + d] = fm * (sL.velocity[d] + n[d] * (-vnL + 2 * sL.csound) / sL.gamma);
            const double t = (sL.gamma - 1) * vnL + 2 * sL.csound;
            f[D + 1] = fm * ((sL.vv - vnL * vnL) / 2 + t * t / (2 * (sL.gamma * sL.gamma - 1)));
        }
        if (mR < -1) {
            f[0] += sR.rho * vnR;
            for (int d = 0; d < D; ++d) f[1 + d] += sR.momentum[d] * vnR + Eu * sR.p * n[d];
            f[D + 1] += vnR * (sR.rhoE + Eu * ec * sR.p);
        } else if (mR <= 1) {
            const double fm = -sR.rho * sR.csound * (mR - 1) * (mR - 1) / 4.0;
            f[0] += fm;
            for (int d = 0; d < D; ++d)
                f[1 + d] += fm * (sR.velocity[d] + n[d] * (-vnR - 2 * sR.csound) / sR.gamma);
            const double t = (sR.gamma - 1) * vnR - 2 * sR.csound;
            f[D + 1] += fm * ((sR.vv - vnR * vnR) / 2 + t * t / (2 * (sR.gamma * sR.gamma - 1)));
        }
    }
};

// DiffusionFlux (navier_stokes.hpp:1211-1301)
template <int D, bool FULL_NS>
struct NSDiffusionFlux {
    static constexpr int neq = D + 2;
    static constexpr bool has_homogeneity = false;
    const NSPhysics<D>* ph;
    bool textbook_energy = false;  // true: heat flux consistent with the volume flux
    void operator()(const double* u, const double* gradu, const double* n, double* f) const {
        for (int k = 0; k < neq; ++k) f[k] = 0;
        if (!FULL_NS) return;
        const auto s = ph->state_cons(u);
        const double Re = ph->nondim.Re, ec = ph->nondim.e_coeff;
        double gv[D][D], gE[D], tau[D][D], q[D];
        ph->state_grads(s, gradu, gv, gE);
        ph->shear(s, gv, tau);
        ph->heat(s, gv, gE, q);
        for (int j = 0; j < D; ++j) {
            double ef;
            if (textbook_energy) {
                ef = q[j] * n[j];
            } else {
                // :1276-1279 the q.n value is overwritten by mu gamma / Pr * dE/dx_j n_j
                const double mu = ph->viscosity(s.T);
                ef = mu * s.gamma / ph->Pr * gE[j] * n[j];
            }
            for (int i = 0; i < D; ++i) {
                f[1 + i] += tau[i][j] / Re * n[j];
                ef += s.velocity[i] * tau[i][j] * n[j];
            }
            ef *= ec / Re;
            f[D + 1] += ef;
        }
    }
    void neumann_flux(const double*, double* out) const { for (int k = 0; k < neq; ++k) out[k] = 0; }
    void homogeneity_tensor(const double*, double*) const {}
};

}  // namespace oracle
