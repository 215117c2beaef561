// CPU ORACLE (test infrastructure only; never linked into the product).
//
// Callers of form_residual restated from the reference:
//  L2 projection + element mass solve (disc/projection.hpp:83-110,
//    solvers/element_linear_solve.hpp:37-90), l2_error (disc/l2_error.hpp:37-94),
//  calculate_mass_matrix (element/finite_element.hpp:586-611),
//  CFLTimestep / FixedTimestep / termination (solvers/explicit_utils.hpp:23-166),
//  RK3SSP::step (solvers/ssp_rk3.hpp:106-195), RK3TVD::step (tvd_rk3.hpp:107-231),
//  ExplicitEuler::step (explicit_euler.hpp:87-138),
//  form_petsc_jacobian_fd DG part (form_petsc_jacobian.hpp:55-317, fd_utils.hpp:10-14),
//  domain_integral_jacobian (conservation_law.hpp:316-432).
#pragma once
#include <map>

#include "disc.hpp"

namespace oracle {

template <int D>
std::vector<double> mass_matrix(const FESpace<D>& fs, int e) {
    const int nb = fs.nb;
    std::vector<double> M(nb * nb, 0.0);
    const auto g = fs.geo(e);
    for (int q = 0; q < fs.nq; ++q) {
        double J[D][D], H[D][D][D];
        g.eval(fs.q1, fs.quad.pts[q].data(), nullptr, J, H);
        const double detJ = determinant<D>(J);
        const double* bi = &fs.B[q * nb];
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) M[i * nb + j] += bi[i] * bi[j] * fs.quad.w[q] * detJ;
    }
    return M;
}

// Projection (projection.hpp:83-110) + ElementLinearSolver solve per element
template <int D>
void project(const FESpace<D>& fs, int neq, const Callback& f, double* u) {
    const int nb = fs.nb;
    std::vector<double> b(nb * neq), bb(nb), xx(nb);
    for (int e = 0; e < fs.mesh.nelem; ++e) {
        std::fill(b.begin(), b.end(), 0.0);
        const auto g = fs.geo(e);
        for (int q = 0; q < fs.nq; ++q) {
            double x[D], J[D][D], H[D][D][D];
            g.eval(fs.q1, fs.quad.pts[q].data(), x, J, H);
            const double detJ = determinant<D>(J);
            double fe[8];
            f(x, fe);
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v) b[i * neq + v] += fe[v] * fs.quad.w[q] * fs.B[q * nb + i] * detJ;
        }
        DenseLU lu;
        lu.factor(mass_matrix(fs, e), nb);
        for (int v = 0; v < neq; ++v) {
            for (int i = 0; i < nb; ++i) bb[i] = b[i * neq + v];
            lu.solve(bb.data(), xx.data());
            for (int i = 0; i < nb; ++i) u[((size_t)e * nb + i) * neq + v] = xx[i];
        }
    }
}

// l2_error (l2_error.hpp:37-94), |detJ|
template <int D>
double l2_error(const FESpace<D>& fs, int neq, const Callback& exact, const double* u) {
    const int nb = fs.nb;
    std::vector<double> l2(neq, 0.0);
    for (int e = 0; e < fs.mesh.nelem; ++e) {
        const auto g = fs.geo(e);
        for (int q = 0; q < fs.nq; ++q) {
            double x[D], J[D][D], H[D][D][D];
            g.eval(fs.q1, fs.quad.pts[q].data(), x, J, H);
            const double detJ = determinant<D>(J);
            double fe[8], uu[8] = {};
            exact(x, fe);
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v) uu[v] += fs.B[q * nb + i] * u[((size_t)e * nb + i) * neq + v];
            for (int v = 0; v < neq; ++v) l2[v] += std::pow(uu[v] - fe[v], 2) * fs.quad.w[q] * std::abs(detJ);
        }
    }
    double s = 0;
    for (int v = 0; v < neq; ++v) s += l2[v];
    return std::sqrt(s);
}

// Per-element LU factors of the mass matrices.  The reference rebuilds and
// refactors them at every stage (ssp_rk3.hpp:142-163); the arithmetic is the
// same so they are cached here.
template <int D>
struct MassInverse {
    std::vector<DenseLU> lu;
    explicit MassInverse(const FESpace<D>& fs) {
        lu.resize(fs.mesh.nelem);
        for (int e = 0; e < fs.mesh.nelem; ++e) lu[e].factor(mass_matrix(fs, e), fs.nb);
    }
    // out += M^-1 r per element and component (singular: 1e-8 * r, tvd_rk3.hpp:190-204)
    void apply_add(const FESpace<D>& fs, int neq, const double* r, double* out) const {
        const int nb = fs.nb;
        std::vector<double> b(nb), x(nb);
        for (int e = 0; e < fs.mesh.nelem; ++e)
            for (int v = 0; v < neq; ++v) {
                for (int i = 0; i < nb; ++i) b[i] = r[((size_t)e * nb + i) * neq + v];
                if (lu[e].singular) {
                    for (int i = 0; i < nb; ++i) out[((size_t)e * nb + i) * neq + v] += 1e-8 * b[i];
                    continue;
                }
                lu[e].solve(b.data(), x.data());
                for (int i = 0; i < nb; ++i) out[((size_t)e * nb + i) * neq + v] += x[i];
            }
    }
};

// CFLTimestep::operator() (explicit_utils.hpp:61-90)
template <int D, class Disc>
double cfl_timestep(const FESpace<D>& fs, Disc& disc, double cfl) {
    double reflen = 1e8;
    for (int e = 0; e < fs.mesh.nelem; ++e) {
        double xi[D] = {}, J[D][D], H[D][D][D];
        fs.geo(e).eval(fs.q1, xi, nullptr, J, H);
        reflen = std::min(reflen, std::pow(determinant<D>(J), 1.0 / D));
    }
    return disc.dt_from_cfl(cfl, reflen) / (2 * std::max(1, fs.p) + 1);
}

enum class RKMode { REFERENCE_COMPAT = 0, TEXTBOOK = 1 };

// RK3SSP::step.  REFERENCE_COMPAT evaluates every stage residual at u^n
// (ssp_rk3.hpp:132-137, form_residual is called with `u`); TEXTBOOK uses the
// stage states.
template <int D, class Disc>
void rk3_ssp_step(const FESpace<D>& fs, const Disc& disc, const MassInverse<D>& minv, double dt, RKMode mode,
                  double* u) {
    constexpr int neq = Disc::neq;
    const size_t N = (size_t)fs.mesh.nelem * fs.nb * neq;
    std::vector<double> res(N), r1(N, 0.0), r2(N, 0.0), r3(N, 0.0), us(N);
    auto stage = [&](const double* ust, std::vector<double>& out) {
        std::fill(out.begin(), out.end(), 0.0);
        form_residual(fs, disc, mode == RKMode::REFERENCE_COMPAT ? u : ust, res.data());
        minv.apply_add(fs, neq, res.data(), out.data());
    };
    stage(u, r1);
    for (size_t k = 0; k < N; ++k) us[k] = u[k] + dt * r1[k];
    stage(us.data(), r2);
    for (size_t k = 0; k < N; ++k) us[k] = u[k] + 0.25 * dt * r1[k] + 0.25 * dt * r2[k];
    stage(us.data(), r3);
    for (size_t k = 0; k < N; ++k) u[k] += 1.0 / 6.0 * dt * r1[k];
    for (size_t k = 0; k < N; ++k) u[k] += 1.0 / 6.0 * dt * r2[k];
    for (size_t k = 0; k < N; ++k) u[k] += 2.0 / 3.0 * dt * r3[k];
}

// RK3TVD::step (tvd_rk3.hpp:107-231): in-place Shu-Osher stages
template <int D, class Disc>
void rk3_tvd_step(const FESpace<D>& fs, const Disc& disc, const MassInverse<D>& minv, double dt, double* u) {
    constexpr int neq = Disc::neq;
    const size_t N = (size_t)fs.mesh.nelem * fs.nb * neq;
    static const double alfa[3] = {0.0, 0.75, 1.0 / 3.0};
    static const double beta[3] = {1.0, 0.25, 2.0 / 3.0};
    std::vector<double> res(N), r1(N), uold(u, u + N);
    for (int s = 0; s < 3; ++s) {
        std::fill(r1.begin(), r1.end(), 0.0);
        form_residual(fs, disc, u, res.data());
        minv.apply_add(fs, neq, res.data(), r1.data());
        for (size_t k = 0; k < N; ++k) u[k] = alfa[s] * uold[k] + beta[s] * u[k];
        for (size_t k = 0; k < N; ++k) u[k] += beta[s] * dt * r1[k];
    }
}

// ExplicitEuler::step (explicit_euler.hpp:87-138)
template <int D, class Disc>
void explicit_euler_step(const FESpace<D>& fs, const Disc& disc, const MassInverse<D>& minv, double dt, double* u) {
    constexpr int neq = Disc::neq;
    const size_t N = (size_t)fs.mesh.nelem * fs.nb * neq;
    std::vector<double> res(N), r1(N, 0.0);
    form_residual(fs, disc, u, res.data());
    minv.apply_add(fs, neq, res.data(), r1.data());
    for (size_t k = 0; k < N; ++k) u[k] += dt * r1[k];
}

// ---------------------------------------------------------------------------
// Block Jacobian by finite differences (form_petsc_jacobian_fd DG part,
// form_petsc_jacobian.hpp:55-317, with scale_fd_epsilon, fd_utils.hpp:10-14)
// and the element block from domain_integral_jacobian (conservation_law.hpp:316-432).
// Blocks are (nb*neq)^2 row-major [i*neq+v][j*neq+w], keyed by (row elem, col elem).

template <int D, class Disc>
void domain_integral_jacobian(const FESpace<D>& fs, const Disc& disc, int e, const double* U, double* jac) {
    constexpr int m = Disc::neq;
    const int nb = fs.nb, bs = nb * m;
    const auto g = fs.geo(e);
    std::vector<double> gx(nb * D);
    const double sqeps = std::sqrt(std::numeric_limits<double>::epsilon());
    for (int q = 0; q < fs.nq; ++q) {
        const double w = fs.quad.w[q];
        double x[D], J[D][D], H[D][D][D];
        g.eval(fs.q1, fs.quad.pts[q].data(), x, J, H);
        double detJ = std::max(0.0, determinant<D>(J));
        const double* bi = &fs.B[q * nb];
        phys_grad<D>(nb, J, &fs.G[q * nb * D], gx.data());
        double u[m] = {}, gu[m * D];
        for (int i = 0; i < nb; ++i)
            for (int v = 0; v < m; ++v) u[v] += U[i * m + v] * bi[i];
        Disc::contract(nb, U, gx.data(), D, gu);
        double F[m * D], Fp[m * D];
        disc.phys_flux(u, gu, F);
        double fro = 0;
        for (int k = 0; k < m * D; ++k) fro += F[k] * F[k];
        const double eps = std::max(sqeps, sqeps * std::sqrt(fro));
        double dFdu[m][D][m], dFdg[m][D][m][D];
        for (int jw = 0; jw < m; ++jw) {
            const double uo = u[jw];
            u[jw] += eps;
            disc.phys_flux(u, gu, Fp);
            for (int ie = 0; ie < m; ++ie)
                for (int id = 0; id < D; ++id) dFdu[ie][id][jw] = (Fp[ie * D + id] - F[ie * D + id]) / eps;
            u[jw] = uo;
            for (int jd = 0; jd < D; ++jd) {
                const double go = gu[jw * D + jd];
                gu[jw * D + jd] += eps;
                disc.phys_flux(u, gu, Fp);
                for (int ie = 0; ie < m; ++ie)
                    for (int id = 0; id < D; ++id) dFdg[ie][id][jw][jd] = (Fp[ie * D + id] - F[ie * D + id]) / eps;
                gu[jw * D + jd] = go;
            }
        }
        for (int it = 0; it < nb; ++it)
            for (int ie = 0; ie < m; ++ie)
                for (int id = 0; id < D; ++id)
                    for (int jd = 0; jd < nb; ++jd)
                        for (int je = 0; je < m; ++je) {
                            double& a = jac[(it * m + ie) * bs + jd * m + je];
                            a += dFdu[ie][id][je] * bi[jd] * gx[it * D + id] * detJ * w;
                            for (int jj = 0; jj < D; ++jj)
                                a += dFdg[ie][id][je][jj] * gx[jd * D + jj] * gx[it * D + id] * detJ * w;
                        }
    }
}

template <int D, class Disc>
void fd_jacobian(const FESpace<D>& fs, const Disc& disc, const double* u, double* res,
                 std::map<std::pair<int, int>, std::vector<double>>& blocks) {
    constexpr int m = Disc::neq;
    const int nb = fs.nb, bs = nb * m;
    const double sqeps = std::sqrt(std::numeric_limits<double>::epsilon());
    auto blk = [&](int r, int c) -> std::vector<double>& {
        auto& b = blocks[{r, c}];
        if (b.empty()) b.assign((size_t)bs * bs, 0.0);
        return b;
    };
    auto nrm2 = [](const std::vector<double>& v) {
        double s = 0;
        for (double x : v) s += x * x;
        return std::sqrt(s);
    };
    std::fill(res, res + (size_t)fs.mesh.nelem * bs, 0.0);
    std::vector<double> uL(bs), uR(bs), rL(bs), rR(bs), rLp(bs), rRp(bs);
    for (const auto& t : fs.boundary) {
        if (fs.coupled_periodic && t.bctype == BC_PERIODIC) continue;
        std::copy(u + (size_t)t.elemL * bs, u + (size_t)(t.elemL + 1) * bs, uL.begin());
        std::copy(u + (size_t)t.elemR * bs, u + (size_t)(t.elemR + 1) * bs, uR.begin());
        std::fill(rL.begin(), rL.end(), 0.0);
        disc.boundary_integral(fs, t, uL.data(), uR.data(), rL.data());
        for (int k = 0; k < bs; ++k) res[(size_t)t.elemL * bs + k] += rL[k];
        const double eps = std::max(sqeps, sqeps * nrm2(rL));
        auto& J = blk(t.elemL, t.elemL);
        for (int c = 0; c < bs; ++c) {
            const double old = uL[c];
            uL[c] += eps;
            std::fill(rLp.begin(), rLp.end(), 0.0);
            disc.boundary_integral(fs, t, uL.data(), uR.data(), rLp.data());
            for (int r = 0; r < bs; ++r) J[(size_t)r * bs + c] += (rLp[r] - rL[r]) / eps;
            uL[c] = old;
        }
    }
    auto interior = [&](const typename FESpace<D>::Trace& t) {
        std::copy(u + (size_t)t.elemL * bs, u + (size_t)(t.elemL + 1) * bs, uL.begin());
        std::copy(u + (size_t)t.elemR * bs, u + (size_t)(t.elemR + 1) * bs, uR.begin());
        std::fill(rL.begin(), rL.end(), 0.0);
        std::fill(rR.begin(), rR.end(), 0.0);
        disc.trace_integral(fs, t, uL.data(), uR.data(), rL.data(), rR.data());
        for (int k = 0; k < bs; ++k) res[(size_t)t.elemL * bs + k] += rL[k];
        for (int k = 0; k < bs; ++k) res[(size_t)t.elemR * bs + k] += rR[k];
        const double eps = std::max(sqeps, sqeps * std::max(nrm2(rL), nrm2(rR)));
        for (int side = 0; side < 2; ++side) {
            std::vector<double>& us = side ? uR : uL;
            const int ec = side ? t.elemR : t.elemL;
            std::vector<double> JL((size_t)bs * bs, 0.0), JR((size_t)bs * bs, 0.0);
            for (int c = 0; c < bs; ++c) {
                const double old = us[c];
                us[c] += eps;
                std::fill(rLp.begin(), rLp.end(), 0.0);
                std::fill(rRp.begin(), rRp.end(), 0.0);
                disc.trace_integral(fs, t, uL.data(), uR.data(), rLp.data(), rRp.data());
                for (int r = 0; r < bs; ++r) {
                    JL[(size_t)r * bs + c] += (rLp[r] - rL[r]) / eps;
                    JR[(size_t)r * bs + c] += (rRp[r] - rR[r]) / eps;
                }
                us[c] = old;
            }
            auto& BL = blk(t.elemL, ec);
            auto& BR = blk(t.elemR, ec);
            for (size_t k = 0; k < JL.size(); ++k) { BL[k] += JL[k]; BR[k] += JR[k]; }
        }
    };
    for (const auto& t : fs.interior) interior(t);
    for (const auto& t : fs.periodic) interior(t);
    for (int e = 0; e < fs.mesh.nelem; ++e) {
        std::copy(u + (size_t)e * bs, u + (size_t)(e + 1) * bs, uL.begin());
        std::fill(rL.begin(), rL.end(), 0.0);
        disc.domain_integral(fs, e, uL.data(), rL.data());
        auto& J = blk(e, e);
        domain_integral_jacobian(fs, disc, e, uL.data(), J.data());
        for (int k = 0; k < bs; ++k) res[(size_t)e * bs + k] += rL[k];
    }
}

// Newton without linesearch (PetscNewton::solve with no_linesearch,
// petsc_newton.hpp:205-329) where the linear solve J du = R is a dense LU of
// the assembled block Jacobian: a direct solve the device GMRES is checked
// against.  Stops when ||R||_2 <= tau_abs + tau_rel ||R(u0)||_2
// (ConvergenceCriteria, nonlinear_solver_utils.hpp:23-60); returns the steps
// taken, or -1 when the dense system is too large.
template <int D, class Disc>
int newton_dense(const FESpace<D>& fs, const Disc& disc, double* u, double tau_abs, double tau_rel, int kmax,
                 double* rhist, int nhist) {
    constexpr int m = Disc::neq;
    const int nb = fs.nb, bs = nb * m, E = fs.mesh.nelem;
    const size_t N = (size_t)E * bs;
    if (N > 6000) return -1;
    std::vector<double> res(N), du(N), A;
    auto nrm2 = [&](const std::vector<double>& v) {
        double s2 = 0;
        for (double x : v) s2 += x * x;
        return std::sqrt(s2);
    };
    std::map<std::pair<int, int>, std::vector<double>> B;
    fd_jacobian(fs, disc, u, res.data(), B);
    const double r0 = nrm2(res);
    int k = 0;
    for (k = 0; k < kmax; ++k) {
        A.assign(N * N, 0.0);
        for (const auto& kv : B) {
            const int re = kv.first.first, ce = kv.first.second;
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j < bs; ++j)
                    A[((size_t)re * bs + i) * N + (size_t)ce * bs + j] += kv.second[(size_t)i * bs + j];
        }
        DenseLU lu;
        lu.factor(A, (int)N);
        if (lu.singular) throw std::runtime_error("singular Jacobian");
        lu.solve(res.data(), du.data());
        for (size_t i = 0; i < N; ++i) u[i] -= du[i];
        B.clear();
        fd_jacobian(fs, disc, u, res.data(), B);
        const double rk = nrm2(res);
        if (k < nhist) rhist[k] = rk;
        if (rk <= tau_abs + tau_rel * r0) {
            ++k;
            break;
        }
    }
    return k;
}

}  // namespace oracle
