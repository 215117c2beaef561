// CPU ORACLE (test infrastructure only; never linked into the product).
//
// FESpace tables, FiniteElement / TraceSpace geometry and the
// ConservationLawDDG integrals, restated from:
//  fespace/fespace.hpp:200-416, element/finite_element.hpp:89-611,
//  element/TraceSpace.hpp:44-585, element/reference_element.hpp:38-288,
//  geometry/face.hpp:279-497, geometry/hypercube_face.hpp:179-274,
//  disc/conservation_law.hpp:92-1176, solvers/form_residual.hpp:56-268.
#pragma once
#include <functional>
#include <limits>
#include <optional>
#include <vector>

#include "physics.hpp"

namespace oracle {

// FiniteElement geometry for a geometry-order-1 hypercube element
// (hypercube::transform/jacobian/hessian, HypercubeTransformations.hpp:52-161).
template <int D>
struct ElementGeo {
    std::array<double, D> X[1 << D];
    void eval(const TensorBasis<D>& q1, const double* xi, double* x, double (&J)[D][D],
              double (&H)[D][D][D]) const {
        constexpr int NV = 1 << D;
        double B[NV], G[NV * D], Hn[NV * D * D];
        q1.eval(xi, B, G, Hn);
        for (int i = 0; i < D; ++i) {
            if (x) x[i] = 0;
            for (int j = 0; j < D; ++j) {
                J[i][j] = 0;
                for (int k = 0; k < D; ++k) H[i][j][k] = 0;
            }
        }
        for (int n = 0; n < NV; ++n)
            for (int i = 0; i < D; ++i) {
                if (x) x[i] += B[n] * X[n][i];
                for (int j = 0; j < D; ++j) {
                    J[i][j] += G[n * D + j] * X[n][i];
                    for (int k = 0; k < D; ++k) H[i][j][k] += Hn[(n * D + j) * D + k] * X[n][i];
                }
            }
    }
};

// eval_phys_grad_basis (finite_element.hpp:220-258): g = gref * adj(J) / det
template <int D>
inline void phys_grad(int nb, const double (&J)[D][D], const double* gref, double* g) {
    double A[D][D];
    adjugate<D>(J, A);
    double det = determinant<D>(J);
    det = (det == 0.0) ? 1.0 : det;
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int k = 0; k < D; ++k) s += gref[i * D + k] * A[k][j];
            g[i * D + j] = s;
        }
    for (int i = 0; i < nb * D; ++i) g[i] /= det;
}

// eval_phys_hess_basis (finite_element.hpp:328-398)
template <int D>
inline void phys_hess(int nb, const double (&J)[D][D], const double (&Hm)[D][D][D], const double* gphys,
                      const double* href, double* h) {
    double A[D][D];
    adjugate<D>(J, A);
    double det = determinant<D>(J);
    det = (det == 0.0) ? 1.0 : det;
    const double det2 = det * det;
    for (int i = 0; i < nb; ++i) {
        double rhs[D][D];
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                rhs[a][b] = href[(i * D + a) * D + b];
                for (int k = 0; k < D; ++k) rhs[a][b] -= Hm[k][a][b] * gphys[i * D + k];
            }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                double s = 0;
                for (int k = 0; k < D; ++k)
                    for (int l = 0; l < D; ++l) s += A[k][b] * A[l][a] * rhs[l][k];
                h[(i * D + a) * D + b] = s / det2;
            }
    }
}

template <int D>
struct FESpace {
    static constexpr int NV = 1 << D;
    static constexpr int NFV = Mesh<D>::NFV;
    static constexpr int TD = (D > 1) ? D - 1 : 0;
    Mesh<D> mesh;
    int p, n, nb, nq, nfq;
    bool coupled_periodic = false;
    TensorBasis<D> basis;
    TensorBasis<D> q1;
    TensorBasis<(D > 1 ? D - 1 : 1)> fq1;
    TensorQuadrature<D> quad;
    TensorQuadrature<TD> fquad;
    std::vector<double> B, G;  // volume tables [nq][nb], [nq][nb][D]

    struct Trace {
        int face = -1;
        int elemL = -1, elemR = -1;
        int fnl = 0, fnr = 0, orient = 0;
        int bctype = BC_INTERIOR, bcflag = 0;
        std::array<int, NFV> nodes{};
        std::array<double, D> shiftR{};  // added to the right centroid (coupled periodic)
        std::vector<double> xiL, xiR;    // [nfq][D]
        std::vector<double> BL, GL, HL, BR, GR, HR;
        // each side's point mirrored through its element along the face
        // normal (the opposite face) and the basis there: the published
        // DDGIC reading's test gradients
        std::vector<double> xiLm, xiRm, GLm, HLm, GRm, HRm;
    };
    std::vector<Trace> interior, boundary, periodic;

    FESpace(Mesh<D> m, int order, bool coupled)
        : mesh(std::move(m)), p(order), n(order + 1), nb(ipow(order + 1, D)), nq(ipow(order + 1, D)),
          nfq(ipow(order + 1, TD)), coupled_periodic(coupled), basis(order), q1(1), fq1(1),
          quad(order + 1), fquad(order + 1) {
        B.resize(nq * nb);
        G.resize(nq * nb * D);
        for (int q = 0; q < nq; ++q) basis.eval(quad.pts[q].data(), &B[q * nb], &G[q * nb * D], nullptr);
        for (int f = 0; f < (int)mesh.faces.size(); ++f) {
            const auto& fc = mesh.faces[f];
            Trace t = make_trace(f, fc.elemL, (f < mesh.n_interior) ? fc.elemR : fc.elemL, fc.fnl, fc.fnr,
                                 fc.orient, fc.nodes);
            t.bctype = fc.bctype;
            t.bcflag = fc.bcflag;
            if (f < mesh.n_interior) interior.push_back(t);
            else boundary.push_back(t);
        }
        if (coupled) {
            // pair periodic boundary faces A (negative side, even offset) with B
            for (int f = mesh.n_interior; f < (int)mesh.faces.size(); ++f) {
                const auto& fa = mesh.faces[f];
                if (fa.bctype != BC_PERIODIC || fa.fnl >= D) continue;
                const auto& fb = mesh.faces[fa.bcflag];
                const int idim = fa.fnl;
                // translate A's vertices to their partners on B's side for the orientation
                std::array<int, NFV> vA{};
                for (int k = 0; k < NFV; ++k) {
                    auto ijk = mesh.node_ijk(fa.nodes[k]);
                    ijk[idim] = mesh.nnode_dir[idim] - 1;
                    vA[k] = mesh.node_index(ijk);
                }
                const int orient = get_orientation<D>(fb.nodes.data(), vA.data());
                Trace t = make_trace(fa.bcflag, fb.elemL, fa.elemL, fb.fnl, fa.fnl, orient, fb.nodes);
                t.bctype = BC_INTERIOR;
                t.shiftR[idim] = mesh.period[idim];
                periodic.push_back(t);
            }
        }
    }

    Trace make_trace(int f, int eL, int eR, int fnl, int fnr, int orient, const std::array<int, NFV>& nodes) {
        Trace t;
        t.face = f; t.elemL = eL; t.elemR = eR; t.fnl = fnl; t.fnr = fnr; t.orient = orient; t.nodes = nodes;
        t.xiL.resize(nfq * D); t.xiR.resize(nfq * D);
        t.BL.resize(nfq * nb); t.GL.resize(nfq * nb * D); t.HL.resize(nfq * nb * D * D);
        t.BR.resize(nfq * nb); t.GR.resize(nfq * nb * D); t.HR.resize(nfq * nb * D * D);
        for (int q = 0; q < nfq; ++q) {
            double s[3] = {0, 0, 0}, sR[3] = {0, 0, 0};
            for (int d = 0; d < TD; ++d) s[d] = fquad.pts[q][d];
            trace_to_ref<D>(fnl, s, &t.xiL[q * D]);
            orient_transform<D>(orient, s, sR);
            trace_to_ref<D>(fnr, sR, &t.xiR[q * D]);
            basis.eval(&t.xiL[q * D], &t.BL[q * nb], &t.GL[q * nb * D], &t.HL[q * nb * D * D]);
            basis.eval(&t.xiR[q * D], &t.BR[q * nb], &t.GR[q * nb * D], &t.HR[q * nb * D * D]);
        }
        t.xiLm = t.xiL; t.xiRm = t.xiR;
        t.GLm.resize(nfq * nb * D); t.HLm.resize(nfq * nb * D * D);
        t.GRm.resize(nfq * nb * D); t.HRm.resize(nfq * nb * D * D);
        std::vector<double> Bm(nb);
        for (int q = 0; q < nfq; ++q) {
            t.xiLm[q * D + fnl % D] = -t.xiL[q * D + fnl % D];
            t.xiRm[q * D + fnr % D] = -t.xiR[q * D + fnr % D];
            basis.eval(&t.xiLm[q * D], Bm.data(), &t.GLm[q * nb * D], &t.HLm[q * nb * D * D]);
            basis.eval(&t.xiRm[q * D], Bm.data(), &t.GRm[q * nb * D], &t.HRm[q * nb * D * D]);
        }
        return t;
    }

    ElementGeo<D> geo(int e) const {
        ElementGeo<D> g;
        for (int k = 0; k < NV; ++k) g.X[k] = mesh.coord[mesh.conn[e][k]];
        return g;
    }
    std::array<double, D> centroid(int e) const {
        double xi[D] = {}, x[D], J[D][D], H[D][D][D];
        geo(e).eval(q1, xi, x, J, H);
        std::array<double, D> c;
        for (int d = 0; d < D; ++d) c[d] = x[d];
        return c;
    }

    // Face geometry: physical point and Jacobian dx/ds from the face nodes
    // (HypercubeTraceTransformation::transform_physical / Jacobian,
    // HypercubeTransformations.hpp:1237-1298; 1D: :1382-1410).
    void face_geo(const Trace& t, int q, double* x, double (&Jf)[D][(D > 1 ? D - 1 : 1)]) const {
        if constexpr (D == 1) {
            x[0] = mesh.coord[t.nodes[0]][0];
            Jf[0][0] = (t.fnl == 0) ? -1.0 : 1.0;
        } else {
            double s[D - 1];
            for (int d = 0; d < D - 1; ++d) s[d] = fquad.pts[q][d];
            double Bf[NFV], Gf[NFV * (D - 1)];
            fq1.eval(s, Bf, Gf, nullptr);
            for (int i = 0; i < D; ++i) {
                x[i] = 0;
                for (int j = 0; j < D - 1; ++j) Jf[i][j] = 0;
            }
            for (int k = 0; k < NFV; ++k) {
                const auto& X = mesh.coord[t.nodes[k]];
                for (int i = 0; i < D; ++i) {
                    x[i] += Bf[k] * X[i];
                    for (int j = 0; j < D - 1; ++j) Jf[i][j] += Gf[k * (D - 1) + j] * X[i];
                }
            }
        }
    }
};

// sqrt(det(J^T J)) (Face::rootRiemannMetric, face.hpp:409-446; 1 for points)
template <int D>
inline double root_riemann(const double (&Jf)[D][(D > 1 ? D - 1 : 1)]) {
    if constexpr (D == 1) return 1.0;
    else if constexpr (D == 2) return std::sqrt(Jf[0][0] * Jf[0][0] + Jf[1][0] * Jf[1][0]);
    else {
        double g[2][2] = {};
        for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l)
                for (int i = 0; i < 3; ++i) g[k][l] += Jf[i][k] * Jf[i][l];
        return std::sqrt(g[0][0] * g[1][1] - g[0][1] * g[1][0]);
    }
}

// calc_ortho + normalize: normal = eps . J (2D (J10, -J00); 3D col0 x col1);
// orientation pinned by the reference-cube normals of
// TestHypercubeTransform.cpp:320-384 and validate_normals (mesh_utils.hpp:200-242).
template <int D>
inline void unit_normal(const double (&Jf)[D][(D > 1 ? D - 1 : 1)], double* nrm) {
    if constexpr (D == 1) {
        nrm[0] = Jf[0][0];
    } else if constexpr (D == 2) {
        nrm[0] = Jf[1][0];
        nrm[1] = -Jf[0][0];
    } else {
        nrm[0] = Jf[1][0] * Jf[2][1] - Jf[2][0] * Jf[1][1];
        nrm[1] = Jf[2][0] * Jf[0][1] - Jf[0][0] * Jf[2][1];
        nrm[2] = Jf[0][0] * Jf[1][1] - Jf[1][0] * Jf[0][1];
    }
    double s = 0;
    for (int d = 0; d < D; ++d) s += nrm[d] * nrm[d];
    s = std::sqrt(s);
    for (int d = 0; d < D; ++d) nrm[d] /= s;
}

using Callback = std::function<void(const double*, double*)>;

template <int D, class PF, class CF, class DF>
struct ConservationLawDDG {
    static constexpr int neq = PF::neq;
    using PFlux = PF;
    PF phys_flux;
    CF conv_nflux;
    DF diff_flux;
    bool interior_penalty = false;
    double sigma_ic = 0.0;
    // reading of the interface correction: false = conservation_law.hpp as
    // written (sigma/2 per side, Dirichlet term, each side's test gradient at
    // the face point); true = sigma per side, no boundary term, and each
    // side's test gradient averaged between the face point and its mirror
    // through the element along the face normal, which reproduces the
    // published DDGIC table (doc/sine_diffusion_1d_study.md:45-55; in 1D the
    // mean of the test function's derivative at the element's two ends)
    bool ddgic_published = false;
    std::vector<Callback> dirichlet_callbacks, neumann_callbacks;
    std::optional<Callback> user_source;

    double dt_from_cfl(double cfl, double reflen) { return phys_flux.dt_from_cfl(cfl, reflen); }

    // conservation_law.hpp:244-314
    void domain_integral(const FESpace<D>& fs, int e, const double* U, double* R) const {
        const auto g = fs.geo(e);
        const int nb = fs.nb;
        std::vector<double> gx(nb * D);
        double u[neq], gradu[neq * D], F[neq * D];
        for (int q = 0; q < fs.nq; ++q) {
            const double* xi = fs.quad.pts[q].data();
            const double w = fs.quad.w[q];
            double x[D], J[D][D], H[D][D][D];
            g.eval(fs.q1, xi, x, J, H);
            double detJ = determinant<D>(J);
            detJ = std::max(0.0, detJ);
            const double* bi = &fs.B[q * nb];
            phys_grad<D>(nb, J, &fs.G[q * nb * D], gx.data());
            for (int v = 0; v < neq; ++v) u[v] = 0;
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v) u[v] += U[i * neq + v] * bi[i];
            for (int v = 0; v < neq; ++v)
                for (int d = 0; d < D; ++d) {
                    double s = 0;
                    for (int i = 0; i < nb; ++i) s += U[i * neq + v] * gx[i * D + d];
                    gradu[v * D + d] = s;
                }
            phys_flux(u, gradu, F);
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v)
                    for (int d = 0; d < D; ++d) R[i * neq + v] += F[v * D + d] * gx[i * D + d] * detJ * w;
            if (user_source) {
                double src[neq];
                (*user_source)(x, src);
                for (int i = 0; i < nb; ++i)
                    for (int v = 0; v < neq; ++v) R[i * neq + v] -= src[v] * bi[i] * detJ * w;
            }
        }
    }

    // PhysDomainEval (finite_element.hpp:497-564) for one side of a trace
    void side_eval(const FESpace<D>& fs, int e, const double* xi, const double* gref, const double* href,
                   double* gx, double* hx) const {
        double J[D][D], H[D][D][D];
        fs.geo(e).eval(fs.q1, xi, nullptr, J, H);
        phys_grad<D>(fs.nb, J, gref, gx);
        phys_hess<D>(fs.nb, J, H, gx, href, hx);
    }

    // conservation_law.hpp:434-612 (also used for coupled periodic pairs)
    void trace_integral(const FESpace<D>& fs, const typename FESpace<D>::Trace& t, const double* UL,
                        const double* UR, double* RL, double* RR) const {
        const int nb = fs.nb;
        auto cL = fs.centroid(t.elemL);
        auto cR = fs.centroid(t.elemR);
        for (int d = 0; d < D; ++d) cR[d] += t.shiftR[d];
        std::vector<double> gL(nb * D), gR(nb * D), hL(nb * D * D), hR(nb * D * D);
        double uL[neq], uR[neq], guL[neq * D], guR[neq * D], huL[neq * D * D], huR[neq * D * D];
        for (int q = 0; q < fs.nfq; ++q) {
            const double w = fs.fquad.w[q];
            double x[D], Jf[D][(D > 1 ? D - 1 : 1)], nrm[D];
            fs.face_geo(t, q, x, Jf);
            const double sqrtg = root_riemann<D>(Jf);
            unit_normal<D>(Jf, nrm);
            const double* biL = &t.BL[q * nb];
            const double* biR = &t.BR[q * nb];
            side_eval(fs, t.elemL, &t.xiL[q * D], &t.GL[q * nb * D], &t.HL[q * nb * D * D], gL.data(), hL.data());
            side_eval(fs, t.elemR, &t.xiR[q * D], &t.GR[q * nb * D], &t.HR[q * nb * D * D], gR.data(), hR.data());
            for (int v = 0; v < neq; ++v) {
                uL[v] = 0; uR[v] = 0;
                for (int i = 0; i < nb; ++i) uL[v] += UL[i * neq + v] * biL[i];
                for (int i = 0; i < nb; ++i) uR[v] += UR[i * neq + v] * biR[i];
            }
            contract(nb, UL, gL.data(), D, guL);
            contract(nb, UR, gR.data(), D, guR);
            contract(nb, UL, hL.data(), D * D, huL);
            contract(nb, UR, hR.data(), D * D, huR);
            double fadv[neq], fvis[neq], gddg[neq * D], uavg[neq];
            conv_nflux(uL, uR, nrm, fadv);
            double h = 0;
            for (int d = 0; d < D; ++d) h += nrm[d] * ((x[d] - cL[d]) + (cR[d] - x[d]));
            h = std::copysign(std::max(std::abs(h), std::numeric_limits<double>::epsilon()), h);
            const int order = fs.p;
            const double beta0 = std::pow(order + 1, 2);
            double beta1 = 1 / std::max((double)(2 * order * (order + 1)), 1.0);
            if (interior_penalty) beta1 = 0.0;
            for (int v = 0; v < neq; ++v) {
                const double jump = uR[v] - uL[v];
                for (int d = 0; d < D; ++d) {
                    gddg[v * D + d] = beta0 * jump / h * nrm[d] + 0.5 * (guL[v * D + d] + guR[v * D + d]);
                    double ht = 0;
                    for (int j = 0; j < D; ++j) ht += (huR[(v * D + j) * D + d] - huL[(v * D + j) * D + d]) * nrm[j];
                    gddg[v * D + d] += beta1 * h * ht;
                }
            }
            for (int v = 0; v < neq; ++v) uavg[v] = 0.5 * (uL[v] + uR[v]);
            diff_flux(uavg, gddg, nrm, fvis);
            for (int v = 0; v < neq; ++v) { fadv[v] *= w * sqrtg; fvis[v] *= w * sqrtg; }
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v) RL[i * neq + v] += (fvis[v] - fadv[v]) * biL[i];
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v) RR[i * neq + v] -= (fvis[v] - fadv[v]) * biR[i];
            if constexpr (DF::has_homogeneity) {
                if (sigma_ic != 0.0) {
                    // SOURCE: the 0.5 of the average operator (:600-603);
                    // PUBLISHED: sigma_ic on each side
                    const double ic_side = ddgic_published ? 1.0 : 0.5;
                    double Gt[neq * D * neq * D];
                    diff_flux.homogeneity_tensor(uavg, Gt);
                    if (ddgic_published) {
                        std::vector<double> gm(nb * D), hm(nb * D * D);
                        side_eval(fs, t.elemL, &t.xiLm[q * D], &t.GLm[q * nb * D], &t.HLm[q * nb * D * D], gm.data(), hm.data());
                        for (int i = 0; i < nb * D; ++i) gL[i] = 0.5 * (gL[i] + gm[i]);
                        side_eval(fs, t.elemR, &t.xiRm[q * D], &t.GRm[q * nb * D], &t.HRm[q * nb * D * D], gm.data(), hm.data());
                        for (int i = 0; i < nb * D; ++i) gR[i] = 0.5 * (gR[i] + gm[i]);
                    }
                    for (int it = 0; it < nb; ++it)
                        for (int ie = 0; ie < neq; ++ie)
                            for (int k = 0; k < D; ++k)
                                for (int r = 0; r < neq; ++r) {
                                    const double jr = uR[r] - uL[r];
                                    for (int s = 0; s < D; ++s) {
                                        const double ic = sigma_ic * Gt[((ie * D + k) * neq + r) * D + s] * nrm[k] * jr * w * sqrtg;
                                        RL[it * neq + ie] -= ic * ic_side * gL[it * D + s];
                                        RR[it * neq + ie] -= ic * ic_side * gR[it * D + s];
                                    }
                                }
                }
            }
        }
    }

    static void contract(int nb, const double* U, const double* g, int K, double* out) {
        for (int v = 0; v < neq; ++v)
            for (int k = 0; k < K; ++k) {
                double s = 0;
                for (int i = 0; i < nb; ++i) s += U[i * neq + v] * g[i * K + k];
                out[v * K + k] = s;
            }
    }

    // conservation_law.hpp:631-1176 (DIRICHLET, NEUMANN, EXTRAPOLATION,
    // SPACETIME_FUTURE, default/apply_bc)
    void boundary_integral(const FESpace<D>& fs, const typename FESpace<D>::Trace& t, const double* UL,
                           const double* /*UR*/, double* RL) const {
        const int nb = fs.nb;
        const auto cL = fs.centroid(t.elemL);
        std::vector<double> gL(nb * D);
        double uL[neq], guL[neq * D];
        for (int q = 0; q < fs.nfq; ++q) {
            const double w = fs.fquad.w[q];
            double x[D], Jf[D][(D > 1 ? D - 1 : 1)], nrm[D];
            fs.face_geo(t, q, x, Jf);
            const double sqrtg = root_riemann<D>(Jf);
            unit_normal<D>(Jf, nrm);
            const double* biL = &t.BL[q * nb];
            if (t.bctype == BC_NEUMANN) {
                double nv[neq], fv[neq];
                neumann_callbacks.at(t.bcflag)(x, nv);
                diff_flux.neumann_flux(nv, fv);
                for (int v = 0; v < neq; ++v) fv[v] *= w * sqrtg;
                for (int i = 0; i < nb; ++i)
                    for (int v = 0; v < neq; ++v) RL[i * neq + v] += fv[v] * biL[i];
                continue;
            }
            const bool general = !(t.bctype == BC_DIRICHLET || t.bctype == BC_EXTRAPOLATION ||
                                   t.bctype == BC_SPACETIME_FUTURE);
            if (general && !PF::implements_bcs) return;  // default branch without apply_bc: nothing
            // eval_phys_grad_basis_l_qp recomputes J at xiL
            {
                double J[D][D], H[D][D][D];
                fs.geo(t.elemL).eval(fs.q1, &t.xiL[q * D], nullptr, J, H);
                phys_grad<D>(nb, J, &t.GL[q * nb * D], gL.data());
            }
            for (int v = 0; v < neq; ++v) {
                uL[v] = 0;
                for (int i = 0; i < nb; ++i) uL[v] += UL[i * neq + v] * biL[i];
            }
            contract(nb, UL, gL.data(), D, guL);
            double uB[neq], guB[neq * D], fadv[neq], fvis[neq], gddg[neq * D], uavg[neq];
            if (t.bctype == BC_EXTRAPOLATION || t.bctype == BC_SPACETIME_FUTURE) {
                conv_nflux(uL, uL, nrm, fadv);
                for (int k = 0; k < neq * D; ++k) gddg[k] = guL[k];
                for (int v = 0; v < neq; ++v) uavg[v] = uL[v];
                diff_flux(uavg, gddg, nrm, fvis);
            } else {
                const bool dir = t.bctype == BC_DIRICHLET;
                if (dir) {
                    for (int v = 0; v < neq; ++v) uB[v] = 0;
                    dirichlet_callbacks.at(t.bcflag)(x, uB);
                } else {
                    phys_flux.apply_bc(uL, guL, nrm, t.bctype, t.bcflag, uB, guB);
                }
                double h = 0;
                for (int d = 0; d < D; ++d) h += std::abs(nrm[d] * (x[d] - cL[d]));
                h = std::copysign(std::max(std::abs(h), std::numeric_limits<double>::epsilon()), h);
                const double beta0 = std::pow(fs.p + 1, 2);
                if (dir) conv_nflux(uL, uB, nrm, fadv);
                for (int v = 0; v < neq; ++v) {
                    const double jump = uB[v] - uL[v];
                    for (int d = 0; d < D; ++d) gddg[v * D + d] = beta0 * jump / h * nrm[d] + guL[v * D + d];
                }
                if (!dir) conv_nflux(uL, uB, nrm, fadv);
                for (int v = 0; v < neq; ++v) uavg[v] = 0.5 * (uL[v] + uB[v]);
                diff_flux(uavg, gddg, nrm, fvis);
            }
            for (int v = 0; v < neq; ++v) { fadv[v] *= w * sqrtg; fvis[v] *= w * sqrtg; }
            for (int i = 0; i < nb; ++i)
                for (int v = 0; v < neq; ++v) RL[i * neq + v] += (fvis[v] - fadv[v]) * biL[i];
            if constexpr (DF::has_homogeneity) {
                if (t.bctype == BC_DIRICHLET && sigma_ic != 0.0 && !ddgic_published) {
                    double Gt[neq * D * neq * D];
                    diff_flux.homogeneity_tensor(uavg, Gt);
                    for (int it = 0; it < nb; ++it)
                        for (int ie = 0; ie < neq; ++ie)
                            for (int k = 0; k < D; ++k)
                                for (int r = 0; r < neq; ++r) {
                                    const double jr = uB[r] - uL[r];
                                    for (int s = 0; s < D; ++s)
                                        RL[it * neq + ie] -= sigma_ic * Gt[((ie * D + k) * neq + r) * D + s] * nrm[k] *
                                                             gL[it * D + s] * jr * w * sqrtg;
                                }
                }
            }
        }
    }
};

// form_residual (form_residual.hpp:56-268): boundary traces, interior traces,
// elements, in that order; coupled periodic pairs follow the interior traces.
template <int D, class Disc>
void form_residual(const FESpace<D>& fs, const Disc& disc, const double* u, double* res) {
    constexpr int neq = Disc::neq;
    const int nb = fs.nb;
    const int esz = nb * neq;
    std::fill(res, res + (size_t)fs.mesh.nelem * esz, 0.0);
    std::vector<double> rL(esz), rR(esz);
    for (const auto& t : fs.boundary) {
        if (fs.coupled_periodic && t.bctype == BC_PERIODIC) continue;
        if (!fs.coupled_periodic && t.bctype == BC_PERIODIC && Disc::PFlux::implements_bcs)
            throw std::runtime_error("PERIODIC boundary with a flux that implements apply_bc is undefined in the reference");
        std::fill(rL.begin(), rL.end(), 0.0);
        disc.boundary_integral(fs, t, u + (size_t)t.elemL * esz, u + (size_t)t.elemR * esz, rL.data());
        for (int k = 0; k < esz; ++k) res[(size_t)t.elemL * esz + k] += rL[k];
    }
    auto do_trace = [&](const typename FESpace<D>::Trace& t) {
        std::fill(rL.begin(), rL.end(), 0.0);
        std::fill(rR.begin(), rR.end(), 0.0);
        disc.trace_integral(fs, t, u + (size_t)t.elemL * esz, u + (size_t)t.elemR * esz, rL.data(), rR.data());
        for (int k = 0; k < esz; ++k) res[(size_t)t.elemL * esz + k] += rL[k];
        for (int k = 0; k < esz; ++k) res[(size_t)t.elemR * esz + k] += rR[k];
    };
    for (const auto& t : fs.interior) do_trace(t);
    for (const auto& t : fs.periodic) do_trace(t);
    for (int e = 0; e < fs.mesh.nelem; ++e) {
        std::fill(rL.begin(), rL.end(), 0.0);
        disc.domain_integral(fs, e, u + (size_t)e * esz, rL.data());
        for (int k = 0; k < esz; ++k) res[(size_t)e * esz + k] += rL[k];
    }
}

}  // namespace oracle
