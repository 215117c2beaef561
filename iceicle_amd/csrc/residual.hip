// DG residual kernels for CDNA4 (gfx950), FP64.
//
// One residual = two launches on the caller's stream:
//  1. face kernels (interior + boundary): one thread per face quadrature
//     point; evaluate both traces (value, physical gradient, normal Hessian
//     term), the numerical fluxes and the DDG single-valued gradient, and
//     write the weighted normal flux Q[face][fq][v] (and the DDGIC lifting
//     vector when sigma_ic != 0).  This is ConservationLawDDG::trace_integral /
//     boundaryIntegral (conservation_law.hpp:434-1176) up to the final
//     multiplication by the trace basis.
//  2. element kernel: a workgroup owns a tile of elements whose coefficients
//     are staged in LDS with one coalesced copy; phase 1 has one thread per
//     (element, volume quadrature point) and evaluates the physical flux
//     contracted into the reference frame; phase 2 has one thread per
//     (element, basis function) and sums the volume term
//     (domain_integral, :244-314), the source term and the lifted face fluxes
//     of the element's 2*ndim faces, then writes R with one coalesced store.
// The face-flux buffer replaces the reference's scatter of per-face residual
// blocks (form_residual.hpp:137-242): every element gathers its own faces, so
// the result is deterministic and needs no atomics.
#include <hip/hip_runtime.h>

#include "fields.h"
#include "kernels.h"
#include "physics_dev.h"

namespace ice {

template <int B, int E>
struct Pow { static constexpr int value = B * Pow<B, E - 1>::value; };
template <int B>
struct Pow<B, 0> { static constexpr int value = 1; };

// ---- geometry helpers (geometry order 1 hypercubes) ----------------------

// Q1 map at xi: point, Jacobian J[i][a] = dx_i/dxi_a and Hessian H[i][a][b]
// (hypercube::transform/jacobian/hessian, HypercubeTransformations.hpp:52-161).
template <int D>
__device__ __forceinline__ void q1_map(const double (&X)[1 << D][D], const double* xi, double* x, double (&J)[D][D],
                                       double (&H)[D][D][D]) {
    double N[D][2], dN[D][2];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        N[d][0] = 0.5 * (1 - xi[d]);
        N[d][1] = 1.0 - N[d][0];
        dN[d][0] = -0.5;
        dN[d][1] = 0.5;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        if (x) x[i] = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            J[i][a] = 0;
#pragma unroll
            for (int b = 0; b < D; ++b) H[i][a][b] = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) {
        int bit[D];
#pragma unroll
        for (int d = 0; d < D; ++d) bit[d] = (k >> (D - 1 - d)) & 1;
        double phi = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) phi *= N[d][bit[d]];
        double g[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double t = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) t *= (d == a) ? dN[d][bit[d]] : N[d][bit[d]];
            g[a] = t;
        }
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (x) x[i] += phi * X[k][i];
#pragma unroll
            for (int a = 0; a < D; ++a) J[i][a] += g[a] * X[k][i];
        }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                if (a == b) continue;
                double t = 1.0;
#pragma unroll
                for (int d = 0; d < D; ++d) t *= (d == a || d == b) ? dN[d][bit[d]] : N[d][bit[d]];
#pragma unroll
                for (int i = 0; i < D; ++i) H[i][a][b] += t * X[k][i];
            }
    }
}

template <int D>
__device__ __forceinline__ double det_adj(const double (&J)[D][D], double (&A)[D][D]) {
    if constexpr (D == 1) {
        A[0][0] = 1.0;
        return J[0][0];
    } else if constexpr (D == 2) {
        A[0][0] = J[1][1]; A[0][1] = -J[0][1];
        A[1][0] = -J[1][0]; A[1][1] = J[0][0];
        return J[0][0] * J[1][1] - J[0][1] * J[1][0];
    } else {
        A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
        A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
        A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
        A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
}

template <int D>
__device__ __forceinline__ void load_vertices(const double* __restrict__ Xall, int e, double (&X)[1 << D][D]) {
    const double* p = Xall + (size_t)e * (1 << D) * D;
#pragma unroll
    for (int k = 0; k < (1 << D); ++k)
#pragma unroll
        for (int d = 0; d < D; ++d) X[k][d] = p[k * D + d];
}

// Face map from the left element's face vertices at face quadrature point
// fq: physical point x, unit normal and sqrt(det(J^T J)).
// (HypercubeFace::transform/Jacobian + calc_ortho/normalize/rootRiemannMetric;
//  1D: PointFace conventions, HypercubeTransformations.hpp:1358-1456.)
template <int D, int N>
__device__ __forceinline__ double face_geometry(const DevSpace& S, const double (&X)[1 << D][D], int fnl, int fq,
                                                double* x, double* nrm) {
    if constexpr (D == 1) {
        const int v = S.fvo[fnl * 4 + 0];
        x[0] = X[v][0];
        nrm[0] = (fnl == 0) ? -1.0 : 1.0;
        return 1.0;
    } else {
        constexpr int TD = D - 1;
        constexpr int NFV = 1 << TD;
        const double* xs = S.tab;
        double s[TD];
        {
            int r = fq;
#pragma unroll
            for (int k = TD - 1; k >= 0; --k) { s[k] = xs[r % N]; r /= N; }
        }
        double Jf[D][TD];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            x[i] = 0;
#pragma unroll
            for (int t = 0; t < TD; ++t) Jf[i][t] = 0;
        }
#pragma unroll
        for (int k = 0; k < NFV; ++k) {
            const int v = S.fvo[fnl * 4 + k];
            double psi = 1.0, dpsi[TD];
#pragma unroll
            for (int t = 0; t < TD; ++t) {
                const int bit = (k >> (TD - 1 - t)) & 1;
                const double n0 = 0.5 * (1 - s[t]);
                psi *= bit ? (1.0 - n0) : n0;
            }
#pragma unroll
            for (int t = 0; t < TD; ++t) {
                double g = 1.0;
#pragma unroll
                for (int u = 0; u < TD; ++u) {
                    const int bit = (k >> (TD - 1 - u)) & 1;
                    const double n0 = 0.5 * (1 - s[u]);
                    g *= (u == t) ? (bit ? 0.5 : -0.5) : (bit ? (1.0 - n0) : n0);
                }
                dpsi[t] = g;
            }
#pragma unroll
            for (int i = 0; i < D; ++i) {
                x[i] += psi * X[v][i];
#pragma unroll
                for (int t = 0; t < TD; ++t) Jf[i][t] += dpsi[t] * X[v][i];
            }
        }
        double sg;
        if constexpr (D == 2) {
            nrm[0] = Jf[1][0];
            nrm[1] = -Jf[0][0];
            sg = sqrt(Jf[0][0] * Jf[0][0] + Jf[1][0] * Jf[1][0]);
        } else {
            nrm[0] = Jf[1][0] * Jf[2][1] - Jf[2][0] * Jf[1][1];
            nrm[1] = Jf[2][0] * Jf[0][1] - Jf[0][0] * Jf[2][1];
            nrm[2] = Jf[0][0] * Jf[1][1] - Jf[1][0] * Jf[0][1];
            double g00 = 0, g01 = 0, g11 = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                g00 += Jf[i][0] * Jf[i][0];
                g01 += Jf[i][0] * Jf[i][1];
                g11 += Jf[i][1] * Jf[i][1];
            }
            sg = sqrt(g00 * g11 - g01 * g01);
        }
        double nn = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) nn += nrm[d] * nrm[d];
        nn = sqrt(nn);
#pragma unroll
        for (int d = 0; d < D; ++d) nrm[d] /= nn;
        return sg;
    }
}

// Trace of one element at a tensor point of S (indices pt[D]): value u[M],
// physical gradient gx[M][D] and, if HESS, the normal Hessian term
// hn[M][d] = sum_j (d2u/dx_j dx_d) n_j of the physical Hessian
// (PhysDomainEval + eval_phys_hess_basis, finite_element.hpp:328-398,497-564).
template <int D, int N, int M, bool HESS>
__device__ __forceinline__ void trace_eval(const DevSpace& S, const double* __restrict__ Ue, const int* pt,
                                           const double (&X)[1 << D][D], const double* nrm, double* u, double* gx,
                                           double* hn) {
    constexpr int NH = D * (D + 1) / 2;
    const double* xs = S.tab;
    const double* L0 = S.tab + (N + 2);
    const double* L1 = L0 + (N + 2) * N;
    const double* L2 = L1 + (N + 2) * N;
    double a0[D][N], a1[D][N], a2[D][N];
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xi[d] = xs[pt[d]];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            a0[d][i] = L0[pt[d] * N + i];
            a1[d][i] = L1[pt[d] * N + i];
            a2[d][i] = HESS ? L2[pt[d] * N + i] : 0.0;
        }
    }
    double gr[M][D], hr[M][NH];
#pragma unroll
    for (int v = 0; v < M; ++v) {
        u[v] = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) gr[v][d] = 0;
#pragma unroll
        for (int h = 0; h < NH; ++h) hr[v][h] = 0;
    }
    constexpr int NB = Pow<N, D>::value;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        int id[D];
        {
            int r = i;
#pragma unroll
            for (int d = D - 1; d >= 0; --d) { id[d] = r % N; r /= N; }
        }
        double c0 = 1.0, cg[D], ch[NH];
#pragma unroll
        for (int d = 0; d < D; ++d) c0 *= a0[d][id[d]];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double t = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) t *= (d == k) ? a1[d][id[d]] : a0[d][id[d]];
            cg[k] = t;
        }
        if constexpr (HESS) {
            int h = 0;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = a; b < D; ++b) {
                    double t = 1.0;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        if (a == b) t *= (d == a) ? a2[d][id[d]] : a0[d][id[d]];
                        else t *= (d == a || d == b) ? a1[d][id[d]] : a0[d][id[d]];
                    }
                    ch[h++] = t;
                }
        }
#pragma unroll
        for (int v = 0; v < M; ++v) {
            const double c = Ue[i * M + v];
            u[v] += c * c0;
#pragma unroll
            for (int k = 0; k < D; ++k) gr[v][k] += c * cg[k];
            if constexpr (HESS) {
#pragma unroll
                for (int h = 0; h < NH; ++h) hr[v][h] += c * ch[h];
            }
        }
    }
    double J[D][D], Hm[D][D][D], A[D][D];
    q1_map<D>(X, xi, nullptr, J, Hm);
    double det = det_adj<D>(J, A);
    det = (det == 0.0) ? 1.0 : det;
    const double idet = 1.0 / det;
#pragma unroll
    for (int v = 0; v < M; ++v)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) s += gr[v][k] * A[k][j];
            gx[v * D + j] = s * idet;
        }
    if constexpr (HESS) {
        const double idet2 = idet * idet;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            double rhs[D][D];
            {
                int h = 0;
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = a; b < D; ++b) {
                        rhs[a][b] = hr[v][h];
                        rhs[b][a] = hr[v][h];
                        ++h;
                    }
            }
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b)
#pragma unroll
                    for (int k = 0; k < D; ++k) rhs[a][b] -= Hm[k][a][b] * gx[v * D + k];
            // Hx[a][b] = sum_{k,l} A[k][b] A[l][a] rhs[l][k] / det^2 ; hn[d] = sum_j Hx[j][d] n_j
            double T[D][D];  // T[l][b] = sum_k rhs[l][k] A[k][b]
#pragma unroll
            for (int l = 0; l < D; ++l)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    double s = 0;
#pragma unroll
                    for (int k = 0; k < D; ++k) s += rhs[l][k] * A[k][b];
                    T[l][b] = s;
                }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double s = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    double hx = 0;
#pragma unroll
                    for (int l = 0; l < D; ++l) hx += A[l][j] * T[l][d];
                    s += hx * nrm[j];
                }
                hn[v * D + d] = s * idet2;
            }
        }
    }
}

__device__ __forceinline__ int info_fnl(int info) { return info & 7; }
__device__ __forceinline__ int info_fnr(int info) { return (info >> 3) & 7; }
__device__ __forceinline__ int info_orient(int info) { return (info >> 6) & 7; }
__device__ __forceinline__ int info_bctype(int info) { return (info >> 9) & 15; }
__device__ __forceinline__ int info_pdim(int info) { return (info >> 13) & 3; }

template <int D>
__device__ __forceinline__ void centroid(const double (&X)[1 << D][D], double* c) {
    const double w = 1.0 / (1 << D);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < (1 << D); ++k) s += w * X[k][d];
        c[d] = s;
    }
}

template <int N, int D>
__device__ __forceinline__ double face_weight(const DevSpace& S, int fq) {
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    if constexpr (D == 1) {
        return 1.0;
    } else {
        double w = 1.0;
        int idx[D - 1];
        int r = fq;
#pragma unroll
        for (int k = D - 2; k >= 0; --k) { idx[k] = r % N; r /= N; }
#pragma unroll
        for (int k = 0; k < D - 1; ++k) w = (k == 0) ? wgl[idx[k]] : w * wgl[idx[k]];
        return w;
    }
}

// ---- interior faces ------------------------------------------------------
template <int D, int P, class PH>
__device__ __forceinline__ void face_interior(const DevSpace& S, const PhysParams& ph, const double* __restrict__ U,
                                              double* __restrict__ Q, double* __restrict__ Pic, int tid) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    const int f = tid / NFQ;
    const int q = tid - f * NFQ;
    if (f >= S.nfi) return;
    const int4 rec = S.fint[f];
    const int eL = rec.x, eR = rec.y, info = rec.z;
    const int fnl = info_fnl(info), fnr = info_fnr(info), orient = info_orient(info), pdim = info_pdim(info);
    double XL[1 << D][D], XR[1 << D][D];
    load_vertices<D>(S.X, eL, XL);
    load_vertices<D>(S.X, eR, XR);
    double x[D], nrm[D];
    const double sqrtg = face_geometry<D, N>(S, XL, fnl, q, x, nrm);
    const double w = face_weight<N, D>(S, q);
    double cL[D], cR[D];
    centroid<D>(XL, cL);
    centroid<D>(XR, cR);
    if (pdim) cR[pdim - 1] += S.period[pdim - 1];
    int ptL[D], ptR[D];
    const signed char* fpL = S.fpt + ((fnl * S.norient + 0) * NFQ + q) * D;
    const signed char* fpR = S.fpt + ((fnr * S.norient + orient) * NFQ + q) * D;
#pragma unroll
    for (int d = 0; d < D; ++d) { ptL[d] = fpL[d]; ptR[d] = fpR[d]; }
    double uL[M], uR[M], gL[M * D], gR[M * D], hL[M * D], hR[M * D];
    trace_eval<D, N, M, true>(S, U + (size_t)eL * NB * M, ptL, XL, nrm, uL, gL, hL);
    trace_eval<D, N, M, true>(S, U + (size_t)eR * NB * M, ptR, XR, nrm, uR, gR, hR);
    double fadv[M], fvis[M], gddg[M * D], uavg[M];
    PH::conv(ph, uL, uR, nrm, fadv);
    double h = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) h += nrm[d] * ((x[d] - cL[d]) + (cR[d] - x[d]));
    h = copysign(fmax(fabs(h), 2.220446049250313e-16), h);
#pragma unroll
    for (int v = 0; v < M; ++v) {
        const double jump = uR[v] - uL[v];
#pragma unroll
        for (int d = 0; d < D; ++d)
            gddg[v * D + d] = ph.beta0 * jump / h * nrm[d] + 0.5 * (gL[v * D + d] + gR[v * D + d]) +
                              ph.beta1 * h * (hR[v * D + d] - hL[v * D + d]);
        uavg[v] = 0.5 * (uL[v] + uR[v]);
    }
    PH::diff(ph, uavg, gddg, nrm, fvis);
    const double ws = w * sqrtg;
    double* Qf = Q + ((size_t)f * NFQ + q) * M;
#pragma unroll
    for (int v = 0; v < M; ++v) Qf[v] = fvis[v] * ws - fadv[v] * ws;
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (Pic) {
            double G[M * D * M * D];
            PH::homogeneity(ph, uavg, G);
            double* Pf = Pic + ((size_t)f * NFQ + q) * M * D;
#pragma unroll
            for (int v = 0; v < M; ++v)
#pragma unroll
                for (int s = 0; s < D; ++s) {
                    double acc = 0;
#pragma unroll
                    for (int k = 0; k < D; ++k)
#pragma unroll
                        for (int r = 0; r < M; ++r) acc += G[((v * D + k) * M + r) * D + s] * nrm[k] * (uR[r] - uL[r]);
                    Pf[v * D + s] = 0.5 * ph.sigma_ic * acc * ws;
                }
        }
    }
}

// ---- boundary faces ------------------------------------------------------
template <int D, int P, class PH>
__device__ __forceinline__ void face_boundary(const DevSpace& S, const PhysParams& ph, const double* __restrict__ bval,
                                              const double* __restrict__ U, double* __restrict__ Q,
                                              double* __restrict__ Pic, int tid) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    const int b = tid / NFQ;
    const int q = tid - b * NFQ;
    if (b >= S.nfb) return;
    const int4 rec = S.fbdy[b];
    const int eL = rec.x, info = rec.y, bcflag = rec.z;
    const int fnl = info_fnl(info), bctype = info_bctype(info);
    const size_t slot = (size_t)S.nfi + b;
    double* Qf = Q + (slot * NFQ + q) * M;
    double* Pf = nullptr;
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (Pic) {
            Pf = Pic + (slot * NFQ + q) * M * D;
#pragma unroll
            for (int k = 0; k < M * D; ++k) Pf[k] = 0.0;
        }
    }
    const bool general = !(bctype == ICE_BC_DIRICHLET || bctype == ICE_BC_NEUMANN || bctype == ICE_BC_EXTRAPOLATION ||
                           bctype == ICE_BC_SPACETIME_FUTURE);
    if (general && !PH::IMPLEMENTS_BCS) {  // boundaryIntegral default branch without apply_bc
#pragma unroll
        for (int v = 0; v < M; ++v) Qf[v] = 0.0;
        return;
    }
    double XL[1 << D][D];
    load_vertices<D>(S.X, eL, XL);
    double x[D], nrm[D];
    const double sqrtg = face_geometry<D, N>(S, XL, fnl, q, x, nrm);
    const double ws = face_weight<N, D>(S, q) * sqrtg;
    const double* bv = bval + ((size_t)b * NFQ + q) * M;
    if (bctype == ICE_BC_NEUMANN) {
        double g[M], fv[M];
#pragma unroll
        for (int v = 0; v < M; ++v) g[v] = bv[v];
        PH::neumann(ph, g, fv);
#pragma unroll
        for (int v = 0; v < M; ++v) Qf[v] = fv[v] * ws;
        return;
    }
    int pt[D];
    const signed char* fp = S.fpt + ((fnl * S.norient + 0) * NFQ + q) * D;
#pragma unroll
    for (int d = 0; d < D; ++d) pt[d] = fp[d];
    double uL[M], gL[M * D];
    trace_eval<D, N, M, false>(S, U + (size_t)eL * NB * M, pt, XL, nrm, uL, gL, nullptr);
    double fadv[M], fvis[M], gddg[M * D], uavg[M], uB[M];
    if (bctype == ICE_BC_EXTRAPOLATION || bctype == ICE_BC_SPACETIME_FUTURE) {
        PH::conv(ph, uL, uL, nrm, fadv);
        PH::diff(ph, uL, gL, nrm, fvis);
    } else {
        if (bctype == ICE_BC_DIRICHLET) {
#pragma unroll
            for (int v = 0; v < M; ++v) uB[v] = bv[v];
        } else {
            PH::apply_bc(ph, uL, gL, nrm, bctype, bcflag, uB);
        }
        double cL[D];
        centroid<D>(XL, cL);
        double h = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) h += fabs(nrm[d] * (x[d] - cL[d]));
        h = copysign(fmax(fabs(h), 2.220446049250313e-16), h);
        PH::conv(ph, uL, uB, nrm, fadv);
#pragma unroll
        for (int v = 0; v < M; ++v) {
            const double jump = uB[v] - uL[v];
#pragma unroll
            for (int d = 0; d < D; ++d) gddg[v * D + d] = ph.beta0_bdy * jump / h * nrm[d] + gL[v * D + d];
            uavg[v] = 0.5 * (uL[v] + uB[v]);
        }
        PH::diff(ph, uavg, gddg, nrm, fvis);
        if constexpr (PH::HAS_HOMOGENEITY) {
            if (Pf && bctype == ICE_BC_DIRICHLET) {
                double G[M * D * M * D];
                PH::homogeneity(ph, uavg, G);
#pragma unroll
                for (int v = 0; v < M; ++v)
#pragma unroll
                    for (int s = 0; s < D; ++s) {
                        double acc = 0;
#pragma unroll
                        for (int k = 0; k < D; ++k)
#pragma unroll
                            for (int r = 0; r < M; ++r) acc += G[((v * D + k) * M + r) * D + s] * nrm[k] * (uB[r] - uL[r]);
                        Pf[v * D + s] = ph.sigma_ic * acc * ws;
                    }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < M; ++v) Qf[v] = fvis[v] * ws - fadv[v] * ws;
}

// Interior and boundary faces in one launch: blocks [0, nbi) take interior
// face points, the rest boundary face points.
template <int D, int P, class PH>
__global__ __launch_bounds__(256) void face_kernel(DevSpace S, PhysParams ph, DiscDev dd, const double* __restrict__ U,
                                                   int nbi) {
    if ((int)blockIdx.x < nbi) {
        face_interior<D, P, PH>(S, ph, U, dd.Q, dd.Pic, blockIdx.x * blockDim.x + threadIdx.x);
    } else {
        face_boundary<D, P, PH>(S, ph, dd.bval, U, dd.Q, dd.Pic, (blockIdx.x - nbi) * blockDim.x + threadIdx.x);
    }
}

// ---- elements: volume integral + face lifting -----------------------------
template <int D, int P>
struct ElemTile {
    static constexpr int N = P + 1;
    static constexpr int NB = Pow<N, D>::value;
    static constexpr int EPB = (256 / NB) > 0 ? (256 / NB) : 1;
    static constexpr int NT = EPB * NB;
};

template <int D, int P, class PH>
__global__ __launch_bounds__(256) void element_kernel(DevSpace S, PhysParams ph, const double* __restrict__ srcload,
                                                      const double* __restrict__ U, const double* __restrict__ Q,
                                                      const double* __restrict__ Pic, double* __restrict__ R,
                                                      unsigned long long* __restrict__ wave) {
    using T = ElemTile<D, P>;
    constexpr int N = T::N, NB = T::NB, EPB = T::EPB;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    __shared__ double sU[EPB * NB * M];
    __shared__ double sG[EPB * NB * M * D];
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / NB;
    const int k = threadIdx.x - le * NB;
    const int e = e0 + le;
    const bool active = (threadIdx.x < T::NT) && (e < S.nelem);
    {
        const size_t base = (size_t)e0 * NB * M;
        const size_t total = (size_t)S.nelem * NB * M;
        for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x)
            sU[idx] = (base + idx < total) ? U[base + idx] : 0.0;
    }
    __syncthreads();
    const double* xs = S.tab;
    const double* L0 = S.tab + (N + 2);
    const double* L1 = L0 + (N + 2) * N;
    const double* wgl = L1 + 2 * (N + 2) * N;
    double X[1 << D][D];
    int kd[D];
    {
        int r = k;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) { kd[d] = r % N; r /= N; }
    }
    double lam_w = 0.0, vis_w = 0.0;
    if (active) {
        load_vertices<D>(S.X, e, X);
        // phase 1: (element, quadrature point kd)
        double xi[D], wq = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xi[d] = xs[kd[d]];
            wq = (d == 0) ? wgl[kd[d]] : wq * wgl[kd[d]];
        }
        double x[D], J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, x, J, Hm);
        const double det = det_adj<D>(J, A);
        const double detc = fmax(0.0, det);
        const double detg = (det == 0.0) ? 1.0 : det;
        const double* Ue = sU + le * NB * M;
        double u[M], gr[M][D];
#pragma unroll
        for (int v = 0; v < M; ++v) {
            u[v] = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) gr[v][d] = 0;
        }
        // 1D basis values/derivatives at this point, one row per axis
        double b0[D][N], b1[D][N];
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                b0[d][i] = L0[kd[d] * N + i];
                b1[d][i] = L1[kd[d] * N + i];
            }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int id[D];
            {
                int r = i;
#pragma unroll
                for (int d = D - 1; d >= 0; --d) { id[d] = r % N; r /= N; }
            }
            double c0 = 1.0, cg[D];
#pragma unroll
            for (int d = 0; d < D; ++d) c0 *= b0[d][id[d]];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double t = 1.0;
#pragma unroll
                for (int d = 0; d < D; ++d) t *= (d == a) ? b1[d][id[d]] : b0[d][id[d]];
                cg[a] = t;
            }
#pragma unroll
            for (int v = 0; v < M; ++v) {
                const double c = Ue[i * M + v];
                u[v] += c * c0;
#pragma unroll
                for (int a = 0; a < D; ++a) gr[v][a] += c * cg[a];
            }
        }
        double gx[M * D], F[M * D];
        const double idetg = 1.0 / detg;
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double s = 0;
#pragma unroll
                for (int a = 0; a < D; ++a) s += gr[v][a] * A[a][j];
                gx[v * D + j] = s * idetg;
            }
        PH::flux(ph, u, gx, F);
        if (wave) {
            // lambda_max / visc_max bookkeeping of the physical flux calls
            lam_w = PH::wave_speed(ph, u);
            vis_w = PH::visc_at(ph, u);
            if (e == S.nelem - 1 && k == NB - 1) wave[1] = (unsigned long long)__double_as_longlong(lam_w);
        }
        const double fac = detc * wq * idetg;
        double* Gq = sG + (le * NB + k) * M * D;
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double s = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) s += F[v * D + j] * A[a][j];
                Gq[v * D + a] = s * fac;
            }
    }
    if (wave) {  // wavefront max, inactive lanes contribute 0
        for (int off = 32; off > 0; off >>= 1) {
            lam_w = fmax(lam_w, __shfl_xor(lam_w, off));
            vis_w = fmax(vis_w, __shfl_xor(vis_w, off));
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMax(&wave[0], (unsigned long long)__double_as_longlong(lam_w));
            atomicMax(&wave[2], (unsigned long long)__double_as_longlong(vis_w));
        }
    }
    __syncthreads();
    if (!active) return;
    // phase 2: (element, basis function kd)
    double r[M];
#pragma unroll
    for (int v = 0; v < M; ++v) r[v] = 0;
    // test function kd at the 1D points: one column per axis
    double t0[D][N], t1[D][N];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int q = 0; q < N; ++q) {
            t0[d][q] = L0[q * N + kd[d]];
            t1[d][q] = L1[q * N + kd[d]];
        }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        int qd[D];
        {
            int rr = q;
#pragma unroll
            for (int d = D - 1; d >= 0; --d) { qd[d] = rr % N; rr /= N; }
        }
        double cg[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double t = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) t *= (d == a) ? t1[d][qd[d]] : t0[d][qd[d]];
            cg[a] = t;
        }
        const double* Gq = sG + (le * NB + q) * M * D;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            double s = 0;
#pragma unroll
            for (int a = 0; a < D; ++a) s += Gq[v * D + a] * cg[a];
            r[v] += s;
        }
    }
    if (srcload) {
        const double* sl = srcload + ((size_t)e * NB + k) * M;
#pragma unroll
        for (int v = 0; v < M; ++v) r[v] -= sl[v];
    }
    // lift the face fluxes of the 2*D faces of this element
#pragma unroll 1
    for (int j = 0; j < 2 * D; ++j) {
        const int ef = S.efaces[(size_t)e * 2 * D + j];
        if (ef < 0) continue;
        const size_t slot = (size_t)(ef >> 4);
        const int side = (ef >> 3) & 1;
        const int orient = ef & 7;
        const double sgn = side ? -1.0 : 1.0;
        const signed char* fp = S.fpt + ((size_t)(j * S.norient + orient) * NFQ) * D;
        const double* Qs = Q + slot * NFQ * M;
#pragma unroll 1
        for (int fq = 0; fq < NFQ; ++fq) {
            double phi = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) phi *= L0[fp[fq * D + d] * N + kd[d]];
#pragma unroll
            for (int v = 0; v < M; ++v) r[v] += sgn * Qs[fq * M + v] * phi;
        }
        if constexpr (PH::HAS_HOMOGENEITY) {
            if (Pic) {
                const double* Ps = Pic + slot * NFQ * M * D;
#pragma unroll 1
                for (int fq = 0; fq < NFQ; ++fq) {
                    double xi[D], cg[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) xi[d] = xs[fp[fq * D + d]];
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        double t = 1.0;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            const int s = fp[fq * D + d];
                            t *= (d == a) ? L1[s * N + kd[d]] : L0[s * N + kd[d]];
                        }
                        cg[a] = t;
                    }
                    double J[D][D], Hm[D][D][D], A[D][D];
                    q1_map<D>(X, xi, nullptr, J, Hm);
                    double det = det_adj<D>(J, A);
                    det = (det == 0.0) ? 1.0 : det;
                    double gphi[D];
#pragma unroll
                    for (int jj = 0; jj < D; ++jj) {
                        double s = 0;
#pragma unroll
                        for (int a = 0; a < D; ++a) s += cg[a] * A[a][jj];
                        gphi[jj] = s / det;
                    }
#pragma unroll
                    for (int v = 0; v < M; ++v) {
                        double s = 0;
#pragma unroll
                        for (int d = 0; d < D; ++d) s += Ps[(fq * M + v) * D + d] * gphi[d];
                        r[v] -= s;
                    }
                }
            }
        }
    }
    double* Re = R + ((size_t)e * NB + k) * M;
#pragma unroll
    for (int v = 0; v < M; ++v) Re[v] = r[v];
}

// ---- one-time tables ----------------------------------------------------
// Boundary values at boundary face quadrature points: Dirichlet values
// (dirichlet_callbacks[bcflag](x), conservation_law.hpp:683) or Neumann data
// (neumann_callbacks[bcflag](x), :810); zero for other boundary types.
template <int D, int P>
__global__ __launch_bounds__(256) void bval_kernel(DevSpace S, const ice_field* __restrict__ fields, int M,
                                                   double* __restrict__ bval) {
    constexpr int N = P + 1;
    constexpr int NFQ = Pow<N, D - 1>::value;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = tid / NFQ, q = tid - b * NFQ;
    if (b >= S.nfb) return;
    const int4 rec = S.fbdy[b];
    const int bctype = info_bctype(rec.y), bcflag = rec.z & (ICE_MAX_BCFLAG - 1);
    double* out = bval + ((size_t)b * NFQ + q) * M;
    double X[1 << D][D], x[D], nrm[D];
    load_vertices<D>(S.X, rec.x, X);
    (void)face_geometry<D, N>(S, X, info_fnl(rec.y), q, x, nrm);
    double v[8];
    for (int k = 0; k < M; ++k) v[k] = 0.0;
    if (bctype == ICE_BC_DIRICHLET) ice_field_eval(&fields[bcflag], D, M, x, v);
    else if (bctype == ICE_BC_NEUMANN) ice_field_eval(&fields[ICE_MAX_BCFLAG + bcflag], D, M, x, v);
    for (int k = 0; k < M; ++k) out[k] = v[k];
}

// Source load vector b[e][i][v] = sum_q s_v(x_q) phi_i(xi_q) max(0, detJ) w_q
// (the user_source term of domain_integral, conservation_law.hpp:300-312).
template <int D, int P>
__global__ __launch_bounds__(256) void srcload_kernel(DevSpace S, const ice_field* __restrict__ field, int M, int clamp,
                                                      double* __restrict__ out) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    const int e = blockIdx.x;
    const int i = threadIdx.x;
    if (i >= NB) return;
    const double* xs = S.tab;
    const double* L0 = S.tab + (N + 2);
    const double* wgl = L0 + 3 * (N + 2) * N;
    double X[1 << D][D];
    load_vertices<D>(S.X, e, X);
    int id[D];
    {
        int r = i;
        for (int d = D - 1; d >= 0; --d) { id[d] = r % N; r /= N; }
    }
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < NB; ++q) {
        int qd[D];
        int r = q;
        for (int d = D - 1; d >= 0; --d) { qd[d] = r % N; r /= N; }
        double xi[D], wq = 1.0, phi = 1.0;
        for (int d = 0; d < D; ++d) {
            xi[d] = xs[qd[d]];
            wq = (d == 0) ? wgl[qd[d]] : wq * wgl[qd[d]];
            phi *= L0[qd[d] * N + id[d]];
        }
        double x[D], J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, x, J, Hm);
        const double det = det_adj<D>(J, A);
        const double detc = clamp ? fmax(0.0, det) : det;
        double src[8];
        ice_field_eval(field, D, M, x, src);
        for (int v = 0; v < M; ++v) acc[v] += src[v] * phi * detc * wq;
    }
    for (int v = 0; v < M; ++v) out[((size_t)e * NB + i) * M + v] = acc[v];
}

template <int D, int P>
int launch_setup(const DevSpace& S, const ice_field* fields, int M, double* bval, double* srcload, hipStream_t st) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    if (bval && S.nfb > 0) {
        const long nt = (long)S.nfb * NFQ;
        bval_kernel<D, P><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(S, fields, M, bval);
    }
    if (srcload)
        srcload_kernel<D, P><<<S.nelem, ((NB + 63) / 64) * 64, 0, st>>>(S, fields + 2 * ICE_MAX_BCFLAG, M, 1, srcload);
    return (int)hipGetLastError();
}

// L2 error integrand per element: sum_q |u_h(x_q) - f(x_q)|^2 w_q |detJ_q|
// summed over components (l2_error.hpp:37-94); one thread per element.
template <int D, int P>
__global__ __launch_bounds__(256) void l2err_kernel(DevSpace S, const ice_field* __restrict__ f, int M,
                                                    const double* __restrict__ U, double* __restrict__ out) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S.nelem) return;
    const double* xs = S.tab;
    const double* L0 = S.tab + (N + 2);
    const double* wgl = L0 + 3 * (N + 2) * N;
    double X[1 << D][D];
    load_vertices<D>(S.X, e, X);
    double acc = 0.0;
    for (int q = 0; q < NB; ++q) {
        int qd[D];
        int r = q;
        for (int d = D - 1; d >= 0; --d) { qd[d] = r % N; r /= N; }
        double xi[D], wq = 1.0;
        for (int d = 0; d < D; ++d) {
            xi[d] = xs[qd[d]];
            wq = (d == 0) ? wgl[qd[d]] : wq * wgl[qd[d]];
        }
        double x[D], J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, x, J, Hm);
        const double adet = fabs(det_adj<D>(J, A));
        double fe[8], uh[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        ice_field_eval(f, D, M, x, fe);
        for (int i = 0; i < NB; ++i) {
            int id[D];
            int t = i;
            for (int d = D - 1; d >= 0; --d) { id[d] = t % N; t /= N; }
            double phi = 1.0;
            for (int d = 0; d < D; ++d) phi *= L0[qd[d] * N + id[d]];
            for (int v = 0; v < M; ++v) uh[v] += phi * U[((size_t)e * NB + i) * M + v];
        }
        for (int v = 0; v < M; ++v) acc += (uh[v] - fe[v]) * (uh[v] - fe[v]) * wq * adet;
    }
    out[e] = acc;
}

template <int D, int P>
int launch_proj(const DevSpace& S, const ice_field* f, int M, double* b, hipStream_t st) {
    constexpr int NB = Pow<P + 1, D>::value;
    srcload_kernel<D, P><<<S.nelem, ((NB + 63) / 64) * 64, 0, st>>>(S, f, M, 0, b);
    return (int)hipGetLastError();
}

template <int D, int P>
int launch_l2(const DevSpace& S, const ice_field* f, int M, const double* U, double* out, hipStream_t st) {
    l2err_kernel<D, P><<<(S.nelem + 255) / 256, 256, 0, st>>>(S, f, M, U, out);
    return (int)hipGetLastError();
}

int projection_rhs(int ndim, int p, int neq, const DevSpace& S, const ice_field* f, double* b, hipStream_t st) {
#define ICE_PCASE(DD, PP) if (ndim == DD && p == PP) return launch_proj<DD, PP>(S, f, neq, b, st);
    ICE_PCASE(1, 1) ICE_PCASE(1, 2) ICE_PCASE(1, 3) ICE_PCASE(1, 4)
    ICE_PCASE(2, 1) ICE_PCASE(2, 2) ICE_PCASE(2, 3) ICE_PCASE(2, 4)
    ICE_PCASE(3, 1) ICE_PCASE(3, 2) ICE_PCASE(3, 3) ICE_PCASE(3, 4)
#undef ICE_PCASE
    return -1;
}

int l2_error_terms(int ndim, int p, int neq, const DevSpace& S, const ice_field* f, const double* U, double* out,
                   hipStream_t st) {
#define ICE_LCASE(DD, PP) if (ndim == DD && p == PP) return launch_l2<DD, PP>(S, f, neq, U, out, st);
    ICE_LCASE(1, 1) ICE_LCASE(1, 2) ICE_LCASE(1, 3) ICE_LCASE(1, 4)
    ICE_LCASE(2, 1) ICE_LCASE(2, 2) ICE_LCASE(2, 3) ICE_LCASE(2, 4)
    ICE_LCASE(3, 1) ICE_LCASE(3, 2) ICE_LCASE(3, 3) ICE_LCASE(3, 4)
#undef ICE_LCASE
    return -1;
}

int setup_disc_tables(int ndim, int p, int neq, const DevSpace& S, const ice_field* fields, double* bval,
                      double* srcload, hipStream_t st) {
#define ICE_SETUP_CASE(DD, PP) \
    if (ndim == DD && p == PP) return launch_setup<DD, PP>(S, fields, neq, bval, srcload, st);
    ICE_SETUP_CASE(1, 1) ICE_SETUP_CASE(1, 2) ICE_SETUP_CASE(1, 3) ICE_SETUP_CASE(1, 4)
    ICE_SETUP_CASE(2, 1) ICE_SETUP_CASE(2, 2) ICE_SETUP_CASE(2, 3) ICE_SETUP_CASE(2, 4)
    ICE_SETUP_CASE(3, 1) ICE_SETUP_CASE(3, 2) ICE_SETUP_CASE(3, 3) ICE_SETUP_CASE(3, 4)
#undef ICE_SETUP_CASE
    return -1;
}

// ---- launcher ------------------------------------------------------------
template <int D, int P, class PH>
int launch_residual(const DevSpace& S, const PhysParams& ph, const DiscDev& dd, const double* U, double* R,
                    hipStream_t st, hipEvent_t* ev) {
    constexpr int N = P + 1;
    constexpr int NFQ = Pow<N, D - 1>::value;
    DiscDev d = dd;
    if (!PH::HAS_HOMOGENEITY) d.Pic = nullptr;
    if (ev) (void)hipEventRecord(ev[0], st);
    const int nbi = (int)(((long)S.nfi * NFQ + 255) / 256);
    const int nbb = (int)(((long)S.nfb * NFQ + 255) / 256);
    if (nbi + nbb > 0) face_kernel<D, P, PH><<<nbi + nbb, 256, 0, st>>>(S, ph, d, U, nbi);
    if (ev) (void)hipEventRecord(ev[1], st);
    using T = ElemTile<D, P>;
    const int nblk = (S.nelem + T::EPB - 1) / T::EPB;
    element_kernel<D, P, PH><<<nblk, 256, 0, st>>>(S, ph, d.srcload, U, d.Q, d.Pic, R, d.wave);
    if (ev) (void)hipEventRecord(ev[3], st);
    return (int)hipGetLastError();
}

template <class PH, int D>
int dispatch_order(int p, const DevSpace& S, const PhysParams& ph, const DiscDev& dd, const double* U, double* R,
                   hipStream_t st, hipEvent_t* ev) {
    switch (p) {
    case 1: return launch_residual<D, 1, PH>(S, ph, dd, U, R, st, ev);
    case 2: return launch_residual<D, 2, PH>(S, ph, dd, U, R, st, ev);
    case 3: return launch_residual<D, 3, PH>(S, ph, dd, U, R, st, ev);
    case 4: return launch_residual<D, 4, PH>(S, ph, dd, U, R, st, ev);
    default: return -1;
    }
}

int residual_fp64(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                  const double* U, double* R, hipStream_t st, hipEvent_t* ev) {
    if (flux == ICE_FLUX_BURGERS) {
        if (ndim == 1) return dispatch_order<Burgers<1>, 1>(p, S, ph, dd, U, R, st, ev);
        if (ndim == 2) return dispatch_order<Burgers<2>, 2>(p, S, ph, dd, U, R, st, ev);
        if (ndim == 3) return dispatch_order<Burgers<3>, 3>(p, S, ph, dd, U, R, st, ev);
    } else if (flux == ICE_FLUX_NAVIER_STOKES) {
        if (ndim == 1) return dispatch_order<NavierStokes<1, true>, 1>(p, S, ph, dd, U, R, st, ev);
        if (ndim == 2) return dispatch_order<NavierStokes<2, true>, 2>(p, S, ph, dd, U, R, st, ev);
        if (ndim == 3) return dispatch_order<NavierStokes<3, true>, 3>(p, S, ph, dd, U, R, st, ev);
    } else if (flux == ICE_FLUX_EULER) {
        if (ndim == 1) return dispatch_order<NavierStokes<1, false>, 1>(p, S, ph, dd, U, R, st, ev);
        if (ndim == 2) return dispatch_order<NavierStokes<2, false>, 2>(p, S, ph, dd, U, R, st, ev);
        if (ndim == 3) return dispatch_order<NavierStokes<3, false>, 3>(p, S, ph, dd, U, R, st, ev);
    }
    return -1;
}

}  // namespace ice
