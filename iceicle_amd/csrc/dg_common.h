// Device helpers shared by the residual, Jacobian and setup kernels:
// geometry-order-1 element and face maps, trace evaluation from basis
// coefficients, and the pointwise numerical face fluxes of ConservationLawDDG
// (conservation_law.hpp:434-1176) written once for all callers.
#pragma once
#include <hip/hip_runtime.h>

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

// Face map from the left element's face vertices FX[k] = X[fvo[fnl][k]]
// (k < 2^(D-1)) at face quadrature point fq: physical point x, unit normal
// and sqrt(det(J^T J)).
// (HypercubeFace::transform/Jacobian + calc_ortho/normalize/rootRiemannMetric;
//  1D: PointFace conventions, HypercubeTransformations.hpp:1358-1456.)
template <int D, int N>
__device__ __forceinline__ double face_geometry_fv(const DevSpace& S, const double (&FX)[1 << (D - 1)][D], int fnl,
                                                   int fq, double* x, double* nrm) {
    if constexpr (D == 1) {
        x[0] = FX[0][0];
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
                x[i] += psi * FX[k][i];
#pragma unroll
                for (int t = 0; t < TD; ++t) Jf[i][t] += dpsi[t] * FX[k][i];
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

// The same from all the element's vertices X.
template <int D, int N>
__device__ __forceinline__ double face_geometry(const DevSpace& S, const double (&X)[1 << D][D], int fnl, int fq,
                                                double* x, double* nrm) {
    double FX[1 << (D - 1)][D];
#pragma unroll
    for (int k = 0; k < (1 << (D - 1)); ++k)
#pragma unroll
        for (int d = 0; d < D; ++d) FX[k][d] = X[S.fvo[fnl * 4 + k]][d];
    return face_geometry_fv<D, N>(S, FX, fnl, fq, x, nrm);
}

// Face vertices of element e's face fnl read straight from the vertex table
// (no element-sized register array indexed at run time).
template <int D>
__device__ __forceinline__ void load_face_vertices(const DevSpace& S, int e, int fnl, double (&FX)[1 << (D - 1)][D]) {
    const double* p = S.X + (size_t)e * (1 << D) * D;
#pragma unroll
    for (int k = 0; k < (1 << (D - 1)); ++k) {
        const int v = S.fvo[fnl * 4 + k];
#pragma unroll
        for (int d = 0; d < D; ++d) FX[k][d] = p[v * D + d];
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


// DDGIC lifting vector at one face point (the interface correction of
// conservation_law.hpp:580-610, Dirichlet faces :750-780):
// P[v][s] = sigma sum_k sum_r G(uavg)[v][k][r][s] n_k jmp_r w sqrt(g); each
// element adjacent to the point subtracts P . grad phi_i there.
template <int D, class PH>
__device__ __forceinline__ void ddgic_lift(const typename PH::Params& ph, typename PH::real sigma,
                                           const typename PH::real* uavg, const typename PH::real* nrm,
                                           const typename PH::real* jmp, typename PH::real ws, typename PH::real* P) {
    constexpr int M = PH::M;
    using R = typename PH::real;
    R G[M * D * M * D];
    PH::homogeneity(ph, uavg, G);
#pragma unroll
    for (int v = 0; v < M; ++v)
#pragma unroll
        for (int s = 0; s < D; ++s) {
            R acc = 0;
#pragma unroll
            for (int k = 0; k < D; ++k)
#pragma unroll
                for (int r = 0; r < M; ++r) acc += G[((v * D + k) * M + r) * D + s] * nrm[k] * jmp[r];
            P[v * D + s] = sigma * acc * ws;
        }
}

// Physical gradient of the basis function with tensor index kd at the point
// whose table rows are pt (rows 0..N-1 the nodes, N and N + 1 the interval
// ends) in the element with vertices X
template <int D, int N>
__device__ __forceinline__ void basis_grad_rows(const double* __restrict__ tab, const double (&X)[1 << D][D],
                                                const int* pt, const int* kd, double* gphi) {
    const double* L0 = tab + (N + 2);
    const double* L1 = L0 + (N + 2) * N;
    double xi[D], cg[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = tab[pt[d]];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double t = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) t *= (d == a) ? L1[pt[d] * N + kd[d]] : L0[pt[d] * N + kd[d]];
        cg[a] = t;
    }
    double J[D][D], Hm[D][D][D], A[D][D];
    q1_map<D>(X, xi, nullptr, J, Hm);
    double det = det_adj<D>(J, A);
    det = (det == 0.0) ? 1.0 : det;
#pragma unroll
    for (int jj = 0; jj < D; ++jj) {
        double sacc = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) sacc += cg[a] * A[a][jj];
        gphi[jj] = sacc / det;
    }
}

// DDGIC test gradient at the face point with table rows fp: the basis
// gradient there, or (mirror, the published reading of the interface
// correction) its average with the gradient at the point reflected through
// the element along the face normal, i.e. the same point of the opposite
// face (the one axis whose row is an interval end, N <-> N + 1)
template <int D, int N>
__device__ __forceinline__ void ic_test_gradient(const double* __restrict__ tab, const double (&X)[1 << D][D],
                                                 const signed char* fp, const int* kd, bool mirror, double* gphi) {
    int pt[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pt[d] = fp[d];
    basis_grad_rows<D, N>(tab, X, pt, kd, gphi);
    if (mirror) {
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (pt[d] >= N) pt[d] = 2 * N + 1 - pt[d];
        double gm[D];
        basis_grad_rows<D, N>(tab, X, pt, kd, gm);
#pragma unroll
        for (int d = 0; d < D; ++d) gphi[d] = 0.5 * (gphi[d] + gm[d]);
    }
}

// DDG single-valued gradient and weighted normal flux at one interior face
// point (conservation_law.hpp:502-575): Q = (F_visc - F_conv) w sqrt(g).
template <int D, class PH>
__device__ __forceinline__ void interior_point_flux(const typename PH::Params& ph, const typename PH::real* uL, const typename PH::real* gL,
                                                    const typename PH::real* hL, const typename PH::real* uR, const typename PH::real* gR,
                                                    const typename PH::real* hR, const typename PH::real* nrm, typename PH::real h, typename PH::real ws,
                                                    typename PH::real* Q) {
    constexpr int M = PH::M;
    using R = typename PH::real;
    R fadv[M], fvis[M], gddg[M * D], uavg[M];
    PH::conv(ph, uL, uR, nrm, fadv);
#pragma unroll
    for (int v = 0; v < M; ++v) {
        const R jump = uR[v] - uL[v];
#pragma unroll
        for (int d = 0; d < D; ++d)
            gddg[v * D + d] = ph.beta0 * jump / h * nrm[d] + R(0.5) * (gL[v * D + d] + gR[v * D + d]) +
                              ph.beta1 * h * (hR[v * D + d] - hL[v * D + d]);
        uavg[v] = R(0.5) * (uL[v] + uR[v]);
    }
    PH::diff(ph, uavg, gddg, nrm, fvis);
#pragma unroll
    for (int v = 0; v < M; ++v) Q[v] = fvis[v] * ws - fadv[v] * ws;
}

// The same flux from the sum gs = gL + gR of the two elements' shares of the
// DDG gradient (each (grad u)/2 -/+ beta1 h (Hess u) n, see trace_kernel).
template <int D, class PH>
__device__ __forceinline__ void interior_point_flux_split(const typename PH::Params& ph, const typename PH::real* uL, const typename PH::real* uR,
                                                          const typename PH::real* gs, const typename PH::real* nrm, typename PH::real h, typename PH::real ws,
                                                          typename PH::real* Q, const typename PH::real* jmp = nullptr) {
    // jmp: uR - uL formed in higher precision by the caller (FP32 fluxes),
    // else the difference of the two arguments
    constexpr int M = PH::M;
    using R = typename PH::real;
    R fadv[M], fvis[M], gddg[M * D], uavg[M];
    PH::conv(ph, uL, uR, nrm, fadv);
#pragma unroll
    for (int v = 0; v < M; ++v) {
        const R jump = jmp ? jmp[v] : uR[v] - uL[v];
#pragma unroll
        for (int d = 0; d < D; ++d) gddg[v * D + d] = ph.beta0 * jump / h * nrm[d] + gs[v * D + d];
        uavg[v] = R(0.5) * (uL[v] + uR[v]);
    }
    PH::diff(ph, uavg, gddg, nrm, fvis);
#pragma unroll
    for (int v = 0; v < M; ++v) Q[v] = fvis[v] * ws - fadv[v] * ws;
}

// Weighted normal flux at one boundary face point (boundaryIntegral,
// conservation_law.hpp:631-1176).  bv: Dirichlet values or Neumann data at
// the point; hb: the guarded boundary DDG distance.  Returns false when the
// boundary adds nothing (default branch of a flux without apply_bc).
template <int D, class PH>
__device__ __forceinline__ bool boundary_point_flux(const typename PH::Params& ph, int bctype, int bcflag, const typename PH::real* uL,
                                                    const typename PH::real* gL, const typename PH::real* bv, const typename PH::real* nrm, typename PH::real hb,
                                                    typename PH::real ws, typename PH::real* Q, typename PH::real* uB_out,
                                                    const typename PH::real* jmp_dirichlet = nullptr) {
    // jmp_dirichlet: bv - uL formed in higher precision by the caller (FP32
    // fluxes on Dirichlet faces), else the difference of the arguments
    constexpr int M = PH::M;
    using R = typename PH::real;
    R fadv[M], fvis[M], gddg[M * D], uavg[M], uB[M];
    if (bctype == ICE_BC_NEUMANN) {
        R g[M];
#pragma unroll
        for (int v = 0; v < M; ++v) g[v] = bv[v];
        PH::neumann(ph, g, fvis);
#pragma unroll
        for (int v = 0; v < M; ++v) Q[v] = fvis[v] * ws;
        return true;
    }
    const bool general = !(bctype == ICE_BC_DIRICHLET || bctype == ICE_BC_EXTRAPOLATION ||
                           bctype == ICE_BC_SPACETIME_FUTURE);
    if (general && !PH::IMPLEMENTS_BCS) {
#pragma unroll
        for (int v = 0; v < M; ++v) Q[v] = R(0);
        return false;
    }
    // One call of each flux on value arrays for every boundary type:
    // extrapolation takes uB = uL, the interior gradient and uL as the state
    // by value selects (identical values to separate calls with uL and gL;
    // calls with different array arguments per branch would be merged into
    // one call through a pointer select and move the arrays to scratch).
    const bool extrap = bctype == ICE_BC_EXTRAPOLATION || bctype == ICE_BC_SPACETIME_FUTURE;
    if (bctype == ICE_BC_DIRICHLET) {
#pragma unroll
        for (int v = 0; v < M; ++v) uB[v] = bv[v];
    } else if (extrap) {
#pragma unroll
        for (int v = 0; v < M; ++v) uB[v] = uL[v];
    } else {
        PH::apply_bc(ph, uL, gL, nrm, bctype, bcflag, uB);
    }
    PH::conv(ph, uL, uB, nrm, fadv);
#pragma unroll
    for (int v = 0; v < M; ++v) {
        const R jump = (jmp_dirichlet && bctype == ICE_BC_DIRICHLET) ? jmp_dirichlet[v] : uB[v] - uL[v];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const R gd = ph.beta0_bdy * jump / hb * nrm[d] + gL[v * D + d];
            gddg[v * D + d] = extrap ? gL[v * D + d] : gd;
        }
        uavg[v] = extrap ? uL[v] : R(0.5) * (uL[v] + uB[v]);
    }
    PH::diff(ph, uavg, gddg, nrm, fvis);
#pragma unroll
    for (int v = 0; v < M; ++v) Q[v] = fvis[v] * ws - fadv[v] * ws;
    if (uB_out) {
#pragma unroll
        for (int v = 0; v < M; ++v) uB_out[v] = uB[v];
    }
    return true;
}

// guarded DDG distances (conservation_law.hpp:518-527, 707-713)
template <int D>
__device__ __forceinline__ double ddg_h_interior(const double* nrm, const double* x, const double* cL, const double* cR) {
    double h = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) h += nrm[d] * ((x[d] - cL[d]) + (cR[d] - x[d]));
    return copysign(fmax(fabs(h), 2.220446049250313e-16), h);
}
template <int D>
__device__ __forceinline__ double ddg_h_boundary(const double* nrm, const double* x, const double* cL) {
    double h = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) h += fabs(nrm[d] * (x[d] - cL[d]));
    return copysign(fmax(fabs(h), 2.220446049250313e-16), h);
}

}  // namespace ice
