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

#include "dg_common.h"
#include "fields.h"
#include "kernels.h"
#include "physics_dev.h"

namespace ice {

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
    const double h = ddg_h_interior<D>(nrm, x, cL, cR);
    const double ws = w * sqrtg;
    double* Qf = Q + ((size_t)f * NFQ + q) * M;
    {
        double Qv[M];
        interior_point_flux<D, PH>(ph, uL, gL, hL, uR, gR, hR, nrm, h, ws, Qv);
#pragma unroll
        for (int v = 0; v < M; ++v) Qf[v] = Qv[v];
    }
    double uavg[M];
#pragma unroll
    for (int v = 0; v < M; ++v) uavg[v] = 0.5 * (uL[v] + uR[v]);
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
    double uL[M] = {}, gL[M * D] = {};
    if (bctype != ICE_BC_NEUMANN) {
        int pt[D];
        const signed char* fp = S.fpt + ((fnl * S.norient + 0) * NFQ + q) * D;
#pragma unroll
        for (int d = 0; d < D; ++d) pt[d] = fp[d];
        trace_eval<D, N, M, false>(S, U + (size_t)eL * NB * M, pt, XL, nrm, uL, gL, nullptr);
    }
    double cL[D];
    centroid<D>(XL, cL);
    const double hb = ddg_h_boundary<D>(nrm, x, cL);
    double Qv[M], uB[M];
    boundary_point_flux<D, PH>(ph, bctype, bcflag, uL, gL, bv, nrm, hb, ws, Qv, uB);
#pragma unroll
    for (int v = 0; v < M; ++v) Qf[v] = Qv[v];
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (Pf && bctype == ICE_BC_DIRICHLET) {
            double uavg[M], G[M * D * M * D];
#pragma unroll
            for (int v = 0; v < M; ++v) uavg[v] = 0.5 * (uL[v] + uB[v]);
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
                                                      unsigned long long* __restrict__ wave,
                                                      double* __restrict__ wave_part) {
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
    __shared__ double red[2 * 4];
    if (wave) {  // block max into wave_part (inactive lanes contribute 0); folded by wave_reduce_kernel
        for (int off = 32; off > 0; off >>= 1) {
            lam_w = fmax(lam_w, __shfl_xor(lam_w, off));
            vis_w = fmax(vis_w, __shfl_xor(vis_w, off));
        }
        if ((threadIdx.x & 63) == 0) {
            red[2 * (threadIdx.x >> 6)] = lam_w;
            red[2 * (threadIdx.x >> 6) + 1] = vis_w;
        }
    }
    __syncthreads();
    if (wave && threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { a = fmax(a, red[2 * w]); b = fmax(b, red[2 * w + 1]); }
        wave_part[2 * blockIdx.x] = a;
        wave_part[2 * blockIdx.x + 1] = b;
    }
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

// Fold the per-block maxima of one residual into the running maxima
// (lambda_max of BurgersFlux never resets, burgers.hpp:89; visc_max, navier_stokes.hpp:964-966).
__global__ void wave_reduce_kernel(const double* __restrict__ part, int nblk, unsigned long long* __restrict__ wave) {
    __shared__ double sa[256], sb[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a = fmax(a, part[2 * i]);
        b = fmax(b, part[2 * i + 1]);
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + s]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double ra = fmax(__longlong_as_double((long long)wave[0]), sa[0]);
        const double rb = fmax(__longlong_as_double((long long)wave[2]), sb[0]);
        wave[0] = (unsigned long long)__double_as_longlong(ra);
        wave[2] = (unsigned long long)__double_as_longlong(rb);
    }
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
    element_kernel<D, P, PH><<<nblk, 256, 0, st>>>(S, ph, d.srcload, U, d.Q, d.Pic, R, d.wave, d.wave_part);
    if (d.wave) wave_reduce_kernel<<<1, 256, 0, st>>>(d.wave_part, nblk, d.wave);
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
