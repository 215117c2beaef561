// Finite-difference element-block Jacobian on the device
// (form_petsc_jacobian_fd DG part, form_petsc_jacobian.hpp:55-317, and
// ConservationLawDDG::domain_integral_jacobian, conservation_law.hpp:316-432).
//
// Storage: block-sparse rows, one block row per element with 2*ndim+1 slots:
// slot 0 = diagonal block, slot 1+j = the neighbour across local face j.
// Blocks are (nb*neq)^2 row-major [i*neq+v][j*neq+w] in FP64.
//
// Face blocks: one workgroup per face.  The reference perturbs one
// coefficient at a time and re-integrates the whole face; a coefficient
// perturbation changes the traces at each face point by eps*phi_i,
// eps*grad(phi_i) and eps*Hess(phi_i).n only, so each (column, face point)
// pair re-evaluates just the pointwise numerical flux from the stored base
// traces.  The flux differences are lifted to block entries in LDS; blocks
// coupling two different elements belong to exactly one face and are
// stored directly.  Diagonal blocks are owned by their element: the volume
// kernel writes them first, then the boundary and interior faces add their
// contributions one face colour per launch (no two faces of a colour share
// an element), so every entry is summed in the same order on every run and
// no atomics are needed.
#include <hip/hip_runtime.h>

#include "dg_common.h"
#include "fields.h"
#include "kernels.h"
#include "physics_dev.h"

namespace ice {

namespace {

constexpr double SQRT_EPS = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)

// phi_i, reference gradient and reference Hessian at S-point pt (tensor products)
template <int D, int N>
__device__ __forceinline__ void basis_at(const double* tab, const int* pt, int i, double& phi, double* gref,
                                         double* href) {
    const double* L0 = tab + (N + 2);
    const double* L1 = L0 + (N + 2) * N;
    const double* L2 = L1 + (N + 2) * N;
    int id[D];
    int r = i;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) { id[d] = r % N; r /= N; }
    phi = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) phi *= L0[pt[d] * N + id[d]];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double t = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) t *= (d == a) ? L1[pt[d] * N + id[d]] : L0[pt[d] * N + id[d]];
        gref[a] = t;
    }
    if (href) {
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                double t = 1.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    if (a == b) t *= (d == a) ? L2[pt[d] * N + id[d]] : L0[pt[d] * N + id[d]];
                    else t *= (d == a || d == b) ? L1[pt[d] * N + id[d]] : L0[pt[d] * N + id[d]];
                }
                href[a * D + b] = t;
            }
    }
}

// per (side, face point) geometry kept in LDS: adj(J)/det, the map Hessian
template <int D>
struct PointGeo {
    double Ai[D][D];    // adj(J) / det
    double Hm[D][D][D];
};

template <int D>
__device__ __forceinline__ void make_geo(const double (&X)[1 << D][D], const double* xi, PointGeo<D>& g) {
    double J[D][D], A[D][D];
    q1_map<D>(X, xi, nullptr, J, g.Hm);
    double det = det_adj<D>(J, A);
    det = (det == 0.0) ? 1.0 : det;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) g.Ai[a][b] = A[a][b] / det;
}

// physical gradient and normal Hessian term of one basis function
template <int D>
__device__ __forceinline__ void phys_basis(const PointGeo<D>& g, const double* gref, const double* href,
                                           const double* nrm, double* gx, double* hn) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) s += gref[k] * g.Ai[k][j];
        gx[j] = s;
    }
    if (hn) {
        double rhs[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                double s = href[a * D + b];
#pragma unroll
                for (int k = 0; k < D; ++k) s -= g.Hm[k][a][b] * gx[k];
                rhs[a][b] = s;
            }
        // Hx = Ai^T rhs Ai ; hn[d] = sum_j Hx[j][d] n_j
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double hx = 0;
#pragma unroll
                for (int l = 0; l < D; ++l)
#pragma unroll
                    for (int k = 0; k < D; ++k) hx += g.Ai[l][j] * rhs[l][k] * g.Ai[k][d];
                s += hx * nrm[j];
            }
            hn[d] = s;
        }
    }
}

}  // namespace

// ---- interior faces --------------------------------------------------------
// One workgroup per face of one colour (jorder[base + blockIdx.x]).
template <int D, int P, class PH>
__global__ __launch_bounds__(256) void jac_face_kernel(DevSpace S, PhysParams ph, const double* __restrict__ U,
                                                       double* __restrict__ J, int bslots, int base) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    constexpr int NBM = NB * M;
    constexpr int NC = 2 * NBM;  // columns: L dofs then R dofs
    __shared__ PointGeo<D> sgeo[2][NFQ];
    __shared__ double su[2][NFQ][M], sg[2][NFQ][M * D], shn[2][NFQ][M * D];
    __shared__ double snrm[NFQ][D], sh[NFQ], sws[NFQ];
    __shared__ int spt[2][NFQ][D];
    // basis value, physical gradient and normal Hessian term of every basis
    // function at every face point, per side
    __shared__ double bphi[2][NFQ][NB], bgx[2][NFQ][NB][D], bhn[2][NFQ][NB][D];
    __shared__ double sQ0[NFQ][M];
    __shared__ double sdQ[NC][NFQ][M];
    // DDGIC lifting vectors (fluxes with a homogeneity tensor): base and
    // per-column differences, lifted with the basis gradients of each side
    constexpr bool IC = PH::HAS_HOMOGENEITY;
    constexpr int MD = M * D;
    __shared__ double sP0[IC ? NFQ : 1][IC ? MD : 1];
    __shared__ double sdP[IC ? NC : 1][IC ? NFQ : 1][IC ? MD : 1];
    const bool ddgic = IC && ph.sigma_face != 0.0;
    // the lift's test gradients (ic_test_gradient): the face point's, or with
    // ic_mirror their average with the mirror point's, whose geometry sgeom holds
    __shared__ double btg[IC ? 2 : 1][IC ? NFQ : 1][IC ? NB : 1][IC ? D : 1];
    __shared__ PointGeo<D> sgeom[IC ? 2 : 1][IC ? NFQ : 1];
    __shared__ double sred[256];
    __shared__ double seps;
    const int f = S.jorder[base + blockIdx.x];
    const int4 rec = S.fint[f];
    const int eL = rec.x, eR = rec.y, info = rec.z;
    const int fnl = info_fnl(info), fnr = info_fnr(info), orient = info_orient(info), pdim = info_pdim(info);
    const int tid = threadIdx.x;
    if (tid < 2 * NFQ) {
        const int side = tid / NFQ, q = tid % NFQ;
        double XL[1 << D][D], XS[1 << D][D];
        load_vertices<D>(S.X, eL, XL);
        load_vertices<D>(S.X, side ? eR : eL, XS);
        double x[D], nrm[D];
        const double sqrtg = face_geometry<D, N>(S, XL, fnl, q, x, nrm);
        int pt[D];
        const signed char* fp =
            S.fpt + (((side ? fnr : fnl) * S.norient + (side ? orient : 0)) * NFQ + q) * D;
        double xi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pt[d] = fp[d];
            spt[side][q][d] = pt[d];
            xi[d] = S.tab[pt[d]];
        }
        make_geo<D>(XS, xi, sgeo[side][q]);
        if constexpr (IC) {
            if (ddgic && ph.ic_mirror) {
                double xim[D];
#pragma unroll
                for (int d = 0; d < D; ++d) xim[d] = S.tab[pt[d] >= N ? 2 * N + 1 - pt[d] : pt[d]];
                make_geo<D>(XS, xim, sgeom[side][q]);
            }
        }
        double u[M], g[M * D], hn[M * D];
        trace_eval<D, N, M, true>(S, U + (size_t)(side ? eR : eL) * NB * M, pt, XS, nrm, u, g, hn);
#pragma unroll
        for (int v = 0; v < M; ++v) su[side][q][v] = u[v];
#pragma unroll
        for (int k = 0; k < M * D; ++k) { sg[side][q][k] = g[k]; shn[side][q][k] = hn[k]; }
        if (side == 0) {
            double XR[1 << D][D], cL[D], cR[D];
            load_vertices<D>(S.X, eR, XR);
            centroid<D>(XL, cL);
            centroid<D>(XR, cR);
            if (pdim) cR[pdim - 1] += S.period[pdim - 1];
#pragma unroll
            for (int d = 0; d < D; ++d) snrm[q][d] = nrm[d];
            sh[q] = ddg_h_interior<D>(nrm, x, cL, cR);
            sws[q] = face_weight<N, D>(S, q) * sqrtg;
        }
    }
    __syncthreads();
    for (int t = tid; t < 2 * NFQ * NB; t += blockDim.x) {
        const int side = t / (NFQ * NB), q = (t / NB) % NFQ, i = t % NB;
        double phi, gref[D], href[D * D], gx[D], hn[D];
        basis_at<D, N>(S.tab, spt[side][q], i, phi, gref, href);
        phys_basis<D>(sgeo[side][q], gref, href, snrm[q], gx, hn);
        bphi[side][q][i] = phi;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            bgx[side][q][i][d] = gx[d];
            bhn[side][q][i][d] = hn[d];
        }
        if constexpr (IC) {
            if (ddgic) {
                if (ph.ic_mirror) {
                    int ptm[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const int pd = spt[side][q][d];
                        ptm[d] = pd >= N ? 2 * N + 1 - pd : pd;
                    }
                    double phm, grm[D], hrm[D * D], gxm[D], hnm[D];
                    basis_at<D, N>(S.tab, ptm, i, phm, grm, hrm);
                    phys_basis<D>(sgeom[side][q], grm, hrm, snrm[q], gxm, hnm);
#pragma unroll
                    for (int d = 0; d < D; ++d) btg[side][q][i][d] = 0.5 * (gx[d] + gxm[d]);
                } else {
#pragma unroll
                    for (int d = 0; d < D; ++d) btg[side][q][i][d] = gx[d];
                }
            }
        }
    }
    if (tid < NFQ) {
        double Q[M];
        interior_point_flux<D, PH>(ph, su[0][tid], sg[0][tid], shn[0][tid], su[1][tid], sg[1][tid], shn[1][tid],
                                   snrm[tid], sh[tid], sws[tid], Q);
#pragma unroll
        for (int v = 0; v < M; ++v) sQ0[tid][v] = Q[v];
        if constexpr (IC) {
            if (ddgic) {
                double uavg[M], jmp[M], Pv[MD];
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    uavg[v] = 0.5 * (su[0][tid][v] + su[1][tid][v]);
                    jmp[v] = su[1][tid][v] - su[0][tid][v];
                }
                ddgic_lift<D, PH>(ph, ph.sigma_face, uavg, snrm[tid], jmp, sws[tid], Pv);
#pragma unroll
                for (int k = 0; k < MD; ++k) sP0[tid][k] = Pv[k];
            }
        }
    }
    __syncthreads();
    // eps = max(sqrt eps, sqrt eps * max(|resL|, |resR|)) (form_petsc_jacobian.hpp:191),
    // the face's residual on either side: +-Q lifted with the trace basis,
    // minus the DDGIC vectors against the side's basis gradients
    double sL = 0, sR = 0;
    for (int r = tid; r < NC; r += blockDim.x) {
        const int side = r / NBM, iv = r % NBM, i = iv / M, v = iv % M;
        double acc = 0;
        for (int q = 0; q < NFQ; ++q) acc += sQ0[q][v] * bphi[side][q][i];
        if (side) acc = -acc;
        if constexpr (IC) {
            if (ddgic)
                for (int q = 0; q < NFQ; ++q)
#pragma unroll
                    for (int d = 0; d < D; ++d) acc -= sP0[q][v * D + d] * btg[side][q][i][d];
        }
        if (side) sR += acc * acc;
        else sL += acc * acc;
    }
    sred[tid] = sL;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) sred[tid] += sred[tid + s];
        __syncthreads();
    }
    const double nL = sqrt(sred[0]);
    __syncthreads();
    sred[tid] = sR;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) sred[tid] += sred[tid + s];
        __syncthreads();
    }
    if (tid == 0) seps = fmax(SQRT_EPS, SQRT_EPS * fmax(nL, sqrt(sred[0])));
    __syncthreads();
    const double eps = seps;
    // flux differences for every (column, face point)
    for (int it = tid; it < NC * NFQ; it += blockDim.x) {
        const int c = it / NFQ, q = it % NFQ;
        const int side = c / NBM, iw = c % NBM, i = iw / M, w = iw % M;
        // both sides' records with column c's perturbation added by selects
        // (a runtime index into the register arrays would move them to scratch)
        double uL[M], gL[M * D], hL[M * D], uR[M], gR[M * D], hR[M * D];
        const double du = eps * bphi[side][q][i];
        double dg[D], dh[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            dg[d] = eps * bgx[side][q][i][d];
            dh[d] = eps * bhn[side][q][i][d];
        }
#pragma unroll
        for (int v = 0; v < M; ++v) {
            const bool hitL = side == 0 && v == w, hitR = side == 1 && v == w;
            uL[v] = su[0][q][v];
            uR[v] = su[1][q][v];
            if (hitL) uL[v] += du;
            if (hitR) uR[v] += du;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                gL[v * D + d] = sg[0][q][v * D + d];
                hL[v * D + d] = shn[0][q][v * D + d];
                gR[v * D + d] = sg[1][q][v * D + d];
                hR[v * D + d] = shn[1][q][v * D + d];
                if (hitL) {
                    gL[v * D + d] += dg[d];
                    hL[v * D + d] += dh[d];
                }
                if (hitR) {
                    gR[v * D + d] += dg[d];
                    hR[v * D + d] += dh[d];
                }
            }
        }
        double Q[M];
        interior_point_flux<D, PH>(ph, uL, gL, hL, uR, gR, hR, snrm[q], sh[q], sws[q], Q);
#pragma unroll
        for (int v = 0; v < M; ++v) sdQ[c][q][v] = Q[v] - sQ0[q][v];
        if constexpr (IC) {
            if (ddgic) {
                double uavg[M], jmp[M], Pv[MD];
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    uavg[v] = 0.5 * (uL[v] + uR[v]);
                    jmp[v] = uR[v] - uL[v];
                }
                ddgic_lift<D, PH>(ph, ph.sigma_face, uavg, snrm[q], jmp, sws[q], Pv);
#pragma unroll
                for (int k = 0; k < MD; ++k) sdP[c][q][k] = Pv[k] - sP0[q][k];
            }
        }
    }
    __syncthreads();
    // lift: row (side rs, i, v), column c; the left rows first, then the
    // right rows (with periodic coupling across a single element both sides
    // can be the same element, whose diagonal block then takes both in turn)
    const size_t bs = NBM;
    const double ieps = 1.0 / eps;
    // each thread takes LU entries a pass and issues the loads of its
    // diagonal-block entries together before adding to them (the adds wait
    // on one round trip instead of LU)
    constexpr int ENT = NBM * NC;
    constexpr int LU = 8;
#pragma unroll 1
    for (int rs = 0; rs < 2; ++rs) {
        const int erow = rs ? eR : eL;
        double* Jd = J + ((size_t)erow * bslots) * bs * bs;
        double* Jo = J + ((size_t)erow * bslots + 1 + (rs ? fnr : fnl)) * bs * bs;
#pragma unroll 1
        for (int e0 = tid; e0 < ENT; e0 += blockDim.x * LU) {
            double val[LU], old[LU];
            int off[LU];
            bool diag[LU], ok[LU];
#pragma unroll
            for (int k = 0; k < LU; ++k) {
                const int ent = e0 + k * (int)blockDim.x;
                ok[k] = ent < ENT;
                const int en = ok[k] ? ent : 0;
                const int iv = en / NC, c = en % NC;
                const int i = iv / M, v = iv % M;
                const int cs = c / NBM, cw = c % NBM;
                double acc = 0;
#pragma unroll
                for (int q = 0; q < NFQ; ++q) acc += sdQ[c][q][v] * bphi[rs][q][i];
                if (rs) acc = -acc;
                if constexpr (IC) {
                    if (ddgic)
#pragma unroll
                        for (int q = 0; q < NFQ; ++q)
#pragma unroll
                            for (int d = 0; d < D; ++d) acc -= sdP[c][q][v * D + d] * btg[rs][q][i][d];
                }
                val[k] = acc * ieps;
                diag[k] = rs == cs;
                off[k] = iv * (int)bs + cw;
                old[k] = (ok[k] && diag[k]) ? Jd[off[k]] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < LU; ++k) {
                if (!ok[k]) continue;
                if (diag[k]) Jd[off[k]] = old[k] + val[k];
                else Jo[off[k]] = val[k];
            }
        }
        __syncthreads();
    }
}

// ---- boundary faces --------------------------------------------------------
template <int D, int P, class PH>
__global__ __launch_bounds__(256) void jac_bface_kernel(DevSpace S, PhysParams ph, const double* __restrict__ bval,
                                                        const double* __restrict__ U, double* __restrict__ J,
                                                        int bslots, int base) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    constexpr int NBM = NB * M;
    __shared__ PointGeo<D> sgeo[NFQ];
    __shared__ double su[NFQ][M], sg[NFQ][M * D];
    __shared__ double snrm[NFQ][D], sh[NFQ], sws[NFQ];
    __shared__ int spt[NFQ][D];
    __shared__ double bphi[NFQ][NB], bgx[NFQ][NB][D];
    __shared__ double sQ0[NFQ][M];
    __shared__ double sdQ[NBM][NFQ][M];
    constexpr bool IC = PH::HAS_HOMOGENEITY;
    constexpr int MD = M * D;
    __shared__ double sP0[IC ? NFQ : 1][IC ? MD : 1];
    __shared__ double sdP[IC ? NBM : 1][IC ? NFQ : 1][IC ? MD : 1];
    __shared__ double sred[256];
    __shared__ double seps;
    const int b = S.jorder[S.nfi + base + blockIdx.x];
    const int4 rec = S.fbdy[b];
    const int eL = rec.x, info = rec.y, bcflag = rec.z;
    const int fnl = info_fnl(info), bctype = info_bctype(info);
    const bool general = !(bctype == ICE_BC_DIRICHLET || bctype == ICE_BC_NEUMANN || bctype == ICE_BC_EXTRAPOLATION ||
                           bctype == ICE_BC_SPACETIME_FUTURE);
    if (bctype == ICE_BC_NEUMANN || (general && !PH::IMPLEMENTS_BCS)) return;  // no dependence on u
    // the DDGIC Dirichlet-face term (the source form's sigma_bdy)
    const bool ddgic = IC && ph.sigma_bdy != 0.0 && bctype == ICE_BC_DIRICHLET;
    const int tid = threadIdx.x;
    if (tid < NFQ) {
        const int q = tid;
        double XL[1 << D][D];
        load_vertices<D>(S.X, eL, XL);
        double x[D], nrm[D];
        const double sqrtg = face_geometry<D, N>(S, XL, fnl, q, x, nrm);
        int pt[D];
        const signed char* fp = S.fpt + ((fnl * S.norient + 0) * NFQ + q) * D;
        double xi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pt[d] = fp[d];
            spt[q][d] = pt[d];
            xi[d] = S.tab[pt[d]];
        }
        make_geo<D>(XL, xi, sgeo[q]);
        double u[M], g[M * D];
        trace_eval<D, N, M, false>(S, U + (size_t)eL * NB * M, pt, XL, nrm, u, g, nullptr);
#pragma unroll
        for (int v = 0; v < M; ++v) su[q][v] = u[v];
#pragma unroll
        for (int k = 0; k < M * D; ++k) sg[q][k] = g[k];
        double cL[D];
        centroid<D>(XL, cL);
#pragma unroll
        for (int d = 0; d < D; ++d) snrm[q][d] = nrm[d];
        sh[q] = ddg_h_boundary<D>(nrm, x, cL);
        sws[q] = face_weight<N, D>(S, q) * sqrtg;
        double Q[M], uB[M];
        boundary_point_flux<D, PH>(ph, bctype, bcflag, u, g, bval + ((size_t)b * NFQ + q) * M, nrm, sh[q], sws[q], Q,
                                   uB);
#pragma unroll
        for (int v = 0; v < M; ++v) sQ0[q][v] = Q[v];
        if constexpr (IC) {
            if (ddgic) {
                double uavg[M], jmp[M], Pv[MD];
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    uavg[v] = 0.5 * (u[v] + uB[v]);
                    jmp[v] = uB[v] - u[v];
                }
                ddgic_lift<D, PH>(ph, ph.sigma_bdy, uavg, nrm, jmp, sws[q], Pv);
#pragma unroll
                for (int k = 0; k < MD; ++k) sP0[q][k] = Pv[k];
            }
        }
    }
    __syncthreads();
    for (int t = tid; t < NFQ * NB; t += blockDim.x) {
        const int q = t / NB, i = t % NB;
        double phi, gref[D], gx[D];
        basis_at<D, N>(S.tab, spt[q], i, phi, gref, nullptr);
        phys_basis<D>(sgeo[q], gref, nullptr, snrm[q], gx, nullptr);
        bphi[q][i] = phi;
#pragma unroll
        for (int d = 0; d < D; ++d) bgx[q][i][d] = gx[d];
    }
    __syncthreads();
    double sL = 0;
    for (int r = tid; r < NBM; r += blockDim.x) {
        const int i = r / M, v = r % M;
        double acc = 0;
        for (int q = 0; q < NFQ; ++q) acc += sQ0[q][v] * bphi[q][i];
        if constexpr (IC) {
            if (ddgic)
                for (int q = 0; q < NFQ; ++q)
#pragma unroll
                    for (int d = 0; d < D; ++d) acc -= sP0[q][v * D + d] * bgx[q][i][d];
        }
        sL += acc * acc;
    }
    sred[tid] = sL;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) sred[tid] += sred[tid + s];
        __syncthreads();
    }
    if (tid == 0) seps = fmax(SQRT_EPS, SQRT_EPS * sqrt(sred[0]));  // form_petsc_jacobian.hpp:122
    __syncthreads();
    const double eps = seps;
    for (int it = tid; it < NBM * NFQ; it += blockDim.x) {
        const int c = it / NFQ, q = it % NFQ;
        const int i = c / M, w = c % M;
        double u[M], g[M * D];
#pragma unroll
        for (int v = 0; v < M; ++v) u[v] = su[q][v];
#pragma unroll
        for (int k = 0; k < M * D; ++k) g[k] = sg[q][k];
        u[w] += eps * bphi[q][i];
#pragma unroll
        for (int d = 0; d < D; ++d) g[w * D + d] += eps * bgx[q][i][d];
        double Q[M], uB[M];
        boundary_point_flux<D, PH>(ph, bctype, bcflag, u, g, bval + ((size_t)b * NFQ + q) * M, snrm[q], sh[q], sws[q],
                                   Q, uB);
#pragma unroll
        for (int v = 0; v < M; ++v) sdQ[c][q][v] = Q[v] - sQ0[q][v];
        if constexpr (IC) {
            if (ddgic) {
                double uavg[M], jmp[M], Pv[MD];
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    uavg[v] = 0.5 * (u[v] + uB[v]);
                    jmp[v] = uB[v] - u[v];
                }
                ddgic_lift<D, PH>(ph, ph.sigma_bdy, uavg, snrm[q], jmp, sws[q], Pv);
#pragma unroll
                for (int k = 0; k < MD; ++k) sdP[c][q][k] = Pv[k] - sP0[q][k];
            }
        }
    }
    __syncthreads();
    const size_t bs = NBM;
    const double ieps = 1.0 / eps;
    for (int ent = tid; ent < NBM * NBM; ent += blockDim.x) {
        const int r = ent / NBM, c = ent % NBM;
        const int i = r / M, v = r % M;
        double acc = 0;
#pragma unroll
        for (int q = 0; q < NFQ; ++q) acc += sdQ[c][q][v] * bphi[q][i];
        if constexpr (IC) {
            if (ddgic)
#pragma unroll
                for (int q = 0; q < NFQ; ++q)
#pragma unroll
                    for (int d = 0; d < D; ++d) acc -= sdP[c][q][v * D + d] * bgx[q][i][d];
        }
        J[((size_t)eL * bslots) * bs * bs + (size_t)r * bs + c] += acc * ieps;
    }
}

// ---- element volume blocks ---------------------------------------------------
template <int D, int P, class PH>
__global__ __launch_bounds__(256) void jac_elem_kernel(DevSpace S, PhysParams ph, const double* __restrict__ U,
                                                       double* __restrict__ J, int bslots) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int M = PH::M;
    constexpr int NBM = NB * M;
    __shared__ double sdFdu[NB][M][D][M];
    __shared__ double sdFdg[NB][M][D][M][D];
    __shared__ double sgx[NB][NB][D];  // [q][i][d] physical basis gradients
    __shared__ double sphi[NB][NB];    // [q][j] basis values
    __shared__ double sdw[NB];
    __shared__ double su[NB][M], sg[NB][M * D], sF[NB][M * D], seps[NB];  // state, gradient, flux, eps per point
    const int e = blockIdx.x;
    const int tid = threadIdx.x;
    const double* xs = S.tab;
    const double* L0 = S.tab + (N + 2);
    const double* wgl = L0 + 3 * (N + 2) * N;
    const double* Ue = U + (size_t)e * NBM;
    if (tid < NB) {
        const int q = tid;
        int qd[D], pt[D];
        {
            int r = q;
#pragma unroll
            for (int d = D - 1; d >= 0; --d) { qd[d] = r % N; r /= N; }
        }
        double xi[D], wq = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pt[d] = qd[d];
            xi[d] = xs[qd[d]];
            wq = (d == 0) ? wgl[qd[d]] : wq * wgl[qd[d]];
        }
        double X[1 << D][D];
        load_vertices<D>(S.X, e, X);
        double Jm[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, nullptr, Jm, Hm);
        const double det = det_adj<D>(Jm, A);
        const double detg = (det == 0.0) ? 1.0 : det;
        sdw[q] = fmax(0.0, det) * wq;
        double u[M] = {}, g[M * D] = {};
        for (int i = 0; i < NB; ++i) {
            double phi, gref[D];
            basis_at<D, N>(S.tab, pt, i, phi, gref, nullptr);
            sphi[q][i] = phi;
            double gxi[D];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double s = 0;
#pragma unroll
                for (int k = 0; k < D; ++k) s += gref[k] * A[k][j];
                gxi[j] = s / detg;
                sgx[q][i][j] = gxi[j];
            }
#pragma unroll
            for (int v = 0; v < M; ++v) {
                const double c = Ue[i * M + v];
                u[v] += c * phi;
#pragma unroll
                for (int j = 0; j < D; ++j) g[v * D + j] += c * gxi[j];
            }
        }
        double F[M * D];
        PH::flux(ph, u, g, F);
        double fro = 0;
#pragma unroll
        for (int k = 0; k < M * D; ++k) fro += F[k] * F[k];
        seps[q] = fmax(SQRT_EPS, SQRT_EPS * sqrt(fro));  // conservation_law.hpp:366-369
#pragma unroll
        for (int v = 0; v < M; ++v) su[q][v] = u[v];
#pragma unroll
        for (int k = 0; k < M * D; ++k) {
            sg[q][k] = g[k];
            sF[q][k] = F[k];
        }
    }
    __syncthreads();
    // one perturbed flux per (point, perturbation): M value perturbations,
    // then M * D gradient perturbations, spread over the workgroup
    constexpr int NPERT = M + M * D;
    for (int t = tid; t < NB * NPERT; t += blockDim.x) {
        const int q = t / NPERT, k = t % NPERT;
        double u[M], g[M * D], Fp[M * D];
#pragma unroll
        for (int v = 0; v < M; ++v) u[v] = su[q][v];
#pragma unroll
        for (int j = 0; j < M * D; ++j) g[j] = sg[q][j];
        const double eps = seps[q];
        const double ieps = 1.0 / eps;
        if (k < M) u[k] += eps;
        else g[k - M] += eps;
        PH::flux(ph, u, g, Fp);
        if (k < M) {
#pragma unroll
            for (int v = 0; v < M; ++v)
#pragma unroll
                for (int d = 0; d < D; ++d) sdFdu[q][v][d][k] = (Fp[v * D + d] - sF[q][v * D + d]) * ieps;
        } else {
            const int w = (k - M) / D, jd = (k - M) % D;
#pragma unroll
            for (int v = 0; v < M; ++v)
#pragma unroll
                for (int d = 0; d < D; ++d) sdFdg[q][v][d][w][jd] = (Fp[v * D + d] - sF[q][v * D + d]) * ieps;
        }
    }
    __syncthreads();
    const size_t bs = NBM;
    for (int ent = tid; ent < NBM * NBM; ent += blockDim.x) {
        const int r = ent / NBM, c = ent % NBM;
        const int i = r / M, v = r % M, j = c / M, w = c % M;
        double acc = 0;
        for (int q = 0; q < NB; ++q) {
            const double phij = sphi[q][j];
            double s = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double t2 = sdFdu[q][v][d][w] * phij;
#pragma unroll
                for (int ee = 0; ee < D; ++ee) t2 += sdFdg[q][v][d][w][ee] * sgx[q][j][ee];
                s += sgx[q][i][d] * t2;
            }
            acc += s * sdw[q];
        }
        // first writer of the diagonal block (the faces add to it afterwards)
        J[((size_t)e * bslots) * bs * bs + (size_t)r * bs + c] = acc;
    }
}

// zero the off-diagonal slots without a neighbour (boundary faces): the
// block products multiply them by zero, so they must hold finite values
__global__ void zero_missing_slots_kernel(int nelem, int bslots, size_t bsq, const int* __restrict__ cols,
                                          double* __restrict__ J) {
    const int e = blockIdx.x;
    for (int s = 1; s < bslots; ++s) {
        if (cols[(size_t)e * bslots + s] >= 0) continue;
        double* blk = J + ((size_t)e * bslots + s) * bsq;
        for (size_t k = threadIdx.x; k < bsq; k += blockDim.x) blk[k] = 0.0;
    }
}

__global__ void to_fp32_kernel(size_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

template <int D, int P, class PH>
int launch_jac(const DevSpace& S, const PhysParams& ph, const DiscDev& dd, const double* U, double* J,
               hipStream_t st) {
    const int bslots = 2 * D + 1;
    constexpr int NBM = Pow<P + 1, D>::value * PH::M;
    zero_missing_slots_kernel<<<S.nelem, 256, 0, st>>>(S.nelem, bslots, (size_t)NBM * NBM, S.jcols, J);
    // diagonal blocks: volume term first, then boundary faces and interior
    // faces one colour at a time (fixed summation order, no atomics)
    jac_elem_kernel<D, P, PH><<<S.nelem, 256, 0, st>>>(S, ph, U, J, bslots);
    for (int c = 0; c < S.ncol_b; ++c) {
        const int n = S.coff_b[c + 1] - S.coff_b[c];
        if (n > 0) jac_bface_kernel<D, P, PH><<<n, 256, 0, st>>>(S, ph, dd.bval, U, J, bslots, S.coff_b[c]);
    }
    for (int c = 0; c < S.ncol_i; ++c) {
        const int n = S.coff_i[c + 1] - S.coff_i[c];
        if (n > 0) jac_face_kernel<D, P, PH><<<n, 256, 0, st>>>(S, ph, U, J, bslots, S.coff_i[c]);
    }
    return (int)hipGetLastError();
}

template <class PH, int D>
int jac_order(int p, const DevSpace& S, const PhysParams& ph, const DiscDev& dd, const double* U, double* J,
              hipStream_t st) {
    switch (p) {
    case 1: return launch_jac<D, 1, PH>(S, ph, dd, U, J, st);
    case 2: return launch_jac<D, 2, PH>(S, ph, dd, U, J, st);
    case 3: return launch_jac<D, 3, PH>(S, ph, dd, U, J, st);
    case 4: return launch_jac<D, 4, PH>(S, ph, dd, U, J, st);
    default: return -1;
    }
}

// 3D: orders whose per-workgroup tables (the face kernel's flux
// differences for all 2 nb m columns, the element kernel's flux derivatives
// and basis gradients) fit the 160 KB of LDS of a workgroup: P1 and P2
template <class PH>
int jac_order3(int p, const DevSpace& S, const PhysParams& ph, const DiscDev& dd, const double* U, double* J,
               hipStream_t st) {
    switch (p) {
    case 1: return launch_jac<3, 1, PH>(S, ph, dd, U, J, st);
    case 2: return launch_jac<3, 2, PH>(S, ph, dd, U, J, st);
    default: return -2;
    }
}

int jacobian_fp64(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                  const double* U, double* J, size_t nvals, hipStream_t st) {
    (void)nvals;
    if (flux == ICE_FLUX_BURGERS) {
        if (ndim == 1) return jac_order<Burgers<1>, 1>(p, S, ph, dd, U, J, st);
        if (ndim == 2) return jac_order<Burgers<2>, 2>(p, S, ph, dd, U, J, st);
        if (ndim == 3) return jac_order3<Burgers<3>>(p, S, ph, dd, U, J, st);
    } else if (flux == ICE_FLUX_NAVIER_STOKES) {
        if (ndim == 1) return jac_order<NavierStokes<1, true>, 1>(p, S, ph, dd, U, J, st);
        if (ndim == 2) return jac_order<NavierStokes<2, true>, 2>(p, S, ph, dd, U, J, st);
        if (ndim == 3) return jac_order3<NavierStokes<3, true>>(p, S, ph, dd, U, J, st);
    } else if (flux == ICE_FLUX_EULER) {
        if (ndim == 1) return jac_order<NavierStokes<1, false>, 1>(p, S, ph, dd, U, J, st);
        if (ndim == 2) return jac_order<NavierStokes<2, false>, 2>(p, S, ph, dd, U, J, st);
        if (ndim == 3) return jac_order3<NavierStokes<3, false>>(p, S, ph, dd, U, J, st);
    }
    return -1;
}

int convert_fp32(size_t n, const double* a, float* b, hipStream_t st) {
    to_fp32_kernel<<<4096, 256, 0, st>>>(n, a, b);
    return (int)hipGetLastError();
}

}  // namespace ice
