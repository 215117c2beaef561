// DG residual kernels for CDNA4 (gfx950), FP64.
//
// One residual = three launches on the caller's stream:
//  1. trace kernel: per element and element face, the trace (value and the
//     element's share of the DDG gradient) at the face quadrature points,
//     from the LDS-staged coefficients by sum factorisation;
//  2. flux kernel (interior + boundary faces): one thread per face
//     quadrature point; the numerical fluxes and the DDG single-valued
//     gradient from the two stored traces, written as the weighted normal
//     flux Q[face][fq][v] (and the DDGIC lifting vector when sigma_ic != 0).
//     Together this is ConservationLawDDG::trace_integral / boundaryIntegral
//     (conservation_law.hpp:434-1176) up to the final multiplication by the
//     trace basis.
//  3. element kernel: a workgroup owns a tile of elements whose coefficients
//     are staged in LDS with one coalesced copy; phase 1 has one thread per
//     (element, volume quadrature point) and evaluates the physical flux
//     contracted into the reference frame; phase 2 has one thread per
//     (element, basis function) and sums the volume term
//     (domain_integral, :244-314), the source term and the lifted face fluxes
//     of the element's 2*ndim faces, then writes R with one coalesced store.
// Mixed precision (FP32 arithmetic) keeps each element's coefficients as FP32
// differences from the element's first coefficient, dU_i = U_i - U_0 (formed
// in FP64), with U_0 in FP64.  Derivatives see only dU (the basis gradients
// sum to zero), values are U_0 + sum dU_i phi_i with the sum in FP32 and the
// addition in FP64, and the face jumps uR - uL are formed in FP64 before the
// DDG penalty multiplies them by beta0 / h.  FP32 rounding then scales with
// the variation of u inside an element (~h |grad u|) instead of |u|, so the
// derivative, jump and penalty terms keep FP64-like accuracy.
// The face-flux buffer replaces the reference's scatter of per-face residual
// blocks (form_residual.hpp:137-242): every element gathers its own faces, so
// the result is deterministic and needs no atomics.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dg_common.h"
#include "fields.h"
#include "kernels.h"
#include "physics_dev.h"

// Minimum waves per SIMD requested from the register allocator for the
// residual kernels (build-time tuning knobs; 1 = no constraint).
#ifndef ICE_TRACE_MINW
#define ICE_TRACE_MINW 1
#endif
#ifndef ICE_FLUX_MINW
#define ICE_FLUX_MINW 1
#endif
// trace-kernel workgroup target (whole waves for one element when an
// element has more trace points)
#ifndef ICE_TRACE_TPB
#define ICE_TRACE_TPB 256
#endif
#ifndef ICE_ELEM_MINW
#define ICE_ELEM_MINW 1
#endif
#ifndef ICE_ELEM3_MINW
#define ICE_ELEM3_MINW 1
#endif
// target threads per workgroup of the 3D element kernel (whole elements per
// workgroup, rounded up to full waves)
#ifndef ICE_ELEM3_TPB
#define ICE_ELEM3_TPB 256
#endif

namespace ice {

// ---- split pipeline: traces per element face, then fluxes per face point ----
//
// trace_kernel: for every element and each of its 2*D faces, the trace at the
// face's quadrature points (in the element's own canonical order) is formed
// once from the LDS-staged coefficients and stored as
//   T[e][j][q'][0][v]      = u_v
//   T[e][j][q'][1 + d][v]  = g_d  (the element's share of the DDG gradient)
// where on interior faces g = grad(u)/2 + s beta1 h (Hess(u) n), s = -1 on
// the left and +1 on the right element, so that the face flux needs
// gL + gR = (grad uL + grad uR)/2 + beta1 h (Hess uR - Hess uL) n
// (conservation_law.hpp:529-552); on boundary faces g = grad(u).
// flux_kernel then evaluates the numerical fluxes at each face point from the
// two stored traces and the precomputed face geometry (normal, w sqrt(g), h).

// 1D tables in the kernels' arithmetic type
template <class VT>
__device__ __forceinline__ const VT* tab_of(const DevSpace& S);
template <>
__device__ __forceinline__ const double* tab_of<double>(const DevSpace& S) { return S.tab; }
template <>
__device__ __forceinline__ const float* tab_of<float>(const DevSpace& S) { return S.tabf; }

// Staged coefficient idx of the tile starting at U[base] ([element][node][var]):
// FP64 as is; FP32 as the difference from the element's first coefficient of
// the same variable (dU = U_i - U_0, the mixed-precision representation)
template <class VT, int NB, int M, class UT>
__device__ __forceinline__ VT coef_stage(const UT* __restrict__ U, size_t base, size_t total, int idx) {
    if (base + idx >= total) return VT(0);
    if constexpr (std::is_same<VT, double>::value) {
        return (VT)U[base + idx];
    } else {
        const int l = idx / (NB * M), v = idx % M;
        return (VT)((double)U[base + idx] - (double)U[base + (size_t)l * NB * M + v]);
    }
}

// Flux parameters of element e: a per-element viscosity (batched instances)
// replaces the discretization's when given
template <class P>
__device__ __forceinline__ P local_params(const P& p, const double* __restrict__ emu, int e) {
    P q = p;
    if (emu) q.mu = (decltype(q.mu))emu[e];
    return q;
}

// Minimum waves per SIMD of the 2D element kernel: the FP64 Burgers P1
// instantiation sits two registers above an occupancy step (82 VGPRs, 5
// waves) and fits 80 (6 waves) without scratch.  Caps that spill were
// slower: 3D Euler face fluxes at 6 waves (83 -> 80 VGPRs, 20 B of scratch:
// 0.375 -> 0.427 ms at 64^3 P3) and 2D Burgers P4 traces at 4 waves
// (140 -> 128 VGPRs, 32 B: the 512-instance P4 group 0.63 -> 0.66 ms).
template <int P, class PH>
constexpr int elem2_min_waves() {
    return (P == 1 && std::is_same<PH, Burgers<2, double>>::value) ? 6 : ICE_ELEM_MINW;
}

template <class PH, int D>
struct TraceRec {
    // Euler needs only values: its diffusive flux is identically zero
    static constexpr bool NEED_GRAD =
        !(std::is_same<PH, NavierStokes<D, false, double>>::value || std::is_same<PH, NavierStokes<D, false, float>>::value);
    static constexpr int NV = NEED_GRAD ? 1 + D : 1;
};

// Trace of one element at point pt of its face j by sum factorisation:
// the coefficients are first contracted along the face normal direction k
// with the 1D basis values, first and second derivatives at the face's
// endpoint, C_r[t] = sum_a U[a along k, t] L_r(+-1)[a], then evaluated along
// the tangential directions at the point.  Gives u, the physical gradient
// and the normal Hessian term hn (as trace_eval, one component at a time to
// keep the register footprint small), and stores the trace record of
// trace_kernel: out[v] = u, out[(1 + d) M + v] = gx/2 + bh hn on interior
// faces, gx on boundary faces.
template <int D, int N, int M, class VT>
__device__ __forceinline__ void face_trace_sf(const DevSpace& S, const VT* Ltab, const VT* Ue, int j, const int* pt,
                                              const double (&X)[1 << D][D], const double* nrm, bool interior,
                                              double bh, VT* __restrict__ out) {
    constexpr int NB = Pow<N, D>::value;
    constexpr int NT = NB / N;  // tangential tensor points
    constexpr int NH = D * (D + 1) / 2;
    const double* xs = S.tab;
    const VT* L0 = Ltab;
    const VT* L1 = L0 + (N + 2) * N;
    const VT* L2 = L1 + (N + 2) * N;
    const int k = j % D;
    const int ep = pt[k];
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = xs[pt[d]];
    double Jm[D][D], Hm[D][D][D], A[D][D];
    q1_map<D>(X, xi, nullptr, Jm, Hm);
    double det = det_adj<D>(Jm, A);
    det = (det == 0.0) ? 1.0 : det;
    const double idet = 1.0 / det;
    VT ln[3][N];  // normal-direction 1D rows at the endpoint
#pragma unroll
    for (int a = 0; a < N; ++a) {
        ln[0][a] = L0[ep * N + a];
        ln[1][a] = L1[ep * N + a];
        ln[2][a] = L2[ep * N + a];
    }
    // tangential dims in ascending order and their 1D rows at the point
    int td[D > 1 ? D - 1 : 1];
    {
        int c = 0;
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (d != k) td[c++] = d;
    }
    // strides of the element's tensor index: dimension 0 slowest
    int stride[D];
    {
        int st = 1;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) { stride[d] = st; st *= N; }
    }
#pragma unroll 1
    for (int v = 0; v < M; ++v) {
        VT C[3][NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            // tangential multi-index of t (first tangential dim slowest)
            int base = 0;
            {
                int r = t;
#pragma unroll
                for (int c = D - 2; c >= 0; --c) {
                    base += (r % N) * stride[td[c]];
                    r /= N;
                }
            }
            VT c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
            for (int a = 0; a < N; ++a) {
                const VT uv = Ue[(base + a * stride[k]) * M + v];
                c0 += uv * ln[0][a];
                c1 += uv * ln[1][a];
                c2 += uv * ln[2][a];
            }
            C[0][t] = c0;
            C[1][t] = c1;
            C[2][t] = c2;
        }
        // reference derivatives: gr[d], hr[(a,b) a <= b]
        double uv, gr[D], hr[NH];
        if constexpr (D == 1) {
            uv = C[0][0];
            gr[0] = C[1][0];
            hr[0] = C[2][0];
        } else if constexpr (D == 2) {
            const int p1 = pt[td[0]];
            VT s0 = 0, sk = 0, s1 = 0, skk = 0, sk1 = 0, s11 = 0;
#pragma unroll
            for (int b = 0; b < N; ++b) {
                const VT l0 = L0[p1 * N + b], l1 = L1[p1 * N + b], l2 = L2[p1 * N + b];
                s0 += C[0][b] * l0;
                sk += C[1][b] * l0;
                s1 += C[0][b] * l1;
                skk += C[2][b] * l0;
                sk1 += C[1][b] * l1;
                s11 += C[0][b] * l2;
            }
            uv = s0;
            // hr order (0,0), (0,1), (1,1)
            const bool k0 = (k == 0);
            gr[0] = k0 ? sk : s1;
            gr[1] = k0 ? s1 : sk;
            hr[0] = k0 ? skk : s11;
            hr[1] = sk1;
            hr[2] = k0 ? s11 : skk;
        } else {
            const int p1 = pt[td[0]], p2 = pt[td[1]];
            VT s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // 0, k, 1, 2, kk, k1, k2, 11, 12, 22
#pragma unroll
            for (int b1 = 0; b1 < N; ++b1) {
                const VT a0 = L0[p1 * N + b1], a1 = L1[p1 * N + b1], a2 = L2[p1 * N + b1];
                VT r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};  // [C index][deriv along dim2]
#pragma unroll
                for (int b2 = 0; b2 < N; ++b2) {
                    const VT c0 = L0[p2 * N + b2], c1 = L1[p2 * N + b2], c2 = L2[p2 * N + b2];
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        const VT cv = C[r][b1 * N + b2];
                        r0[r] += cv * c0;
                        r1[r] += cv * c1;
                        r2[r] += cv * c2;
                    }
                }
                s[0] += r0[0] * a0;
                s[1] += r0[1] * a0;
                s[2] += r0[0] * a1;
                s[3] += r1[0] * a0;
                s[4] += r0[2] * a0;
                s[5] += r0[1] * a1;
                s[6] += r1[1] * a0;
                s[7] += r0[0] * a2;
                s[8] += r1[0] * a1;
                s[9] += r2[0] * a0;
            }
            uv = s[0];
            // (k, o1, o2) = (0,1,2), (1,0,2) or (2,0,1); hr order 00 01 02 11 12 22
            if (k == 0) {
                gr[0] = s[1]; gr[1] = s[2]; gr[2] = s[3];
                hr[0] = s[4]; hr[1] = s[5]; hr[2] = s[6]; hr[3] = s[7]; hr[4] = s[8]; hr[5] = s[9];
            } else if (k == 1) {
                gr[1] = s[1]; gr[0] = s[2]; gr[2] = s[3];
                hr[3] = s[4]; hr[1] = s[5]; hr[4] = s[6]; hr[0] = s[7]; hr[2] = s[8]; hr[5] = s[9];
            } else {
                gr[2] = s[1]; gr[0] = s[2]; gr[1] = s[3];
                hr[5] = s[4]; hr[2] = s[5]; hr[4] = s[6]; hr[0] = s[7]; hr[1] = s[8]; hr[3] = s[9];
            }
        }
        // physical gradient and normal Hessian term (finite_element.hpp:220-258, 328-398)
        double gx[D];
#pragma unroll
        for (int jj = 0; jj < D; ++jj) {
            double sacc = 0;
#pragma unroll
            for (int kk = 0; kk < D; ++kk) sacc += gr[kk] * A[kk][jj];
            gx[jj] = sacc * idet;
        }
        double rhs[D][D];
        {
            int h = 0;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = a; b < D; ++b) {
                    rhs[a][b] = hr[h];
                    rhs[b][a] = hr[h];
                    ++h;
                }
        }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b)
#pragma unroll
                for (int kk = 0; kk < D; ++kk) rhs[a][b] -= Hm[kk][a][b] * gx[kk];
        double Tm[D][D];
#pragma unroll
        for (int l = 0; l < D; ++l)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                double sacc = 0;
#pragma unroll
                for (int kk = 0; kk < D; ++kk) sacc += rhs[l][kk] * A[kk][b];
                Tm[l][b] = sacc;
            }
        const double idet2 = idet * idet;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double sacc = 0;
#pragma unroll
            for (int jj = 0; jj < D; ++jj) {
                double hx = 0;
#pragma unroll
                for (int l = 0; l < D; ++l) hx += A[l][jj] * Tm[l][d];
                sacc += hx * nrm[jj];
            }
            const double hn = sacc * idet2;
            out[(1 + d) * M + v] = (VT)(interior ? 0.5 * gx[d] + bh * hn : gx[d]);
        }
        out[v] = (VT)uv;
    }
}

// Trace record of one element at its tensor point pt on its face j (Ue the
// element's coefficients, L the 1D value / derivative tables): u only when
// the flux has no diffusive part, else as face_trace_sf.
template <int D, int N, int M, class VT, bool GRAD>
__device__ __forceinline__ void element_point_trace(const DevSpace& S, const VT* L, const VT* Ue, int j, const int* pt,
                                                    const double (&X)[1 << D][D], const double* nrm, bool interior,
                                                    double bh, VT* __restrict__ out) {
    constexpr int NB = Pow<N, D>::value;
    if constexpr (!GRAD) {
        // The nodes include the interval ends, so along the face normal only
        // the node on the face has a nonzero basis value at the face (the
        // others vanish at -1 exactly and are O(1e-16) at +1, where the
        // cumulative equispaced nodes end one ulp short of 1): the sum runs
        // over the N^(D-1) face nodes.
        constexpr int NFN = NB / N;
        const int kn = j % D;
        const int inode = (j < D) ? 0 : N - 1;
        VT a0[D][N];
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int i = 0; i < N; ++i) a0[d][i] = L[pt[d] * N + i];
        VT u[M];
#pragma unroll
        for (int v = 0; v < M; ++v) u[v] = 0;
        // the normal axis as a compile-time constant, so every node offset
        // but the face's is an immediate
        auto face_sum = [&](auto KN) {
            constexpr int K = decltype(KN)::value;
            constexpr int SK = Pow<N, D - 1 - K>::value;  // node stride along the normal axis
            const VT en = L[pt[K] * N + inode];
            const VT* Ub = Ue + (size_t)inode * SK * M;
            if constexpr (D == 3) {
                // one tangential index per pass (a fully unrolled face issues
                // all its loads at once and halves the occupancy)
                constexpr int TA = (K == 0) ? 1 : 0, TB = (K == 2) ? 1 : 2;
                constexpr int SA = Pow<N, 2 - TA>::value, SB = Pow<N, 2 - TB>::value;
#pragma unroll 1
                for (int ia = 0; ia < N; ++ia) {
                    const VT ca = en * L[pt[TA] * N + ia];
                    const VT* Ua = Ub + ia * SA * M;
#pragma unroll
                    for (int ib = 0; ib < N; ++ib) {
                        const VT c0 = ca * a0[TB][ib];
#pragma unroll
                        for (int v = 0; v < M; ++v) u[v] += Ua[ib * SB * M + v] * c0;
                    }
                }
                return;
            }
#pragma unroll
            for (int t = 0; t < NFN; ++t) {
                // t enumerates the tangential indices in tensor order
                int node = 0, r = t;
                VT c0 = en;
#pragma unroll
                for (int d = D - 1; d >= 0; --d) {
                    if (d == K) continue;
                    const int id = r % N;
                    r /= N;
                    int sd = 1;
#pragma unroll
                    for (int dd = d + 1; dd < D; ++dd) sd *= N;
                    node += id * sd;
                    c0 *= a0[d][id];
                }
#pragma unroll
                for (int v = 0; v < M; ++v) u[v] += Ub[node * M + v] * c0;
            }
        };
        if (kn == 0) face_sum(std::integral_constant<int, 0>{});
        if constexpr (D > 1) if (kn == 1) face_sum(std::integral_constant<int, (D > 1 ? 1 : 0)>{});
        if constexpr (D > 2) if (kn == 2) face_sum(std::integral_constant<int, (D > 2 ? 2 : 0)>{});
#pragma unroll
        for (int v = 0; v < M; ++v) out[v] = u[v];
    } else {
        face_trace_sf<D, N, M, VT>(S, L, Ue, j, pt, X, nrm, interior, bh, out);
    }
}

template <int D, int P, class PH, int TPB, class UT = double>
__global__ __launch_bounds__(TPB, ICE_TRACE_MINW) void trace_kernel(DevSpace S, typename PH::Params ph,
                                                                     const UT* __restrict__ U,
                                                                     typename PH::real* __restrict__ T) {
    using VT = typename PH::real;
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    constexpr int TPE = 2 * D * NFQ;                       // trace points per element
    static_assert(TPB >= TPE, "one thread per trace point of an element");
    constexpr int EPB = TPB / TPE;  // elements per block
    constexpr bool GRAD = TraceRec<PH, D>::NEED_GRAD;
    constexpr int NV = TraceRec<PH, D>::NV;
    __shared__ VT sU[EPB * NB * M];
    __shared__ VT sL[3 * (N + 2) * N];  // 1D values, first and second derivatives
    __shared__ VT sOut[EPB * TPE * NV * M];  // the tile's trace records, written out contiguously
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / TPE;
    const int t = threadIdx.x - le * TPE;
    const int e = e0 + le;
    const bool active = threadIdx.x < EPB * TPE && e < S.nelem;
    const int j = t / NFQ, qp = t - j * NFQ;
    // prologue: all global loads (coefficients, tables, this point's face
    // data and element vertices) issued before the first barrier
    {
        const size_t base = (size_t)e0 * NB * M;
        const size_t total = (size_t)S.nelem * NB * M;
        for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x)
            sU[idx] = coef_stage<VT, NB, M, UT>(U, base, total, idx);
        const VT* tv = tab_of<VT>(S) + (N + 2);
        for (int idx = threadIdx.x; idx < 3 * (N + 2) * N; idx += blockDim.x) sL[idx] = tv[idx];
    }
    int pt[D] = {};
    double X[1 << D][D];
    double nrm[D] = {}, h = 1.0;
    int ef = -1, q = qp;
    if (active) {
        const signed char* fp = S.fpt + ((j * S.norient + 0) * NFQ + qp) * D;
#pragma unroll
        for (int d = 0; d < D; ++d) pt[d] = fp[d];
        if constexpr (GRAD) {
            load_vertices<D>(S.X, e, X);
            ef = S.efaces[(size_t)e * 2 * D + j];
            // the face point q of this element point qp (inverse orientation permutation)
            if (ef >= 0 && (ef >> 4) < S.nfi) q = S.finv[(j * S.norient + (ef & 7)) * NFQ + qp];
            if (ef >= 0) {
                const double* g = S.fgeo + ((size_t)(ef >> 4) * NFQ + q) * (D + 2);
#pragma unroll
                for (int d = 0; d < D; ++d) nrm[d] = g[d];
                h = g[D + 1];
            }
        }
    }
    __syncthreads();
    if (active) {
        const bool interior = ef >= 0 && (ef >> 4) < S.nfi;
        const int side = (ef >> 3) & 1;
        const double bh = (side ? 1.0 : -1.0) * (double)ph.beta1 * h;
        element_point_trace<D, N, M, VT, GRAD>(S, sL, sU + le * NB * M, j, pt, X, nrm, interior, bh,
                                               sOut + (size_t)threadIdx.x * NV * M);
    }
    __syncthreads();
    // records of elements e0 .. e0 + ne - 1 are one contiguous range of T
    const int ne = min(EPB, S.nelem - e0);
    VT* dst = T + (size_t)e0 * TPE * NV * M;
    for (int idx = threadIdx.x; idx < ne * TPE * NV * M; idx += blockDim.x) dst[idx] = sOut[idx];
}

// 3D P3 traces of fluxes without a gradient, two points of a face row per
// thread: along the face normal only the face's N^2 nodes contribute, and the
// contraction along the face's first tangential axis, t[ib] = sum_ia
// l_ia(s_a) U[ia][ib], is shared by the points of a row (canonical face
// points run row by row).  Per point: 40 LDS reads and 50 FMA per variable
// set instead of 80 and 80 with one point per thread.  The element tile's
// coefficients and its trace records share one LDS buffer.
template <class PH, int EPB, class UT = double>
__global__ __launch_bounds__(((EPB * 48 + 63) / 64) * 64) void trace_kernel_rows(DevSpace S, typename PH::Params ph,
                                                                                 const UT* __restrict__ U,
                                                                                 typename PH::real* __restrict__ T) {
    using VT = typename PH::real;
    constexpr int D = 3, N = 4, NN = 16, NB = 64, NFQ = 16;
    constexpr int M = PH::M;
    constexpr int TPE = 2 * D * NN / 2;  // threads per element: two points each
    constexpr int PTE = 2 * D * NFQ;     // trace points per element
    static_assert(!TraceRec<PH, D>::NEED_GRAD, "fluxes without a gradient");
    constexpr int BUF = (EPB * PTE * M > EPB * NB * M) ? EPB * PTE * M : EPB * NB * M;
    __shared__ VT sbuf[BUF];
    __shared__ VT sL[(N + 2) * N];
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / TPE;
    const int t = threadIdx.x - le * TPE;
    const int e = e0 + le;
    const bool active = threadIdx.x < EPB * TPE && e < S.nelem;
    const int j = t / (NN / 2), rem = t - j * (NN / 2);
    const int s0 = rem / 2, s1 = 2 * (rem - s0 * 2);  // canonical points q0 = s0 N + s1, q0 + 1
    {
        const size_t base = (size_t)e0 * NB * M;
        const size_t total = (size_t)S.nelem * NB * M;
        for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x)
            sbuf[idx] = coef_stage<VT, NB, M, UT>(U, base, total, idx);
        const VT* tv = tab_of<VT>(S) + (N + 2);
        for (int idx = threadIdx.x; idx < (N + 2) * N; idx += blockDim.x) sL[idx] = tv[idx];
    }
    int pa[D] = {}, pb[D] = {};
    if (active) {
        const signed char* fp = S.fpt + ((size_t)(j * S.norient + 0) * NFQ + s0 * N + s1) * D;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pa[d] = fp[d];
            pb[d] = fp[D + d];
        }
    }
    __syncthreads();
    VT out[2][M];
    if (active) {
        const int kn = j % D;
        const int inode = (j < D) ? 0 : N - 1;
        const VT* Ue = sbuf + (size_t)le * NB * M;
        auto row = [&](auto KN) {
            constexpr int K = decltype(KN)::value;
            constexpr int TA = (K == 0) ? 1 : 0, TB = (K == 2) ? 1 : 2;
            constexpr int SK = Pow<N, 2 - K>::value, SA = Pow<N, 2 - TA>::value, SB = Pow<N, 2 - TB>::value;
            const VT en = sL[pa[K] * N + inode];
            VT la[N];
#pragma unroll
            for (int i = 0; i < N; ++i) la[i] = sL[pa[TA] * N + i] * en;
            const VT* Ub = Ue + inode * SK * M;
#pragma unroll
            for (int v = 0; v < M; ++v) out[0][v] = out[1][v] = VT(0);
            // one tangential column per pass (unrolled, the face's 80 values
            // would be loaded at once and halve the occupancy)
#pragma unroll 1
            for (int ib = 0; ib < N; ++ib) {
                const VT* Uc = Ub + ib * SB * M;
                VT tv[M];
#pragma unroll
                for (int v = 0; v < M; ++v) tv[v] = VT(0);
#pragma unroll
                for (int ia = 0; ia < N; ++ia)
#pragma unroll
                    for (int v = 0; v < M; ++v) tv[v] += Uc[ia * SA * M + v] * la[ia];
                const VT b0 = sL[pa[TB] * N + ib], b1 = sL[pb[TB] * N + ib];
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    out[0][v] += tv[v] * b0;
                    out[1][v] += tv[v] * b1;
                }
            }
        };
        if (kn == 0) row(std::integral_constant<int, 0>{});
        else if (kn == 1) row(std::integral_constant<int, 1>{});
        else row(std::integral_constant<int, 2>{});
    }
    __syncthreads();  // the coefficients are read: the buffer takes the records
    if (active) {
        VT* o = sbuf + ((size_t)le * PTE + j * NFQ + s0 * N + s1) * M;
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int v = 0; v < M; ++v) o[k * M + v] = out[k][v];
    }
    __syncthreads();
    // records of elements e0 .. e0 + ne - 1 are one contiguous range of T
    const int ne = min(EPB, S.nelem - e0);
    VT* dst = T + (size_t)e0 * PTE * M;
    for (int idx = threadIdx.x; idx < ne * PTE * M; idx += blockDim.x) dst[idx] = sbuf[idx];
}

// Numerical flux (and DDGIC lifting vector) at point q of interior face f
// from the two trace records of the point (left element tL, right tR), ph
// the left element's flux parameters.
// g: the point's unit normal, w sqrt(g) and DDG distance h (a row of the
// face geometry table, or the same quantities formed from the vertices).
template <int D, class PH, class UT>
__device__ __forceinline__ void interior_face_point(const DevSpace& S, const typename PH::Params& ph, const DiscDev& dd,
                                                    int f, int q, int NFQ, const typename PH::real* tL,
                                                    const typename PH::real* tR, const UT* U0L,
                                                    const UT* U0R, const double* g) {
    using VT = typename PH::real;
    constexpr int M = PH::M;
    constexpr int NV = TraceRec<PH, D>::NV;
    constexpr bool DELTA = std::is_same<VT, float>::value;  // FP32 traces hold u - U_0
    VT* const Qb = reinterpret_cast<VT*>(dd.Q);
    VT* const Pb = reinterpret_cast<VT*>(dd.Pic);
    VT nrm[D];
#pragma unroll
    for (int d = 0; d < D; ++d) nrm[d] = (VT)g[d];
    const VT ws = (VT)g[D], h = (VT)g[D + 1];
    VT uL[M], uR[M], gs[M * D], jmp[M];
#pragma unroll
    for (int v = 0; v < M; ++v) {
        if constexpr (DELTA) {
            const double a = (double)U0L[v] + (double)tL[v], b = (double)U0R[v] + (double)tR[v];
            uL[v] = (VT)a;
            uR[v] = (VT)b;
            jmp[v] = (VT)(b - a);
        } else {
            uL[v] = tL[v];
            uR[v] = tR[v];
            jmp[v] = uR[v] - uL[v];
        }
    }
    if constexpr (NV > 1) {
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int d = 0; d < D; ++d) gs[v * D + d] = tL[(1 + d) * M + v] + tR[(1 + d) * M + v];
    } else {
#pragma unroll
        for (int k = 0; k < M * D; ++k) gs[k] = VT(0);
    }
    VT Qv[M];
    interior_point_flux_split<D, PH>(ph, uL, uR, gs, nrm, h, ws, Qv, jmp);
    VT* Qf = Qb + ((size_t)f * NFQ + q) * M;
#pragma unroll
    for (int v = 0; v < M; ++v) Qf[v] = Qv[v];
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (dd.Pic) {
            VT uavg[M];
#pragma unroll
            for (int v = 0; v < M; ++v) uavg[v] = VT(0.5) * (uL[v] + uR[v]);
            ddgic_lift<D, PH>(ph, ph.sigma_face, uavg, nrm, jmp, ws, Pb + ((size_t)f * NFQ + q) * M * D);
        }
    }
}

// Numerical flux (and DDGIC lifting vector) at point q of boundary face b
// from the interior element's trace record tL (not read on Neumann faces).
template <int D, class PH, class UT>
__device__ __forceinline__ void boundary_face_point(const DevSpace& S, const typename PH::Params& ph, const DiscDev& dd,
                                                    int b, int q, int NFQ, int bctype, int bcflag,
                                                    const typename PH::real* tL, const UT* U0L) {
    using VT = typename PH::real;
    constexpr int M = PH::M;
    constexpr int NV = TraceRec<PH, D>::NV;
    constexpr bool DELTA = std::is_same<VT, float>::value;  // FP32 traces hold u - U_0
    VT* const Qb = reinterpret_cast<VT*>(dd.Q);
    VT* const Pb = reinterpret_cast<VT*>(dd.Pic);
    const size_t slot = (size_t)S.nfi + b;
    VT* Qf = Qb + (slot * NFQ + q) * M;
    VT* Pf = nullptr;
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (dd.Pic) {
            Pf = Pb + (slot * NFQ + q) * M * D;
#pragma unroll
            for (int k = 0; k < M * D; ++k) Pf[k] = VT(0);
        }
    }
    const bool general = !(bctype == ICE_BC_DIRICHLET || bctype == ICE_BC_NEUMANN ||
                           bctype == ICE_BC_EXTRAPOLATION || bctype == ICE_BC_SPACETIME_FUTURE);
    if (general && !PH::IMPLEMENTS_BCS) {
#pragma unroll
        for (int v = 0; v < M; ++v) Qf[v] = VT(0);
        return;
    }
    const double* g = S.fgeo + (slot * NFQ + q) * (D + 2);
    VT nrm[D];
#pragma unroll
    for (int d = 0; d < D; ++d) nrm[d] = (VT)g[d];
    const VT ws = (VT)g[D], hb = (VT)g[D + 1];
    VT uL[M] = {}, gL[M * D] = {}, jmpD[M] = {};
    if (bctype != ICE_BC_NEUMANN) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
            if constexpr (DELTA) {
                const double a = (double)U0L[v] + (double)tL[v];
                uL[v] = (VT)a;
                jmpD[v] = (VT)(dd.bval[((size_t)b * NFQ + q) * M + v] - a);
            } else {
                uL[v] = tL[v];
            }
        }
        if constexpr (NV > 1) {
#pragma unroll
            for (int v = 0; v < M; ++v)
#pragma unroll
                for (int d = 0; d < D; ++d) gL[v * D + d] = tL[(1 + d) * M + v];
        }
    }
    VT bv[M];
#pragma unroll
    for (int v = 0; v < M; ++v) bv[v] = (VT)dd.bval[((size_t)b * NFQ + q) * M + v];
    VT Qv[M], uB[M];
    boundary_point_flux<D, PH>(ph, bctype, bcflag, uL, gL, bv, nrm, hb, ws, Qv, uB, DELTA ? jmpD : nullptr);
#pragma unroll
    for (int v = 0; v < M; ++v) Qf[v] = Qv[v];
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (Pf && bctype == ICE_BC_DIRICHLET) {
            VT uavg[M], jb[M];
#pragma unroll
            for (int v = 0; v < M; ++v) {
                uavg[v] = VT(0.5) * (uL[v] + uB[v]);
                jb[v] = DELTA ? jmpD[v] : uB[v] - uL[v];
            }
            ddgic_lift<D, PH>(ph, ph.sigma_bdy, uavg, nrm, jb, ws, Pf);
        }
    }
}

// Numerical fluxes at face points from stored traces: interior points first
// (blocks [0, nbi)), then boundary points.  BDY = false compiles the boundary
// branch out (meshes without boundary faces): the boundary-condition code
// holds more registers than the interior flux (Euler: 75 against 50 VGPRs).
template <int D, int P, class PH, bool BDY, class UT = double>
__global__ __launch_bounds__(256, ICE_FLUX_MINW) void flux_kernel(DevSpace S, typename PH::Params ph0, DiscDev dd,
                                                                   const typename PH::real* __restrict__ T, int nbi,
                                                                   const UT* __restrict__ U) {
    using VT = typename PH::real;
    constexpr int N = P + 1;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int NBM = Pow<N, D>::value * PH::M;  // one element's coefficients (U_0 = the first M)
    constexpr int M = PH::M;
    constexpr int NV = TraceRec<PH, D>::NV;
    if ((int)blockIdx.x < nbi) {
        const int tid = blockIdx.x * blockDim.x + threadIdx.x;
        const int f = tid / NFQ, q = tid - f * NFQ;
        if (f >= S.nfi) return;
        const int4 rec = S.fint[f];
        const int eL = rec.x, eR = rec.y, info = rec.z;
        const int fnl = info_fnl(info), fnr = info_fnr(info), orient = info_orient(info);
        const int qr = S.fperm[(fnr * S.norient + orient) * NFQ + q];
        const VT* tL = T + (((size_t)eL * 2 * D + fnl) * NFQ + q) * NV * M;
        const VT* tR = T + (((size_t)eR * 2 * D + fnr) * NFQ + qr) * NV * M;
        // 3D fluxes without a diffusive part need only the normal and w sqrt(g),
        // formed here from the left element's face vertices (96 B per face,
        // shared by its points) rather than read from the table (40 B per point)
        double geo[D + 2];
        if constexpr (D == 3 && NV == 1) {
            double FX[1 << (D - 1)][D], x[D];
            load_face_vertices<D>(S, eL, fnl, FX);
            const double sg = face_geometry_fv<D, N>(S, FX, fnl, q, x, geo);
            geo[D] = face_weight<N, D>(S, q) * sg;
            geo[D + 1] = 1.0;  // the DDG distance multiplies the (zero) diffusive flux only
        } else {
            const double* g = S.fgeo + ((size_t)f * NFQ + q) * (D + 2);
#pragma unroll
            for (int k = 0; k < D + 2; ++k) geo[k] = g[k];
        }
        interior_face_point<D, PH, UT>(S, local_params(ph0, dd.emu, eL), dd, f, q, NFQ, tL, tR, U + (size_t)eL * NBM,
                                   U + (size_t)eR * NBM, geo);
    } else if constexpr (BDY) {
        const int tid = (blockIdx.x - nbi) * blockDim.x + threadIdx.x;
        const int b = tid / NFQ, q = tid - b * NFQ;
        if (b >= S.nfb) return;
        const int4 rec = S.fbdy[b];
        const int eL = rec.x, info = rec.y;
        const VT* tL = T + (((size_t)eL * 2 * D + info_fnl(info)) * NFQ + q) * NV * M;
        boundary_face_point<D, PH, UT>(S, local_params(ph0, dd.emu, eL), dd, b, q, NFQ, info_bctype(info), rec.z, tL,
                                   U + (size_t)eL * NBM);
    }
}

// Face geometry table (u-independent): per face row and face point the unit
// normal, w sqrt(g) and the guarded DDG distance h (conservation_law.hpp:
// 484-527 interior, 700-713 boundary).
template <int D, int P>
__global__ __launch_bounds__(256) void fgeo_kernel(DevSpace S, double* __restrict__ out) {
    constexpr int N = P + 1;
    constexpr int NFQ = Pow<N, D - 1>::value;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = tid / NFQ, q = tid - row * NFQ;
    if (row >= S.nfi + S.nfb) return;
    double XL[1 << D][D], x[D], nrm[D], cL[D], h, sqrtg;
    if (row < S.nfi) {
        const int4 rec = S.fint[row];
        double XR[1 << D][D], cR[D];
        load_vertices<D>(S.X, rec.x, XL);
        load_vertices<D>(S.X, rec.y, XR);
        sqrtg = face_geometry<D, N>(S, XL, info_fnl(rec.z), q, x, nrm);
        centroid<D>(XL, cL);
        centroid<D>(XR, cR);
        const int pdim = info_pdim(rec.z);
        if (pdim) cR[pdim - 1] += S.period[pdim - 1];
        h = ddg_h_interior<D>(nrm, x, cL, cR);
    } else {
        const int4 rec = S.fbdy[row - S.nfi];
        load_vertices<D>(S.X, rec.x, XL);
        sqrtg = face_geometry<D, N>(S, XL, info_fnl(rec.y), q, x, nrm);
        centroid<D>(XL, cL);
        h = ddg_h_boundary<D>(nrm, x, cL);
    }
    double* o = out + ((size_t)row * NFQ + q) * (D + 2);
#pragma unroll
    for (int d = 0; d < D; ++d) o[d] = nrm[d];
    o[D] = face_weight<N, D>(S, q) * sqrtg;
    o[D + 1] = h;
}

template <int D, int P>
int launch_fgeo(const DevSpace& S, double* out, hipStream_t st) {
    constexpr int NFQ = Pow<P + 1, D - 1>::value;
    const long nt = (long)(S.nfi + S.nfb) * NFQ;
    if (nt > 0) fgeo_kernel<D, P><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(S, out);
    return (int)hipGetLastError();
}

int face_geometry_table(int ndim, int p, const DevSpace& S, double* out, hipStream_t st) {
#define ICE_GCASE(DD, PP) if (ndim == DD && p == PP) return launch_fgeo<DD, PP>(S, out, st);
    ICE_GCASE(1, 1) ICE_GCASE(1, 2) ICE_GCASE(1, 3) ICE_GCASE(1, 4)
    ICE_GCASE(2, 1) ICE_GCASE(2, 2) ICE_GCASE(2, 3) ICE_GCASE(2, 4)
    ICE_GCASE(3, 1) ICE_GCASE(3, 2) ICE_GCASE(3, 3) ICE_GCASE(3, 4)
#undef ICE_GCASE
    return -1;
}

size_t trace_buffer_size(int ndim, int p, int flux, int nelem) {
    int nfq = 1;
    for (int d = 0; d < ndim - 1; ++d) nfq *= p + 1;
    const int m = flux == ICE_FLUX_BURGERS ? 1 : ndim + 2;
    const int nv = flux == ICE_FLUX_EULER ? 1 : 1 + ndim;
    return (size_t)nelem * 2 * ndim * nfq * nv * m;
}

// ---- elements: volume integral + face lifting -----------------------------
// Whole elements per workgroup for a target of TPB threads, rounded up to
// full waves
template <int D, int P, int TPB = 256>
struct ElemTile {
    static constexpr int N = P + 1;
    static constexpr int NB = Pow<N, D>::value;
    static constexpr int EPB = (TPB / NB) > 0 ? (TPB / NB) : 1;
    static constexpr int NT = EPB * NB;
    static constexpr int BLOCK = ((NT + 63) / 64) * 64;
};
// workgroup size targets selected by DiscDev::tile (TILE_MFMA: the matrix-core
// 3D P3 kernel of the mixed residual where it applies, else the default tile)
constexpr int tile_tpb(int tile) { return tile == 1 ? 128 : tile == 2 ? 64 : 256; }
constexpr int TILE_MFMA = 3;

template <int D, int P, class PH>
__global__ __launch_bounds__(256) void element_kernel(DevSpace S, PhysParams ph0, const double* __restrict__ emu,
                                                      const double* __restrict__ srcload,
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
    const PhysParams ph = local_params(ph0, emu, active ? e : 0);
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
                    double gphi[D];
                    ic_test_gradient<D, N>(S.tab, X, fp + fq * D, kd, ph.ic_mirror != 0, gphi);
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

// 2D element kernel by sum factorisation.  Thread k of an element is the
// tensor pair (k0, k1) (dimension 0 slowest), taken in four phases:
//  1a (a, q1):  T0[a][q1] = sum_b U[a][b] l_b(x_q1),  T1 = sum_b U[a][b] l_b'(x_q1)
//  1b (q0, q1): u, du/dxi from T, geometry, physical flux -> reference-frame flux G
//  2a (q0, i1): S0 = sum_q1 G0[q0][q1] l_i1(x_q1),  S1 = sum_q1 G1[q0][q1] l_i1'(x_q1)
//  2b (i0, i1): sum_q0 S0[q0][i1] l_i0'(x_q0) + S1[q0][i1] l_i0(x_q0), minus the
//               source load, plus the lifted face fluxes; R written once.
// Same terms as element_kernel (domain_integral, conservation_law.hpp:244-314,
// and the trace_integral lifting), 2N instead of N^2 multiply-adds per
// point and basis function.  The lifting skips faces on which the basis
// function's node does not lie (its trace there is zero).
template <int P, class PH, int TPB, class UT = double>
__global__ __launch_bounds__((ElemTile<2, P, TPB>::BLOCK), (elem2_min_waves<P, PH>())) void element_kernel_2d(DevSpace S, typename PH::Params ph0,
                                                                        const double* __restrict__ emu,
                                                                        const double* __restrict__ srcload,
                                                                        const UT* __restrict__ U,
                                                                        const typename PH::real* __restrict__ Q,
                                                                        const typename PH::real* __restrict__ Pic,
                                                                        double* __restrict__ R,
                                                         unsigned long long* __restrict__ wave,
                                                         double* __restrict__ wave_part) {
    constexpr int D = 2;
    using TL = ElemTile<D, P, TPB>;
    constexpr int N = TL::N, NB = TL::NB, EPB = TL::EPB;
    constexpr int NFQ = N;
    constexpr int M = PH::M;
    using VT = typename PH::real;
    // The staged coefficients are dead after phase 1a, so the reference-frame
    // fluxes G of phase 1b reuse their storage sUG.  FP64 (LOWLDS): sUG also
    // holds the phase-1a derivative contractions T1 (second half) until phase
    // 1b has read them, sT holds T0 then S0, the second test contraction is
    // summed from G in phase 2b and the source load is read from global
    // memory (34 KB for a 16-element P3 workgroup instead of 50 KB).  FP32:
    // sT holds T0, T1 then S0, S1 and the source load is staged.
    // FP64: node strides are padded to an odd number of entries and element
    // strides by one more, so that the lanes of a phase, which differ in
    // node and element, fall on different LDS banks (power-of-two strides
    // put every element's row on the same bank; 2D NS 128 x 128 P3: 21.4 ->
    // 20.1 us).  The
    // FP32 tiles measured slightly slower padded and keep the dense layout.
    constexpr bool LOWLDS = std::is_same<VT, double>::value;
    constexpr bool PAD = LOWLDS;
    constexpr int MS = (PAD && M % 2 == 0) ? M + 1 : M;              // node stride of sU, T0, T1
    constexpr int GS = (PAD && (M * D) % 2 == 0) ? M * D + 1 : M * D;  // node stride of G
    constexpr int ESU = NB * MS + (PAD ? 1 : 0), ESG = NB * GS + (PAD ? 1 : 0);  // element strides
    constexpr int UGN = (LOWLDS ? 2 * ESU : ESU) > ESG ? (LOWLDS ? 2 * ESU : ESU) : ESG;
    __shared__ VT sUG[EPB * UGN];
    __shared__ VT sT[EPB * (LOWLDS ? 1 : 2) * ESU];
    VT* sU = sUG;
    VT* sG = sUG;
    // face fluxes in each face's tensor order, vertices, face slots, FP64
    // tables and the source load, all staged before the first barrier
    constexpr int LIFT = 2 * D * NFQ * M;
    constexpr int BLOCK = TL::BLOCK;
    constexpr int QIT = (EPB * LIFT + BLOCK - 1) / BLOCK;
    constexpr int NT2 = (N + 2) * N;
    __shared__ VT sQ[EPB * LIFT];
    __shared__ double sX[EPB * 4 * D];
    __shared__ int sEF[EPB * 2 * D];
    __shared__ double sLd[NT2];
    __shared__ signed char sFT[2 * D * 2 * NFQ];  // two orientation codes in 2D
    __shared__ double sSL[LOWLDS ? 1 : EPB * NB * M];
    __shared__ double sU0[LOWLDS ? 1 : EPB * M];  // FP32: each element's first coefficients
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / NB;
    const int k = threadIdx.x - le * NB;
    const int k0 = k / N, k1 = k - (k / N) * N;
    const int e = e0 + le;
    const bool active = (threadIdx.x < TL::NT) && (e < S.nelem);
    const typename PH::Params ph = local_params(ph0, emu, active ? e : 0);
    {
        const size_t base = (size_t)e0 * NB * M;
        const size_t total = (size_t)S.nelem * NB * M;
        for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x) {
            const int l = idx / (NB * M), r = idx - l * (NB * M), nd = r / M, v = r - nd * M;
            sU[l * ESU + nd * MS + v] = coef_stage<VT, NB, M, UT>(U, base, total, idx);
        }
        if (!LOWLDS)
            for (int idx = threadIdx.x; idx < EPB * M; idx += blockDim.x) {
                const size_t g = base + (size_t)(idx / M) * NB * M + idx % M;
                sU0[idx] = g < total ? (double)U[g] : 0.0;
            }
        if (!LOWLDS && srcload)
            for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x)
                sSL[idx] = (base + idx < total) ? srcload[base + idx] : 0.0;
        const size_t xb = (size_t)e0 * 4 * D, xt = (size_t)S.nelem * 4 * D;
        for (int idx = threadIdx.x; idx < EPB * 4 * D; idx += blockDim.x) sX[idx] = (xb + idx < xt) ? S.X[xb + idx] : 0.0;
        const size_t fb = (size_t)e0 * 2 * D, ft = (size_t)S.nelem * 2 * D;
        for (int idx = threadIdx.x; idx < EPB * 2 * D; idx += blockDim.x)
            sEF[idx] = (fb + idx < ft) ? S.efaces[fb + idx] : -1;
        for (int idx = threadIdx.x; idx < NT2; idx += blockDim.x) sLd[idx] = S.tab[(N + 2) + idx];
        for (int idx = threadIdx.x; idx < 2 * D * 2 * NFQ && idx < 2 * D * S.norient * NFQ; idx += blockDim.x)
            sFT[idx] = S.ftens[idx];
    }
    __syncthreads();
    // the element face fluxes, loaded now and staged after phase 1a
    VT qv[QIT];
    int qdst[QIT];
#pragma unroll
    for (int it = 0; it < QIT; ++it) {
        const int idx = threadIdx.x + it * BLOCK;
        const int l = idx / LIFT, rem = idx - l * LIFT;
        const int j = rem / (NFQ * M), w = rem - j * (NFQ * M);
        const int fq = w / M, v = w - fq * M;
        const int ef = (idx < EPB * LIFT) ? sEF[l * 2 * D + j] : -1;
        qv[it] = (ef >= 0) ? Q[(size_t)(ef >> 4) * NFQ * M + w] : (VT)0;
        qdst[it] = (ef >= 0) ? l * LIFT + (j * NFQ + sFT[(j * S.norient + (ef & 7)) * NFQ + fq]) * M + v : -1;
    }
    const double* xs = S.tab;
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    const VT* L0 = tab_of<VT>(S) + (N + 2);
    const VT* L1 = L0 + (N + 2) * N;
    VT* T0 = sT + (size_t)le * (LOWLDS ? 1 : 2) * ESU;  // [a][q1][v], node stride MS
    VT* T1 = LOWLDS ? sUG + (size_t)EPB * ESU + (size_t)le * ESU : T0 + ESU;
    // phase 1a: (a = k0, q1 = k1)
    if (active) {
        const VT* Ua = sU + (size_t)le * ESU + k0 * N * MS;
        VT l0[N], l1[N];
#pragma unroll
        for (int b = 0; b < N; ++b) {
            l0[b] = L0[k1 * N + b];
            l1[b] = L1[k1 * N + b];
        }
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT t0 = 0, t1 = 0;
#pragma unroll
            for (int b = 0; b < N; ++b) {
                const VT c = Ua[b * MS + v];
                t0 += c * l0[b];
                t1 += c * l1[b];
            }
            T0[k * MS + v] = t0;
            T1[k * MS + v] = t1;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < QIT; ++it)
        if (qdst[it] >= 0) sQ[qdst[it]] = qv[it];
    double X[1 << D][D];
    {
        const int lec = le < EPB ? le : 0;
#pragma unroll
        for (int vv = 0; vv < (1 << D); ++vv)
#pragma unroll
            for (int d = 0; d < D; ++d) X[vv][d] = sX[(lec * 4 + vv) * D + d];
    }
    double lam_w = 0.0, vis_w = 0.0;
    VT gq[M * D];  // this point's reference-frame fluxes (FP64: stored once T1 is dead)
    // phase 1b: (q0 = k0, q1 = k1)
    if (active) {
        double xi[D] = {xs[k0], xs[k1]};
        const double wq = wgl[k0] * wgl[k1];
        double x[D], J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, x, J, Hm);
        const double det = det_adj<D>(J, A);
        const double detc = fmax(0.0, det);
        const double detg = (det == 0.0) ? 1.0 : det;
        VT l0[N], l1[N];
#pragma unroll
        for (int a = 0; a < N; ++a) {
            l0[a] = L0[k0 * N + a];
            l1[a] = L1[k0 * N + a];
        }
        VT u[M], gx[M * D], F[M * D];
        const double idetg = 1.0 / detg;
        const VT a00 = (VT)(A[0][0] * idetg), a01 = (VT)(A[0][1] * idetg), a10 = (VT)(A[1][0] * idetg),
                 a11 = (VT)(A[1][1] * idetg);
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT s0 = 0, s1 = 0, sg0 = 0;
#pragma unroll
            for (int a = 0; a < N; ++a) {
                const VT t0 = T0[(a * N + k1) * MS + v];
                s0 += t0 * l0[a];
                sg0 += t0 * l1[a];
                s1 += T1[(a * N + k1) * MS + v] * l0[a];
            }
            if constexpr (LOWLDS) u[v] = s0;
            else u[v] = (VT)((double)s0 + sU0[le * M + v]);  // U_0 + sum dU_i phi_i
            // grad_x = grad_xi A / det (finite_element.hpp:220-258)
            gx[v * D + 0] = sg0 * a00 + s1 * a10;
            gx[v * D + 1] = sg0 * a01 + s1 * a11;
        }
        PH::flux(ph, u, gx, F);
        if (wave) {
            lam_w = PH::wave_speed(ph, u);
            vis_w = PH::visc_at(ph, u);
            if (e == S.nelem - 1 && k == NB - 1) wave[1] = (unsigned long long)__double_as_longlong(lam_w);
        }
        const double fac = detc * wq * idetg;
        const VT ga[D][D] = {{(VT)(A[0][0] * fac), (VT)(A[0][1] * fac)}, {(VT)(A[1][0] * fac), (VT)(A[1][1] * fac)}};
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int a = 0; a < D; ++a) {
                VT sacc = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) sacc += F[v * D + j] * ga[a][j];
                gq[v * D + a] = sacc;
            }
    }
    if (!LOWLDS && active) {
        VT* Gq = sG + (size_t)le * ESG + k * GS;
#pragma unroll
        for (int v = 0; v < M * D; ++v) Gq[v] = gq[v];
    }
    __shared__ double red[2 * 4];
    if (wave) {
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
    if constexpr (LOWLDS) {
        if (active) {
            VT* Gq = sG + (size_t)le * ESG + k * GS;
#pragma unroll
            for (int v = 0; v < M * D; ++v) Gq[v] = gq[v];
        }
        __syncthreads();
    }
    if (wave && threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { a = fmax(a, red[2 * w]); b = fmax(b, red[2 * w + 1]); }
        wave_part[2 * blockIdx.x] = a;
        wave_part[2 * blockIdx.x + 1] = b;
    }
    // the source load (FP64: read now, subtracted in phase 2b)
    double sl[M];
#pragma unroll
    for (int v = 0; v < M; ++v)
        sl[v] = !srcload       ? 0.0
                : !LOWLDS      ? sSL[(le * NB + k) * M + v]
                : active       ? srcload[((size_t)e * NB + k) * M + v]
                               : 0.0;
    // phase 2a: (q0 = k0, i1 = k1); T0 (and T1 in FP32) reused as S0 (S1) [q0][i1][v]
    if (active) {
        VT l0[N], l1[N];
#pragma unroll
        for (int q1 = 0; q1 < N; ++q1) {
            l0[q1] = L0[q1 * N + k1];
            l1[q1] = L1[q1 * N + k1];
        }
        const VT* Gr = sG + (size_t)le * ESG + k0 * N * GS;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT s0 = 0, s1 = 0;
#pragma unroll
            for (int q1 = 0; q1 < N; ++q1) {
                s0 += Gr[q1 * GS + v * D + 0] * l0[q1];
                if constexpr (!LOWLDS) s1 += Gr[q1 * GS + v * D + 1] * l1[q1];
            }
            T0[k * MS + v] = s0;
            if constexpr (!LOWLDS) T1[k * MS + v] = s1;
        }
    }
    __syncthreads();
    if (!active) return;
    // phase 2b: (i0 = k0, i1 = k1)
    double r[M];
    {
        VT l0[N], l1[N];
#pragma unroll
        for (int q0 = 0; q0 < N; ++q0) {
            l0[q0] = L0[q0 * N + k0];
            l1[q0] = L1[q0 * N + k0];
        }
        // FP64: S1[q0][i1] = sum_q1 G1[q0][q1] L1[q1][i1], summed here from G
        VT m1[N];
#pragma unroll
        for (int q1 = 0; q1 < N; ++q1) m1[q1] = L1[q1 * N + k1];
        const VT* Ge = sG + (size_t)le * ESG;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT sacc = 0;
#pragma unroll
            for (int q0 = 0; q0 < N; ++q0) {
                VT s1 = 0;
                if constexpr (LOWLDS) {
#pragma unroll
                    for (int q1 = 0; q1 < N; ++q1) s1 += Ge[(q0 * N + q1) * GS + v * D + 1] * m1[q1];
                } else {
                    s1 = T1[(q0 * N + k1) * MS + v];
                }
                sacc += T0[(q0 * N + k1) * MS + v] * l1[q0] + s1 * l0[q0];
            }
            r[v] = sacc;
        }
    }
    if (srcload) {
#pragma unroll
        for (int v = 0; v < M; ++v) r[v] -= sl[v];
    }
    const int kd[D] = {k0, k1};
    // lifting of the faces this thread's node lies on, from the staged
    // fluxes: the trace basis at the node factorises into the normal-axis
    // endpoint value and l_s(x_p) along the face
    unsigned mask = 0;
#pragma unroll
    for (int j = 0; j < 2 * D; ++j)
        if (kd[j % D] == ((j < D) ? 0 : N - 1)) mask |= 1u << j;
    while (mask) {
        const int j = __builtin_ctz(mask);
        mask &= mask - 1;
        const int ef = sEF[le * 2 * D + j];
        if (ef < 0) continue;
        const int kn = j % D;
        const int st = kd[1 - kn];
        const double en = sLd[((j < D) ? N : N + 1) * N + kd[kn]];
        const double sgn = ((ef >> 3) & 1) ? -en : en;
        const VT* Qf = sQ + le * LIFT + j * NFQ * M;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int p = 0; p < NFQ; ++p) acc += (double)Qf[p * M + v] * sLd[p * N + st];
            r[v] += sgn * acc;
        }
    }
#pragma unroll 1
    for (int j = 0; j < 2 * D; ++j) {
        const int ef = sEF[le * 2 * D + j];
        if (ef < 0) continue;
        const size_t slot = (size_t)(ef >> 4);
        const int orient = ef & 7;
        const signed char* fp = S.fpt + ((size_t)(j * S.norient + orient) * NFQ) * D;
        if constexpr (PH::HAS_HOMOGENEITY) {
            if (Pic) {
                const VT* Ps = Pic + slot * NFQ * M * D;
#pragma unroll 1
                for (int fq = 0; fq < NFQ; ++fq) {
                    double gphi[D];
                    ic_test_gradient<D, N>(S.tab, X, fp + fq * D, kd, ph.ic_mirror != 0, gphi);
#pragma unroll
                    for (int v = 0; v < M; ++v) {
                        double sacc = 0;
#pragma unroll
                        for (int d = 0; d < D; ++d) sacc += (double)Ps[(fq * M + v) * D + d] * gphi[d];
                        r[v] -= sacc;
                    }
                }
            }
        }
    }
    double* Re = R + ((size_t)e * NB + k) * M;
#pragma unroll
    for (int v = 0; v < M; ++v) Re[v] = r[v];
}

// 3D element kernel by sum factorisation.  Thread k of an element is the
// tensor triple (k0, k1, k2) (dimension 0 slowest).  Each element owns one
// LDS buffer of three [v][point] arrays that every stage reuses in place: a
// stage reads its inputs into registers, the workgroup synchronises, then
// the stage writes its outputs.
//  1a (a, b, q2):   A0 = sum_c U l_c(x_q2),  A1 = sum_c U l_c'(x_q2)
//  1b (a, q1, q2):  B00 = sum_b A0 l_b(x_q1), B01 = sum_b A0 l_b'(x_q1), B10 = sum_b A1 l_b(x_q1)
//  1c (q0, q1, q2): u = sum_a B00 l_a(x_q0), du/dxi = (sum_a B00 l_a', sum_a B01 l_a, sum_a B10 l_a);
//                   geometry, physical flux, reference-frame flux G_a = w det F A^T / det
//  2a (q0, q1, i2): C0 = sum_q2 G0 l_i2, C1 = sum_q2 G1 l_i2, C2 = sum_q2 G2 l_i2'
//  2b (q0, i1, i2): E0 = sum_q1 C0 l_i1, E1 = sum_q1 C1 l_i1' + C2 l_i1
//  2c (i0, i1, i2): sum_q0 E0 l_i0' + E1 l_i0 (the domain integral,
//                   conservation_law.hpp:244-314), minus the source load,
//                   plus the lifted face fluxes.
// Fluxes that use no gradient (Euler) skip A1, B01, B10 and the derivatives.
// Lifting: the element's face fluxes are staged in LDS in the element's own
// tensor order of each face's tangential axes (ftens), so the trace basis at
// a face node factorises: sum_pa l_sa(x_pa) sum_pb Q l_sb(x_pb).  The inner
// sums replace the face-flux rows in place during phase 1b (one thread per
// row), and a thread lifts only the faces its node lies on (the trace of
// every other basis function vanishes there).  All global loads that depend on nothing else
// (coefficients, vertices, face slots, tables) are issued before the first
// barrier; the face fluxes, whose addresses come from the face slots, are
// loaded into registers before phase 1a and stored to LDS after it.
template <int P, int TPB>
using Elem3Tile = ElemTile<3, P, TPB>;

template <int P, class PH, int TPB, class UT = double>
__global__ __launch_bounds__((Elem3Tile<P, TPB>::BLOCK), ICE_ELEM3_MINW) void element_kernel_3d(DevSpace S, typename PH::Params ph0,
                                                                         const double* __restrict__ emu,
                                                                         const double* __restrict__ srcload,
                                                                         const UT* __restrict__ U,
                                                                         const typename PH::real* __restrict__ Q,
                                                                         const typename PH::real* __restrict__ Pic,
                                                                         double* __restrict__ R,
                                                                         unsigned long long* __restrict__ wave,
                                                                         double* __restrict__ wave_part) {
    constexpr int D = 3;
    using TL = Elem3Tile<P, TPB>;
    constexpr int N = TL::N, NB = TL::NB, EPB = TL::EPB, BLOCK = TL::BLOCK;
    constexpr int NN = N * N;
    constexpr int NFQ = NN;
    constexpr int M = PH::M;
    constexpr bool GRAD = TraceRec<PH, D>::NEED_GRAD;
    constexpr int NBP = NB + 1;  // padded point stride of one [v][point] array
    constexpr int ARR = M * NBP;
    constexpr int EBUF = 3 * ARR;
    constexpr int LIFT = 2 * D * NFQ * M;
    constexpr int NV = 1 << D;
    constexpr int NT2 = (N + 2) * N;
    constexpr int QIT = (EPB * LIFT + BLOCK - 1) / BLOCK;
    using VT = typename PH::real;
    __shared__ VT sbuf[EPB * EBUF];
    // Fluxes without a gradient leave arrays 1 and 2 of the element buffer
    // unused until phase 1c: the face fluxes are staged there and lifted
    // before 1c, which keeps the workgroup's LDS to the element buffers
    // (3D P3 FP64: 4 waves per SIMD instead of 3)
    constexpr bool QALIAS = !GRAD && LIFT <= 2 * ARR && !PH::HAS_HOMOGENEITY;
    constexpr int QST = QALIAS ? EBUF : LIFT;  // element stride of the staged fluxes
    __shared__ VT sQarr[QALIAS ? 1 : EPB * LIFT];
    __shared__ VT sL[2 * NT2];
    __shared__ double sLd[NT2];
    __shared__ double sX[EPB * NV * D];
    __shared__ int sEF[EPB * 2 * D];
    __shared__ signed char sFT[2 * D * 8 * NFQ];
    constexpr bool DELTA = std::is_same<VT, float>::value;  // coefficients staged as U_i - U_0
    __shared__ double sU0[DELTA ? EPB * M : 1];
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / NB;
    const int k = threadIdx.x - le * NB;
    const int k0 = k / NN, k1 = (k / N) % N, k2 = k % N;
    const int e = e0 + le;
    const bool active = (threadIdx.x < TL::NT) && (e < S.nelem);
    const int lec = le < EPB ? le : 0;
    const typename PH::Params ph = local_params(ph0, emu, active ? e : 0);
    VT* B = sbuf + (size_t)lec * EBUF;
    VT* const sQ = QALIAS ? sbuf + ARR : sQarr;
    // prologue: every global load that depends on nothing else
    {
        const size_t base = (size_t)e0 * NB * M;
        const size_t total = (size_t)S.nelem * NB * M;
        for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x) {
            const int l = idx / (NB * M), rem = idx - l * (NB * M);
            const int node = rem / M, v = rem - node * M;
            sbuf[l * EBUF + v * NBP + node] = coef_stage<VT, NB, M, UT>(U, base, total, idx);
        }
        if (DELTA)
            for (int idx = threadIdx.x; idx < EPB * M; idx += blockDim.x) {
                const size_t g = base + (size_t)(idx / M) * NB * M + idx % M;
                sU0[idx] = g < total ? (double)U[g] : 0.0;
            }
        const size_t xb = (size_t)e0 * NV * D, xt = (size_t)S.nelem * NV * D;
        for (int idx = threadIdx.x; idx < EPB * NV * D; idx += blockDim.x)
            sX[idx] = (xb + idx < xt) ? S.X[xb + idx] : 0.0;
        const size_t fb = (size_t)e0 * 2 * D, ft = (size_t)S.nelem * 2 * D;
        for (int idx = threadIdx.x; idx < EPB * 2 * D; idx += blockDim.x)
            sEF[idx] = (fb + idx < ft) ? S.efaces[fb + idx] : -1;
        const VT* tv = tab_of<VT>(S) + (N + 2);
        for (int idx = threadIdx.x; idx < 2 * NT2; idx += blockDim.x) sL[idx] = tv[idx];
        for (int idx = threadIdx.x; idx < NT2; idx += blockDim.x) sLd[idx] = S.tab[(N + 2) + idx];
        for (int idx = threadIdx.x; idx < 2 * D * 8 * NFQ && idx < 2 * D * S.norient * NFQ; idx += blockDim.x)
            sFT[idx] = S.ftens[idx];
    }
    __syncthreads();
    // the element face fluxes, loaded now and staged after phase 1a:
    // sQ[l][j][t][v] with t the point's tensor index on the face
    VT qv[QIT];
    int qdst[QIT];
#pragma unroll
    for (int it = 0; it < QIT; ++it) {
        const int idx = threadIdx.x + it * BLOCK;
        const int l = idx / LIFT, rem = idx - l * LIFT;
        const int j = rem / (NFQ * M), w = rem - j * (NFQ * M);
        const int fq = w / M, v = w - fq * M;
        const int ef = (idx < EPB * LIFT) ? sEF[l * 2 * D + j] : -1;
        qv[it] = (ef >= 0) ? Q[(size_t)(ef >> 4) * NFQ * M + w] : (VT)0;
        qdst[it] = (ef >= 0) ? l * QST + (j * NFQ + sFT[(j * S.norient + (ef & 7)) * NFQ + fq]) * M + v : -1;
    }
    const double* xs = S.tab;
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    const VT* L0 = sL;
    const VT* L1 = sL + NT2;
    // lifting of the faces this thread's node lies on, added to r
    const int kd[D] = {k0, k1, k2};
    auto lift = [&](double (&r)[M]) {
        unsigned mask = 0;
#pragma unroll
        for (int j = 0; j < 2 * D; ++j)
            if (kd[j % D] == ((j < D) ? 0 : N - 1)) mask |= 1u << j;
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask &= mask - 1;
            const int ef = sEF[lec * 2 * D + j];
            if (ef < 0) continue;
            const int kn = j % D;
            const int sa = (kn == 0) ? k1 : k0, sb = (kn == 2) ? k1 : k2;
            // the node's normal-axis basis function at the face's endpoint (S index n or n + 1)
            const double en = sLd[((j < D) ? N : N + 1) * N + kd[kn]];
            const double sgn = ((ef >> 3) & 1) ? -en : en;
            double la[N];
#pragma unroll
            for (int p = 0; p < N; ++p) la[p] = sLd[p * N + sa];
            const VT* Zf = sQ + lec * QST + j * NFQ * M + sb * M;  // rows contracted along pb
#pragma unroll
            for (int v = 0; v < M; ++v) {
                double acc = 0.0;
#pragma unroll
                for (int pa = 0; pa < N; ++pa) acc += (double)Zf[pa * N * M + v] * la[pa];
                r[v] += sgn * acc;
            }
        }
    };
    double r[M];
#pragma unroll
    for (int v = 0; v < M; ++v) r[v] = 0.0;
    VT o0[M], o1[M], o2[M];  // stage outputs held across the barrier
    // 1a: (a = k0, b = k1, q2 = k2)
    if (active) {
        VT l0[N], l1[N];
#pragma unroll
        for (int c = 0; c < N; ++c) {
            l0[c] = L0[k2 * N + c];
            l1[c] = L1[k2 * N + c];
        }
        const VT* Ua = B + k0 * NN + k1 * N;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT t0 = 0, t1 = 0;
#pragma unroll
            for (int c = 0; c < N; ++c) {
                const VT x = Ua[v * NBP + c];
                t0 += x * l0[c];
                if constexpr (GRAD) t1 += x * l1[c];
            }
            o0[v] = t0;
            o1[v] = t1;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < QIT; ++it)
        if (qdst[it] >= 0) sQ[qdst[it]] = qv[it];
    if (active) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
            B[v * NBP + k] = o0[v];
            if constexpr (GRAD) B[ARR + v * NBP + k] = o1[v];
        }
    }
    __syncthreads();
    // face-flux rows contracted along the face's second tangential axis, in
    // place: sQ[l][j][pa][sb] = sum_pb Q[pa][pb] l_sb(x_pb); the lifting then
    // sums over pa only
    for (int tz = threadIdx.x; tz < EPB * 2 * D * N; tz += BLOCK) {
        const int l = tz / (2 * D * N), rem = tz - l * (2 * D * N);
        VT* Qr = sQ + l * QST + rem * N * M;  // rem = j * N + pa
        VT row[N][M];
#pragma unroll
        for (int pb = 0; pb < N; ++pb)
#pragma unroll
            for (int v = 0; v < M; ++v) row[pb][v] = Qr[pb * M + v];
#pragma unroll
        for (int sb = 0; sb < N; ++sb)
#pragma unroll
            for (int v = 0; v < M; ++v) {
                double z = 0.0;
#pragma unroll
                for (int pb = 0; pb < N; ++pb) z += (double)row[pb][v] * sLd[pb * N + sb];
                Qr[sb * M + v] = (VT)z;
            }
    }
    // 1b: (a = k0, q1 = k1, q2 = k2)
    if (active) {
        VT l0[N], l1[N];
#pragma unroll
        for (int b = 0; b < N; ++b) {
            l0[b] = L0[k1 * N + b];
            l1[b] = L1[k1 * N + b];
        }
        const VT* A0 = B + k0 * NN + k2;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT s00 = 0, s01 = 0, s10 = 0;
#pragma unroll
            for (int b = 0; b < N; ++b) {
                const VT x = A0[v * NBP + b * N];
                s00 += x * l0[b];
                if constexpr (GRAD) {
                    s01 += x * l1[b];
                    s10 += A0[ARR + v * NBP + b * N] * l0[b];
                }
            }
            o0[v] = s00;
            o1[v] = s01;
            o2[v] = s10;
        }
    }
    __syncthreads();
    if (active) {
        if constexpr (QALIAS) lift(r);  // the staged fluxes are overwritten in 1c
#pragma unroll
        for (int v = 0; v < M; ++v) {
            B[v * NBP + k] = o0[v];
            if constexpr (GRAD) {
                B[ARR + v * NBP + k] = o1[v];
                B[2 * ARR + v * NBP + k] = o2[v];
            }
        }
    }
    __syncthreads();
    // 1c: (q0 = k0, q1 = k1, q2 = k2)
    double lam_w = 0.0, vis_w = 0.0;
    if (active) {
        double X[NV][D];
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int d = 0; d < D; ++d) X[i][d] = sX[(lec * NV + i) * D + d];
        const double xi[D] = {xs[k0], xs[k1], xs[k2]};
        const double wq = wgl[k0] * wgl[k1] * wgl[k2];
        double J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, nullptr, J, Hm);
        const double det = det_adj<D>(J, A);
        const double detc = fmax(0.0, det);
        const double detg = (det == 0.0) ? 1.0 : det;
        const double idetg = 1.0 / detg;
        VT l0[N], l1[N];
#pragma unroll
        for (int a = 0; a < N; ++a) {
            l0[a] = L0[k0 * N + a];
            l1[a] = L1[k0 * N + a];
        }
        VT ai[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int j = 0; j < D; ++j) ai[a][j] = (VT)(A[a][j] * idetg);
        const VT* Bq = B + k1 * N + k2;
        VT u[M], gx[M * D], F[M * D];
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT s = 0, g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
            for (int a = 0; a < N; ++a) {
                const VT x = Bq[v * NBP + a * NN];
                s += x * l0[a];
                if constexpr (GRAD) {
                    g0 += x * l1[a];
                    g1 += Bq[ARR + v * NBP + a * NN] * l0[a];
                    g2 += Bq[2 * ARR + v * NBP + a * NN] * l0[a];
                }
            }
            if constexpr (DELTA) u[v] = (VT)((double)s + sU0[lec * M + v]);  // U_0 + sum dU_i phi_i
            else u[v] = s;
            // grad_x = grad_xi A / det (finite_element.hpp:220-258)
#pragma unroll
            for (int j = 0; j < D; ++j) gx[v * D + j] = g0 * ai[0][j] + g1 * ai[1][j] + g2 * ai[2][j];
        }
        PH::flux(ph, u, gx, F);
        if (wave) {
            lam_w = PH::wave_speed(ph, u);
            vis_w = PH::visc_at(ph, u);
            if (e == S.nelem - 1 && k == NB - 1) wave[1] = (unsigned long long)__double_as_longlong(lam_w);
        }
        const double fac = detc * wq * idetg;
        VT ga[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int j = 0; j < D; ++j) ga[a][j] = (VT)(A[a][j] * fac);
#pragma unroll
        for (int v = 0; v < M; ++v) {
            o0[v] = F[v * D + 0] * ga[0][0] + F[v * D + 1] * ga[0][1] + F[v * D + 2] * ga[0][2];
            o1[v] = F[v * D + 0] * ga[1][0] + F[v * D + 1] * ga[1][1] + F[v * D + 2] * ga[1][2];
            o2[v] = F[v * D + 0] * ga[2][0] + F[v * D + 1] * ga[2][1] + F[v * D + 2] * ga[2][2];
        }
    }
    __shared__ double red[2 * 4];
    if (wave) {
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
    if (active) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
            B[v * NBP + k] = o0[v];
            B[ARR + v * NBP + k] = o1[v];
            B[2 * ARR + v * NBP + k] = o2[v];
        }
    }
    __syncthreads();
    // 2a: (q0 = k0, q1 = k1, i2 = k2)
    if (active) {
        VT l0[N], l1[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            l0[q] = L0[q * N + k2];
            l1[q] = L1[q * N + k2];
        }
        const VT* G = B + k0 * NN + k1 * N;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
            for (int q = 0; q < N; ++q) {
                c0 += G[v * NBP + q] * l0[q];
                c1 += G[ARR + v * NBP + q] * l0[q];
                c2 += G[2 * ARR + v * NBP + q] * l1[q];
            }
            o0[v] = c0;
            o1[v] = c1;
            o2[v] = c2;
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
            B[v * NBP + k] = o0[v];
            B[ARR + v * NBP + k] = o1[v];
            B[2 * ARR + v * NBP + k] = o2[v];
        }
    }
    __syncthreads();
    // 2b: (q0 = k0, i1 = k1, i2 = k2)
    if (active) {
        VT l0[N], l1[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            l0[q] = L0[q * N + k1];
            l1[q] = L1[q * N + k1];
        }
        const VT* C = B + k0 * NN + k2;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT s0 = 0, s1 = 0;
#pragma unroll
            for (int q = 0; q < N; ++q) {
                s0 += C[v * NBP + q * N] * l0[q];
                s1 += C[ARR + v * NBP + q * N] * l1[q] + C[2 * ARR + v * NBP + q * N] * l0[q];
            }
            o0[v] = s0;
            o1[v] = s1;
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
            B[v * NBP + k] = o0[v];
            B[ARR + v * NBP + k] = o1[v];
        }
    }
    __syncthreads();
    if (!active) return;
    // 2c: (i0 = k0, i1 = k1, i2 = k2)
    {
        VT l0[N], l1[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            l0[q] = L0[q * N + k0];
            l1[q] = L1[q * N + k0];
        }
        const VT* E = B + k1 * N + k2;
#pragma unroll
        for (int v = 0; v < M; ++v) {
            VT s = 0;
#pragma unroll
            for (int q = 0; q < N; ++q) s += E[v * NBP + q * NN] * l1[q] + E[ARR + v * NBP + q * NN] * l0[q];
            r[v] += (double)s;
        }
        if (srcload) {
            const double* sl = srcload + ((size_t)e * NB + k) * M;
#pragma unroll
            for (int v = 0; v < M; ++v) r[v] -= sl[v];
        }
    }
    if constexpr (!QALIAS) lift(r);
    if constexpr (PH::HAS_HOMOGENEITY) {
        if (Pic) {
            double X[NV][D];
#pragma unroll
            for (int i = 0; i < NV; ++i)
#pragma unroll
                for (int d = 0; d < D; ++d) X[i][d] = sX[(lec * NV + i) * D + d];
#pragma unroll 1
            for (int j = 0; j < 2 * D; ++j) {
                const int ef = sEF[lec * 2 * D + j];
                if (ef < 0) continue;
                const size_t slot = (size_t)(ef >> 4);
                const signed char* fp = S.fpt + ((size_t)(j * S.norient + (ef & 7)) * NFQ) * D;
                const VT* Ps = Pic + slot * NFQ * M * D;
#pragma unroll 1
                for (int fq = 0; fq < NFQ; ++fq) {
                    double gphi[D];
                    ic_test_gradient<D, N>(S.tab, X, fp + fq * D, kd, ph.ic_mirror != 0, gphi);
#pragma unroll
                    for (int v = 0; v < M; ++v) {
                        double sacc = 0;
#pragma unroll
                        for (int d = 0; d < D; ++d) sacc += (double)Ps[(fq * M + v) * D + d] * gphi[d];
                        r[v] -= sacc;
                    }
                }
            }
        }
    }
    double* Re = R + ((size_t)e * NB + k) * M;
#pragma unroll
    for (int v = 0; v < M; ++v) Re[v] = r[v];
}

// ---- 3D P3 element kernel on the matrix cores (FP32, fluxes without a
// gradient) ------------------------------------------------------------------
// The sum-factorised contractions of element_kernel_3d are 4 x 4 operators
// applied along one axis of each element's [var][4][4][4] tensors.  Here
// one lane owns one tensor line (4 entries along the contracted axis) and
// the wave applies the operator to 64 lines at once with
// v_mfma_f32_4x4x1f32 (16 blocks of 4 x 4 outer products): lane 4b + r
// supplies operator row r, lane l its line's entry c, and after the four
// K = 1 steps lane l holds its line's 4 outputs (register r = output r).
// Each phase writes its outputs rotated, [ev][r][x][y] for the input line
// [ev][x][y][:], so the next phase's contraction axis is again contiguous
// (one 16-byte LDS read per line) and three rotations restore the natural
// point order.  Interpolation to the quadrature points (3 phases), the flux
// at each point (one thread per point, FP64 geometry), the test
// contractions (2a: three arrays, 2b: two with the sum of the second and
// third, 2c: one), then per node the source load, the face lifting and the
// FP64 store, as element_kernel_3d.
typedef float ice_floatx4 __attribute__((ext_vector_type(4)));

template <int M>
struct Elem3Mfma {
    static constexpr int EPB = 4;                          // elements per workgroup
    static constexpr int LINES = EPB * M * 16;             // tensor lines per array
    static constexpr int NWAVE = (LINES / 64 > EPB) ? LINES / 64 : EPB;  // line waves, or a wave per element
    static constexpr int BLOCK = 64 * NWAVE;
    // LDS stride of one [ev] tensor: a multiple of 16 floats that is 16 mod
    // 64, so the rotated stores of the 4 tensors a wave covers fall on
    // distinct banks
    static constexpr int EVS = 80;
    static constexpr int ARR = EPB * M * EVS;              // one array of the workgroup
};

// 4 x 4 contraction of the lane's line x: returns out[r] = sum_c A[r][c] x[c]
// when lane l passes a[c] = A[l % 4][c]
__device__ __forceinline__ ice_floatx4 mfma_line(const float (&a)[4], const float (&x)[4], ice_floatx4 acc) {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a[c], x[c], acc, 0, 0, 0);
    return acc;
}

// line L = (ev, x, y) of an array with [ev] stride EVS: its 4 entries start at
__device__ __forceinline__ int mfma_line_base(int L, int evs) { return (L >> 4) * evs + (L & 15) * 4; }
// rotated output slot r of line L: [ev][r][x][y]
__device__ __forceinline__ int mfma_rot_slot(int L, int r, int evs) { return (L >> 4) * evs + r * 16 + (L & 15); }

template <class PH, class UT = double>
__global__ __launch_bounds__(Elem3Mfma<PH::M>::BLOCK) void element_kernel_3d_mfma(
    DevSpace S, typename PH::Params ph0, const double* __restrict__ emu, const double* __restrict__ srcload,
    const UT* __restrict__ U, const float* __restrict__ Q, double* __restrict__ R,
    unsigned long long* __restrict__ wave, double* __restrict__ wave_part) {
    constexpr int D = 3, N = 4, NN = 16, NB = 64, NFQ = 16;
    constexpr int M = PH::M;
    using TL = Elem3Mfma<M>;
    constexpr int EPB = TL::EPB, LINES = TL::LINES, EVS = TL::EVS, ARR = TL::ARR, BLOCK = TL::BLOCK;
    constexpr int NV = 1 << D;
    constexpr int NT2 = (N + 2) * N;
    constexpr int LIFT = 2 * D * NFQ * M;
    static_assert(!TraceRec<PH, D>::NEED_GRAD, "fluxes without a gradient");
    static_assert((LINES % 64) == 0, "whole waves of tensor lines");
    __shared__ __align__(16) float s0[3 * ARR];
    __shared__ __align__(16) float s1[3 * ARR];
    __shared__ float sQ[EPB * LIFT];
    __shared__ float sL[2 * NT2];
    __shared__ double sLd[NT2];
    __shared__ double sX[EPB * NV * D];
    __shared__ int sEF[EPB * 2 * D];
    __shared__ signed char sFT[2 * D * 8 * NFQ];
    __shared__ double red[2 * (BLOCK / 64)];
    __shared__ double sU0[EPB * M];  // each element's first coefficients (coefficients staged as U_i - U_0)
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int e0 = blockIdx.x * EPB;
    // prologue: coefficients as [ev][node] (ev = element * M + var), tables,
    // vertices, face slots
    {
        const size_t base = (size_t)e0 * NB * M, total = (size_t)S.nelem * NB * M;
        for (int idx = tid; idx < EPB * NB * M; idx += BLOCK) {
            const int l = idx / (NB * M), rem = idx - l * (NB * M);
            const int node = rem / M, v = rem - node * M;
            s0[(l * M + v) * EVS + node] = coef_stage<float, NB, M, UT>(U, base, total, idx);
        }
        for (int idx = tid; idx < EPB * M; idx += BLOCK) {
            const size_t g = base + (size_t)(idx / M) * NB * M + idx % M;
            sU0[idx] = g < total ? (double)U[g] : 0.0;
        }
        const size_t xb = (size_t)e0 * NV * D, xt = (size_t)S.nelem * NV * D;
        for (int idx = tid; idx < EPB * NV * D; idx += BLOCK) sX[idx] = (xb + idx < xt) ? S.X[xb + idx] : 0.0;
        const size_t fb = (size_t)e0 * 2 * D, ft = (size_t)S.nelem * 2 * D;
        for (int idx = tid; idx < EPB * 2 * D; idx += BLOCK) sEF[idx] = (fb + idx < ft) ? S.efaces[fb + idx] : -1;
        for (int idx = tid; idx < 2 * NT2; idx += BLOCK) sL[idx] = S.tabf[(N + 2) + idx];
        for (int idx = tid; idx < NT2; idx += BLOCK) sLd[idx] = S.tab[(N + 2) + idx];
        for (int idx = tid; idx < 2 * D * 8 * NFQ && idx < 2 * D * S.norient * NFQ; idx += BLOCK)
            sFT[idx] = S.ftens[idx];
    }
    __syncthreads();
    // face fluxes in the face's tensor order (as element_kernel_3d): loaded
    // into registers now, stored to LDS after the first contraction phase,
    // so their latency overlaps it
    constexpr int QIT = (EPB * LIFT + BLOCK - 1) / BLOCK;
    float qv[QIT];
    int qdst[QIT];
#pragma unroll
    for (int it = 0; it < QIT; ++it) {
        const int idx = tid + it * BLOCK;
        const int l = idx / LIFT, rem = idx - l * LIFT;
        const int j = rem / (NFQ * M), w = rem - j * (NFQ * M);
        const int fq = w / M, v = w - fq * M;
        const int ef = (idx < EPB * LIFT) ? sEF[l * 2 * D + j] : -1;  // faces without a slot are not lifted
        qv[it] = (ef >= 0) ? Q[(size_t)(ef >> 4) * NFQ * M + w] : 0.f;
        qdst[it] = (ef >= 0) ? l * LIFT + (j * NFQ + sFT[(j * S.norient + (ef & 7)) * NFQ + fq]) * M + v : -1;
    }
    // per-lane operator rows: interpolation a[c] = L0[point r][basis c],
    // test contractions with L0^T and L1^T: a[c] = L[point c][basis r]
    const int r4 = lane & 3;
    float aI[4], aT0[4], aT1[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        aI[c] = sL[r4 * N + c];
        aT0[c] = sL[c * N + r4];
        aT1[c] = sL[NT2 + c * N + r4];
    }
    // one rotation pass: lines of `in` (any count, uniform per wave) -> `out`
    auto rotate = [&](const float* in, float* out, const float (&a)[4]) {
#pragma unroll 1
        for (int g = wv; g < LINES / 64; g += BLOCK / 64) {
            const int L = g * 64 + lane;
            const float4 xv = *reinterpret_cast<const float4*>(in + mfma_line_base(L, EVS));
            const float x[4] = {xv.x, xv.y, xv.z, xv.w};
            ice_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma_line(a, x, acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) out[mfma_rot_slot(L, r, EVS)] = acc[r];
        }
    };
    // 1a, 1b, 1c: u at the quadrature points, natural order, in s1[0]
    rotate(s0, s1, aI);
#pragma unroll
    for (int it = 0; it < QIT; ++it)
        if (qdst[it] >= 0) sQ[qdst[it]] = qv[it];
    __syncthreads();
    rotate(s1, s0, aI);
    __syncthreads();
    rotate(s0, s1, aI);
    __syncthreads();
    // flux at each point: contravariant fluxes G_a = F . adj(J)[a] detc w / det into s0[a]
    const double* xs = S.tab;
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    double lam_w = 0.0, vis_w = 0.0;
    for (int t = tid; t < EPB * NB; t += BLOCK) {
        const int l = t / NB, k = t - l * NB;
        const int e = e0 + l;
        if (e >= S.nelem) continue;
        const int k0 = k / NN, k1 = (k / N) % N, k2 = k % N;
        const typename PH::Params ph = local_params(ph0, emu, e);
        double X[NV][D];
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int d = 0; d < D; ++d) X[i][d] = sX[(l * NV + i) * D + d];
        const double xi[D] = {xs[k0], xs[k1], xs[k2]};
        const double wq = wgl[k0] * wgl[k1] * wgl[k2];
        double J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, nullptr, J, Hm);
        const double det = det_adj<D>(J, A);
        const double detc = fmax(0.0, det);
        const double detg = (det == 0.0) ? 1.0 : det;
        const double fac = detc * wq / detg;
        float u[M], gx[M * D] = {}, F[M * D];
#pragma unroll
        for (int v = 0; v < M; ++v) u[v] = (float)((double)s1[(l * M + v) * EVS + k] + sU0[l * M + v]);
        PH::flux(ph, u, gx, F);
        if (wave) {
            lam_w = fmax(lam_w, (double)PH::wave_speed(ph, u));
            vis_w = fmax(vis_w, (double)PH::visc_at(ph, u));
            if (e == S.nelem - 1 && k == NB - 1) wave[1] = (unsigned long long)__double_as_longlong((double)PH::wave_speed(ph, u));
        }
        float ga[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int j = 0; j < D; ++j) ga[a][j] = (float)(A[a][j] * fac);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int v = 0; v < M; ++v)
                s0[a * ARR + (l * M + v) * EVS + k] = F[v * D + 0] * ga[a][0] + F[v * D + 1] * ga[a][1] +
                                                     F[v * D + 2] * ga[a][2];
    }
    if (wave) {
        for (int off = 32; off > 0; off >>= 1) {
            lam_w = fmax(lam_w, __shfl_xor(lam_w, off));
            vis_w = fmax(vis_w, __shfl_xor(vis_w, off));
        }
        if (lane == 0) {
            red[2 * wv] = lam_w;
            red[2 * wv + 1] = vis_w;
        }
    }
    __syncthreads();
    if (wave && tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) { a = fmax(a, red[2 * w]); b = fmax(b, red[2 * w + 1]); }
        wave_part[2 * blockIdx.x] = a;
        wave_part[2 * blockIdx.x + 1] = b;
    }
    // 2a: contract q2 of G0, G1 (values) and G2 (derivatives) -> s1[a]
#pragma unroll 1
    for (int g = wv; g < 3 * LINES / 64; g += BLOCK / 64) {
        const int a = g / (LINES / 64), L = (g - a * (LINES / 64)) * 64 + lane;
        const float4 xv = *reinterpret_cast<const float4*>(s0 + a * ARR + mfma_line_base(L, EVS));
        const float x[4] = {xv.x, xv.y, xv.z, xv.w};
        ice_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = (a == 2) ? mfma_line(aT1, x, acc) : mfma_line(aT0, x, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) s1[a * ARR + mfma_rot_slot(L, r, EVS)] = acc[r];
    }
    __syncthreads();
    // 2b: contract q1: T0 = L0^T G0' -> s0[0]; T12 = L1^T G1' + L0^T G2' -> s0[1]
#pragma unroll 1
    for (int g = wv; g < 2 * LINES / 64; g += BLOCK / 64) {
        const int o = g / (LINES / 64), L = (g - o * (LINES / 64)) * 64 + lane;
        ice_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        const float4 xv = *reinterpret_cast<const float4*>(s1 + o * ARR + mfma_line_base(L, EVS));
        const float x[4] = {xv.x, xv.y, xv.z, xv.w};
        if (o == 0) {
            acc = mfma_line(aT0, x, acc);
        } else {
            acc = mfma_line(aT1, x, acc);
            const float4 yv = *reinterpret_cast<const float4*>(s1 + 2 * ARR + mfma_line_base(L, EVS));
            const float y[4] = {yv.x, yv.y, yv.z, yv.w};
            acc = mfma_line(aT0, y, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) s0[o * ARR + mfma_rot_slot(L, r, EVS)] = acc[r];
    }
    __syncthreads();
    // 2c: contract q0: R = L1^T T0' + L0^T T12' -> s1[0], natural node order
#pragma unroll 1
    for (int g = wv; g < LINES / 64; g += BLOCK / 64) {
        const int L = g * 64 + lane;
        const float4 xv = *reinterpret_cast<const float4*>(s0 + mfma_line_base(L, EVS));
        const float4 yv = *reinterpret_cast<const float4*>(s0 + ARR + mfma_line_base(L, EVS));
        const float x[4] = {xv.x, xv.y, xv.z, xv.w}, y[4] = {yv.x, yv.y, yv.z, yv.w};
        ice_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_line(aT1, x, acc);
        acc = mfma_line(aT0, y, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) s1[mfma_rot_slot(L, r, EVS)] = acc[r];
    }
    __syncthreads();
    // per node: volume term, source load, lifting, FP64 store (element_kernel_3d)
    for (int t = tid; t < EPB * NB; t += BLOCK) {
        const int l = t / NB, k = t - l * NB;
        const int e = e0 + l;
        if (e >= S.nelem) continue;
        const int k0 = k / NN, k1 = (k / N) % N, k2 = k % N;
        double r[M];
#pragma unroll
        for (int v = 0; v < M; ++v) r[v] = (double)s1[(l * M + v) * EVS + k];
        if (srcload) {
            const double* sl = srcload + ((size_t)e * NB + k) * M;
#pragma unroll
            for (int v = 0; v < M; ++v) r[v] -= sl[v];
        }
        const int kd[D] = {k0, k1, k2};
        unsigned mask = 0;
#pragma unroll
        for (int j = 0; j < 2 * D; ++j)
            if (kd[j % D] == ((j < D) ? 0 : N - 1)) mask |= 1u << j;
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask &= mask - 1;
            const int ef = sEF[l * 2 * D + j];
            if (ef < 0) continue;
            const int kn = j % D;
            const int sa = (kn == 0) ? k1 : k0, sb = (kn == 2) ? k1 : k2;
            const double en = sLd[((j < D) ? N : N + 1) * N + kd[kn]];
            const double sgn = ((ef >> 3) & 1) ? -en : en;
            double la[N], lb[N];
#pragma unroll
            for (int p = 0; p < N; ++p) {
                la[p] = sLd[p * N + sa];
                lb[p] = sLd[p * N + sb];
            }
            const float* Qf = sQ + l * LIFT + j * NFQ * M;
#pragma unroll
            for (int v = 0; v < M; ++v) {
                double acc = 0.0;
#pragma unroll
                for (int pa = 0; pa < N; ++pa) {
                    double tt = 0.0;
#pragma unroll
                    for (int pb = 0; pb < N; ++pb) tt += (double)Qf[(pa * N + pb) * M + v] * lb[pb];
                    acc += tt * la[pa];
                }
                r[v] += sgn * acc;
            }
        }
        double* Re = R + ((size_t)e * NB + k) * M;
#pragma unroll
        for (int v = 0; v < M; ++v) Re[v] = r[v];
    }
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

// The same load vector from values at the volume quadrature points,
// vals[e][q][v] in the order of ice_space_qp_coords (clamp: max(0, detJ) as
// the source load; 0: the raw detJ of the projection).
template <int D, int P>
__global__ __launch_bounds__(256) void load_values_kernel(DevSpace S, const double* __restrict__ vals, int M,
                                                          int clamp, double* __restrict__ out) {
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
        double J[D][D], Hm[D][D][D], A[D][D];
        q1_map<D>(X, xi, nullptr, J, Hm);
        const double det = det_adj<D>(J, A);
        const double detc = clamp ? fmax(0.0, det) : det;
        const double* src = vals + ((size_t)e * NB + q) * M;
        for (int v = 0; v < M; ++v) acc[v] += src[v] * phi * detc * wq;
    }
    for (int v = 0; v < M; ++v) out[((size_t)e * NB + i) * M + v] = acc[v];
}

// Physical coordinates of the boundary face quadrature points, out[b][q][d]
// in the device boundary face order (the points bval_kernel evaluates at).
template <int D, int P>
__global__ __launch_bounds__(256) void bface_coords_kernel(DevSpace S, double* __restrict__ out) {
    constexpr int N = P + 1;
    constexpr int NFQ = Pow<N, D - 1>::value;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = tid / NFQ, q = tid - b * NFQ;
    if (b >= S.nfb) return;
    const int4 rec = S.fbdy[b];
    double X[1 << D][D], x[D], nrm[D];
    load_vertices<D>(S.X, rec.x, X);
    (void)face_geometry<D, N>(S, X, info_fnl(rec.y), q, x, nrm);
    for (int d = 0; d < D; ++d) out[((size_t)b * NFQ + q) * D + d] = x[d];
}

template <int D, int P>
int launch_load_values(const DevSpace& S, const double* vals, int M, int clamp, double* out, hipStream_t st) {
    constexpr int NB = Pow<P + 1, D>::value;
    load_values_kernel<D, P><<<S.nelem, ((NB + 63) / 64) * 64, 0, st>>>(S, vals, M, clamp, out);
    return (int)hipGetLastError();
}

template <int D, int P>
int launch_bface_coords(const DevSpace& S, double* out, hipStream_t st) {
    constexpr int NFQ = Pow<P + 1, D - 1>::value;
    const long nt = (long)S.nfb * NFQ;
    if (nt > 0) bface_coords_kernel<D, P><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(S, out);
    return (int)hipGetLastError();
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

int load_from_values(int ndim, int p, int neq, const DevSpace& S, const double* vals, int clamp, double* out,
                     hipStream_t st) {
    if (neq > 8) return -1;
#define ICE_VCASE(DD, PP) if (ndim == DD && p == PP) return launch_load_values<DD, PP>(S, vals, neq, clamp, out, st);
    ICE_VCASE(1, 1) ICE_VCASE(1, 2) ICE_VCASE(1, 3) ICE_VCASE(1, 4)
    ICE_VCASE(2, 1) ICE_VCASE(2, 2) ICE_VCASE(2, 3) ICE_VCASE(2, 4)
    ICE_VCASE(3, 1) ICE_VCASE(3, 2) ICE_VCASE(3, 3) ICE_VCASE(3, 4)
#undef ICE_VCASE
    return -1;
}

int bface_coords(int ndim, int p, const DevSpace& S, double* out, hipStream_t st) {
#define ICE_BCASE(DD, PP) if (ndim == DD && p == PP) return launch_bface_coords<DD, PP>(S, out, st);
    ICE_BCASE(1, 1) ICE_BCASE(1, 2) ICE_BCASE(1, 3) ICE_BCASE(1, 4)
    ICE_BCASE(2, 1) ICE_BCASE(2, 2) ICE_BCASE(2, 3) ICE_BCASE(2, 4)
    ICE_BCASE(3, 1) ICE_BCASE(3, 2) ICE_BCASE(3, 3) ICE_BCASE(3, 4)
#undef ICE_BCASE
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
// PH fixes the arithmetic of the trace, flux and 2D element kernels (double,
// or float for the mixed-precision residual); geometry, the source load and
// the residual itself stay FP64.
template <int D, int P, class PH, class UT>
int launch_residual(const DevSpace& S, const typename PH::Params& ph, const DiscDev& dd, const UT* U, double* R,
                    hipStream_t st, hipEvent_t* ev) {
    using VT = typename PH::real;
    constexpr int N = P + 1;
    constexpr int NFQ = Pow<N, D - 1>::value;
    DiscDev d = dd;
    if (!PH::HAS_HOMOGENEITY) d.Pic = nullptr;
    VT* T = reinterpret_cast<VT*>(d.T);
    if (ev) (void)hipEventRecord(ev[0], st);
    const int nbi = (int)(((long)S.nfi * NFQ + 255) / 256);
    const int nbb = (int)(((long)S.nfb * NFQ + 255) / 256);
    if constexpr (D == 3 && P == 3 && !TraceRec<PH, D>::NEED_GRAD) {
        constexpr int EPB = 4;  // 192 threads
        trace_kernel_rows<PH, EPB, UT><<<(S.nelem + EPB - 1) / EPB, ((EPB * 48 + 63) / 64) * 64, 0, st>>>(S, ph, U, T);
    } else {
        constexpr int TPE = 2 * D * NFQ;
        // whole elements per workgroup, rounded up to whole waves only once:
        // a wave that would hold no trace point is not launched (3D P3:
        // 2 x 96 points in 192 threads rather than 256)
        constexpr int EPW = (TPE <= ICE_TRACE_TPB) ? ICE_TRACE_TPB / TPE : 1;
        constexpr int TTPB = ((EPW * TPE + 63) / 64) * 64;
        constexpr int EPB = TTPB / TPE;
        trace_kernel<D, P, PH, TTPB, UT><<<(S.nelem + EPB - 1) / EPB, TTPB, 0, st>>>(S, ph, U, T);
    }
    if (ev) (void)hipEventRecord(ev[1], st);
    if (nbb > 0)
        flux_kernel<D, P, PH, true, UT><<<nbi + nbb, 256, 0, st>>>(S, ph, d, T, nbi, U);
    else if (nbi > 0)
        flux_kernel<D, P, PH, false, UT><<<nbi, 256, 0, st>>>(S, ph, d, T, nbi, U);
    if (ev) (void)hipEventRecord(ev[2], st);
    using TL = ElemTile<D, P>;
    int nblk = (S.nelem + TL::EPB - 1) / TL::EPB;
    if constexpr (D == 2 || D == 3) {
        auto go = [&](auto tpb) {
            constexpr int T = decltype(tpb)::value;
            using TT = ElemTile<D, P, T>;
            nblk = (S.nelem + TT::EPB - 1) / TT::EPB;
            if constexpr (D == 2)
                element_kernel_2d<P, PH, T, UT><<<nblk, TT::BLOCK, 0, st>>>(
                    S, ph, d.emu, d.srcload, U, reinterpret_cast<const VT*>(d.Q), reinterpret_cast<const VT*>(d.Pic), R,
                    d.wave, d.wave_part);
            else
                element_kernel_3d<P, PH, T, UT><<<nblk, TT::BLOCK, 0, st>>>(
                    S, ph, d.emu, d.srcload, U, reinterpret_cast<const VT*>(d.Q), reinterpret_cast<const VT*>(d.Pic), R,
                    d.wave, d.wave_part);
        };
        const int tile = std::is_same<VT, double>::value ? d.tile : d.tile_mixed;
        if constexpr (D == 3 && P == 3 && std::is_same<VT, float>::value && !TraceRec<PH, D>::NEED_GRAD &&
                      !PH::HAS_HOMOGENEITY) {
            if (tile == TILE_MFMA) {
                using TM = Elem3Mfma<PH::M>;
                nblk = (S.nelem + TM::EPB - 1) / TM::EPB;
                element_kernel_3d_mfma<PH, UT><<<nblk, TM::BLOCK, 0, st>>>(
                    S, ph, d.emu, d.srcload, U, reinterpret_cast<const float*>(d.Q), R, d.wave, d.wave_part);
                if (d.wave) wave_reduce_kernel<<<1, 256, 0, st>>>(d.wave_part, nblk, d.wave);
                if (ev) (void)hipEventRecord(ev[3], st);
                return (int)hipGetLastError();
            }
        }
        switch (tile) {
        case 1: go(std::integral_constant<int, tile_tpb(1)>{}); break;
        case 2: go(std::integral_constant<int, tile_tpb(2)>{}); break;
        default: go(std::integral_constant<int, (D == 3 ? ICE_ELEM3_TPB : 256)>{}); break;
        }
    } else {
        static_assert(std::is_same<VT, double>::value || D > 1, "the mixed-precision residual is 2D and 3D");
        static_assert(std::is_same<UT, double>::value, "FP32 states are 2D and 3D");
        if constexpr (std::is_same<VT, double>::value)
            element_kernel<D, P, PH><<<nblk, 256, 0, st>>>(S, ph, d.emu, d.srcload, U, d.Q, d.Pic, R, d.wave,
                                                            d.wave_part);
    }
    if (d.wave) wave_reduce_kernel<<<1, 256, 0, st>>>(d.wave_part, nblk, d.wave);
    if (ev) (void)hipEventRecord(ev[3], st);
    return (int)hipGetLastError();
}

template <class PH, int D, class UT = double>
int dispatch_order(int p, const DevSpace& S, const typename PH::Params& ph, const DiscDev& dd, const UT* U,
                   double* R, hipStream_t st, hipEvent_t* ev) {
    switch (p) {
    case 1: return launch_residual<D, 1, PH, UT>(S, ph, dd, U, R, st, ev);
    case 2: return launch_residual<D, 2, PH, UT>(S, ph, dd, U, R, st, ev);
    case 3: return launch_residual<D, 3, PH, UT>(S, ph, dd, U, R, st, ev);
    case 4: return launch_residual<D, 4, PH, UT>(S, ph, dd, U, R, st, ev);
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

// Mixed precision: FP32 interpolation, fluxes and face data, FP64 geometry,
// source load and residual (2D and 3D).
template <int D, class UT>
int residual_mixed_dim(int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd, const UT* U,
                       double* R, hipStream_t st, hipEvent_t* ev) {
    const PhysT<float> pf = phys_cast<float>(ph);
    if (flux == ICE_FLUX_BURGERS) return dispatch_order<Burgers<D, float>, D, UT>(p, S, pf, dd, U, R, st, ev);
    if (flux == ICE_FLUX_NAVIER_STOKES)
        return dispatch_order<NavierStokes<D, true, float>, D, UT>(p, S, pf, dd, U, R, st, ev);
    if (flux == ICE_FLUX_EULER) return dispatch_order<NavierStokes<D, false, float>, D, UT>(p, S, pf, dd, U, R, st, ev);
    return -1;
}

int residual_mixed(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                   const double* U, double* R, hipStream_t st, hipEvent_t* ev) {
    if (ndim == 2) return residual_mixed_dim<2, double>(p, flux, S, ph, dd, U, R, st, ev);
    if (ndim == 3) return residual_mixed_dim<3, double>(p, flux, S, ph, dd, U, R, st, ev);
    return -2;
}

// The mixed residual of an FP32 state (the same kernels reading FP32
// coefficients; differences and element values still formed in FP64)
int residual_mixed_f32state(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                            const float* U, double* R, hipStream_t st) {
    if (ndim == 2) return residual_mixed_dim<2, float>(p, flux, S, ph, dd, U, R, st, nullptr);
    if (ndim == 3) return residual_mixed_dim<3, float>(p, flux, S, ph, dd, U, R, st, nullptr);
    return -2;
}

}  // namespace ice
