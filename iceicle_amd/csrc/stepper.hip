// Explicit time stepping on the device: inverse mass matrix, CFL time step
// and the vector updates of RK3SSP / RK3TVD / ExplicitEuler
// (solvers/ssp_rk3.hpp:106-195, tvd_rk3.hpp:107-231, explicit_euler.hpp:87-138,
// explicit_utils.hpp:61-160).
//
// Inverse mass matrix: the reference builds M_e = sum_q phi_i phi_j w_q detJ_q
// and LU-solves it at every stage (ssp_rk3.hpp:142-163).  With as many
// quadrature points as basis functions, M_e = B^T W_e B where B = B1 x .. x B1
// is the square tensor-product Vandermonde of the Lagrange basis at the
// Gauss points and W_e = diag(w_q detJ_q), so M_e^-1 = B^-1 W_e^-1 B^-T: two
// sum-factorised passes with the N x N table B1^-1 and one scaling per
// quadrature point, with no per-element storage.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "physics_dev.h"

namespace ice {

template <int B, int E>
struct PowS { static constexpr int value = B * PowS<B, E - 1>::value; };
template <int B>
struct PowS<B, 0> { static constexpr int value = 1; };

template <int D>
__device__ __forceinline__ double q1_det(const double* __restrict__ Xe, const double* xi) {
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int a = 0; a < D; ++a) J[i][a] = 0;
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double t = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int bit = (k >> (D - 1 - d)) & 1;
                const double n0 = 0.5 * (1 - xi[d]);
                t *= (d == a) ? (bit ? 0.5 : -0.5) : (bit ? 1.0 - n0 : n0);
            }
#pragma unroll
            for (int i = 0; i < D; ++i) J[i][a] += t * Xe[k * D + i];
        }
    }
    if constexpr (D == 1) return J[0][0];
    else if constexpr (D == 2) return J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else
        return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// Optional stage combination formed by the kernel from the increment k_new
// it computes: out = base + dt (c0 k0 + c1 k1 + cn k_new) over the nprev
// earlier increments k0, k1, with the step of rk_combine_kernel (the
// stepper's dt, or the element's, none once the run is done); k_new is
// rounded to KT first, as when it is stored and read back.
template <class KT>
struct StageComb {
    KT* out = nullptr;  // null: k = M^-1 r only
    const KT* base = nullptr;
    const KT* k0 = nullptr;
    const KT* k1 = nullptr;
    int nprev = 0;
    double c0 = 0, c1 = 0, cn = 0;
    const double* dtp = nullptr;
    const double* dte = nullptr;
    const StepState* stp = nullptr;
    bool store_k = true;  // also store k_new (a later stage reads it)
};

// out = base + dt (c0 k0 + c1 k1 + cn kn) over the NL contiguous values of a
// line at o, with NPV earlier increments: straight-line loads of each input
// (so they merge into wide loads), all read before out, which may be base,
// is stored; then kn itself if the stage keeps it
template <int NPV, int NL, class KT>
__device__ __forceinline__ void stage_combine_line(const StageComb<KT>& cb, KT* __restrict__ k, size_t o, double dt,
                                                   const KT* kn) {
    double sum[NL], bv[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        double acc = 0.0;
        if constexpr (NPV > 0) acc = cb.c0 * (double)cb.k0[o + i];
        if constexpr (NPV > 1) acc += cb.c1 * (double)cb.k1[o + i];
        sum[i] = acc;
        bv[i] = (double)cb.base[o + i];
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        sum[i] += cb.cn * (double)kn[i];
        cb.out[o + i] = (KT)(bv[i] + dt * sum[i]);
    }
    if (cb.store_k) {
#pragma unroll
        for (int i = 0; i < NL; ++i) k[o + i] = kn[i];
    }
}

template <int NL, class KT>
__device__ __forceinline__ void stage_combine_line(const StageComb<KT>& cb, KT* __restrict__ k, size_t o, double dt,
                                                   const KT* kn) {
    if (cb.nprev == 0) stage_combine_line<0, NL, KT>(cb, k, o, dt, kn);
    else if (cb.nprev == 1) stage_combine_line<1, NL, KT>(cb, k, o, dt, kn);
    else stage_combine_line<2, NL, KT>(cb, k, o, dt, kn);
}

// k = M^-1 r for every element and component.  Elements whose mass matrix is
// singular (a zero detJ at a quadrature point) get 1e-8 r, the fallback of
// tvd_rk3.hpp:190-204.
// inverse-mass workgroup target (3D P3 Euler: 128 threads 633 us, 256
// threads 728 us per application on 64^3 hexes)
#ifndef ICE_MINV_TPB
#define ICE_MINV_TPB 128
#endif
template <int D, int P, int M, int TPB, class KT = double>
__global__ __launch_bounds__(TPB) void minv_kernel(DevSpace S, const double* __restrict__ r, KT* __restrict__ k,
                                                   StageComb<KT> cb) {
    constexpr int N = P + 1;
    constexpr int NB = PowS<N, D>::value;
    static_assert(TPB >= NB, "one thread per node of an element");
    constexpr int EPB = TPB / NB;
    __shared__ double bufA[EPB * NB * M];
    __shared__ double bufB[EPB * NB * M];
    __shared__ int sing[EPB];
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / NB;
    const int kk = threadIdx.x - le * NB;
    const int e = e0 + le;
    const bool active = threadIdx.x < EPB * NB && e < S.nelem;
    const double* xs = S.tab;
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    const double* Binv = wgl + N;  // Binv[i][q]
    if (threadIdx.x < EPB) sing[threadIdx.x] = 0;
    {
        const size_t base = (size_t)e0 * NB * M, total = (size_t)S.nelem * NB * M;
        for (int idx = threadIdx.x; idx < EPB * NB * M; idx += blockDim.x)
            bufA[idx] = (base + idx < total) ? r[base + idx] : 0.0;
    }
    const size_t o = ((size_t)(active ? e : 0) * NB + kk) * M;
    // the stage combination's base and earlier increments, loaded here so
    // that their latency hides behind the contractions (every value is read
    // before out, which may be base, is stored)
    double sum[M], bv[M];
    if (cb.out && active) {
#pragma unroll
        for (int v = 0; v < M; ++v) {
            sum[v] = cb.nprev > 0 ? cb.c0 * (double)cb.k0[o + v] : 0.0;
            if (cb.nprev > 1) sum[v] += cb.c1 * (double)cb.k1[o + v];
            bv[v] = (double)cb.base[o + v];
        }
    }
    int kd[D];
    {
        int t = kk;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) { kd[d] = t % N; t /= N; }
    }
    __syncthreads();
    double* cur = bufA;
    double* nxt = bufB;
    // y = B^-T r: contract axis d with Binv[i_d][q_d]
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (active) {
            int stride = 1;
#pragma unroll
            for (int dd = D - 1; dd > d; --dd) stride *= N;
            const int base = kk - kd[d] * stride;
            double acc[M];
#pragma unroll
            for (int v = 0; v < M; ++v) acc[v] = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double c = Binv[j * N + kd[d]];
                const double* src = cur + (le * NB + base + j * stride) * M;
#pragma unroll
                for (int v = 0; v < M; ++v) acc[v] += c * src[v];
            }
#pragma unroll
            for (int v = 0; v < M; ++v) nxt[(le * NB + kk) * M + v] = acc[v];
        }
        __syncthreads();
        double* t = cur; cur = nxt; nxt = t;
    }
    // z = y / (w_q detJ_q)
    if (active) {
        double xi[D], w = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xi[d] = xs[kd[d]];
            w = (d == 0) ? wgl[kd[d]] : w * wgl[kd[d]];
        }
        const double det = q1_det<D>(S.X + (size_t)e * (1 << D) * D, xi);
        if (det == 0.0) sing[le] = 1;
        const double s = 1.0 / (w * det);
#pragma unroll
        for (int v = 0; v < M; ++v) cur[(le * NB + kk) * M + v] *= s;
    }
    __syncthreads();
    // out = B^-1 z: contract axis d with Binv[i_d][q_d]
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (active) {
            int stride = 1;
#pragma unroll
            for (int dd = D - 1; dd > d; --dd) stride *= N;
            const int base = kk - kd[d] * stride;
            double acc[M];
#pragma unroll
            for (int v = 0; v < M; ++v) acc[v] = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double c = Binv[kd[d] * N + j];
                const double* src = cur + (le * NB + base + j * stride) * M;
#pragma unroll
                for (int v = 0; v < M; ++v) acc[v] += c * src[v];
            }
#pragma unroll
            for (int v = 0; v < M; ++v) nxt[(le * NB + kk) * M + v] = acc[v];
        }
        __syncthreads();
        double* t = cur; cur = nxt; nxt = t;
    }
    if (!active) return;
    const bool sg = sing[le] != 0;
    KT kn[M];
#pragma unroll
    for (int v = 0; v < M; ++v) kn[v] = sg ? (KT)(1e-8 * r[o + v]) : (KT)cur[(le * NB + kk) * M + v];
    if (cb.out) {
        const double dt = (cb.stp && cb.stp->done) ? 0.0 : (cb.dte ? cb.dte[e] : *cb.dtp);
#pragma unroll
        for (int v = 0; v < M; ++v) {
            sum[v] += cb.cn * (double)kn[v];
            cb.out[o + v] = (KT)(bv[v] + dt * sum[v]);
        }
        if (!cb.store_k) return;
    }
#pragma unroll
    for (int v = 0; v < M; ++v) k[o + v] = kn[v];
}

// The same in 3D with one tensor line of N nodes per thread: the passes run
// along axes 2, 1, 0 (B^-T), scale, then 0, 1, 2 (B^-1), so the two passes
// meet on axis 0 in one thread's registers and four LDS transposes (in
// place: a thread rewrites the entries of its own line) replace the six
// LDS round trips of the node-per-thread kernel; the residual is read and
// the result written as contiguous lines of N nodes.
// target threads per workgroup of the line kernels (whole elements)
#ifndef ICE_MINV_LINES_THREADS
#define ICE_MINV_LINES_THREADS 128
#endif
#ifndef ICE_MINV_LINES2_THREADS
#define ICE_MINV_LINES2_THREADS 128
#endif
// minimum waves per SIMD the 3D line kernel's registers must allow where a
// line holds at most this many values (its LDS allows three; longer lines
// spill under the cap)
#ifndef ICE_MINV_LINES_MINW
#define ICE_MINV_LINES_MINW 3
#endif
// plane padding of the N = 4 line kernel's LDS buffer (4: conflict-free
// transposes)
#ifndef ICE_MINV_LINES_PAD4
#define ICE_MINV_LINES_PAD4 4
#endif
#ifndef ICE_MINV_LINES_CAPVALS
#define ICE_MINV_LINES_CAPVALS 20
#endif
template <int P, int M, class KT = double>
struct MinvLines {
    static constexpr int N = P + 1, NN = N * N, NB = NN * N;
    static constexpr int PL = NN + (N == 4 ? ICE_MINV_LINES_PAD4 : 1);  // plane stride (padding: conflict-free transposes)
    static constexpr int VS = N * PL;
    static constexpr int EBUF = M * VS;
    static constexpr int EPB = ICE_MINV_LINES_THREADS / NN;
    static constexpr int BLOCK = ((EPB * NN + 63) / 64) * 64;
    static constexpr int MINW = (N * M <= ICE_MINV_LINES_CAPVALS) ? ICE_MINV_LINES_MINW : 1;
};

template <int P, int M, class KT = double>
__global__ __launch_bounds__((MinvLines<P, M, KT>::BLOCK), (MinvLines<P, M, KT>::MINW)) void minv_lines_kernel(DevSpace S,
                                                                                  const double* __restrict__ r,
                                                                                  KT* __restrict__ k,
                                                                                  StageComb<KT> cb) {
    constexpr int D = 3;
    using LL = MinvLines<P, M, KT>;
    constexpr int N = LL::N, NN = LL::NN, NB = LL::NB, PL = LL::PL, VS = LL::VS, EBUF = LL::EBUF, EPB = LL::EPB;
    constexpr int NV = 1 << D;
    __shared__ double buf[EPB * EBUF];
    __shared__ double sX[EPB * NV * D];
    __shared__ double sB[N * N];
    __shared__ int sing[EPB];
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / NN;
    const int t = threadIdx.x - le * NN;
    const int a = t / N, b = t - (t / N) * N;  // the line's two fixed indices
    const int e = e0 + le;
    const bool active = threadIdx.x < EPB * NN && e < S.nelem;
    const int lec = le < EPB ? le : 0;
    double* const B = buf + (size_t)lec * EBUF;
    const double* xs = S.tab;
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    double y[N][M];
    {
        const size_t xb = (size_t)e0 * NV * D, xt = (size_t)S.nelem * NV * D;
        for (int idx = threadIdx.x; idx < EPB * NV * D; idx += blockDim.x)
            sX[idx] = (xb + idx < xt) ? S.X[xb + idx] : 0.0;
        for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) sB[idx] = wgl[N + idx];  // Binv[i][q]
        if (threadIdx.x < EPB) sing[threadIdx.x] = 0;
        // the line (a, b, c), c < N, of the residual
        const double* rl = r + ((size_t)(active ? e : 0) * NB + a * NN + b * N) * M;
        double in[N][M];
#pragma unroll
        for (int c = 0; c < N; ++c)
#pragma unroll
            for (int v = 0; v < M; ++v) in[c][v] = active ? rl[c * M + v] : 0.0;
        __syncthreads();
        // B^-T along axis 2: y[q] = sum_i Binv[i][q] in[i]
#pragma unroll
        for (int q = 0; q < N; ++q)
#pragma unroll
            for (int v = 0; v < M; ++v) {
                double acc = 0;
#pragma unroll
                for (int i = 0; i < N; ++i) acc += sB[i * N + q] * in[i][v];
                y[q][v] = acc;
            }
    }
    if (active) {
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int q = 0; q < N; ++q) B[v * VS + a * PL + b * N + q] = y[q][v];
    }
    __syncthreads();
    // B^-T along axis 1: line (a, ., b)
    if (active) {
        double in[N][M];
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) in[i][v] = B[v * VS + a * PL + i * N + b];
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int q = 0; q < N; ++q) {
                double acc = 0;
#pragma unroll
                for (int i = 0; i < N; ++i) acc += sB[i * N + q] * in[i][v];
                B[v * VS + a * PL + q * N + b] = acc;
            }
    }
    __syncthreads();
    // B^-T along axis 0: line (., a, b); scale by 1 / (w_q detJ_q); B^-1 along
    // axis 0 in the same registers
    if (active) {
        double in[N][M];
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) in[i][v] = B[v * VS + i * PL + a * N + b];
        const double* Xe = sX + lec * NV * D;
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const double xi[D] = {xs[q], xs[a], xs[b]};
            const double det = q1_det<D>(Xe, xi);
            if (det == 0.0) sing[lec] = 1;
            const double sc = 1.0 / (wgl[q] * wgl[a] * wgl[b] * det);
#pragma unroll
            for (int v = 0; v < M; ++v) {
                double acc = 0;
#pragma unroll
                for (int i = 0; i < N; ++i) acc += sB[i * N + q] * in[i][v];
                y[q][v] = acc * sc;
            }
        }
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double acc = 0;
#pragma unroll
                for (int q = 0; q < N; ++q) acc += sB[i * N + q] * y[q][v];
                B[v * VS + i * PL + a * N + b] = acc;
            }
    }
    __syncthreads();
    // B^-1 along axis 1: line (a, ., b)
    if (active) {
        double in[N][M];
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int q = 0; q < N; ++q) in[q][v] = B[v * VS + a * PL + q * N + b];
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double acc = 0;
#pragma unroll
                for (int q = 0; q < N; ++q) acc += sB[i * N + q] * in[q][v];
                B[v * VS + a * PL + i * N + b] = acc;
            }
    }
    __syncthreads();
    if (!active) return;
    // B^-1 along axis 2: line (a, b, .), written out as a line of the result
    double in[N][M];
#pragma unroll
    for (int v = 0; v < M; ++v)
#pragma unroll
        for (int q = 0; q < N; ++q) in[q][v] = B[v * VS + a * PL + b * N + q];
    const size_t o = ((size_t)e * NB + a * NN + b * N) * M;
    const bool sg = sing[lec] != 0;
    KT kn[N][M];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int v = 0; v < M; ++v) {
            double acc = 0;
#pragma unroll
            for (int q = 0; q < N; ++q) acc += sB[i * N + q] * in[q][v];
            kn[i][v] = sg ? (KT)(1e-8 * r[o + i * M + v]) : (KT)acc;
        }
    if (!cb.out) {
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int v = 0; v < M; ++v) k[o + i * M + v] = kn[i][v];
        return;
    }
    // (the last stage updates the state in place: out is base)
    const double dt = (cb.stp && cb.stp->done) ? 0.0 : (cb.dte ? cb.dte[e] : *cb.dtp);
    stage_combine_line<N * M, KT>(cb, k, o, dt, &kn[0][0]);
}

// The same in 2D: a thread holds the line (t, .) of N nodes, applies B^-T
// along axis 1 in registers, and after one LDS transpose the column (., t):
// B^-T along axis 0, the scaling and B^-1 along axis 0 without leaving
// registers; a transpose back and B^-1 along axis 1 end on the line it
// read.  Two barriers instead of the node-per-thread kernel's four LDS
// passes, and N M loads and stores in flight per thread.
template <int P, int M>
struct MinvLines2 {
    static constexpr int N = P + 1, NB = N * N;
    static constexpr int PL = N + (N % 2 == 0 ? 1 : 0);  // odd line stride
    static constexpr int VS = N * PL;
    static constexpr int EBUF = M * VS;
    static constexpr int EPB = ICE_MINV_LINES2_THREADS / N;
    static constexpr int BLOCK = ((EPB * N + 63) / 64) * 64;
    // waves per SIMD the registers must allow with several variables (the
    // three stage-combination paths otherwise cost P2-P3 a wave)
    static constexpr int MINW = (M > 1 && P <= 2) ? 5 : (M > 1 && P == 3) ? 4 : 1;
};

template <int P, int M, class KT = double>
__global__ __launch_bounds__((MinvLines2<P, M>::BLOCK), (MinvLines2<P, M>::MINW)) void minv_lines2_kernel(DevSpace S,
                                                                                const double* __restrict__ r,
                                                                                KT* __restrict__ k,
                                                                                StageComb<KT> cb) {
    constexpr int D = 2;
    using LL = MinvLines2<P, M>;
    constexpr int N = LL::N, NB = LL::NB, PL = LL::PL, VS = LL::VS, EBUF = LL::EBUF, EPB = LL::EPB;
    constexpr int NV = 1 << D;
    __shared__ double buf[EPB * EBUF];
    __shared__ double sX[EPB * NV * D];
    __shared__ double sB[N * N];
    __shared__ int sing[EPB];
    const int e0 = blockIdx.x * EPB;
    const int le = threadIdx.x / N;
    const int t = threadIdx.x - le * N;
    const int e = e0 + le;
    const bool active = threadIdx.x < EPB * N && e < S.nelem;
    const int lec = le < EPB ? le : 0;
    double* const B = buf + (size_t)lec * EBUF;
    const double* xs = S.tab;
    const double* wgl = S.tab + (N + 2) + 3 * (N + 2) * N;
    const size_t o = ((size_t)(active ? e : 0) * NB + t * N) * M;  // line (t, .)
    {
        const size_t xb = (size_t)e0 * NV * D, xt = (size_t)S.nelem * NV * D;
        for (int idx = threadIdx.x; idx < EPB * NV * D; idx += blockDim.x)
            sX[idx] = (xb + idx < xt) ? S.X[xb + idx] : 0.0;
        for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) sB[idx] = wgl[N + idx];  // Binv[i][q]
        if (threadIdx.x < EPB) sing[threadIdx.x] = 0;
        double in[N][M];
#pragma unroll
        for (int c = 0; c < N; ++c)
#pragma unroll
            for (int v = 0; v < M; ++v) in[c][v] = active ? r[o + c * M + v] : 0.0;
        __syncthreads();
        // B^-T along axis 1: y[q] = sum_i Binv[i][q] in[i]
        if (active) {
#pragma unroll
            for (int v = 0; v < M; ++v)
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    double acc = 0;
#pragma unroll
                    for (int i = 0; i < N; ++i) acc += sB[i * N + q] * in[i][v];
                    B[v * VS + t * PL + q] = acc;
                }
        }
    }
    __syncthreads();
    // column (., t): B^-T along axis 0, 1 / (w_q detJ_q), B^-1 along axis 0
    if (active) {
        double in[N][M], z[N][M];
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) in[i][v] = B[v * VS + i * PL + t];
        const double* Xe = sX + lec * NV * D;
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const double xi[D] = {xs[q], xs[t]};
            const double det = q1_det<D>(Xe, xi);
            if (det == 0.0) sing[lec] = 1;
            const double sc = 1.0 / (wgl[q] * wgl[t] * det);
#pragma unroll
            for (int v = 0; v < M; ++v) {
                double acc = 0;
#pragma unroll
                for (int i = 0; i < N; ++i) acc += sB[i * N + q] * in[i][v];
                z[q][v] = acc * sc;
            }
        }
#pragma unroll
        for (int v = 0; v < M; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double acc = 0;
#pragma unroll
                for (int q = 0; q < N; ++q) acc += sB[i * N + q] * z[q][v];
                B[v * VS + i * PL + t] = acc;
            }
    }
    __syncthreads();
    if (!active) return;
    // line (t, .): B^-1 along axis 1
    double in[N][M];
#pragma unroll
    for (int v = 0; v < M; ++v)
#pragma unroll
        for (int q = 0; q < N; ++q) in[q][v] = B[v * VS + t * PL + q];
    const bool sg = sing[lec] != 0;
    KT kn[N][M];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int v = 0; v < M; ++v) {
            double acc = 0;
#pragma unroll
            for (int q = 0; q < N; ++q) acc += sB[i * N + q] * in[q][v];
            kn[i][v] = sg ? (KT)(1e-8 * r[o + i * M + v]) : (KT)acc;
        }
    if (cb.out) {
        const double dt = (cb.stp && cb.stp->done) ? 0.0 : (cb.dte ? cb.dte[e] : *cb.dtp);
        stage_combine_line<N * M, KT>(cb, k, o, dt, &kn[0][0]);
        return;
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int v = 0; v < M; ++v) k[o + i * M + v] = kn[i][v];
}

// y += coef * (*dt) * x   (axpy, fespan.hpp:283-300, with the step size read on the device)
__global__ void axpy_dt_kernel(size_t n, double coef, const double* __restrict__ dtp, const double* __restrict__ x,
                               double* __restrict__ y) {
    const double a = coef * (dtp ? *dtp : 1.0);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

// y += coef * dt_e * x with a step size per element (batched instances);
// per = entries per element; nothing once the stepper's run is done
__global__ void axpy_edt_kernel(size_t n, double coef, const double* __restrict__ dte, size_t per,
                                const StepState* __restrict__ stp, const double* __restrict__ x,
                                double* __restrict__ y) {
    if (stp->done) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] += coef * dte[i / per] * x[i];
}

// out = base + dt_e (c0 k0 + c1 k1 + c2 k2): one pass for a whole RK stage
// combination (nk = 1..3 stage vectors), the step size read on the device
// (the stepper's, or each element's; zero once the run is done).  out may
// alias base.
// (T = float for FP32 states: the combination is formed in FP64 and rounded once)
template <class T>
__global__ void rk_combine_kernel(size_t n, const T* __restrict__ base, int nk, const T* __restrict__ k0,
                                  const T* __restrict__ k1, const T* __restrict__ k2, double c0, double c1,
                                  double c2, const double* __restrict__ dtp, const double* __restrict__ dte,
                                  size_t per, const StepState* __restrict__ stp, T* out) {
    const bool done = stp && stp->done;
    const double dts = dtp ? *dtp : 1.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double dt = done ? 0.0 : (dte ? dte[i / per] : dts);
        double s = c0 * (double)k0[i];
        if (nk > 1) s += c1 * (double)k1[i];
        if (nk > 2) s += c2 * (double)k2[i];
        out[i] = (T)((double)base[i] + dt * s);
    }
}

// y = a x + b y  (axpby, fespan.hpp:312-330)
__global__ void axpby_kernel(size_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

// Time-step control on the device, so steps run without host round trips.
// CFLTimestep (explicit_utils.hpp:61-90) with the flux family's dt_from_cfl
// (burgers.hpp:103-108 / navier_stokes.hpp:1191-1203), then the termination
// condition's limit_dt (explicit_utils.hpp:123-160).
__global__ void step_begin_kernel(StepState* st, unsigned long long* wave, StepControl c) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const bool done = c.use_tfinal ? (st->time >= c.tfinal) : (st->itime >= c.ntime);
    st->done = done ? 1 : 0;
    double dt = c.dt_fixed;
    if (c.use_cfl) {
        const double lam_run = __longlong_as_double((long long)wave[0]);
        const double lam_last = __longlong_as_double((long long)wave[1]);
        const double visc_max = __longlong_as_double((long long)wave[2]);
        if (c.flux == ICE_FLUX_BURGERS) {
            dt = (c.reflen * c.cfl) / (c.mu / c.reflen + lam_run);
        } else {
            dt = (c.reflen * c.cfl) / lam_last;
            if (c.flux == ICE_FLUX_NAVIER_STOKES) dt = fmin(dt, c.reflen * c.reflen * c.cfl / visc_max);
            wave[1] = 0ull;  // dt_from_cfl resets lambda_max and visc_max
            wave[2] = 0ull;
        }
        dt /= (2 * c.pmax + 1);
    }
    if (c.use_tfinal) dt = fmin(dt, c.tfinal - st->time);
    st->dt = done ? 0.0 : dt;
}

__global__ void step_end_kernel(StepState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!st->done) {
        st->time += st->dt;
        st->itime += 1;
    }
}

// k = M^-1 r, and with cb.out set the stage combination in the same launch
// (k is then stored only if cb.store_k)
template <int D, int P, int M, class KT>
int launch_minv_t(const DevSpace& S, const double* r, KT* k, const StageComb<KT>& cb, hipStream_t st) {
    if constexpr (D == 3 && P >= 2) {
        using LL = MinvLines<P, M, KT>;
        minv_lines_kernel<P, M, KT><<<(S.nelem + LL::EPB - 1) / LL::EPB, LL::BLOCK, 0, st>>>(S, r, k, cb);
        return (int)hipGetLastError();
    }
    else if constexpr (D == 2) {
        using LL = MinvLines2<P, M>;
        minv_lines2_kernel<P, M, KT><<<(S.nelem + LL::EPB - 1) / LL::EPB, LL::BLOCK, 0, st>>>(S, r, k, cb);
        return (int)hipGetLastError();
    } else {
        constexpr int NB = PowS<P + 1, D>::value;
        // workgroup: ICE_MINV_TPB threads, or whole waves for one element
        // when an element has more nodes
        constexpr int TPB = (NB <= ICE_MINV_TPB) ? ICE_MINV_TPB : ((NB + 63) / 64) * 64;
        constexpr int EPB = TPB / NB;
        minv_kernel<D, P, M, TPB, KT><<<(S.nelem + EPB - 1) / EPB, TPB, 0, st>>>(S, r, k, cb);
        return (int)hipGetLastError();
    }
}

template <int D, int M, class KT>
int minv_order(int p, const DevSpace& S, const double* r, KT* k, const StageComb<KT>& cb, hipStream_t st) {
    switch (p) {
    case 1: return launch_minv_t<D, 1, M>(S, r, k, cb, st);
    case 2: return launch_minv_t<D, 2, M>(S, r, k, cb, st);
    case 3: return launch_minv_t<D, 3, M>(S, r, k, cb, st);
    case 4: return launch_minv_t<D, 4, M>(S, r, k, cb, st);
    default: return -1;
    }
}

template <class KT>
int minv_dispatch(int ndim, int p, int neq, const DevSpace& S, const double* r, KT* k, hipStream_t st,
                  const StageComb<KT>& cb = StageComb<KT>{}) {
    if (ndim == 1 && neq == 1) return minv_order<1, 1>(p, S, r, k, cb, st);
    if (ndim == 1 && neq == 3) return minv_order<1, 3>(p, S, r, k, cb, st);
    if (ndim == 2 && neq == 1) return minv_order<2, 1>(p, S, r, k, cb, st);
    if (ndim == 2 && neq == 4) return minv_order<2, 4>(p, S, r, k, cb, st);
    if (ndim == 3 && neq == 1) return minv_order<3, 1>(p, S, r, k, cb, st);
    if (ndim == 3 && neq == 5) return minv_order<3, 5>(p, S, r, k, cb, st);
    return -1;
}

int launch_minv(int ndim, int p, int neq, const DevSpace& S, const double* r, double* k, hipStream_t st) {
    return minv_dispatch(ndim, p, neq, S, r, k, st);
}

int launch_minv_f32(int ndim, int p, int neq, const DevSpace& S, const double* r, float* k, hipStream_t st) {
    if (ndim == 1) return -1;  // FP32 states are 2D and 3D
    return minv_dispatch(ndim, p, neq, S, r, k, st);
}

// k = M^-1 r and out = base + dt (c[0] k0 + .. + c[nprev] k) in one launch;
// k receives the increment only if store_k (a later stage reads it).
template <class KT>
int minv_combine_t(int ndim, int p, int neq, const DevSpace& S, const double* r, KT* k, const KT* base, int nprev,
                   const KT* const* kk, const double* c, const double* dtp, const double* dte,
                   const StepState* stp, KT* out, bool store_k, hipStream_t st) {
    StageComb<KT> cb;
    cb.out = out;
    cb.base = base;
    cb.k0 = nprev > 0 ? kk[0] : nullptr;
    cb.k1 = nprev > 1 ? kk[1] : nullptr;
    cb.nprev = nprev;
    cb.c0 = nprev > 0 ? c[0] : 0.0;
    cb.c1 = nprev > 1 ? c[1] : 0.0;
    cb.cn = c[nprev];
    cb.dtp = dtp;
    cb.dte = dte;
    cb.stp = stp;
    cb.store_k = store_k;
    return minv_dispatch(ndim, p, neq, S, r, k, st, cb);
}

int launch_minv_combine(int ndim, int p, int neq, const DevSpace& S, const double* r, double* k, const double* base,
                        int nprev, const double* const* kk, const double* c, const double* dtp, const double* dte,
                        const StepState* stp, double* out, bool store_k, hipStream_t st) {
    return minv_combine_t<double>(ndim, p, neq, S, r, k, base, nprev, kk, c, dtp, dte, stp, out, store_k, st);
}

int launch_minv_combine_f32(int ndim, int p, int neq, const DevSpace& S, const double* r, float* k, const float* base,
                            int nprev, const float* const* kk, const double* c, const double* dtp, const double* dte,
                            const StepState* stp, float* out, bool store_k, hipStream_t st) {
    return minv_combine_t<float>(ndim, p, neq, S, r, k, base, nprev, kk, c, dtp, dte, stp, out, store_k, st);
}

static unsigned grid_for(size_t n) {
    const size_t b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

int launch_axpy_dt(size_t n, double coef, const double* dtp, const double* x, double* y, hipStream_t st) {
    axpy_dt_kernel<<<grid_for(n), 256, 0, st>>>(n, coef, dtp, x, y);
    return (int)hipGetLastError();
}

int launch_axpy_edt(size_t n, double coef, const double* dte, size_t per, const StepState* stp, const double* x,
                    double* y, hipStream_t st) {
    axpy_edt_kernel<<<grid_for(n), 256, 0, st>>>(n, coef, dte, per, stp, x, y);
    return (int)hipGetLastError();
}

int launch_rk_combine(size_t n, const double* base, int nk, const double* const* k, const double* c, const double* dtp,
                      const double* dte, size_t per, const StepState* stp, double* out, hipStream_t st) {
    rk_combine_kernel<double><<<grid_for(n), 256, 0, st>>>(n, base, nk, k[0], nk > 1 ? k[1] : nullptr,
                                                            nk > 2 ? k[2] : nullptr, c[0], nk > 1 ? c[1] : 0.0,
                                                            nk > 2 ? c[2] : 0.0, dtp, dte, per, stp, out);
    return (int)hipGetLastError();
}

int launch_rk_combine_f32(size_t n, const float* base, int nk, const float* const* k, const double* c,
                          const double* dtp, const double* dte, size_t per, const StepState* stp, float* out,
                          hipStream_t st) {
    rk_combine_kernel<float><<<grid_for(n), 256, 0, st>>>(n, base, nk, k[0], nk > 1 ? k[1] : nullptr,
                                                           nk > 2 ? k[2] : nullptr, c[0], nk > 1 ? c[1] : 0.0,
                                                           nk > 2 ? c[2] : 0.0, dtp, dte, per, stp, out);
    return (int)hipGetLastError();
}

__global__ void widen_kernel(size_t n, const float* __restrict__ x, double* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = (double)x[i];
}

int launch_widen(size_t n, const float* x, double* y, hipStream_t st) {
    widen_kernel<<<grid_for(n), 256, 0, st>>>(n, x, y);
    return (int)hipGetLastError();
}

int launch_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t st) {
    axpby_kernel<<<grid_for(n), 256, 0, st>>>(n, a, x, b, y);
    return (int)hipGetLastError();
}

int launch_step_begin(StepState* s, unsigned long long* wave, const StepControl& c, hipStream_t st) {
    step_begin_kernel<<<1, 64, 0, st>>>(s, wave, c);
    return (int)hipGetLastError();
}

int launch_step_end(StepState* s, hipStream_t st) {
    step_end_kernel<<<1, 64, 0, st>>>(s);
    return (int)hipGetLastError();
}

}  // namespace ice
