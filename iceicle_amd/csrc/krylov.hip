// Linear algebra of the Newton solve on the device: block-sparse products
// with the element-block Jacobian, block-Jacobi factors, and the vector
// kernels of restarted GMRES.  Replaces the PETSc Mat/KSP/PC objects the
// reference's PetscNewton drives (petsc_newton.hpp:154-244).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace ice {

namespace {

// y[e] = sum_s J[e][s] x[cols[e][s]]  (or b[e] - that).  Lanes of a wave are
// split into groups of G (a power of two >= min(bs, 64)); each group owns one
// block row at a time and reads it contiguously, so a wave reads 64/G rows
// with one coalesced load per slot.  The x segments of the row's column
// elements are staged in LDS.
template <class T, class V, int NS>
__global__ __launch_bounds__(256) void bsr_kernel(int nelem, int bs, int G, int epb, const int* __restrict__ cols,
                                                  const T* __restrict__ J, const V* __restrict__ x,
                                                  const V* __restrict__ b, V* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* sx = reinterpret_cast<V*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wpe = 4 / epb;
    const int le = wave / wpe, wi = wave % wpe;
    const int e0 = blockIdx.x * epb;
    const int nstage = epb * NS * bs;
    for (int k = threadIdx.x; k < nstage; k += blockDim.x) {
        const int lel = k / (NS * bs), rem = k % (NS * bs), s = rem / bs, c = rem % bs;
        const int ee = e0 + lel;
        V v = 0;
        if (ee < nelem) {
            const int col = cols[(size_t)ee * NS + s];
            if (col >= 0) v = x[(size_t)col * bs + c];
        }
        sx[k] = v;
    }
    __syncthreads();
    const int e = e0 + le;
    if (e >= nelem) return;
    int ecol[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) ecol[s] = cols[(size_t)e * NS + s];
    const int rpw = 64 / G;
    const int sub = lane / G, gl = lane % G;
    const V* xs = sx + (size_t)le * NS * bs;
    for (int r0 = wi * rpw; r0 < bs; r0 += wpe * rpw) {
        const int r = r0 + sub;
        V acc = 0;
        if (r < bs) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (ecol[s] < 0) continue;
                const T* Jr = J + (((size_t)e * NS + s) * bs + r) * bs;
                for (int c = gl; c < bs; c += G) acc += (V)Jr[c] * xs[s * bs + c];
            }
        }
        for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (gl == 0 && r < bs) {
            const size_t i = (size_t)e * bs + r;
            y[i] = b ? b[i] - acc : acc;
        }
    }
}

int group_width(int bs) {
    int g = 1;
    while (g < bs && g < 64) g <<= 1;
    return g;
}

}  // namespace

template <class T, class V>
int bsr_apply(int nelem, int nslots, int bs, const int* cols, const T* J, const V* x, const V* b, V* y,
              hipStream_t st) {
    const int G = group_width(bs);
    const int epb = (bs >= 32) ? 1 : 4;
    const int grid = (nelem + epb - 1) / epb;
    const size_t sh = (size_t)epb * nslots * bs * sizeof(V);
    switch (nslots) {
    case 3: bsr_kernel<T, V, 3><<<grid, 256, sh, st>>>(nelem, bs, G, epb, cols, J, x, b, y); break;
    case 5: bsr_kernel<T, V, 5><<<grid, 256, sh, st>>>(nelem, bs, G, epb, cols, J, x, b, y); break;
    case 7: bsr_kernel<T, V, 7><<<grid, 256, sh, st>>>(nelem, bs, G, epb, cols, J, x, b, y); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}

template int bsr_apply<double, double>(int, int, int, const int*, const double*, const double*, const double*,
                                       double*, hipStream_t);
template int bsr_apply<float, float>(int, int, int, const int*, const float*, const float*, const float*, float*,
                                     hipStream_t);
template int bsr_apply<float, double>(int, int, int, const int*, const float*, const double*, const double*,
                                      double*, hipStream_t);


// ---------------------------------------------------------------------------
// Block-Jacobi factors: in-place Gauss-Jordan inversion with partial pivoting
// of every diagonal block (slot 0 of each block row), one workgroup per
// block, the block held in LDS.  A block whose pivot vanishes is replaced by
// the identity and counted in *nsingular.

namespace {

// In-place inverse of the bs x bs matrix A in LDS by Gauss-Jordan elimination
// with partial pivoting (largest |A[i][k]|, lowest row on ties), by the whole
// workgroup; perm (LDS, bs ints) records the row interchanges.  Returns
// false, uniformly across the workgroup, when a pivot is zero.
__device__ bool gj_inverse_lds(double* A, int* perm, int bs, int* s_piv, int* s_bad) {
    const int tid = threadIdx.x;
    if (tid == 0) *s_bad = 0;
    __syncthreads();
    for (int k = 0; k < bs; ++k) {
        if (tid < 64) {
            double best = -1.0;
            int bi = k;
            for (int i = k + tid; i < bs; i += 64) {
                const double a = fabs(A[(size_t)i * bs + k]);
                if (a > best) { best = a; bi = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (tid == 0) {
                *s_piv = bi;
                perm[k] = bi;
                if (!(best > 0.0)) *s_bad = 1;
            }
        }
        __syncthreads();
        if (*s_bad) return false;
        const int pr = *s_piv;
        if (pr != k)
            for (int c = tid; c < bs; c += blockDim.x) {
                const double t = A[(size_t)k * bs + c];
                A[(size_t)k * bs + c] = A[(size_t)pr * bs + c];
                A[(size_t)pr * bs + c] = t;
            }
        __syncthreads();
        const double ipiv = 1.0 / A[(size_t)k * bs + k];
        __syncthreads();
        for (int c = tid; c < bs; c += blockDim.x) {
            const double v = (c == k) ? 1.0 : A[(size_t)k * bs + c];
            A[(size_t)k * bs + c] = v * ipiv;
        }
        __syncthreads();
        // eliminate column k from every other row
        for (int idx = tid; idx < bs * bs; idx += blockDim.x) {
            const int i = idx / bs, c = idx % bs;
            if (i == k || c == k) continue;
            A[(size_t)i * bs + c] -= A[(size_t)i * bs + k] * A[(size_t)k * bs + c];
        }
        __syncthreads();
        for (int i = tid; i < bs; i += blockDim.x)
            if (i != k) A[(size_t)i * bs + k] = -A[(size_t)i * bs + k] * A[(size_t)k * bs + k];
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int k = bs - 1; k >= 0; --k) {
        const int pr = perm[k];
        if (pr != k)
            for (int i = tid; i < bs; i += blockDim.x) {
                const double t = A[(size_t)i * bs + k];
                A[(size_t)i * bs + k] = A[(size_t)i * bs + pr];
                A[(size_t)i * bs + pr] = t;
            }
        __syncthreads();
    }
    return true;
}

template <class T>
__global__ __launch_bounds__(256) void block_inverse_kernel(int nelem, int nslots, int bs, const double* __restrict__ J,
                                                            T* __restrict__ Dinv, int* __restrict__ nsingular,
                                                            int transposed) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* A = reinterpret_cast<double*>(smem_raw);  // bs * bs
    int* perm = reinterpret_cast<int*>(A + (size_t)bs * bs);
    __shared__ int s_piv;
    __shared__ int s_bad;
    const int e = blockIdx.x;
    const int tid = threadIdx.x;
    const double* src = J + (size_t)e * nslots * bs * bs;
    for (int k = tid; k < bs * bs; k += blockDim.x) A[k] = src[k];
    __syncthreads();
    const bool ok = gj_inverse_lds(A, perm, bs, &s_piv, &s_bad);
    T* dst = Dinv + (size_t)e * bs * bs;
    if (!ok) {
        for (int k = tid; k < bs * bs; k += blockDim.x) dst[k] = (T)((k / bs == k % bs) ? 1.0 : 0.0);
        if (tid == 0) atomicAdd(nsingular, 1);
        return;
    }
    if (transposed) {
        for (int k = tid; k < bs * bs; k += blockDim.x) dst[k] = (T)A[(size_t)(k % bs) * bs + k / bs];
    } else {
        for (int k = tid; k < bs * bs; k += blockDim.x) dst[k] = (T)A[k];
    }
}

// Block LU of a block-tridiagonal matrix (a chain of elements e = 0..n-1,
// slot 1 = element e-1, slot 2 = element e+1), one wave walking the chain:
// S_0 = D_0, S_e = D_e - L_e W_{e-1}, W_e = S_e^-1 U_e, with L_e = J[e][1],
// U_e = J[e][2].  Stores S_e^-1 and W_e.  A singular S_e is replaced by the
// identity and counted.  The chain is sequential, so its cost is latency:
// the three blocks of element e+1 are loaded into registers while element e
// is factored and stored to LDS (double buffer) at the end of the step.
constexpr int LINE_STAGE = 12;  // staged values per lane: 3 bs^2 <= 64 * 12 (bs <= 16)
__global__ __launch_bounds__(64) void line_factor_kernel(int nelem, int nslots, int bs, const double* __restrict__ J,
                                                          double* __restrict__ Sinv, double* __restrict__ W,
                                                          int* __restrict__ nsingular) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int bb = bs * bs;
    double* A = reinterpret_cast<double*>(smem_raw);
    double* Wp = A + bb;
    double* blk = Wp + bb;  // [2][D | L | U]
    int* perm = reinterpret_cast<int*>(blk + 6 * bb);
    __shared__ int s_piv, s_bad;
    const int tid = threadIdx.x;
    auto src = [&](int e, int k) {  // k-th of element e's D, L, U values (slots 0, 1, 2)
        return J[(size_t)e * nslots * bb + k];
    };
    for (int k = tid; k < 3 * bb; k += 64) blk[k] = src(0, k);
    __syncthreads();
    double stage[LINE_STAGE];
    for (int e = 0; e < nelem; ++e) {
        const double* cur = blk + (e & 1) * 3 * bb;
        if (e + 1 < nelem) {
#pragma unroll
            for (int r = 0; r < LINE_STAGE; ++r) {
                const int k = tid + r * 64;
                if (k < 3 * bb) stage[r] = src(e + 1, k);
            }
        }
        for (int idx = tid; idx < bb; idx += 64) {
            const int i = idx / bs, c = idx % bs;
            double sacc = cur[idx];
            if (e > 0)
                for (int k = 0; k < bs; ++k) sacc -= cur[bb + i * bs + k] * Wp[k * bs + c];
            A[idx] = sacc;
        }
        __syncthreads();
        if (!gj_inverse_lds(A, perm, bs, &s_piv, &s_bad)) {
            for (int idx = tid; idx < bb; idx += 64) A[idx] = (idx / bs == idx % bs) ? 1.0 : 0.0;
            if (tid == 0) atomicAdd(nsingular, 1);
            __syncthreads();
        }
        for (int idx = tid; idx < bb; idx += 64) Sinv[(size_t)e * bb + idx] = A[idx];
        if (e + 1 < nelem) {
            for (int idx = tid; idx < bb; idx += 64) {
                const int i = idx / bs, c = idx % bs;
                double sacc = 0.0;
                for (int k = 0; k < bs; ++k) sacc += A[i * bs + k] * cur[2 * bb + k * bs + c];
                Wp[idx] = sacc;
                W[(size_t)e * bb + idx] = sacc;
            }
            double* nxt = blk + ((e + 1) & 1) * 3 * bb;
#pragma unroll
            for (int r = 0; r < LINE_STAGE; ++r) {
                const int k = tid + r * 64;
                if (k < 3 * bb) nxt[k] = stage[r];
            }
        }
        __syncthreads();
    }
}

// x = (L U)^-1 b with the factors of line_factor_kernel: forward
// y_e = S_e^-1 (b_e - L_e y_{e-1}), backward x_e = y_e - W_e x_{e+1}.
// One wave; lane i owns row i of the current block (bs <= 16).  The next
// element's blocks are loaded into registers a step ahead, as in the
// factorisation.
__global__ __launch_bounds__(64) void line_solve_kernel(int nelem, int nslots, int bs, const double* __restrict__ J,
                                                        const double* __restrict__ Sinv, const double* __restrict__ W,
                                                        const double* __restrict__ b, double* __restrict__ x) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int bb = bs * bs;
    double* t = reinterpret_cast<double*>(smem_raw);
    double* yp = t + bs;
    const int tid = threadIdx.x;
    const bool row = tid < bs;
    // lane i holds row i of L_e and S_e^-1 (forward) or W_e (backward), and b_e[i]
    double Lr[16], Sr[16], bi = 0.0;
    double nLr[16], nSr[16], nbi = 0.0;
    auto load_fwd = [&](int e, double* l, double* sr, double& bv) {
        if (!row) return;
        const double* L = J + ((size_t)e * nslots + 1) * bb + tid * bs;
        const double* Si = Sinv + (size_t)e * bb + tid * bs;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < bs) {
                l[k] = L[k];
                sr[k] = Si[k];
            }
        bv = b[(size_t)e * bs + tid];
    };
    load_fwd(0, Lr, Sr, bi);
    for (int e = 0; e < nelem; ++e) {
        if (e + 1 < nelem) load_fwd(e + 1, nLr, nSr, nbi);
        if (row) {
            double sacc = bi;
            if (e > 0)
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (k < bs) sacc -= Lr[k] * yp[k];
            t[tid] = sacc;
        }
        __syncthreads();
        if (row) {
            double sacc = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k < bs) sacc += Sr[k] * t[k];
            x[(size_t)e * bs + tid] = sacc;
            t[bs + tid] = sacc;  // yp
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            Lr[k] = nLr[k];
            Sr[k] = nSr[k];
        }
        bi = nbi;
    }
    // yp holds x_{n-1} = y_{n-1}; lane i re-reads the y_e[i] it wrote itself
    auto load_bwd = [&](int e, double* w, double& yv) {
        if (!row || e < 0) return;
        const double* We = W + (size_t)e * bb + tid * bs;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < bs) w[k] = We[k];
        yv = x[(size_t)e * bs + tid];
    };
    load_bwd(nelem - 2, Lr, bi);
    for (int e = nelem - 2; e >= 0; --e) {
        load_bwd(e - 1, nLr, nbi);
        if (row) {
            double sacc = bi;
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k < bs) sacc -= Lr[k] * yp[k];
            t[tid] = sacc;
        }
        __syncthreads();
        if (row) {
            x[(size_t)e * bs + tid] = t[tid];
            yp[tid] = t[tid];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) Lr[k] = nLr[k];
        bi = nbi;
    }
}

// The same factorisation and solve for small blocks (bs <= 5, e.g. scalar
// 1D problems) by a single lane holding everything in registers: the chain
// is a sequence of dependent steps, and without LDS round trips and barriers
// each step is a few hundred register operations.  The next element's blocks
// are loaded a step ahead.  Gauss-Jordan on [S | I] with partial pivoting
// (largest |S[i][k]|, lowest row on ties) by selects, so no register array
// is indexed at run time.
template <int BS>
__global__ __launch_bounds__(64) void line_factor_small_kernel(int nelem, int nslots, const double* __restrict__ J,
                                                               double* __restrict__ Sinv, double* __restrict__ W,
                                                               int* __restrict__ nsingular) {
    if (threadIdx.x != 0) return;
    constexpr int BB = BS * BS;
    double Wp[BB] = {};
    double cD[BB], cL[BB], cU[BB], nD[BB], nL[BB], nU[BB];
    auto load = [&](int e, double* d, double* l, double* u) {
        const double* Je = J + (size_t)e * nslots * BB;
#pragma unroll
        for (int k = 0; k < BB; ++k) {
            d[k] = Je[k];
            l[k] = Je[BB + k];
            u[k] = Je[2 * BB + k];
        }
    };
    load(0, cD, cL, cU);
    for (int e = 0; e < nelem; ++e) {
        if (e + 1 < nelem) load(e + 1, nD, nL, nU);
        double A[BS][2 * BS];
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                double sacc = cD[i * BS + c];
                if (e > 0)
#pragma unroll
                    for (int k = 0; k < BS; ++k) sacc -= cL[i * BS + k] * Wp[k * BS + c];
                A[i][c] = sacc;
                A[i][BS + c] = (i == c) ? 1.0 : 0.0;
            }
        bool bad = false;
#pragma unroll
        for (int k = 0; k < BS; ++k) {
            int piv = k;
            double best = fabs(A[k][k]);
#pragma unroll
            for (int i = k + 1; i < BS; ++i) {
                const double a = fabs(A[i][k]);
                if (a > best) { best = a; piv = i; }
            }
            if (!(best > 0.0)) bad = true;
#pragma unroll
            for (int i = k + 1; i < BS; ++i)
#pragma unroll
                for (int c = 0; c < 2 * BS; ++c) {
                    const double tk = A[k][c], ti = A[i][c];
                    A[k][c] = (piv == i) ? ti : tk;
                    A[i][c] = (piv == i) ? tk : ti;
                }
            const double ip = 1.0 / A[k][k];
#pragma unroll
            for (int c = 0; c < 2 * BS; ++c) A[k][c] *= ip;
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                if (i == k) continue;
                const double f = A[i][k];
#pragma unroll
                for (int c = 0; c < 2 * BS; ++c) A[i][c] -= f * A[k][c];
            }
        }
        if (bad) {
#pragma unroll
            for (int i = 0; i < BS; ++i)
#pragma unroll
                for (int c = 0; c < BS; ++c) A[i][BS + c] = (i == c) ? 1.0 : 0.0;
            atomicAdd(nsingular, 1);
        }
        double* So = Sinv + (size_t)e * BB;
#pragma unroll
        for (int i = 0; i < BS; ++i)
#pragma unroll
            for (int c = 0; c < BS; ++c) So[i * BS + c] = A[i][BS + c];
        if (e + 1 < nelem) {
            double* Wo = W + (size_t)e * BB;
#pragma unroll
            for (int i = 0; i < BS; ++i)
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    double sacc = 0.0;
#pragma unroll
                    for (int k = 0; k < BS; ++k) sacc += A[i][BS + k] * cU[k * BS + c];
                    Wp[i * BS + c] = sacc;
                    Wo[i * BS + c] = sacc;
                }
#pragma unroll
            for (int k = 0; k < BB; ++k) {
                cD[k] = nD[k];
                cL[k] = nL[k];
                cU[k] = nU[k];
            }
        }
    }
}

template <int BS>
__global__ __launch_bounds__(64) void line_solve_small_kernel(int nelem, int nslots, const double* __restrict__ J,
                                                              const double* __restrict__ Sinv,
                                                              const double* __restrict__ W, const double* __restrict__ b,
                                                              double* __restrict__ x) {
    if (threadIdx.x != 0) return;
    constexpr int BB = BS * BS;
    double yp[BS] = {};
    double cL[BB], cS[BB], cb[BS], nL[BB], nS[BB], nb[BS];
    auto load = [&](int e, double* l, double* si, double* bv) {
        const double* L = J + ((size_t)e * nslots + 1) * BB;
#pragma unroll
        for (int k = 0; k < BB; ++k) {
            l[k] = L[k];
            si[k] = Sinv[(size_t)e * BB + k];
        }
#pragma unroll
        for (int k = 0; k < BS; ++k) bv[k] = b[(size_t)e * BS + k];
    };
    load(0, cL, cS, cb);
    for (int e = 0; e < nelem; ++e) {
        if (e + 1 < nelem) load(e + 1, nL, nS, nb);
        double t[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            double sacc = cb[i];
            if (e > 0)
#pragma unroll
                for (int k = 0; k < BS; ++k) sacc -= cL[i * BS + k] * yp[k];
            t[i] = sacc;
        }
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            double sacc = 0.0;
#pragma unroll
            for (int k = 0; k < BS; ++k) sacc += cS[i * BS + k] * t[k];
            yp[i] = sacc;
            x[(size_t)e * BS + i] = sacc;
        }
#pragma unroll
        for (int k = 0; k < BB; ++k) {
            cL[k] = nL[k];
            cS[k] = nS[k];
        }
#pragma unroll
        for (int k = 0; k < BS; ++k) cb[k] = nb[k];
    }
    // backward: x_e = y_e - W_e x_{e+1}; yp holds x_{n-1}
    double cW[BB], cy[BS], nW[BB], ny[BS];
    auto loadb = [&](int e, double* w, double* yv) {
#pragma unroll
        for (int k = 0; k < BB; ++k) w[k] = W[(size_t)e * BB + k];
#pragma unroll
        for (int k = 0; k < BS; ++k) yv[k] = x[(size_t)e * BS + k];
    };
    if (nelem >= 2) loadb(nelem - 2, cW, cy);
    for (int e = nelem - 2; e >= 0; --e) {
        if (e >= 1) loadb(e - 1, nW, ny);
        double xe[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            double sacc = cy[i];
#pragma unroll
            for (int k = 0; k < BS; ++k) sacc -= cW[i * BS + k] * yp[k];
            xe[i] = sacc;
        }
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            yp[i] = xe[i];
            x[(size_t)e * BS + i] = xe[i];
        }
#pragma unroll
        for (int k = 0; k < BB; ++k) cW[k] = nW[k];
#pragma unroll
        for (int k = 0; k < BS; ++k) cy[k] = ny[k];
    }
}

// y[e] = D[e] x[e] for dense bs x bs blocks (block-Jacobi application), one
// block row per lane group as in bsr_kernel
template <class T, class V>
__global__ __launch_bounds__(256) void block_diag_kernel(int nelem, int bs, int G, int epb, const T* __restrict__ D,
                                                         const V* __restrict__ x, V* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* sx = reinterpret_cast<V*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wpe = 4 / epb;
    const int le = wave / wpe, wi = wave % wpe;
    const int e0 = blockIdx.x * epb;
    for (int k = threadIdx.x; k < epb * bs; k += blockDim.x) {
        const int ee = e0 + k / bs;
        sx[k] = (ee < nelem) ? x[(size_t)ee * bs + k % bs] : (V)0;
    }
    __syncthreads();
    const int e = e0 + le;
    if (e >= nelem) return;
    const int rpw = 64 / G;
    const int sub = lane / G, gl = lane % G;
    const V* xs = sx + (size_t)le * bs;
    for (int r0 = wi * rpw; r0 < bs; r0 += wpe * rpw) {
        const int r = r0 + sub;
        V acc = 0;
        if (r < bs) {
            const T* Dr = D + ((size_t)e * bs + r) * bs;
            for (int c = gl; c < bs; c += G) acc += (V)Dr[c] * xs[c];
        }
        for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (gl == 0 && r < bs) y[(size_t)e * bs + r] = acc;
    }
}

__global__ void convert_slots_kernel(size_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

}  // namespace

template <class T>
int block_inverse(int nelem, int nslots, int bs, const double* J, T* Dinv, int* nsingular, bool transposed,
                  hipStream_t st) {
    const size_t sh = (size_t)bs * bs * sizeof(double) + (size_t)bs * sizeof(int);
    if (sh > 160 * 1024) return -1;
    block_inverse_kernel<T><<<nelem, 256, sh, st>>>(nelem, nslots, bs, J, Dinv, nsingular, transposed ? 1 : 0);
    return (int)hipGetLastError();
}
template int block_inverse<float>(int, int, int, const double*, float*, int*, bool, hipStream_t);
template int block_inverse<double>(int, int, int, const double*, double*, int*, bool, hipStream_t);

int line_factor(int nelem, int nslots, int bs, const double* J, double* Sinv, double* W, int* nsingular,
                hipStream_t st) {
    const size_t sh = 8 * (size_t)bs * bs * sizeof(double) + (size_t)bs * sizeof(int);
    if (bs > 16 || nslots < 3 || nelem < 1) return -1;
    switch (bs) {
    case 1: line_factor_small_kernel<1><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, nsingular); break;
    case 2: line_factor_small_kernel<2><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, nsingular); break;
    case 3: line_factor_small_kernel<3><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, nsingular); break;
    case 4: line_factor_small_kernel<4><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, nsingular); break;
    case 5: line_factor_small_kernel<5><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, nsingular); break;
    default: line_factor_kernel<<<1, 64, sh, st>>>(nelem, nslots, bs, J, Sinv, W, nsingular); break;
    }
    return (int)hipGetLastError();
}

int line_solve(int nelem, int nslots, int bs, const double* J, const double* Sinv, const double* W, const double* b,
               double* x, hipStream_t st) {
    if (bs > 16 || nslots < 3 || nelem < 1) return -1;
    switch (bs) {
    case 1: line_solve_small_kernel<1><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, b, x); break;
    case 2: line_solve_small_kernel<2><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, b, x); break;
    case 3: line_solve_small_kernel<3><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, b, x); break;
    case 4: line_solve_small_kernel<4><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, b, x); break;
    case 5: line_solve_small_kernel<5><<<1, 64, 0, st>>>(nelem, nslots, J, Sinv, W, b, x); break;
    default:
        line_solve_kernel<<<1, 64, 2 * (size_t)bs * sizeof(double), st>>>(nelem, nslots, bs, J, Sinv, W, b, x);
        break;
    }
    return (int)hipGetLastError();
}

template <class T, class V>
int block_diag_apply(int nelem, int bs, const T* D, const V* x, V* y, hipStream_t st) {
    const int G = group_width(bs);
    const int epb = (bs >= 32) ? 1 : 4;
    const int grid = (nelem + epb - 1) / epb;
    block_diag_kernel<T, V><<<grid, 256, (size_t)epb * bs * sizeof(V), st>>>(nelem, bs, G, epb, D, x, y);
    return (int)hipGetLastError();
}
template int block_diag_apply<float, double>(int, int, const float*, const double*, double*, hipStream_t);
template int block_diag_apply<double, double>(int, int, const double*, const double*, double*, hipStream_t);

int convert_to_fp32(size_t n, const double* a, float* b, hipStream_t st) {
    convert_slots_kernel<<<4096, 256, 0, st>>>(n, a, b);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Vector kernels of GMRES.  Reductions are two-stage and deterministic:
// per-workgroup partial sums in a fixed order, then one workgroup per
// result summing the partials in a fixed order.

namespace {

constexpr int VEC_EPT = 8;  // entries per thread in the reduction kernels

__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
}

// partial[blk][j] = sum_i V_j[i] w[i] for j < nv, partial[blk][nv] = w.w
__global__ __launch_bounds__(256) void mdot_partial_kernel(size_t n, int nv, const double* __restrict__ V, size_t ldv,
                                                           const double* __restrict__ w, double* __restrict__ part,
                                                           int npart) {
    __shared__ double red[4];
    const size_t base = (size_t)blockIdx.x * blockDim.x * VEC_EPT + threadIdx.x;
    double wr[VEC_EPT];
#pragma unroll
    for (int e = 0; e < VEC_EPT; ++e) {
        const size_t i = base + (size_t)e * blockDim.x;
        wr[e] = (i < n) ? w[i] : 0.0;
    }
    for (int j = 0; j <= nv; ++j) {
        double s = 0.0;
        if (j < nv) {
            const double* vj = V + (size_t)j * ldv;
#pragma unroll
            for (int e = 0; e < VEC_EPT; ++e) {
                const size_t i = base + (size_t)e * blockDim.x;
                if (i < n) s += vj[i] * wr[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < VEC_EPT; ++e) s += wr[e] * wr[e];
        }
        s = block_sum(s, red);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * npart + j] = s;
    }
}

// out[j] = sum_blk part[blk][j], one workgroup per j
__global__ __launch_bounds__(256) void reduce_partials_kernel(int nblk, const double* __restrict__ part, int npart,
                                                              double* __restrict__ out) {
    __shared__ double red[4];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(size_t)b * npart + j];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[j] = s;
}

// w -= sum_j h[j] V_j ; partial[blk] = w.w afterwards
__global__ __launch_bounds__(256) void maxpy_norm_kernel(size_t n, int nv, const double* __restrict__ V, size_t ldv,
                                                         const double* __restrict__ h, double* __restrict__ w,
                                                         double* __restrict__ part) {
    __shared__ double red[4];
    const size_t base = (size_t)blockIdx.x * blockDim.x * VEC_EPT + threadIdx.x;
    double acc[VEC_EPT];
#pragma unroll
    for (int e = 0; e < VEC_EPT; ++e) {
        const size_t i = base + (size_t)e * blockDim.x;
        acc[e] = (i < n) ? w[i] : 0.0;
    }
    for (int j = 0; j < nv; ++j) {
        const double hj = h[j];
        const double* vj = V + (size_t)j * ldv;
#pragma unroll
        for (int e = 0; e < VEC_EPT; ++e) {
            const size_t i = base + (size_t)e * blockDim.x;
            if (i < n) acc[e] -= hj * vj[i];
        }
    }
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < VEC_EPT; ++e) {
        const size_t i = base + (size_t)e * blockDim.x;
        if (i < n) {
            w[i] = acc[e];
            s += acc[e] * acc[e];
        }
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0 && part) part[blockIdx.x] = s;
}

// x += sum_j y[j] V_j
__global__ __launch_bounds__(256) void combine_kernel(size_t n, int nv, const double* __restrict__ V, size_t ldv,
                                                      const double* __restrict__ y, double* __restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double s = x[i];
        for (int j = 0; j < nv; ++j) s += y[j] * V[(size_t)j * ldv + i];
        x[i] = s;
    }
}

// x *= *s (or 1 / *s when inv)
__global__ void scale_kernel(size_t n, const double* __restrict__ s, int inv, double* __restrict__ x) {
    const double f = inv ? ((*s != 0.0) ? 1.0 / *s : 0.0) : *s;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] *= f;
}

// y = a x + b y
// y = a x + b y; a zero coefficient means the operand is not read (BLAS
// convention), so (0, y, 0, y) clears y whatever it held, NaN included
__global__ void axpby_kernel2(size_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = (a == 0.0 ? 0.0 : a * x[i]) + (b == 0.0 ? 0.0 : b * y[i]);
}

// Arnoldi bookkeeping of step k (one thread): column k of the Hessenberg
// matrix from the two Gram-Schmidt passes, the norm of the new vector,
// Givens rotations, residual estimate.  sc layout: see GmresScalars.
__global__ void arnoldi_kernel(int k, int m, const double* __restrict__ h1, const double* __restrict__ h2,
                               int refine, double* __restrict__ sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double* H = sc;                       // (m+1) x m, column-major: H[j + (m+1) k]
    double* cs = H + (size_t)(m + 1) * m;
    double* sn = cs + m;
    double* g = sn + m;                   // m+1
    double* hn = g + m + 1;               // [0] = ||v_{k+1}|| before normalisation, [1] = |g_{k+1}|
    double* col = H + (size_t)(m + 1) * k;
    for (int j = 0; j <= k; ++j) col[j] = refine ? h1[j] + h2[j] : h1[j];
    const double hk1 = sqrt(fmax(h2[k + 1], 0.0));  // h2[k+1] = squared norm after the last pass
    col[k + 1] = hk1;
    hn[0] = hk1;
    for (int j = 0; j < k; ++j) {
        const double a = col[j], b = col[j + 1];
        col[j] = cs[j] * a + sn[j] * b;
        col[j + 1] = -sn[j] * a + cs[j] * b;
    }
    const double a = col[k], b = col[k + 1];
    const double r = hypot(a, b);
    const double c = (r > 0.0) ? a / r : 1.0, s = (r > 0.0) ? b / r : 0.0;
    cs[k] = c;
    sn[k] = s;
    col[k] = r;
    col[k + 1] = 0.0;
    g[k + 1] = -s * g[k];
    g[k] = c * g[k];
    hn[1] = fabs(g[k + 1]);
}

// back substitution H(0:k,0:k) y = g(0:k), result into y
__global__ void hessenberg_solve_kernel(int k, int m, const double* __restrict__ sc, double* __restrict__ y) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double* H = sc;
    const double* g = sc + (size_t)(m + 1) * m + 2 * m;
    for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int j = i + 1; j < k; ++j) s -= H[i + (size_t)(m + 1) * j] * y[j];
        const double d = H[i + (size_t)(m + 1) * i];
        y[i] = (d != 0.0) ? s / d : 0.0;
    }
}

// g = (beta, 0, ...), beta = sqrt(*nrm2)
__global__ void gmres_init_kernel(int m, const double* __restrict__ nrm2, double* __restrict__ sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double* g = sc + (size_t)(m + 1) * m + 2 * m;
    g[0] = sqrt(fmax(*nrm2, 0.0));
    for (int j = 1; j <= m; ++j) g[j] = 0.0;
    double* hn = g + m + 1;
    hn[0] = g[0];
    hn[1] = g[0];
}

int vec_grid(size_t n) { return (int)((n + 256 * VEC_EPT - 1) / (256 * VEC_EPT)); }

}  // namespace

int gmres_partials(size_t n) { return vec_grid(n); }

int launch_mdot(size_t n, int nv, const double* V, size_t ldv, const double* w, double* part, int npart, double* out,
                hipStream_t st) {
    const int nb = vec_grid(n);
    mdot_partial_kernel<<<nb, 256, 0, st>>>(n, nv, V, ldv, w, part, npart);
    reduce_partials_kernel<<<nv + 1, 256, 0, st>>>(nb, part, npart, out);
    return (int)hipGetLastError();
}

int launch_maxpy_norm(size_t n, int nv, const double* V, size_t ldv, const double* h, double* w, double* part,
                      double* out, hipStream_t st) {
    const int nb = vec_grid(n);
    maxpy_norm_kernel<<<nb, 256, 0, st>>>(n, nv, V, ldv, h, w, part);
    if (out) reduce_partials_kernel<<<1, 256, 0, st>>>(nb, part, 1, out);
    return (int)hipGetLastError();
}

int launch_norm2(size_t n, const double* w, double* part, double* out, hipStream_t st) {
    return launch_mdot(n, 0, nullptr, 0, w, part, 1, out, st);
}

int launch_combine(size_t n, int nv, const double* V, size_t ldv, const double* y, double* x, hipStream_t st) {
    combine_kernel<<<2048, 256, 0, st>>>(n, nv, V, ldv, y, x);
    return (int)hipGetLastError();
}

int launch_scale(size_t n, const double* s, int inv, double* x, hipStream_t st) {
    scale_kernel<<<2048, 256, 0, st>>>(n, s, inv, x);
    return (int)hipGetLastError();
}

int launch_axpby2(size_t n, double a, const double* x, double b, double* y, hipStream_t st) {
    axpby_kernel2<<<2048, 256, 0, st>>>(n, a, x, b, y);
    return (int)hipGetLastError();
}

int launch_arnoldi(int k, int m, const double* h1, const double* h2, int refine, double* sc, hipStream_t st) {
    arnoldi_kernel<<<1, 64, 0, st>>>(k, m, h1, h2, refine, sc);
    return (int)hipGetLastError();
}

int launch_hessenberg_solve(int k, int m, const double* sc, double* y, hipStream_t st) {
    hessenberg_solve_kernel<<<1, 64, 0, st>>>(k, m, sc, y);
    return (int)hipGetLastError();
}

int launch_gmres_init(int m, const double* nrm2, double* sc, hipStream_t st) {
    gmres_init_kernel<<<1, 64, 0, st>>>(m, nrm2, sc);
    return (int)hipGetLastError();
}

}  // namespace ice

// ---------------------------------------------------------------------------
// Multicolour block symmetric Gauss-Seidel, M = (D + L) D^-1 (D + U) with the
// elements ordered colour by colour (no two face neighbours share a colour,
// so the rows of one colour are independent).  One launch per colour and
// sweep direction:
//   forward  (colour c):  y_e = D_e^-1 (r_e - sum_{col colour < c} J_e,col y_col)
//   backward (colour c):  y_e = y_e - D_e^-1 sum_{col colour > c} J_e,col y_col
// The row's partial sum is kept in LDS and multiplied by the factored
// diagonal block in the same workgroup.

namespace ice {
namespace {

template <class T, class V, int NS, bool FWD>
__global__ __launch_bounds__(256) void sgs_sweep_kernel(int nrows, const int* __restrict__ elems,
                                                        const int* __restrict__ ecolor, int c, int bs, int G, int epb,
                                                        const int* __restrict__ cols, const T* __restrict__ J,
                                                        const T* __restrict__ Dinv, const V* __restrict__ r,
                                                        V* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* sx = reinterpret_cast<V*>(smem_raw);        // epb * NS * bs neighbour values
    V* st = sx + (size_t)epb * NS * bs;             // epb * bs partial sums
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wpe = 4 / epb;
    const int le = wave / wpe, wi = wave % wpe;
    const int k0 = blockIdx.x * epb;
    for (int k = threadIdx.x; k < epb * NS * bs; k += blockDim.x) {
        const int lel = k / (NS * bs), rem = k % (NS * bs), s = rem / bs, cc = rem % bs;
        V v = 0;
        if (k0 + lel < nrows && s > 0) {
            const int e = elems[k0 + lel];
            const int col = cols[(size_t)e * NS + s];
            if (col >= 0) {
                const int cc2 = ecolor[col];
                if (FWD ? (cc2 < c) : (cc2 > c)) v = y[(size_t)col * bs + cc];
            }
        }
        sx[k] = v;
    }
    __syncthreads();
    const bool active = (k0 + le) < nrows;
    const int e = active ? elems[k0 + le] : 0;
    const int rpw = 64 / G;
    const int sub = lane / G, gl = lane % G;
    const V* xs = sx + (size_t)le * NS * bs;
    if (active) {
        int ecol[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int col = cols[(size_t)e * NS + s];
            ecol[s] = (s > 0 && col >= 0 && (FWD ? (ecolor[col] < c) : (ecolor[col] > c))) ? col : -1;
        }
        for (int r0 = wi * rpw; r0 < bs; r0 += wpe * rpw) {
            const int row = r0 + sub;
            V acc = 0;
            if (row < bs) {
#pragma unroll
                for (int s = 1; s < NS; ++s) {
                    if (ecol[s] < 0) continue;
                    const T* Jr = J + (((size_t)e * NS + s) * bs + row) * bs;
                    for (int cc = gl; cc < bs; cc += G) acc += (V)Jr[cc] * xs[s * bs + cc];
                }
            }
            for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
            if (gl == 0 && row < bs) st[le * bs + row] = FWD ? r[(size_t)e * bs + row] - acc : acc;
        }
    }
    __syncthreads();
    if (!active) return;
    const V* ts = st + le * bs;
    for (int r0 = wi * rpw; r0 < bs; r0 += wpe * rpw) {
        const int row = r0 + sub;
        V acc = 0;
        if (row < bs) {
            const T* Dr = Dinv + ((size_t)e * bs + row) * bs;
            for (int cc = gl; cc < bs; cc += G) acc += (V)Dr[cc] * ts[cc];
        }
        for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (gl == 0 && row < bs) {
            const size_t i = (size_t)e * bs + row;
            y[i] = FWD ? acc : y[i] - acc;
        }
    }
}

}  // namespace

template <class T, class V>
int sgs_sweep(bool forward, int nrows, const int* elems, const int* ecolor, int c, int nslots, int bs,
              const int* cols, const T* J, const T* Dinv, const V* r, V* y, hipStream_t st) {
    if (nrows <= 0) return 0;
    const int G = group_width(bs);
    const int epb = (bs >= 32) ? 1 : 4;
    const int grid = (nrows + epb - 1) / epb;
    const size_t sh = ((size_t)epb * nslots * bs + (size_t)epb * bs) * sizeof(V);
#define ICE_SGS(NSV)                                                                                               \
    case NSV:                                                                                                      \
        if (forward)                                                                                               \
            sgs_sweep_kernel<T, V, NSV, true><<<grid, 256, sh, st>>>(nrows, elems, ecolor, c, bs, G, epb, cols, J, \
                                                                     Dinv, r, y);                                  \
        else                                                                                                       \
            sgs_sweep_kernel<T, V, NSV, false><<<grid, 256, sh, st>>>(nrows, elems, ecolor, c, bs, G, epb, cols, \
                                                                      J, Dinv, r, y);                              \
        break;
    switch (nslots) {
        ICE_SGS(3)
        ICE_SGS(5)
        ICE_SGS(7)
    default: return -1;
    }
#undef ICE_SGS
    return (int)hipGetLastError();
}
template int sgs_sweep<float, double>(bool, int, const int*, const int*, int, int, int, const int*, const float*,
                                      const float*, const double*, double*, hipStream_t);
template int sgs_sweep<double, double>(bool, int, const int*, const int*, int, int, int, const int*, const double*,
                                       const double*, const double*, double*, hipStream_t);

}  // namespace ice

// ---------------------------------------------------------------------------
// Column-major ("transposed") block kernels for the Krylov operator and the
// preconditioner.  A block's entry (r, c) sits at [c][r], so the threads of
// one block row set own one row each and read a column with one coalesced
// load per column step, accumulating without any cross-lane reduction.
// Threads map to (element, row) pairs; the x segments the workgroup's rows
// need are staged in LDS.

namespace ice {
namespace {

constexpr int RT_THREADS = 256;

// elements [e_first, e_last] touched by the threads of this workgroup
__device__ __forceinline__ void rows_span(int nrows_total, int bs, int& r0, int& e_first, int& e_last) {
    r0 = blockIdx.x * RT_THREADS;
    const int r1 = min(r0 + RT_THREADS, nrows_total) - 1;
    e_first = r0 / bs;
    e_last = r1 / bs;
}

template <class T, class V, int NS>
__global__ __launch_bounds__(RT_THREADS) void bsrT_kernel(int nelem, int bs, const int* __restrict__ cols,
                                                          const T* __restrict__ JT, const V* __restrict__ x,
                                                          const V* __restrict__ b, V* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* sx = reinterpret_cast<V*>(smem_raw);  // [ne][NS][bs]
    const int ntot = nelem * bs;
    int r0, ef, el;
    rows_span(ntot, bs, r0, ef, el);
    const int ne = el - ef + 1;
    for (int k = threadIdx.x; k < ne * NS * bs; k += blockDim.x) {
        const int le = k / (NS * bs), rem = k % (NS * bs), s = rem / bs, c = rem % bs;
        const int col = cols[(size_t)(ef + le) * NS + s];
        sx[k] = (col >= 0) ? x[(size_t)col * bs + c] : (V)0;
    }
    __syncthreads();
    const int g = r0 + threadIdx.x;
    if (g >= ntot) return;
    const int e = g / bs, r = g - e * bs;
    const V* xs = sx + (size_t)(e - ef) * NS * bs;
    V acc0 = 0, acc1 = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (cols[(size_t)e * NS + s] < 0) continue;
        const T* Jc = JT + ((size_t)e * NS + s) * bs * bs + r;
        const V* xv = xs + s * bs;
        int c = 0;
        for (; c + 1 < bs; c += 2) {
            acc0 += (V)Jc[(size_t)c * bs] * xv[c];
            acc1 += (V)Jc[(size_t)(c + 1) * bs] * xv[c + 1];
        }
        if (c < bs) acc0 += (V)Jc[(size_t)c * bs] * xv[c];
    }
    const V acc = acc0 + acc1;
    y[g] = b ? b[g] - acc : acc;
}

// The same product with one wave per element block row, for block sizes bs
// = VEC * RG with RG a power of two (VEC = the 16 bytes of one load in T
// entries): lane (rg, cg) owns rows VEC rg .. VEC rg + VEC - 1 and columns
// cg, cg + CG, ... (CG = 64 / RG), so each load instruction of the wave reads
// CG whole consecutive columns of the column-major block; the partial sums
// are reduced across the column groups with cross-lane shuffles.
template <class T, class V, int NS>
__global__ __launch_bounds__(256) void bsrT_wave_kernel(int nelem, int bs, int lrg, const int* __restrict__ cols,
                                                        const T* __restrict__ JT, const V* __restrict__ x,
                                                        const V* __restrict__ b, V* __restrict__ y) {
    constexpr int VEC = 16 / (int)sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= nelem) return;  // uniform across the wave
    const int rg = lane & ((1 << lrg) - 1), cg = lane >> lrg;
    const int CG = 64 >> lrg;
    V acc[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) acc[k] = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int col = cols[(size_t)e * NS + s];
        if (col < 0) continue;
        const T* Jb = JT + ((size_t)e * NS + s) * bs * bs + rg * VEC;
        const V* xv = x + (size_t)col * bs;
#pragma unroll 4
        for (int c = cg; c < bs; c += CG) {
            const V xc = xv[c];
            const vec_t jv = *reinterpret_cast<const vec_t*>(Jb + (size_t)c * bs);
#pragma unroll
            for (int k = 0; k < VEC; ++k) acc[k] += (V)jv[k] * xc;
        }
    }
    for (int off = 1 << lrg; off < 64; off <<= 1)
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[k] += __shfl_xor(acc[k], off);
    if (cg == 0) {
        const size_t g = (size_t)e * bs + rg * VEC;
#pragma unroll
        for (int k = 0; k < VEC; ++k) y[g + k] = b ? b[g + k] - acc[k] : acc[k];
    }
}

// y = D x with column-major dense blocks D[e][c][r]
template <class T, class V>
__global__ __launch_bounds__(RT_THREADS) void diagT_kernel(int nelem, int bs, const T* __restrict__ DT,
                                                           const V* __restrict__ x, V* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* sx = reinterpret_cast<V*>(smem_raw);
    const int ntot = nelem * bs;
    int r0, ef, el;
    rows_span(ntot, bs, r0, ef, el);
    const int ne = el - ef + 1;
    for (int k = threadIdx.x; k < ne * bs; k += blockDim.x) sx[k] = x[(size_t)ef * bs + k];
    __syncthreads();
    const int g = r0 + threadIdx.x;
    if (g >= ntot) return;
    const int e = g / bs, r = g - e * bs;
    const V* xv = sx + (size_t)(e - ef) * bs;
    const T* Dc = DT + (size_t)e * bs * bs + r;
    V acc0 = 0, acc1 = 0;
    int c = 0;
    for (; c + 1 < bs; c += 2) {
        acc0 += (V)Dc[(size_t)c * bs] * xv[c];
        acc1 += (V)Dc[(size_t)(c + 1) * bs] * xv[c + 1];
    }
    if (c < bs) acc0 += (V)Dc[(size_t)c * bs] * xv[c];
    y[g] = acc0 + acc1;
}

// one colour of a multicolour block Gauss-Seidel sweep over the rows of the
// elements elems[0..nel), column-major blocks (see sgs_sweep_kernel).  A
// workgroup owns whole elements (epb of them; its threads loop over their
// rows), because the multiplication by the factored diagonal block needs the
// element's complete partial-sum vector.
// The same colour pass with one wave per element (block sizes bs = VEC RG,
// RG a power of two, as bsrT_wave_kernel): the off-diagonal products with
// 16-byte loads of whole columns, reduced across the column groups with
// shuffles into the wave's partial-sum vector in LDS, then the product with
// the factored diagonal block the same way.  Neighbour values are read from
// y directly: the colour test excludes every element written in this pass.
template <class T, class V, int NS, bool FWD, class TD>
__global__ __launch_bounds__(256) void sgsT_wave_kernel(int nel, int bs, int lrg, int lrgd,
                                                        const int* __restrict__ elems, const int* __restrict__ ecolor,
                                                        int colour, const int* __restrict__ cols,
                                                        const T* __restrict__ JT, const TD* __restrict__ DT,
                                                        const V* __restrict__ rhs, V* __restrict__ y) {
    constexpr int VEC = 16 / (int)sizeof(T);
    constexpr int VECD = 16 / (int)sizeof(TD);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    typedef TD vecd_t __attribute__((ext_vector_type(VECD)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    V* stv = reinterpret_cast<V*>(smem_raw) + (size_t)w * bs;  // this wave's partial sums
    const int l = blockIdx.x * 4 + w;
    const bool live = l < nel;  // uniform across the wave
    const int e = live ? elems[l] : 0;
    if (live) {
        const int rg = lane & ((1 << lrg) - 1), cg = lane >> lrg;
        const int CG = 64 >> lrg;
        V acc[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[k] = 0;
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            const int col = cols[(size_t)e * NS + s];
            if (col < 0 || !(FWD ? ecolor[col] < colour : ecolor[col] > colour)) continue;
            const T* Jb = JT + ((size_t)e * NS + s) * bs * bs + rg * VEC;
            const V* xv = y + (size_t)col * bs;
#pragma unroll 4
            for (int c = cg; c < bs; c += CG) {
                const V xc = xv[c];
                const vec_t jv = *reinterpret_cast<const vec_t*>(Jb + (size_t)c * bs);
#pragma unroll
                for (int k = 0; k < VEC; ++k) acc[k] += (V)jv[k] * xc;
            }
        }
        for (int off = 1 << lrg; off < 64; off <<= 1)
#pragma unroll
            for (int k = 0; k < VEC; ++k) acc[k] += __shfl_xor(acc[k], off);
        if (cg == 0) {
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                const int r = rg * VEC + k;
                stv[r] = FWD ? rhs[(size_t)e * bs + r] - acc[k] : acc[k];
            }
        }
    }
    __syncthreads();
    if (live) {
        const int rg = lane & ((1 << lrgd) - 1), cg = lane >> lrgd;
        const int CG = 64 >> lrgd;
        V acc[VECD];
#pragma unroll
        for (int k = 0; k < VECD; ++k) acc[k] = 0;
        const TD* Db = DT + (size_t)e * bs * bs + rg * VECD;
#pragma unroll 4
        for (int c = cg; c < bs; c += CG) {
            const V tc = stv[c];
            const vecd_t dv = *reinterpret_cast<const vecd_t*>(Db + (size_t)c * bs);
#pragma unroll
            for (int k = 0; k < VECD; ++k) acc[k] += (V)dv[k] * tc;
        }
        for (int off = 1 << lrgd; off < 64; off <<= 1)
#pragma unroll
            for (int k = 0; k < VECD; ++k) acc[k] += __shfl_xor(acc[k], off);
        if (cg == 0) {
#pragma unroll
            for (int k = 0; k < VECD; ++k) {
                const size_t i = (size_t)e * bs + rg * VECD + k;
                y[i] = FWD ? acc[k] : y[i] - acc[k];
            }
        }
    }
}

template <class T, class V, int NS, bool FWD, class TD>
__global__ __launch_bounds__(RT_THREADS) void sgsT_kernel(int nel, int epb, const int* __restrict__ elems,
                                                          const int* __restrict__ ecolor, int colour, int bs,
                                                          const int* __restrict__ cols, const T* __restrict__ JT,
                                                          const TD* __restrict__ DT, const V* __restrict__ rhs,
                                                          V* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* sx = reinterpret_cast<V*>(smem_raw);  // [epb][NS][bs] neighbour values
    V* st = sx + (size_t)epb * NS * bs;      // [epb][bs] partial sums
    const int l0 = blockIdx.x * epb;
    const int ne = min(epb, nel - l0);
    for (int k = threadIdx.x; k < ne * NS * bs; k += blockDim.x) {
        const int le = k / (NS * bs), rem = k % (NS * bs), s = rem / bs, c = rem % bs;
        V v = 0;
        if (s > 0) {
            const int e = elems[l0 + le];
            const int col = cols[(size_t)e * NS + s];
            if (col >= 0 && (FWD ? ecolor[col] < colour : ecolor[col] > colour)) v = y[(size_t)col * bs + c];
        }
        sx[k] = v;
    }
    __syncthreads();
    for (int g = threadIdx.x; g < ne * bs; g += blockDim.x) {
        const int le = g / bs, r = g - le * bs;
        const int e = elems[l0 + le];
        const V* xs = sx + (size_t)le * NS * bs;
        V acc0 = 0, acc1 = 0;
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            const int col = cols[(size_t)e * NS + s];
            if (col < 0 || !(FWD ? ecolor[col] < colour : ecolor[col] > colour)) continue;
            const T* Jc = JT + ((size_t)e * NS + s) * bs * bs + r;
            const V* xv = xs + s * bs;
            int c = 0;
            for (; c + 1 < bs; c += 2) {
                acc0 += (V)Jc[(size_t)c * bs] * xv[c];
                acc1 += (V)Jc[(size_t)(c + 1) * bs] * xv[c + 1];
            }
            if (c < bs) acc0 += (V)Jc[(size_t)c * bs] * xv[c];
        }
        const V acc = acc0 + acc1;
        st[g] = FWD ? rhs[(size_t)e * bs + r] - acc : acc;
    }
    __syncthreads();
    for (int g = threadIdx.x; g < ne * bs; g += blockDim.x) {
        const int le = g / bs, r = g - le * bs;
        const int e = elems[l0 + le];
        const V* tv = st + (size_t)le * bs;
        const TD* Dc = DT + (size_t)e * bs * bs + r;
        V acc0 = 0, acc1 = 0;
        int c = 0;
        for (; c + 1 < bs; c += 2) {
            acc0 += (V)Dc[(size_t)c * bs] * tv[c];
            acc1 += (V)Dc[(size_t)(c + 1) * bs] * tv[c + 1];
        }
        if (c < bs) acc0 += (V)Dc[(size_t)c * bs] * tv[c];
        const size_t i = (size_t)e * bs + r;
        y[i] = FWD ? acc0 + acc1 : y[i] - (acc0 + acc1);
    }
}

template <class TO>
__global__ void convert_transpose_kernel(size_t nblocks, int bs, const double* __restrict__ a, TO* __restrict__ b) {
    const size_t bb = (size_t)bs * bs;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nblocks * bb;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t blk = i / bb, k = i - blk * bb;
        const int c = (int)(k / bs), r = (int)(k - (size_t)c * bs);  // destination [c][r]
        b[i] = (TO)a[blk * bb + (size_t)r * bs + c];
    }
}

// elements spanned by at most RT_THREADS consecutive rows
int span_elems(int bs) { return (RT_THREADS + bs - 2) / bs + 1; }

}  // namespace

template <class T, class V>
int bsrT_apply(int nelem, int nslots, int bs, const int* cols, const T* JT, const V* x, const V* b, V* y,
               hipStream_t st) {
    constexpr int VEC = 16 / (int)sizeof(T);
    const int rgn = bs / VEC;
    if (bs >= 16 && bs % VEC == 0 && rgn <= 64 && (rgn & (rgn - 1)) == 0) {
        const int lrg = __builtin_ctz(rgn);
        const int wgrid = (nelem + 3) / 4;
        switch (nslots) {
        case 3: bsrT_wave_kernel<T, V, 3><<<wgrid, 256, 0, st>>>(nelem, bs, lrg, cols, JT, x, b, y); break;
        case 5: bsrT_wave_kernel<T, V, 5><<<wgrid, 256, 0, st>>>(nelem, bs, lrg, cols, JT, x, b, y); break;
        case 7: bsrT_wave_kernel<T, V, 7><<<wgrid, 256, 0, st>>>(nelem, bs, lrg, cols, JT, x, b, y); break;
        default: return -1;
        }
        return (int)hipGetLastError();
    }
    const long ntot = (long)nelem * bs;
    const int grid = (int)((ntot + RT_THREADS - 1) / RT_THREADS);
    const size_t sh = (size_t)span_elems(bs) * nslots * bs * sizeof(V);
    switch (nslots) {
    case 3: bsrT_kernel<T, V, 3><<<grid, RT_THREADS, sh, st>>>(nelem, bs, cols, JT, x, b, y); break;
    case 5: bsrT_kernel<T, V, 5><<<grid, RT_THREADS, sh, st>>>(nelem, bs, cols, JT, x, b, y); break;
    case 7: bsrT_kernel<T, V, 7><<<grid, RT_THREADS, sh, st>>>(nelem, bs, cols, JT, x, b, y); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}
template int bsrT_apply<float, double>(int, int, int, const int*, const float*, const double*, const double*, double*,
                                       hipStream_t);
template int bsrT_apply<_Float16, double>(int, int, int, const int*, const _Float16*, const double*, const double*, double*,
                                       hipStream_t);

template <class T, class V>
int diagT_apply(int nelem, int bs, const T* DT, const V* x, V* y, hipStream_t st) {
    const long ntot = (long)nelem * bs;
    const int grid = (int)((ntot + RT_THREADS - 1) / RT_THREADS);
    diagT_kernel<T, V><<<grid, RT_THREADS, (size_t)span_elems(bs) * bs * sizeof(V), st>>>(nelem, bs, DT, x, y);
    return (int)hipGetLastError();
}
template int diagT_apply<float, double>(int, int, const float*, const double*, double*, hipStream_t);
template int diagT_apply<double, double>(int, int, const double*, const double*, double*, hipStream_t);

template <class T, class V, class TD>
int sgsT_sweep(bool forward, int nel, const int* elems, const int* ecolor, int colour, int nslots, int bs,
               const int* cols, const T* JT, const TD* DT, const V* rhs, V* y, hipStream_t st) {
    if (nel <= 0) return 0;
    {
        constexpr int VEC = 16 / (int)sizeof(T), VECD = 16 / (int)sizeof(TD);
        const int rgn = bs / VEC, rgd = bs / VECD;
        auto pow2 = [](int v) { return v > 0 && v <= 64 && (v & (v - 1)) == 0; };
        if (bs >= 16 && bs % VEC == 0 && bs % VECD == 0 && pow2(rgn) && pow2(rgd)) {
            const int lrg = __builtin_ctz(rgn), lrgd = __builtin_ctz(rgd);
            const int wgrid = (nel + 3) / 4;
            const size_t wsh = (size_t)4 * bs * sizeof(V);
#define ICE_SGSW(NSV)                                                                                           \
    case NSV:                                                                                                   \
        if (forward)                                                                                            \
            sgsT_wave_kernel<T, V, NSV, true, TD><<<wgrid, 256, wsh, st>>>(nel, bs, lrg, lrgd, elems, ecolor,   \
                                                                         colour, cols, JT, DT, rhs, y);        \
        else                                                                                                    \
            sgsT_wave_kernel<T, V, NSV, false, TD><<<wgrid, 256, wsh, st>>>(nel, bs, lrg, lrgd, elems, ecolor,  \
                                                                          colour, cols, JT, DT, rhs, y);       \
        break;
            switch (nslots) {
                ICE_SGSW(3)
                ICE_SGSW(5)
                ICE_SGSW(7)
            default: return -1;
            }
#undef ICE_SGSW
            return (int)hipGetLastError();
        }
    }
    const int epb = std::max(1, RT_THREADS / bs);
    const int grid = (nel + epb - 1) / epb;
    const size_t sh = (size_t)epb * (nslots + 1) * bs * sizeof(V);
    if (sh > 160 * 1024) return -1;
#define ICE_SGST(NSV)                                                                                               \
    case NSV:                                                                                                       \
        if (forward)                                                                                                \
            sgsT_kernel<T, V, NSV, true, TD><<<grid, RT_THREADS, sh, st>>>(nel, epb, elems, ecolor, colour, bs, cols,  \
                                                                       JT, DT, rhs, y);                             \
        else                                                                                                        \
            sgsT_kernel<T, V, NSV, false, TD><<<grid, RT_THREADS, sh, st>>>(nel, epb, elems, ecolor, colour, bs, cols, \
                                                                        JT, DT, rhs, y);                            \
        break;
    switch (nslots) {
        ICE_SGST(3)
        ICE_SGST(5)
        ICE_SGST(7)
    default: return -1;
    }
#undef ICE_SGST
    return (int)hipGetLastError();
}
template int sgsT_sweep<float, double, float>(bool, int, const int*, const int*, int, int, int, const int*,
                                              const float*, const float*, const double*, double*, hipStream_t);
// FP16 off-diagonal blocks, FP32 factored diagonal blocks (inverse entries can exceed the FP16 range)
template int sgsT_sweep<_Float16, double, float>(bool, int, const int*, const int*, int, int, int, const int*,
                                                 const _Float16*, const float*, const double*, double*, hipStream_t);
template int sgsT_sweep<double, double, double>(bool, int, const int*, const int*, int, int, int, const int*,
                                                const double*, const double*, const double*, double*, hipStream_t);

int convert_transpose_fp32(size_t nblocks, int bs, const double* a, float* b, hipStream_t st) {
    convert_transpose_kernel<float><<<8192, 256, 0, st>>>(nblocks, bs, a, b);
    return (int)hipGetLastError();
}

// FP32 (and optionally FP16) column-major copies of row-major FP64 blocks,
// one workgroup per block: the block is read once, coalesced, into LDS and
// both copies are written coalesced from it.
__global__ __launch_bounds__(256) void convert_transpose_tile_kernel(int bs, const double* __restrict__ a,
                                                                      float* __restrict__ b32,
                                                                      _Float16* __restrict__ b16) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* t = reinterpret_cast<double*>(smem_raw);
    const size_t bb = (size_t)bs * bs;
    const double* src = a + (size_t)blockIdx.x * bb;
    for (int k = threadIdx.x; k < (int)bb; k += blockDim.x) t[k] = src[k];
    __syncthreads();
    float* d32 = b32 + (size_t)blockIdx.x * bb;
    for (int k = threadIdx.x; k < (int)bb; k += blockDim.x) {
        const int c = k / bs, r = k - c * bs;  // destination [c][r]
        const float v = (float)t[r * bs + c];
        d32[k] = v;
        if (b16) b16[(size_t)blockIdx.x * bb + k] = (_Float16)v;
    }
}

// x = A^-1 b with the dense inverse stored column-major (the multigrid's
// coarsest level): a workgroup owns 64 rows, its 16 waves split the columns
// (coalesced column reads, one lane per row), and the 16 partial sums of a
// row are added in a fixed order from LDS
__global__ __launch_bounds__(1024) void dense_inverse_apply_kernel(int n, const double* __restrict__ inv,
                                                                   const double* __restrict__ b,
                                                                   double* __restrict__ x) {
    __shared__ double part[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const int len = (n + 15) / 16;
    const int j0 = w * len, j1 = min(n, j0 + len);
    double s0 = 0.0, s1 = 0.0;
    if (i < n) {
        int j = j0;
        for (; j + 1 < j1; j += 2) {
            s0 += inv[(size_t)j * n + i] * b[j];
            s1 += inv[(size_t)(j + 1) * n + i] * b[j + 1];
        }
        if (j < j1) s0 += inv[(size_t)j * n + i] * b[j];
    }
    part[w][lane] = s0 + s1;
    __syncthreads();
    if (w == 0 && i < n) {
        double sacc = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) sacc += part[k][lane];
        x[i] = sacc;
    }
}

int dense_inverse_apply(int n, const double* inv, const double* b, double* x, hipStream_t st) {
    dense_inverse_apply_kernel<<<(n + 63) / 64, 1024, 0, st>>>(n, inv, b, x);
    return (int)hipGetLastError();
}

int convert_transpose_blocks(size_t nblocks, int bs, const double* a, float* b32, _Float16* b16, hipStream_t st) {
    const size_t sh = (size_t)bs * bs * sizeof(double);
    if (sh > 160 * 1024 || nblocks == 0) return nblocks == 0 ? 0 : -1;
    convert_transpose_tile_kernel<<<(unsigned)nblocks, 256, sh, st>>>(bs, a, b32, b16);
    return (int)hipGetLastError();
}

}  // namespace ice

// ---------------------------------------------------------------------------
// Geometric multigrid on element agglomerates (preconditioner of the Newton
// solve): level 0 holds one constant per element and component (P0), the
// Galerkin projection of the DG Jacobian onto the piecewise constants; each
// further level merges 2 x 2 (x 2) neighbouring cells.  All levels keep the
// block-row layout of the Jacobian (slot 0 the cell itself, slot 1 + j the
// neighbour in the direction of local face j), with m x m blocks.

namespace ice {
namespace {

// A0[e][s][v][w] = sum_{i,j} J[e][s][i m + v][j m + w]  (R J P with P the
// injection of element constants, R = P^T); row-major FP64
__global__ void galerkin_p0_kernel(int nelem, int nslots, int nb, int m, const double* __restrict__ J,
                                   double* __restrict__ A0) {
    const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long total = (long)nelem * nslots * m * m;
    if (tid >= total) return;
    const int w = (int)(tid % m), v = (int)((tid / m) % m);
    const long es = tid / (m * m);
    const int bs = nb * m;
    const double* B = J + es * (long)bs * bs;
    double acc = 0.0;
    for (int i = 0; i < nb; ++i) {
        const double* row = B + (long)(i * m + v) * bs + w;
        for (int j = 0; j < nb; ++j) acc += row[j * m];
    }
    A0[tid] = acc;
}

// Ac[C][s][v][w] = sum over children c of C and slots t of c whose column
// cell maps to slot s of C (the diagonal when it is a sibling)
__global__ void galerkin_agg_kernel(int ncoarse, int nslots, int m, const int* __restrict__ children,
                                    const int* __restrict__ fcols, const int* __restrict__ parent,
                                    const double* __restrict__ Af, double* __restrict__ Ac) {
    const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long total = (long)ncoarse * nslots * m * m;
    if (tid >= total) return;
    const int vw = (int)(tid % (m * m));
    const int s = (int)((tid / (m * m)) % nslots);
    const int C = (int)(tid / ((long)m * m * nslots));
    double acc = 0.0;
    for (int k = 0; k < 8; ++k) {
        const int c = children[C * 8 + k];
        if (c < 0) continue;
        for (int t = 0; t < nslots; ++t) {
            const int col = fcols[(long)c * nslots + t];
            if (col < 0) continue;
            const int target = (parent[col] == C) ? 0 : t;
            if (target == s) acc += Af[(((long)c * nslots + t) * m * m) + vw];
        }
    }
    Ac[tid] = acc;
}

// b0[e][v] = sum_i r[e][i][v]
__global__ void restrict_nodes_kernel(int nelem, int nb, int m, const double* __restrict__ r, double* __restrict__ b0) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nelem * m) return;
    const int e = tid / m, v = tid - e * m;
    const double* re = r + (size_t)e * nb * m + v;
    double acc = 0.0;
    for (int i = 0; i < nb; ++i) acc += re[i * m];
    b0[tid] = acc;
}

// x[e][i][v] += x0[e][v]
__global__ void prolong_nodes_kernel(int nelem, int nb, int m, const double* __restrict__ x0, double* __restrict__ x) {
    const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long total = (long)nelem * nb * m;
    if (tid >= total) return;
    const int v = (int)(tid % m);
    const long e = tid / ((long)nb * m);
    x[tid] += x0[e * m + v];
}

// bc[C][v] = sum over children of r[c][v]
__global__ void restrict_agg_kernel(int ncoarse, int m, const int* __restrict__ children, const double* __restrict__ r,
                                    double* __restrict__ bc) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= ncoarse * m) return;
    const int C = tid / m, v = tid - C * m;
    double acc = 0.0;
    for (int k = 0; k < 8; ++k) {
        const int c = children[C * 8 + k];
        if (c >= 0) acc += r[(size_t)c * m + v];
    }
    bc[tid] = acc;
}

// x[c][v] += xc[parent(c)][v]
__global__ void prolong_agg_kernel(int nfine, int m, const int* __restrict__ parent, const double* __restrict__ xc,
                                   double* __restrict__ x) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nfine * m) return;
    const int c = tid / m, v = tid - c * m;
    x[tid] += xc[(size_t)parent[c] * m + v];
}

inline int grid_of(long n) { return (int)((n + 255) / 256); }

}  // namespace

int mg_galerkin_p0(int nelem, int nslots, int nb, int m, const double* J, double* A0, hipStream_t st) {
    const long total = (long)nelem * nslots * m * m;
    galerkin_p0_kernel<<<grid_of(total), 256, 0, st>>>(nelem, nslots, nb, m, J, A0);
    return (int)hipGetLastError();
}
int mg_galerkin_agg(int ncoarse, int nslots, int m, const int* children, const int* fcols, const int* parent,
                    const double* Af, double* Ac, hipStream_t st) {
    const long total = (long)ncoarse * nslots * m * m;
    galerkin_agg_kernel<<<grid_of(total), 256, 0, st>>>(ncoarse, nslots, m, children, fcols, parent, Af, Ac);
    return (int)hipGetLastError();
}
int mg_restrict_nodes(int nelem, int nb, int m, const double* r, double* b0, hipStream_t st) {
    restrict_nodes_kernel<<<grid_of((long)nelem * m), 256, 0, st>>>(nelem, nb, m, r, b0);
    return (int)hipGetLastError();
}
int mg_prolong_nodes(int nelem, int nb, int m, const double* x0, double* x, hipStream_t st) {
    prolong_nodes_kernel<<<grid_of((long)nelem * nb * m), 256, 0, st>>>(nelem, nb, m, x0, x);
    return (int)hipGetLastError();
}
int mg_restrict_agg(int ncoarse, int m, const int* children, const double* r, double* bc, hipStream_t st) {
    restrict_agg_kernel<<<grid_of((long)ncoarse * m), 256, 0, st>>>(ncoarse, m, children, r, bc);
    return (int)hipGetLastError();
}
int mg_prolong_agg(int nfine, int m, const int* parent, const double* xc, double* x, hipStream_t st) {
    prolong_agg_kernel<<<grid_of((long)nfine * m), 256, 0, st>>>(nfine, m, parent, xc, x);
    return (int)hipGetLastError();
}

}  // namespace ice
