// Linear algebra of the Newton solve on the device: block-sparse products
// with the element-block Jacobian, block-Jacobi factors, and the vector
// kernels of restarted GMRES.  Replaces the PETSc Mat/KSP/PC objects the
// reference's PetscNewton drives (petsc_newton.hpp:154-244).
#include <hip/hip_runtime.h>

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

}  // namespace ice
