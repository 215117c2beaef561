// Warm-start network: a fully connected tanh MLP trained and evaluated on
// the device by the kernels below (no library GEMM).  It maps problem
// parameters to DG coefficients of a steady solution; Newton's FP64 residual
// remains the judge of whatever it predicts.  No reference counterpart.
//
// Layout (all FP32, row-major):
//   activations H_l [batch][w_l]; layer l maps H_l to
//   H_{l+1} = act(H_l W_l + b_l), W_l [w_l][w_{l+1}], b_l [w_{l+1}],
//   act = tanh on hidden layers, identity on the output layer;
//   parameters packed layer by layer as W_0, b_0, W_1, b_1, ...
// Loss: mean squared error over batch x outputs.
//
// GEMMs run on the matrix cores (v_mfma_f32_16x16x4f32, FP32 accumulate):
// a 256-thread workgroup owns a 64 x 64 tile of C, each wave a 32 x 32
// quarter of it as 2 x 2 MFMA blocks; 64 x 16 slices of op(A) and op(B) are
// staged in LDS per K step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/iceicle_amd.h"
#include "engine.h"

namespace ice {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GBM = 64, GBN = 64, GBK = 16;

// C[M][N] = op(A)[M][K] op(B)[K][N] (+ bias[j]) (tanh if act).
// op(A)[i][k] = TA ? A[k lda + i] : A[i lda + k]; op(B)[k][j] = TB ? B[j ldb + k] : B[k ldb + j].
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f32_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                       const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                       int ldc, const float* __restrict__ bias, int act) {
    __shared__ float sA[GBK][GBM + 4];
    __shared__ float sB[GBK][GBN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * GBM, j0 = blockIdx.x * GBN;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += GBK) {
#pragma unroll
        for (int t = tid; t < GBM * GBK; t += 256) {
            int i, k;
            if (TA) { i = t % GBM; k = t / GBM; } else { k = t % GBK; i = t / GBK; }
            const int gi = i0 + i, gk = k0 + k;
            float v = 0.f;
            if (gi < M && gk < K) v = TA ? A[(size_t)gk * lda + gi] : A[(size_t)gi * lda + gk];
            sA[k][i] = v;
        }
#pragma unroll
        for (int t = tid; t < GBK * GBN; t += 256) {
            int k, j;
            if (TB) { k = t % GBK; j = t / GBK; } else { j = t % GBN; k = t / GBN; }
            const int gk = k0 + k, gj = j0 + j;
            float v = 0.f;
            if (gk < K && gj < N) v = TB ? B[(size_t)gj * ldb + gk] : B[(size_t)gk * ldb + gj];
            sB[k][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            // A operand: lane holds A[row = lane % 16][k = lane / 16];
            // B operand: lane holds B[k = lane / 16][col = lane % 16]
            const int kl = kk + (lane >> 4);
            float a[2], b[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                a[q] = sA[kl][wi + q * 16 + (lane & 15)];
                b[q] = sB[kl][wj + q * 16 + (lane & 15)];
            }
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[p], b[q], acc[p][q], 0, 0, 0);
        }
        __syncthreads();
    }
    // result: lane holds C[row = 4 (lane / 16) + r][col = lane % 16], r = 0..3
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + p * 16 + 4 * (lane >> 4) + r;
                const int j = j0 + wj + q * 16 + (lane & 15);
                if (i < M && j < N) {
                    float v = acc[p][q][r];
                    if (bias) v += bias[j];
                    if (act) v = tanhf(v);
                    C[(size_t)i * ldc + j] = v;
                }
            }
}

// dZ = dH (1 - H^2) (backward through tanh, H = tanh(Z))
__global__ void tanh_backward_kernel(size_t n, const float* __restrict__ H, float* __restrict__ d) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] *= 1.f - H[i] * H[i];
}

// db[j] = sum_i dY[i][j], rows summed in order (deterministic)
__global__ void colsum_kernel(int rows, int cols, const float* __restrict__ d, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; ++i) s += d[(size_t)i * cols + j];
    out[j] = s;
}

// dY = 2 (Y - T) / n and per-block partial sums of (Y - T)^2
constexpr int LOSS_BLOCKS = 256;
__global__ __launch_bounds__(256) void mse_kernel(size_t n, const float* __restrict__ Y, const float* __restrict__ T,
                                                  float* __restrict__ dY, double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    const float sc = 2.f / (float)n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float e = Y[i] - T[i];
        if (dY) dY[i] = sc * e;
        s += (double)e * e;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void mse_final_kernel(int nb, size_t n, const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    out[0] = s / (double)n;
}

// Adam (Kingma & Ba 2015) with bias-corrected moments
__global__ void adam_kernel(size_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                            const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                            float* __restrict__ p) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows of X (and T) selected by idx: Xb[i] = X[idx[i]]
__global__ void gather_rows_kernel(int rows, int cols, const int* __restrict__ idx, const float* __restrict__ X,
                                   float* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)rows * cols) return;
    const int i = (int)(t / cols), c = (int)(t % cols);
    out[t] = X[(size_t)idx[i] * cols + c];
}

template <bool TA, bool TB>
void gemm(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc, const float* bias,
          int act, hipStream_t st) {
    dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM);
    gemm_f32_kernel<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, act);
}

inline unsigned nblk(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace ice

struct ice_mlp {
    int L = 0;
    std::vector<int> w;
    std::vector<size_t> woff, boff;
    size_t np = 0;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Ve = nullptr;
    int cap = 0;
    std::vector<float*> H;  // H[1..L] [cap][w[l]]
    float *dA = nullptr, *dB = nullptr, *Xb = nullptr, *Tb = nullptr;
    int* idx = nullptr;
    double *part = nullptr, *loss = nullptr;
    long long t = 0;
    void free_acts() {
        for (float* p : H)
            if (p) (void)hipFree(p);
        H.assign(L + 1, nullptr);
        for (void* p : {(void*)dA, (void*)dB, (void*)Xb, (void*)Tb, (void*)idx})
            if (p) (void)hipFree(p);
        dA = dB = Xb = Tb = nullptr;
        idx = nullptr;
        cap = 0;
    }
    ~ice_mlp() {
        free_acts();
        for (void* p : {(void*)P, (void*)G, (void*)Mo, (void*)Ve, (void*)part, (void*)loss})
            if (p) (void)hipFree(p);
    }
    int maxw() const { return *std::max_element(w.begin(), w.end()); }
    bool ensure(int batch) {
        if (batch <= cap) return true;
        free_acts();
        const size_t mw = (size_t)maxw();
        for (int l = 1; l <= L; ++l)
            if (hipMalloc(&H[l], (size_t)batch * w[l] * sizeof(float)) != hipSuccess) return false;
        if (hipMalloc(&dA, (size_t)batch * mw * sizeof(float)) != hipSuccess) return false;
        if (hipMalloc(&dB, (size_t)batch * mw * sizeof(float)) != hipSuccess) return false;
        if (hipMalloc(&Xb, (size_t)batch * w[0] * sizeof(float)) != hipSuccess) return false;
        if (hipMalloc(&Tb, (size_t)batch * w[L] * sizeof(float)) != hipSuccess) return false;
        if (hipMalloc(&idx, (size_t)batch * sizeof(int)) != hipSuccess) return false;
        cap = batch;
        return true;
    }
    // forward into H[1..L] (the last layer into `y` if given)
    void forward(const float* x, int batch, float* y, hipStream_t st) {
        const float* in = x;
        for (int l = 0; l < L; ++l) {
            float* out = (l == L - 1 && y) ? y : H[l + 1];
            ice::gemm<false, false>(batch, w[l + 1], w[l], in, w[l], P + woff[l], w[l + 1], out, w[l + 1], P + boff[l],
                                    l < L - 1 ? 1 : 0, st);
            in = out;
        }
    }
    // gradient of the mean squared error into G; loss into `loss` (device)
    void gradient(const float* x, const float* tgt, int batch, hipStream_t st) {
        forward(x, batch, nullptr, st);
        const size_t n = (size_t)batch * w[L];
        ice::mse_kernel<<<ice::LOSS_BLOCKS, 256, 0, st>>>(n, H[L], tgt, dA, part);
        ice::mse_final_kernel<<<1, 64, 0, st>>>(ice::LOSS_BLOCKS, n, part, loss);
        float* dcur = dA;
        float* dnext = dB;
        for (int l = L - 1; l >= 0; --l) {
            const float* in = (l == 0) ? x : H[l];
            ice::gemm<true, false>(w[l], w[l + 1], batch, in, w[l], dcur, w[l + 1], G + woff[l], w[l + 1], nullptr, 0,
                                   st);
            ice::colsum_kernel<<<ice::nblk((size_t)w[l + 1]), 256, 0, st>>>(batch, w[l + 1], dcur, G + boff[l]);
            if (l > 0) {
                ice::gemm<false, true>(batch, w[l], w[l + 1], dcur, w[l + 1], P + woff[l], w[l + 1], dnext, w[l],
                                       nullptr, 0, st);
                const size_t nh = (size_t)batch * w[l];
                ice::tanh_backward_kernel<<<ice::nblk(nh), 256, 0, st>>>(nh, H[l], dnext);
                std::swap(dcur, dnext);
            }
        }
    }
};

namespace {
int mlp_fail(int code, const std::string& m) {
    ice::set_error(m);
    return code;
}
int hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return ICE_OK;
    return mlp_fail(ICE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

extern "C" {

int ice_mlp_create(int nlayers, const int* widths, unsigned long long seed, ice_mlp** out) {
    if (!widths || !out || nlayers < 1 || nlayers > 16) return mlp_fail(ICE_ERR_ARG, "bad layer description");
    *out = nullptr;
    for (int l = 0; l <= nlayers; ++l)
        if (widths[l] < 1 || widths[l] > (1 << 20)) return mlp_fail(ICE_ERR_ARG, "layer width out of range");
    auto m = new ice_mlp();
    m->L = nlayers;
    m->w.assign(widths, widths + nlayers + 1);
    m->H.assign(nlayers + 1, nullptr);
    size_t off = 0;
    for (int l = 0; l < nlayers; ++l) {
        m->woff.push_back(off);
        off += (size_t)widths[l] * widths[l + 1];
        m->boff.push_back(off);
        off += widths[l + 1];
    }
    m->np = off;
    // Glorot-uniform weights, zero biases
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < nlayers; ++l) {
        const double a = std::sqrt(6.0 / (widths[l] + widths[l + 1]));
        std::uniform_real_distribution<double> U(-a, a);
        for (size_t k = 0; k < (size_t)widths[l] * widths[l + 1]; ++k) h[m->woff[l] + k] = (float)U(rng);
    }
    int rc = ICE_OK;
    for (float** p : {&m->P, &m->G, &m->Mo, &m->Ve})
        if (rc == ICE_OK) rc = hip_check(hipMalloc(p, off * sizeof(float)), "hipMalloc");
    if (rc == ICE_OK) rc = hip_check(hipMalloc(&m->part, ice::LOSS_BLOCKS * sizeof(double)), "hipMalloc");
    if (rc == ICE_OK) rc = hip_check(hipMalloc(&m->loss, sizeof(double)), "hipMalloc");
    if (rc == ICE_OK) rc = hip_check(hipMemcpy(m->P, h.data(), off * sizeof(float), hipMemcpyHostToDevice), "upload");
    if (rc == ICE_OK) rc = hip_check(hipMemset(m->Mo, 0, off * sizeof(float)), "hipMemset");
    if (rc == ICE_OK) rc = hip_check(hipMemset(m->Ve, 0, off * sizeof(float)), "hipMemset");
    if (rc != ICE_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return ICE_OK;
}

int ice_mlp_destroy(ice_mlp* m) {
    delete m;
    return ICE_OK;
}

int ice_mlp_num_params(ice_mlp* m, long long* n) {
    if (!m || !n) return mlp_fail(ICE_ERR_ARG, "null argument");
    *n = (long long)m->np;
    return ICE_OK;
}

int ice_mlp_get_params(ice_mlp* m, float* host) {
    if (!m || !host) return mlp_fail(ICE_ERR_ARG, "null argument");
    return hip_check(hipMemcpy(host, m->P, m->np * sizeof(float), hipMemcpyDeviceToHost), "download");
}

int ice_mlp_set_params(ice_mlp* m, const float* host) {
    if (!m || !host) return mlp_fail(ICE_ERR_ARG, "null argument");
    m->t = 0;
    int rc = hip_check(hipMemcpy(m->P, host, m->np * sizeof(float), hipMemcpyHostToDevice), "upload");
    if (rc == ICE_OK) rc = hip_check(hipMemset(m->Mo, 0, m->np * sizeof(float)), "hipMemset");
    if (rc == ICE_OK) rc = hip_check(hipMemset(m->Ve, 0, m->np * sizeof(float)), "hipMemset");
    return rc;
}

int ice_mlp_forward(ice_mlp* m, const float* x, int batch, float* y, void* stream) {
    if (!m || !x || !y || batch < 1) return mlp_fail(ICE_ERR_ARG, "bad argument");
    if (!m->ensure(batch)) return mlp_fail(ICE_ERR_ALLOC, "activation buffers");
    m->forward(x, batch, y, (hipStream_t)stream);
    return hip_check(hipGetLastError(), "mlp forward launch");
}

int ice_mlp_gradient(ice_mlp* m, const float* x, const float* tgt, int batch, double* loss, float* grad,
                     void* stream) {
    if (!m || !x || !tgt || batch < 1) return mlp_fail(ICE_ERR_ARG, "bad argument");
    if (!m->ensure(batch)) return mlp_fail(ICE_ERR_ALLOC, "activation buffers");
    hipStream_t st = (hipStream_t)stream;
    m->gradient(x, tgt, batch, st);
    int rc = hip_check(hipGetLastError(), "mlp gradient launch");
    if (rc == ICE_OK && grad)
        rc = hip_check(hipMemcpyAsync(grad, m->G, m->np * sizeof(float), hipMemcpyDeviceToDevice, st), "copy");
    if (rc == ICE_OK && loss) {
        rc = hip_check(hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, st), "copy");
        if (rc == ICE_OK) rc = hip_check(hipStreamSynchronize(st), "sync");
    }
    return rc;
}

int ice_mlp_train(ice_mlp* m, const float* X, const float* T, int n, const ice_mlp_train_opts* o, double* loss_hist,
                  void* stream) {
    if (!m || !X || !T || !o || n < 1 || o->batch < 1 || o->epochs < 0) return mlp_fail(ICE_ERR_ARG, "bad argument");
    const int batch = std::min(o->batch, n);
    if (!m->ensure(batch)) return mlp_fail(ICE_ERR_ALLOC, "activation buffers");
    hipStream_t st = (hipStream_t)stream;
    std::mt19937_64 rng(o->seed);
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    const int win = m->w[0], wout = m->w[m->L];
    for (int ep = 0; ep < o->epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        double ep_loss = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            int rc = hip_check(hipMemcpyAsync(m->idx, perm.data() + b0, batch * sizeof(int), hipMemcpyHostToDevice, st),
                               "index upload");
            if (rc != ICE_OK) return rc;
            ice::gather_rows_kernel<<<ice::nblk((size_t)batch * win), 256, 0, st>>>(batch, win, m->idx, X, m->Xb);
            ice::gather_rows_kernel<<<ice::nblk((size_t)batch * wout), 256, 0, st>>>(batch, wout, m->idx, T, m->Tb);
            m->gradient(m->Xb, m->Tb, batch, st);
            m->t += 1;
            const float c1 = 1.f - std::pow((float)o->beta1, (float)m->t);
            const float c2 = 1.f - std::pow((float)o->beta2, (float)m->t);
            const float lr = (float)(o->lr * std::pow(o->lr_decay, (double)ep));
            ice::adam_kernel<<<ice::nblk(m->np), 256, 0, st>>>(m->np, lr, (float)o->beta1, (float)o->beta2,
                                                                (float)o->eps, c1, c2, m->G, m->Mo, m->Ve, m->P);
            double l = 0.0;
            rc = hip_check(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, st), "copy");
            if (rc == ICE_OK) rc = hip_check(hipStreamSynchronize(st), "sync");
            if (rc != ICE_OK) return rc;
            ep_loss += l;
            ++nb;
        }
        if (loss_hist) loss_hist[ep] = nb ? ep_loss / nb : 0.0;
    }
    return hip_check(hipGetLastError(), "mlp train launch");
}

}  // extern "C"
