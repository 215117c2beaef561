// Newton-linesearch steady solve on the device (PetscNewton::solve,
// petsc_newton.hpp:205-329) with the device block Jacobian and a device
// GMRES in place of PETSc KSP.  The control flow (steps, linesearch,
// convergence test) runs on the host; every vector operation and reduction
// runs on the caller's stream, and the host reads back only scalars.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/iceicle_amd.h"
#include "engine.h"
#include "kernels.h"
#include "linesearch.h"

namespace ice {
// krylov.hip
template <class T>
int block_inverse(int nelem, int nslots, int bs, const double* J, T* Dinv, int* nsingular, bool transposed,
                  hipStream_t st);
template <class T, class V>
int bsrT_apply(int nelem, int nslots, int bs, const int* cols, const T* JT, const V* x, const V* b, V* y,
               hipStream_t st);
template <class T, class V>
int diagT_apply(int nelem, int bs, const T* DT, const V* x, V* y, hipStream_t st);
template <class T, class V, class TD>
int sgsT_sweep(bool forward, int nel, const int* elems, const int* ecolor, int colour, int nslots, int bs,
               const int* cols, const T* JT, const TD* DT, const V* rhs, V* y, hipStream_t st);
int convert_transpose_fp32(size_t nblocks, int bs, const double* a, float* b, hipStream_t st);
int convert_transpose_blocks(size_t nblocks, int bs, const double* a, float* b32, _Float16* b16, hipStream_t st);
int dense_inverse_apply(int n, const double* inv, const double* b, double* x, hipStream_t st);
int line_factor(int nelem, int nslots, int bs, const double* J, double* Sinv, double* W, int* nsingular,
                hipStream_t st);
int line_solve(int nelem, int nslots, int bs, const double* J, const double* Sinv, const double* W, const double* b,
               double* x, hipStream_t st);
int mg_galerkin_p0(int nelem, int nslots, int nb, int m, const double* J, double* A0, hipStream_t st);
int mg_galerkin_agg(int ncoarse, int nslots, int m, const int* children, const int* fcols, const int* parent,
                    const double* Af, double* Ac, hipStream_t st);
int mg_restrict_nodes(int nelem, int nb, int m, const double* r, double* b0, hipStream_t st);
int mg_prolong_nodes(int nelem, int nb, int m, const double* x0, double* x, hipStream_t st);
int mg_restrict_agg(int ncoarse, int m, const int* children, const double* r, double* bc, hipStream_t st);
int mg_prolong_agg(int nfine, int m, const int* parent, const double* xc, double* x, hipStream_t st);
template <class T, class V>
int block_diag_apply(int nelem, int bs, const T* D, const V* x, V* y, hipStream_t st);
int convert_to_fp32(size_t n, const double* a, float* b, hipStream_t st);
int gmres_partials(size_t n);
int launch_mdot(size_t n, int nv, const double* V, size_t ldv, const double* w, double* part, int npart, double* out,
                hipStream_t st);
int launch_maxpy_norm(size_t n, int nv, const double* V, size_t ldv, const double* h, double* w, double* part,
                      double* out, hipStream_t st);
int launch_norm2(size_t n, const double* w, double* part, double* out, hipStream_t st);
int launch_combine(size_t n, int nv, const double* V, size_t ldv, const double* y, double* x, hipStream_t st);
int launch_scale(size_t n, const double* s, int inv, double* x, hipStream_t st);
int launch_axpby2(size_t n, double a, const double* x, double b, double* y, hipStream_t st);
int launch_arnoldi(int k, int m, const double* h1, const double* h2, int refine, double* sc, hipStream_t st);
int launch_hessenberg_solve(int k, int m, const double* sc, double* y, hipStream_t st);
int launch_gmres_init(int m, const double* nrm2, double* sc, hipStream_t st);
template <class T, class V>
int sgs_sweep(bool forward, int nrows, const int* elems, const int* ecolor, int c, int nslots, int bs,
              const int* cols, const T* J, const T* Dinv, const V* r, V* y, hipStream_t st);
}  // namespace ice

namespace {

struct Fail {
    int code;
    std::string msg;
};

#define NT_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw Fail{ICE_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)}; \
    } while (0)
#define NT_LAUNCH(call)                                                        \
    do {                                                                       \
        if ((call) != 0) throw Fail{ICE_ERR_HIP, std::string("launch: ") + #call}; \
    } while (0)

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace


// One level of the agglomeration multigrid: n cells of an nx x ny x nz grid
// (x fastest), m x m blocks in the Jacobian's slot layout.
struct MGLevel {
    int n = 0;
    int dims[3] = {1, 1, 1};
    std::vector<int> hcols;        // [n][nslots]
    const int* cols = nullptr;     // device copy of hcols
    int* parent = nullptr;         // [n] cell of the next coarser level (null on the coarsest)
    int* children = nullptr;       // [n][8] cells of the next finer level (level 0: none)
    double* A64 = nullptr;         // [n][nslots][m][m] row-major
    float* A32 = nullptr;          // column-major FP32 copy
    float* Dinv = nullptr;         // column-major factored diagonal blocks
    int* ecolor = nullptr;
    int* celems = nullptr;
    std::vector<int> coff;
    double *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;
};

struct ice_newton {
    ice_disc* disc = nullptr;
    ice_newton_opts opt{};
    int nelem = 0, nslots = 0, bs = 0;
    size_t n = 0, nblkvals = 0;
    int m = 30, npart = 0, nred = 0;
    // device memory
    double* J64 = nullptr;
    float* J32 = nullptr;       // FP32 copy, each block column-major ([c][r])
    void* Dinv = nullptr;       // [nelem][bs][bs] factored diagonal blocks: FP32 column-major, or FP64 row-major
    double* V = nullptr;        // (m+1) x n Krylov basis
    double* z = nullptr;        // n
    double* t = nullptr;        // n
    double* res = nullptr;      // n, R(u)
    double* du = nullptr;       // n
    double* ustep = nullptr;    // n
    double* rwork = nullptr;    // n
    double* xr = nullptr;       // n, refinement residual
    double* part = nullptr;     // nred x (m+2)
    double* h1 = nullptr;       // m+2
    double* h2 = nullptr;       // m+2
    double* sc = nullptr;       // Hessenberg, rotations, g, norms
    double* y = nullptr;        // m
    int* nsing = nullptr;
    _Float16* J16 = nullptr;      // FP16 copy of the blocks for the smoother (column-major per block)
    double* line_Sinv = nullptr;  // block-tridiagonal factors (line preconditioner)
    double* line_W = nullptr;
    std::vector<MGLevel> mg;    // multigrid levels (level 0 = element constants)
    double *mg_r = nullptr, *mg_t = nullptr;  // fine-level work vectors of the V-cycle
    double* coarse_inv = nullptr;  // dense inverse of the coarsest level (column-major), when it is small
    int coarse_n = 0;
    bool coarse_direct = false;    // this factorisation's coarsest level is solved exactly
    int neq = 0, nbasis = 0;
    int* ecolor = nullptr;      // element colour (multicolour Gauss-Seidel)
    int* celems = nullptr;      // elements grouped by colour
    std::vector<int> coff;      // colour offsets into celems
    double* host = nullptr;     // pinned scalars
    bool factored = false;
    std::vector<void*> owned;

    ~ice_newton() {
        for (void* p : owned)
            if (p) (void)hipFree(p);
        if (host) (void)hipHostFree(host);
    }
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        NT_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        owned.push_back(p);
        return static_cast<T*>(p);
    }

    bool fp32() const { return opt.jacobian_precision == ICE_PREC_FP32; }
    double* Vj(int j) const { return V + (size_t)j * n; }

    double read_scalar(const double* d, hipStream_t st) {
        NT_HIP(hipMemcpyAsync(host, d, sizeof(double), hipMemcpyDeviceToHost, st));
        NT_HIP(hipStreamSynchronize(st));
        return host[0];
    }
    double norm(const double* v, hipStream_t st) {
        NT_LAUNCH(ice::launch_norm2(n, v, part, h1, st));
        return std::sqrt(std::max(read_scalar(h1, st), 0.0));
    }

    // Krylov operator w = A z and preconditioner z = M^-1 v
    void apply_A(const double* x, double* w, hipStream_t st) {
        const int* cols = ice::disc_jcols(disc);
        if (fp32()) NT_LAUNCH((ice::bsrT_apply<float, double>(nelem, nslots, bs, cols, J32, x, nullptr, w, st)));
        else NT_LAUNCH((ice::bsr_apply<double, double>(nelem, nslots, bs, cols, J64, x, nullptr, w, st)));
    }
    void apply_sgs_fine(const double* v, double* out, hipStream_t st) {
        const int nc = (int)coff.size() - 1;
        const int* cols = ice::disc_jcols(disc);
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < nc; ++k) {
                const bool fwd = pass == 0;
                const int c = fwd ? k : nc - 1 - k;
                if (!fwd && c == nc - 1) continue;  // last colour: backward sweep leaves it unchanged
                const int* el = celems + coff[c];
                const int nr = coff[c + 1] - coff[c];
                if (J16)
                    NT_LAUNCH((ice::sgsT_sweep<_Float16, double, float>(fwd, nr, el, ecolor, c, nslots, bs, cols,
                                                                        J16, (const float*)Dinv, v, out, st)));
                else if (fp32())
                    NT_LAUNCH((ice::sgsT_sweep<float, double, float>(fwd, nr, el, ecolor, c, nslots, bs, cols, J32,
                                                              (const float*)Dinv, v, out, st)));
                else
                    NT_LAUNCH((ice::sgs_sweep<double, double>(fwd, nr, el, ecolor, c, nslots, bs, cols, J64,
                                                              (const double*)Dinv, v, out, st)));
            }
    }

    void apply_M(const double* v, double* out, hipStream_t st) {
        if (opt.preconditioner == ICE_PC_LINE) {
            NT_LAUNCH(ice::line_solve(nelem, nslots, bs, J64, line_Sinv, line_W, v, out, st));
            return;
        }
        if (opt.preconditioner == ICE_PC_MULTIGRID) {
            apply_mg(v, out, st);
            return;
        }
        if (opt.preconditioner == ICE_PC_BLOCK_SGS) {
            apply_sgs_fine(v, out, st);
            return;
        }
        if (fp32()) NT_LAUNCH((ice::diagT_apply<float, double>(nelem, bs, (const float*)Dinv, v, out, st)));
        else NT_LAUNCH((ice::block_diag_apply<double, double>(nelem, bs, (const double*)Dinv, v, out, st)));
    }

    // greedy colouring of a block graph (face neighbours differ); returns the
    // colour offsets, fills the device colour and grouped-cell arrays
    std::vector<int> colour_graph(int ncell, const std::vector<int>& cols, int** dcolour, int** dcells) {
        std::vector<int> col(ncell, -1);
        int nc = 0;
        for (int e = 0; e < ncell; ++e) {
            unsigned used = 0;
            for (int s2 = 1; s2 < nslots; ++s2) {
                const int nb = cols[(size_t)e * nslots + s2];
                if (nb >= 0 && nb != e && col[nb] >= 0) used |= 1u << col[nb];
            }
            int c = 0;
            while (used & (1u << c)) ++c;
            col[e] = c;
            nc = std::max(nc, c + 1);
        }
        std::vector<int> off(nc + 1, 0);
        for (int e = 0; e < ncell; ++e) ++off[col[e] + 1];
        for (int c = 0; c < nc; ++c) off[c + 1] += off[c];
        std::vector<int> order(ncell), fill(off.begin(), off.end() - 1);
        for (int e = 0; e < ncell; ++e) order[fill[col[e]]++] = e;
        *dcolour = alloc<int>(ncell);
        *dcells = alloc<int>(ncell);
        NT_HIP(hipMemcpy(*dcolour, col.data(), ncell * sizeof(int), hipMemcpyHostToDevice));
        NT_HIP(hipMemcpy(*dcells, order.data(), ncell * sizeof(int), hipMemcpyHostToDevice));
        return off;
    }

    void colour() {
        std::vector<int> cols((size_t)nelem * nslots);
        int ns = 0, b = 0;
        if (ice_jacobian_layout(disc, &ns, &b, cols.data()) != ICE_OK) throw Fail{ICE_ERR_STATE, "no Jacobian layout"};
        coff = colour_graph(nelem, cols, &ecolor, &celems);
    }

    // out = M^-1 rhs for one multicolour block SGS sweep with m x m blocks on level L
    void sgs_level(const MGLevel& L, const double* rhs, double* out, hipStream_t st) {
        const int nc = (int)L.coff.size() - 1;
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < nc; ++k) {
                const bool fwd = pass == 0;
                const int c = fwd ? k : nc - 1 - k;
                if (!fwd && c == nc - 1) continue;
                NT_LAUNCH((ice::sgsT_sweep<float, double, float>(fwd, L.coff[c + 1] - L.coff[c], L.celems + L.coff[c],
                                                          L.ecolor, c, nslots, neq, L.cols, L.A32, L.Dinv, rhs, out,
                                                          st)));
            }
    }

    // x = V-cycle approximation of A_l^-1 b (x need not be initialised)
    void vcycle(int l, const double* b, double* x, hipStream_t st) {
        MGLevel& L = mg[l];
        const size_t nl = (size_t)L.n * neq;
        if (l + 1 == (int)mg.size() && coarse_direct) {
            NT_LAUNCH(ice::dense_inverse_apply(coarse_n, coarse_inv, b, x, st));
            return;
        }
        sgs_level(L, b, x, st);
        if (l + 1 == (int)mg.size()) {
            for (int it = 0; it < 8; ++it) {
                NT_LAUNCH((ice::bsrT_apply<float, double>(L.n, nslots, neq, L.cols, L.A32, x, b, L.r, st)));
                sgs_level(L, L.r, L.t, st);
                NT_LAUNCH(ice::launch_axpby2(nl, 1.0, L.t, 1.0, x, st));
            }
            return;
        }
        MGLevel& Cn = mg[l + 1];
        NT_LAUNCH((ice::bsrT_apply<float, double>(L.n, nslots, neq, L.cols, L.A32, x, b, L.r, st)));
        NT_LAUNCH(ice::mg_restrict_agg(Cn.n, neq, Cn.children, L.r, Cn.b, st));
        vcycle(l + 1, Cn.b, Cn.x, st);
        NT_LAUNCH(ice::mg_prolong_agg(L.n, neq, L.parent, Cn.x, x, st));
        NT_LAUNCH((ice::bsrT_apply<float, double>(L.n, nslots, neq, L.cols, L.A32, x, b, L.r, st)));
        sgs_level(L, L.r, L.t, st);
        NT_LAUNCH(ice::launch_axpby2(nl, 1.0, L.t, 1.0, x, st));
    }

    // hierarchy structure: level 0 = the elements, then 2^ndim agglomerates
    // down to at most 64 cells
    void build_hierarchy() {
        int nel[3], ne = 0, nbs = 0;
        ice::disc_grid(disc, nel, &ne, &nbs);
        if (nel[0] <= 0) throw Fail{ICE_ERR_UNSUPPORTED, "the multigrid preconditioner needs a single structured mesh"};
        neq = ne;
        nbasis = nbs;
        const int D = (nslots - 1) / 2;
        MGLevel L0;
        L0.n = nelem;
        for (int k = 0; k < 3; ++k) L0.dims[k] = nel[k];
        L0.hcols.resize((size_t)nelem * nslots);
        int ns = 0, b = 0;
        if (ice_jacobian_layout(disc, &ns, &b, L0.hcols.data()) != ICE_OK) throw Fail{ICE_ERR_STATE, "no layout"};
        L0.cols = ice::disc_jcols(disc);
        mg.push_back(std::move(L0));
        while (mg.back().n > 64 && mg.size() < 16) {
            MGLevel& F = mg.back();
            MGLevel C;
            for (int k = 0; k < 3; ++k) C.dims[k] = (k < D) ? (F.dims[k] + 1) / 2 : 1;
            C.n = C.dims[0] * C.dims[1] * C.dims[2];
            std::vector<int> par(F.n), ch((size_t)C.n * 8, -1), cnt(C.n, 0);
            for (int c = 0; c < F.n; ++c) {
                const int i = c % F.dims[0], j = (c / F.dims[0]) % F.dims[1], k = c / (F.dims[0] * F.dims[1]);
                const int P = i / 2 + C.dims[0] * (j / 2 + C.dims[1] * (k / 2));
                par[c] = P;
                ch[(size_t)P * 8 + cnt[P]++] = c;
            }
            C.hcols.assign((size_t)C.n * nslots, -1);
            for (int P = 0; P < C.n; ++P) C.hcols[(size_t)P * nslots] = P;
            for (int c = 0; c < F.n; ++c)
                for (int t = 1; t < nslots; ++t) {
                    const int col = F.hcols[(size_t)c * nslots + t];
                    if (col < 0 || par[col] == par[c]) continue;
                    C.hcols[(size_t)par[c] * nslots + t] = par[col];
                }
            F.parent = alloc<int>(F.n);
            NT_HIP(hipMemcpy(F.parent, par.data(), F.n * sizeof(int), hipMemcpyHostToDevice));
            C.children = alloc<int>((size_t)C.n * 8);
            NT_HIP(hipMemcpy(C.children, ch.data(), ch.size() * sizeof(int), hipMemcpyHostToDevice));
            int* dc = alloc<int>(C.hcols.size());
            NT_HIP(hipMemcpy(dc, C.hcols.data(), C.hcols.size() * sizeof(int), hipMemcpyHostToDevice));
            C.cols = dc;
            mg.push_back(std::move(C));
        }
        for (MGLevel& L : mg) {
            const size_t nb2 = (size_t)L.n * nslots * neq * neq;
            L.A64 = alloc<double>(nb2);
            L.A32 = alloc<float>(nb2);
            L.Dinv = alloc<float>((size_t)L.n * neq * neq);
            for (double** v : {&L.x, &L.b, &L.r, &L.t}) *v = alloc<double>((size_t)L.n * neq);
            L.coff = colour_graph(L.n, L.hcols, &L.ecolor, &L.celems);
        }
        mg_r = alloc<double>(n);
        mg_t = alloc<double>(n);
        coarse_n = mg.back().n * neq;
        if (coarse_n <= 1024) coarse_inv = alloc<double>((size_t)coarse_n * coarse_n);
    }

    // dense inverse of the coarsest level by Gauss-Jordan with partial
    // pivoting on the host (at most 1024 unknowns), uploaded column-major;
    // false if the level is singular (the V-cycle then keeps its SGS sweeps)
    bool factor_coarse(hipStream_t st) {
        if (!coarse_inv) return false;
        const MGLevel& L = mg.back();
        const int cn = coarse_n, mm = neq * neq;
        std::vector<double> blk((size_t)L.n * nslots * mm);
        NT_HIP(hipMemcpyAsync(blk.data(), L.A64, blk.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        NT_HIP(hipStreamSynchronize(st));
        std::vector<double> a((size_t)cn * cn, 0.0), inv((size_t)cn * cn, 0.0);
        for (int P = 0; P < L.n; ++P)
            for (int t = 0; t < nslots; ++t) {
                const int col = L.hcols[(size_t)P * nslots + t];
                if (col < 0) continue;
                const double* B = blk.data() + ((size_t)P * nslots + t) * mm;
                for (int r = 0; r < neq; ++r)
                    for (int c = 0; c < neq; ++c) a[(size_t)(P * neq + r) * cn + col * neq + c] += B[r * neq + c];
            }
        for (int i = 0; i < cn; ++i) inv[(size_t)i * cn + i] = 1.0;
        for (int k = 0; k < cn; ++k) {
            int p = k;
            for (int i = k + 1; i < cn; ++i)
                if (std::abs(a[(size_t)i * cn + k]) > std::abs(a[(size_t)p * cn + k])) p = i;
            if (!(std::abs(a[(size_t)p * cn + k]) > 0.0)) return false;
            if (p != k)
                for (int j = 0; j < cn; ++j) {
                    std::swap(a[(size_t)k * cn + j], a[(size_t)p * cn + j]);
                    std::swap(inv[(size_t)k * cn + j], inv[(size_t)p * cn + j]);
                }
            const double d = 1.0 / a[(size_t)k * cn + k];
            for (int j = 0; j < cn; ++j) {
                a[(size_t)k * cn + j] *= d;
                inv[(size_t)k * cn + j] *= d;
            }
            for (int i = 0; i < cn; ++i) {
                if (i == k) continue;
                const double f = a[(size_t)i * cn + k];
                if (f == 0.0) continue;
                for (int j = 0; j < cn; ++j) {
                    a[(size_t)i * cn + j] -= f * a[(size_t)k * cn + j];
                    inv[(size_t)i * cn + j] -= f * inv[(size_t)k * cn + j];
                }
            }
        }
        std::vector<double> cm((size_t)cn * cn);
        for (int i = 0; i < cn; ++i)
            for (int j = 0; j < cn; ++j) cm[(size_t)j * cn + i] = inv[(size_t)i * cn + j];
        NT_HIP(hipMemcpyAsync(coarse_inv, cm.data(), cm.size() * sizeof(double), hipMemcpyHostToDevice, st));
        NT_HIP(hipStreamSynchronize(st));
        return true;
    }

    // numeric setup of the hierarchy from the FP64 Jacobian
    void factor_mg(hipStream_t st) {
        for (size_t l = 0; l < mg.size(); ++l) {
            MGLevel& L = mg[l];
            if (l == 0)
                NT_LAUNCH(ice::mg_galerkin_p0(nelem, nslots, nbasis, neq, J64, L.A64, st));
            else
                NT_LAUNCH(ice::mg_galerkin_agg(L.n, nslots, neq, L.children, mg[l - 1].cols, mg[l - 1].parent,
                                               mg[l - 1].A64, L.A64, st));
            NT_LAUNCH(ice::convert_transpose_fp32((size_t)L.n * nslots, neq, L.A64, L.A32, st));
            NT_LAUNCH(ice::block_inverse<float>(L.n, nslots, neq, L.A64, L.Dinv, nsing, true, st));
        }
        coarse_direct = factor_coarse(st);
    }

    // fine-level V-cycle: SGS, coarse correction through the element
    // constants, SGS
    void apply_mg(const double* v, double* out, hipStream_t st) {
        const int* cols = ice::disc_jcols(disc);
        // the fine residuals inside the preconditioner read the smoother's blocks
        auto fine_residual = [&](const double* x) {
            if (J16) NT_LAUNCH((ice::bsrT_apply<_Float16, double>(nelem, nslots, bs, cols, J16, x, v, mg_r, st)));
            else NT_LAUNCH((ice::bsrT_apply<float, double>(nelem, nslots, bs, cols, J32, x, v, mg_r, st)));
        };
        apply_sgs_fine(v, out, st);
        fine_residual(out);
        NT_LAUNCH(ice::mg_restrict_nodes(nelem, nbasis, neq, mg_r, mg[0].b, st));
        vcycle(0, mg[0].b, mg[0].x, st);
        NT_LAUNCH(ice::mg_prolong_nodes(nelem, nbasis, neq, mg[0].x, out, st));
        fine_residual(out);
        apply_sgs_fine(mg_r, mg_t, st);
        NT_LAUNCH(ice::launch_axpby2(n, 1.0, mg_t, 1.0, out, st));
    }

    void factor(hipStream_t st) {
        NT_HIP(hipMemsetAsync(nsing, 0, sizeof(int), st));
        if (opt.preconditioner == ICE_PC_LINE) {
            if (fp32()) NT_LAUNCH(ice::convert_transpose_fp32((size_t)nelem * nslots, bs, J64, J32, st));
            NT_LAUNCH(ice::line_factor(nelem, nslots, bs, J64, line_Sinv, line_W, nsing, st));
            factored = true;
            return;
        }
        if (fp32()) {
            NT_LAUNCH(ice::convert_transpose_blocks((size_t)nelem * nslots, bs, J64, J32, J16, st));
            NT_LAUNCH(ice::block_inverse<float>(nelem, nslots, bs, J64, (float*)Dinv, nsing, true, st));
        } else {
            NT_LAUNCH(ice::block_inverse<double>(nelem, nslots, bs, J64, (double*)Dinv, nsing, false, st));
        }
        if (opt.preconditioner == ICE_PC_MULTIGRID) factor_mg(st);
        factored = true;
    }

    // Restarted GMRES (right preconditioning, classical Gram-Schmidt: one
    // pass as PETSc's default, or with one reorthogonalisation pass when
    // gmres_cgs_refine = 1): x += correction solving A x = b to
    // ||b - A x|| <= rtol ||b|| or maxit iterations.  x must hold 0 on entry.
    int gmres(const double* b, double* x, double rtol, int maxit, hipStream_t st, double* est) {
        int total = 0;
        NT_HIP(hipMemcpyAsync(Vj(0), b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        NT_LAUNCH(ice::launch_norm2(n, Vj(0), part, h1, st));
        NT_LAUNCH(ice::launch_gmres_init(m, h1, sc, st));
        const double* beta_d = sc + (size_t)(m + 1) * m + 2 * m;  // g[0]
        const double beta0 = read_scalar(beta_d, st);
        *est = 0.0;
        if (!(beta0 > 0.0)) return 0;
        const double target = rtol * beta0;
        double* hn = sc + (size_t)(m + 1) * m + 2 * m + m + 1;
        bool done = false;
        while (!done) {
            NT_LAUNCH(ice::launch_scale(n, beta_d, 1, Vj(0), st));
            int k = 0;
            double r = 0.0;
            for (; k < m;) {
                apply_M(Vj(k), z, st);
                apply_A(z, Vj(k + 1), st);
                double* w = Vj(k + 1);
                NT_LAUNCH(ice::launch_mdot(n, k + 1, V, n, w, part, npart, h1, st));
                if (opt.gmres_cgs_refine) {
                    NT_LAUNCH(ice::launch_maxpy_norm(n, k + 1, V, n, h1, w, part, nullptr, st));
                    NT_LAUNCH(ice::launch_mdot(n, k + 1, V, n, w, part, npart, h2, st));
                    NT_LAUNCH(ice::launch_maxpy_norm(n, k + 1, V, n, h2, w, part, h2 + k + 1, st));
                } else {
                    NT_LAUNCH(ice::launch_maxpy_norm(n, k + 1, V, n, h1, w, part, h2 + k + 1, st));
                }
                NT_LAUNCH(ice::launch_arnoldi(k, m, h1, h2, opt.gmres_cgs_refine, sc, st));
                NT_LAUNCH(ice::launch_scale(n, hn, 1, w, st));
                ++k;
                ++total;
                r = read_scalar(hn + 1, st);
                if (!(r > target) || total >= maxit || !std::isfinite(r)) break;
            }
            // x += M^-1 V_k y_k
            NT_LAUNCH(ice::launch_hessenberg_solve(k, m, sc, y, st));
            NT_LAUNCH(ice::launch_axpby2(n, 0.0, t, 0.0, t, st));
            NT_LAUNCH(ice::launch_combine(n, k, V, n, y, t, st));
            apply_M(t, z, st);
            NT_LAUNCH(ice::launch_axpby2(n, 1.0, z, 1.0, x, st));
            *est = r / beta0;
            if (!(r > target) || total >= maxit || !std::isfinite(r)) break;
            // restart from the true residual of the Krylov operator
            apply_A(x, z, st);
            NT_HIP(hipMemcpyAsync(Vj(0), b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            NT_LAUNCH(ice::launch_axpby2(n, -1.0, z, 1.0, Vj(0), st));
            NT_LAUNCH(ice::launch_norm2(n, Vj(0), part, h1, st));
            NT_LAUNCH(ice::launch_gmres_init(m, h1, sc, st));
            const double beta = read_scalar(beta_d, st);
            if (!(beta > target)) {
                *est = beta / beta0;
                break;
            }
        }
        return total;
    }

    // J x = b: GMRES on the Krylov operator, then FP64 refinement passes
    // xr = b - J64 x, x += GMRES(xr) until ||xr|| <= linear_rtol ||b||
    int linear_solve(const double* b, double* x, double rtol, hipStream_t st, int* refinements, double* relres) {
        const double bn = norm(b, st);
        NT_LAUNCH(ice::launch_axpby2(n, 0.0, x, 0.0, x, st));
        *refinements = 0;
        if (!(bn > 0.0)) {
            *relres = 0.0;
            return 0;
        }
        const int* cols = ice::disc_jcols(disc);
        double est = 0.0;
        int its = gmres(b, x, rtol, opt.gmres_maxit, st, &est);
        double rel = 1.0;
        for (int pass = 0;; ++pass) {
            NT_LAUNCH((ice::bsr_apply<double, double>(nelem, nslots, bs, cols, J64, x, b, xr, st)));
            rel = norm(xr, st) / bn;
            if (!fp32() || rel <= rtol || pass >= opt.refine_max || its >= opt.gmres_maxit ||
                !std::isfinite(rel))
                break;
            NT_LAUNCH(ice::launch_axpby2(n, 0.0, rwork, 0.0, rwork, st));
            its += gmres(xr, rwork, std::min(0.1, rtol / rel), opt.gmres_maxit - its, st, &est);
            NT_LAUNCH(ice::launch_axpby2(n, 1.0, rwork, 1.0, x, st));
            ++*refinements;
        }
        *relres = rel;
        return its;
    }
};

namespace ice {
int newton_residual_norm(ice_newton* s, const double* u, double* r, hipStream_t st, double* nrm) {
    if (disc_residual(s->disc, u, r, st) != 0) return -1;
    *nrm = s->norm(r, st);
    return 0;
}
}  // namespace ice

extern "C" {

int ice_newton_opts_default(ice_newton_opts* o) {
    if (!o) return ICE_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    o->tau_abs = 1e-8;
    o->tau_rel = 1e-8;
    o->kmax = 100;
    o->linesearch = ICE_LS_NONE;
    o->ls_kmax = 20;
    o->ls_alpha_initial = 1.0;
    o->ls_alpha_max = 10.0;
    o->ls_c1 = 1e-4;
    o->ls_c2 = 0.9;
    o->ls_alpha_min = 0.0;
    o->jacobian_precision = ICE_PREC_FP32;
    o->gmres_restart = 30;
    o->gmres_maxit = 10000;
    o->linear_rtol = 1e-5;
    o->refine_max = 3;
    o->verbose = 0;
    o->preconditioner = ICE_PC_BLOCK_SGS;
    o->forcing = ICE_FORCING_FIXED;
    o->smoother_precision = ICE_PREC_FP32;
    o->gmres_cgs_refine = 0;
    return ICE_OK;
}

int ice_newton_create(ice_disc* d, const ice_newton_opts* opts, ice_newton** out) {
    if (!d || !out) {
        ice::set_error("null argument");
        return ICE_ERR_ARG;
    }
    *out = nullptr;
    ice_newton_opts o;
    if (opts) o = *opts;
    else ice_newton_opts_default(&o);
    if (o.gmres_restart < 1 || o.gmres_restart > 200 || o.kmax < 0 || o.gmres_maxit < 1 ||
        (o.jacobian_precision != ICE_PREC_FP32 && o.jacobian_precision != ICE_PREC_FP64) ||
        (o.linesearch < ICE_LS_NONE || o.linesearch > ICE_LS_CORRIGAN) ||
        (o.preconditioner != ICE_PC_BLOCK_JACOBI && o.preconditioner != ICE_PC_BLOCK_SGS &&
         o.preconditioner != ICE_PC_MULTIGRID && o.preconditioner != ICE_PC_LINE) ||
        (o.forcing != ICE_FORCING_FIXED && o.forcing != ICE_FORCING_EW) ||
        (o.smoother_precision != ICE_PREC_FP32 && o.smoother_precision != ICE_PREC_FP16) ||
        (o.gmres_cgs_refine != 0 && o.gmres_cgs_refine != 1) ||
        (o.smoother_precision == ICE_PREC_FP16 &&
         (o.jacobian_precision != ICE_PREC_FP32 ||
          (o.preconditioner != ICE_PC_BLOCK_SGS && o.preconditioner != ICE_PC_MULTIGRID)))) {
        ice::set_error("invalid Newton options");
        return ICE_ERR_ARG;
    }
    try {
        auto s = std::make_unique<ice_newton>();
        s->disc = d;
        s->opt = o;
        s->m = o.gmres_restart;
        if (ice::disc_dims(d, &s->nelem, &s->nslots, &s->bs) != 0) throw Fail{ICE_ERR_UNSUPPORTED, "no Jacobian for this discretization"};
        s->n = (size_t)s->nelem * s->bs;
        s->nblkvals = (size_t)s->nelem * s->nslots * s->bs * s->bs;
        s->nred = ice::gmres_partials(s->n);
        s->npart = s->m + 2;
        s->J64 = s->alloc<double>(s->nblkvals);
        if (s->fp32()) {
            s->J32 = s->alloc<float>(s->nblkvals);
            if (o.smoother_precision == ICE_PREC_FP16) s->J16 = s->alloc<_Float16>(s->nblkvals);
            s->Dinv = s->alloc<float>((size_t)s->nelem * s->bs * s->bs);
        } else {
            s->Dinv = s->alloc<double>((size_t)s->nelem * s->bs * s->bs);
        }
        s->V = s->alloc<double>((size_t)(s->m + 1) * s->n);
        for (double** p : {&s->z, &s->t, &s->res, &s->du, &s->ustep, &s->rwork, &s->xr}) *p = s->alloc<double>(s->n);
        s->part = s->alloc<double>((size_t)s->nred * s->npart);
        s->h1 = s->alloc<double>(s->m + 2);
        s->h2 = s->alloc<double>(s->m + 2);
        s->sc = s->alloc<double>((size_t)(s->m + 1) * s->m + 2 * s->m + (s->m + 1) + 2);
        s->y = s->alloc<double>(s->m);
        s->nsing = s->alloc<int>(1);
        if (o.preconditioner == ICE_PC_LINE) {
            // the elements must form one open chain in index order: slot 1
            // (face 0) is element e - 1 and slot 2 (face 1) element e + 1
            std::vector<int> hc((size_t)s->nelem * s->nslots);
            NT_HIP(hipMemcpy(hc.data(), ice::disc_jcols(d), hc.size() * sizeof(int), hipMemcpyDeviceToHost));
            bool chain = s->nslots == 3;
            for (int e = 0; chain && e < s->nelem; ++e) {
                const int l = hc[(size_t)e * 3 + 1], r = hc[(size_t)e * 3 + 2];
                chain = (l == (e > 0 ? e - 1 : -1)) && (r == (e + 1 < s->nelem ? e + 1 : -1));
            }
            if (!chain)
                throw Fail{ICE_ERR_UNSUPPORTED, "the line preconditioner needs a non-periodic 1D mesh"};
            s->line_Sinv = s->alloc<double>((size_t)s->nelem * s->bs * s->bs);
            s->line_W = s->alloc<double>((size_t)s->nelem * s->bs * s->bs);
        } else if (o.preconditioner != ICE_PC_BLOCK_JACOBI) {
            s->colour();
        }
        if (o.preconditioner == ICE_PC_MULTIGRID) {
            if (!s->fp32()) throw Fail{ICE_ERR_UNSUPPORTED, "the multigrid preconditioner runs with FP32 blocks"};
            s->build_hierarchy();
        }
        NT_HIP(hipHostMalloc((void**)&s->host, 8 * sizeof(double)));
        *out = s.release();
        return ICE_OK;
    } catch (const Fail& f) {
        ice::set_error(f.msg);
        return f.code;
    }
}

int ice_newton_destroy(ice_newton* s) {
    delete s;
    return ICE_OK;
}

int ice_newton_linear_solve(ice_newton* s, const double* u, const double* b, double* x, int* iterations,
                            double* rel_residual, void* stream) {
    if (!s || !b || !x) {
        ice::set_error("null argument");
        return ICE_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    try {
        if (u) {
            if (ice::disc_jacobian(s->disc, u, nullptr, s->J64, st) != 0) throw Fail{ICE_ERR_HIP, "jacobian failed"};
            s->factor(st);
        }
        if (!s->factored) throw Fail{ICE_ERR_STATE, "no Jacobian assembled yet"};
        int ref = 0;
        double rel = 0.0;
        const int its = s->linear_solve(b, x, s->opt.linear_rtol, st, &ref, &rel);
        if (iterations) *iterations = its;
        if (rel_residual) *rel_residual = rel;
        NT_HIP(hipStreamSynchronize(st));
        return ICE_OK;
    } catch (const Fail& f) {
        ice::set_error(f.msg);
        return f.code;
    }
}

int ice_newton_solve(ice_newton* s, double* u, ice_newton_report* rep, void* stream) {
    if (!s || !u || !rep) {
        ice::set_error("null argument");
        return ICE_ERR_ARG;
    }
    std::memset(rep, 0, sizeof(*rep));
    hipStream_t st = (hipStream_t)stream;
    const ice_newton_opts& o = s->opt;
    const size_t n = s->n;
    try {
        const double t0 = now();
        double tj = 0, tl = 0, tls = 0;
        // residual and Jacobian at u0
        double ta = now();
        if (ice::disc_jacobian(s->disc, u, s->res, s->J64, st) != 0) throw Fail{ICE_ERR_HIP, "jacobian failed"};
        const double r0 = s->norm(s->res, st);
        tj += now() - ta;
        rep->r0 = r0;
        double rk = r0;
        bool have_jac = true;
        // linear tolerance of each step: fixed, or Eisenstat-Walker choice 2
        // (gamma 0.9, alpha 2, safeguarded, capped at 0.1; never tighter than
        // needed to reach the absolute target)
        double eta = (o.forcing == ICE_FORCING_EW) ? 0.1 : o.linear_rtol;
        double r_prev = r0;
        int k = 0;
        bool converged = false;
        for (k = 0; k < o.kmax; ++k) {
            ta = now();
            if (!have_jac) {
                if (ice::disc_jacobian(s->disc, u, nullptr, s->J64, st) != 0) throw Fail{ICE_ERR_HIP, "jacobian failed"};
            }
            s->factor(st);
            NT_HIP(hipStreamSynchronize(st));
            tj += now() - ta;
            ta = now();
            int ref = 0;
            double rel = 0.0;
            const int its = s->linear_solve(s->res, s->du, eta, st, &ref, &rel);
            tl += now() - ta;
            rep->linear_iterations += its;
            rep->refinements += ref;
            NT_HIP(hipMemcpyAsync(s->host + 1, s->nsing, sizeof(int), hipMemcpyDeviceToHost, st));
            NT_HIP(hipStreamSynchronize(st));
            rep->singular_blocks = std::max(rep->singular_blocks, *reinterpret_cast<int*>(s->host + 1));
            // linesearch on phi(alpha) = ||R(u - alpha du)||_2
            ta = now();
            double alpha = 1.0;
            if (o.linesearch != ICE_LS_NONE) {
                std::map<double, double> memo;
                memo[0.0] = rk;
                auto phi = [&](double a) -> double {
                    auto it = memo.find(a);
                    if (it != memo.end()) return it->second;
                    NT_HIP(hipMemcpyAsync(s->ustep, u, n * sizeof(double), hipMemcpyDeviceToDevice, st));
                    NT_LAUNCH(ice::launch_axpby2(n, -a, s->du, 1.0, s->ustep, st));
                    double v = 0.0;
                    if (ice::newton_residual_norm(s, s->ustep, s->rwork, st, &v) != 0)
                        throw Fail{ICE_ERR_HIP, "residual failed"};
                    v = ice::ls_merit(v);
                    if (o.verbose) std::printf("linesearch: alpha = %.6e | linesearch residual = %.6e\n", a, v);
                    memo[a] = v;
                    return v;
                };
                if (o.linesearch == ICE_LS_WOLFE) alpha = ice::wolfe_linesearch(phi, o);
                else alpha = ice::corrigan_linesearch(phi, o);
            }
            NT_LAUNCH(ice::launch_axpby2(n, -alpha, s->du, 1.0, u, st));
            tls += now() - ta;
            // residual at the new u; the Jacobian there is assembled at the
            // start of the next step only if the solve goes on
            ta = now();
            if (ice::newton_residual_norm(s, u, s->res, st, &rk) != 0) throw Fail{ICE_ERR_HIP, "residual failed"};
            tj += now() - ta;
            have_jac = false;
            if (o.forcing == ICE_FORCING_EW) {
                const double ratio = rk / r_prev;
                double e_new = 0.9 * ratio * ratio;
                const double safe = 0.9 * eta * eta;
                if (safe > 0.1) e_new = std::max(e_new, safe);
                e_new = std::min(e_new, 0.1);
                const double target = o.tau_abs + o.tau_rel * r0;
                e_new = std::max(e_new, 0.5 * target / std::max(rk, 1e-300));
                eta = std::max(e_new, o.linear_rtol * 1e-3);
            }
            r_prev = rk;
            if (k < ICE_NEWTON_MAX_HIST) {
                rep->rhist[k] = rk;
                rep->alpha[k] = alpha;
                rep->linear_its[k] = its;
                rep->nhist = k + 1;
            }
            if (o.verbose)
                std::printf("newton %d: ||R|| = %.6e alpha = %.4f gmres its = %d (rel %.2e, refinements %d)\n", k, rk,
                            alpha, its, rel, ref);
            if (rk <= o.tau_abs + o.tau_rel * r0) {  // ConservationCriteria::done_callback
                converged = true;
                ++k;
                break;
            }
        }
        rep->iterations = k;
        rep->converged = converged ? 1 : 0;
        rep->rnorm = rk;
        NT_HIP(hipStreamSynchronize(st));
        rep->seconds = now() - t0;
        rep->jacobian_seconds = tj;
        rep->linear_seconds = tl;
        rep->linesearch_seconds = tls;
        return ICE_OK;
    } catch (const Fail& f) {
        ice::set_error(f.msg);
        return f.code;
    }
}

}  // extern "C"
