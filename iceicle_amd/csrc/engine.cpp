// Host engine behind the C-ABI: FE space construction (reference tables,
// trace maps, device upload) and discretization handles.  Replaces the
// reference's FESpace constructor (fespace.hpp:200-416) and the
// ConservationLawDDG object (conservation_law.hpp:92-230) for the device path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/iceicle_amd.h"
#include "engine.h"
#include "kernels.h"
#include "mesh.hpp"
#include "physics_dev.h"
#include "fields.h"

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define ICE_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) throw ice::Error{ICE_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)}; \
    } while (0)

template <class T>
T* upload(const std::vector<T>& h) {
    T* d = nullptr;
    if (h.empty()) return nullptr;
    ICE_HIP(hipMalloc(&d, h.size() * sizeof(T)));
    ICE_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Gauss-Legendre points in the reference's order (+x1, -x1, +x2, -x2, ..., 0)
// with Newton from cos(pi (i - 1/4) / (n + 1/2)) (quadrules_1d.hpp:36-76).
void gauss_legendre(int n, double* x, double* w) {
    const int m = n / 2;
    double wc = 2.0;
    for (int i = 1; i <= m; ++i) {
        double z = std::cos(M_PI * (i - 0.25) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 50; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 1; k < n; ++k) {
                const double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::abs(dz) < 1e-16) break;
        }
        const double wi = 2.0 / ((1 - z * z) * dp * dp);
        if (n % 2) wc -= 2 * wi;
        x[2 * (i - 1)] = z;
        w[2 * (i - 1)] = wi;
        x[2 * (i - 1) + 1] = -z;
        w[2 * (i - 1) + 1] = wi;
    }
    if (n % 2) { x[n - 1] = 0.0; w[n - 1] = wc; }
}

// Equispaced Lagrange basis on [-1,1] (lagrange_1d.hpp:19-212): value,
// first and second derivative of all p+1 polynomials at xi.
void lagrange(int p, double xi, double* N, double* dN, double* d2N) {
    double xn[8], wb[8];
    xn[0] = -1.0;
    for (int j = 1; j <= p; ++j) xn[j] = xn[j - 1] + 2.0 / p;
    for (int j = 0; j <= p; ++j) {
        wb[j] = 1.0;
        for (int k = 0; k <= p; ++k)
            if (k != j) wb[j] *= (xn[j] - xn[k]);
        wb[j] = 1.0 / wb[j];
    }
    if (p == 1) {
        N[0] = 0.5 * (1 - xi);
        N[1] = 1.0 - N[0];
        dN[0] = -0.5;
        dN[1] = 0.5;
        d2N[0] = d2N[1] = 0.0;
        return;
    }
    for (int j = 0; j <= p; ++j) {
        // products over subsets excluding j: value, first and second derivative
        double v = 1.0, d1 = 0.0, d2 = 0.0;
        for (int k = 0; k <= p; ++k) {
            if (k == j) continue;
            const double f = xi - xn[k];
            d2 = d2 * f + 2.0 * d1;
            d1 = d1 * f + v;
            v = v * f;
        }
        N[j] = v * wb[j];
        dN[j] = d1 * wb[j];
        d2N[j] = d2 * wb[j];
    }
}

}  // namespace

// ---------------------------------------------------------------- handles
struct ice_space {
    ice::HostMesh mesh;
    int ndim = 0, p = 0, n = 0, nb = 0, nq = 0, nfq = 0;
    bool coupled = false;
    int batch = 1;                    // meshes concatenated into this space
    std::vector<double> tab;          // host copy of the S tables
    std::vector<int4> fint, fbdy;
    std::vector<int> efaces;
    std::vector<signed char> fpt, fvo, fperm, finv, ftens;
    std::vector<double> X;
    std::vector<int> jcols;           // [nelem][2*ndim+1] block columns of the element-block Jacobian
    ice::DevSpace dev{};
    void* dmem[14] = {};
    ~ice_space() {
        for (void* p : dmem)
            if (p) (void)hipFree(p);
    }
};

struct ice_disc {
    std::vector<hipEvent_t> prof_ev;  // 4 per profiled residual
    int prof_max = 0, prof_n = 0;
    bool profiling = false;
    ice_space* space = nullptr;
    ice_disc_params prm{};
    ice::PhysParams ph{};
    int neq = 0;
    bool ddgic = false;
    int has_source = 0;
    ice_field* d_fields = nullptr;
    double* d_Q = nullptr;
    double* d_Pic = nullptr;
    double* d_T = nullptr;
    double* d_bval = nullptr;
    double* d_srcload = nullptr;
    unsigned long long* d_wave = nullptr;
    double* d_wave_part = nullptr;
    double* d_emu = nullptr;  // per-element viscosity (batched instances)
    double* d_u64 = nullptr;  // FP64 copy of an FP32 state for the FP64 residual check, made on first use
    bool bval_values = false; // boundary data set from point values (not recomputed from the fields)
    bool src_values = false;  // source load set from point values (likewise)
    int wave_users = 0;       // CFL steppers bound to this discretization
    ice::DiscDev dd{};
    ~ice_disc() {
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        for (void* p : {(void*)d_fields, (void*)d_Q, (void*)d_Pic, (void*)d_bval, (void*)d_srcload, (void*)d_wave,
                        (void*)d_wave_part, (void*)d_T, (void*)d_emu, (void*)d_u64})
            if (p) (void)hipFree(p);
    }
};

namespace ice {

const DevSpace& space_dev(const ice_space* s) { return s->dev; }
const HostMesh& space_mesh(const ice_space* s) { return s->mesh; }
int space_batch(const ice_space* s) { return s->batch; }
bool space_coupled(const ice_space* s) { return s->coupled; }
const PhysParams& disc_phys(const ice_disc* d) { return d->ph; }
int disc_flux(const ice_disc* d) { return d->prm.flux; }

int disc_residual(ice_disc* d, const double* u, double* r, hipStream_t st) {
    const ice_space* s = d->space;
    return residual_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, st, nullptr);
}

int disc_refresh_geometry(ice_disc* d, hipStream_t st) {
    const ice_space* s = d->space;
    int rc = face_geometry_table(s->ndim, s->p, s->dev, const_cast<double*>(s->dev.fgeo), st);
    double* bv = d->bval_values ? nullptr : d->d_bval;
    double* sl = d->src_values ? nullptr : d->d_srcload;
    if (rc == 0 && (bv || sl)) rc = setup_disc_tables(s->ndim, s->p, d->neq, s->dev, d->d_fields, bv, sl, st);
    return rc;
}

// FP64 residual whose per-call buffers (element face traces, face fluxes,
// DDGIC lifting vectors, per-workgroup wave-speed maxima) are carved from
// `region` (cap doubles, 256-byte aligned) instead of the discretization's
// own.  *need receives the doubles required; with region == nullptr only
// the size is computed.  Every carved entry the kernels read is written
// earlier in the same call, so the region needs no clearing.
int disc_residual_in(ice_disc* d, const double* u, double* r, hipStream_t st, double* region, size_t cap,
                     size_t* need) {
    const ice_space* s = d->space;
    const int D = s->ndim;
    auto up = [](size_t n) { return (n + 31) / 32 * 32; };  // 256-byte sub-buffers
    const size_t rows = (size_t)(s->dev.nfi + s->dev.nfb) * s->nfq * d->neq;
    const size_t nT = up(std::max<size_t>(trace_buffer_size(D, s->p, d->prm.flux, s->mesh.nelem), 1));
    const size_t nQ = up(std::max<size_t>(rows, 1));
    const size_t nP = d->ddgic ? up(rows * D) : 0;
    const size_t nW = d->dd.wave ? up(2 * std::max<size_t>(s->mesh.nelem, 1)) : 0;
    const size_t total = nT + nQ + nP + nW;
    if (need) *need = total;
    if (!region) return 0;
    if (total > cap) return -2;
    DiscDev dd = d->dd;
    dd.T = region;
    dd.Q = region + nT;
    dd.Pic = d->ddgic ? region + nT + nQ : nullptr;
    if (d->dd.wave) dd.wave_part = region + nT + nQ + nP;
    return residual_fp64(D, s->p, d->prm.flux, s->dev, d->ph, dd, u, r, st, nullptr);
}

// residual (if r) and FD block Jacobian at u
int disc_jacobian(ice_disc* d, const double* u, double* r, double* J, hipStream_t st) {
    const ice_space* s = d->space;
    if (r && disc_residual(d, u, r, st) != 0) return -1;
    const size_t bs = (size_t)s->nb * d->neq;
    const size_t nvals = (size_t)s->mesh.nelem * (2 * s->ndim + 1) * bs * bs;
    return jacobian_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, J, nvals, st);
}

// Jacobian layout; nonzero when the device Jacobian does not cover the discretization
int disc_dims(const ice_disc* d, int* nelem, int* nslots, int* bs) {
    const ice_space* s = d->space;
    *nelem = s->mesh.nelem;
    *nslots = 2 * s->ndim + 1;
    *bs = s->nb * d->neq;
    // the FD Jacobian kernels use the discretization's viscosity, not per-element ones
    return ((s->ndim > 2 && s->p > 2) || d->d_emu) ? -1 : 0;
}

const int* disc_jcols(const ice_disc* d) { return d->space->dev.jcols; }

// elements per direction of the structured mesh (x fastest), neq and nbasis
void disc_grid(const ice_disc* d, int* nel, int* neq, int* nbasis) {
    const ice_space* s = d->space;
    for (int k = 0; k < 3; ++k) nel[k] = s->mesh.nel[k];
    *neq = d->neq;
    *nbasis = s->nb;
}
const ice_space* disc_space(const ice_disc* d) { return d->space; }
int disc_neq(const ice_disc* d) { return d->neq; }
void set_error(const std::string& m) { g_err = m; }

}  // namespace ice

extern "C" {

int ice_version(void) { return ICE_API_VERSION; }
const char* ice_last_error(void) { return g_err.c_str(); }

int ice_disc_params_default(int flux, ice_disc_params* out) {
    if (!out) return fail(ICE_ERR_ARG, "null output");
    std::memset(out, 0, sizeof(*out));
    out->flux = flux;
    out->mu = 0.001;
    out->gamma = 1.4;
    out->Rgas = 287.052874;
    out->Pr = 0.72;
    for (double& r : out->ref) r = 1.0;
    out->viscosity_kind = ICE_VISC_CONSTANT;
    out->visc_mu = 1.0;
    for (double& t : out->isothermal_T) t = 1.0;
    out->freestream[0] = 1.0;
    out->freestream[2] = 1.0;
    out->freestream[5] = 1.0;
    return ICE_OK;
}

int ice_mesh_export(const ice_mesh_desc* d, ice_mesh_counts* counts, double* coord, int* conn, int* faces) {
    if (!d || !counts) return fail(ICE_ERR_ARG, "null argument");
    try {
        const ice::HostMesh m = ice::build_mesh(*d, false);
        const int D = m.ndim;
        const int nfv = (D > 1) ? (1 << (D - 1)) : 1;
        counts->ndim = D;
        counts->nelem = m.nelem;
        counts->nnode = m.nnode;
        counts->nvert_el = 1 << D;
        counts->nfacenode = nfv;
        counts->nface = (int)m.faces.size();
        counts->nface_interior = m.n_interior;
        if (coord) std::memcpy(coord, m.coord.data(), m.coord.size() * sizeof(double));
        if (conn) std::memcpy(conn, m.conn.data(), m.conn.size() * sizeof(int));
        if (faces) {
            int k = 0;
            for (const auto& f : m.faces) {
                faces[k++] = f.elemL;
                faces[k++] = f.elemR;
                faces[k++] = f.fnl * 512;
                faces[k++] = f.fnr * 512 + f.orient;
                faces[k++] = f.bctype;
                faces[k++] = f.bcflag;
                for (int j = 0; j < nfv; ++j) faces[k++] = f.nodes[j];
            }
        }
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
}

// Interior (and coupled periodic pair) faces and boundary faces of one
// structured mesh as device records, element ids local to the mesh.
static void face_lists(const ice::HostMesh& m, bool coupled, std::vector<int4>& fint, std::vector<int4>& fbdy) {
    const int D = m.ndim;
    for (int f = 0; f < m.n_interior; ++f) {
        const auto& fc = m.faces[f];
        fint.push_back(int4{fc.elemL, fc.elemR, ice::pack_face_info(fc.fnl, fc.fnr, fc.orient, ICE_BC_INTERIOR, 0), 0});
    }
    if (coupled) {
        for (int f = m.n_interior; f < (int)m.faces.size(); ++f) {
            const auto& fa = m.faces[f];
            if (fa.bctype != ICE_BC_PERIODIC || fa.fnl >= D) continue;
            const auto& fb = m.faces[fa.bcflag];
            const int idim = fa.fnl;
            const int nfv = (D > 1) ? (1 << (D - 1)) : 1;
            int va[4];
            for (int k = 0; k < nfv; ++k) {
                // partner of A's node on B's side: same ordinates, last index along idim
                int r = fa.nodes[k], ijk[3] = {0, 0, 0}, stride = 1, idx = 0;
                for (int d = 0; d < D; ++d) { ijk[d] = r % m.nnd[d]; r /= m.nnd[d]; }
                ijk[idim] = m.nnd[idim] - 1;
                for (int d = 0; d < D; ++d) { idx += ijk[d] * stride; stride *= m.nnd[d]; }
                va[k] = idx;
            }
            const int orient = ice::face_orientation(D, fb.nodes, va);
            if (orient < 0) throw ice::Error{ICE_ERR_ARG, "periodic faces do not match"};
            fint.push_back(int4{fb.elemL, fa.elemL, ice::pack_face_info(fb.fnl, fa.fnl, orient, ICE_BC_INTERIOR, idim + 1), 0});
        }
    }
    for (int f = m.n_interior; f < (int)m.faces.size(); ++f) {
        const auto& fc = m.faces[f];
        if (coupled && fc.bctype == ICE_BC_PERIODIC) continue;
        fbdy.push_back(int4{fc.elemL, ice::pack_face_info(fc.fnl, 0, 0, fc.bctype, 0), fc.bcflag, 0});
    }
}

// Several meshes of the same dimension and periodic mode concatenated into
// one space (elements of mesh k after those of meshes 0..k-1); one mesh is
// the ordinary space.
static int space_create(const ice_mesh_desc* mds, int nmesh, const ice_space_desc* sd, ice_space** out) {
    if (!mds || !sd || !out || nmesh < 1) return fail(ICE_ERR_ARG, "null argument");
    *out = nullptr;
    if (sd->order < 1 || sd->order > ICE_MAX_ORDER) return fail(ICE_ERR_UNSUPPORTED, "order must be 1..ICE_MAX_ORDER");
    if (sd->geometry_order != 1) return fail(ICE_ERR_UNSUPPORTED, "only geometry order 1 is supported");
    for (int k = 0; k < nmesh; ++k) {
        if (mds[k].periodic_mode != ICE_PERIODIC_REFERENCE && mds[k].periodic_mode != ICE_PERIODIC_COUPLED)
            return fail(ICE_ERR_ARG, "bad periodic_mode");
        if (mds[k].periodic_mode != mds[0].periodic_mode || mds[k].ndim != mds[0].ndim)
            return fail(ICE_ERR_ARG, "batched meshes must share ndim and periodic_mode");
    }
    try {
        auto s = std::make_unique<ice_space>();
        s->coupled = mds[0].periodic_mode == ICE_PERIODIC_COUPLED;
        std::vector<ice::HostMesh> meshes;
        for (int k = 0; k < nmesh; ++k) meshes.push_back(ice::build_mesh(mds[k], s->coupled));
        if (nmesh == 1) {
            s->mesh = meshes[0];
        } else {
            // merged host mesh: element, node and face ids offset; the
            // structured dimensions are left 0 (no single grid)
            ice::HostMesh& mm = s->mesh;
            mm.ndim = meshes[0].ndim;
            for (int d = 0; d < 3; ++d) {
                mm.nel[d] = 0;
                mm.nnd[d] = 0;
                mm.period[d] = meshes[0].period[d];
            }
            int noff = 0, eoff = 0, foff = 0;
            for (const auto& mk : meshes) {
                for (int d = 0; d < mm.ndim; ++d)
                    if (s->coupled && std::abs(mk.period[d] - mm.period[d]) > 1e-14 * std::max(1.0, std::abs(mm.period[d])))
                        throw ice::Error{ICE_ERR_ARG, "coupled periodic batched meshes must share the box extents"};
                mm.coord.insert(mm.coord.end(), mk.coord.begin(), mk.coord.end());
                for (int c : mk.conn) mm.conn.push_back(c + noff);
                for (auto f : mk.faces) {
                    f.elemL += eoff;
                    if (f.elemR >= 0) f.elemR += eoff;
                    if (f.bctype == ICE_BC_PERIODIC) f.bcflag += foff;
                    for (int v = 0; v < ice::MAX_NFV; ++v) f.nodes[v] += noff;
                    mm.faces.push_back(f);
                }
                mm.n_interior += mk.n_interior;
                noff += mk.nnode;
                eoff += mk.nelem;
                foff += (int)mk.faces.size();
            }
            mm.nelem = eoff;
            mm.nnode = noff;
        }
        s->batch = nmesh;
        const ice::HostMesh& m = s->mesh;
        const int D = m.ndim, p = sd->order, n = p + 1;
        s->ndim = D;
        s->p = p;
        s->n = n;
        s->nb = 1;
        for (int d = 0; d < D; ++d) s->nb *= n;
        s->nq = s->nb;
        s->nfq = s->nb / n;
        const int NV = 1 << D;
        // element vertex coordinates, element-local tensor order
        s->X.resize((size_t)m.nelem * NV * D);
        for (int e = 0; e < m.nelem; ++e)
            for (int v = 0; v < NV; ++v)
                for (int d = 0; d < D; ++d)
                    s->X[((size_t)e * NV + v) * D + d] = m.coord[(size_t)m.conn[(size_t)e * NV + v] * D + d];
        // S tables: GL points, then -1, +1
        double gx[8], gw[8];
        gauss_legendre(n, gx, gw);
        const int ns = n + 2;
        s->tab.assign(ns + 3 * ns * n + n + n * n, 0.0);
        double* xs = s->tab.data();
        double* L0 = xs + ns;
        double* L1 = L0 + ns * n;
        double* L2 = L1 + ns * n;
        double* wg = L2 + ns * n;
        for (int k = 0; k < n; ++k) { xs[k] = gx[k]; wg[k] = gw[k]; }
        xs[n] = -1.0;
        xs[n + 1] = 1.0;
        for (int k = 0; k < ns; ++k) lagrange(p, xs[k], L0 + k * n, L1 + k * n, L2 + k * n);
        // Binv = B1^-1 with B1[q][i] = l_i(x_q): the mass matrix of any element is
        // B^T W B with B = B1 x ... x B1 square, so M^-1 = B^-1 W^-1 B^-T
        {
            double* Binv = wg + n;
            std::vector<double> a(L0, L0 + n * n), inv(n * n, 0.0);
            for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
            for (int c = 0; c < n; ++c) {  // Gauss-Jordan with partial pivoting
                int piv = c;
                for (int r = c + 1; r < n; ++r)
                    if (std::abs(a[r * n + c]) > std::abs(a[piv * n + c])) piv = r;
                for (int j = 0; j < n; ++j) {
                    std::swap(a[c * n + j], a[piv * n + j]);
                    std::swap(inv[c * n + j], inv[piv * n + j]);
                }
                const double d = a[c * n + c];
                for (int j = 0; j < n; ++j) { a[c * n + j] /= d; inv[c * n + j] /= d; }
                for (int r = 0; r < n; ++r) {
                    if (r == c) continue;
                    const double f = a[r * n + c];
                    for (int j = 0; j < n; ++j) { a[r * n + j] -= f * a[c * n + j]; inv[r * n + j] -= f * inv[c * n + j]; }
                }
            }
            for (int k = 0; k < n * n; ++k) Binv[k] = inv[k];  // Binv[i][q]
        }
        // trace point and face vertex tables
        const int no = ice::n_orientations(D);
        s->fpt.resize((size_t)2 * D * no * s->nfq * D);
        for (int j = 0; j < 2 * D; ++j)
            for (int o = 0; o < no; ++o)
                for (int q = 0; q < s->nfq; ++q) {
                    int pt[3];
                    ice::trace_point_indices(D, n, j, o, q, pt);
                    for (int d = 0; d < D; ++d) s->fpt[(((size_t)j * no + o) * s->nfq + q) * D + d] = (signed char)pt[d];
                }
        // face point q of orientation o -> the same point in the element's canonical (orientation 0) order
        s->fperm.assign((size_t)2 * D * no * s->nfq, -1);
        s->finv.assign((size_t)2 * D * no * s->nfq, -1);
        for (int j = 0; j < 2 * D; ++j)
            for (int o = 0; o < no; ++o)
                for (int q = 0; q < s->nfq; ++q) {
                    const signed char* a = &s->fpt[(((size_t)j * no + o) * s->nfq + q) * D];
                    int hit = -1;
                    for (int q2 = 0; q2 < s->nfq && hit < 0; ++q2) {
                        const signed char* b = &s->fpt[(((size_t)j * no + 0) * s->nfq + q2) * D];
                        if (std::memcmp(a, b, D) == 0) hit = q2;
                    }
                    if (hit < 0) throw ice::Error{ICE_ERR_ARG, "orientation does not permute the face points"};
                    s->fperm[((size_t)j * no + o) * s->nfq + q] = (signed char)hit;
                    s->finv[((size_t)j * no + o) * s->nfq + hit] = (signed char)q;
                }
        // face point q of orientation o -> its index in the element's tensor
        // order of the face's tangential axes (increasing axis, slowest first)
        s->ftens.assign((size_t)2 * D * no * s->nfq, 0);
        for (int j = 0; j < 2 * D; ++j)
            for (int o = 0; o < no; ++o)
                for (int q = 0; q < s->nfq; ++q) {
                    const signed char* a = &s->fpt[(((size_t)j * no + o) * s->nfq + q) * D];
                    int idx = 0;
                    for (int d = 0; d < D; ++d)
                        if (d != j % D) idx = idx * n + a[d];
                    s->ftens[((size_t)j * no + o) * s->nfq + q] = (signed char)idx;
                }
        s->fvo.assign(2 * D * 4, 0);
        for (int j = 0; j < 2 * D; ++j) {
            int v[4];
            ice::face_vertex_order(D, j, v);
            for (int k = 0; k < ((D > 1) ? (1 << (D - 1)) : 1); ++k) s->fvo[j * 4 + k] = (signed char)v[k];
        }
        // device face lists: every mesh's interior faces (and coupled
        // periodic pairs), then every mesh's boundary faces, element ids
        // offset by the elements of the meshes before it
        {
            int eoff = 0;
            std::vector<std::vector<int4>> bd(meshes.size());
            for (size_t k = 0; k < meshes.size(); ++k) {
                std::vector<int4> fi;
                face_lists(meshes[k], s->coupled, fi, bd[k]);
                for (int4 r : fi) s->fint.push_back(int4{r.x + eoff, r.y + eoff, r.z, r.w});
                for (int4& r : bd[k]) r.x += eoff;
                eoff += meshes[k].nelem;
            }
            for (const auto& b : bd) s->fbdy.insert(s->fbdy.end(), b.begin(), b.end());
        }
        s->efaces.assign((size_t)m.nelem * 2 * D, -1);
        const int nfi = (int)s->fint.size(), nfb = (int)s->fbdy.size();
        for (int f = 0; f < nfi; ++f) {
            const int4 r = s->fint[f];
            const int fnl = r.z & 7, fnr = (r.z >> 3) & 7, orient = (r.z >> 6) & 7;
            s->efaces[(size_t)r.x * 2 * D + fnl] = (f << 4);
            s->efaces[(size_t)r.y * 2 * D + fnr] = (f << 4) | (1 << 3) | orient;
        }
        for (int b = 0; b < nfb; ++b) {
            const int4 r = s->fbdy[b];
            s->efaces[(size_t)r.x * 2 * D + (r.y & 7)] = ((nfi + b) << 4);
        }
        // Jacobian block columns: the element itself, then the element across each local face
        {
            const int ns = 2 * D + 1;
            s->jcols.assign((size_t)m.nelem * ns, -1);
            for (int e = 0; e < m.nelem; ++e) {
                s->jcols[(size_t)e * ns] = e;
                for (int j = 0; j < 2 * D; ++j) {
                    const int v = s->efaces[(size_t)e * 2 * D + j];
                    if (v < 0 || (v >> 4) >= nfi) continue;
                    const int4 r = s->fint[v >> 4];
                    s->jcols[(size_t)e * ns + 1 + j] = ((v >> 3) & 1) ? r.x : r.y;
                }
            }
        }
        // Jacobian face colouring (greedy, in face order): the lowest colour
        // not yet used by a face of either element
        std::vector<int> jorder(std::max(nfi + nfb, 1));
        auto colour_faces = [&](int nf, auto elems, int* coff, int& ncol, int base) {
            std::vector<unsigned> used((size_t)m.nelem, 0u);
            std::vector<int> col(nf);
            ncol = 0;
            for (int f = 0; f < nf; ++f) {
                const auto [a, b] = elems(f);
                const unsigned busy = used[a] | used[b];
                int c = 0;
                while (busy & (1u << c)) ++c;
                if (c >= ice::DevSpace::MAX_FACE_COLOURS) throw ice::Error{ICE_ERR_UNSUPPORTED, "face colouring overflow"};
                col[f] = c;
                used[a] |= 1u << c;
                used[b] |= 1u << c;
                ncol = std::max(ncol, c + 1);
            }
            std::fill(coff, coff + ice::DevSpace::MAX_FACE_COLOURS + 1, 0);
            for (int f = 0; f < nf; ++f) ++coff[col[f] + 1];
            for (int c = 0; c < ice::DevSpace::MAX_FACE_COLOURS; ++c) coff[c + 1] += coff[c];
            std::vector<int> fill(coff, coff + ice::DevSpace::MAX_FACE_COLOURS);
            for (int f = 0; f < nf; ++f) jorder[base + fill[col[f]]++] = f;
        };
        ice::DevSpace& dvc = s->dev;
        colour_faces(nfi, [&](int f) { return std::pair<int, int>(s->fint[f].x, s->fint[f].y); }, dvc.coff_i,
                     dvc.ncol_i, 0);
        colour_faces(nfb, [&](int f) { return std::pair<int, int>(s->fbdy[f].x, s->fbdy[f].x); }, dvc.coff_b,
                     dvc.ncol_b, nfi);
        s->dmem[13] = upload(jorder);
        dvc.jorder = (const int*)s->dmem[13];
        // upload
        s->dmem[7] = upload(s->jcols);
        s->dmem[0] = upload(s->X);
        s->dmem[1] = upload(s->fint);
        s->dmem[2] = upload(s->fbdy);
        s->dmem[3] = upload(s->efaces);
        s->dmem[4] = upload(s->tab);
        s->dmem[5] = upload(s->fpt);
        s->dmem[6] = upload(s->fvo);
        ice::DevSpace& dv = s->dev;
        dv.ndim = D;
        dv.p = p;
        dv.n = n;
        dv.nelem = m.nelem;
        dv.nfi = nfi;
        dv.nfb = nfb;
        dv.norient = no;
        for (int d = 0; d < 3; ++d) dv.period[d] = m.period[d];
        dv.X = (const double*)s->dmem[0];
        dv.fint = (const int4*)s->dmem[1];
        dv.fbdy = (const int4*)s->dmem[2];
        dv.efaces = (const int*)s->dmem[3];
        dv.tab = (const double*)s->dmem[4];
        dv.fpt = (const signed char*)s->dmem[5];
        dv.fvo = (const signed char*)s->dmem[6];
        dv.jcols = (const int*)s->dmem[7];
        s->dmem[8] = upload(s->fperm);
        s->dmem[9] = upload(s->finv);
        {
            std::vector<float> tf(s->tab.begin(), s->tab.end());
            s->dmem[11] = upload(tf);
            dv.tabf = (const float*)s->dmem[11];
        }
        dv.fperm = (const signed char*)s->dmem[8];
        dv.finv = (const signed char*)s->dmem[9];
        s->dmem[12] = upload(s->ftens);
        dv.ftens = (const signed char*)s->dmem[12];
        {
            const size_t ng = (size_t)(nfi + nfb) * s->nfq * (D + 2);
            ICE_HIP(hipMalloc(&s->dmem[10], std::max<size_t>(ng, 1) * sizeof(double)));
            dv.fgeo = (const double*)s->dmem[10];
            if (ice::face_geometry_table(D, p, dv, (double*)s->dmem[10], nullptr) != 0)
                throw ice::Error{ICE_ERR_HIP, "face geometry launch failed"};
            ICE_HIP(hipDeviceSynchronize());
        }
        *out = s.release();
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    } catch (const std::exception& e) {
        return fail(ICE_ERR_ALLOC, e.what());
    }
}

int ice_space_create(const ice_mesh_desc* md, const ice_space_desc* sd, ice_space** out) {
    return space_create(md, 1, sd, out);
}

int ice_space_create_batch(const ice_mesh_desc* mds, int nmesh, const ice_space_desc* sd, ice_space** out) {
    return space_create(mds, nmesh, sd, out);
}

int ice_space_destroy(ice_space* s) {
    delete s;
    return ICE_OK;
}

int ice_space_query(const ice_space* s, ice_space_info* info) {
    if (!s || !info) return fail(ICE_ERR_ARG, "null argument");
    info->ndim = s->ndim;
    info->order = s->p;
    info->nbasis = s->nb;
    info->nqp = s->nq;
    info->nfqp = s->nfq;
    info->nelem = s->mesh.nelem;
    info->nface = s->dev.nfi + s->dev.nfb;
    info->nface_interior = s->dev.nfi;
    info->nface_boundary = s->dev.nfb;
    info->ndof = (int64_t)s->mesh.nelem * s->nb;
    return ICE_OK;
}

int ice_space_qp_coords(const ice_space* s, double* xq) {
    if (!s || !xq) return fail(ICE_ERR_ARG, "null argument");
    const int D = s->ndim, n = s->n, NV = 1 << D;
    for (int e = 0; e < s->mesh.nelem; ++e)
        for (int q = 0; q < s->nq; ++q) {
            int r = q, qd[3];
            for (int d = D - 1; d >= 0; --d) { qd[d] = r % n; r /= n; }
            double x[3] = {0, 0, 0};
            for (int v = 0; v < NV; ++v) {
                double phi = 1.0;
                for (int d = 0; d < D; ++d) {
                    const double n0 = 0.5 * (1 - s->tab[qd[d]]);
                    phi *= ((v >> (D - 1 - d)) & 1) ? (1.0 - n0) : n0;
                }
                for (int d = 0; d < D; ++d) x[d] += phi * s->X[((size_t)e * NV + v) * D + d];
            }
            for (int d = 0; d < D; ++d) xq[((size_t)e * s->nq + q) * D + d] = x[d];
        }
    return ICE_OK;
}

// Flux parameters of a discretization (reference nondimensionalisation,
// navier_stokes.hpp:129-213; DDG coefficients, conservation_law.hpp:518-527).
static void derive_phys(const ice_disc_params* p, int order, ice::PhysParams& ph) {
    std::memset(&ph, 0, sizeof(ph));
    ph.mu = p->mu;
    for (int k = 0; k < 3; ++k) { ph.a[k] = p->a[k]; ph.b[k] = p->b[k]; }
    ph.gamma = p->gamma;
    ph.Pr = p->Pr;
    const double* r = p->ref;  // t, l, rho, u, e, p, T, mu
    ph.Re = r[2] * r[3] * r[1] / r[7];
    ph.Eu = r[5] / (r[2] * r[3] * r[3]);
    ph.e_coeff = r[3] * r[3] / r[4];
    const double cp = p->gamma / (p->gamma - 1) * p->Rgas;
    ph.T_coeff = cp * r[6] * r[2] / r[5];
    ph.visc_kind = p->viscosity_kind;
    ph.visc_mu = p->visc_mu;
    ph.mu_ratio = 1.716e-5 / r[7];
    ph.T0r = 273.1 / r[6];
    ph.Tsr = 110.5 / r[6];
    for (int k = 0; k < ICE_MAX_BCFLAG; ++k) ph.isoT[k] = p->isothermal_T[k];
    ph.fs_rho = p->freestream[0];
    ph.fs_u = p->freestream[1];
    for (int k = 0; k < 3; ++k) ph.fs_dir[k] = p->freestream[2 + k];
    ph.fs_T = p->freestream[5];
    ph.textbook_energy = p->ns_face_energy_textbook;
    ph.sigma_ic = p->sigma_ic;
    // interface correction (conservation_law.hpp:580-610, :750-780): SOURCE
    // keeps the average operator's 1/2 per side and the Dirichlet term,
    // PUBLISHED applies sigma per side, no boundary term, and each side's
    // test gradient averaged between the face point and its mirror through
    // the element along the face normal (the published DDGIC table,
    // doc/sine_diffusion_1d_study.md:45-55)
    const bool pub = p->ddgic_form == ICE_DDGIC_PUBLISHED;
    ph.sigma_face = pub ? p->sigma_ic : 0.5 * p->sigma_ic;
    ph.sigma_bdy = pub ? 0.0 : p->sigma_ic;
    ph.ic_mirror = pub ? 1 : 0;
    ph.beta0 = (double)(order + 1) * (order + 1);
    ph.beta1 = p->interior_penalty ? 0.0 : 1.0 / std::max((double)(2 * order * (order + 1)), 1.0);
    ph.beta0_bdy = ph.beta0;
    ice::set_derived(ph);
}

int ice_disc_create(ice_space* s, const ice_disc_params* p, ice_disc** out) {
    if (!s || !p || !out) return fail(ICE_ERR_ARG, "null argument");
    *out = nullptr;
    const int D = s->ndim;
    if (p->flux != ICE_FLUX_BURGERS && p->flux != ICE_FLUX_NAVIER_STOKES && p->flux != ICE_FLUX_EULER)
        return fail(ICE_ERR_ARG, "unknown flux family");
    if (p->ddgic_form != ICE_DDGIC_SOURCE && p->ddgic_form != ICE_DDGIC_PUBLISHED)
        return fail(ICE_ERR_ARG, "ddgic_form must be ICE_DDGIC_SOURCE or ICE_DDGIC_PUBLISHED");
    const bool ns = p->flux != ICE_FLUX_BURGERS;
    // boundary types the reference defines for this flux family
    for (const auto& f : s->mesh.faces) {
        if (f.bctype == ICE_BC_INTERIOR) continue;
        const int bt = f.bctype;
        if ((bt == ICE_BC_DIRICHLET || bt == ICE_BC_NEUMANN || bt == ICE_BC_NO_SLIP_ISOTHERMAL) &&
            (f.bcflag < 0 || f.bcflag >= ICE_MAX_BCFLAG))
            return fail(ICE_ERR_ARG, "bcflag out of range for a Dirichlet/Neumann/isothermal boundary");
        if (ns) {
            const bool ok = bt == ICE_BC_DIRICHLET || bt == ICE_BC_NEUMANN || bt == ICE_BC_EXTRAPOLATION ||
                            bt == ICE_BC_SPACETIME_FUTURE || bt == ICE_BC_SLIP_WALL || bt == ICE_BC_NO_SLIP_ISOTHERMAL ||
                            bt == ICE_BC_WALL_GENERAL || bt == ICE_BC_RIEMANN || (bt == ICE_BC_PERIODIC && s->coupled);
            if (!ok)
                return fail(ICE_ERR_UNSUPPORTED,
                            "boundary condition has no Navier-Stokes implementation in the reference (apply_bc logs an "
                            "anomaly and returns an uninitialised state)");
        }
    }
    try {
        auto d = std::make_unique<ice_disc>();
        d->space = s;
        d->prm = *p;
        d->neq = ns ? D + 2 : 1;
        derive_phys(p, s->p, d->ph);
        d->ddgic = (!ns) && p->sigma_ic != 0.0;
        d->has_source = p->source.kind != ICE_FIELD_NONE;
        std::vector<ice_field> fields(2 * ICE_MAX_BCFLAG + 1);
        for (int k = 0; k < ICE_MAX_BCFLAG; ++k) {
            fields[k] = p->dirichlet[k];
            fields[ICE_MAX_BCFLAG + k] = p->neumann[k];
        }
        fields[2 * ICE_MAX_BCFLAG] = p->source;
        d->d_fields = upload(fields);
        const size_t rows = (size_t)(s->dev.nfi + s->dev.nfb) * s->nfq * d->neq;
        ICE_HIP(hipMalloc(&d->d_Q, std::max<size_t>(rows, 1) * sizeof(double)));
        if (d->ddgic) ICE_HIP(hipMalloc(&d->d_Pic, std::max<size_t>(rows * D, 1) * sizeof(double)));
        ICE_HIP(hipMalloc(&d->d_T, std::max<size_t>(ice::trace_buffer_size(D, s->p, p->flux, s->mesh.nelem), 1) *
                                     sizeof(double)));
        if (s->dev.nfb > 0) ICE_HIP(hipMalloc(&d->d_bval, (size_t)s->dev.nfb * s->nfq * d->neq * sizeof(double)));
        if (d->has_source) ICE_HIP(hipMalloc(&d->d_srcload, (size_t)s->mesh.nelem * s->nb * d->neq * sizeof(double)));
        if (ice::setup_disc_tables(D, s->p, d->neq, s->dev, d->d_fields, d->d_bval, d->d_srcload, nullptr) != 0)
            throw ice::Error{ICE_ERR_HIP, "table setup launch failed"};
        ICE_HIP(hipMalloc(&d->d_wave, 4 * sizeof(unsigned long long)));
        ICE_HIP(hipMemset(d->d_wave, 0, 4 * sizeof(unsigned long long)));
        // per-workgroup maxima of the element kernels: at most one workgroup per element
        ICE_HIP(hipMalloc(&d->d_wave_part, 2 * std::max<size_t>(s->mesh.nelem, 1) * sizeof(double)));
        ICE_HIP(hipDeviceSynchronize());
        d->dd.wave = nullptr;  // wave-speed bookkeeping only while a CFL stepper uses it
        d->dd.wave_part = d->d_wave_part;
        d->dd.fields = d->d_fields;
        d->dd.bval = d->d_bval;
        d->dd.srcload = d->d_srcload;
        d->dd.Q = d->d_Q;
        d->dd.Pic = d->d_Pic;
        d->dd.T = d->d_T;
        *out = d.release();
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
}

int ice_disc_destroy(ice_disc* d) {
    delete d;
    return ICE_OK;
}

int ice_disc_query(const ice_disc* d, int* neq, int64_t* size) {
    if (!d) return fail(ICE_ERR_ARG, "null discretization");
    if (neq) *neq = d->neq;
    if (size) *size = (int64_t)d->space->dev.nelem * d->space->nb * d->neq;
    return ICE_OK;
}

int ice_disc_update(ice_disc* d, const ice_disc_params* p) {
    if (!d || !p) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    if (p->flux != d->prm.flux) return fail(ICE_ERR_ARG, "the flux family of a discretization is fixed");
    const bool ddgic = p->flux == ICE_FLUX_BURGERS && p->sigma_ic != 0.0;
    const int has_source = p->source.kind != ICE_FIELD_NONE;
    if (ddgic != d->ddgic || has_source != d->has_source)
        return fail(ICE_ERR_ARG, "DDGIC and the source term must stay on or off as at creation");
    try {
        derive_phys(p, s->p, d->ph);
        d->prm = *p;
        std::vector<ice_field> fields(2 * ICE_MAX_BCFLAG + 1);
        for (int k = 0; k < ICE_MAX_BCFLAG; ++k) {
            fields[k] = p->dirichlet[k];
            fields[ICE_MAX_BCFLAG + k] = p->neumann[k];
        }
        fields[2 * ICE_MAX_BCFLAG] = p->source;
        ICE_HIP(hipMemcpy(d->d_fields, fields.data(), fields.size() * sizeof(ice_field), hipMemcpyHostToDevice));
        if (ice::setup_disc_tables(s->ndim, s->p, d->neq, s->dev, d->d_fields, d->bval_values ? nullptr : d->d_bval,
                                   d->src_values ? nullptr : d->d_srcload, nullptr) != 0)
            throw ice::Error{ICE_ERR_HIP, "table setup launch failed"};
        ICE_HIP(hipDeviceSynchronize());
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
}

int ice_disc_set_element_viscosity(ice_disc* d, const double* mu) {
    if (!d) return fail(ICE_ERR_ARG, "null argument");
    if (d->prm.flux != ICE_FLUX_BURGERS) return fail(ICE_ERR_UNSUPPORTED, "per-element viscosity is a Burgers option");
    if (d->wave_users > 0) return fail(ICE_ERR_STATE, "a CFL stepper uses this discretization's viscosity");
    const size_t ne = (size_t)d->space->mesh.nelem;
    try {
        if (!mu) {
            if (d->d_emu) ICE_HIP(hipFree(d->d_emu));
            d->d_emu = nullptr;
        } else {
            if (!d->d_emu) ICE_HIP(hipMalloc(&d->d_emu, std::max<size_t>(ne, 1) * sizeof(double)));
            ICE_HIP(hipMemcpy(d->d_emu, mu, ne * sizeof(double), hipMemcpyHostToDevice));
        }
        d->dd.emu = d->d_emu;
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
}

int ice_disc_set_tile(ice_disc* d, int tile) {
    if (!d) return fail(ICE_ERR_ARG, "null argument");
    if (tile < 0 || tile > 3) return fail(ICE_ERR_ARG, "tile must be 0, 1, 2 or 3");
    d->dd.tile = tile;
    d->dd.tile_mixed = tile;
    return ICE_OK;
}

int ice_disc_autotune(ice_disc* d, const double* u, double* r, int reps, int* best, double* ms, void* stream) {
    if (!d || !u || !r || reps < 1) return fail(ICE_ERR_ARG, "bad argument");
    const ice_space* s = d->space;
    if (s->ndim < 2) {  // one element kernel shape in 1D
        if (best) *best = 0;
        return ICE_OK;
    }
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(ICE_ERR_HIP, "event create failed");
    int pick = 0, pick_mixed = 0;
    int rc = ICE_OK;
    // FP64 residual, then the mixed one (its own tile: FP32 tiles hold half
    // the bytes, so the fastest shape can differ)
    for (int prec = 0; prec < 2 && rc == ICE_OK; ++prec) {
        double tbest = 1e300;
        // tile 3 (matrix cores) exists for the mixed residual of 3D P3 Euler only
        const bool mfma = prec == 1 && s->ndim == 3 && s->p == 3 && d->prm.flux == ICE_FLUX_EULER;
        for (int tile = 0; tile < (mfma ? 4 : 3) && rc == ICE_OK; ++tile) {
            d->dd.tile = tile;
            d->dd.tile_mixed = tile;
            auto once = [&]() {
                return (prec == 0 ? ice::residual_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, st,
                                                       nullptr)
                                  : ice::residual_mixed(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, st,
                                                        nullptr)) == 0;
            };
            float t = 0.f;
            bool ok = true;
            for (int k = 0; k < 2 && ok; ++k) ok = once();
            ok = ok && hipEventRecord(e0, st) == hipSuccess;
            for (int k = 0; k < reps && ok; ++k) ok = once();
            ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&t, e0, e1) == hipSuccess;
            if (!ok) {
                rc = fail(ICE_ERR_HIP, "autotune timing failed");
                break;
            }
            const double per = (double)t / reps;
            if (ms && prec == 0) ms[tile] = per;
            if (per < tbest) {
                tbest = per;
                (prec == 0 ? pick : pick_mixed) = tile;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    d->dd.tile = pick;
    d->dd.tile_mixed = pick_mixed;
    if (best) *best = pick;
    return rc;
}

int ice_residual(ice_disc* d, const double* u, double* r, void* stream) {
    if (!d || !u || !r) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    hipEvent_t* ev = nullptr;
    if (d->profiling && d->prof_n < d->prof_max) ev = &d->prof_ev[(size_t)4 * d->prof_n++];
    const int rc = ice::residual_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, (hipStream_t)stream, ev);
    if (rc != 0) return fail(ICE_ERR_HIP, std::string("residual launch failed: ") + hipGetErrorString((hipError_t)rc));
    return ICE_OK;
}

int ice_residual_mixed(ice_disc* d, const double* u, double* r, void* stream) {
    if (!d || !u || !r) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    if (s->ndim < 2) return fail(ICE_ERR_UNSUPPORTED, "the mixed-precision residual covers ndim = 2 and 3");
    hipEvent_t* ev = nullptr;
    if (d->profiling && d->prof_n < d->prof_max) ev = &d->prof_ev[(size_t)4 * d->prof_n++];
    const int rc = ice::residual_mixed(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, (hipStream_t)stream, ev);
    if (rc != 0) return fail(ICE_ERR_HIP, std::string("residual launch failed: ") + hipGetErrorString((hipError_t)rc));
    return ICE_OK;
}

int ice_mass_inverse_apply(ice_disc* d, const double* r, double* out, void* stream) {
    if (!d || !r || !out) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    const int rc = ice::launch_minv(s->ndim, s->p, d->neq, s->dev, r, out, (hipStream_t)stream);
    if (rc != 0) return fail(ICE_ERR_HIP, "mass inverse launch failed");
    return ICE_OK;
}

int ice_project(ice_disc* d, const ice_field* f, double* u, void* stream) {
    if (!d || !f || !u) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)s->mesh.nelem * s->nb * d->neq;
    double* b = nullptr;
    ice_field* df = nullptr;
    int rc = 0;
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess || hipMalloc(&df, sizeof(ice_field)) != hipSuccess) {
        if (b) (void)hipFree(b);
        return fail(ICE_ERR_ALLOC, "projection scratch allocation failed");
    }
    rc |= (int)hipMemcpyAsync(df, f, sizeof(ice_field), hipMemcpyHostToDevice, st);
    rc |= ice::projection_rhs(s->ndim, s->p, d->neq, s->dev, df, b, st);
    rc |= ice::launch_minv(s->ndim, s->p, d->neq, s->dev, b, u, st);
    rc |= (int)hipStreamSynchronize(st);
    (void)hipFree(b);
    (void)hipFree(df);
    if (rc != 0) return fail(ICE_ERR_HIP, "projection failed");
    return ICE_OK;
}

int ice_project_values(ice_disc* d, const double* vq, double* u, void* stream) {
    if (!d || !vq || !u) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)s->mesh.nelem * s->nb * d->neq;
    double* b = nullptr;
    if (hipMalloc(&b, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess)
        return fail(ICE_ERR_ALLOC, "projection scratch allocation failed");
    int rc = ice::load_from_values(s->ndim, s->p, d->neq, s->dev, vq, 0, b, st);
    if (rc == 0) rc = ice::launch_minv(s->ndim, s->p, d->neq, s->dev, b, u, st);
    if (rc == 0) rc = (int)hipStreamSynchronize(st);
    (void)hipFree(b);
    if (rc != 0) return fail(ICE_ERR_HIP, "projection failed");
    return ICE_OK;
}

int ice_disc_set_source_values(ice_disc* d, const double* vq, void* stream) {
    if (!d || !vq) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    hipStream_t st = (hipStream_t)stream;
    if (!d->d_srcload &&
        hipMalloc(&d->d_srcload, std::max<size_t>((size_t)s->mesh.nelem * s->nb * d->neq, 1) * sizeof(double)) !=
            hipSuccess) {
        d->d_srcload = nullptr;
        return fail(ICE_ERR_ALLOC, "source load allocation failed");
    }
    int rc = ice::load_from_values(s->ndim, s->p, d->neq, s->dev, vq, 1, d->d_srcload, st);
    if (rc == 0) rc = (int)hipStreamSynchronize(st);
    if (rc != 0) return fail(ICE_ERR_HIP, "source load failed");
    d->src_values = true;
    d->dd.srcload = d->d_srcload;
    return ICE_OK;
}

int ice_disc_set_boundary_values(ice_disc* d, const double* bv, void* stream) {
    if (!d || !bv) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    const size_t n = (size_t)s->dev.nfb * s->nfq * d->neq;
    if (n == 0) return ICE_OK;
    hipStream_t st = (hipStream_t)stream;
    int rc = (int)hipMemcpyAsync(d->d_bval, bv, n * sizeof(double), hipMemcpyDeviceToDevice, st);
    if (rc == 0) rc = (int)hipStreamSynchronize(st);
    if (rc != 0) return fail(ICE_ERR_HIP, "boundary value copy failed");
    d->bval_values = true;
    return ICE_OK;
}

int ice_space_bface_coords(const ice_space* s, double* xb, int* bctype, int* bcflag) {
    if (!s) return fail(ICE_ERR_ARG, "null argument");
    const int nfb = s->dev.nfb;
    if (bctype || bcflag)
        for (int b = 0; b < nfb; ++b) {
            if (bctype) bctype[b] = (s->fbdy[b].y >> 9) & 15;
            if (bcflag) bcflag[b] = s->fbdy[b].z;
        }
    if (!xb || nfb == 0) return ICE_OK;
    const size_t n = (size_t)nfb * s->nfq * s->ndim;
    double* dx = nullptr;
    if (hipMalloc(&dx, n * sizeof(double)) != hipSuccess) return fail(ICE_ERR_ALLOC, "allocation failed");
    int rc = ice::bface_coords(s->ndim, s->p, s->dev, dx, nullptr);
    if (rc == 0) rc = (int)hipMemcpy(xb, dx, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    if (rc != 0) return fail(ICE_ERR_HIP, "boundary point evaluation failed");
    return ICE_OK;
}

int ice_l2_error(ice_disc* d, const ice_field* f, const double* u, double* err, void* stream) {
    if (!d || !f || !u || !err) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    hipStream_t st = (hipStream_t)stream;
    double* terms = nullptr;
    ice_field* df = nullptr;
    if (hipMalloc(&terms, (size_t)s->mesh.nelem * sizeof(double)) != hipSuccess ||
        hipMalloc(&df, sizeof(ice_field)) != hipSuccess) {
        if (terms) (void)hipFree(terms);
        return fail(ICE_ERR_ALLOC, "l2 scratch allocation failed");
    }
    int rc = (int)hipMemcpyAsync(df, f, sizeof(ice_field), hipMemcpyHostToDevice, st);
    rc |= ice::l2_error_terms(s->ndim, s->p, d->neq, s->dev, df, u, terms, st);
    std::vector<double> h(s->mesh.nelem);
    rc |= (int)hipMemcpyAsync(h.data(), terms, h.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    rc |= (int)hipStreamSynchronize(st);
    (void)hipFree(terms);
    (void)hipFree(df);
    if (rc != 0) return fail(ICE_ERR_HIP, "l2 error failed");
    double acc = 0.0;
    for (double v : h) acc += v;
    *err = std::sqrt(acc);
    return ICE_OK;
}

int ice_jacobian_layout(const ice_disc* d, int* nslots, int* block_size, int* cols) {
    if (!d || !nslots || !block_size) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    *nslots = 2 * s->ndim + 1;
    *block_size = s->nb * d->neq;
    if (cols) std::memcpy(cols, s->jcols.data(), s->jcols.size() * sizeof(int));
    return ICE_OK;
}

int ice_jacobian(ice_disc* d, const double* u, double* r, double* blocks, void* stream) {
    if (!d || !u || !blocks) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    if (s->ndim > 2 && s->p > 2)
        return fail(ICE_ERR_UNSUPPORTED, "the device Jacobian covers P1 and P2 in 3D (per-face tables in LDS)");
    hipStream_t st = (hipStream_t)stream;
    if (r) {
        const int rc = ice::residual_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, st, nullptr);
        if (rc != 0) return fail(ICE_ERR_HIP, "residual launch failed");
    }
    const size_t bs = (size_t)s->nb * d->neq;
    const size_t nvals = (size_t)s->mesh.nelem * (2 * s->ndim + 1) * bs * bs;
    const int rc = ice::jacobian_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, blocks, nvals, st);
    if (rc != 0) return fail(ICE_ERR_HIP, std::string("jacobian launch failed: ") + hipGetErrorString((hipError_t)rc));
    return ICE_OK;
}

int ice_jacobian_apply(const ice_disc* d, const double* blocks, const double* x, double* y, void* stream) {
    if (!d || !blocks || !x || !y) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    const int rc = ice::bsr_apply<double, double>(s->mesh.nelem, 2 * s->ndim + 1, s->nb * d->neq, s->dev.jcols, blocks,
                                                  x, nullptr, y, (hipStream_t)stream);
    if (rc != 0) return fail(ICE_ERR_HIP, "block product launch failed");
    return ICE_OK;
}

int ice_profile_begin(ice_disc* d, int max_calls) {
    if (!d || max_calls < 1) return fail(ICE_ERR_ARG, "bad argument");
    try {
        while ((int)d->prof_ev.size() < 4 * max_calls) {
            hipEvent_t e;
            ICE_HIP(hipEventCreate(&e));
            d->prof_ev.push_back(e);
        }
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
    d->prof_max = max_calls;
    d->prof_n = 0;
    d->profiling = true;
    return ICE_OK;
}

int ice_profile_end(ice_disc* d, double* stage_ms, int* ncalls) {
    if (!d || !stage_ms || !ncalls) return fail(ICE_ERR_ARG, "null argument");
    if (!d->profiling) return fail(ICE_ERR_STATE, "profiling not active");
    d->profiling = false;
    for (int k = 0; k < 3; ++k) stage_ms[k] = 0.0;
    for (int c = 0; c < d->prof_n; ++c) {
        hipEvent_t* ev = &d->prof_ev[(size_t)4 * c];
        if (hipEventSynchronize(ev[3]) != hipSuccess) return fail(ICE_ERR_HIP, "event synchronize failed");
        float ms0 = 0.f, ms1 = 0.f, ms2 = 0.f;
        if (hipEventElapsedTime(&ms0, ev[0], ev[1]) != hipSuccess || hipEventElapsedTime(&ms1, ev[1], ev[2]) != hipSuccess ||
            hipEventElapsedTime(&ms2, ev[2], ev[3]) != hipSuccess)
            return fail(ICE_ERR_HIP, "event timing failed");
        stage_ms[0] += ms0;
        stage_ms[1] += ms1;
        stage_ms[2] += ms2;
    }
    *ncalls = d->prof_n;
    return ICE_OK;
}

int ice_field_eval_points(const ice_field* f, int ndim, int neq, int npts, const double* x, double* out) {
    if (!f || !x || !out || ndim < 1 || ndim > 3 || neq < 1 || neq > 5) return fail(ICE_ERR_ARG, "bad argument");
    for (int i = 0; i < npts; ++i) ice_field_eval(f, ndim, neq, x + (size_t)i * ndim, out + (size_t)i * neq);
    return ICE_OK;
}

int ice_space_node_coords(const ice_space* s, double* xn) {
    if (!s || !xn) return fail(ICE_ERR_ARG, "null argument");
    const int D = s->ndim, n = s->n, p = s->p, NV = 1 << D;
    for (int e = 0; e < s->mesh.nelem; ++e)
        for (int i = 0; i < s->nb; ++i) {
            int r = i, id[3];
            for (int d = D - 1; d >= 0; --d) { id[d] = r % n; r /= n; }
            double xi[3];
            for (int d = 0; d < D; ++d) {
                double x = -1.0;  // equispaced nodes by cumulative steps (lagrange_1d.hpp:40-45)
                for (int j = 0; j < id[d]; ++j) x += 2.0 / p;
                xi[d] = x;
            }
            double x[3] = {0, 0, 0};
            for (int v = 0; v < NV; ++v) {
                double phi = 1.0;
                for (int d = 0; d < D; ++d) {
                    const double n0 = 0.5 * (1 - xi[d]);
                    phi *= ((v >> (D - 1 - d)) & 1) ? (1.0 - n0) : n0;
                }
                for (int d = 0; d < D; ++d) x[d] += phi * s->X[((size_t)e * NV + v) * D + d];
            }
            for (int d = 0; d < D; ++d) xn[((size_t)e * s->nb + i) * D + d] = x[d];
        }
    return ICE_OK;
}

int ice_trace_points(int ndim, int order, int face_nr, int orientation, double* xi) {
    if (ndim < 1 || ndim > 3 || order < 1 || order > ICE_MAX_ORDER || face_nr < 0 || face_nr >= 2 * ndim || !xi ||
        orientation < 0 || orientation >= ice::n_orientations(ndim))
        return fail(ICE_ERR_ARG, "bad argument");
    const int n = order + 1;
    double gx[8], gw[8];
    gauss_legendre(n, gx, gw);
    int nfq = 1;
    for (int d = 0; d < ndim - 1; ++d) nfq *= n;
    for (int q = 0; q < nfq; ++q) {
        int pt[3];
        ice::trace_point_indices(ndim, n, face_nr, orientation, q, pt);
        for (int d = 0; d < ndim; ++d) xi[q * ndim + d] = (pt[d] < n) ? gx[pt[d]] : (pt[d] == n ? -1.0 : 1.0);
    }
    return ICE_OK;
}

}  // extern "C"

// ------------------------------------------------------------ explicit steppers
struct ice_stepper {
    bool tracks_wave = false;
    ice_disc* disc = nullptr;
    ice_stepper_desc desc{};
    ice::StepControl ctl{};
    size_t n = 0;
    double* buf[5] = {};      // res, k1, k2, k3, ustage / uold
    float* fbuf[4] = {};      // FP32 states: k1, k2, k3, ustage (allocated by the first FP32 step)
    ice::StepState* st = nullptr;
    double* dte = nullptr;    // per-element step sizes (batched instances), or null
    ~ice_stepper() {
        if (tracks_wave && disc && --disc->wave_users == 0) disc->dd.wave = nullptr;
        for (double* b : buf)
            if (b) (void)hipFree(b);
        for (float* b : fbuf)
            if (b) (void)hipFree(b);
        if (st) (void)hipFree(st);
        if (dte) (void)hipFree(dte);
    }
};

namespace {

int step_once(ice_stepper* S, double* u, hipStream_t st) {
    ice_disc* d = S->disc;
    const ice_space* sp = d->space;
    const size_t n = S->n;
    double *res = S->buf[0], *k1 = S->buf[1], *k2 = S->buf[2], *k3 = S->buf[3], *us = S->buf[4];
    const double* dt = &S->st->dt;
    const size_t per = (size_t)sp->nb * d->neq;
    int rc = 0;
    auto R = [&](const double* x) {
        rc |= ice::residual_fp64(sp->ndim, sp->p, d->prm.flux, sp->dev, d->ph, d->dd, x, res, st, nullptr);
    };
    auto Minv = [&](double* k) { rc |= ice::launch_minv(sp->ndim, sp->p, d->neq, sp->dev, res, k, st); };
    auto copy = [&](const double* x, double* y) {
        rc |= (int)hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, st);
    };
    // y += c dt x with the stepper's step size, or each element's own
    auto axpy = [&](double c, const double* x, double* y) {
        if (S->dte) rc |= ice::launch_axpy_edt(n, c, S->dte, per, S->st, x, y, st);
        else rc |= ice::launch_axpy_dt(n, c, dt, x, y, st);
    };
    rc |= ice::launch_step_begin(S->st, d->d_wave, S->ctl, st);
    switch (S->desc.kind) {
    case ICE_STEP_RK3_SSP:
        if (S->desc.reference_compat) {
            // ssp_rk3.hpp:132-190 with every stage residual evaluated at u^n
            R(u);
            Minv(k1);
            axpy(1.0 / 6.0, k1, u);
            axpy(1.0 / 6.0, k1, u);
            axpy(2.0 / 3.0, k1, u);
        } else {
            // ssp_rk3.hpp:164-190; each stage's increment and combination
            // u1 = u + dt k1, u2 = u + dt (k1 + k2)/4, u = u + dt (k1/6 + k2/6 + 2 k3/3)
            // in one pass over the state
            auto minv_combine = [&](double* k, int nprev, const double* const* kk, const double* c, double* out,
                                    bool store_k) {
                rc |= ice::launch_minv_combine(sp->ndim, sp->p, d->neq, sp->dev, res, k, u, nprev, kk, c, dt, S->dte,
                                               S->st, out, store_k, st);
            };
            const double* k12[2] = {k1, k2};
            const double c1[1] = {1.0}, c2[2] = {0.25, 0.25}, c3[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
            R(u);
            minv_combine(k1, 0, k12, c1, us, true);
            R(us);
            minv_combine(k2, 1, k12, c2, us, true);
            R(us);
            minv_combine(k3, 2, k12, c3, u, false);
        }
        break;
    case ICE_STEP_RK3_TVD: {
        // tvd_rk3.hpp:117-226: Shu-Osher stages in place
        static const double alfa[3] = {0.0, 0.75, 1.0 / 3.0};
        static const double beta[3] = {1.0, 0.25, 2.0 / 3.0};
        copy(u, us);
        for (int s = 0; s < 3; ++s) {
            R(u);
            Minv(k1);
            rc |= ice::launch_axpby(n, alfa[s], us, beta[s], u, st);
            axpy(beta[s], k1, u);
        }
        break;
    }
    default:  // ExplicitEuler (explicit_euler.hpp:87-138)
        R(u);
        Minv(k1);
        axpy(1.0, k1, u);
        break;
    }
    rc |= ice::launch_step_end(S->st, st);
    return rc;
}

// FP32 state: the stage residuals are the mixed-precision residual read from
// the FP32 coefficients (FP64 output), the stage increments M^-1 R are kept
// in FP32 and every stage combination is formed in FP64 and rounded once.
// Textbook SSP-RK3 and explicit Euler.
int step_once(ice_stepper* S, float* u, hipStream_t st) {
    ice_disc* d = S->disc;
    const ice_space* sp = d->space;
    double* res = S->buf[0];
    float *k1 = S->fbuf[0], *k2 = S->fbuf[1], *k3 = S->fbuf[2], *us = S->fbuf[3];
    const double* dt = &S->st->dt;
    int rc = 0;
    auto R = [&](const float* x) {
        rc |= ice::residual_mixed_f32state(sp->ndim, sp->p, d->prm.flux, sp->dev, d->ph, d->dd, x, res, st);
    };
    // each stage's increment and combination in one pass over the state
    auto minv_combine = [&](float* k, int nprev, const float* const* kk, const double* c, float* out, bool store_k) {
        rc |= ice::launch_minv_combine_f32(sp->ndim, sp->p, d->neq, sp->dev, res, k, u, nprev, kk, c, dt, S->dte,
                                           S->st, out, store_k, st);
    };
    const float* kk[2] = {k1, k2};
    const double c1[1] = {1.0}, c2[2] = {0.25, 0.25}, c3[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
    rc |= ice::launch_step_begin(S->st, d->d_wave, S->ctl, st);
    R(u);
    if (S->desc.kind == ICE_STEP_RK3_SSP) {
        minv_combine(k1, 0, kk, c1, us, true);
        R(us);
        minv_combine(k2, 1, kk, c2, us, true);
        R(us);
        minv_combine(k3, 2, kk, c3, u, false);
    } else {
        minv_combine(k1, 0, kk, c1, u, false);
    }
    rc |= ice::launch_step_end(S->st, st);
    return rc;
}

// FP32-state steps: the stepper kinds and spaces they cover, and the FP32
// buffers, allocated once.
int prepare_f32(ice_stepper* S) {
    const ice_space* sp = S->disc->space;
    if (sp->ndim < 2) return fail(ICE_ERR_UNSUPPORTED, "FP32 states cover ndim = 2 and 3");
    const bool ok = (S->desc.kind == ICE_STEP_RK3_SSP && !S->desc.reference_compat) ||
                    S->desc.kind == ICE_STEP_EXPLICIT_EULER;
    if (!ok) return fail(ICE_ERR_UNSUPPORTED, "FP32 states step with textbook RK3-SSP or explicit Euler");
    if (!S->fbuf[0]) {
        try {
            for (float*& b : S->fbuf) ICE_HIP(hipMalloc(&b, std::max<size_t>(S->n, 1) * sizeof(float)));
        } catch (const ice::Error& e) {
            for (float*& b : S->fbuf) {
                if (b) (void)hipFree(b);
                b = nullptr;
            }
            return fail(e.code, e.msg);
        }
    }
    return ICE_OK;
}

}  // namespace

extern "C" {

int ice_stepper_create(ice_disc* d, const ice_stepper_desc* desc, ice_stepper** out) {
    if (!d || !desc || !out) return fail(ICE_ERR_ARG, "null argument");
    *out = nullptr;
    if (desc->kind < ICE_STEP_RK3_SSP || desc->kind > ICE_STEP_EXPLICIT_EULER) return fail(ICE_ERR_ARG, "bad stepper kind");
    if (!desc->use_cfl && !(desc->dt > 0)) return fail(ICE_ERR_ARG, "fixed time step must be positive");
    if (desc->use_cfl && d->d_emu) return fail(ICE_ERR_UNSUPPORTED, "CFL steps with per-element viscosity");
    try {
        auto S = std::make_unique<ice_stepper>();
        S->disc = d;
        S->desc = *desc;
        const ice_space* sp = d->space;
        S->n = (size_t)sp->mesh.nelem * sp->nb * d->neq;
        for (double*& b : S->buf) ICE_HIP(hipMalloc(&b, S->n * sizeof(double)));
        ICE_HIP(hipMalloc(&S->st, sizeof(ice::StepState)));
        ice::StepState zero{0.0, 0.0, 0, 0};
        ICE_HIP(hipMemcpy(S->st, &zero, sizeof(zero), hipMemcpyHostToDevice));
        // reference length: min over elements of det(J(centroid))^(1/ndim) (explicit_utils.hpp:70-83)
        const int D = sp->ndim, NV = 1 << D;
        double reflen = 1e8;
        for (int e = 0; e < sp->mesh.nelem; ++e) {
            double J[3][3] = {};
            for (int k = 0; k < NV; ++k)
                for (int a = 0; a < D; ++a) {
                    double t = 1.0;
                    for (int dd = 0; dd < D; ++dd) {
                        const int bit = (k >> (D - 1 - dd)) & 1;
                        t *= (dd == a) ? (bit ? 0.5 : -0.5) : 0.5;
                    }
                    for (int i = 0; i < D; ++i) J[i][a] += t * sp->X[((size_t)e * NV + k) * D + i];
                }
            double det = (D == 1) ? J[0][0]
                         : (D == 2) ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                    : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                          J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            reflen = std::min(reflen, std::pow(det, 1.0 / D));
        }
        ice::StepControl& c = S->ctl;
        c.use_cfl = desc->use_cfl;
        c.dt_fixed = desc->dt;
        c.cfl = desc->cfl;
        c.reflen = reflen;
        c.mu = d->prm.mu;
        c.pmax = std::max(1, sp->p);
        c.flux = d->prm.flux;
        c.use_tfinal = desc->use_tfinal;
        c.tfinal = desc->tfinal;
        c.ntime = desc->ntime;
        if (desc->use_cfl) {
            S->tracks_wave = true;
            d->wave_users++;
            d->dd.wave = d->d_wave;
        }
        *out = S.release();
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
}

int ice_stepper_set_element_dt(ice_stepper* s, const double* dt) {
    if (!s) return fail(ICE_ERR_ARG, "null argument");
    if (s->ctl.use_cfl) return fail(ICE_ERR_STATE, "a CFL stepper chooses its own step");
    const size_t ne = (size_t)s->disc->space->mesh.nelem;
    try {
        if (!dt) {
            if (s->dte) ICE_HIP(hipFree(s->dte));
            s->dte = nullptr;
            return ICE_OK;
        }
        for (size_t e = 0; e < ne; ++e)
            if (!(dt[e] > 0)) return fail(ICE_ERR_ARG, "element step sizes must be positive");
        if (!s->dte) ICE_HIP(hipMalloc(&s->dte, std::max<size_t>(ne, 1) * sizeof(double)));
        ICE_HIP(hipMemcpy(s->dte, dt, ne * sizeof(double), hipMemcpyHostToDevice));
        return ICE_OK;
    } catch (const ice::Error& e) {
        return fail(e.code, e.msg);
    }
}

// ---- step graphs -------------------------------------------------------
struct ice_step_graph_impl {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // extern "C"

namespace {

// Capture nsteps steps of every stepper into one graph.  With concurrent set
// and every stepper on its own discretization and state, each stepper's
// steps form a branch of their own (forked from and joined to the capture
// stream by events), so the graph may run the branches side by side;
// otherwise all steps are one chain in stepper order.
template <class UT>
int make_step_graph(ice_stepper* const* steppers, UT* const* u, int n, int nsteps, int concurrent,
                    ice_step_graph** out) {
    if (!steppers || !u || !out || n < 1 || nsteps < 1) return fail(ICE_ERR_ARG, "bad argument");
    *out = nullptr;
    for (int k = 0; k < n; ++k) {
        if (!steppers[k] || !u[k]) return fail(ICE_ERR_ARG, "null stepper or state");
        if constexpr (std::is_same<UT, float>::value) {
            const int rc = prepare_f32(steppers[k]);
            if (rc != ICE_OK) return rc;
        }
    }
    bool branches = concurrent && n > 1;
    for (int a = 0; a < n && branches; ++a)
        for (int b = a + 1; b < n && branches; ++b)
            if (steppers[a]->disc == steppers[b]->disc || u[a] == u[b] || steppers[a] == steppers[b])
                branches = false;
    std::vector<hipStream_t> streams(branches ? n + 1 : 1, nullptr);
    std::vector<hipEvent_t> events(branches ? n + 1 : 0, nullptr);
    auto release = [&]() {
        for (hipEvent_t e : events)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t s : streams)
            if (s) (void)hipStreamDestroy(s);
    };
    for (hipStream_t& s : streams)
        if (hipStreamCreate(&s) != hipSuccess) {
            release();
            return fail(ICE_ERR_HIP, "stream create failed");
        }
    for (hipEvent_t& e : events)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            release();
            return fail(ICE_ERR_HIP, "event create failed");
        }
    hipStream_t cs = streams[0];
    auto g = std::make_unique<ice_step_graph_impl>();
    int rc = 0;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        release();
        return fail(ICE_ERR_HIP, "stream capture failed to start");
    }
    if (branches) {
        rc |= (int)hipEventRecord(events[0], cs);
        for (int k = 0; k < n; ++k) {
            hipStream_t bs = streams[k + 1];
            rc |= (int)hipStreamWaitEvent(bs, events[0], 0);
            for (int t = 0; t < nsteps; ++t) rc |= step_once(steppers[k], u[k], bs);
            rc |= (int)hipEventRecord(events[k + 1], bs);
            rc |= (int)hipStreamWaitEvent(cs, events[k + 1], 0);
        }
    } else {
        for (int t = 0; t < nsteps; ++t)
            for (int k = 0; k < n; ++k) rc |= step_once(steppers[k], u[k], cs);
    }
    const hipError_t ec = hipStreamEndCapture(cs, &g->graph);
    release();
    if (rc != 0 || ec != hipSuccess || !g->graph) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        return fail(ICE_ERR_HIP, "step capture failed");
    }
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g->graph);
        return fail(ICE_ERR_HIP, "graph instantiation failed");
    }
    // upload the executable graph now rather than on its first launch (an
    // optimisation only: a failure leaves the upload to the launch)
    if (hipGraphUpload(g->exec, nullptr) == hipSuccess) (void)hipStreamSynchronize(nullptr);
    else (void)hipGetLastError();
    *out = reinterpret_cast<ice_step_graph*>(g.release());
    return ICE_OK;
}

}  // namespace

extern "C" {

int ice_step_graph_create(ice_stepper* const* steppers, double* const* u, int n, int nsteps, ice_step_graph** out) {
    return make_step_graph(steppers, u, n, nsteps, 0, out);
}

int ice_step_graph_create_ex(ice_stepper* const* steppers, void* const* u, int state_fp32, int n, int nsteps,
                             int concurrent, ice_step_graph** out) {
    if (!u) return fail(ICE_ERR_ARG, "bad argument");
    if (state_fp32)
        return make_step_graph(steppers, reinterpret_cast<float* const*>(u), n, nsteps, concurrent, out);
    return make_step_graph(steppers, reinterpret_cast<double* const*>(u), n, nsteps, concurrent, out);
}

int ice_step_graph_launch(ice_step_graph* g, void* stream) {
    if (!g) return fail(ICE_ERR_ARG, "null argument");
    auto* G = reinterpret_cast<ice_step_graph_impl*>(g);
    if (hipGraphLaunch(G->exec, (hipStream_t)stream) != hipSuccess) return fail(ICE_ERR_HIP, "graph launch failed");
    return ICE_OK;
}

int ice_step_graph_destroy(ice_step_graph* g) {
    if (!g) return ICE_OK;
    auto* G = reinterpret_cast<ice_step_graph_impl*>(g);
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    delete G;
    return ICE_OK;
}

int ice_stepper_destroy(ice_stepper* s) {
    delete s;
    return ICE_OK;
}

int ice_stepper_step(ice_stepper* s, double* u, void* stream) {
    if (!s || !u) return fail(ICE_ERR_ARG, "null argument");
    if (step_once(s, u, (hipStream_t)stream) != 0) return fail(ICE_ERR_HIP, "step launch failed");
    return ICE_OK;
}

int ice_stepper_step_f32(ice_stepper* s, float* u, void* stream) {
    if (!s || !u) return fail(ICE_ERR_ARG, "null argument");
    const int rc = prepare_f32(s);
    if (rc != ICE_OK) return rc;
    if (step_once(s, u, (hipStream_t)stream) != 0) return fail(ICE_ERR_HIP, "step launch failed");
    return ICE_OK;
}

int ice_residual_f32state(ice_disc* d, const float* u, double* r, int check_fp64, void* stream) {
    if (!d || !u || !r) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = d->space;
    if (s->ndim < 2) return fail(ICE_ERR_UNSUPPORTED, "FP32 states cover ndim = 2 and 3");
    hipStream_t st = (hipStream_t)stream;
    int rc = 0;
    if (check_fp64) {
        const size_t n = (size_t)s->mesh.nelem * s->nb * d->neq;
        if (!d->d_u64) {
            if (hipMalloc(&d->d_u64, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) {
                d->d_u64 = nullptr;
                return fail(ICE_ERR_HIP, "allocation of the FP64 copy of the state failed");
            }
        }
        rc = ice::launch_widen(n, u, d->d_u64, st);
        if (rc == 0) rc = ice::residual_fp64(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, d->d_u64, r, st, nullptr);
    } else {
        rc = ice::residual_mixed_f32state(s->ndim, s->p, d->prm.flux, s->dev, d->ph, d->dd, u, r, st);
    }
    if (rc != 0) return fail(ICE_ERR_HIP, std::string("residual launch failed: ") + hipGetErrorString((hipError_t)rc));
    return ICE_OK;
}

int ice_stepper_status(ice_stepper* s, double* time, double* dt, long long* itime, int* done) {
    if (!s) return fail(ICE_ERR_ARG, "null argument");
    ice::StepState h{};
    if (hipMemcpy(&h, s->st, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return fail(ICE_ERR_HIP, "status copy failed");
    if (time) *time = h.time;
    if (dt) *dt = h.dt;
    if (itime) *itime = h.itime;
    if (done) {
        const ice::StepControl& c = s->ctl;
        *done = c.use_tfinal ? (h.time >= c.tfinal) : (h.itime >= c.ntime);
    }
    return ICE_OK;
}

int ice_stepper_solve(ice_stepper* s, double* u, void* stream, long long* nsteps, double* time) {
    if (!s || !u) return fail(ICE_ERR_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    ice_disc* d = s->disc;
    const ice_space* sp = d->space;
    // RK3SSP::solve / RK3TVD::solve: one residual first for the wave speeds
    if (ice::residual_fp64(sp->ndim, sp->p, d->prm.flux, sp->dev, d->ph, d->dd, u, s->buf[0], st, nullptr) != 0)
        return fail(ICE_ERR_HIP, "residual launch failed");
    const long long chunk = s->ctl.use_tfinal ? 32 : std::max<long long>(1, s->ctl.ntime);
    for (;;) {
        for (long long k = 0; k < chunk; ++k)
            if (step_once(s, u, st) != 0) return fail(ICE_ERR_HIP, "step launch failed");
        if (hipStreamSynchronize(st) != hipSuccess) return fail(ICE_ERR_HIP, "stream synchronize failed");
        int done = 0;
        long long it = 0;
        double t = 0;
        int rc = ice_stepper_status(s, &t, nullptr, &it, &done);
        if (rc != ICE_OK) return rc;
        if (done || (!s->ctl.use_tfinal)) {
            if (nsteps) *nsteps = it;
            if (time) *time = t;
            return ICE_OK;
        }
    }
}

}  // extern "C"
