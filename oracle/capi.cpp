// CPU ORACLE (test infrastructure only; never linked into the product).
//
// Flat C entry points over the oracle templates so that tests/ (ctypes),
// __graft_entry__.smoke() and bench.py's CPU baseline can drive the oracle
// with the same descriptors (include/iceicle_amd.h) as the product.
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "mdg.hpp"
#include "parallel.hpp"
#include "solvers.hpp"
#include "../iceicle_amd/csrc/fields.h"

using namespace oracle;

namespace {

thread_local std::string g_err;

struct ProblemBase {
    virtual ~ProblemBase() = default;
    virtual void info(int* out) const = 0;
    virtual void residual(const double* u, double* r) = 0;
    virtual void residual_parallel(const double* u, double* r, int nthreads) = 0;
    virtual void faces(int* recs) const = 0;
    virtual void mesh_arrays(double* coord, int* conn) const = 0;
    virtual void project(const ice_field* f, double* u) const = 0;
    virtual double l2_error(const ice_field* f, const double* u) const = 0;
    virtual void rk3_ssp(double* u, double dt, int nsteps, int mode) = 0;
    virtual int rk3_tvd_tfinal(double* u, double dt, double tfinal) = 0;
    virtual double cfl_dt(double cfl) = 0;
    virtual void trace_points(double* xiL, double* xiR) const = 0;
    virtual void face_geometry(double* x, double* nrm, double* sqrtg) const = 0;
    virtual void qp_coords(double* xq) const = 0;
    virtual void mass_apply(const double* r, double* out) const = 0;
    virtual void rk_solve(double* u, int kind, int compat, int use_cfl, double dt, double cfl, int use_tfinal,
                          double tfinal, long long ntime, long long* nsteps, double* time) = 0;
    // blocks in BSR slot order: per row e, slot 0 = (e,e), slot 1+j = neighbour across local face j
    virtual void jacobian(const double* u, double* res, int* cols, double* blocks) = 0;
    virtual int newton(double* u, double tau_abs, double tau_rel, int kmax, double* rhist, int nhist) = 0;
    // MDG interface conservation on selected traces: the selected nodes
    // (count in *nsel, ids in nodes when non-null), and the residual
    virtual void mdg_nodeset(const int* traces, int ntr, int* nsel, int* nodes) const = 0;
    virtual void mdg_residual(const int* traces, int ntr, const double* u, double* res) const = 0;
};

template <int D, class PF, class CF, class DF>
struct Problem final : ProblemBase {
    using Disc = ConservationLawDDG<D, PF, CF, DF>;
    static constexpr int neq = PF::neq;
    FESpace<D> fs;
    Disc disc;
    std::unique_ptr<MassInverse<D>> minv;
    ice_disc_params prm;

    Problem(FESpace<D>&& space, Disc&& d, const ice_disc_params& p) : fs(std::move(space)), disc(std::move(d)), prm(p) {
        disc.interior_penalty = p.interior_penalty != 0;
        disc.sigma_ic = p.sigma_ic;
        disc.ddgic_published = p.ddgic_form == ICE_DDGIC_PUBLISHED;
        for (int k = 0; k < ICE_MAX_BCFLAG; ++k) {
            ice_field fd = p.dirichlet[k], fn = p.neumann[k];
            disc.dirichlet_callbacks.push_back([fd](const double* x, double* out) { ice_field_eval(&fd, D, neq, x, out); });
            disc.neumann_callbacks.push_back([fn](const double* x, double* out) { ice_field_eval(&fn, D, neq, x, out); });
        }
        if (p.source.kind != ICE_FIELD_NONE) {
            ice_field fsrc = p.source;
            disc.user_source = [fsrc](const double* x, double* out) { ice_field_eval(&fsrc, D, neq, x, out); };
        }
    }
    void info(int* out) const override {
        out[0] = D; out[1] = fs.mesh.nelem; out[2] = fs.nb; out[3] = neq;
        out[4] = (int)fs.mesh.faces.size(); out[5] = fs.mesh.n_interior; out[6] = fs.nfq; out[7] = fs.nq;
        out[8] = fs.mesh.nnode; out[9] = (int)fs.periodic.size();
    }
    void residual(const double* u, double* r) override { form_residual(fs, disc, u, r); }
    ParallelPlan<D> plan;
    std::vector<double> fbuf;
    void residual_parallel(const double* u, double* r, int nthreads) override {
        form_residual_parallel(fs, disc, u, r, plan, fbuf, nthreads);
    }
    void faces(int* recs) const override {
        constexpr int NFV = Mesh<D>::NFV;
        int k = 0;
        for (const auto& f : fs.mesh.faces) {
            recs[k++] = f.elemL; recs[k++] = f.elemR;
            recs[k++] = f.fnl * 512; recs[k++] = f.fnr * 512 + f.orient;
            recs[k++] = f.bctype; recs[k++] = f.bcflag;
            for (int j = 0; j < NFV; ++j) recs[k++] = f.nodes[j];
        }
    }
    void mesh_arrays(double* coord, int* conn) const override {
        if (coord)
            for (int n = 0; n < fs.mesh.nnode; ++n)
                for (int d = 0; d < D; ++d) coord[n * D + d] = fs.mesh.coord[n][d];
        if (conn)
            for (int e = 0; e < fs.mesh.nelem; ++e)
                for (int k = 0; k < (1 << D); ++k) conn[e * (1 << D) + k] = fs.mesh.conn[e][k];
    }
    void project(const ice_field* f, double* u) const override {
        ice_field ff = *f;
        oracle::project<D>(fs, neq, [ff](const double* x, double* out) { ice_field_eval(&ff, D, neq, x, out); }, u);
    }
    double l2_error(const ice_field* f, const double* u) const override {
        ice_field ff = *f;
        return oracle::l2_error<D>(fs, neq, [ff](const double* x, double* out) { ice_field_eval(&ff, D, neq, x, out); }, u);
    }
    MassInverse<D>& mi() {
        if (!minv) minv = std::make_unique<MassInverse<D>>(fs);
        return *minv;
    }
    void rk3_ssp(double* u, double dt, int nsteps, int mode) override {
        for (int s = 0; s < nsteps; ++s) rk3_ssp_step(fs, disc, mi(), dt, mode == 0 ? RKMode::REFERENCE_COMPAT : RKMode::TEXTBOOK, u);
    }
    // RK3TVD::solve with FixedTimestep and TfinalTermination (tvd_rk3.hpp:240-270,
    // explicit_utils.hpp:148-160): an initial residual, then steps until t >= tfinal
    int rk3_tvd_tfinal(double* u, double dt, double tfinal) override {
        double t = 0;
        int it = 0;
        while (!(t >= tfinal)) {
            const double dts = std::min(dt, tfinal - t);
            rk3_tvd_step(fs, disc, mi(), dts, u);
            t += dts;
            ++it;
        }
        return it;
    }
    double cfl_dt(double cfl) override { return cfl_timestep(fs, disc, cfl); }
    void trace_points(double* xiL, double* xiR) const override {
        size_t k = 0;
        for (const auto* list : {&fs.interior, &fs.boundary})
            for (const auto& t : *list)
                for (int q = 0; q < fs.nfq; ++q) {
                    for (int d = 0; d < D; ++d) { xiL[k * D + d] = t.xiL[q * D + d]; xiR[k * D + d] = t.xiR[q * D + d]; }
                    ++k;
                }
    }
    void face_geometry(double* x, double* nrm, double* sqrtg) const override {
        size_t k = 0;
        for (const auto* list : {&fs.interior, &fs.boundary})
            for (const auto& t : *list)
                for (int q = 0; q < fs.nfq; ++q) {
                    double xx[D], Jf[D][(D > 1 ? D - 1 : 1)], nn[D];
                    fs.face_geo(t, q, xx, Jf);
                    unit_normal<D>(Jf, nn);
                    for (int d = 0; d < D; ++d) { x[k * D + d] = xx[d]; nrm[k * D + d] = nn[d]; }
                    sqrtg[k] = root_riemann<D>(Jf);
                    ++k;
                }
    }
    void qp_coords(double* xq) const override {
        for (int e = 0; e < fs.mesh.nelem; ++e) {
            const auto g = fs.geo(e);
            for (int q = 0; q < fs.nq; ++q) {
                double x[D], J[D][D], H[D][D][D];
                g.eval(fs.q1, fs.quad.pts[q].data(), x, J, H);
                for (int d = 0; d < D; ++d) xq[((size_t)e * fs.nq + q) * D + d] = x[d];
            }
        }
    }
    void mass_apply(const double* r, double* out) const override {
        const_cast<Problem*>(this)->mi().apply_add(fs, neq, r, out);
    }
    void mdg_nodeset(const int* traces, int ntr, int* nsel, int* nodes) const override {
        NodeSet<D> ns(fs, std::vector<int>(traces, traces + ntr));
        *nsel = (int)ns.selected.size();
        if (nodes) std::copy(ns.selected.begin(), ns.selected.end(), nodes);
    }
    void mdg_residual(const int* traces, int ntr, const double* u, double* res) const override {
        NodeSet<D> ns(fs, std::vector<int>(traces, traces + ntr));
        form_mdg_residual(fs, disc, ns, u, res);
    }
    int newton(double* u, double tau_abs, double tau_rel, int kmax, double* rhist, int nhist) override {
        return newton_dense(fs, disc, u, tau_abs, tau_rel, kmax, rhist, nhist);
    }
    void jacobian(const double* u, double* res, int* cols, double* blocks) override {
        std::map<std::pair<int, int>, std::vector<double>> B;
        fd_jacobian(fs, disc, u, res, B);
        const int E = fs.mesh.nelem, S = 2 * D + 1, bs = fs.nb * neq;
        std::vector<int> nbr((size_t)E * S, -1);
        for (int e = 0; e < E; ++e) nbr[(size_t)e * S] = e;
        for (const auto* list : {&fs.interior, &fs.periodic})
            for (const auto& t : *list) {
                nbr[(size_t)t.elemL * S + 1 + t.fnl] = t.elemR;
                nbr[(size_t)t.elemR * S + 1 + t.fnr] = t.elemL;
            }
        // the (row, column) block goes to the first slot naming that column:
        // where two faces of an element meet the same neighbour (a periodic
        // direction two elements wide, or one element wide for the diagonal)
        // the later slots hold zeros, so the slots of a row sum to the row's
        // blocks and the product over slots is the matrix product
        for (int e = 0; e < E; ++e)
            for (int sl = 0; sl < S; ++sl) {
                const int c = nbr[(size_t)e * S + sl];
                cols[(size_t)e * S + sl] = c;
                bool first = true;
                for (int s0 = 0; s0 < sl; ++s0) first = first && nbr[(size_t)e * S + s0] != c;
                double* dst = blocks + ((size_t)e * S + sl) * bs * bs;
                auto it = (c >= 0 && first) ? B.find({e, c}) : B.end();
                for (int k = 0; k < bs * bs; ++k) dst[k] = (it != B.end()) ? it->second[k] : 0.0;
            }
    }
    // RK3SSP/RK3TVD/ExplicitEuler ::solve (ssp_rk3.hpp:208-233, tvd_rk3.hpp:240-270,
    // explicit_euler.hpp:150-170): an initial residual, then steps until the
    // termination condition holds; dt from FixedTimestep or CFLTimestep,
    // limited by TfinalTermination::limit_dt
    void rk_solve(double* u, int kind, int compat, int use_cfl, double dt, double cfl, int use_tfinal, double tfinal,
                  long long ntime, long long* nsteps, double* time) override {
        std::vector<double> r((size_t)fs.mesh.nelem * fs.nb * neq);
        form_residual(fs, disc, u, r.data());
        double t = 0;
        long long it = 0;
        while (!(use_tfinal ? t >= tfinal : it >= ntime)) {
            double dtk = use_cfl ? cfl_timestep(fs, disc, cfl) : dt;
            if (use_tfinal) dtk = std::min(dtk, tfinal - t);
            if (kind == 0) rk3_ssp_step(fs, disc, mi(), dtk, compat ? RKMode::REFERENCE_COMPAT : RKMode::TEXTBOOK, u);
            else if (kind == 1) rk3_tvd_step(fs, disc, mi(), dtk, u);
            else explicit_euler_step(fs, disc, mi(), dtk, u);
            t += dtk;
            ++it;
        }
        if (nsteps) *nsteps = it;
        if (time) *time = t;
    }
};

template <int D>
std::unique_ptr<ProblemBase> make(const ice_mesh_desc* m, const ice_space_desc* s, const ice_disc_params* p, bool fix3d,
                                  BurgersCoefficients* bc, NSPhysics<D>* ph) {
    auto mesh = build_uniform_mesh<D>(m->nelem, m->xmin, m->xmax, m->bctype, m->bcflag, m->coord, fix3d);
    FESpace<D> fs(std::move(mesh), s->order, m->periodic_mode == ICE_PERIODIC_COUPLED);
    if (p->flux == ICE_FLUX_BURGERS) {
        using Dsc = ConservationLawDDG<D, BurgersFlux<D>, BurgersUpwind<D>, BurgersDiffusionFlux<D>>;
        Dsc d{BurgersFlux<D>{bc}, BurgersUpwind<D>{bc}, BurgersDiffusionFlux<D>{bc}};
        return std::make_unique<Problem<D, BurgersFlux<D>, BurgersUpwind<D>, BurgersDiffusionFlux<D>>>(std::move(fs), std::move(d), *p);
    } else if (p->flux == ICE_FLUX_NAVIER_STOKES) {
        using Dsc = ConservationLawDDG<D, NSFlux<D, true>, VanLeer<D>, NSDiffusionFlux<D, true>>;
        Dsc d{NSFlux<D, true>{ph}, VanLeer<D>{ph}, NSDiffusionFlux<D, true>{ph, p->ns_face_energy_textbook != 0}};
        return std::make_unique<Problem<D, NSFlux<D, true>, VanLeer<D>, NSDiffusionFlux<D, true>>>(std::move(fs), std::move(d), *p);
    } else {
        using Dsc = ConservationLawDDG<D, NSFlux<D, false>, VanLeer<D>, NSDiffusionFlux<D, false>>;
        Dsc d{NSFlux<D, false>{ph}, VanLeer<D>{ph}, NSDiffusionFlux<D, false>{ph, false}};
        return std::make_unique<Problem<D, NSFlux<D, false>, VanLeer<D>, NSDiffusionFlux<D, false>>>(std::move(fs), std::move(d), *p);
    }
}

// The flux functors hold pointers to their coefficient objects (the reference
// holds references, burgers.hpp:58); the handle owns both.
struct Handle {
    BurgersCoefficients bc;
    NSPhysics<1> ph1;
    NSPhysics<2> ph2;
    NSPhysics<3> ph3;
    std::unique_ptr<ProblemBase> prob;
};

template <int D>
void setup_physics(NSPhysics<D>& ph, const ice_disc_params* p) {
    ph.gamma = p->gamma; ph.Rgas = p->Rgas; ph.Pr = p->Pr;
    ph.ref = ReferenceParameters{p->ref[0], p->ref[1], p->ref[2], p->ref[3], p->ref[4], p->ref[5], p->ref[6], p->ref[7]};
    ph.nondim = create_nondim(ph.ref);
    ph.visc_kind = p->viscosity_kind; ph.visc_mu = p->visc_mu;
    for (int k = 0; k < 4; ++k) ph.isothermal_T[k] = p->isothermal_T[k];
    ph.fs_rho = p->freestream[0]; ph.fs_u = p->freestream[1];
    for (int d = 0; d < 3; ++d) ph.fs_dir[d] = p->freestream[2 + d];
    ph.fs_T = p->freestream[5];
}

}  // namespace

extern "C" {

const char* ice_oracle_last_error(void) { return g_err.c_str(); }

void* ice_oracle_create(const ice_mesh_desc* m, const ice_space_desc* s, const ice_disc_params* p, int fix3d) {
    try {
        if (!m || !s || !p) throw std::runtime_error("null descriptor");
        if (m->ndim < 1 || m->ndim > 3) throw std::runtime_error("ndim must be 1..3");
        if (s->order < 1 || s->order > MAXP - 1) throw std::runtime_error("order out of range");
        auto h = std::make_unique<Handle>();
        h->bc.mu = p->mu;
        for (int d = 0; d < 3; ++d) { h->bc.a[d] = p->a[d]; h->bc.b[d] = p->b[d]; }
        setup_physics<1>(h->ph1, p);
        setup_physics<2>(h->ph2, p);
        setup_physics<3>(h->ph3, p);
        if (m->ndim == 1) h->prob = make<1>(m, s, p, fix3d != 0, &h->bc, &h->ph1);
        else if (m->ndim == 2) h->prob = make<2>(m, s, p, fix3d != 0, &h->bc, &h->ph2);
        else h->prob = make<3>(m, s, p, fix3d != 0, &h->bc, &h->ph3);
        return h.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ice_oracle_destroy(void* h) { delete static_cast<Handle*>(h); }

#define ORACLE_TRY(body) \
    try { body; return 0; } catch (const std::exception& e) { g_err = e.what(); return -1; }

int ice_oracle_info(void* h, int* out) { ORACLE_TRY(static_cast<Handle*>(h)->prob->info(out)) }
int ice_oracle_residual(void* h, const double* u, double* r) { ORACLE_TRY(static_cast<Handle*>(h)->prob->residual(u, r)) }
int ice_oracle_residual_parallel(void* h, const double* u, double* r, int nthreads) {
    ORACLE_TRY(static_cast<Handle*>(h)->prob->residual_parallel(u, r, nthreads))
}
int ice_oracle_max_threads() { return omp_get_max_threads(); }
int ice_oracle_faces(void* h, int* recs) { ORACLE_TRY(static_cast<Handle*>(h)->prob->faces(recs)) }
int ice_oracle_mesh(void* h, double* coord, int* conn) { ORACLE_TRY(static_cast<Handle*>(h)->prob->mesh_arrays(coord, conn)) }
int ice_oracle_project(void* h, const ice_field* f, double* u) { ORACLE_TRY(static_cast<Handle*>(h)->prob->project(f, u)) }
double ice_oracle_l2_error(void* h, const ice_field* f, const double* u) {
    try { return static_cast<Handle*>(h)->prob->l2_error(f, u); } catch (const std::exception& e) { g_err = e.what(); return -1.0; }
}
int ice_oracle_rk3_ssp(void* h, double* u, double dt, int nsteps, int mode) {
    ORACLE_TRY(static_cast<Handle*>(h)->prob->rk3_ssp(u, dt, nsteps, mode))
}
int ice_oracle_rk3_tvd_tfinal(void* h, double* u, double dt, double tfinal) {
    try { return static_cast<Handle*>(h)->prob->rk3_tvd_tfinal(u, dt, tfinal); } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
double ice_oracle_cfl_dt(void* h, double cfl) {
    try { return static_cast<Handle*>(h)->prob->cfl_dt(cfl); } catch (const std::exception& e) { g_err = e.what(); return -1.0; }
}
int ice_oracle_trace_points(void* h, double* xiL, double* xiR) { ORACLE_TRY(static_cast<Handle*>(h)->prob->trace_points(xiL, xiR)) }
int ice_oracle_face_geometry(void* h, double* x, double* n, double* sg) { ORACLE_TRY(static_cast<Handle*>(h)->prob->face_geometry(x, n, sg)) }
int ice_oracle_qp_coords(void* h, double* xq) { ORACLE_TRY(static_cast<Handle*>(h)->prob->qp_coords(xq)) }
int ice_oracle_mass_apply(void* h, const double* r, double* out) { ORACLE_TRY(static_cast<Handle*>(h)->prob->mass_apply(r, out)) }
int ice_oracle_jacobian(void* h, const double* u, double* res, int* cols, double* blocks) {
    ORACLE_TRY(static_cast<Handle*>(h)->prob->jacobian(u, res, cols, blocks))
}
int ice_oracle_mdg_nodeset(void* h, const int* traces, int ntr, int* nsel, int* nodes) {
    ORACLE_TRY(static_cast<Handle*>(h)->prob->mdg_nodeset(traces, ntr, nsel, nodes))
}
int ice_oracle_mdg_residual(void* h, const int* traces, int ntr, const double* u, double* res) {
    ORACLE_TRY(static_cast<Handle*>(h)->prob->mdg_residual(traces, ntr, u, res))
}
int ice_oracle_newton(void* h, double* u, double tau_abs, double tau_rel, int kmax, double* rhist, int nhist) {
    try {
        return static_cast<Handle*>(h)->prob->newton(u, tau_abs, tau_rel, kmax, rhist, nhist);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}
int ice_oracle_rk_solve(void* h, double* u, int kind, int compat, int use_cfl, double dt, double cfl, int use_tfinal,
                        double tfinal, long long ntime, long long* nsteps, double* time) {
    ORACLE_TRY(static_cast<Handle*>(h)->prob->rk_solve(u, kind, compat, use_cfl, dt, cfl, use_tfinal, tfinal, ntime, nsteps, time))
}

// reference-element primitives for closed-form checks
int ice_oracle_lagrange_1d(int p, double xi, double* N, double* dN, double* d2N) {
    if (p < 0 || p > MAXP) return -1;
    Lagrange1D l(p);
    l.d2(xi, N, dN, d2N);
    return 0;
}
int ice_oracle_gauss_legendre(int n, double* x, double* w) {
    if (n < 1 || n > MAXP + 1) return -1;
    GaussLegendre1D g(n);
    for (int i = 0; i < n; ++i) { x[i] = g.x[i]; w[i] = g.w[i]; }
    return 0;
}
int ice_oracle_tensor_basis(int ndim, int p, const double* xi, double* B, double* G, double* H) {
    if (ndim == 1) { TensorBasis<1>(p).eval(xi, B, G, H); return 0; }
    if (ndim == 2) { TensorBasis<2>(p).eval(xi, B, G, H); return 0; }
    if (ndim == 3) { TensorBasis<3>(p).eval(xi, B, G, H); return 0; }
    return -1;
}
// element map Jacobian of a geometry-order-1 hypercube at xi (vertices in tensor order)
int ice_oracle_element_jacobian(int ndim, const double* X, const double* xi, double* J) {
    auto run = [&](auto dimc) {
        constexpr int D = decltype(dimc)::value;
        ElementGeo<D> g;
        for (int k = 0; k < (1 << D); ++k)
            for (int d = 0; d < D; ++d) g.X[k][d] = X[k * D + d];
        double JJ[D][D], H[D][D][D];
        g.eval(TensorBasis<D>(1), xi, nullptr, JJ, H);
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) J[i * D + j] = JJ[i][j];
    };
    if (ndim == 1) run(std::integral_constant<int, 1>{});
    else if (ndim == 2) run(std::integral_constant<int, 2>{});
    else if (ndim == 3) run(std::integral_constant<int, 3>{});
    else return -1;
    return 0;
}
// trace coordinates -> element reference point for face `face_nr` seen with
// `orient` (orient_transform then trace_to_ref), for arbitrary s
int ice_oracle_trace_ref(int ndim, int face_nr, int orient, const double* s, double* xi) {
    double sr[3] = {0, 0, 0};
    if (ndim == 1) { trace_to_ref<1>(face_nr, s, xi); return 0; }
    if (ndim == 2) { orient_transform<2>(orient, s, sr); trace_to_ref<2>(face_nr, sr, xi); return 0; }
    if (ndim == 3) { orient_transform<3>(orient, s, sr); trace_to_ref<3>(face_nr, sr, xi); return 0; }
    return -1;
}
// NS physical flux (zero gradient allowed) and VanLeer, 1..3D, for the
// known answers of test_ns.cpp:66-107
int ice_oracle_ns_flux(const ice_disc_params* p, int ndim, const double* u, const double* gradu, double* F) {
    auto run = [&](auto dimc) {
        constexpr int D = decltype(dimc)::value;
        NSPhysics<D> ph; setup_physics<D>(ph, p);
        NSFlux<D, true> f{&ph};
        f(u, gradu, F);
    };
    if (ndim == 1) run(std::integral_constant<int, 1>{});
    else if (ndim == 2) run(std::integral_constant<int, 2>{});
    else if (ndim == 3) run(std::integral_constant<int, 3>{});
    else return -1;
    return 0;
}
int ice_oracle_vanleer(const ice_disc_params* p, int ndim, const double* uL, const double* uR, const double* n, double* f) {
    auto run = [&](auto dimc) {
        constexpr int D = decltype(dimc)::value;
        NSPhysics<D> ph; setup_physics<D>(ph, p);
        VanLeer<D> vl{&ph};
        vl(uL, uR, n, f);
    };
    if (ndim == 1) run(std::integral_constant<int, 1>{});
    else if (ndim == 2) run(std::integral_constant<int, 2>{});
    else if (ndim == 3) run(std::integral_constant<int, 3>{});
    else return -1;
    return 0;
}

// Diffusive numerical flux at a face point (navier_stokes.hpp:1241-1289 for
// NS, burgers.hpp:206-220 for Burgers) for the consistency checks against
// the physical flux.
int ice_oracle_diff_flux(const ice_disc_params* p, int ndim, const double* u, const double* gradu, const double* n,
                         double* f) {
    auto run = [&](auto dimc) {
        constexpr int D = decltype(dimc)::value;
        if (p->flux == ICE_FLUX_BURGERS) {
            BurgersCoefficients c; c.mu = p->mu;
            for (int d = 0; d < 3; ++d) { c.a[d] = p->a[d]; c.b[d] = p->b[d]; }
            BurgersDiffusionFlux<D>{&c}(u, gradu, n, f);
        } else {
            NSPhysics<D> ph; setup_physics<D>(ph, p);
            NSDiffusionFlux<D, true> df{&ph};
            df.textbook_energy = p->ns_face_energy_textbook != 0;
            df(u, gradu, n, f);
        }
    };
    if (ndim == 1) run(std::integral_constant<int, 1>{});
    else if (ndim == 2) run(std::integral_constant<int, 2>{});
    else if (ndim == 3) run(std::integral_constant<int, 3>{});
    else return -1;
    return 0;
}
// Burgers physical flux (burgers.hpp:72-91); lambda_out gets the wave speed norm
int ice_oracle_burgers_flux(const ice_disc_params* p, int ndim, const double* u, const double* gradu, double* F,
                            double* lambda_out) {
    auto run = [&](auto dimc) {
        constexpr int D = decltype(dimc)::value;
        BurgersCoefficients c; c.mu = p->mu;
        for (int d = 0; d < 3; ++d) { c.a[d] = p->a[d]; c.b[d] = p->b[d]; }
        BurgersFlux<D> fl{&c};
        fl(u, gradu, F);
        *lambda_out = fl.lambda_max;
    };
    if (ndim == 1) run(std::integral_constant<int, 1>{});
    else if (ndim == 2) run(std::integral_constant<int, 2>{});
    else if (ndim == 3) run(std::integral_constant<int, 3>{});
    else return -1;
    return 0;
}
// NS boundary state (apply_bc, navier_stokes.hpp:975-1189); FULL_NS picks the
// wall-general mapping
int ice_oracle_apply_bc(const ice_disc_params* p, int ndim, const double* uL, const double* graduL, const double* n,
                        int bctype, int bcflag, double* uR, double* graduR) {
    try {
        auto run = [&](auto dimc) {
            constexpr int D = decltype(dimc)::value;
            NSPhysics<D> ph; setup_physics<D>(ph, p);
            double ul[D + 2];
            for (int k = 0; k < D + 2; ++k) ul[k] = uL[k];
            if (p->flux == ICE_FLUX_EULER) NSFlux<D, false>{&ph}.apply_bc(ul, graduL, n, bctype, bcflag, uR, graduR);
            else NSFlux<D, true>{&ph}.apply_bc(ul, graduL, n, bctype, bcflag, uR, graduR);
        };
        if (ndim == 1) run(std::integral_constant<int, 1>{});
        else if (ndim == 2) run(std::integral_constant<int, 2>{});
        else if (ndim == 3) run(std::integral_constant<int, 3>{});
        else return -1;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
