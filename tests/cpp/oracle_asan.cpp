// The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
// (the reference's sanitizer build, CMakeLists.txt:38,52-56): one translation
// unit with the oracle's C API, driven over every flux family, boundary
// type, dimension and solver entry point on small meshes.  Any invalid
// access or undefined operation aborts with the sanitizer's report; a clean
// run prints "OK".
#include "../../oracle/capi.cpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

int failures = 0;

void expect(bool c, const char* what) {
    if (!c) {
        std::printf("FAILED: %s (%s)\n", what, ice_oracle_last_error());
        ++failures;
    }
}

// the reference defaults (burgers.hpp:14-40, navier_stokes.hpp:129-213)
void ice_disc_params_default_oracle(int flux, ice_disc_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->flux = flux;
    p->mu = 0.001;
    p->gamma = 1.4;
    p->Rgas = 287.052874;
    p->Pr = 0.72;
    for (double& r : p->ref) r = 1.0;
    p->visc_mu = 1.0;
    for (double& t : p->isothermal_T) t = 1.0;
    p->freestream[0] = 1.0;
    p->freestream[2] = 1.0;
    p->freestream[5] = 1.0;
}

ice_mesh_desc box(int ndim, int n, const int* bct, const int* flag, int periodic_mode) {
    ice_mesh_desc m{};
    m.ndim = ndim;
    for (int d = 0; d < ndim; ++d) {
        m.nelem[d] = n + d % 2;
        m.xmax[d] = 1.0;
    }
    for (int k = 0; k < 2 * ndim; ++k) {
        m.bctype[k] = bct[k];
        m.bcflag[k] = flag ? flag[k] : 0;
    }
    m.periodic_mode = periodic_mode;
    return m;
}

std::vector<double> state(int flux, int ndim, int n) {
    std::vector<double> u;
    const int m = flux == ICE_FLUX_BURGERS ? 1 : ndim + 2;
    for (int i = 0; i < n / m; ++i) {
        const double s = 0.1 * std::sin(0.37 * i);
        if (flux == ICE_FLUX_BURGERS) {
            u.push_back(s);
            continue;
        }
        const double rho = 1.0 + s;
        u.push_back(rho);
        for (int d = 0; d < ndim; ++d) u.push_back(rho * (0.3 - 0.1 * d + s));
        u.push_back(1.0 / 0.4 / 1.4 + 0.1 + s);
    }
    return u;
}

void run(const char* name, const ice_mesh_desc& m, int order, const ice_disc_params& p, bool solvers) {
    ice_space_desc s{order, 1};
    void* h = ice_oracle_create(&m, &s, &p, 1);
    expect(h != nullptr, name);
    if (!h) return;
    int info[10];
    ice_oracle_info(h, info);
    const int n = info[1] * info[2] * info[3];
    std::vector<double> u = state(p.flux, m.ndim, n), r(n), out(n);
    expect(ice_oracle_residual(h, u.data(), r.data()) == 0, name);
    // the all-core variant is bit-identical to the serial loop
    expect(ice_oracle_residual_parallel(h, u.data(), out.data(), 4) == 0, name);
    expect(std::memcmp(r.data(), out.data(), n * sizeof(double)) == 0, name);
    expect(ice_oracle_mass_apply(h, r.data(), out.data()) == 0, name);
    std::vector<int> recs(info[4] * (6 + (1 << (m.ndim - 1))));
    expect(ice_oracle_faces(h, recs.data()) == 0, name);
    std::vector<double> xq((size_t)info[1] * info[7] * m.ndim);
    expect(ice_oracle_qp_coords(h, xq.data()) == 0, name);
    ice_field c{};
    c.kind = ICE_FIELD_CONSTANT;
    for (int k = 0; k < info[3]; ++k) c.p[k] = u[k];
    expect(ice_oracle_project(h, &c, out.data()) == 0, name);
    expect(ice_oracle_l2_error(h, &c, u.data()) >= 0.0, name);
    long long nsteps = 0;
    double t = 0;
    std::vector<double> w = u;
    expect(ice_oracle_rk_solve(h, w.data(), 0, 0, 1, 0.0, 0.1, 0, 0.0, 2, &nsteps, &t) == 0, name);
    if (solvers) {
        const int S = 2 * m.ndim + 1, bs = info[2] * info[3];
        std::vector<int> cols((size_t)info[1] * S);
        std::vector<double> blocks((size_t)info[1] * S * bs * bs);
        expect(ice_oracle_jacobian(h, u.data(), r.data(), cols.data(), blocks.data()) == 0, name);
        std::vector<double> hist(8);
        w = u;
        ice_oracle_newton(h, w.data(), 1e-10, 0.0, 2, hist.data(), 8);
    }
    ice_oracle_destroy(h);
    std::printf("%s: ok\n", name);
}

}  // namespace

int main() {
    ice_disc_params b;
    ice_disc_params_default_oracle(ICE_FLUX_BURGERS, &b);
    b.mu = 0.02;
    b.b[0] = b.b[1] = b.b[2] = 1.0;
    b.a[0] = 0.3;
    b.sigma_ic = 0.5;
    b.dirichlet[0].kind = ICE_FIELD_CONSTANT;
    b.dirichlet[0].p[0] = 0.25;
    b.neumann[0].kind = ICE_FIELD_CONSTANT;
    b.neumann[0].p[0] = -0.5;
    const int bc1[2] = {ICE_BC_DIRICHLET, ICE_BC_NEUMANN};
    run("burgers 1D P3 DDGIC", box(1, 4, bc1, nullptr, ICE_PERIODIC_REFERENCE), 3, b, true);
    const int bc2[4] = {ICE_BC_DIRICHLET, ICE_BC_NEUMANN, ICE_BC_EXTRAPOLATION, ICE_BC_DIRICHLET};
    run("burgers 2D P2 mixed BCs", box(2, 3, bc2, nullptr, ICE_PERIODIC_REFERENCE), 2, b, true);
    b.ddgic_form = ICE_DDGIC_PUBLISHED;
    const int bcp[6] = {ICE_BC_PERIODIC, ICE_BC_PERIODIC, ICE_BC_PERIODIC, ICE_BC_PERIODIC, ICE_BC_PERIODIC,
                        ICE_BC_PERIODIC};
    run("burgers 3D P2 coupled periodic", box(3, 2, bcp, nullptr, ICE_PERIODIC_COUPLED), 2, b, false);
    run("burgers 2D P1 reference periodic", box(2, 3, bcp, nullptr, ICE_PERIODIC_REFERENCE), 1, b, false);

    ice_disc_params ns;
    ice_disc_params_default_oracle(ICE_FLUX_NAVIER_STOKES, &ns);
    ns.visc_mu = 0.003;
    ns.dirichlet[0].kind = ICE_FIELD_NS_MMS_STATE;
    ns.source.kind = ICE_FIELD_NS_MMS_SOURCE;
    ns.isothermal_T[0] = (1.0 / 1.4) / 287.052874;
    ns.freestream[1] = 0.3;
    ns.freestream[5] = (1.0 / 1.4) / 287.052874;
    const int dir4[4] = {ICE_BC_DIRICHLET, ICE_BC_DIRICHLET, ICE_BC_DIRICHLET, ICE_BC_DIRICHLET};
    run("NS 2D P3 manufactured", box(2, 2, dir4, nullptr, ICE_PERIODIC_REFERENCE), 3, ns, true);
    const int walls[4] = {ICE_BC_RIEMANN, ICE_BC_SLIP_WALL, ICE_BC_EXTRAPOLATION, ICE_BC_NO_SLIP_ISOTHERMAL};
    ns.viscosity_kind = ICE_VISC_DIMENSIONLESS_SUTHERLAND;
    run("NS 2D P2 walls, Riemann, Sutherland", box(2, 3, walls, nullptr, ICE_PERIODIC_REFERENCE), 2, ns, false);
    ns.viscosity_kind = ICE_VISC_CONSTANT;
    ns.dirichlet[0].kind = ICE_FIELD_CONSTANT;
    ns.dirichlet[0].p[0] = 1.0;
    ns.dirichlet[0].p[1] = 0.3;
    ns.dirichlet[0].p[4] = 1.0 / 0.4 / 1.4 + 0.045;
    ns.source.kind = ICE_FIELD_NONE;
    const int w3[6] = {ICE_BC_DIRICHLET, ICE_BC_SLIP_WALL, ICE_BC_RIEMANN, ICE_BC_EXTRAPOLATION,
                       ICE_BC_NO_SLIP_ISOTHERMAL, ICE_BC_DIRICHLET};
    run("NS 3D P2 mixed BCs", box(3, 2, w3, nullptr, ICE_PERIODIC_REFERENCE), 2, ns, false);
    ice_disc_params eu;
    ice_disc_params_default_oracle(ICE_FLUX_EULER, &eu);
    run("Euler 3D P3 coupled periodic", box(3, 2, bcp, nullptr, ICE_PERIODIC_COUPLED), 3, eu, false);
    if (failures == 0) std::printf("OK\n");
    return failures == 0 ? 0 : 1;
}
