// Test infrastructure (the oracle): the MDG interface-conservation residual
// on selected traces, restated from the reference.
//  * nodeset_dof_map (fe_function/node_set_layout.hpp:46-83): the nodes of
//    the selected traces, minus every node of a boundary face, in ascending
//    global node order;
//  * ConservationLawDDG::interface_conservation (disc/conservation_law.hpp:
//    1178-1288): per face quadrature point the jump of the normal physical
//    flux, (F(uR, grad uR) - F(uL, grad uL)) . n, times w sqrt(g), tested
//    against the face's geometry-order Lagrange basis (the trace basis,
//    element/reference_element.hpp:228-236); Dirichlet boundary traces take
//    uR from the Dirichlet callback, other boundary types contribute nothing;
//    with P1 on both sides the gradients are dropped (:1256-1260);
//  * form_mdg_residual (solvers/form_residual.hpp:270-327): zero, then per
//    selected trace in selection order scatter the face residual onto the
//    selected nodes (scatter_facspan: trace basis function k <-> face node k).
// Trace index convention (FESpace::traces): interior faces first, then the
// boundary faces, in mesh face order.
#pragma once
#include <vector>

#include "disc.hpp"

namespace oracle {

template <int D>
struct NodeSet {
    std::vector<int> traces;         // selected trace indices
    std::vector<int> selected;       // global node ids of the MDG degrees of freedom
    std::vector<int> inv;            // node -> index in selected, or selected.size()

    NodeSet(const FESpace<D>& fs, const std::vector<int>& tr) : traces(tr) {
        std::vector<char> pick(fs.mesh.nnode, 0);
        const int ni = (int)fs.interior.size();
        auto trace_of = [&](int t) -> const typename FESpace<D>::Trace& {
            return t < ni ? fs.interior[t] : fs.boundary[t - ni];
        };
        for (int t : tr)
            for (int nd : trace_of(t).nodes) pick[nd] = 1;
        for (const auto& b : fs.boundary)
            for (int nd : b.nodes) pick[nd] = 0;
        for (int nd = 0; nd < fs.mesh.nnode; ++nd)
            if (pick[nd]) selected.push_back(nd);
        inv.assign(fs.mesh.nnode, (int)selected.size());
        for (int i = 0; i < (int)selected.size(); ++i) inv[selected[i]] = i;
    }
};

// interface_conservation of one trace: res[k][v], k over the face nodes
template <int D, class Disc>
void interface_conservation(const FESpace<D>& fs, const Disc& disc, const typename FESpace<D>::Trace& t,
                            const double* UL, const double* UR, double* res) {
    constexpr int neq = Disc::neq;
    constexpr int NFV = FESpace<D>::NFV;
    const int nb = fs.nb;
    std::vector<double> gL(nb * D), gR(nb * D), hL(nb * D * D), hR(nb * D * D);
    for (int k = 0; k < NFV * neq; ++k) res[k] = 0.0;
    for (int q = 0; q < fs.nfq; ++q) {
        const double w = fs.fquad.w[q];
        double x[D], Jf[D][(D > 1 ? D - 1 : 1)], nrm[D];
        fs.face_geo(t, q, x, Jf);
        const double sqrtg = root_riemann<D>(Jf);
        unit_normal<D>(Jf, nrm);
        const double* biL = &t.BL[q * nb];
        const double* biR = &t.BR[q * nb];
        disc.side_eval(fs, t.elemL, &t.xiL[q * D], &t.GL[q * nb * D], &t.HL[q * nb * D * D], gL.data(), hL.data());
        disc.side_eval(fs, t.elemR, &t.xiR[q * D], &t.GR[q * nb * D], &t.HR[q * nb * D * D], gR.data(), hR.data());
        double uL[neq], uR[neq], guL[neq * D], guR[neq * D];
        for (int v = 0; v < neq; ++v) {
            uL[v] = 0;
            uR[v] = 0;
            for (int i = 0; i < nb; ++i) uL[v] += UL[i * neq + v] * biL[i];
            for (int i = 0; i < nb; ++i) uR[v] += UR[i * neq + v] * biR[i];
        }
        Disc::contract(nb, UL, gL.data(), D, guL);
        Disc::contract(nb, UR, gR.data(), D, guR);
        if (t.bctype != BC_INTERIOR) {
            if (t.bctype == BC_DIRICHLET) {
                disc.dirichlet_callbacks[t.bcflag](x, uR);
            } else {
                for (int k = 0; k < NFV * neq; ++k) res[k] = 0.0;
                return;
            }
        }
        if (fs.p == 1) {  // P1 on both sides: no diffusive part (:1256-1260)
            for (int k = 0; k < neq * D; ++k) guL[k] = guR[k] = 0.0;
        }
        double FL[neq * D], FR[neq * D];
        disc.phys_flux(uL, guL, FL);
        disc.phys_flux(uR, guR, FR);
        // the trace basis: geometry-order-1 Lagrange on the face reference domain
        double s[3] = {0, 0, 0};
        for (int d = 0; d < FESpace<D>::TD; ++d) s[d] = fs.fquad.pts[q][d];
        double bt[NFV], gt[NFV * (D > 1 ? D - 1 : 1)];
        if constexpr (D == 1) bt[0] = 1.0;
        else fs.fq1.eval(s, bt, gt, nullptr);
        for (int v = 0; v < neq; ++v) {
            double jn = 0.0;
            for (int d = 0; d < D; ++d) jn += FR[v * D + d] * nrm[d] - FL[v * D + d] * nrm[d];
            const double ic = jn * sqrtg * w;
            for (int k = 0; k < NFV; ++k) res[k * neq + v] -= ic * bt[k];
        }
    }
}

// form_mdg_residual: out[selected node][v]
template <int D, class Disc>
void form_mdg_residual(const FESpace<D>& fs, const Disc& disc, const NodeSet<D>& ns, const double* u, double* out) {
    constexpr int neq = Disc::neq;
    constexpr int NFV = FESpace<D>::NFV;
    const int esz = fs.nb * neq;
    std::fill(out, out + ns.selected.size() * neq, 0.0);
    const int ni = (int)fs.interior.size();
    std::vector<double> res(NFV * neq);
    for (int tr : ns.traces) {
        const auto& t = tr < ni ? fs.interior[tr] : fs.boundary[tr - ni];
        interface_conservation(fs, disc, t, u + (size_t)t.elemL * esz, u + (size_t)t.elemR * esz, res.data());
        for (int k = 0; k < NFV; ++k) {
            const int sel = ns.inv[t.nodes[k]];
            if (sel >= (int)ns.selected.size()) continue;
            for (int v = 0; v < neq; ++v) out[(size_t)sel * neq + v] += res[k * neq + v];
        }
    }
}

}  // namespace oracle
