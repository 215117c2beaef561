// All-core variant of the oracle's form_residual (test infrastructure: the
// CPU baseline timed on every host core).  Same arithmetic as
// form_residual in disc.hpp (form_residual.hpp:56-268), bit for bit:
//  1. every boundary and interior face integral is evaluated in parallel
//     into its own buffer (faces are independent; each thread works on a
//     private copy of the discretization, whose wave-speed bookkeeping is
//     mutable);
//  2. every element, in parallel, sums its contributions in the order the
//     serial loop adds them (boundary faces, interior faces, coupled
//     periodic faces, each in list order, then its own volume integral),
//     starting from zero, so each entry sees the same sequence of additions.
// The per-element contribution lists are built once per space.
#pragma once
#include <omp.h>

#include <vector>

#include "disc.hpp"

namespace oracle {

template <int D>
struct ParallelPlan {
    // contribution c of element e: buffer offset (in units of one element
    // block) into the face buffer, listed in serial order
    std::vector<int> start;  // [nelem + 1]
    std::vector<long> src;   // block offsets
    int nb_faces = 0, ni_faces = 0;
    std::vector<const typename FESpace<D>::Trace*> bnd, itf;

    template <class Disc>
    void build(const FESpace<D>& fs) {
        bnd.clear();
        itf.clear();
        for (const auto& t : fs.boundary) {
            if (fs.coupled_periodic && t.bctype == BC_PERIODIC) continue;
            bnd.push_back(&t);
        }
        for (const auto& t : fs.interior) itf.push_back(&t);
        for (const auto& t : fs.periodic) itf.push_back(&t);
        nb_faces = (int)bnd.size();
        ni_faces = (int)itf.size();
        const int ne = fs.mesh.nelem;
        std::vector<std::vector<long>> lists(ne);
        // face buffer layout: boundary face f -> block f; interior face f ->
        // blocks nb_faces + 2f (left) and nb_faces + 2f + 1 (right)
        for (int f = 0; f < nb_faces; ++f) lists[bnd[f]->elemL].push_back(f);
        for (int f = 0; f < ni_faces; ++f) {
            lists[itf[f]->elemL].push_back(nb_faces + 2L * f);
            lists[itf[f]->elemR].push_back(nb_faces + 2L * f + 1);
        }
        start.assign(ne + 1, 0);
        src.clear();
        for (int e = 0; e < ne; ++e) {
            start[e] = (int)src.size();
            src.insert(src.end(), lists[e].begin(), lists[e].end());
        }
        start[ne] = (int)src.size();
    }
};

template <int D, class Disc>
void form_residual_parallel(const FESpace<D>& fs, const Disc& disc, const double* u, double* res,
                            ParallelPlan<D>& plan, std::vector<double>& fbuf, int nthreads) {
    constexpr int neq = Disc::neq;
    const int esz = fs.nb * neq;
    for (const auto& t : fs.boundary)
        if (!fs.coupled_periodic && t.bctype == BC_PERIODIC && Disc::PFlux::implements_bcs)
            throw std::runtime_error("PERIODIC boundary with a flux that implements apply_bc is undefined in the reference");
    if (plan.start.empty()) plan.template build<Disc>(fs);
    const long nblk = plan.nb_faces + 2L * plan.ni_faces;
    fbuf.resize((size_t)std::max<long>(nblk, 1) * esz);
    const int ne = fs.mesh.nelem;
#pragma omp parallel num_threads(nthreads)
    {
        Disc dl = disc;  // private wave-speed bookkeeping
#pragma omp for schedule(static)
        for (int f = 0; f < plan.nb_faces; ++f) {
            const auto& t = *plan.bnd[f];
            double* rL = fbuf.data() + (size_t)f * esz;
            std::fill(rL, rL + esz, 0.0);
            dl.boundary_integral(fs, t, u + (size_t)t.elemL * esz, u + (size_t)t.elemR * esz, rL);
        }
#pragma omp for schedule(static)
        for (int f = 0; f < plan.ni_faces; ++f) {
            const auto& t = *plan.itf[f];
            double* rL = fbuf.data() + (size_t)(plan.nb_faces + 2L * f) * esz;
            double* rR = rL + esz;
            std::fill(rL, rL + 2 * esz, 0.0);
            dl.trace_integral(fs, t, u + (size_t)t.elemL * esz, u + (size_t)t.elemR * esz, rL, rR);
        }
        std::vector<double> rv(esz);
#pragma omp for schedule(static)
        for (int e = 0; e < ne; ++e) {
            double* re = res + (size_t)e * esz;
            std::fill(re, re + esz, 0.0);
            for (int c = plan.start[e]; c < plan.start[e + 1]; ++c) {
                const double* b = fbuf.data() + (size_t)plan.src[c] * esz;
                for (int k = 0; k < esz; ++k) re[k] += b[k];
            }
            std::fill(rv.begin(), rv.end(), 0.0);
            dl.domain_integral(fs, e, u + (size_t)e * esz, rv.data());
            for (int k = 0; k < esz; ++k) re[k] += rv[k];
        }
    }
}

}  // namespace oracle
