// MDG interface-conservation residual on selected traces
// (form_mdg_residual, form_residual.hpp:270-327, over
// ConservationLawDDG::interface_conservation, conservation_law.hpp:1178-1288).
//
// The degrees of freedom are the nodes of the selected traces minus every
// node of a boundary face (nodeset_dof_map, node_set_layout.hpp:46-83), in
// ascending node order.  A selected boundary trace therefore adds nothing:
// its nodes are all deselected and its residual is dropped by the scatter.
// Two launches:
//  1. one thread per (selected interior face, face point): both traces
//     (value and physical gradient, FP64; gradients dropped at P1 as the
//     reference does), the physical fluxes on both sides, and the jump of
//     the normal flux times w sqrt(g) -> J[face][q][v];
//  2. one thread per (selected node, variable): the node's contributions in
//     selection order, sum_q J[face][q][v] * psi_k(s_q) with psi_k the P1
//     trace basis function of the node on that face; no atomics, so the
//     summation order is fixed.
// The combined residual (form_residual over a nodeset, form_residual.hpp:
// 380-424) first moves the selected nodes to the coordinates that follow the
// DG coefficients in u: one thread per selected node writes them into every
// element vertex slot of that node, then the face geometry table (and the
// source load) are rebuilt from the moved vertices, and the DG and MDG
// residuals are formed as above.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "dg_common.h"
#include "engine.h"
#include "mesh.hpp"
#include "physics_dev.h"

struct ice_mdg_sel {
    const ice_space* space = nullptr;
    int nnode = 0;                 // selected nodes
    int nface = 0;                 // selected interior faces (device face rows)
    std::vector<int> nodes;        // global ids of the selected nodes
    int* d_face = nullptr;         // [nface] face rows
    int* d_start = nullptr;        // [nnode + 1] contribution ranges
    int* d_contrib = nullptr;      // [ncontrib] (selected face index << 2) | face-local node
    double* d_jump = nullptr;      // [nface][nfq][neq_max]
    int* d_vstart = nullptr;       // [nnode + 1] ranges of d_vslot
    int* d_vslot = nullptr;        // element vertex slots (element * 2^ndim + local vertex) of each selected node
    ~ice_mdg_sel() {
        for (void* p : {(void*)d_face, (void*)d_start, (void*)d_contrib, (void*)d_jump, (void*)d_vstart,
                        (void*)d_vslot})
            if (p) (void)hipFree(p);
    }
};

namespace ice {
namespace {

constexpr int MDG_MAX_NEQ = 5;

// DROP: drop the gradients at P1 as interface_conservation does; frows
// null: every interior face row in order (the interface error).
template <int D, int P, class PH, bool DROP = true>
__global__ __launch_bounds__(256) void mdg_jump_kernel(DevSpace S, typename PH::Params ph, const double* __restrict__ U,
                                                       const int* __restrict__ frows, int nface,
                                                       double* __restrict__ jump) {
    constexpr int N = P + 1;
    constexpr int NB = Pow<N, D>::value;
    constexpr int NFQ = Pow<N, D - 1>::value;
    constexpr int M = PH::M;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = tid / NFQ, q = tid - s * NFQ;
    if (s >= nface) return;
    const int f = frows ? frows[s] : s;
    const int4 rec = S.fint[f];
    const int eL = rec.x, eR = rec.y, info = rec.z;
    const int fnl = info_fnl(info), fnr = info_fnr(info), orient = info_orient(info);
    const double* g = S.fgeo + ((size_t)f * NFQ + q) * (D + 2);
    double nrm[D];
#pragma unroll
    for (int d = 0; d < D; ++d) nrm[d] = g[d];
    const double ws = g[D];
    double uS[2][M], gS[2][M * D];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int e = side ? eR : eL;
        double X[1 << D][D];
        load_vertices<D>(S.X, e, X);
        const signed char* fp = S.fpt + (((side ? fnr : fnl) * S.norient + (side ? orient : 0)) * NFQ + q) * D;
        int pt[D];
#pragma unroll
        for (int d = 0; d < D; ++d) pt[d] = fp[d];
        trace_eval<D, N, M, false>(S, U + (size_t)e * NB * M, pt, X, nrm, uS[side], gS[side], nullptr);
        if (DROP && P == 1) {  // P1 on both sides: no diffusive part (conservation_law.hpp:1256-1260)
#pragma unroll
            for (int k = 0; k < M * D; ++k) gS[side][k] = 0.0;
        }
    }
    double FL[M * D], FR[M * D];
    PH::flux(ph, uS[0], gS[0], FL);
    PH::flux(ph, uS[1], gS[1], FR);
#pragma unroll
    for (int v = 0; v < M; ++v) {
        double jn = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) jn += FR[v * D + d] * nrm[d] - FL[v * D + d] * nrm[d];
        jump[((size_t)s * NFQ + q) * M + v] = jn * ws;
    }
}

// res[node][v] = -sum over the node's faces of sum_q J[face][q][v] psi_k(s_q)
template <int D, int P, int M>
__global__ __launch_bounds__(256) void mdg_gather_kernel(DevSpace S, const int* __restrict__ start,
                                                         const int* __restrict__ contrib, int nnode,
                                                         const double* __restrict__ jump, double* __restrict__ res) {
    constexpr int N = P + 1;
    constexpr int TD = D - 1;
    constexpr int NFQ = Pow<N, TD>::value;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int node = tid / M, v = tid - node * M;
    if (node >= nnode) return;
    const double* xs = S.tab;
    double acc = 0.0;
    for (int c = start[node]; c < start[node + 1]; ++c) {
        const int s = contrib[c] >> 2, k = contrib[c] & 3;
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < NFQ; ++q) {
            // P1 trace basis of face node k at face point q (dimension 0 slowest)
            double psi = 1.0;
            int r = q;
#pragma unroll
            for (int a = TD - 1; a >= 0; --a) {
                const double sa = xs[r % N];
                r /= N;
                const int bit = (k >> (TD - 1 - a)) & 1;
                psi *= bit ? 0.5 * (1.0 + sa) : 0.5 * (1.0 - sa);
            }
            t += jump[((size_t)s * NFQ + q) * M + v] * psi;
        }
        acc -= t;
    }
    res[(size_t)node * M + v] = acc;
}

template <int D, int P, class PH>
int launch_mdg(const DevSpace& S, const PhysParams& ph, const ice_mdg_sel* sel, const double* U, double* res,
               hipStream_t st) {
    constexpr int NFQ = Pow<P + 1, D - 1>::value;
    constexpr int M = PH::M;
    const long nt = (long)sel->nface * NFQ;
    if (nt > 0)
        mdg_jump_kernel<D, P, PH><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(S, ph, U, sel->d_face, sel->nface,
                                                                               sel->d_jump);
    const long ng = (long)sel->nnode * M;
    if (ng > 0)
        mdg_gather_kernel<D, P, M><<<(unsigned)((ng + 255) / 256), 256, 0, st>>>(S, sel->d_start, sel->d_contrib,
                                                                                 sel->nnode, sel->d_jump, res);
    return (int)hipGetLastError();
}

// the interface error's jumps: every interior face, gradients kept
template <int D, int P, class PH>
int launch_ic_jumps(const DevSpace& S, const PhysParams& ph, const double* U, double* jump, hipStream_t st) {
    constexpr int NFQ = Pow<P + 1, D - 1>::value;
    const long nt = (long)S.nfi * NFQ;
    if (nt > 0)
        mdg_jump_kernel<D, P, PH, false><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(S, ph, U, nullptr, S.nfi, jump);
    return (int)hipGetLastError();
}

template <class PH, int D>
int ic_order(int p, const DevSpace& S, const PhysParams& ph, const double* U, double* jump, hipStream_t st) {
    switch (p) {
    case 1: return launch_ic_jumps<D, 1, PH>(S, ph, U, jump, st);
    case 2: return launch_ic_jumps<D, 2, PH>(S, ph, U, jump, st);
    case 3: return launch_ic_jumps<D, 3, PH>(S, ph, U, jump, st);
    case 4: return launch_ic_jumps<D, 4, PH>(S, ph, U, jump, st);
    default: return -1;
    }
}

int ic_dispatch(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const double* U, double* jump,
                hipStream_t st) {
    if (flux == ICE_FLUX_BURGERS) {
        if (ndim == 1) return ic_order<Burgers<1>, 1>(p, S, ph, U, jump, st);
        if (ndim == 2) return ic_order<Burgers<2>, 2>(p, S, ph, U, jump, st);
        if (ndim == 3) return ic_order<Burgers<3>, 3>(p, S, ph, U, jump, st);
    } else if (flux == ICE_FLUX_NAVIER_STOKES) {
        if (ndim == 1) return ic_order<NavierStokes<1, true>, 1>(p, S, ph, U, jump, st);
        if (ndim == 2) return ic_order<NavierStokes<2, true>, 2>(p, S, ph, U, jump, st);
        if (ndim == 3) return ic_order<NavierStokes<3, true>, 3>(p, S, ph, U, jump, st);
    } else if (flux == ICE_FLUX_EULER) {
        if (ndim == 1) return ic_order<NavierStokes<1, false>, 1>(p, S, ph, U, jump, st);
        if (ndim == 2) return ic_order<NavierStokes<2, false>, 2>(p, S, ph, U, jump, st);
        if (ndim == 3) return ic_order<NavierStokes<3, false>, 3>(p, S, ph, U, jump, st);
    }
    return -1;
}

template <class PH, int D>
int mdg_order(int p, const DevSpace& S, const PhysParams& ph, const ice_mdg_sel* sel, const double* U, double* res,
              hipStream_t st) {
    switch (p) {
    case 1: return launch_mdg<D, 1, PH>(S, ph, sel, U, res, st);
    case 2: return launch_mdg<D, 2, PH>(S, ph, sel, U, res, st);
    case 3: return launch_mdg<D, 3, PH>(S, ph, sel, U, res, st);
    case 4: return launch_mdg<D, 4, PH>(S, ph, sel, U, res, st);
    default: return -1;
    }
}

int mdg_dispatch(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const ice_mdg_sel* sel,
                 const double* U, double* res, hipStream_t st) {
    if (flux == ICE_FLUX_BURGERS) {
        if (ndim == 1) return mdg_order<Burgers<1>, 1>(p, S, ph, sel, U, res, st);
        if (ndim == 2) return mdg_order<Burgers<2>, 2>(p, S, ph, sel, U, res, st);
        if (ndim == 3) return mdg_order<Burgers<3>, 3>(p, S, ph, sel, U, res, st);
    } else if (flux == ICE_FLUX_NAVIER_STOKES) {
        if (ndim == 1) return mdg_order<NavierStokes<1, true>, 1>(p, S, ph, sel, U, res, st);
        if (ndim == 2) return mdg_order<NavierStokes<2, true>, 2>(p, S, ph, sel, U, res, st);
        if (ndim == 3) return mdg_order<NavierStokes<3, true>, 3>(p, S, ph, sel, U, res, st);
    } else if (flux == ICE_FLUX_EULER) {
        if (ndim == 1) return mdg_order<NavierStokes<1, false>, 1>(p, S, ph, sel, U, res, st);
        if (ndim == 2) return mdg_order<NavierStokes<2, false>, 2>(p, S, ph, sel, U, res, st);
        if (ndim == 3) return mdg_order<NavierStokes<3, false>, 3>(p, S, ph, sel, U, res, st);
    }
    return -1;
}

int fail(int code, const std::string& m) {
    set_error(m);
    return code;
}

// selected node s to xnew[s][:] in every element vertex slot it occupies
__global__ __launch_bounds__(256) void mdg_move_nodes_kernel(const double* __restrict__ xnew, int nnode, int ndim,
                                                             const int* __restrict__ vstart,
                                                             const int* __restrict__ vslot, double* __restrict__ X) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnode) return;
    for (int i = vstart[s]; i < vstart[s + 1]; ++i)
        for (int d = 0; d < ndim; ++d) X[(size_t)vslot[i] * ndim + d] = xnew[(size_t)s * ndim + d];
}

template <class T>
T* upload_vec(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)) != hipSuccess) throw Error{ICE_ERR_ALLOC, "hipMalloc failed"};
    if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        throw Error{ICE_ERR_HIP, "hipMemcpy failed"};
    }
    return d;
}

}  // namespace
}  // namespace ice

extern "C" int ice_mdg_select(const ice_space* space, const int* traces, int ntraces, ice_mdg_sel** out) {
    using namespace ice;
    if (!space || !out || ntraces < 0 || (ntraces > 0 && !traces)) return fail(ICE_ERR_ARG, "bad argument");
    *out = nullptr;
    const HostMesh& m = space_mesh(space);
    if (space_coupled(space) || space_batch(space) != 1)
        return fail(ICE_ERR_UNSUPPORTED, "MDG selections cover single meshes without coupled periodic faces");
    const int nface = (int)m.faces.size();
    for (int k = 0; k < ntraces; ++k)
        if (traces[k] < 0 || traces[k] >= nface) return fail(ICE_ERR_ARG, "trace index out of range");
    try {
        auto sel = std::make_unique<ice_mdg_sel>();
        sel->space = space;
        const int nfv = 1 << (m.ndim - 1), nv = 1 << m.ndim;
        // face nodes in the order the device face geometry uses (the left
        // element's face vertex order, face_vertex_order), which is the order
        // of the face's trace basis functions
        auto face_node = [&](int f, int j) {
            int fvo[4];
            face_vertex_order(m.ndim, m.faces[f].fnl, fvo);
            return m.conn[(size_t)m.faces[f].elemL * nv + fvo[j]];
        };
        std::vector<char> pick(m.nnode, 0);
        for (int k = 0; k < ntraces; ++k)
            for (int j = 0; j < nfv; ++j) pick[face_node(traces[k], j)] = 1;
        for (int f = m.n_interior; f < nface; ++f)
            for (int j = 0; j < nfv; ++j) pick[face_node(f, j)] = 0;
        std::vector<int> inv(m.nnode, -1);
        for (int nd = 0; nd < m.nnode; ++nd)
            if (pick[nd]) {
                inv[nd] = (int)sel->nodes.size();
                sel->nodes.push_back(nd);
            }
        sel->nnode = (int)sel->nodes.size();
        // interior faces of the selection (device face row = mesh face index
        // on meshes without coupled pairs), in selection order
        std::vector<int> rows;
        std::vector<std::vector<int>> lists(sel->nnode);
        for (int k = 0; k < ntraces; ++k) {
            const int f = traces[k];
            if (f >= m.n_interior) continue;
            const int s = (int)rows.size();
            rows.push_back(f);
            for (int j = 0; j < nfv; ++j) {
                const int nd = inv[face_node(f, j)];
                if (nd >= 0) lists[nd].push_back((s << 2) | j);
            }
        }
        sel->nface = (int)rows.size();
        std::vector<int> start(sel->nnode + 1, 0), contrib;
        for (int i = 0; i < sel->nnode; ++i) {
            start[i] = (int)contrib.size();
            contrib.insert(contrib.end(), lists[i].begin(), lists[i].end());
        }
        start[sel->nnode] = (int)contrib.size();
        int nfq = 1;
        for (int d = 0; d < m.ndim - 1; ++d) nfq *= space_dev(space).n;
        // element vertex slots of each selected node
        std::vector<std::vector<int>> vl(sel->nnode);
        for (int e = 0; e < m.nelem; ++e)
            for (int k = 0; k < nv; ++k) {
                const int s = inv[m.conn[(size_t)e * nv + k]];
                if (s >= 0) vl[s].push_back(e * nv + k);
            }
        std::vector<int> vstart(sel->nnode + 1, 0), vslot;
        for (int i = 0; i < sel->nnode; ++i) {
            vstart[i] = (int)vslot.size();
            vslot.insert(vslot.end(), vl[i].begin(), vl[i].end());
        }
        vstart[sel->nnode] = (int)vslot.size();
        sel->d_face = upload_vec(rows);
        sel->d_start = upload_vec(start);
        sel->d_contrib = upload_vec(contrib);
        sel->d_vstart = upload_vec(vstart);
        sel->d_vslot = upload_vec(vslot);
        if (hipMalloc(&sel->d_jump, std::max<size_t>((size_t)sel->nface * nfq * MDG_MAX_NEQ, 1) * sizeof(double)) !=
            hipSuccess)
            throw Error{ICE_ERR_ALLOC, "hipMalloc failed"};
        *out = sel.release();
        return ICE_OK;
    } catch (const Error& e) {
        return fail(e.code, e.msg);
    }
}

extern "C" int ice_mdg_query(const ice_mdg_sel* sel, int* nnodes, int* nodes) {
    if (!sel) return ice::fail(ICE_ERR_ARG, "null selection");
    if (nnodes) *nnodes = sel->nnode;
    if (nodes) std::copy(sel->nodes.begin(), sel->nodes.end(), nodes);
    return ICE_OK;
}

extern "C" int ice_mdg_residual(ice_disc* disc, const ice_mdg_sel* sel, const double* u_dev, double* res_dev,
                                void* stream) {
    using namespace ice;
    if (!disc || !sel || !u_dev || (sel->nnode > 0 && !res_dev)) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = disc_space(disc);
    if (s != sel->space) return fail(ICE_ERR_ARG, "the selection belongs to another space");
    const DevSpace& S = space_dev(s);
    const int rc = mdg_dispatch(S.ndim, S.p, disc_flux(disc), S, disc_phys(disc), sel, u_dev, res_dev,
                                (hipStream_t)stream);
    if (rc != 0) return fail(ICE_ERR_HIP, std::string("MDG residual launch failed: ") + hipGetErrorString((hipError_t)rc));
    return ICE_OK;
}

extern "C" int ice_mdg_form_residual(ice_disc* disc, const ice_mdg_sel* sel, const double* u_dev, double* res_dev,
                                     void* stream) {
    using namespace ice;
    if (!disc || !sel || !u_dev || !res_dev) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = disc_space(disc);
    if (s != sel->space) return fail(ICE_ERR_ARG, "the selection belongs to another space");
    const DevSpace& S = space_dev(s);
    hipStream_t st = (hipStream_t)stream;
    int nel[3] = {}, neq = 0, nb = 0;  // elements per direction, equations, basis functions
    disc_grid(disc, nel, &neq, &nb);
    const size_t ndg = (size_t)S.nelem * nb * neq;
    int rc = 0;
    if (sel->nnode > 0) {
        // the vertex table is the space's own device array
        mdg_move_nodes_kernel<<<(sel->nnode + 255) / 256, 256, 0, st>>>(
            u_dev + ndg, sel->nnode, S.ndim, sel->d_vstart, sel->d_vslot, const_cast<double*>(S.X));
        rc = (int)hipGetLastError();
        if (rc == 0) rc = disc_refresh_geometry(disc, st);
    }
    if (rc == 0) rc = disc_residual(disc, u_dev, res_dev, st);
    if (rc == 0)
        rc = mdg_dispatch(S.ndim, S.p, disc_flux(disc), S, disc_phys(disc), sel, u_dev, res_dev + ndg, st);
    if (rc != 0)
        return fail(ICE_ERR_HIP, std::string("MDG form_residual launch failed: ") + hipGetErrorString((hipError_t)rc));
    return ICE_OK;
}

extern "C" int ice_interface_error(ice_disc* disc, const double* u_dev, double* err, void* stream) {
    using namespace ice;
    if (!disc || !u_dev || !err) return fail(ICE_ERR_ARG, "null argument");
    const ice_space* s = disc_space(disc);
    const DevSpace& S = space_dev(s);
    hipStream_t st = (hipStream_t)stream;
    int nel[3] = {}, neq = 0, nb = 0;
    disc_grid(disc, nel, &neq, &nb);
    int nfq = 1;
    for (int d = 0; d < S.ndim - 1; ++d) nfq *= S.n;
    const size_t n = (size_t)S.nfi * nfq * neq;
    double* jump = nullptr;
    if (hipMalloc(&jump, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess)
        return fail(ICE_ERR_ALLOC, "allocation failed");
    std::vector<double> h(n);
    int rc = ic_dispatch(S.ndim, S.p, disc_flux(disc), S, disc_phys(disc), u_dev, jump, st);
    if (rc == 0 && n > 0) rc = (int)hipMemcpyAsync(h.data(), jump, n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (rc == 0) rc = (int)hipStreamSynchronize(st);
    (void)hipFree(jump);
    if (rc != 0) return fail(ICE_ERR_HIP, "interface error launch failed");
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += std::fabs(h[i]);
    *err = sum;
    return ICE_OK;
}

extern "C" int ice_mdg_destroy(ice_mdg_sel* sel) {
    delete sel;
    return ICE_OK;
}
