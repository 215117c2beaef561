// CPU ORACLE (test infrastructure only; never linked into the product).
//
// Uniform hypercube mesh, face enumeration and trace maps restated from the
// reference: AbstractMesh uniform constructor (mesh/mesh.hpp:553-913),
// generate_directional_nodes (mesh.hpp:238-268), HypercubeElementTransformation
// face utilities (transformations/HypercubeTransformations.hpp:408-984),
// HypercubeTraceOrientTransformation (:995-1150), HypercubeTraceTransformation
// (:1152-1456), HypercubeFace (geometry/hypercube_face.hpp:179-274).
#pragma once
#include <array>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "basis.hpp"

namespace oracle {

enum BC : int {
    BC_PERIODIC = 0, BC_PARALLEL_COM, BC_NEUMANN, BC_DIRICHLET, BC_EXTRAPOLATION,
    BC_RIEMANN, BC_NO_SLIP_ISOTHERMAL, BC_SLIP_WALL, BC_WALL_GENERAL, BC_INLET,
    BC_OUTLET, BC_SPACETIME_PAST, BC_SPACETIME_FUTURE, BC_INTERIOR
};

// first_dim_sign (HypercubeTransformations.hpp:176-209): Levi-Civita sign of
// (trace_coord, remaining axes ascending), negated on negative faces.
template <int D>
inline double first_dim_sign(int face_nr) {
    if constexpr (D == 1) return 1.0;
    const int c = face_nr % D;
    const bool neg = face_nr / D == 0;
    double lc = (c % 2 == 0) ? 1.0 : -1.0;  // parity of moving axis c to the front
    return neg ? -lc : lc;
}

// get_face_nodes for geometry order 1 (HypercubeTransformations.hpp:830-890):
// element-local vertex indices of face `face_nr` in face-node order.
template <int D>
inline std::array<int, (D > 1 ? (1 << (D - 1)) : 1)> face_local_nodes(int face_nr) {
    constexpr int NFV = (D > 1 ? (1 << (D - 1)) : 1);
    std::array<int, NFV> out{};
    const int c = face_nr % D;
    const bool neg = face_nr / D == 0;
    const int tfd = (c == 0) ? 1 : 0;
    const double sgn = first_dim_sign<D>(face_nr);
    std::array<int, D> ijk{};
    if (sgn < 0 && D > 1) ijk[tfd] = 1;
    ijk[c] = neg ? 0 : 1;
    auto convert = [](const std::array<int, D>& a) {
        int r = 0;
        for (int d = 0; d < D; ++d) r = r * 2 + a[d];
        return r;
    };
    auto next = [&](std::array<int, D>& a) -> bool {
        for (int d = D - 1; d >= 0; --d) {
            if (d == c) continue;
            if (d == tfd && sgn < 0) {
                if (a[d] == 0) a[d] = 1;
                else { a[d]--; return true; }
            } else {
                if (a[d] == 1) a[d] = 0;
                else { a[d]++; return true; }
            }
        }
        return false;
    };
    int k = 0;
    do { out[k++] = convert(ijk); } while (next(ijk) && k < NFV);
    return out;
}

// Lexicographic permutations of {0..n-1} (HypercubeTraceOrientTransformation ctor).
inline std::vector<std::vector<int>> axis_permutations(int n) {
    std::vector<std::vector<int>> out;
    std::vector<int> p(n);
    for (int i = 0; i < n; ++i) p[i] = i;
    do { out.push_back(p); } while (std::next_permutation(p.begin(), p.end()));
    return out;
}

// getOrientation (HypercubeTransformations.hpp:1063-1112)
template <int D>
inline int get_orientation(const int* vL, const int* vR) {
    constexpr int TD = D - 1;
    if constexpr (TD <= 0) return 0;
    else {
        constexpr int NV = 1 << TD;
        int sign_code = 0;
        std::vector<int> left_dirs(TD);
        const int g0 = vR[0];
        for (int idim = 0; idim < TD; ++idim) {
            const int g1 = vR[1 << (TD - idim - 1)];
            int l0 = -1, l1 = -1;
            for (int iv = 0; iv < NV; ++iv) {
                if (vL[iv] == g0) l0 = iv;
                else if (vL[iv] == g1) l1 = iv;
                if (l0 >= 0 && l1 >= 0) break;
            }
            if (l0 < 0 || l1 < 0) return -1;
            if (l1 < l0) sign_code |= (1 << idim);
            const int diff = std::abs(l1 - l0);
            int lg = 0;
            while ((1 << lg) < diff) ++lg;
            left_dirs[idim] = TD - lg - 1;
        }
        auto perms = axis_permutations(TD);
        int iperm = 0;
        for (size_t i = 0; i < perms.size(); ++i)
            if (perms[i] == left_dirs) { iperm = (int)i; break; }
        return iperm * NV + sign_code;
    }
}

// HypercubeTraceOrientTransformation::transform (:1128-1148)
template <int D>
inline void orient_transform(int orient, const double* s, double* sR) {
    constexpr int TD = D - 1;
    if constexpr (TD <= 0) return;
    else {
        constexpr int NV = 1 << TD;
        const int iperm = orient / NV, sign = orient % NV;
        auto perms = axis_permutations(TD);
        for (int i = 0; i < TD; ++i) {
            const double v = s[perms[iperm][i]];
            sR[i] = (sign & (1 << i)) ? -v : v;
        }
    }
}

// HypercubeTraceTransformation::transform (:1215-1235) and the 1D
// specialisation (:1370-1380): trace coordinates -> element reference point.
template <int D>
inline void trace_to_ref(int face_nr, const double* s, double* xi) {
    if constexpr (D == 1) {
        xi[0] = (face_nr == 0) ? -1.0 : 1.0;
    } else {
        const int c = face_nr % D;
        const bool neg = face_nr / D == 0;
        xi[c] = neg ? -1.0 : 1.0;
        for (int d = 0; d < c; ++d) xi[d] = s[d];
        for (int d = c + 1; d < D; ++d) xi[d] = s[d - 1];
        if (c == 0) xi[1] *= first_dim_sign<D>(face_nr);
        else xi[0] *= first_dim_sign<D>(face_nr);
    }
}

template <int D>
struct Mesh {
    static constexpr int NV = 1 << D;
    static constexpr int NFV = (D > 1 ? (1 << (D - 1)) : 1);
    struct Face {
        int elemL, elemR, fnl, fnr, orient, bctype, bcflag;
        std::array<int, NFV> nodes;
    };
    std::array<int, D> nel{};
    std::array<int, D> nnode_dir{};
    std::array<double, D> period{};
    int nelem = 0, nnode = 0;
    std::vector<std::array<double, D>> coord;
    std::vector<std::array<int, NV>> conn;
    std::vector<Face> faces;
    int n_interior = 0;

    int node_index(const std::array<int, D>& ijk) const {
        int r = 0, s = 1;
        for (int d = 0; d < D; ++d) { r += ijk[d] * s; s *= nnode_dir[d]; }
        return r;
    }
    std::array<int, D> node_ijk(int n) const {
        std::array<int, D> ijk{};
        for (int d = 0; d < D; ++d) { ijk[d] = n % nnode_dir[d]; n /= nnode_dir[d]; }
        return ijk;
    }
};

// Uniform mesh constructor.  `fix3d_boundary_order` repairs the reference's
// ordinate increment in the boundary-face loop (mesh.hpp:879-890), which in 3D
// carries into the boundary-normal axis when idim == 1 and so duplicates some
// y-boundary faces and omits others; in 1D and 2D both variants are identical.
template <int D>
Mesh<D> build_uniform_mesh(const int* nelem, const double* xmin, const double* xmax,
                           const int* bctypes, const int* bcflags,
                           const double* coord_override, bool fix3d_boundary_order = true) {
    using M = Mesh<D>;
    M mesh;
    for (int d = 0; d < D; ++d) {
        mesh.nel[d] = nelem[d];
        mesh.nnode_dir[d] = nelem[d] + 1;
        mesh.period[d] = xmax[d] - xmin[d];
    }
    // generate_directional_nodes (mesh.hpp:238-268): cumulative dx sums
    std::array<std::vector<double>, D> nodes1d;
    for (int d = 0; d < D; ++d) {
        const double dx = (xmax[d] - xmin[d]) / (nelem[d] * 1);
        nodes1d[d].resize(nelem[d] + 1);
        nodes1d[d][0] = xmin[d];
        for (int i = 0; i < nelem[d]; ++i) nodes1d[d][i + 1] = nodes1d[d][i] + dx;
    }
    mesh.nnode = 1;
    mesh.nelem = 1;
    for (int d = 0; d < D; ++d) { mesh.nnode *= mesh.nnode_dir[d]; mesh.nelem *= nelem[d]; }
    mesh.coord.resize(mesh.nnode);
    for (int n = 0; n < mesh.nnode; ++n) {
        auto ijk = mesh.node_ijk(n);
        for (int d = 0; d < D; ++d)
            mesh.coord[n][d] = coord_override ? coord_override[n * D + d] : nodes1d[d][ijk[d]];
    }
    // element strides (x fastest) and connectivity via ijk_poin (mesh.hpp:623-661)
    std::array<int, D> stride{};
    for (int d = 0; d < D; ++d) { stride[d] = 1; for (int j = 0; j < d; ++j) stride[d] *= nelem[j]; }
    mesh.conn.resize(mesh.nelem);
    for (int e = 0; e < mesh.nelem; ++e) {
        std::array<int, D> eijk{};
        int r = e;
        for (int d = 0; d < D; ++d) { eijk[d] = r % nelem[d]; r /= nelem[d]; }
        for (int k = 0; k < M::NV; ++k) {
            auto lijk = tensor_ijk<D>(k, 2);
            std::array<int, D> g{};
            for (int d = 0; d < D; ++d) g[d] = eijk[d] + lijk[d];
            mesh.conn[e][k] = mesh.node_index(g);
        }
    }
    auto face_nodes = [&](int e, int fnr) {
        std::array<int, M::NFV> out{};
        auto loc = face_local_nodes<D>(fnr);
        for (int k = 0; k < M::NFV; ++k) out[k] = mesh.conn[e][loc[k]];
        return out;
    };
    // interior faces (mesh.hpp:679-753)
    for (int idir = 0; idir < D; ++idir) {
        std::array<int, D> ijk{};
        auto next_ijk = [&](std::array<int, D>& a) -> bool {
            for (int d = 0; d < D; ++d) {
                if (d == idir) {
                    if (a[d] >= nelem[d] - 2) a[d] = 0;
                    else { a[d]++; return true; }
                } else {
                    if (a[d] >= nelem[d] - 1) a[d] = 0;
                    else { a[d]++; return true; }
                }
            }
            return false;
        };
        if (nelem[idir] > 1) do {
            auto ijkr = ijk;
            ijkr[idir]++;
            int iel = 0, ier = 0;
            for (int d = 0; d < D; ++d) { iel += ijk[d] * stride[d]; ier += ijkr[d] * stride[d]; }
            typename M::Face f{};
            f.elemL = iel; f.elemR = ier;
            f.fnl = D + idir; f.fnr = idir;
            f.nodes = face_nodes(iel, f.fnl);
            auto vr = face_nodes(ier, f.fnr);
            f.orient = get_orientation<D>(f.nodes.data(), vr.data());
            f.bctype = BC_INTERIOR; f.bcflag = 0;
            mesh.faces.push_back(f);
        } while (next_ijk(ijk));
    }
    mesh.n_interior = (int)mesh.faces.size();
    // boundary faces (mesh.hpp:759-893)
    for (int idim = 0; idim < D; ++idim) {
        int nbfac = 1;
        for (int j = 0; j < D; ++j) if (j != idim) nbfac *= nelem[j];
        std::array<int, D> ijk{};
        for (int ifac = 0; ifac < nbfac; ++ifac) {
            int iel = 0;
            for (int d = 0; d < D; ++d) iel += ijk[d] * stride[d];
            typename M::Face fa{};
            fa.elemL = iel; fa.elemR = -1; fa.fnl = idim; fa.fnr = 0; fa.orient = 0;
            fa.nodes = face_nodes(iel, fa.fnl);
            fa.bctype = bctypes[idim]; fa.bcflag = bcflags[idim];
            ijk[idim] = nelem[idim] - 1;
            iel = 0;
            for (int d = 0; d < D; ++d) iel += ijk[d] * stride[d];
            typename M::Face fb{};
            fb.elemL = iel; fb.elemR = -1; fb.fnl = idim + D; fb.fnr = 0; fb.orient = 0;
            fb.nodes = face_nodes(iel, fb.fnl);
            fb.bctype = bctypes[D + idim]; fb.bcflag = bcflags[D + idim];
            if (bctypes[idim] == BC_PERIODIC) {
                const int ia = (int)mesh.faces.size();
                fa.bcflag = ia + 1;
                fb.bcflag = ia;
            }
            mesh.faces.push_back(fa);
            mesh.faces.push_back(fb);
            ijk[idim] = 0;
            if (fix3d_boundary_order) {
                for (int d = 0; d < D; ++d) {
                    if (d == idim) continue;
                    if (++ijk[d] < nelem[d]) break;
                    ijk[d] = 0;
                }
            } else {
                const int first = (idim == 0) ? 1 : 0;
                if (D > 1) ++ijk[first];
                for (int j = first; j < D - 1; ++j) {
                    if (j == idim) {
                    } else if (ijk[j] == nelem[j]) {
                        ijk[j] = 0;
                        ++ijk[j + 1];
                    } else break;
                }
            }
        }
    }
    return mesh;
}

}  // namespace oracle
