// Host-side structured mesh, face list and trace-point tables (see mesh.hpp).
#include "mesh.hpp"

#include <algorithm>
#include <cmath>

namespace ice {

namespace {

inline int vertex_index(int ndim, const int* ijk) {
    int r = 0;
    for (int d = 0; d < ndim; ++d) r = 2 * r + ijk[d];
    return r;
}

// lexicographic rank of a permutation of {0..n-1}
int perm_rank(const int* p, int n) {
    int rank = 0;
    for (int i = 0; i < n; ++i) {
        int smaller = 0;
        for (int j = i + 1; j < n; ++j)
            if (p[j] < p[i]) ++smaller;
        int f = 1;
        for (int k = 2; k <= n - 1 - i; ++k) f *= k;
        rank += smaller * f;
    }
    return rank;
}

// permutation of {0..n-1} with lexicographic rank r
void perm_unrank(int r, int n, int* p) {
    int pool[3] = {0, 1, 2};
    int m = n;
    for (int i = 0; i < n; ++i) {
        int f = 1;
        for (int k = 2; k <= n - 1 - i; ++k) f *= k;
        const int k = r / f;
        r %= f;
        p[i] = pool[k];
        for (int j = k; j < m - 1; ++j) pool[j] = pool[j + 1];
        --m;
    }
}

inline int gl_mirror(int n, int k) { return (n % 2 == 1 && k == n - 1) ? k : (k ^ 1); }

}  // namespace

int face_first_sign(int ndim, int j) {
    if (ndim == 1) return 1;
    const int c = j % ndim;
    const int lc = (c % 2 == 0) ? 1 : -1;
    return (j < ndim) ? -lc : lc;
}

void face_vertex_order(int ndim, int j, int* out) {
    if (ndim == 1) { out[0] = (j == 0) ? 0 : 1; return; }
    const int c = j % ndim;
    const int tfd = (c == 0) ? 1 : 0;
    const bool rev = face_first_sign(ndim, j) < 0;
    const int td = ndim - 1;
    int axes[2], na = 0;
    for (int d = 0; d < ndim; ++d)
        if (d != c) axes[na++] = d;
    for (int t = 0; t < (1 << td); ++t) {
        int ijk[3] = {0, 0, 0};
        ijk[c] = (j < ndim) ? 0 : 1;
        for (int k = 0; k < td; ++k) {
            int bit = (t >> (td - 1 - k)) & 1;
            if (axes[k] == tfd && rev) bit = 1 - bit;
            ijk[axes[k]] = bit;
        }
        out[t] = vertex_index(ndim, ijk);
    }
}

int n_orientations(int ndim) {
    const int td = ndim - 1;
    if (td <= 0) return 1;
    int f = 1;
    for (int k = 2; k <= td; ++k) f *= k;
    return f << td;
}

int face_orientation(int ndim, const int* vl, const int* vr) {
    const int td = ndim - 1;
    if (td <= 0) return 0;
    const int nv = 1 << td;
    int sign = 0, dirs[2] = {0, 0};
    for (int i = 0; i < td; ++i) {
        const int g1 = vr[1 << (td - 1 - i)];
        int l0 = -1, l1 = -1;
        for (int k = 0; k < nv; ++k) {
            if (vl[k] == vr[0]) l0 = k;
            else if (vl[k] == g1) l1 = k;
        }
        if (l0 < 0 || l1 < 0) return -1;
        if (l1 < l0) sign |= 1 << i;
        const int diff = std::abs(l1 - l0);
        int lg = 0;
        while ((1 << lg) < diff) ++lg;
        dirs[i] = td - 1 - lg;
    }
    return perm_rank(dirs, td) * nv + sign;
}

void trace_point_indices(int ndim, int n, int face_nr, int orient, int fq, int* out) {
    const int td = ndim - 1;
    int s[2] = {0, 0}, sr[2] = {0, 0};
    for (int k = td - 1; k >= 0; --k) { s[k] = fq % n; fq /= n; }
    if (td > 0) {
        const int nv = 1 << td;
        int perm[2] = {0, 1};
        perm_unrank(orient / nv, td, perm);
        const int sign = orient % nv;
        for (int i = 0; i < td; ++i) sr[i] = (sign >> i & 1) ? gl_mirror(n, s[perm[i]]) : s[perm[i]];
    }
    const int c = face_nr % ndim;
    out[c] = (face_nr < ndim) ? n : n + 1;  // S index of -1 / +1
    int k = 0;
    for (int d = 0; d < ndim; ++d)
        if (d != c) out[d] = sr[k++];
    if (ndim > 1 && face_first_sign(ndim, face_nr) < 0) {
        const int a = (c == 0) ? 1 : 0;
        out[a] = gl_mirror(n, out[a]);
    }
}

HostMesh build_mesh(const ice_mesh_desc& d, bool coupled) {
    HostMesh m;
    if (d.ndim < 1 || d.ndim > 3) throw Error{ICE_ERR_ARG, "ndim must be 1, 2 or 3"};
    const int D = d.ndim;
    m.ndim = D;
    m.nelem = 1;
    m.nnode = 1;
    for (int k = 0; k < D; ++k) {
        if (d.nelem[k] < 1) throw Error{ICE_ERR_ARG, "nelem must be >= 1 in every direction"};
        if (!(d.xmax[k] > d.xmin[k])) throw Error{ICE_ERR_ARG, "xmax must exceed xmin"};
        m.nel[k] = d.nelem[k];
        m.nnd[k] = d.nelem[k] + 1;
        m.nelem *= m.nel[k];
        m.nnode *= m.nnd[k];
        m.period[k] = d.xmax[k] - d.xmin[k];
    }
    for (int k = 0; k < 2 * D; ++k) {
        if (d.bctype[k] < 0 || d.bctype[k] > ICE_BC_INTERIOR) throw Error{ICE_ERR_ARG, "bad boundary condition code"};
        if (d.bctype[k] == ICE_BC_PERIODIC && d.bctype[(k + D) % (2 * D)] != ICE_BC_PERIODIC)
            throw Error{ICE_ERR_ARG, "periodic boundaries must come in opposite pairs"};
    }
    // node coordinates: per-axis cumulative spacing (generate_directional_nodes)
    std::vector<double> axis[3];
    for (int k = 0; k < D; ++k) {
        const double dx = (d.xmax[k] - d.xmin[k]) / m.nel[k];
        axis[k].resize(m.nnd[k]);
        axis[k][0] = d.xmin[k];
        for (int i = 0; i < m.nel[k]; ++i) axis[k][i + 1] = axis[k][i] + dx;
    }
    m.coord.resize((size_t)m.nnode * D);
    for (int nd = 0; nd < m.nnode; ++nd) {
        int r = nd;
        for (int k = 0; k < D; ++k) {
            const int i = r % m.nnd[k];
            r /= m.nnd[k];
            m.coord[(size_t)nd * D + k] = d.coord ? d.coord[(size_t)nd * D + k] : axis[k][i];
        }
    }
    const int NV = 1 << D;
    const int NFV = (D > 1) ? (1 << (D - 1)) : 1;
    auto node_id = [&](const int* ijk) {
        int r = 0, s = 1;
        for (int k = 0; k < D; ++k) { r += ijk[k] * s; s *= m.nnd[k]; }
        return r;
    };
    auto elem_id = [&](const int* ijk) {
        int r = 0, s = 1;
        for (int k = 0; k < D; ++k) { r += ijk[k] * s; s *= m.nel[k]; }
        return r;
    };
    m.conn.resize((size_t)m.nelem * NV);
    for (int e = 0; e < m.nelem; ++e) {
        int eijk[3], r = e;
        for (int k = 0; k < D; ++k) { eijk[k] = r % m.nel[k]; r /= m.nel[k]; }
        for (int v = 0; v < NV; ++v) {
            int g[3];
            for (int k = 0; k < D; ++k) g[k] = eijk[k] + ((v >> (D - 1 - k)) & 1);
            m.conn[(size_t)e * NV + v] = node_id(g);
        }
    }
    int fvo[2 * 3][MAX_NFV];
    for (int j = 0; j < 2 * D; ++j) face_vertex_order(D, j, fvo[j]);
    auto fill_nodes = [&](int e, int j, int* out) {
        for (int k = 0; k < NFV; ++k) out[k] = m.conn[(size_t)e * NV + fvo[j][k]];
    };
    // interior faces, direction by direction, left element ordinates with axis 0 fastest
    for (int idir = 0; idir < D; ++idir) {
        if (m.nel[idir] < 2) continue;
        int range[3];
        long total = 1;
        for (int k = 0; k < D; ++k) { range[k] = (k == idir) ? m.nel[k] - 1 : m.nel[k]; total *= range[k]; }
        for (long t = 0; t < total; ++t) {
            int ijk[3], r = (int)t;
            for (int k = 0; k < D; ++k) { ijk[k] = r % range[k]; r /= range[k]; }
            int ijr[3] = {ijk[0], ijk[1], ijk[2]};
            ijr[idir] += 1;
            HostFace f{};
            f.elemL = elem_id(ijk);
            f.elemR = elem_id(ijr);
            f.fnl = D + idir;
            f.fnr = idir;
            fill_nodes(f.elemL, f.fnl, f.nodes);
            int vr[MAX_NFV];
            fill_nodes(f.elemR, f.fnr, vr);
            f.orient = face_orientation(D, f.nodes, vr);
            f.bctype = ICE_BC_INTERIOR;
            f.bcflag = 0;
            m.faces.push_back(f);
        }
    }
    m.n_interior = (int)m.faces.size();
    // boundary faces: per axis, remaining ordinates with the lowest axis fastest, A (-) then B (+)
    for (int idim = 0; idim < D; ++idim) {
        int range[3];
        long total = 1;
        for (int k = 0; k < D; ++k) { range[k] = (k == idim) ? 1 : m.nel[k]; total *= range[k]; }
        for (long t = 0; t < total; ++t) {
            int ijk[3], r = (int)t;
            for (int k = 0; k < D; ++k) { ijk[k] = r % range[k]; r /= range[k]; }
            HostFace fa{}, fb{};
            ijk[idim] = 0;
            fa.elemL = elem_id(ijk);
            fa.elemR = -1;
            fa.fnl = idim;
            fill_nodes(fa.elemL, fa.fnl, fa.nodes);
            fa.bctype = d.bctype[idim];
            fa.bcflag = d.bcflag[idim];
            ijk[idim] = m.nel[idim] - 1;
            fb.elemL = elem_id(ijk);
            fb.elemR = -1;
            fb.fnl = idim + D;
            fill_nodes(fb.elemL, fb.fnl, fb.nodes);
            fb.bctype = d.bctype[idim + D];
            fb.bcflag = d.bcflag[idim + D];
            if (d.bctype[idim] == ICE_BC_PERIODIC) {
                const int ia = (int)m.faces.size();
                fa.bcflag = ia + 1;
                fb.bcflag = ia;
            }
            m.faces.push_back(fa);
            m.faces.push_back(fb);
        }
    }
    (void)coupled;
    return m;
}

}  // namespace ice
