// Host-side mesh and trace-map construction for the device engine.
//
// Produces the same node, element and face order as the reference's uniform
// AbstractMesh constructor (mesh/mesh.hpp:553-913) and the same trace
// conventions as HypercubeElementTransformation / HypercubeTraceOrientTransformation
// (transformations/HypercubeTransformations.hpp:408-1150), but computes them
// from closed-form structured-grid index formulas and small lookup tables
// instead of the reference's per-face virtual transformations.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/iceicle_amd.h"

namespace ice {

struct Error {
    int code;
    std::string msg;
};

constexpr int MAX_NV = 8;    // 2^3 element vertices
constexpr int MAX_NFV = 4;   // 2^2 face vertices

struct HostFace {
    int elemL, elemR;
    int fnl, fnr, orient;
    int bctype, bcflag;
    int pdim;                 // coupled periodic pair: axis + 1, else 0
    int nodes[MAX_NFV];
};

struct HostMesh {
    int ndim = 0;
    int nel[3] = {1, 1, 1};
    int nnd[3] = {1, 1, 1};   // nodes per direction
    int nelem = 0, nnode = 0;
    double period[3] = {0, 0, 0};
    std::vector<double> coord;        // [nnode][ndim]
    std::vector<int> conn;            // [nelem][2^ndim], element-local tensor order
    std::vector<HostFace> faces;      // reference order: interior, then boundary
    int n_interior = 0;
};

// Sign that flips the first trace axis so that face maps keep outward normals
// (first_dim_sign, HypercubeTransformations.hpp:176-209).
int face_first_sign(int ndim, int face_nr);

// Element-local vertex indices of face `face_nr` in face-node order
// (get_face_nodes / get_face_vert for geometry order 1, :830-946).
void face_vertex_order(int ndim, int face_nr, int* out);

// Orientation code of the right face given both faces' vertex ids
// (getOrientation, :1063-1112); -1 if the vertex sets differ.
int face_orientation(int ndim, const int* vl, const int* vr);

// Build the mesh; throws ice::Error on bad input.
HostMesh build_mesh(const ice_mesh_desc& d, bool coupled_periodic);

// Trace point tables: every reference point used by the engine is a tensor
// product of 1D points from S = {GL_0..GL_{n-1}, -1, +1}.  For a face number j,
// orientation o and face quadrature point q, write the S index of each axis of
// the element-side reference point (HypercubeTraceTransformation::transform
// after HypercubeTraceOrientTransformation::transform).
void trace_point_indices(int ndim, int n, int face_nr, int orient, int fq, int* out);

// Number of orientation codes for ndim (trace_ndim! * 2^trace_ndim).
int n_orientations(int ndim);

}  // namespace ice
