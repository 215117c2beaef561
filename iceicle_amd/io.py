"""On-disk formats of the reference: restart files and field output.

* write_restart / read_restart: the text restart format of
  fe_function/restart.hpp:22-110 (file RESTART/restart<k>_<rank>: the mesh
  node coordinates, then every coefficient u[e, i, v] in fe_layout_right
  order, both at 16 significant digits, space separated).
* DatWriter: the 1D gnuplot output of io/dat_writer.hpp (per element 30
  equispaced points of each field set, and its two end points).
* PVDWriter.write_vtu / write_mesh: the discontinuous (one point set per
  element) ASCII VTU of io/pvd_writer.hpp:985-1186 with the header of
  io/pvd_writer.cpp:33-47, for 2D quads and 3D hexes.

Everything here runs on the host from the mesh built by ice_mesh_export
(the same node and element order as the reference's AbstractMesh) and a
host copy of the coefficients; a device tensor is copied to the host first.
"""
import os

import numpy as np

from . import engine


def _host(u):
    """Host float64 copy of a numpy array or a (device) torch tensor."""
    if hasattr(u, "detach"):
        u = u.detach().cpu().numpy()
    return np.asarray(u, dtype=np.float64)


class HostGeometry:
    """Element vertices and the Q1 map x(xi) of every element, the tensor
    Lagrange basis of order p, both with dimension 0 slowest
    (tensor_product.hpp:321-638, HypercubeTransformations.hpp:87-112)."""

    def __init__(self, mesh, order):
        coord, conn, _, cnt = engine.mesh_export(mesh)
        if cnt.nvert_el != 2 ** cnt.ndim:
            raise ValueError("geometry order 1 hypercube meshes only")
        self.ndim, self.order = cnt.ndim, int(order)
        self.coord, self.conn = coord, conn
        self.X = coord[conn]  # [nelem][2^d][d]
        self.nelem = cnt.nelem
        self.nbasis = (self.order + 1) ** self.ndim

    def _tensor(self, xi, rows):
        """Tensor product of per-dimension 1D values rows[d](x) [npts][n]."""
        xi = np.atleast_2d(np.asarray(xi, dtype=np.float64))
        out = np.ones((xi.shape[0], 1))
        for d in range(self.ndim):
            r = rows(xi[:, d])
            out = (out[:, :, None] * r[:, None, :]).reshape(xi.shape[0], -1)
        return out

    def q1_shape(self, xi):
        return self._tensor(xi, lambda s: np.stack([(1.0 - s) / 2.0, (1.0 + s) / 2.0], axis=1))

    def transform(self, xi):
        """Physical points of reference points xi [npts][d] in every element: [nelem][npts][d]."""
        return np.einsum("pn,end->epd", self.q1_shape(xi), self.X)

    def det_jacobian(self, xi):
        """det(dx/dxi) of the Q1 map at reference points xi in every element: [nelem][npts]."""
        xi = np.atleast_2d(np.asarray(xi, dtype=np.float64))
        npts, D = xi.shape[0], self.ndim
        J = np.empty((self.nelem, npts, D, D))
        for a in range(D):
            sh = np.ones((npts, 1))
            for d in range(D):
                s = xi[:, d]
                r = (np.stack([np.full_like(s, -0.5), np.full_like(s, 0.5)], axis=1) if d == a
                     else np.stack([(1.0 - s) / 2.0, (1.0 + s) / 2.0], axis=1))
                sh = (sh[:, :, None] * r[:, None, :]).reshape(npts, -1)
            J[:, :, :, a] = np.einsum("pn,end->epd", sh, self.X)
        return np.linalg.det(J)

    def basis(self, xi):
        """Lagrange basis on the equispaced nodes -1 + 2j/p (lagrange_1d.hpp:19-87) at xi: [npts][nbasis]."""
        p = self.order
        nodes = np.array([-1.0 + 2.0 * j / p for j in range(p + 1)])

        def rows(s):
            r = np.ones((s.shape[0], p + 1))
            for j in range(p + 1):
                for k in range(p + 1):
                    if k != j:
                        r[:, j] *= (s - nodes[k]) / (nodes[j] - nodes[k])
            return r

        return self._tensor(xi, rows)

    def evaluate(self, u, neq, xi):
        """Field values at reference points xi of every element: [nelem][npts][neq]."""
        U = _host(u).reshape(self.nelem, self.nbasis, neq)
        return np.einsum("pi,eiv->epv", self.basis(xi), U)


def _g(v, prec):
    """A double as an ostream with setprecision(prec) prints it (%g)."""
    return f"{v:.{prec}g}"


def write_restart(geo, u, k, directory=None, rank=0):
    """write_restart (restart.hpp:22-61): RESTART/restart<k>_<rank> under
    `directory` (default: the working directory).  Returns the path."""
    d = os.path.join(directory or os.getcwd(), "RESTART")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"restart{k}_{rank}")
    U = _host(u).ravel()
    with open(path, "w") as fh:
        fh.write("".join(_g(x, 16) + " " for x in geo.coord.ravel()))
        fh.write("\n")
        fh.write("".join(_g(x, 16) + " " for x in U))
    return path


def read_restart(name, ndim, nnode, size, directory=None, rank=0):
    """read_restart (restart.hpp:70-110): the node coordinates [nnode][ndim]
    (which replace the mesh's) and the coefficients (size values) of
    RESTART/<name>_<rank>."""
    path = os.path.join(directory or os.getcwd(), "RESTART", f"{name}_{rank}")
    with open(path) as fh:
        vals = np.array(fh.read().split(), dtype=np.float64)
    if vals.size != nnode * ndim + size:
        raise ValueError(f"{path}: {vals.size} values, expected {nnode * ndim} node coordinates + {size} coefficients")
    return vals[:nnode * ndim].reshape(nnode, ndim), vals[nnode * ndim:]


class DatWriter:
    """io/dat_writer.hpp: 1D field sets as gnuplot text.  Each registered
    field set writes <collection>_fieldset<i>_rank<r>_i<itime>.dat (30
    equispaced points per element, a blank line after each element) and
    <collection>_endpoints<i>_rank<r>_i<itime>.dat (the element ends)."""

    WIDTH, PREC = 18, 10

    def __init__(self, geo, directory=None, collection_name="iceicle_data"):
        if geo.ndim != 1:
            raise ValueError("the .dat writer covers 1D meshes")
        self.geo = geo
        self.collection_name = collection_name
        self.data_directory = os.path.join(directory or os.getcwd(), "iceicle_data")
        self.fields = []

    def register_fields(self, u, neq, names=None):
        names = list(names) if names is not None else []
        if len(names) != neq:  # dat_writer.hpp:213-220
            names = [f"u{i}" for i in range(neq)]
        self.fields.append((u, neq, names))

    def _write(self, path, u, neq, names, npoin):
        W, P = self.WIDTH, self.PREC
        xi = np.array([[-1.0 + 2.0 / (npoin - 1) * i] for i in range(npoin)])
        x = self.geo.transform(xi)[:, :, 0]
        val = self.geo.evaluate(u, neq, xi)
        with open(path, "w") as fh:
            fh.write(f"{'x':>{W}}" + "".join(" " + f"{s:>{W}}" for s in names) + "\n")
            for e in range(self.geo.nelem):
                for i in range(npoin):
                    fh.write(f"{x[e, i]:{W}.{P}e}" + "".join(" " + f"{val[e, i, f]:>{W}.{P}e}"
                                                           for f in range(neq)) + "\n")
                fh.write("\n")

    def write_dat(self, itime, time=0.0, rank=0):
        os.makedirs(self.data_directory, exist_ok=True)
        paths = []
        for i, (u, neq, names) in enumerate(self.fields):
            for kind, npoin in (("fieldset", 30), ("endpoints", 2)):
                p = os.path.join(self.data_directory,
                                 f"{self.collection_name}_{kind}{i}_rank{rank}_i{itime}.dat")
                self._write(p, u, neq, names, npoin)
                paths.append(p)
        return paths


# VTK cells of io/pvd_writer.hpp:64-168: reference nodes and cell type
_VTK_QUAD = ([(-1, -1), (1, -1), (1, 1), (-1, 1)], 9)
_VTK_QUADRATIC_QUAD = ([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)], 23)
_VTK_BIQUADRATIC_QUAD = ([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)], 28)
_VTK_HEX_NODES = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
VTK_HEXAHEDRON_ID = 12
# the reference's table gives the hexahedron the tetra's id (pvd_writer.hpp:154-167)
REFERENCE_HEXAHEDRON_ID = 10


def vtk_cell(ndim, order, reference_hex_id=False):
    """get_vtk_element (pvd_writer.hpp:172-241) for hypercubes: the cell for
    max(geometry order 1, basis order)."""
    if ndim == 2:
        return _VTK_QUAD if order <= 1 else (_VTK_BIQUADRATIC_QUAD if order == 2 else _VTK_QUADRATIC_QUAD)
    if ndim == 3:
        return _VTK_HEX_NODES, (REFERENCE_HEXAHEDRON_ID if reference_hex_id else VTK_HEXAHEDRON_ID)
    raise ValueError("the VTU writer covers 2D and 3D meshes")


def _open(name, fields=()):
    return "<" + name + "".join(f' {k}="{v}"' for k, v in fields) + ">\n"


def _close(name):
    return f"</{name}>\n"


_HEADER = _open("VTKFile", [("type", "UnstructuredGrid"), ("version", "0.1"), ("byte_order", "LittleEndian"),
                            ("header_type", "UInt64")])


class PVDWriter:
    """io/pvd_writer.hpp: ASCII VTU output of DG fields (every element its
    own copy of the VTK cell's points).  `reference_hex_id` writes hexahedra
    with the cell type the reference's table holds (10) instead of VTK's 12."""

    def __init__(self, geo, directory=None, collection_name="data", print_precision=8, reference_hex_id=False):
        self.geo = geo
        self.collection_name = collection_name
        self.print_precision = print_precision
        self.data_directory = os.path.join(directory or os.getcwd(), "iceicle_data")
        self.reference_hex_id = reference_hex_id
        self.fields = []

    def register_fields(self, u, neq, names):
        if len(names) != neq:
            raise ValueError("one name per field component")
        self.fields.append((u, neq, list(names)))

    def _cell(self, basis_order):
        nodes, vid = vtk_cell(self.geo.ndim, basis_order, self.reference_hex_id)
        return np.array(nodes, dtype=np.float64), vid

    def _grid(self, basis_order, with_fields):
        g, P = self.geo, self.print_precision
        nodes, vid = self._cell(basis_order)
        nn = nodes.shape[0]
        out = [_open("Piece", [("NumberOfPoints", g.nelem * nn), ("NumberOfCells", g.nelem)])]
        out.append(_open("Points"))
        out.append(_open("DataArray", [("type", "Float64"), ("Name", "Points"), ("NumberOfComponents", "3"),
                                       ("format", "ascii")]))
        xp = g.transform(nodes)
        for e in range(g.nelem):
            for i in range(nn):
                out.append("".join(_g(xp[e, i, d], P) + " " for d in range(g.ndim)) + "0 " * (3 - g.ndim) + "\n")
        out += [_close("DataArray"), _close("Points"), _open("Cells")]
        out.append(_open("DataArray", [("type", "Int64"), ("Name", "connectivity"), ("format", "ascii")]))
        for e in range(g.nelem):
            out.append("".join(f"{e * nn + i} " for i in range(nn)) + "\n")
        out.append(_close("DataArray"))
        out.append(_open("DataArray", [("type", "Int64"), ("Name", "offsets"), ("format", "ascii")]))
        out.append("".join(f"{(e + 1) * nn} " for e in range(g.nelem)) + "\n")
        out.append(_close("DataArray"))
        out.append(_open("DataArray", [("type", "Int64"), ("Name", "types"), ("format", "ascii")]))
        out.append("".join(f"{vid} " for _ in range(g.nelem)) + "\n")
        out += [_close("DataArray"), _close("Cells")]
        if with_fields:
            out.append(_open("PointData"))
            for u, neq, names in self.fields:
                val = g.evaluate(u, neq, nodes)
                for f, name in enumerate(names):
                    out.append(_open("DataArray", [("type", "Float64"), ("Name", name), ("format", "ascii")]))
                    for e in range(g.nelem):
                        out.append("".join(" " + _g(val[e, i, f], P) for i in range(nn)) + "\n")
                    out.append(_close("DataArray"))
            out.append(_close("PointData"))
        out.append(_close("Piece"))
        return out

    def write_vtu(self, itime, time=0.0):
        """<collection_name><itime>.vtu with the registered fields (pvd_writer.hpp:1108-1146)."""
        os.makedirs(self.data_directory, exist_ok=True)
        path = os.path.join(self.data_directory, f"{self.collection_name}{itime}.vtu")
        body = [_HEADER, _open("UnstructuredGrid"), *self._grid(self.geo.order, True), _close("UnstructuredGrid"),
                "</VTKFile>\n"]
        with open(path, "w") as fh:
            fh.write("".join(body))
        return path

    def write_mesh(self):
        """mesh.vtu: the geometry alone, at geometry order 1 (pvd_writer.hpp:1156-1186)."""
        os.makedirs(self.data_directory, exist_ok=True)
        path = os.path.join(self.data_directory, "mesh.vtu")
        body = [_HEADER, _open("UnstructuredGrid"), *self._grid(1, False), _close("UnstructuredGrid"),
                "</VTKFile>\n"]
        with open(path, "w") as fh:
            fh.write("".join(body))
        return path


def read_vtu_points(path):
    """Points, connectivity, offsets, types and point-data arrays of an ASCII
    VTU written above (used to check files)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arr = {}
    for da in piece.iter("DataArray"):
        typ = np.int64 if da.get("type") == "Int64" else np.float64
        arr[da.get("Name")] = np.array((da.text or "").split(), dtype=typ)
    arr["Points"] = arr["Points"].reshape(-1, 3)
    return arr, int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))

