"""Python front end of libiceicle_amd.so (ctypes over the C-ABI in
include/iceicle_amd.h).

Mirrors the reference's C++ entry points for the residual path:
  FESpace(mesh, order)                 ~ FESpace ctor (fespace.hpp:200-416)
  ConservationLawDDG(space, params)    ~ ConservationLawDDG (conservation_law.hpp:92-230)
  form_residual(space, disc, u, res)   ~ solvers::form_residual (form_residual.hpp:56-268)
Device memory and streams come from torch (ROCm); `u` and `res` are CUDA
tensors in fe_layout_right order.  There is no CPU fallback: without the
built library or a GPU these raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiceicle_amd.so")
_LIB = None

# every symbol declared in include/iceicle_amd.h, with its ctypes signature
P = C.c_void_p
_SIGS = {
    "ice_version": ([], C.c_int),
    "ice_last_error": ([], C.c_char_p),
    "ice_disc_params_default": ([C.c_int, C.POINTER(abi.DiscParams)], C.c_int),
    "ice_mesh_export": ([C.POINTER(abi.MeshDesc), C.POINTER(abi.MeshCounts), C.POINTER(C.c_double),
                         C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "ice_space_create": ([C.POINTER(abi.MeshDesc), C.POINTER(abi.SpaceDesc), C.POINTER(P)], C.c_int),
    "ice_space_destroy": ([P], C.c_int),
    "ice_space_query": ([P, C.POINTER(abi.SpaceInfo)], C.c_int),
    "ice_space_qp_coords": ([P, C.POINTER(C.c_double)], C.c_int),
    "ice_disc_create": ([P, C.POINTER(abi.DiscParams), C.POINTER(P)], C.c_int),
    "ice_disc_destroy": ([P], C.c_int),
    "ice_disc_query": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int64)], C.c_int),
    "ice_residual": ([P, P, P, P], C.c_int),
    "ice_residual_mixed": ([P, P, P, P], C.c_int),
    "ice_residual_noalloc": ([P, P, P, P], C.c_int),
    "ice_noalloc_scratch_bytes": ([P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "ice_mdg_select": ([P, C.POINTER(C.c_int), C.c_int, C.POINTER(P)], C.c_int),
    "ice_mdg_query": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "ice_mdg_residual": ([P, P, P, P, P], C.c_int),
    "ice_mdg_form_residual": ([P, P, P, P, P], C.c_int),
    "ice_interface_error": ([P, P, C.POINTER(C.c_double), P], C.c_int),
    "ice_mdg_destroy": ([P], C.c_int),
    "ice_disc_update": ([P, C.POINTER(abi.DiscParams)], C.c_int),
    "ice_space_create_batch": ([C.POINTER(abi.MeshDesc), C.c_int, C.POINTER(abi.SpaceDesc), C.POINTER(P)], C.c_int),
    "ice_disc_set_element_viscosity": ([P, P], C.c_int),
    "ice_disc_set_tile": ([P, C.c_int], C.c_int),
    "ice_disc_autotune": ([P, P, P, C.c_int, C.POINTER(C.c_int), P, P], C.c_int),
    "ice_stepper_set_element_dt": ([P, P], C.c_int),
    "ice_step_graph_create": ([P, P, C.c_int, C.c_int, C.POINTER(P)], C.c_int),
    "ice_step_graph_create_ex": ([P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(P)], C.c_int),
    "ice_step_graph_launch": ([P, P], C.c_int),
    "ice_step_graph_destroy": ([P], C.c_int),
    "ice_profile_begin": ([P, C.c_int], C.c_int),
    "ice_profile_end": ([P, C.POINTER(C.c_double), C.POINTER(C.c_int)], C.c_int),
    "ice_field_eval_points": ([C.POINTER(abi.Field), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)], C.c_int),
    "ice_space_node_coords": ([P, C.POINTER(C.c_double)], C.c_int),
    "ice_trace_points": ([C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "ice_mass_inverse_apply": ([P, P, P, P], C.c_int),
    "ice_project": ([P, C.POINTER(abi.Field), P, P], C.c_int),
    "ice_l2_error": ([P, C.POINTER(abi.Field), P, C.POINTER(C.c_double), P], C.c_int),
    "ice_project_values": ([P, P, P, P], C.c_int),
    "ice_disc_set_source_values": ([P, P, P], C.c_int),
    "ice_space_bface_coords": ([P, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "ice_disc_set_boundary_values": ([P, P, P], C.c_int),
    "ice_stepper_create": ([P, C.POINTER(abi.StepperDesc), C.POINTER(P)], C.c_int),
    "ice_stepper_destroy": ([P], C.c_int),
    "ice_stepper_step": ([P, P, P], C.c_int),
    "ice_stepper_step_f32": ([P, P, P], C.c_int),
    "ice_residual_f32state": ([P, P, P, C.c_int, P], C.c_int),
    "ice_stepper_solve": ([P, P, P, C.POINTER(C.c_longlong), C.POINTER(C.c_double)], C.c_int),
    "ice_stepper_status": ([P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                            C.POINTER(C.c_int)], C.c_int),
    "ice_jacobian_layout": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "ice_jacobian": ([P, P, P, P, P], C.c_int),
    "ice_jacobian_apply": ([P, P, P, P, P], C.c_int),
    "ice_newton_opts_default": ([C.POINTER(abi.NewtonOpts)], C.c_int),
    "ice_newton_create": ([P, C.POINTER(abi.NewtonOpts), C.POINTER(P)], C.c_int),
    "ice_newton_destroy": ([P], C.c_int),
    "ice_newton_solve": ([P, P, C.POINTER(abi.NewtonReport), P], C.c_int),
    "ice_newton_linear_solve": ([P, P, P, P, C.POINTER(C.c_int), C.POINTER(C.c_double), P], C.c_int),
    "ice_mlp_create": ([C.c_int, C.POINTER(C.c_int), C.c_ulonglong, C.POINTER(P)], C.c_int),
    "ice_mlp_destroy": ([P], C.c_int),
    "ice_mlp_num_params": ([P, C.POINTER(C.c_longlong)], C.c_int),
    "ice_mlp_get_params": ([P, P], C.c_int),
    "ice_mlp_set_params": ([P, P], C.c_int),
    "ice_mlp_forward": ([P, P, C.c_int, P, P], C.c_int),
    "ice_mlp_gradient": ([P, P, P, C.c_int, C.POINTER(C.c_double), P, P], C.c_int),
    "ice_mlp_train": ([P, P, P, C.c_int, C.POINTER(abi.MlpTrainOpts), P, P], C.c_int),
}


class IceError(RuntimeError):
    pass


def lib():
    """Load the engine library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise IceError(f"{LIB_PATH} is missing: build it with `python -m iceicle_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(rc):
    if rc != abi.ICE_OK:
        raise IceError(f"iceicle_amd error {rc}: {lib().ice_last_error().decode()}")


def mesh_export(mesh):
    """Host-only mesh build: (coord, conn, faces) in reference order."""
    L = lib()
    cnt = abi.MeshCounts()
    check(L.ice_mesh_export(C.byref(mesh), C.byref(cnt), None, None, None))
    coord = np.empty((cnt.nnode, cnt.ndim))
    conn = np.empty((cnt.nelem, cnt.nvert_el), dtype=np.int32)
    faces = np.empty((cnt.nface, abi.FACE_REC_FIXED + cnt.nfacenode), dtype=np.int32)
    check(L.ice_mesh_export(C.byref(mesh), C.byref(cnt), coord.ctypes.data_as(C.POINTER(C.c_double)),
                            conn.ctypes.data_as(C.POINTER(C.c_int)), faces.ctypes.data_as(C.POINTER(C.c_int))))
    return coord, conn, faces, cnt


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def field_eval(field, ndim, neq, x):
    """Evaluate an analytic field at points x[npts][ndim] on the host."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, ndim)
    out = np.empty((x.shape[0], neq))
    check(lib().ice_field_eval_points(C.byref(field), ndim, neq, x.shape[0], _dp(x), _dp(out)))
    return out


def trace_points(ndim, order, face_nr, orientation):
    nfq = (order + 1) ** (ndim - 1)
    xi = np.empty((nfq, ndim))
    check(lib().ice_trace_points(ndim, order, face_nr, orientation, _dp(xi)))
    return xi


def _require_cuda(t, name, dtype=None):
    import torch
    dtype = torch.float64 if dtype is None else dtype
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise IceError(f"{name} must be a contiguous {str(dtype).replace('torch.', '')} CUDA tensor")


class FESpace:
    """Mesh + reference tables + trace maps resident on the current GPU."""

    def __init__(self, mesh, order):
        """`mesh` is one MeshDesc, or a list of them for a batched space
        (the meshes concatenated, element ids of mesh k after meshes 0..k-1)."""
        L = lib()
        self._mesh = mesh
        self._space_desc = abi.space_desc(order)
        h = P()
        if isinstance(mesh, (list, tuple)):
            arr = (abi.MeshDesc * len(mesh))(*mesh)
            self._mesh_array = arr
            check(L.ice_space_create_batch(arr, len(mesh), C.byref(self._space_desc), C.byref(h)))
        else:
            check(L.ice_space_create(C.byref(mesh), C.byref(self._space_desc), C.byref(h)))
        self.h = h
        info = abi.SpaceInfo()
        check(L.ice_space_query(self.h, C.byref(info)))
        self.info = info
        self.ndim, self.order, self.nbasis = info.ndim, info.order, info.nbasis
        self.nelem, self.nqp, self.nfqp = info.nelem, info.nqp, info.nfqp

    def qp_coords(self):
        xq = np.empty((self.nelem, self.nqp, self.ndim))
        check(lib().ice_space_qp_coords(self.h, xq.ctypes.data_as(C.POINTER(C.c_double))))
        return xq

    def node_coords(self):
        xn = np.empty((self.nelem, self.nbasis, self.ndim))
        check(lib().ice_space_node_coords(self.h, _dp(xn)))
        return xn

    def bface_coords(self):
        """(xb [nfb][nfqp][ndim], bctype [nfb], bcflag [nfb]) of the boundary
        faces in device order (ice_space_bface_coords)."""
        nfb = self.info.nface_boundary
        xb = np.empty((nfb, self.nfqp, self.ndim))
        bt = np.empty(nfb, dtype=np.int32)
        bf = np.empty(nfb, dtype=np.int32)
        check(lib().ice_space_bface_coords(self.h, _dp(xb), bt.ctypes.data_as(C.POINTER(C.c_int)),
                                           bf.ctypes.data_as(C.POINTER(C.c_int))))
        return xb, bt, bf

    def interpolate(self, field, neq):
        """Nodal interpolant of an analytic field (host array, fe_layout_right)."""
        xn = self.node_coords().reshape(-1, self.ndim)
        return field_eval(field, self.ndim, neq, xn).ravel()

    def __del__(self):
        # at interpreter exit the module may be torn down first; the driver
        # releases device memory with the process then
        try:
            if getattr(self, "h", None):
                lib().ice_space_destroy(self.h)
                self.h = None
        except Exception:
            pass


class ConservationLawDDG:
    """Discretization bound to an FESpace (flux family + DDG parameters)."""

    def __init__(self, space, params):
        L = lib()
        self.space = space
        self.params = params
        h = P()
        check(L.ice_disc_create(space.h, C.byref(params), C.byref(h)))
        self.h = h
        self.neq = 1 if params.flux == abi.FLUX_BURGERS else space.ndim + 2
        self.size = space.nelem * space.nbasis * self.neq

    def set_element_viscosity(self, mu):
        """Per-element viscosity (Burgers; host array [nelem]) or None."""
        if mu is None:
            check(lib().ice_disc_set_element_viscosity(self.h, None))
            return
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        if mu.size != self.space.nelem:
            raise IceError(f"expected {self.space.nelem} viscosities")
        check(lib().ice_disc_set_element_viscosity(self.h, mu.ctypes.data_as(P)))

    def set_tile(self, tile):
        check(lib().ice_disc_set_tile(self.h, int(tile)))

    def set_source_values(self, vq, stream=None):
        """Source term from values at the volume quadrature points
        (float64 CUDA tensor, [nelem][nqp][neq] in qp_coords order)."""
        _require_cuda(vq, "vq")
        if vq.numel() != self.space.nelem * self.space.nqp * self.neq:
            raise IceError("source values must be [nelem][nqp][neq]")
        check(lib().ice_disc_set_source_values(self.h, _ptr(vq), stream or current_stream()))

    def set_boundary_values(self, bv, stream=None):
        """Dirichlet states / Neumann data at the boundary face quadrature
        points (float64 CUDA tensor, [nfb][nfqp][neq] in bface_coords order)."""
        _require_cuda(bv, "bv")
        if bv.numel() != self.space.info.nface_boundary * self.space.nfqp * self.neq:
            raise IceError("boundary values must be [nface_boundary][nfqp][neq]")
        check(lib().ice_disc_set_boundary_values(self.h, _ptr(bv), stream or current_stream()))

    def autotune(self, u, r, reps=10, stream=None):
        """Time the element-kernel tiles on this state and keep the fastest,
        for the FP64 residual and (2D, 3D) separately for the mixed one;
        returns the FP64 residual's (best tile, [ms per residual for tiles
        0, 1, 2])."""
        _require_cuda(u, "u")
        _require_cuda(r, "r")
        best = C.c_int()
        ms = np.zeros(3)
        check(lib().ice_disc_autotune(self.h, _ptr(u), _ptr(r), reps, C.byref(best), ms.ctypes.data_as(P),
                                      stream or current_stream()))
        return best.value, ms.tolist()

    def update(self, params):
        """Replace the parameters in place (same flux family; DDGIC and the
        source term stay on or off as at creation)."""
        check(lib().ice_disc_update(self.h, C.byref(params)))
        self.params = params

    def profile_begin(self, max_calls):
        check(lib().ice_profile_begin(self.h, max_calls))

    def profile_end(self):
        """Summed milliseconds of the (face trace, face flux, element) kernels
        and the number of residuals profiled since profile_begin."""
        ms = (C.c_double * 3)()
        n = C.c_int()
        check(lib().ice_profile_end(self.h, ms, C.byref(n)))
        return [ms[0], ms[1], ms[2]], n.value

    def __del__(self):
        # at interpreter exit the module may be torn down first; the driver
        # releases device memory with the process then
        try:
            if getattr(self, "h", None):
                lib().ice_disc_destroy(self.h)
                self.h = None
        except Exception:
            pass


def current_stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def form_residual(space, disc, u, res, stream=None, precision="fp64"):
    """res <- R(u) on the GPU; precision "fp64" (the reference's arithmetic)
    or "mixed" (FP32 interpolation and fluxes, FP64 geometry and output, 2D and 3D).
    u may be an FP32 state (2D and 3D): "mixed" is then the residual FP32
    steps use, "fp64" the FP64 residual of the widened state (the check of an
    FP32 solution against the FP64 discretization).  res is float64."""
    import torch
    _require_cuda(res, "res")
    if isinstance(u, torch.Tensor) and u.dtype == torch.float32:
        _require_cuda(u, "u", torch.float32)
    else:
        _require_cuda(u, "u")
    if u.numel() != disc.size or res.numel() != disc.size:
        raise IceError(f"u/res must have {disc.size} entries")
    if precision not in ("fp64", "mixed"):
        raise IceError(f"unknown precision {precision!r}")
    st = current_stream() if stream is None else stream
    if u.dtype == torch.float32:
        check(lib().ice_residual_f32state(disc.h, _ptr(u), _ptr(res), 1 if precision == "fp64" else 0, st))
        return
    fn = lib().ice_residual if precision == "fp64" else lib().ice_residual_mixed
    check(fn(disc.h, C.c_void_p(u.data_ptr()), C.c_void_p(res.data_ptr()), st))


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class MDGSelection:
    """Selected traces of the MDG interface-conservation residual
    (nodeset_dof_map, node_set_layout.hpp:46-83): trace t < number of
    interior faces is interior face t, else boundary face t - that number."""

    def __init__(self, space, traces):
        self.space = space
        tr = np.ascontiguousarray(traces, dtype=np.int32)
        h = P()
        check(lib().ice_mdg_select(space.h, tr.ctypes.data_as(C.POINTER(C.c_int)), tr.size, C.byref(h)))
        self.h = h
        n = C.c_int()
        check(lib().ice_mdg_query(self.h, C.byref(n), None))
        self.nodes = np.empty(n.value, dtype=np.int32)
        check(lib().ice_mdg_query(self.h, C.byref(n), self.nodes.ctypes.data_as(C.POINTER(C.c_int))))

    def residual(self, disc, u, res=None, stream=None):
        """form_mdg_residual: res[node][neq] on the device."""
        import torch
        _require_cuda(u, "u")
        if res is None:
            res = torch.empty((self.nodes.size, disc.neq), dtype=torch.float64, device=u.device)
        check(lib().ice_mdg_residual(disc.h, self.h, _ptr(u), _ptr(res), stream or current_stream()))
        return res

    def form_residual(self, disc, u, res=None, stream=None):
        """The combined residual over the selection (form_residual with a
        nodeset, form_residual.hpp:380-424): u = [DG coefficients | selected
        node coordinates [nnodes][ndim]] (float64 CUDA), returns [DG residual
        | MDG residual [nnodes][neq]].  The selected nodes of the space's
        device geometry move to the given coordinates and stay there."""
        import torch
        _require_cuda(u, "u")
        nin = disc.size + self.nodes.size * self.space.ndim
        if u.numel() != nin:
            raise IceError(f"u must have {nin} entries (DG coefficients, then node coordinates)")
        nout = disc.size + self.nodes.size * disc.neq
        if res is None:
            res = torch.empty(nout, dtype=torch.float64, device=u.device)
        _require_cuda(res, "res")
        if res.numel() != nout:
            raise IceError(f"res must have {nout} entries")
        check(lib().ice_mdg_form_residual(disc.h, self.h, _ptr(u), _ptr(res), stream or current_stream()))
        return res

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ice_mdg_destroy(self.h)
                self.h = None
        except Exception:
            pass


def mass_inverse_apply(disc, r, out, stream=None):
    """out <- M^-1 r per element (inverse mass matrix of the explicit steppers)."""
    _require_cuda(r, "r")
    _require_cuda(out, "out")
    check(lib().ice_mass_inverse_apply(disc.h, _ptr(r), _ptr(out), stream or current_stream()))


def interface_error(disc, u, stream=None):
    """ic_error (l2_error.hpp:279-410): sum over interior faces of |jump of
    the normal physical flux| w sqrt(g), over points and equations."""
    _require_cuda(u, "u")
    e = C.c_double()
    check(lib().ice_interface_error(disc.h, _ptr(u), C.byref(e), stream or current_stream()))
    return e.value


def project_values(disc, vq, u, stream=None):
    """u <- L2 projection of values at the volume quadrature points
    (float64 CUDA tensors; vq [nelem][nqp][neq] in qp_coords order)."""
    _require_cuda(vq, "vq")
    _require_cuda(u, "u")
    if vq.numel() != disc.space.nelem * disc.space.nqp * disc.neq or u.numel() != disc.size:
        raise IceError("vq must be [nelem][nqp][neq] and u the discretization's size")
    check(lib().ice_project_values(disc.h, _ptr(vq), _ptr(u), stream or current_stream()))


def project(disc, field, u, stream=None):
    """u <- L2 projection of an analytic field (synchronous)."""
    _require_cuda(u, "u")
    check(lib().ice_project(disc.h, C.byref(field), _ptr(u), stream or current_stream()))


def l2_error(disc, field, u, stream=None):
    _require_cuda(u, "u")
    err = C.c_double()
    check(lib().ice_l2_error(disc.h, C.byref(field), _ptr(u), C.byref(err), stream or current_stream()))
    return err.value


class ExplicitStepper:
    """RK3SSP / RK3TVD / ExplicitEuler on the device (abi.stepper_desc)."""

    def __init__(self, disc, desc):
        self.disc = disc
        self.desc = desc
        h = P()
        check(lib().ice_stepper_create(disc.h, C.byref(desc), C.byref(h)))
        self.h = h

    def step(self, u, stream=None):
        """One step of u: a float64 state, or a float32 one (2D and 3D,
        textbook RK3-SSP or explicit Euler; ice_stepper_step_f32)."""
        import torch
        if isinstance(u, torch.Tensor) and u.dtype == torch.float32:
            _require_cuda(u, "u", torch.float32)
            check(lib().ice_stepper_step_f32(self.h, _ptr(u), stream or current_stream()))
            return
        _require_cuda(u, "u")
        check(lib().ice_stepper_step(self.h, _ptr(u), stream or current_stream()))

    def solve(self, u, stream=None):
        _require_cuda(u, "u")
        n = C.c_longlong()
        t = C.c_double()
        check(lib().ice_stepper_solve(self.h, _ptr(u), stream or current_stream(), C.byref(n), C.byref(t)))
        return n.value, t.value

    def set_element_dt(self, dt):
        """Per-element step sizes (host array [nelem]) or None."""
        if dt is None:
            check(lib().ice_stepper_set_element_dt(self.h, None))
            return
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        if dt.size != self.disc.space.nelem:
            raise IceError(f"expected {self.disc.space.nelem} step sizes")
        check(lib().ice_stepper_set_element_dt(self.h, dt.ctypes.data_as(P)))

    def status(self):
        t, dt, it, done = C.c_double(), C.c_double(), C.c_longlong(), C.c_int()
        check(lib().ice_stepper_status(self.h, C.byref(t), C.byref(dt), C.byref(it), C.byref(done)))
        return dict(time=t.value, dt=dt.value, itime=it.value, done=bool(done.value))

    def __del__(self):
        # at interpreter exit the module may be torn down first; the driver
        # releases device memory with the process then
        try:
            if getattr(self, "h", None):
                lib().ice_stepper_destroy(self.h)
                self.h = None
        except Exception:
            pass


class BlockJacobian:
    """Element-block Jacobian of a discretization on the device
    (form_petsc_jacobian_fd, form_petsc_jacobian.hpp:55-317).

    `blocks` is a float64 CUDA tensor [nelem][nslots][bs][bs]; slot 0 is the
    diagonal block, slot 1+j couples an element with the element across its
    local face j; `cols[e][s]` names that element (-1 on a boundary face)."""

    def __init__(self, disc):
        import torch
        self.disc = disc
        ns, bs = C.c_int(), C.c_int()
        check(lib().ice_jacobian_layout(disc.h, C.byref(ns), C.byref(bs), None))
        self.nslots, self.bs = ns.value, bs.value
        nelem = disc.space.nelem
        self.cols = np.empty((nelem, self.nslots), dtype=np.int32)
        check(lib().ice_jacobian_layout(disc.h, C.byref(ns), C.byref(bs),
                                        self.cols.ctypes.data_as(C.POINTER(C.c_int))))
        self.blocks = torch.empty((nelem, self.nslots, self.bs, self.bs), dtype=torch.float64, device="cuda")

    def assemble(self, u, res=None, stream=None):
        """blocks <- dR/du at u; res <- R(u) when given."""
        _require_cuda(u, "u")
        if res is not None:
            _require_cuda(res, "res")
        check(lib().ice_jacobian(self.disc.h, _ptr(u), _ptr(res) if res is not None else None,
                                 _ptr(self.blocks), stream or current_stream()))

    def apply(self, x, y, stream=None):
        """y <- J x."""
        _require_cuda(x, "x")
        _require_cuda(y, "y")
        check(lib().ice_jacobian_apply(self.disc.h, _ptr(self.blocks), _ptr(x), _ptr(y), stream or current_stream()))


def newton_opts(**kw):
    """Newton options with the reference's defaults, overridden by keyword."""
    o = abi.NewtonOpts()
    check(lib().ice_newton_opts_default(C.byref(o)))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise IceError(f"unknown Newton option {k}")
        setattr(o, k, v)
    return o


class PetscNewton:
    """Newton-linesearch steady solver on the device, named after the
    reference's PetscNewton (petsc_newton.hpp:124-329) whose role it takes;
    the linear algebra is the library's own GMRES + block Jacobi."""

    def __init__(self, disc, opts=None, **kw):
        self.disc = disc
        self.opts = opts if opts is not None else newton_opts(**kw)
        h = P()
        check(lib().ice_newton_create(disc.h, C.byref(self.opts), C.byref(h)))
        self.h = h

    def solve(self, u, stream=None):
        """Solve R(u) = 0 in place from the initial guess u; returns the report as a dict."""
        _require_cuda(u, "u")
        rep = abi.NewtonReport()
        check(lib().ice_newton_solve(self.h, _ptr(u), C.byref(rep), stream or current_stream()))
        return rep.as_dict()

    def linear_solve(self, b, x, u=None, stream=None):
        """x <- J^-1 b for the Jacobian at u (or the last one assembled)."""
        _require_cuda(b, "b")
        _require_cuda(x, "x")
        its, rel = C.c_int(), C.c_double()
        check(lib().ice_newton_linear_solve(self.h, _ptr(u) if u is not None else None, _ptr(b), _ptr(x),
                                            C.byref(its), C.byref(rel), stream or current_stream()))
        return its.value, rel.value

    def __del__(self):
        # at interpreter exit the module may be torn down first; the driver
        # releases device memory with the process then
        try:
            if getattr(self, "h", None):
                lib().ice_newton_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _require_f32(t, name, cols=None):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise IceError(f"{name} must be a contiguous float32 CUDA tensor")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise IceError(f"{name} must have shape [batch, {cols}]")


class MLP:
    """Fully connected tanh network trained and evaluated by the library's
    device kernels (warm-start predictor; no reference counterpart)."""

    def __init__(self, widths, seed=0):
        self.widths = [int(w) for w in widths]
        arr = (C.c_int * len(self.widths))(*self.widths)
        h = P()
        check(lib().ice_mlp_create(len(self.widths) - 1, arr, seed, C.byref(h)))
        self.h = h
        n = C.c_longlong()
        check(lib().ice_mlp_num_params(self.h, C.byref(n)))
        self.nparams = n.value

    def get_params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        check(lib().ice_mlp_get_params(self.h, out.ctypes.data_as(P)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise IceError(f"expected {self.nparams} parameters")
        check(lib().ice_mlp_set_params(self.h, p.ctypes.data_as(P)))

    def forward(self, x, y=None, stream=None):
        import torch
        _require_f32(x, "x", self.widths[0])
        if y is None:
            y = torch.empty((x.shape[0], self.widths[-1]), dtype=torch.float32, device=x.device)
        _require_f32(y, "y", self.widths[-1])
        check(lib().ice_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(y), stream or current_stream()))
        return y

    def gradient(self, x, t, stream=None):
        """(loss, gradient as a device float32 vector in the packed layout)."""
        import torch
        _require_f32(x, "x", self.widths[0])
        _require_f32(t, "t", self.widths[-1])
        g = torch.empty(self.nparams, dtype=torch.float32, device=x.device)
        loss = C.c_double()
        check(lib().ice_mlp_gradient(self.h, _ptr(x), _ptr(t), x.shape[0], C.byref(loss), _ptr(g),
                                     stream or current_stream()))
        return loss.value, g

    def train(self, x, t, opts, stream=None):
        """Adam training; returns the per-epoch mean loss."""
        _require_f32(x, "x", self.widths[0])
        _require_f32(t, "t", self.widths[-1])
        if x.shape[0] != t.shape[0]:
            raise IceError("x and t must have the same number of rows")
        hist = np.zeros(max(opts.epochs, 1))
        check(lib().ice_mlp_train(self.h, _ptr(x), _ptr(t), x.shape[0], C.byref(opts),
                                  hist.ctypes.data_as(P), stream or current_stream()))
        return hist[:opts.epochs]

    def __del__(self):
        # at interpreter exit the module may be torn down first; the driver
        # releases device memory with the process then
        try:
            if getattr(self, "h", None):
                lib().ice_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass


class StepGraph:
    """nsteps steps of several steppers (each on its own device state)
    captured into one HIP graph; launch() replays them all."""

    def __init__(self, steppers, states, nsteps, concurrent=False):
        """states: all float64, or all float32 (FP32-state steps).  With
        concurrent, steppers on distinct discretizations and states become
        independent branches of the graph (ice_step_graph_create_ex)."""
        import torch
        if len(steppers) != len(states):
            raise IceError("one state per stepper")
        fp32 = bool(states) and states[0].dtype == torch.float32
        for u in states:
            _require_cuda(u, "u", torch.float32 if fp32 else torch.float64)
        self._keep = (list(steppers), list(states))
        sa = (P * len(steppers))(*[s.h for s in steppers])
        ua = (P * len(states))(*[u.data_ptr() for u in states])
        h = P()
        check(lib().ice_step_graph_create_ex(sa, ua, int(fp32), len(steppers), nsteps, int(bool(concurrent)),
                                             C.byref(h)))
        self.h = h
        self.nsteps = nsteps

    def launch(self, stream=None):
        check(lib().ice_step_graph_launch(self.h, stream or current_stream()))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ice_step_graph_destroy(self.h)
                self.h = None
        except Exception:
            pass
