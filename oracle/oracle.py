"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module, and only as the checker.  The oracle is a serial FP64
restatement of the reference's form_residual path (see oracle/*.hpp for the
file:line of every function it follows).
"""
import ctypes as C
import os

import numpy as np

from iceicle_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def max_threads():
    """Host threads OpenMP would use (OMP_NUM_THREADS, else the cores)."""
    return int(lib().ice_oracle_max_threads())


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.ice_oracle_create.restype = P
        L.ice_oracle_create.argtypes = [C.POINTER(abi.MeshDesc), C.POINTER(abi.SpaceDesc), C.POINTER(abi.DiscParams), C.c_int]
        L.ice_oracle_destroy.argtypes = [P]
        L.ice_oracle_last_error.restype = C.c_char_p
        for name, args in [
            ("ice_oracle_info", [P, ip]),
            ("ice_oracle_residual", [P, dp, dp]),
            ("ice_oracle_residual_parallel", [P, dp, dp, C.c_int]),
            ("ice_oracle_mdg_nodeset", [P, ip, C.c_int, ip, ip]),
            ("ice_oracle_mdg_residual", [P, ip, C.c_int, dp, dp]),
            ("ice_oracle_faces", [P, ip]),
            ("ice_oracle_mesh", [P, dp, ip]),
            ("ice_oracle_project", [P, C.POINTER(abi.Field), dp]),
            ("ice_oracle_rk3_ssp", [P, dp, C.c_double, C.c_int, C.c_int]),
            ("ice_oracle_rk3_tvd_tfinal", [P, dp, C.c_double, C.c_double]),
            ("ice_oracle_trace_points", [P, dp, dp]),
            ("ice_oracle_face_geometry", [P, dp, dp, dp]),
            ("ice_oracle_qp_coords", [P, dp]),
            ("ice_oracle_mass_apply", [P, dp, dp]),
            ("ice_oracle_lagrange_1d", [C.c_int, C.c_double, dp, dp, dp]),
            ("ice_oracle_gauss_legendre", [C.c_int, dp, dp]),
            ("ice_oracle_tensor_basis", [C.c_int, C.c_int, dp, dp, dp, dp]),
            ("ice_oracle_element_jacobian", [C.c_int, dp, dp, dp]),
            ("ice_oracle_ns_flux", [C.POINTER(abi.DiscParams), C.c_int, dp, dp, dp]),
            ("ice_oracle_vanleer", [C.POINTER(abi.DiscParams), C.c_int, dp, dp, dp, dp]),
            ("ice_oracle_diff_flux", [C.POINTER(abi.DiscParams), C.c_int, dp, dp, dp, dp]),
            ("ice_oracle_burgers_flux", [C.POINTER(abi.DiscParams), C.c_int, dp, dp, dp, dp]),
            ("ice_oracle_apply_bc", [C.POINTER(abi.DiscParams), C.c_int, dp, dp, dp, C.c_int, C.c_int, dp, dp]),
            ("ice_oracle_trace_ref", [C.c_int, C.c_int, C.c_int, dp, dp]),
            ("ice_oracle_jacobian", [P, dp, dp, ip, dp]),
            ("ice_oracle_rk_solve", [P, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.c_double, C.c_longlong, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]),
        ]:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.ice_oracle_l2_error.argtypes = [P, C.POINTER(abi.Field), dp]
        L.ice_oracle_l2_error.restype = C.c_double
        L.ice_oracle_newton.argtypes = [P, dp, C.c_double, C.c_double, C.c_int, dp, C.c_int]
        L.ice_oracle_newton.restype = C.c_int
        L.ice_oracle_max_threads.restype = C.c_int
        L.ice_oracle_cfl_dt.argtypes = [P, C.c_double]
        L.ice_oracle_cfl_dt.restype = C.c_double
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().ice_oracle_last_error().decode())


class Oracle:
    """One mesh + FE space + ConservationLawDDG on the CPU."""

    def __init__(self, mesh, space, params, fix3d_boundary_order=True):
        L = lib()
        self._mesh, self._space, self._params = mesh, space, params
        self.h = L.ice_oracle_create(C.byref(mesh), C.byref(space), C.byref(params), int(fix3d_boundary_order))
        if not self.h:
            raise RuntimeError("oracle: " + L.ice_oracle_last_error().decode())
        info = np.zeros(10, dtype=np.int32)
        _check(L.ice_oracle_info(self.h, _ip(info)))
        (self.ndim, self.nelem, self.nbasis, self.neq, self.nface, self.nface_interior,
         self.nfqp, self.nqp, self.nnode, self.nperiodic) = (int(v) for v in info)
        self.size = self.nelem * self.nbasis * self.neq

    def __del__(self):
        if getattr(self, "h", None):
            lib().ice_oracle_destroy(self.h)
            self.h = None

    def residual(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.empty(self.size)
        _check(lib().ice_oracle_residual(self.h, _dp(u), _dp(r)))
        return r

    def mdg_nodeset(self, traces):
        """Global node ids of the MDG degrees of freedom of the selected traces
        (node_set_layout.hpp:46-83); trace t < n_interior is interior face t,
        else boundary face t - n_interior."""
        tr = np.ascontiguousarray(traces, dtype=np.int32)
        n = C.c_int()
        _check(lib().ice_oracle_mdg_nodeset(self.h, _ip(tr), tr.size, C.byref(n), None))
        out = np.empty(n.value, dtype=np.int32)
        _check(lib().ice_oracle_mdg_nodeset(self.h, _ip(tr), tr.size, C.byref(n), _ip(out)))
        return out

    def mdg_residual(self, traces, u):
        """Interface-conservation residual [selected node][neq] (form_mdg_residual,
        form_residual.hpp:270-327)."""
        tr = np.ascontiguousarray(traces, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        nsel = self.mdg_nodeset(tr).size
        out = np.empty(nsel * self.neq)
        _check(lib().ice_oracle_mdg_residual(self.h, _ip(tr), tr.size, _dp(u), _dp(out)))
        return out.reshape(nsel, self.neq)

    def residual_parallel(self, u, nthreads=0):
        """The same residual, bit for bit, from the element/face-parallel loop
        on nthreads host threads (0: every available core)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.empty(self.size)
        n = nthreads if nthreads > 0 else max_threads()
        _check(lib().ice_oracle_residual_parallel(self.h, _dp(u), _dp(r), int(n)))
        return r

    def faces(self):
        nfv = 1 if self.ndim == 1 else 2 ** (self.ndim - 1)
        rec = np.empty((self.nface, abi.FACE_REC_FIXED + nfv), dtype=np.int32)
        _check(lib().ice_oracle_faces(self.h, _ip(rec)))
        return rec

    def mesh(self):
        coord = np.empty((self.nnode, self.ndim))
        conn = np.empty((self.nelem, 2 ** self.ndim), dtype=np.int32)
        _check(lib().ice_oracle_mesh(self.h, _dp(coord), _ip(conn)))
        return coord, conn

    def project(self, field):
        u = np.empty(self.size)
        _check(lib().ice_oracle_project(self.h, C.byref(field), _dp(u)))
        return u

    def l2_error(self, field, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().ice_oracle_l2_error(self.h, C.byref(field), _dp(u))

    def rk3_ssp(self, u, dt, nsteps, textbook=True):
        u = np.array(u, dtype=np.float64)
        _check(lib().ice_oracle_rk3_ssp(self.h, _dp(u), dt, nsteps, 1 if textbook else 0))
        return u

    def rk3_tvd_tfinal(self, u, dt, tfinal):
        u = np.array(u, dtype=np.float64)
        n = lib().ice_oracle_rk3_tvd_tfinal(self.h, _dp(u), dt, tfinal)
        if n < 0:
            _check(-1)
        return u, n

    def rk_solve(self, u, kind=0, compat=False, dt=0.0, cfl=None, tfinal=None, ntime=0):
        """Explicit solve as the reference's solve(): kind 0 RK3-SSP, 1 RK3-TVD,
        2 explicit Euler; CFL time step if cfl is given, else fixed dt."""
        u = np.array(u, dtype=np.float64)
        n = C.c_longlong()
        t = C.c_double()
        _check(lib().ice_oracle_rk_solve(self.h, _dp(u), kind, int(compat), int(cfl is not None), dt,
                                         cfl or 0.0, int(tfinal is not None), tfinal or 0.0, ntime,
                                         C.byref(n), C.byref(t)))
        return u, n.value, t.value

    def jacobian(self, u):
        """(res, cols[nelem][2 ndim + 1], blocks[nelem][2 ndim + 1][bs][bs]) by the
        reference's finite-difference block Jacobian.  Slot 0 is the diagonal,
        slot 1 + f the neighbour across face f; a neighbour met across several
        faces has its whole block in the first such slot and zeros in the
        others."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        S = 2 * self.ndim + 1
        bs = self.nbasis * self.neq
        res = np.empty(self.size)
        cols = np.empty((self.nelem, S), dtype=np.int32)
        blocks = np.empty((self.nelem, S, bs, bs))
        _check(lib().ice_oracle_jacobian(self.h, _dp(u), _dp(res), _ip(cols), _dp(blocks)))
        return res, cols, blocks

    def newton(self, u, tau_abs=1e-10, tau_rel=0.0, kmax=20):
        """Newton with a dense LU linear solve; returns (u, steps, ||R|| history)."""
        u = np.array(u, dtype=np.float64)
        hist = np.zeros(64)
        k = lib().ice_oracle_newton(self.h, _dp(u), tau_abs, tau_rel, kmax, _dp(hist), 64)
        if k < 0:
            _check(-1)
        return u, k, hist[:min(k, 64)]

    def cfl_dt(self, cfl):
        return lib().ice_oracle_cfl_dt(self.h, cfl)

    def trace_points(self):
        n = self.nface * self.nfqp
        xl = np.empty((n, self.ndim))
        xr = np.empty((n, self.ndim))
        _check(lib().ice_oracle_trace_points(self.h, _dp(xl), _dp(xr)))
        return xl.reshape(self.nface, self.nfqp, self.ndim), xr.reshape(self.nface, self.nfqp, self.ndim)

    def face_geometry(self):
        n = self.nface * self.nfqp
        x = np.empty((n, self.ndim))
        nrm = np.empty((n, self.ndim))
        sg = np.empty(n)
        _check(lib().ice_oracle_face_geometry(self.h, _dp(x), _dp(nrm), _dp(sg)))
        sh = (self.nface, self.nfqp)
        return x.reshape(*sh, self.ndim), nrm.reshape(*sh, self.ndim), sg.reshape(sh)

    def qp_coords(self):
        xq = np.empty((self.nelem, self.nqp, self.ndim))
        _check(lib().ice_oracle_qp_coords(self.h, _dp(xq)))
        return xq

    def mass_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(self.size)
        _check(lib().ice_oracle_mass_apply(self.h, _dp(r), _dp(out)))
        return out


def lagrange_1d(p, xi):
    N, dN, d2N = (np.zeros(p + 1) for _ in range(3))
    _check(lib().ice_oracle_lagrange_1d(p, xi, _dp(N), _dp(dN), _dp(d2N)))
    return N, dN, d2N


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    _check(lib().ice_oracle_gauss_legendre(n, _dp(x), _dp(w)))
    return x, w


def tensor_basis(ndim, p, xi):
    nb = (p + 1) ** ndim
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    B, G, H = np.zeros(nb), np.zeros(nb * ndim), np.zeros(nb * ndim * ndim)
    _check(lib().ice_oracle_tensor_basis(ndim, p, _dp(xi), _dp(B), _dp(G), _dp(H)))
    return B, G.reshape(nb, ndim), H.reshape(nb, ndim, ndim)


def element_jacobian(ndim, X, xi):
    X = np.ascontiguousarray(X, dtype=np.float64)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    J = np.zeros(ndim * ndim)
    _check(lib().ice_oracle_element_jacobian(ndim, _dp(X), _dp(xi), _dp(J)))
    return J.reshape(ndim, ndim)


def trace_ref(ndim, face_nr, orient, s):
    s = np.ascontiguousarray(np.atleast_1d(s), dtype=np.float64)
    if s.size == 0:
        s = np.zeros(1)
    xi = np.zeros(ndim)
    _check(lib().ice_oracle_trace_ref(ndim, face_nr, orient, _dp(s), _dp(xi)))
    return xi


def ns_flux(params, ndim, u, gradu):
    u = np.ascontiguousarray(u, dtype=np.float64)
    gradu = np.ascontiguousarray(gradu, dtype=np.float64)
    F = np.zeros((ndim + 2) * ndim)
    _check(lib().ice_oracle_ns_flux(C.byref(params), ndim, _dp(u), _dp(gradu), _dp(F)))
    return F.reshape(ndim + 2, ndim)


def vanleer(params, ndim, uL, uR, n):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (uL, uR, n)]
    f = np.zeros(ndim + 2)
    _check(lib().ice_oracle_vanleer(C.byref(params), ndim, _dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), _dp(f)))
    return f


def diff_flux(params, ndim, u, gradu, n):
    neq = 1 if params.flux == abi.FLUX_BURGERS else ndim + 2
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (u, gradu, n)]
    f = np.zeros(neq)
    _check(lib().ice_oracle_diff_flux(C.byref(params), ndim, _dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]), _dp(f)))
    return f


def burgers_flux(params, ndim, u, gradu):
    """Physical Burgers flux F[d] and the wave-speed norm it records."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (u, gradu)]
    F = np.zeros(ndim)
    lam = np.zeros(1)
    _check(lib().ice_oracle_burgers_flux(C.byref(params), ndim, _dp(arrs[0]), _dp(arrs[1]), _dp(F), _dp(lam)))
    return F, float(lam[0])


def apply_bc(params, ndim, uL, graduL, n, bctype, bcflag=0):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (uL, graduL, n)]
    uR = np.zeros(ndim + 2)
    gR = np.zeros((ndim + 2) * ndim)
    _check(lib().ice_oracle_apply_bc(C.byref(params), ndim, _dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]),
                                     abi.BC[bctype] if isinstance(bctype, str) else bctype, bcflag, _dp(uR), _dp(gR)))
    return uR, gR.reshape(ndim + 2, ndim)
