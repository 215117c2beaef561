"""ctypes mirror of the plain-C descriptors in include/iceicle_amd.h.

Both the product library (libiceicle_amd.so) and the CPU oracle take these
exact structs, so a test builds one descriptor and hands it to both.
"""
import ctypes as C

import numpy as np

ICE_MAX_BCFLAG = 4

# status codes
ICE_OK = 0
ICE_ERR_ARG = -1
ICE_ERR_UNSUPPORTED = -2
ICE_ERR_HIP = -3
ICE_ERR_ALLOC = -4
ICE_ERR_STATE = -5
ICE_ERR_NUMERIC = -6

# boundary conditions (reference enum BOUNDARY_CONDITIONS, geometry/face.hpp:23-103)
BC = dict(PERIODIC=0, PARALLEL_COM=1, NEUMANN=2, DIRICHLET=3, EXTRAPOLATION=4, RIEMANN=5,
          NO_SLIP_ISOTHERMAL=6, SLIP_WALL=7, WALL_GENERAL=8, INLET=9, OUTLET=10,
          SPACETIME_PAST=11, SPACETIME_FUTURE=12, INTERIOR=13)

PERIODIC_REFERENCE = 0
PERIODIC_COUPLED = 1

FLUX_BURGERS = 0
FLUX_NAVIER_STOKES = 1
FLUX_EULER = 2

DDGIC_SOURCE = 0
DDGIC_PUBLISHED = 1

VISC_CONSTANT = 0
VISC_DIMENSIONLESS_SUTHERLAND = 1

FIELD = dict(NONE=0, CONSTANT=1, SINE=2, SINUSOIDS=3, NS_MMS_STATE=4, NS_MMS_SOURCE=5,
             ISENTROPIC_VORTEX=6, EXP_DECAY_SINE=7)

FACE_REC_FIXED = 6


class Field(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_double * 32)]

    @staticmethod
    def make(kind, params=()):
        f = Field()
        f.kind = FIELD[kind] if isinstance(kind, str) else int(kind)
        for i, v in enumerate(params):
            f.p[i] = float(v)
        return f


class MeshDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int),
        ("nelem", C.c_int * 3),
        ("xmin", C.c_double * 3),
        ("xmax", C.c_double * 3),
        ("bctype", C.c_int * 6),
        ("bcflag", C.c_int * 6),
        ("coord", C.POINTER(C.c_double)),
        ("periodic_mode", C.c_int),
    ]


class SpaceDesc(C.Structure):
    _fields_ = [("order", C.c_int), ("geometry_order", C.c_int)]


class MeshCounts(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ndim", "nelem", "nnode", "nvert_el", "nfacenode", "nface",
                                       "nface_interior")]


class SpaceInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ndim", "order", "nbasis", "nqp", "nfqp", "nelem", "nface",
                                       "nface_interior", "nface_boundary")] + [("ndof", C.c_int64)]


class DiscParams(C.Structure):
    _fields_ = [
        ("flux", C.c_int),
        ("mu", C.c_double),
        ("a", C.c_double * 3),
        ("b", C.c_double * 3),
        ("gamma", C.c_double),
        ("Rgas", C.c_double),
        ("Pr", C.c_double),
        ("ref", C.c_double * 8),
        ("viscosity_kind", C.c_int),
        ("visc_mu", C.c_double),
        ("isothermal_T", C.c_double * ICE_MAX_BCFLAG),
        ("freestream", C.c_double * 6),
        ("ns_face_energy_textbook", C.c_int),
        ("sigma_ic", C.c_double),
        ("interior_penalty", C.c_int),
        ("ddgic_form", C.c_int),
        ("dirichlet", Field * ICE_MAX_BCFLAG),
        ("neumann", Field * ICE_MAX_BCFLAG),
        ("source", Field),
    ]


STEP_RK3_SSP = 0
STEP_RK3_TVD = 1
STEP_EXPLICIT_EULER = 2


class StepperDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("reference_compat", C.c_int), ("use_cfl", C.c_int), ("dt", C.c_double),
                ("cfl", C.c_double), ("use_tfinal", C.c_int), ("tfinal", C.c_double), ("ntime", C.c_longlong)]


def stepper_desc(kind=STEP_RK3_SSP, reference_compat=False, dt=None, cfl=None, tfinal=None, ntime=0):
    s = StepperDesc()
    s.kind = kind
    s.reference_compat = int(reference_compat)
    s.use_cfl = int(cfl is not None)
    s.dt = dt or 0.0
    s.cfl = cfl or 0.0
    s.use_tfinal = int(tfinal is not None)
    s.tfinal = tfinal or 0.0
    s.ntime = ntime
    return s


def default_params(flux):
    """Reference defaults (burgers.hpp:14-40, navier_stokes.hpp:129-213, :321-340)."""
    p = DiscParams()
    p.flux = flux
    p.mu = 0.001
    p.gamma = 1.4
    p.Rgas = 287.052874
    p.Pr = 0.72
    for i in range(8):
        p.ref[i] = 1.0
    p.viscosity_kind = VISC_CONSTANT
    p.visc_mu = 1.0
    for i in range(ICE_MAX_BCFLAG):
        p.isothermal_T[i] = 1.0
    p.freestream[0] = 1.0
    p.freestream[2] = 1.0
    p.freestream[5] = 1.0
    return p


def mesh_desc(ndim, nelem, xmin, xmax, bctype, bcflag=None, coord=None, periodic_mode=PERIODIC_REFERENCE):
    """Build a MeshDesc.  `bctype` entries may be names from BC or ints.
    `coord` (optional) is a C-contiguous float64 numpy array [nnode][ndim];
    the descriptor keeps a reference to it, so it lives as long as the
    descriptor (the C struct holds only its address)."""
    m = MeshDesc()
    m.ndim = ndim
    for d in range(ndim):
        m.nelem[d] = int(nelem[d])
        m.xmin[d] = float(xmin[d])
        m.xmax[d] = float(xmax[d])
    for k in range(2 * ndim):
        b = bctype[k]
        m.bctype[k] = BC[b] if isinstance(b, str) else int(b)
        m.bcflag[k] = 0 if bcflag is None else int(bcflag[k])
    if coord is not None:
        coord = np.ascontiguousarray(coord, dtype=np.float64)
        m.coord = coord.ctypes.data_as(C.POINTER(C.c_double))
        m._coord_owner = coord
    m.periodic_mode = periodic_mode
    return m


def space_desc(order):
    s = SpaceDesc()
    s.order = order
    s.geometry_order = 1
    return s


LS_NONE = 0
LS_WOLFE = 1
LS_CORRIGAN = 2
PREC_FP64 = 0
PREC_FP32 = 1
PC_BLOCK_JACOBI = 0
PC_BLOCK_SGS = 1
PC_MULTIGRID = 2
PC_LINE = 3
PREC_FP16 = 2
FORCING_FIXED = 0
FORCING_EW = 1
NEWTON_MAX_HIST = 64


class NewtonOpts(C.Structure):
    _fields_ = [("tau_abs", C.c_double), ("tau_rel", C.c_double), ("kmax", C.c_int), ("linesearch", C.c_int),
                ("ls_kmax", C.c_int), ("ls_alpha_initial", C.c_double), ("ls_alpha_max", C.c_double),
                ("ls_c1", C.c_double), ("ls_c2", C.c_double), ("ls_alpha_min", C.c_double),
                ("jacobian_precision", C.c_int), ("gmres_restart", C.c_int), ("gmres_maxit", C.c_int),
                ("linear_rtol", C.c_double), ("refine_max", C.c_int), ("verbose", C.c_int),
                ("preconditioner", C.c_int), ("forcing", C.c_int), ("smoother_precision", C.c_int),
                ("gmres_cgs_refine", C.c_int)]


class NewtonReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("r0", C.c_double), ("rnorm", C.c_double),
                ("linear_iterations", C.c_longlong), ("refinements", C.c_int), ("singular_blocks", C.c_int),
                ("seconds", C.c_double), ("jacobian_seconds", C.c_double), ("linear_seconds", C.c_double),
                ("linesearch_seconds", C.c_double), ("nhist", C.c_int),
                ("rhist", C.c_double * NEWTON_MAX_HIST), ("alpha", C.c_double * NEWTON_MAX_HIST),
                ("linear_its", C.c_int * NEWTON_MAX_HIST)]

    def as_dict(self):
        k = self.nhist
        return dict(iterations=self.iterations, converged=bool(self.converged), r0=self.r0, rnorm=self.rnorm,
                    linear_iterations=self.linear_iterations, refinements=self.refinements,
                    singular_blocks=self.singular_blocks, seconds=self.seconds,
                    jacobian_seconds=self.jacobian_seconds, linear_seconds=self.linear_seconds,
                    linesearch_seconds=self.linesearch_seconds, rhist=list(self.rhist[:k]),
                    alpha=list(self.alpha[:k]), linear_its=list(self.linear_its[:k]))


class MlpTrainOpts(C.Structure):
    _fields_ = [("batch", C.c_int), ("epochs", C.c_int), ("lr", C.c_double), ("lr_decay", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double), ("seed", C.c_ulonglong)]


def mlp_train_opts(batch=256, epochs=100, lr=1e-3, lr_decay=1.0, beta1=0.9, beta2=0.999, eps=1e-8, seed=0):
    o = MlpTrainOpts()
    o.batch, o.epochs, o.lr, o.lr_decay = batch, epochs, lr, lr_decay
    o.beta1, o.beta2, o.eps, o.seed = beta1, beta2, eps, seed
    return o
