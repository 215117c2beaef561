"""Named synthetic workloads (mesh + space + discretization parameters).

Each builder returns (mesh_desc, order, params, keepalive) where keepalive
holds the numpy arrays the descriptor points into.  Seeds are
0x1CE1C1E + workload index so every run sees the same perturbed mesh.
"""
import math

import numpy as np

from . import abi

SEED0 = 0x1CE1C1E


def uniform_axes(mesh):
    axes = []
    for d in range(mesh.ndim):
        n = mesh.nelem[d]
        dx = (mesh.xmax[d] - mesh.xmin[d]) / n
        x = [mesh.xmin[d]]
        for _ in range(n):
            x.append(x[-1] + dx)  # cumulative spacing, as the mesh generator
        axes.append(np.array(x))
    return axes


def perturbed_nodes(mesh, frac, seed):
    """Node coordinates (x fastest) with interior nodes moved by U(-frac h, frac h)
    per axis; boundary nodes stay on the box."""
    ndim = mesh.ndim
    grids = np.meshgrid(*uniform_axes(mesh), indexing="ij")
    coords = np.stack([g.transpose(*reversed(range(ndim))).ravel() for g in grids], axis=1)
    rng = np.random.default_rng(seed)
    h = np.array([(mesh.xmax[d] - mesh.xmin[d]) / mesh.nelem[d] for d in range(ndim)])
    shift = rng.uniform(-frac, frac, size=coords.shape) * h
    interior = np.ones(coords.shape[0], dtype=bool)
    for d in range(ndim):
        interior &= (coords[:, d] > mesh.xmin[d] + 1e-12) & (coords[:, d] < mesh.xmax[d] - 1e-12)
    coords[interior] += shift[interior]
    return np.ascontiguousarray(coords)


def ns_mms_params(mu=0.003):
    """2D compressible NS, gamma 1.4, Pr 0.72, unit reference parameters,
    constant viscosity mu (Re = 100 at U = 0.3, L = 1), manufactured source,
    Dirichlet = manufactured state."""
    p = abi.default_params(abi.FLUX_NAVIER_STOKES)
    p.visc_mu = mu
    p.dirichlet[0] = abi.Field.make("NS_MMS_STATE")
    p.source = abi.Field.make("NS_MMS_SOURCE")
    return p


def ns_mms_2d(n=128, order=3, perturb=0.05):
    """Workload: 2D NS manufactured steady problem on n x n perturbed quads."""
    m = abi.mesh_desc(2, [n, n], [0.0, 0.0], [1.0, 1.0], ["DIRICHLET"] * 4)
    coords = perturbed_nodes(m, perturb, SEED0 + 1)
    m = abi.mesh_desc(2, [n, n], [0.0, 0.0], [1.0, 1.0], ["DIRICHLET"] * 4, coord=coords)
    return m, order, ns_mms_params(), coords


def burgers_2d(n=32, order=2, mu=0.01, periodic_mode=abi.PERIODIC_COUPLED, seed=SEED0):
    """Workload: 2D viscous Burgers (b = (1,1)) on a periodic n x n grid."""
    m = abi.mesh_desc(2, [n, n], [0.0, 0.0], [1.0, 1.0], ["PERIODIC"] * 4, periodic_mode=periodic_mode)
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = mu
    p.b[0] = p.b[1] = 1.0
    rng = np.random.default_rng(seed)
    amps = rng.uniform(-1.0, 1.0, size=4)
    modes = [(1, 0), (0, 1), (1, 1), (2, 1)]
    ic = [4]
    for a, (mx, my) in zip(amps, modes):
        ic += [a, mx, my, 0]
    return m, order, p, abi.Field.make("SINUSOIDS", ic)


def euler_vortex_3d(n=64, order=3, half=20.0):
    """Workload: 3D Euler on n^3 hexes of [-half, half]^3 with periodic faces
    coupled, P{order}; isentropic vortex (examples/conservation_law/
    isentropic_vortex.lua coefficients: K = 5, centre (-10, -10), free stream
    (2, 2), c_inf = 1) extruded in z.  Returns (mesh, order, params, ic)."""
    m = abi.mesh_desc(3, [n, n, n], [-half] * 3, [half] * 3, ["PERIODIC"] * 6,
                      periodic_mode=abi.PERIODIC_COUPLED)
    p = abi.default_params(abi.FLUX_EULER)
    return m, order, p, abi.Field.make("ISENTROPIC_VORTEX", [-10.0, -10.0, 5.0, 1.0, 2.0, 2.0])


def density_perturbed(u, neq, seed, amp=0.01):
    """Multiply the density coefficients by 1 + U(-amp, amp) (seeded)."""
    u = np.array(u, dtype=np.float64).reshape(-1, neq)
    rng = np.random.default_rng(seed)
    u[:, 0] *= 1.0 + rng.uniform(-amp, amp, size=u.shape[0])
    return u.ravel()


def mms_state_perturbed(space, neq, seed, amp=1e-3):
    """Nodal interpolant of the manufactured state with a small seeded
    perturbation, so the residual is far from equilibrium."""
    from . import engine
    u = space.interpolate(abi.Field.make("NS_MMS_STATE"), neq)
    rng = np.random.default_rng(seed)
    return u * (1.0 + amp * rng.uniform(-1.0, 1.0, size=u.shape))


def residual_bytes(ndim, order, neq, nelem, nnode, nface, s=8):
    """Algorithmic HBM bytes of one residual: state read + residual write,
    vertex coordinates, P1 connectivity and face records."""
    nb = (order + 1) ** ndim
    return 2 * s * neq * nelem * nb + 8 * ndim * nnode + 4 * (2 ** ndim) * nelem + 12 * nface
