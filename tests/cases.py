"""Shared problem builders for the parity tests (CPU oracle vs GPU engine)."""
import math

import numpy as np

from iceicle_amd import abi, workloads


def perturbed_coords(mesh, frac, seed):
    """Interior nodes moved by U(-frac h, frac h) per axis (see workloads)."""
    return workloads.perturbed_nodes(mesh, frac, seed)


def random_state(flux, ndim, nelem, nbasis, seed):
    rng = np.random.default_rng(seed)
    if flux == abi.FLUX_BURGERS:
        return rng.uniform(-1.0, 1.0, size=nelem * nbasis)
    m = ndim + 2
    u = np.empty((nelem * nbasis, m))
    rho = 1.0 + 0.2 * rng.uniform(-1, 1, size=nelem * nbasis)
    vel = 0.3 * rng.uniform(-1, 1, size=(nelem * nbasis, ndim))
    p = 1.0 / 1.4 * (1.0 + 0.2 * rng.uniform(-1, 1, size=nelem * nbasis))
    u[:, 0] = rho
    u[:, 1:1 + ndim] = rho[:, None] * vel
    u[:, 1 + ndim] = p / 0.4 + 0.5 * rho * (vel ** 2).sum(axis=1)
    return u.ravel()


def burgers_params(ndim, mu=0.01, a=(0, 0, 0), b=(1, 1, 1), sigma_ic=0.0, interior_penalty=False):
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = mu
    for d in range(3):
        p.a[d] = a[d]
        p.b[d] = b[d]
    p.sigma_ic = sigma_ic
    p.interior_penalty = int(interior_penalty)
    p.dirichlet[0] = abi.Field.make("SINUSOIDS", [2, 0.7, 1, 0, 0, -0.4, 1, 1, 0])
    p.dirichlet[1] = abi.Field.make("CONSTANT", [0.25])
    p.neumann[0] = abi.Field.make("SINUSOIDS", [1, 0.3, 0, 1, 0])
    p.neumann[1] = abi.Field.make("CONSTANT", [-0.5])
    return p


def ns_params(ndim, flux=abi.FLUX_NAVIER_STOKES, mu=0.003, source=False, textbook=False):
    p = abi.default_params(flux)
    p.visc_mu = mu
    p.ns_face_energy_textbook = int(textbook)
    p.isothermal_T[0] = (1.0 / 1.4) / 287.052874  # T = p / (rho R) for p = 1/gamma, rho = 1
    p.freestream[0] = 1.0
    p.freestream[1] = 0.3
    p.freestream[2] = 1.0
    p.freestream[5] = (1.0 / 1.4) / 287.052874
    if ndim == 2:
        p.dirichlet[0] = abi.Field.make("NS_MMS_STATE")
        if source:
            p.source = abi.Field.make("NS_MMS_SOURCE")
    else:
        vals = [1.0, 0.3] + [0.0] * (ndim - 1) + [1.0 / 1.4 / 0.4 + 0.5 * 0.09]
        p.dirichlet[0] = abi.Field.make("CONSTANT", vals)
    return p


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
