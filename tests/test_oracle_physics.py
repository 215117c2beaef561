"""Pins of the oracle's flux functors against the reference's known answers."""
import numpy as np
import pytest

from iceicle_amd import abi
from oracle import oracle as O


@pytest.mark.parametrize("ndim", [2, 3])
def test_vanleer_consistency(ndim):
    # test_ns.cpp:66-107: VanLeer(u, u, e_k) equals the physical Euler flux
    # column k (zero gradient) to 1e-10, DimensionlessSutherlands viscosity
    p = abi.default_params(abi.FLUX_NAVIER_STOKES)
    p.viscosity_kind = abi.VISC_DIMENSIONLESS_SUTHERLAND
    for uval in (-1.0, -0.001, 0.0, 0.001, 1.0):
        u = np.array([1.0] + [uval] * ndim + [1.4])
        F = O.ns_flux(p, ndim, u, np.zeros((ndim + 2, ndim)))
        for k in range(ndim):
            n = np.zeros(ndim)
            n[k] = 1.0
            np.testing.assert_allclose(O.vanleer(p, ndim, u, u, n), F[:, k], atol=1e-10)


def test_vanleer_supersonic_upwinding():
    p = abi.default_params(abi.FLUX_EULER)
    uL = np.array([1.0, 3.0, 0.0, 1.0 / 0.4 / 1.4 + 4.5])  # Mach ~ 3 to the right
    # right state also supersonic to the right (Mach ~ 2.7): its split flux is zero
    uR = np.array([0.8, 0.8 * 2.5, 0.08, 0.5 / 0.4 + 0.5 * 0.8 * (2.5 ** 2 + 0.1 ** 2)])
    F = O.ns_flux(p, 2, uL, np.zeros((4, 2)))
    np.testing.assert_allclose(O.vanleer(p, 2, uL, uR, [1.0, 0.0]), F[:, 0], atol=1e-13)


def test_ns_viscous_flux_linear_in_gradient():
    # viscous part of the physical flux is linear in grad u at fixed u
    p = abi.default_params(abi.FLUX_NAVIER_STOKES)
    p.visc_mu = 0.01
    rng = np.random.default_rng(1)
    u = np.array([1.1, 0.2, -0.1, 2.7])
    g1, g2 = rng.normal(size=(4, 2)), rng.normal(size=(4, 2))
    F0 = O.ns_flux(p, 2, u, np.zeros((4, 2)))
    Fa, Fb = O.ns_flux(p, 2, u, g1), O.ns_flux(p, 2, u, g2)
    Fab = O.ns_flux(p, 2, u, 2 * g1 - 0.5 * g2)
    np.testing.assert_allclose(Fab - F0, 2 * (Fa - F0) - 0.5 * (Fb - F0), atol=1e-14)
    # mass flux has no viscous part
    np.testing.assert_allclose(Fa[0], F0[0], atol=0)
