"""Pins of the oracle's residual and time stepping: the reference's published
1D accuracy table (doc/sine_diffusion_1d_study.md:35-43), free-stream
preservation, conservation and linearity invariants, and RK3 properties."""
import math

import numpy as np
import pytest

from iceicle_amd import abi, workloads
from oracle import oracle as O

# doc/sine_diffusion_1d_study.md:35-43 (Standard DDG), rows nx = 2, 4, 8, 16
PUBLISHED_DDG = {
    2: [0.262893571, 0.0412573649, 0.0243959323, 0.000824445723],
    4: [0.0636081844, 0.0142770782, 0.00121340947, 9.13430722e-05],
    8: [0.0165579431, 0.00190644295, 6.954314e-05, 3.05927449e-06],
    16: [0.00417741741, 0.000243693789, 4.21576248e-06, 9.76500519e-08],
}
# relative agreement reached by the FP64 restatement (see the design notes):
# P1/P2 agree to the printed digits except on the 2-element mesh; the smallest
# P4 errors sit at the time-integration / round-off floor.
PIN_TOL = {1: 4e-7, 2: 3e-7, 3: 1e-6, 4: 1e-4}


# doc/sine_diffusion_1d_study.md:45-55 (DDGIC, --ddgic_mult=0.5), rows nx = 2..16
PUBLISHED_DDGIC = {
    2: [0.265591462, 0.0419605773, 0.024368788, 0.000803681355],
    4: [0.0824449545, 0.0145741899, 0.00120722602, 9.18430908e-05],
    8: [0.027170573, 0.00193508108, 6.88325795e-05, 3.07358828e-06],
    16: [0.00765793611, 0.000245728865, 4.17464082e-06, 9.79118744e-08],
}


# doc/sine_diffusion_1d_study.md:57-66 (Interior Penalty, --interior_penalty:
# the DDG flux without its Hessian term, beta_1 = 0), rows nx = 2..16
PUBLISHED_IP = {
    2: [0.262893571, 0.0475606773, 0.0290106132, 0.000952407674],
    4: [0.0636081844, 0.0175930309, 0.00140523714, 0.000111911005],
    8: [0.0165579431, 0.00390484158, 8.63017364e-05, 6.16326363e-06],
    16: [0.00417741741, 0.000942627192, 5.38785297e-06, 3.73127053e-07],
}
IP_TOL = {1: 4e-7, 2: 1e-6, 3: 1e-6, 4: 1e-4}


def sine_diffusion_l2(nx, p, sigma=0.0, ddgic_form=abi.DDGIC_SOURCE, interior_penalty=False):
    """sine_diffusion_1d.cpp:23-184: heat equation through Burgers (mu = 1),
    Dirichlet 0 on [0, 2 pi], IC sin x (L2 projection), RK3-TVD with
    dt = 1e-4 (2 pi / nx)^2 to t = 1, L2 error against exp(-1) sin x."""
    m = abi.mesh_desc(1, [nx], [0.0], [2 * math.pi], ["DIRICHLET", "DIRICHLET"])
    prm = abi.default_params(abi.FLUX_BURGERS)
    prm.mu = 1.0
    prm.sigma_ic = sigma
    prm.ddgic_form = ddgic_form
    prm.interior_penalty = int(interior_penalty)
    prm.dirichlet[0] = abi.Field.make("CONSTANT", [0.0])
    o = O.Oracle(m, abi.space_desc(p), prm)
    u = o.project(abi.Field.make("SINE", [1.0, 1.0, 0.0]))
    u, _ = o.rk3_tvd_tfinal(u, 1e-4 * (2 * math.pi / nx) ** 2, 1.0)
    return o.l2_error(abi.Field.make("EXP_DECAY_SINE", [1.0, 1.0]), u)


@pytest.mark.parametrize("nx", [2, 4, 8, 16])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_published_sine_diffusion_table(nx, p):
    err = sine_diffusion_l2(nx, p)
    ref = PUBLISHED_DDG[nx][p - 1]
    assert abs(err - ref) / ref <= PIN_TOL[p], (nx, p, err, ref)


@pytest.mark.parametrize("nx", [2, 4, 8, 16])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_published_interior_penalty_table(nx, p):
    """The interior-penalty column (beta_1 = 0: no Hessian term in the DDG
    gradient; P1 coincides with standard DDG) to the same digits as the
    standard-DDG column."""
    err = sine_diffusion_l2(nx, p, interior_penalty=True)
    ref = PUBLISHED_IP[nx][p - 1]
    assert abs(err - ref) / ref <= IP_TOL[p], (nx, p, err, ref)
    if p >= 2 and nx >= 8:
        assert abs(err - PUBLISHED_DDG[nx][p - 1]) / ref > 1e-2  # the Hessian term matters


@pytest.mark.parametrize("nx", [2, 4, 8, 16])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_published_ddgic_table(nx, p):
    """The DDGIC column (sigma_ic = 0.5) to the digits of the standard-DDG
    column, by the published reading (abi.DDGIC_PUBLISHED): sigma on each
    side of interior faces, no Dirichlet-face term, and each side's test
    gradient averaged between the face point and its mirror through the
    element (in 1D the mean of the test function's derivative at the
    element's two ends; P1 gradients are constant, so only P2-P4 see the
    mirror).  The form as conservation_law.hpp:580-610 and :750-780 now read
    (sigma/2 per side at the face point, plus the Dirichlet term) misses the
    column by 1-20 %."""
    err = sine_diffusion_l2(nx, p, sigma=0.5, ddgic_form=abi.DDGIC_PUBLISHED)
    ref = PUBLISHED_DDGIC[nx][p - 1]
    assert abs(err - ref) / ref <= PIN_TOL[p], (nx, p, err, ref)
    if nx in (4, 8):
        src = sine_diffusion_l2(nx, p, sigma=0.5, ddgic_form=abi.DDGIC_SOURCE)
        assert abs(src - ref) / ref > 1e-2, (nx, p, src, ref)


def _oracle(ndim, nel, order, params, bct, perturb, coupled=True, seed=5):
    m = abi.mesh_desc(ndim, nel, [0.0] * ndim, [1.0] * ndim, bct,
                      periodic_mode=abi.PERIODIC_COUPLED if coupled else abi.PERIODIC_REFERENCE)
    coords = workloads.perturbed_nodes(m, perturb, seed)
    m = abi.mesh_desc(ndim, nel, [0.0] * ndim, [1.0] * ndim, bct, coord=coords,
                      periodic_mode=m.periodic_mode)
    return O.Oracle(m, abi.space_desc(order), params), (m, coords)


@pytest.mark.parametrize("ndim,order", [(1, 3), (2, 2), (2, 4), (3, 2)])
def test_burgers_free_stream_and_conservation(ndim, order):
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = 0.05
    for d in range(3):
        p.a[d], p.b[d] = 0.3 - 0.2 * d, 1.0
    o, keep = _oracle(ndim, [3] * ndim, order, p, ["PERIODIC"] * (2 * ndim), 0.2)
    # a constant state is an exact steady solution on any mesh
    r = o.residual(np.full(o.size, 0.7))
    assert np.abs(r).max() < 1e-13
    # the Lagrange basis sums to one: the total residual is the sum of all
    # fluxes through periodic faces and vanishes for any state
    rng = np.random.default_rng(ndim)
    r = o.residual(rng.uniform(-1, 1, o.size))
    assert abs(r.sum()) < 1e-12 * np.abs(r).max() * o.size


@pytest.mark.parametrize("flux,ndim", [(abi.FLUX_EULER, 3), (abi.FLUX_NAVIER_STOKES, 2)])
def test_ns_free_stream(flux, ndim):
    p = abi.default_params(flux)
    p.visc_mu = 0.01
    o, keep = _oracle(ndim, [3] * ndim, 2, p, ["PERIODIC"] * (2 * ndim), 0.15)
    state = np.array([1.2] + [0.3, -0.2, 0.1][:ndim] + [2.9])
    u = np.tile(state, o.nelem * o.nbasis)
    assert np.abs(o.residual(u)).max() < 1e-12


def test_linear_advection_diffusion_is_linear():
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = 0.02
    p.a[0], p.a[1] = 0.7, -0.4
    o, keep = _oracle(2, [4, 3], 3, p, ["PERIODIC"] * 4, 0.2)
    rng = np.random.default_rng(3)
    u, v = rng.normal(size=o.size), rng.normal(size=o.size)
    lhs = o.residual(2.0 * u - 3.0 * v)
    rhs = 2.0 * o.residual(u) - 3.0 * o.residual(v)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(rhs).max()


def test_periodic_reference_mode_has_no_boundary_flux():
    # conservation_law.hpp:1043-1051: PERIODIC falls to the default branch, and
    # BurgersFlux has no apply_bc, so periodic boundary faces add nothing
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = 0.01
    p.b[0] = p.b[1] = 1.0
    ref, keep1 = _oracle(2, [3, 3], 2, p, ["PERIODIC"] * 4, 0.1, coupled=False)
    u = np.random.default_rng(0).uniform(-1, 1, ref.size)
    r = ref.residual(u)
    # same mesh with EXTRAPOLATION everywhere differs exactly by boundary terms
    m = abi.mesh_desc(2, [3, 3], [0, 0], [1, 1], ["EXTRAPOLATION"] * 4, coord=keep1[1])
    r2 = O.Oracle(m, abi.space_desc(2), p).residual(u)
    interior = np.ones(ref.nelem, dtype=bool)
    for e in range(ref.nelem):
        i, j = e % 3, e // 3
        interior[e] = 0 < i < 2 and 0 < j < 2
    rr = r.reshape(ref.nelem, -1)
    rr2 = r2.reshape(ref.nelem, -1)
    np.testing.assert_allclose(rr[interior], rr2[interior], atol=1e-14)
    assert np.abs(rr[~interior] - rr2[~interior]).max() > 1e-6


def test_rk3_ssp_reference_mode_is_forward_euler():
    # ssp_rk3.hpp:132-137 evaluates all stages at u^n: u^{n+1} = u^n + dt M^-1 R(u^n)
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = 0.01
    p.b[0] = p.b[1] = 1.0
    o, keep = _oracle(2, [4, 4], 2, p, ["PERIODIC"] * 4, 0.1)
    u0 = np.random.default_rng(2).uniform(-1, 1, o.size)
    dt = 1e-3
    u1 = o.rk3_ssp(u0, dt, 1, textbook=False)
    fe = u0 + dt * o.mass_apply(o.residual(u0))
    np.testing.assert_allclose(u1, fe, atol=1e-14)


def test_rk3_ssp_textbook_third_order():
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = 0.0
    p.a[0], p.a[1] = 1.0, 0.5
    o, keep = _oracle(2, [4, 4], 2, p, ["PERIODIC"] * 4, 0.0)
    u0 = o.project(abi.Field.make("SINUSOIDS", [1, 1.0, 1, 1, 0]))
    T = 0.04
    ref = o.rk3_ssp(u0, T / 64, 64)
    errs = [np.abs(o.rk3_ssp(u0, T / n, n) - ref).max() for n in (4, 8)]
    rate = math.log2(errs[0] / errs[1])
    assert 2.7 < rate < 3.4, (errs, rate)


@pytest.mark.parametrize("case", ["ns2d_dirichlet", "burgers_ddgic_mixed", "burgers_periodic_ref", "euler3d"])
def test_parallel_oracle_bit_identical_to_serial(case):
    """The all-core oracle (the CPU baseline on every host core) sums each
    element's face and volume contributions in the serial loop's order, so
    its residual equals the serial one bit for bit."""
    from iceicle_amd import workloads
    from tests import cases
    if case == "ns2d_dirichlet":
        mesh, order, params, _ = workloads.ns_mms_2d(6, 3)
        o = O.Oracle(mesh, abi.space_desc(order), params)
    elif case == "burgers_ddgic_mixed":
        p = cases.burgers_params(2, sigma_ic=0.5, a=(0.3, -0.2, 0))
        m = abi.mesh_desc(2, [5, 4], [0, 0], [1, 1], ["DIRICHLET", "NEUMANN", "EXTRAPOLATION", "DIRICHLET"],
                          bcflag=[0, 1, 0, 1])
        o = O.Oracle(m, abi.space_desc(2), p)
    elif case == "burgers_periodic_ref":
        m = abi.mesh_desc(2, [4, 5], [0, 0], [1, 1], ["PERIODIC"] * 4)
        o = O.Oracle(m, abi.space_desc(3), cases.burgers_params(2))
    else:
        mesh, order, params, _ = workloads.euler_vortex_3d(3, 2)
        o = O.Oracle(mesh, abi.space_desc(order), params)
    neq = o.size // (o.nelem * o.nbasis)
    ndim = 3 if case == "euler3d" else 2
    u = cases.random_state(abi.FLUX_BURGERS if neq == 1 else abi.FLUX_EULER, ndim, o.nelem, o.nbasis, 4)
    r1 = o.residual(u)
    for nt in (1, 3, 8):
        assert np.array_equal(o.residual_parallel(u, nt), r1), (case, nt)
