"""GPU explicit time integration, inverse mass matrix, projection and L2
error against the CPU oracle."""
import math

import numpy as np
import pytest

from iceicle_amd import abi, workloads
from tests import cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _setup(mesh, order, params):
    from iceicle_amd import engine
    from oracle.oracle import Oracle
    o = Oracle(mesh, abi.space_desc(order), params)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    return o, space, disc


@pytest.mark.parametrize("ndim,order,neq_flux", [(1, 3, abi.FLUX_BURGERS), (2, 2, abi.FLUX_BURGERS),
                                                 (2, 4, abi.FLUX_NAVIER_STOKES), (3, 2, abi.FLUX_EULER)])
def test_mass_inverse(ndim, order, neq_flux):
    from iceicle_amd import engine
    m = abi.mesh_desc(ndim, [3] * ndim, [0] * ndim, [1] * ndim, ["EXTRAPOLATION"] * (2 * ndim))
    coords = workloads.perturbed_nodes(m, 0.2, 4)
    m = abi.mesh_desc(ndim, [3] * ndim, [0] * ndim, [1] * ndim, ["EXTRAPOLATION"] * (2 * ndim), coord=coords)
    o, space, disc = _setup(m, order, abi.default_params(neq_flux))
    r = np.random.default_rng(1).normal(size=o.size)
    ref = o.mass_apply(r)
    out = torch.empty(o.size, dtype=torch.float64, device="cuda")
    engine.mass_inverse_apply(disc, torch.from_numpy(r).cuda(), out)
    torch.cuda.synchronize()
    err = cases.rel_inf(out.cpu().numpy(), ref)
    print(f"mass inverse ndim={ndim} p={order} err={err:.2e}")
    assert err <= 1e-12


def test_projection_and_l2_error():
    from iceicle_amd import engine
    m = abi.mesh_desc(2, [5, 4], [0, 0], [1, 1], ["DIRICHLET"] * 4)
    coords = workloads.perturbed_nodes(m, 0.15, 9)
    m = abi.mesh_desc(2, [5, 4], [0, 0], [1, 1], ["DIRICHLET"] * 4, coord=coords)
    o, space, disc = _setup(m, 3, workloads.ns_mms_params())
    f = abi.Field.make("NS_MMS_STATE")
    ref = o.project(f)
    u = torch.empty(o.size, dtype=torch.float64, device="cuda")
    engine.project(disc, f, u)
    assert cases.rel_inf(u.cpu().numpy(), ref) <= 1e-12
    e_gpu = engine.l2_error(disc, abi.Field.make("CONSTANT", [1.0, 0.3, 0.0, 2.0]), u)
    e_ref = o.l2_error(abi.Field.make("CONSTANT", [1.0, 0.3, 0.0, 2.0]), ref)
    assert abs(e_gpu - e_ref) <= 1e-12 * e_ref


@pytest.mark.parametrize("compat", [True, False])
@pytest.mark.parametrize("periodic_mode", [abi.PERIODIC_REFERENCE, abi.PERIODIC_COUPLED])
def test_burgers_rk3_ssp_cfl_32x32_p2(compat, periodic_mode):
    """2D viscous Burgers, 32x32 periodic, P2, RK3-SSP CFL 0.2, 100 steps."""
    from iceicle_amd import engine
    mesh, order, params, ic = workloads.burgers_2d(32, 2, periodic_mode=periodic_mode)
    o, space, disc = _setup(mesh, order, params)
    u0 = o.project(ic)
    ref, n_ref, t_ref = o.rk_solve(u0, kind=0, compat=compat, cfl=0.2, ntime=100)
    u = torch.from_numpy(u0).cuda()
    st = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_SSP, reference_compat=compat, cfl=0.2, ntime=100))
    n, t = st.solve(u)
    err = cases.rel_inf(u.cpu().numpy(), ref)
    print(f"Burgers 32x32 P2 compat={compat} mode={periodic_mode} steps={n} t={t:.6e} t_ref={t_ref:.6e} err={err:.2e}")
    assert n == n_ref == 100
    assert abs(t - t_ref) <= 1e-13 * t_ref
    assert err <= 1e-11


def test_rk3_tvd_heat_1d_tfinal():
    """sine_diffusion_1d driver: RK3-TVD, fixed dt, TfinalTermination, and the
    published L2 error (doc/sine_diffusion_1d_study.md:40, nx = 8, P3)."""
    from iceicle_amd import engine
    nx, p = 8, 3
    m = abi.mesh_desc(1, [nx], [0.0], [2 * math.pi], ["DIRICHLET", "DIRICHLET"])
    prm = abi.default_params(abi.FLUX_BURGERS)
    prm.mu = 1.0
    prm.dirichlet[0] = abi.Field.make("CONSTANT", [0.0])
    o, space, disc = _setup(m, p, prm)
    u = torch.empty(o.size, dtype=torch.float64, device="cuda")
    engine.project(disc, abi.Field.make("SINE", [1, 1, 0]), u)
    dt = 1e-4 * (2 * math.pi / nx) ** 2
    st = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_TVD, dt=dt, tfinal=1.0))
    n, t = st.solve(u)
    err = engine.l2_error(disc, abi.Field.make("EXP_DECAY_SINE", [1.0, 1.0]), u)
    print(f"1D heat nx={nx} P{p}: steps={n} t={t} L2={err:.9g} (published 6.954314e-05)")
    assert abs(t - 1.0) < 1e-12
    assert abs(err - 6.954314e-05) / 6.954314e-05 < 1e-6


def test_euler_2d_cfl_explicit():
    """NS lambda_max bookkeeping (last quadrature point, reset by dt_from_cfl)."""
    from iceicle_amd import engine
    m = abi.mesh_desc(2, [6, 5], [0, 0], [1, 1], ["PERIODIC"] * 4, periodic_mode=abi.PERIODIC_COUPLED)
    prm = abi.default_params(abi.FLUX_NAVIER_STOKES)
    prm.visc_mu = 0.01
    o, space, disc = _setup(m, 2, prm)
    u0 = o.project(abi.Field.make("ISENTROPIC_VORTEX", [0.5, 0.5, 2.0, 1.0, 0.3, 0.1]))
    ref, n_ref, t_ref = o.rk_solve(u0, kind=0, cfl=0.3, ntime=20)
    u = torch.from_numpy(u0).cuda()
    st = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_SSP, cfl=0.3, ntime=20))
    n, t = st.solve(u)
    err = cases.rel_inf(u.cpu().numpy(), ref)
    print(f"NS vortex steps={n} t={t:.6e} t_ref={t_ref:.6e} err={err:.2e}")
    assert abs(t - t_ref) <= 1e-12 * t_ref and err <= 1e-11




def test_cfl_step_graph_matches_eager():
    """100 RK3-SSP steps of the 2D Burgers 32x32 P2 workload at CFL 0.2 (the
    step size recomputed on the device from the wave speeds every step; the
    wave-speed maxima are reset by the step kernels, not by memsets)
    replayed from one HIP graph equal the eager solve bit for bit, with the
    same step count and final time; a second replay changes nothing, since
    the device-side termination flag is set."""
    import torch
    from iceicle_amd import engine
    mesh, order, params, ic = workloads.burgers_2d(32, 2)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    u0 = torch.empty(disc.size, dtype=torch.float64, device="cuda")
    engine.project(disc, ic, u0)
    desc = abi.stepper_desc(abi.STEP_RK3_SSP, cfl=0.2, ntime=100)
    ua = u0.clone()
    sa = engine.ExplicitStepper(disc, desc)
    n, t = sa.solve(ua)
    ub = u0.clone()
    sb = engine.ExplicitStepper(disc, desc)
    g = engine.StepGraph([sb], [ub], 100)
    g.launch()
    torch.cuda.synchronize()
    stb = sb.status()
    assert n == 100 and stb["itime"] == 100 and stb["done"]
    assert stb["time"] == t
    assert torch.equal(ua, ub)
    keep = ub.clone()
    g.launch()
    torch.cuda.synchronize()
    assert torch.equal(keep, ub)
