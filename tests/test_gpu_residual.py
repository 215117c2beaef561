"""GPU residual parity: libiceicle_amd (HIP, FP64) against the CPU oracle on
the same mesh, coefficients and state.  Bar: relative max-norm error
<= 1e-12 (the FP64 parity tolerance of the residual path)."""
import numpy as np
import pytest

from iceicle_amd import abi
from tests import cases

TOL = 1e-12

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _run(ndim, nel, order, params, bct, perturb=0.0, coupled=False, seed=0, bcflag=None):
    from iceicle_amd import engine
    from oracle.oracle import Oracle

    m = abi.mesh_desc(ndim, nel, [0.0] * ndim, [1.0] * ndim, bct, bcflag=bcflag,
                      periodic_mode=abi.PERIODIC_COUPLED if coupled else abi.PERIODIC_REFERENCE)
    coords = None
    if perturb > 0:
        coords = cases.perturbed_coords(m, perturb, seed + 11)
        m = abi.mesh_desc(ndim, nel, [0.0] * ndim, [1.0] * ndim, bct, bcflag=bcflag, coord=coords,
                          periodic_mode=m.periodic_mode)
    o = Oracle(m, abi.space_desc(order), params)
    u = cases.random_state(params.flux, ndim, o.nelem, o.nbasis, seed)
    r_ref = o.residual(u)
    space = engine.FESpace(m, order)
    disc = engine.ConservationLawDDG(space, params)
    ud = torch.from_numpy(u).cuda()
    rd = torch.full_like(ud, float("nan"))
    engine.form_residual(space, disc, ud, rd)
    torch.cuda.synchronize()
    r = rd.cpu().numpy()
    err = cases.rel_inf(r, r_ref)
    print(f"ndim={ndim} nel={nel} p={order} flux={params.flux} bc={bct} err={err:.3e}", flush=True)
    return err


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_burgers_2d_periodic_reference(order):
    p = cases.burgers_params(2)
    assert _run(2, [5, 4], order, p, ["PERIODIC"] * 4, perturb=0.15) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_burgers_2d_periodic_coupled(order):
    p = cases.burgers_params(2)
    assert _run(2, [4, 5], order, p, ["PERIODIC"] * 4, perturb=0.15, coupled=True) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_burgers_2d_mixed_bcs_ddgic(order):
    p = cases.burgers_params(2, sigma_ic=0.5, a=(0.3, -0.2, 0))
    bct = ["DIRICHLET", "NEUMANN", "EXTRAPOLATION", "DIRICHLET"]
    assert _run(2, [4, 3], order, p, bct, perturb=0.2, bcflag=[0, 1, 0, 1]) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_burgers_1d(order):
    p = cases.burgers_params(1, sigma_ic=0.5)
    assert _run(1, [6], order, p, ["DIRICHLET", "DIRICHLET"], perturb=0.2, bcflag=[0, 1]) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_burgers_3d(order):
    p = cases.burgers_params(3, a=(0.2, 0.1, -0.3))
    bct = ["PERIODIC", "DIRICHLET", "NEUMANN", "PERIODIC", "EXTRAPOLATION", "DIRICHLET"]
    assert _run(3, [3, 2, 3], order, p, bct, perturb=0.15, coupled=True, bcflag=[0, 0, 0, 0, 1, 1]) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_ns_2d_dirichlet_mms_source(order):
    p = cases.ns_params(2, source=True)
    assert _run(2, [4, 4], order, p, ["DIRICHLET"] * 4, perturb=0.05) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_ns_2d_walls_riemann(order):
    p = cases.ns_params(2)
    bct = ["RIEMANN", "SLIP_WALL", "EXTRAPOLATION", "NO_SLIP_ISOTHERMAL"]
    assert _run(2, [4, 3], order, p, bct, perturb=0.1) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_euler_3d_coupled(order):
    p = cases.ns_params(3, flux=abi.FLUX_EULER)
    bct = ["PERIODIC", "SLIP_WALL", "PERIODIC", "PERIODIC", "SLIP_WALL", "PERIODIC"]
    assert _run(3, [3, 3, 2], order, p, bct, perturb=0.1, coupled=True) <= TOL


@pytest.mark.parametrize("order", [1, 2])
def test_ns_3d(order):
    p = cases.ns_params(3)
    bct = ["DIRICHLET", "SLIP_WALL", "RIEMANN", "EXTRAPOLATION", "NO_SLIP_ISOTHERMAL", "DIRICHLET"]
    assert _run(3, [2, 3, 2], order, p, bct, perturb=0.1) <= TOL
