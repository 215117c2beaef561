"""The matrix-core element kernel of the mixed-precision residual (3D P3
Euler, tile 3): against the CPU oracle and the FP64 residual, with the same
FP32 bar as the VALU mixed kernel (MIXED_TOL of test_gpu_residual), and
against the VALU mixed kernel (same arithmetic, different summation order).
Meshes whose element count is not a multiple of the 4-element workgroup,
periodic coupling, every Euler boundary type, perturbed nodes, and the
isentropic vortex state of the 3D Euler workload."""
import numpy as np
import pytest

from iceicle_amd import abi, workloads
from tests import cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MIXED_TOL = 2e-5
TILE_MFMA = 3


def _space(nel, bct, coupled, perturb, seed=7):
    mode = abi.PERIODIC_COUPLED if coupled else abi.PERIODIC_REFERENCE
    m = abi.mesh_desc(3, nel, [0.0] * 3, [1.0] * 3, bct, periodic_mode=mode)
    if perturb:
        c = cases.perturbed_coords(m, perturb, seed)
        m = abi.mesh_desc(3, nel, [0.0] * 3, [1.0] * 3, bct, coord=c, periodic_mode=mode)
    return m


def _compare(m, params, u_host):
    from iceicle_amd import engine
    from oracle.oracle import Oracle
    space = engine.FESpace(m, 3)
    disc = engine.ConservationLawDDG(space, params)
    u = torch.from_numpy(u_host).cuda()
    r64 = torch.full_like(u, float("nan"))
    r0 = torch.full_like(u, float("nan"))
    r3 = torch.full_like(u, float("nan"))
    engine.form_residual(space, disc, u, r64)
    disc.set_tile(0)
    engine.form_residual(space, disc, u, r0, precision="mixed")
    disc.set_tile(TILE_MFMA)
    engine.form_residual(space, disc, u, r3, precision="mixed")
    # the FP64 residual is unaffected by tile 3
    r64b = torch.full_like(u, float("nan"))
    engine.form_residual(space, disc, u, r64b)
    torch.cuda.synchronize()
    assert torch.equal(r64, r64b)
    ref = Oracle(m, abi.space_desc(3), params).residual(u_host)
    e64 = cases.rel_inf(r64.cpu().numpy(), ref)
    e0 = cases.rel_inf(r0.cpu().numpy(), ref)
    e3 = cases.rel_inf(r3.cpu().numpy(), ref)
    d30 = cases.rel_inf(r3.cpu().numpy(), r0.cpu().numpy())
    print(f"nelem={space.nelem}: fp64 {e64:.2e}, mixed VALU {e0:.2e}, mixed MFMA {e3:.2e}, MFMA vs VALU {d30:.2e}",
          flush=True)
    assert e64 <= 1e-12
    assert bool(torch.isfinite(r3).all())
    assert e3 <= MIXED_TOL and d30 <= MIXED_TOL
    return e0, e3


def test_mfma_coupled_periodic_partial_block():
    p = cases.ns_params(3, flux=abi.FLUX_EULER)
    m = _space([5, 3, 3], ["PERIODIC"] * 6, True, 0.1)
    u = cases.random_state(abi.FLUX_EULER, 3, 45, 64, 3)
    _compare(m, p, u)


def test_mfma_boundary_types():
    p = cases.ns_params(3, flux=abi.FLUX_EULER)
    bct = ["SLIP_WALL", "RIEMANN", "DIRICHLET", "EXTRAPOLATION", "SLIP_WALL", "DIRICHLET"]
    m = _space([3, 2, 3], bct, False, 0.1)
    u = cases.random_state(abi.FLUX_EULER, 3, 18, 64, 4)
    _compare(m, p, u)


def test_mfma_vortex_state():
    """The 3D Euler workload's state (isentropic vortex, density perturbed
    by 1%) on a coarse 6^3 version of its mesh."""
    from iceicle_amd import engine
    mesh, order, params, ic = workloads.euler_vortex_3d(6, 3)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    u = torch.empty(disc.size, dtype=torch.float64, device="cuda")
    engine.project(disc, ic, u)
    uh = workloads.density_perturbed(u.cpu().numpy(), disc.neq, workloads.SEED0 + 3)
    del disc, space
    _compare(mesh, params, uh)


def test_mfma_tile_rejected_elsewhere_is_default():
    """Tile 3 on a discretization without the matrix-core kernel (2D) gives
    the default tile's residual bit for bit."""
    from iceicle_amd import engine
    p = cases.ns_params(2, flux=abi.FLUX_EULER)
    m = abi.mesh_desc(2, [4, 3], [0.0, 0.0], [1.0, 1.0], ["SLIP_WALL", "RIEMANN", "DIRICHLET", "EXTRAPOLATION"])
    space = engine.FESpace(m, 3)
    disc = engine.ConservationLawDDG(space, p)
    u = torch.from_numpy(cases.random_state(abi.FLUX_EULER, 2, 12, 16, 2)).cuda()
    a, b = torch.empty_like(u), torch.empty_like(u)
    disc.set_tile(0)
    engine.form_residual(space, disc, u, a, precision="mixed")
    disc.set_tile(TILE_MFMA)
    engine.form_residual(space, disc, u, b, precision="mixed")
    torch.cuda.synchronize()
    assert torch.equal(a, b)
