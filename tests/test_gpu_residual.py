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


@pytest.mark.parametrize("form", [abi.DDGIC_SOURCE, abi.DDGIC_PUBLISHED])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_burgers_2d_mixed_bcs_ddgic(order, form):
    """Both readings of the interface correction on a perturbed mesh; the
    published one averages each side's test gradient with the mirror
    point's, whose geometry differs from the face point's here."""
    p = cases.burgers_params(2, sigma_ic=0.5, a=(0.3, -0.2, 0))
    p.ddgic_form = form
    bct = ["DIRICHLET", "NEUMANN", "EXTRAPOLATION", "DIRICHLET"]
    assert _run(2, [4, 3], order, p, bct, perturb=0.2, bcflag=[0, 1, 0, 1]) <= TOL


@pytest.mark.parametrize("order", [2, 3, 4])
def test_burgers_1d_ddgic_published(order):
    p = cases.burgers_params(1, sigma_ic=0.5)
    p.ddgic_form = abi.DDGIC_PUBLISHED
    assert _run(1, [6], order, p, ["DIRICHLET", "DIRICHLET"], perturb=0.2, bcflag=[0, 1]) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_burgers_1d(order):
    p = cases.burgers_params(1, sigma_ic=0.5)
    assert _run(1, [6], order, p, ["DIRICHLET", "DIRICHLET"], perturb=0.2, bcflag=[0, 1]) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3, 4])
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


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_euler_3d_coupled(order):
    p = cases.ns_params(3, flux=abi.FLUX_EULER)
    bct = ["PERIODIC", "SLIP_WALL", "PERIODIC", "PERIODIC", "SLIP_WALL", "PERIODIC"]
    assert _run(3, [3, 3, 2], order, p, bct, perturb=0.1, coupled=True) <= TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_ns_3d(order):
    p = cases.ns_params(3)
    bct = ["DIRICHLET", "SLIP_WALL", "RIEMANN", "EXTRAPOLATION", "NO_SLIP_ISOTHERMAL", "DIRICHLET"]
    assert _run(3, [2, 3, 2], order, p, bct, perturb=0.1) <= TOL


@pytest.mark.parametrize("form", [abi.DDGIC_SOURCE, abi.DDGIC_PUBLISHED])
@pytest.mark.parametrize("order", [2, 3])
def test_burgers_3d_ddgic(order, form):
    """The DDGIC lifting term (sigma_ic != 0) of the 3D element kernel, both
    readings."""
    p = cases.burgers_params(3, mu=0.05, a=(0.3, 0.0, -0.2), sigma_ic=0.5)
    p.ddgic_form = form
    bct = ["DIRICHLET", "PERIODIC", "NEUMANN", "DIRICHLET", "PERIODIC", "EXTRAPOLATION"]
    assert _run(3, [2, 3, 3], order, p, bct, perturb=0.1, coupled=True, bcflag=[0, 0, 0, 1, 0, 0]) <= TOL


# ---- mixed precision (FP32 interpolation, fluxes and face data; FP64
# geometry, source load and output) against the FP64 oracle.  FP32 rounding
# of O(1) flux terms leaves ~1e-7 absolute error; the bound below is relative
# to the largest residual entry of a far-from-equilibrium state.
MIXED_TOL = 2e-5


def _run_mixed(nel, order, params, bct, seed=0, state=None):
    from iceicle_amd import engine
    from oracle.oracle import Oracle
    nd = len(nel)
    m = abi.mesh_desc(nd, nel, [0.0] * nd, [1.0] * nd, bct)
    coords = cases.perturbed_coords(m, 0.1, seed + 11)
    m = abi.mesh_desc(nd, nel, [0.0] * nd, [1.0] * nd, bct, coord=coords)
    o = Oracle(m, abi.space_desc(order), params)
    u = cases.random_state(params.flux, nd, o.nelem, o.nbasis, seed) if state is None else state
    r_ref = o.residual(u)
    space = engine.FESpace(m, order)
    disc = engine.ConservationLawDDG(space, params)
    ud = torch.from_numpy(u).cuda()
    rd = torch.full_like(ud, float("nan"))
    engine.form_residual(space, disc, ud, rd, precision="mixed")
    torch.cuda.synchronize()
    err = cases.rel_inf(rd.cpu().numpy(), r_ref)
    print(f"mixed ndim={nd} p={order} flux={params.flux} err={err:.3e}", flush=True)
    return err


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_mixed_ns_mms(order):
    from iceicle_amd import workloads
    assert _run_mixed([4, 3], order, workloads.ns_mms_params(), ["DIRICHLET"] * 4) <= MIXED_TOL


def test_mixed_euler_walls():
    p = cases.ns_params(2, flux=abi.FLUX_EULER)
    assert _run_mixed([3, 4], 3, p, ["SLIP_WALL", "RIEMANN", "DIRICHLET", "EXTRAPOLATION"]) <= MIXED_TOL


@pytest.mark.parametrize("order", [2, 3])
def test_mixed_burgers(order):
    p = cases.burgers_params(2, mu=0.02, a=(0.3, -0.2, 0))
    assert _run_mixed([4, 4], order, p, ["DIRICHLET", "NEUMANN", "EXTRAPOLATION", "DIRICHLET"]) <= MIXED_TOL


@pytest.mark.parametrize("order", [2, 3])
def test_mixed_euler_3d(order):
    p = cases.ns_params(3, flux=abi.FLUX_EULER)
    bct = ["SLIP_WALL", "RIEMANN", "DIRICHLET", "EXTRAPOLATION", "SLIP_WALL", "DIRICHLET"]
    assert _run_mixed([3, 2, 2], order, p, bct) <= MIXED_TOL


def test_mixed_ns_3d():
    p = cases.ns_params(3)
    bct = ["DIRICHLET", "SLIP_WALL", "RIEMANN", "EXTRAPOLATION", "NO_SLIP_ISOTHERMAL", "DIRICHLET"]
    assert _run_mixed([2, 2, 3], 2, p, bct) <= MIXED_TOL


def test_mixed_rejects_1d():
    from iceicle_amd import engine
    m = abi.mesh_desc(1, [4], [0], [1], ["EXTRAPOLATION"] * 2)
    disc = engine.ConservationLawDDG(engine.FESpace(m, 1), abi.default_params(abi.FLUX_EULER))
    u = torch.ones(disc.size, dtype=torch.float64, device="cuda")
    with pytest.raises(engine.IceError, match="ndim = 2 and 3"):
        engine.form_residual(disc.space, disc, u, torch.empty_like(u), precision="mixed")


def test_mixed_ns_mms_128_p3_against_fp64():
    """Full-size 128 x 128 P3 workload: mixed vs FP64 device residual (the
    FP64 one matches the oracle to 1e-12 in test_gpu_parity_full).  From the
    uniform free-stream state (the Newton initial guess) and at the
    manufactured state perturbed by 1e-3, where R is ~1e-3 of the flux terms
    it is summed from, the bound is the small-case one: the mixed kernels
    carry the coefficients as differences from the element's first node and
    form the face jumps in FP64 (measured 2e-7 and 7e-6 of max |R|)."""
    from iceicle_amd import engine, workloads
    m, order, p, keep = workloads.ns_mms_2d()
    space = engine.FESpace(m, order)
    disc = engine.ConservationLawDDG(space, p)
    fs = np.tile([1.0, 0.3, 0.0, 1.0 / 1.4 / 0.4 + 0.045], disc.size // 4)
    for name, state, tol in [("free stream", fs, MIXED_TOL),
                             ("perturbed MMS", workloads.mms_state_perturbed(space, 4, workloads.SEED0 + 2), MIXED_TOL)]:
        u = torch.from_numpy(state).cuda()
        r64, r32 = torch.empty_like(u), torch.empty_like(u)
        engine.form_residual(space, disc, u, r64)
        engine.form_residual(space, disc, u, r32, precision="mixed")
        torch.cuda.synchronize()
        err = float((r32 - r64).abs().max() / r64.abs().max())
        print(f"NS MMS 128x128 P3 {name}: mixed vs fp64 rel err {err:.3e}", flush=True)
        assert err <= tol
