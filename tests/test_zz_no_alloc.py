"""ice_residual_noalloc (the allocation-free FP64 residual) against
ice_residual through the C ABI: the same kernels and summation order, so the
results must be bit-identical.  Two calls in a row on different
discretizations, the second smaller, then the first again (the shared
scratch array keeps the previous call's contents), plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from iceicle_amd import abi, workloads
from tests import cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _problem(kind):
    from iceicle_amd import engine
    if kind == "ns_mms_48_p3":
        mesh, order, params, _ = workloads.ns_mms_2d(48, 3)
        space = engine.FESpace(mesh, order)
        disc = engine.ConservationLawDDG(space, params)
        u = workloads.mms_state_perturbed(space, disc.neq, 7)
    elif kind == "burgers_ddgic_6x5_p2":
        p = cases.burgers_params(2, sigma_ic=0.5, a=(0.3, -0.2, 0))
        m = abi.mesh_desc(2, [6, 5], [0.0, 0.0], [1.0, 1.0], ["DIRICHLET", "NEUMANN", "EXTRAPOLATION", "DIRICHLET"],
                          bcflag=[0, 1, 0, 1])
        space = engine.FESpace(m, 2)
        disc = engine.ConservationLawDDG(space, p)
        u = cases.random_state(p.flux, 2, space.nelem, space.nbasis, 3)
    else:  # 3D Euler, coupled periodic faces
        mesh, order, params, _ = workloads.euler_vortex_3d(5, 2)
        space = engine.FESpace(mesh, order)
        disc = engine.ConservationLawDDG(space, params)
        u = cases.random_state(params.flux, 3, space.nelem, space.nbasis, 5)
    return engine, space, disc, torch.from_numpy(np.ascontiguousarray(u)).cuda()


def _compare(kind):
    engine, space, disc, u = _problem(kind)
    L = engine.lib()
    need, mx = C.c_int64(), C.c_int64()
    engine.check(L.ice_noalloc_scratch_bytes(disc.h, C.byref(need), C.byref(mx)))
    assert 0 < need.value <= mx.value == 64 << 20, (kind, need.value, mx.value)
    ref = torch.full_like(u, float("nan"))
    out = torch.full_like(u, float("nan"))
    engine.form_residual(space, disc, u, ref)
    st = engine.current_stream()
    print(f"noalloc: {kind}, {u.numel()} unknowns, scratch {need.value} bytes", flush=True)
    engine.check(L.ice_residual_noalloc(disc.h, C.c_void_p(u.data_ptr()), C.c_void_p(out.data_ptr()), st))
    torch.cuda.synchronize()
    assert bool(torch.isfinite(ref).all())
    assert torch.equal(out, ref), (kind, float((out - ref).abs().max()))
    return need.value


def test_noalloc_matches_residual_two_calls_second_smaller():
    big = _compare("ns_mms_48_p3")
    small = _compare("burgers_ddgic_6x5_p2")
    assert small < big
    _compare("euler_3d_5_p2")
    _compare("ns_mms_48_p3")


def test_noalloc_rejects_oversized_discretization():
    from iceicle_amd import engine
    mesh, order, params, _ = workloads.ns_mms_2d(200, 3)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    L = engine.lib()
    need, mx = C.c_int64(), C.c_int64()
    engine.check(L.ice_noalloc_scratch_bytes(disc.h, C.byref(need), C.byref(mx)))
    assert need.value > mx.value
    u = torch.zeros(disc.size, dtype=torch.float64, device="cuda")
    r = torch.zeros_like(u)
    rc = L.ice_residual_noalloc(disc.h, C.c_void_p(u.data_ptr()), C.c_void_p(r.data_ptr()), engine.current_stream())
    assert rc == abi.ICE_ERR_ARG
    assert b"scratch" in L.ice_last_error()
