"""In-place parameter updates, the block-tridiagonal (line) preconditioner
and the learned warm start of the steady 1D viscous Burgers family."""
import numpy as np
import pytest

from iceicle_amd import abi, workloads
from tests import cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_update_equals_fresh_discretization():
    """A discretization whose parameters were replaced in place gives the
    residual of one created with those parameters, bit for bit."""
    from iceicle_amd import engine, warmstart
    fam = warmstart.BurgersLayer1D(nelem=32)
    u = torch.from_numpy(np.random.default_rng(0).uniform(0.5, 2.0, fam.size)).cuda()
    for prm in [(2e-3, 1.7, 0.6), (8e-3, 0.9, 1.9)]:
        fam.set(*prm)
        fresh = engine.ConservationLawDDG(fam.space, fam.params(*prm))
        r1, r2 = torch.empty_like(u), torch.empty_like(u)
        engine.form_residual(fam.space, fam.disc, u, r1)
        engine.form_residual(fam.space, fresh, u, r2)
        assert torch.equal(r1, r2)
    # Navier-Stokes: viscosity and Dirichlet data replaced
    m, order, p, keep = workloads.ns_mms_2d(8, 2)
    space = engine.FESpace(m, order)
    disc = engine.ConservationLawDDG(space, p)
    p2 = workloads.ns_mms_params(mu=0.01)
    disc.update(p2)
    fresh = engine.ConservationLawDDG(space, p2)
    u = torch.from_numpy(workloads.mms_state_perturbed(space, 4, 3, amp=1e-2)).cuda()
    r1, r2 = torch.empty_like(u), torch.empty_like(u)
    engine.form_residual(space, disc, u, r1)
    engine.form_residual(space, fresh, u, r2)
    assert torch.equal(r1, r2)
    p3 = workloads.ns_mms_params()
    p3.flux = abi.FLUX_EULER
    with pytest.raises(engine.IceError, match="flux family"):
        disc.update(p3)


def test_line_preconditioner_solves_exactly():
    """On a 1D chain the block-tridiagonal factorisation is the exact
    inverse: GMRES needs one iteration."""
    from iceicle_amd import engine
    p = cases.burgers_params(1, mu=0.004, a=(0.0, 0, 0), b=(1.0, 0, 0))
    m = abi.mesh_desc(1, [40], [0.0], [1.0], ["DIRICHLET", "DIRICHLET"], bcflag=[0, 1])
    space = engine.FESpace(m, 3)
    disc = engine.ConservationLawDDG(space, p)
    u = torch.from_numpy(np.random.default_rng(1).uniform(0.5, 1.5, disc.size)).cuda()
    jac = engine.BlockJacobian(disc)
    jac.assemble(u)
    x_true = torch.from_numpy(np.random.default_rng(2).normal(size=disc.size)).cuda()
    b = torch.empty_like(u)
    jac.apply(x_true, b)
    solver = engine.PetscNewton(disc, jacobian_precision=abi.PREC_FP64, preconditioner=abi.PC_LINE,
                                linear_rtol=1e-12)
    # the output buffer's old contents must not leak into the solve (NaN included)
    x = torch.full_like(u, float("nan"))
    its, rel = solver.linear_solve(b, x, u=u)
    err = float((x - x_true).norm() / x_true.norm())
    print(f"line preconditioner: {its} iterations, rel {rel:.2e}, x err {err:.2e}", flush=True)
    assert its <= 2 and rel <= 1e-12 and err <= 1e-9


def test_line_preconditioner_ns_1d():
    """Navier-Stokes 1D, P3: 12 x 12 blocks."""
    from iceicle_amd import engine
    p = cases.ns_params(1)
    m = abi.mesh_desc(1, [30], [0.0], [1.0], ["DIRICHLET", "EXTRAPOLATION"])
    disc = engine.ConservationLawDDG(engine.FESpace(m, 3), p)
    u = torch.from_numpy(cases.random_state(p.flux, 1, 30, 4, 3)).cuda()
    jac = engine.BlockJacobian(disc)
    jac.assemble(u)
    x_true = torch.from_numpy(np.random.default_rng(4).normal(size=disc.size)).cuda()
    b = torch.empty_like(u)
    jac.apply(x_true, b)
    solver = engine.PetscNewton(disc, jacobian_precision=abi.PREC_FP64, preconditioner=abi.PC_LINE,
                                linear_rtol=1e-12)
    x = torch.full_like(u, float("nan"))
    its, rel = solver.linear_solve(b, x, u=u)
    err = float((x - x_true).norm() / x_true.norm())
    print(f"line preconditioner NS 1D: {its} iterations, rel {rel:.2e}, x err {err:.2e}", flush=True)
    assert its <= 2 and rel <= 1e-12 and err <= 1e-8


def test_line_preconditioner_rejects_non_chains():
    from iceicle_amd import engine
    p = cases.burgers_params(1, mu=0.01, a=(1.0, 0, 0))
    m = abi.mesh_desc(1, [8], [0.0], [1.0], ["PERIODIC", "PERIODIC"], periodic_mode=abi.PERIODIC_COUPLED)
    disc = engine.ConservationLawDDG(engine.FESpace(m, 2), p)
    with pytest.raises(engine.IceError, match="non-periodic 1D"):
        engine.PetscNewton(disc, jacobian_precision=abi.PREC_FP64, preconditioner=abi.PC_LINE)
    m2 = abi.mesh_desc(2, [3, 3], [0, 0], [1, 1], ["DIRICHLET"] * 4)
    disc2 = engine.ConservationLawDDG(engine.FESpace(m2, 1), cases.burgers_params(2))
    with pytest.raises(engine.IceError, match="non-periodic 1D"):
        engine.PetscNewton(disc2, jacobian_precision=abi.PREC_FP64, preconditioner=abi.PC_LINE)


def test_family_newton_matches_dense_oracle():
    """Newton with the line preconditioner reaches the oracle's converged
    state (dense LU Newton) on a member of the family."""
    from iceicle_amd import warmstart
    from oracle.oracle import Oracle
    fam = warmstart.BurgersLayer1D(nelem=24)
    prm = (5e-3, 1.6, 0.8)
    fam.set(*prm)
    o = Oracle(fam.mesh, abi.space_desc(3), fam.params(*prm))
    u0 = fam.plain_guess(prm[1], prm[2])
    u_ref, k_ref, _ = o.newton(u0, tau_abs=1e-12, kmax=40)
    u = torch.from_numpy(u0).cuda()
    rep = fam.solve(u)
    assert rep["converged"]
    err = cases.rel_inf(u.cpu().numpy(), u_ref)
    print(f"line-preconditioned Newton {rep['iterations']} steps, oracle {k_ref}; state rel diff {err:.2e}", flush=True)
    assert err <= 1e-9


def test_warm_start_cuts_newton_iterations():
    """Train on 192 GPU-generated solves, then solve 12 unseen problems from
    the network's guess: every solve converges to ||R||_2 < 1e-10 and the
    mean Newton step count falls below the plain initial guess's."""
    from iceicle_amd import engine, warmstart
    fam = warmstart.BurgersLayer1D()
    params = warmstart.sample_params(204, seed=workloads.SEED0 + 5)
    train, test = params[:192], params[192:]
    sols, reps = warmstart.generate(fam, train)
    assert np.all(np.isfinite(sols))
    ws = warmstart.WarmStart(fam, seed=2)
    hist, secs = ws.train(train, sols, abi.mlp_train_opts(batch=32, epochs=200, lr=2e-3, lr_decay=0.985, seed=3))
    print(f"training {secs:.2f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.05 * hist[0]
    plain_its, warm_its, used = [], [], []
    for prm in test:
        fam.set(*prm)
        u = torch.from_numpy(fam.plain_guess(prm[1], prm[2])).cuda()
        plain_its.append(fam.solve(u)["iterations"])
        uw, rep, how = ws.solve(*prm)
        assert rep["converged"] and rep["rnorm"] < 1e-10
        r = torch.empty_like(uw)
        engine.form_residual(fam.space, fam.disc, uw, r)
        assert float(r.norm()) < 1e-10
        warm_its.append(rep["iterations"])
        used.append(how)
    print(f"Newton steps plain {plain_its} (mean {np.mean(plain_its):.2f}), warm {warm_its} "
          f"(mean {np.mean(warm_its):.2f}); guess used {used}", flush=True)
    assert np.mean(warm_its) < np.mean(plain_its)


def test_warm_start_2d_cuts_newton_iterations():
    """The 2D family (16 x 16, P3, multigrid Newton): train on 96 solves,
    solve 6 unseen problems from the network's guess."""
    from iceicle_amd import warmstart
    fam = warmstart.BurgersLayer2D(nelem=16)
    params = warmstart.sample_params(102, seed=workloads.SEED0 + 7)
    train, test = params[:96], params[96:]
    sols, reps = warmstart.generate(fam, train)
    assert np.all(np.isfinite(sols))
    ws = warmstart.WarmStart(fam, seed=4)
    hist, secs = ws.train(train, sols, abi.mlp_train_opts(batch=16, epochs=200, lr=2e-3, lr_decay=0.985, seed=5))
    plain_its, warm_its = [], []
    for prm in test:
        fam.set(*prm)
        u = torch.from_numpy(fam.plain_guess(prm[1], prm[2])).cuda()
        plain_its.append(fam.solve(u)["iterations"])
        uw, rep, how = ws.solve(*prm)
        assert rep["converged"] and rep["rnorm"] < 1e-10
        warm_its.append(rep["iterations"])
    print(f"2D: loss {hist[0]:.3e} -> {hist[-1]:.3e}; Newton steps plain {plain_its}, warm {warm_its}", flush=True)
    assert np.mean(warm_its) < np.mean(plain_its)
