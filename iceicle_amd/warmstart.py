"""Learned warm start for steady Newton solves (no reference counterpart).

Problem family: steady viscous Burgers, F(u) = u^2/2 - nu u_x, on [0, 1]
with Dirichlet states u(0) = uL, u(1) = uR (uL, uR > 0): the flow runs to
the right and the solution is a viscous layer against the outflow boundary,
whose thickness ~ nu / u and jump uR - uL vary with the parameters.  The
discretization is the library's ConservationLawDDG (Burgers flux with
a = 0, b = 1) on a uniform mesh, solved by the device Newton
(PetscNewton) to ||R||_2 < tau.

Pipeline:
  1. generate: solve sampled problems on the GPU from the plain initial
     guess (the linear profile between uL and uR at the Lagrange nodes) and
     keep the converged coefficient vectors;
  2. train: an MLP (library kernels, FP32) maps scaled parameters
     (log10 nu, uL, uR) to the coefficients' departure from the plain guess;
  3. solve: predict and run the same Newton solve from the prediction; if
     the prediction is not finite or Newton from it does not reach the
     tolerance, the plain guess is solved instead (the fallback).  Newton's
     FP64 residual is the judge: the prediction's own residual norm is
     typically larger than the plain guess's (its errors sit in the layer,
     where the residual weights them most), yet it lies in Newton's basin.
"""
import math
import time

import numpy as np

from . import abi, engine

NU_RANGE = (1e-3, 1e-2)
U_RANGE = (0.5, 2.0)


def sample_params(n, seed):
    """n parameter triples (nu log-uniform, uL and uR uniform), seeded."""
    rng = np.random.default_rng(seed)
    nu = 10.0 ** rng.uniform(math.log10(NU_RANGE[0]), math.log10(NU_RANGE[1]), size=n)
    ul = rng.uniform(*U_RANGE, size=n)
    ur = rng.uniform(*U_RANGE, size=n)
    return np.stack([nu, ul, ur], axis=1)


def features(params):
    """Parameters scaled to [-1, 1]^3 (FP32)."""
    p = np.atleast_2d(np.asarray(params, dtype=np.float64))
    lo, hi = math.log10(NU_RANGE[0]), math.log10(NU_RANGE[1])
    f = np.empty_like(p)
    f[:, 0] = (np.log10(p[:, 0]) - lo) / (hi - lo) * 2.0 - 1.0
    for k in (1, 2):
        f[:, k] = (p[:, k] - U_RANGE[0]) / (U_RANGE[1] - U_RANGE[0]) * 2.0 - 1.0
    return f.astype(np.float32)


class BurgersLayer1D:
    """One space, discretization and Newton solver reused for every member of
    the family (ConservationLawDDG.update swaps the parameters in place)."""

    def __init__(self, nelem=256, order=3, tau_abs=1e-10, kmax=60, linesearch=abi.LS_WOLFE):
        self.mesh = abi.mesh_desc(1, [nelem], [0.0], [1.0], ["DIRICHLET", "DIRICHLET"], bcflag=[0, 1])
        self.space = engine.FESpace(self.mesh, order)
        self.xn = self.space.node_coords().reshape(-1)
        self.disc = engine.ConservationLawDDG(self.space, self.params(NU_RANGE[0], 1.0, 1.0))
        self.size = self.disc.size
        # the exact block-tridiagonal solve of the 1D chain as preconditioner
        self.solver = engine.PetscNewton(self.disc, tau_abs=tau_abs, tau_rel=0.0, kmax=kmax, linesearch=linesearch,
                                         linear_rtol=1e-10, jacobian_precision=abi.PREC_FP64,
                                         preconditioner=abi.PC_LINE)

    @staticmethod
    def params(nu, ul, ur):
        p = abi.default_params(abi.FLUX_BURGERS)
        p.mu = float(nu)
        for d in range(3):
            p.a[d] = 0.0
            p.b[d] = 1.0 if d == 0 else 0.0
        p.dirichlet[0] = abi.Field.make("CONSTANT", [float(ul)])
        p.dirichlet[1] = abi.Field.make("CONSTANT", [float(ur)])
        return p

    def set(self, nu, ul, ur):
        self.disc.update(self.params(nu, ul, ur))

    def plain_guess(self, ul, ur):
        return ul + (ur - ul) * self.xn

    def residual_norm(self, u):
        import torch
        r = torch.empty_like(u)
        engine.form_residual(self.space, self.disc, u, r)
        return float(r.norm())

    def solve(self, u):
        """Newton in place from u (device, FP64); the report dict."""
        return self.solver.solve(u)


class BurgersLayer2D(BurgersLayer1D):
    """The 2D member: F(u) = (u^2/2, u^2/2) - nu grad u on [0, 1]^2 with
    u = uL on the inflow sides (x = 0, y = 0) and u = uR on the outflow sides
    (x = 1, y = 1); the layers form against the outflow sides and the
    corner.  Newton with GMRES and the agglomeration multigrid (FP32 blocks,
    FP64 refinement), Eisenstat-Walker forcing."""

    def __init__(self, nelem=64, order=3, tau_abs=1e-10, kmax=60, linesearch=abi.LS_WOLFE):
        self.mesh = abi.mesh_desc(2, [nelem, nelem], [0.0, 0.0], [1.0, 1.0], ["DIRICHLET"] * 4,
                                  bcflag=[0, 0, 1, 1])
        self.space = engine.FESpace(self.mesh, order)
        xn = self.space.node_coords()
        self.xn = 0.5 * (xn[..., 0] + xn[..., 1]).reshape(-1)  # the plain guess ramps along x + y
        self.disc = engine.ConservationLawDDG(self.space, self.params(NU_RANGE[0], 1.0, 1.0))
        self.size = self.disc.size
        self.solver = engine.PetscNewton(self.disc, tau_abs=tau_abs, tau_rel=0.0, kmax=kmax, linesearch=linesearch,
                                         preconditioner=abi.PC_MULTIGRID, forcing=abi.FORCING_EW)

    @staticmethod
    def params(nu, ul, ur):
        p = BurgersLayer1D.params(nu, ul, ur)
        p.b[1] = 1.0
        return p


def generate(fam, params, verbose=False):
    """Solve each problem from the plain guess.  Returns (solutions [n][size]
    FP64 with NaN rows for failures, reports)."""
    import torch
    sols = np.full((len(params), fam.size), np.nan)
    reps = []
    for i, (nu, ul, ur) in enumerate(params):
        fam.set(nu, ul, ur)
        u = torch.from_numpy(fam.plain_guess(ul, ur)).cuda()
        rep = fam.solve(u)
        reps.append(rep)
        if rep["converged"]:
            sols[i] = u.cpu().numpy()
        if verbose and (i + 1) % 100 == 0:
            print(f"generated {i + 1}/{len(params)}", flush=True)
    return sols, reps


class WarmStart:
    """Trained predictor of the departure from the plain guess."""

    def __init__(self, fam, hidden=(256, 256), seed=0):
        self.fam = fam
        self.mlp = engine.MLP([3, *hidden, fam.size], seed=seed)
        self.scale = 1.0

    def targets(self, params, sols):
        d = np.stack([s - self.fam.plain_guess(ul, ur) for (nu, ul, ur), s in zip(params, sols)])
        return d

    def train(self, params, sols, opts):
        import torch
        ok = np.all(np.isfinite(sols), axis=1)
        params, sols = params[ok], sols[ok]
        d = self.targets(params, sols)
        self.scale = float(np.sqrt(np.mean(d ** 2))) or 1.0
        x = torch.from_numpy(features(params)).cuda()
        t = torch.from_numpy((d / self.scale).astype(np.float32)).cuda()
        t0 = time.perf_counter()
        hist = self.mlp.train(x, t, opts)
        torch.cuda.synchronize()
        return hist, time.perf_counter() - t0

    def predict(self, nu, ul, ur):
        """Predicted coefficients (device, FP64)."""
        import torch
        x = torch.from_numpy(features([[nu, ul, ur]])).cuda()
        y = self.mlp.forward(x)
        plain = torch.from_numpy(self.fam.plain_guess(ul, ur)).cuda()
        return plain + self.scale * y[0].double()

    def solve(self, nu, ul, ur):
        """Warm-started Newton with the plain guess as fallback.  Returns
        (u, report, used) where used is "network" or "plain"; the report's
        seconds_total covers the prediction, every Newton solve and the
        fallback."""
        import torch
        fam = self.fam
        fam.set(nu, ul, ur)
        t0 = time.perf_counter()
        u = self.predict(nu, ul, ur)
        rep = None
        if bool(torch.isfinite(u).all()):
            rep = fam.solve(u)
        used = "network"
        if rep is None or not rep["converged"]:
            used = "plain"
            u = torch.from_numpy(fam.plain_guess(ul, ur)).cuda()
            rep = fam.solve(u)
        torch.cuda.synchronize()
        rep["seconds_total"] = time.perf_counter() - t0
        return u, rep, used
