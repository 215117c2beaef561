"""Batched independent instances (no reference counterpart: the reference
solves one mesh per process).

Workload: many independent 2D viscous Burgers problems, F(u) = (b/2) u^2 -
nu grad u with b = (1, 1), each on its own periodic (coupled) N x N mesh of
[0, 1]^2 with its own order p, viscosity nu and seeded sinusoidal initial
condition.  Instances of the same order are concatenated into one batched
space (ice_space_create_batch), so one residual launch sequence covers all of
them; each keeps its viscosity (per-element viscosity) and its own CFL step
(per-element step sizes).  A step of the whole batch is one RK3-SSP step of
each order group, and any number of such steps can be captured into one HIP
graph (StepGraph).  Each group's element-kernel workgroup shape is chosen by
timing (ConservationLawDDG.autotune).
"""
import math

import numpy as np

from . import abi, engine

N_RANGE = (16, 96)
ORDERS = (1, 2, 3, 4)
NU_RANGE = (1e-3, 1e-1)
MODES = ((1, 0), (0, 1), (1, 1), (2, 1))


def sample_instances(n, seed):
    """n instances: (N, p, nu, amplitudes[4]) with N ~ U{16..96}, p ~ U{1..4},
    nu log-uniform on [1e-3, 1e-1], amplitudes U(-1, 1); seeded."""
    rng = np.random.default_rng(seed)
    N = rng.integers(N_RANGE[0], N_RANGE[1] + 1, size=n)
    p = rng.integers(ORDERS[0], ORDERS[-1] + 1, size=n)
    nu = 10.0 ** rng.uniform(math.log10(NU_RANGE[0]), math.log10(NU_RANGE[1]), size=n)
    amp = rng.uniform(-1.0, 1.0, size=(n, len(MODES)))
    return [dict(N=int(N[i]), p=int(p[i]), nu=float(nu[i]), amp=amp[i]) for i in range(n)]


def initial_condition(x, y, amp):
    """u0 = sum_k A_k sin(2 pi (m_k x + n_k y))."""
    u = np.zeros_like(x)
    for a, (m, k) in zip(amp, MODES):
        u += a * np.sin(2.0 * np.pi * (m * x + k * y))
    return u


def params(nu):
    p = abi.default_params(abi.FLUX_BURGERS)
    p.mu = float(nu)
    for d in range(3):
        p.a[d] = 0.0
        p.b[d] = 1.0 if d < 2 else 0.0
    return p


def mesh(N):
    return abi.mesh_desc(2, [N, N], [0.0, 0.0], [1.0, 1.0], ["PERIODIC"] * 4, periodic_mode=abi.PERIODIC_COUPLED)


def cfl_dt(inst, u0, cfl=0.2):
    """Explicit step of one instance: cfl / ((2p + 1) lambda / h + (p + 1)^4
    nu / h^2), h = 1/N, lambda = |b| max|u0|.  The advective part is the
    reference's CFLTimestep scaling (explicit_utils.hpp:61-90); the diffusive
    part grows as (p + 1)^4 / h^2 as the DG diffusion operator's spectrum
    does.  The reference's Burgers dt_from_cfl (burgers.hpp:103-108, divided
    by 2p + 1) scales the viscous limit only as 1 / ((2p + 1) h^2) and is
    unstable for the viscous, high-order members of this family (nu = 0.1,
    p = 4)."""
    h = 1.0 / inst["N"]
    lam = math.sqrt(2.0) * float(np.max(np.abs(u0)))
    p = inst["p"]
    return cfl / ((2 * p + 1) * lam / h + (p + 1) ** 4 * inst["nu"] / h ** 2)


class Group:
    """All instances of one order in one batched space."""

    def __init__(self, insts, order, cfl=0.2):
        import torch
        self.insts = insts
        self.order = order
        self.meshes = [mesh(i["N"]) for i in insts]
        self.space = engine.FESpace(self.meshes, order)
        self.disc = engine.ConservationLawDDG(self.space, params(insts[0]["nu"]))
        nel = [i["N"] ** 2 for i in insts]
        self.offsets = np.concatenate([[0], np.cumsum(nel)])
        mu = np.repeat([i["nu"] for i in insts], nel)
        self.disc.set_element_viscosity(mu)
        xn = self.space.node_coords()  # [nelem][nb][2]
        u0 = np.empty(xn.shape[:2])
        dts = []
        for k, inst in enumerate(insts):
            a, b = self.offsets[k], self.offsets[k + 1]
            u0[a:b] = initial_condition(xn[a:b, :, 0], xn[a:b, :, 1], inst["amp"])
            dts.append(cfl_dt(inst, u0[a:b], cfl))
        self.dt = np.array(dts)
        self.u0 = u0.ravel()
        self.u = torch.from_numpy(self.u0.copy()).cuda()
        self.stepper = engine.ExplicitStepper(self.disc, abi.stepper_desc(abi.STEP_RK3_SSP, dt=float(min(dts)),
                                                                          ntime=1 << 40))
        self.stepper.set_element_dt(np.repeat(self.dt, nel))

    @property
    def size(self):
        return self.disc.size

    def instance_slice(self, k):
        nb = self.space.nbasis
        return slice(int(self.offsets[k]) * nb, int(self.offsets[k + 1]) * nb)


class Batch:
    """Instances grouped by order; one batched space, discretization and
    stepper per group."""

    def __init__(self, insts, cfl=0.2):
        self.insts = insts
        self.groups = []
        self.index = {}  # instance -> (group, position)
        for p in ORDERS:
            ids = [i for i, inst in enumerate(insts) if inst["p"] == p]
            if not ids:
                continue
            g = Group([insts[i] for i in ids], p, cfl)
            for pos, i in enumerate(ids):
                self.index[i] = (len(self.groups), pos)
            self.groups.append(g)
        self.size = sum(g.size for g in self.groups)

    def autotune(self, reps=10):
        import torch
        out = {}
        for g in self.groups:
            r = torch.empty_like(g.u)
            out[g.order] = g.disc.autotune(g.u, r, reps)
        return out

    def residual(self, outs, precision="fp64"):
        for g, r in zip(self.groups, outs):
            engine.form_residual(g.space, g.disc, g.u, r, precision=precision)

    def step(self):
        for g in self.groups:
            g.stepper.step(g.u)

    def graph(self, nsteps, concurrent=False):
        return engine.StepGraph([g.stepper for g in self.groups], [g.u for g in self.groups], nsteps, concurrent)

    def reset(self):
        import torch
        for g in self.groups:
            g.u.copy_(torch.from_numpy(g.u0))

    def instance_state(self, i):
        gi, pos = self.index[i]
        g = self.groups[gi]
        return g.u[g.instance_slice(pos)]

    # ---- FP32 states ----------------------------------------------------
    # Each group also holds an FP32 copy of its state (u32), stepped by the
    # same stepper with FP32 stage buffers; the FP64 residual of the widened
    # FP32 state is the check against the FP64 discretization.

    def reset_f32(self):
        import torch
        for g in self.groups:
            g.u32 = torch.from_numpy(g.u0).to(device=g.u.device, dtype=torch.float32)

    def step_f32(self):
        for g in self.groups:
            g.stepper.step(g.u32)

    def graph_f32(self, nsteps, concurrent=False):
        return engine.StepGraph([g.stepper for g in self.groups], [g.u32 for g in self.groups], nsteps, concurrent)

    def residual_f32(self, outs, check_fp64=False):
        for g, r in zip(self.groups, outs):
            engine.form_residual(g.space, g.disc, g.u32, r, precision="fp64" if check_fp64 else "mixed")

    def instance_state_f32(self, i):
        gi, pos = self.index[i]
        g = self.groups[gi]
        return g.u32[g.instance_slice(pos)]
