"""Benchmark: DG residual throughput on one MI355X.

Workload (the configuration the headline metric is quoted on): 2D compressible
Navier-Stokes manufactured steady problem, 128 x 128 quads with seeded 5%
interior node perturbation, P3 Lagrange, 4 x 4 Gauss-Legendre, 4 conserved
variables, VanLeer + DDG viscous fluxes, Dirichlet boundaries, manufactured
source.  One step = one FP64 residual evaluation R(u) (form_residual) over
the whole mesh with u and R resident in HBM.  After the timed steps, one
Newton-linesearch solve of the same problem from a uniform free-stream state
to ||R||_2 < 1e-10 is timed (the metric's second part, under "newton").

Prints one JSON line: value = residual GDOF/s of the timed pass (K
residuals, no per-kernel events); roofline = the slowest kernel's
algorithmic bytes over its duration from HIP events in a second pass of the
same K residuals, against 8 TB/s;
cpu_baseline = the CPU oracle on full residuals of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the FP64 residual kernels on the default workload
# (128 x 128, P3), from rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by the
# same counters on a 256 MiB copy (x2 and x1): profiles/residual_ns2d_p3_128_hbm_counters.md
# (element kernel: per workgroup tile 0 / 1 / 2 = 256 / 128 / 64 threads)
MEASURED_TRAFFIC_128_P3 = {"trace": (6893 * 2 + 24576) * 1024, "flux": (14738 * 2 + 4368) * 1024,
                           "element": {0: (11040 * 2 + 8192) * 1024, 1: (11103 * 2 + 8192) * 1024,
                                       2: int((11359.5 * 2 + 8192) * 1024)}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=128, help="elements per direction")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-newton", action="store_true", help="skip the Newton time-to-solution leg")
    ap.add_argument("--workload", default="ns2d", choices=["ns2d", "euler3d", "warmstart", "warmstart2d", "batch"],
                    help="ns2d: the headline 2D NS MMS workload (with the 3D Euler leg appended); "
                         "euler3d: only the 3D Euler vortex 64^3 P3 leg")
    ap.add_argument("--no-euler3d", action="store_true", help="skip the 3D Euler vortex leg")
    ap.add_argument("--no-warmstart", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances leg")
    ap.add_argument("--warmstart-train", type=int, default=4000, help="GPU-generated training solves (1D)")
    ap.add_argument("--warmstart2d-train", type=int, default=500, help="GPU-generated training solves (2D)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="only run 20 device copies of an MB-sized buffer (HBM counter calibration)")
    return ap.parse_args()


def traffic_of(kernel, tile):
    t = MEASURED_TRAFFIC_128_P3[kernel]
    return t[tile] if isinstance(t, dict) else t


def kernel_event_pass(engine, space, disc, u, r, precision, nres):
    """The per-kernel durations: nres more residuals of the same state with
    HIP events recorded on the launch stream before and between the three
    kernels.  Each event costs the stream a few microseconds, so the timed
    pass that gives the throughput carries none.  Returns (summed per-kernel
    ms [trace, flux, element], residuals profiled, ms per residual of this
    pass)."""
    import torch
    disc.profile_begin(nres)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(nres):
        engine.form_residual(space, disc, u, r, precision=precision)
    stop.record()
    torch.cuda.synchronize()
    stage_ms, ncalls = disc.profile_end()
    return stage_ms, ncalls, start.elapsed_time(stop) / nres


def cpu_baseline(mesh, order, params, u, seconds):
    """The oracle (serial FP64 restatement of the reference path) on the same
    workload, timed on this host: repeated full residuals until `seconds`."""
    from iceicle_amd import abi
    from oracle.oracle import Oracle
    o = Oracle(mesh, abi.space_desc(order), params)
    o.residual(u)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        o.residual(u)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 100:
            break
    return n, el, o.size


def euler3d(args, n=64, order=3, nres=30, nsteps=50):
    """3D Euler isentropic vortex on 64^3 hexes, P3 (83,886,080 unknowns):
    FP64 and mixed residual throughput, then RK3-SSP steps (textbook stages,
    fixed dt from CFL 0.2 at the initial state) timed over `nsteps` steps."""
    import torch
    from iceicle_amd import abi, engine, workloads
    print(f"euler3d {n}^3 P{order} ...", file=sys.stderr, flush=True)
    mesh, order, params, ic = workloads.euler_vortex_3d(n, order)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    u = torch.empty(disc.size, dtype=torch.float64, device="cuda")
    engine.project(disc, ic, u)
    u = torch.from_numpy(workloads.density_perturbed(u.cpu().numpy(), disc.neq, workloads.SEED0 + 3)).cuda()
    r = torch.empty_like(u)
    for _ in range(3):
        engine.form_residual(space, disc, u, r)
    torch.cuda.synchronize()
    tile, tile_ms = disc.autotune(u, r, reps=3)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(nres):
        engine.form_residual(space, disc, u, r)
    stop.record()
    torch.cuda.synchronize()
    res_ms = start.elapsed_time(stop) / nres
    stage_ms, ncalls, _ = kernel_event_pass(engine, space, disc, u, r, "fp64", nres)
    assert bool(torch.isfinite(r).all())
    # mixed-precision residual (FP32 interpolation, fluxes and face data)
    r32 = torch.empty_like(u)
    for _ in range(3):
        engine.form_residual(space, disc, u, r32, precision="mixed")
    torch.cuda.synchronize()
    start.record()
    for _ in range(nres):
        engine.form_residual(space, disc, u, r32, precision="mixed")
    stop.record()
    torch.cuda.synchronize()
    mixed_ms = start.elapsed_time(stop) / nres
    mixed_err = float((r32 - r).abs().max() / r.abs().max())
    # explicit steps at a fixed dt (CFL 0.2 over the element size and the
    # initial maximum wave speed |v| + c ~ 2.9 + 1.1)
    h = 40.0 / n
    dt = 0.2 * h / (2 * order + 1) / 4.0
    stepper = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_SSP, dt=dt, ntime=1 << 40))
    stepper.step(u)
    torch.cuda.synchronize()
    start.record()
    for _ in range(nsteps):
        stepper.step(u)
    stop.record()
    torch.cuda.synchronize()
    step_ms = start.elapsed_time(stop) / nsteps
    assert bool(torch.isfinite(u).all())
    ndof = disc.size
    del stepper, disc, space
    return {
        "metric": "residual GDOF/s, 3D Euler vortex (FP64, 1 GPU)", "value": round(ndof / (res_ms * 1e-3) / 1e9, 4),
        "unit": "GDOF/s", "n_gpus": 1, "ms_per_residual": round(res_ms, 4),
        "kernels_ms": {k: round(v / ncalls, 4) for k, v in zip(("trace", "flux", "element"), stage_ms)},
        "mixed": {"value": round(ndof / (mixed_ms * 1e-3) / 1e9, 4), "ms_per_residual": round(mixed_ms, 4),
                  "rel_err_vs_fp64": mixed_err},
        "rk3_ssp_step_ms": round(step_ms, 4), "rk3_steps_timed": nsteps, "dt": dt,
        "element_tile": {"best": tile, "ms": [round(t, 4) for t in tile_ms]},
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Euler isentropic vortex {n}^3 hexes, P{order}, {order + 1}^3 GL, "
                               f"coupled periodic faces, density perturbed U(-0.01, 0.01)",
                   "elements": n ** 3, "unknowns": ndof},
    }


def warmstart_leg(ntrain, ntest=100, epochs=300, dim=1):
    """Learned warm start on the steady viscous Burgers layer family (nu
    log-uniform in [1e-3, 1e-2], uL, uR uniform in [0.5, 2]).  1D: 256
    elements, P3, 1024 unknowns, Newton with the exact block-tridiagonal
    preconditioner; 2D: 64 x 64, P3, 65,536 unknowns, Newton with the
    agglomeration multigrid and Eisenstat-Walker forcing.  ntrain GPU Newton
    solves from the plain guess, Adam training of the 3 -> 256 -> 256 ->
    unknowns network on the device, then ntest unseen problems solved from
    the plain guess and from the network's guess (plain-guess fallback),
    both to ||R||_2 < 1e-10 with the same Newton (Wolfe linesearch)."""
    import torch
    from iceicle_amd import abi, warmstart as W, workloads
    print(f"warm start {dim}D: generating {ntrain} solves ...", file=sys.stderr, flush=True)
    params = W.sample_params(ntrain + ntest, seed=workloads.SEED0 + 4 + 10 * (dim - 1))
    train, test = params[:ntrain], params[ntrain:]
    Fam = W.BurgersLayer1D if dim == 1 else W.BurgersLayer2D
    # 1D generation runs without a linesearch (every member converges, in fewer steps)
    gen = Fam(linesearch=abi.LS_NONE) if dim == 1 else Fam()
    t0 = time.perf_counter()
    sols, reps = W.generate(gen, train)
    t_gen = time.perf_counter() - t0
    ok = np.all(np.isfinite(sols), axis=1)
    print(f"warm start: {int(ok.sum())}/{ntrain} converged in {t_gen:.1f} s; first report "
          f"{ {k: reps[0][k] for k in ('converged', 'iterations', 'rhist')} }", file=sys.stderr, flush=True)
    if not ok.any():
        raise RuntimeError("no training solve converged")
    fam = Fam() if dim == 1 else gen
    ws = W.WarmStart(fam, seed=1)
    hist, t_train = ws.train(train, sols, abi.mlp_train_opts(batch=64, epochs=epochs, lr=2e-3, lr_decay=0.99, seed=3))
    print(f"warm start: trained in {t_train:.1f} s, evaluating {ntest} ...", file=sys.stderr, flush=True)
    plain_its, plain_s, plain_ok, warm_its, warm_s, warm_ok, used = [], [], 0, [], [], 0, 0
    for prm in test:
        fam.set(*prm)
        u = torch.from_numpy(fam.plain_guess(prm[1], prm[2])).cuda()
        t1 = time.perf_counter()
        rep = fam.solve(u)
        plain_s.append(time.perf_counter() - t1)
        plain_its.append(rep["iterations"])
        plain_ok += int(rep["converged"])
        uw, rep, how = ws.solve(*prm)
        warm_s.append(rep["seconds_total"])
        warm_its.append(rep["iterations"])
        warm_ok += int(rep["converged"])
        used += int(how == "network")
    problem = ("steady 1D viscous Burgers layer, 256 elements, P3, 1024 unknowns, Dirichlet uL/uR" if dim == 1 else
               "steady 2D viscous Burgers layers, 64x64 quads, P3, 65536 unknowns, Dirichlet uL inflow / uR outflow")
    return {
        "problem": problem,
        "network": f"3 -> 256 -> 256 -> {fam.size}, tanh, FP32, device GEMMs on matrix cores, Adam",
        "train_solves": int(ntrain), "train_converged": int(ok.sum()), "generate_seconds": round(t_gen, 2),
        "train_seconds": round(t_train, 2), "epochs": epochs,
        "loss_first_last": [float(f"{hist[0]:.4e}"), float(f"{hist[-1]:.4e}")],
        "test_problems": int(ntest),
        "plain": {"converged": plain_ok, "mean_newton_steps": round(float(np.mean(plain_its)), 3),
                  "mean_seconds": round(float(np.mean(plain_s)), 5)},
        "warm": {"converged": warm_ok, "network_guess_used": used,
                 "mean_newton_steps": round(float(np.mean(warm_its)), 3),
                 "mean_seconds": round(float(np.mean(warm_s)), 5)},
        "time_ratio_warm_over_plain": round(float(np.mean(warm_s) / np.mean(plain_s)), 4),
    }


def batch_leg(ninst=512, reps=20, nsteps=10):
    """512 independent 2D viscous Burgers instances (N ~ U{16..96}, p ~
    U{1..4}, nu log-uniform on [1e-3, 1e-1], seeded sinusoidal ICs), one
    batched space per order with per-element viscosity and CFL step: FP64 and
    mixed residual throughput over the whole batch, RK3-SSP steps eager and
    as one captured HIP graph, and the same residuals evaluated instance by
    instance (one space per instance) for comparison."""
    import torch
    from iceicle_amd import batch as B, engine, workloads
    print(f"batched: {ninst} instances ...", file=sys.stderr, flush=True)
    insts = B.sample_instances(ninst, seed=workloads.SEED0 + 3)
    t0 = time.perf_counter()
    bt = B.Batch(insts)
    t_setup = time.perf_counter() - t0
    tiles = bt.autotune(reps=5)
    outs = [torch.empty_like(g.u) for g in bt.groups]
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, n):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(n):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / n

    ms64 = timed(lambda: bt.residual(outs), reps)
    r64 = [o.clone() for o in outs]
    ms32 = timed(lambda: bt.residual(outs, precision="mixed"), reps)
    err32 = max(float((o - r).abs().max() / r.abs().max()) for o, r in zip(outs, r64))
    # per-instance evaluation of the same residuals: one space per instance
    singles = []
    for inst in insts:
        sp = engine.FESpace(B.mesh(inst["N"]), inst["p"])
        dc = engine.ConservationLawDDG(sp, B.params(inst["nu"]))
        uu = torch.zeros(dc.size, dtype=torch.float64, device="cuda")
        singles.append((sp, dc, uu, torch.empty_like(uu)))
    for k, (sp, dc, uu, rr) in enumerate(singles):
        uu.copy_(bt.instance_state(k))
    ms_single = timed(lambda: [engine.form_residual(sp, dc, uu, rr) for sp, dc, uu, rr in singles], 3)
    same = all(torch.equal(rr, outs_k) for (sp, dc, uu, rr), outs_k in
               zip(singles, [r64[bt.index[k][0]][bt.groups[bt.index[k][0]].instance_slice(bt.index[k][1])]
                             for k in range(ninst)]))
    del singles
    # RK3-SSP steps of the whole batch: eager launches, then one graph
    ms_eager = timed(bt.step, nsteps)
    graph = bt.graph(nsteps)
    ms_graph = timed(graph.launch, 3) / nsteps
    finite = all(bool(torch.isfinite(g.u).all()) for g in bt.groups)
    return {
        "instances": ninst, "unknowns": int(bt.size), "groups": {g.order: len(g.insts) for g in bt.groups},
        "setup_seconds": round(t_setup, 2),
        "tiles": {str(k): {"best": v[0], "ms": [round(x, 4) for x in v[1]]} for k, v in tiles.items()},
        "residual_fp64": {"ms": round(ms64, 4), "GDOF_s": round(bt.size / (ms64 * 1e-3) / 1e9, 3)},
        "residual_mixed": {"ms": round(ms32, 4), "GDOF_s": round(bt.size / (ms32 * 1e-3) / 1e9, 3),
                           "rel_err_vs_fp64": float(f"{err32:.3e}")},
        "residual_per_instance": {"ms": round(ms_single, 3), "GDOF_s": round(bt.size / (ms_single * 1e-3) / 1e9, 3),
                                  "bit_identical_to_batched": bool(same)},
        "rk3_ssp_step_ms": {"eager": round(ms_eager, 4), "graph": round(ms_graph, 4), "graph_steps": nsteps},
        "state_finite": finite,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    from iceicle_amd import abi, engine, workloads

    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 8 * n, "calibrate_copy_bytes_written": 8 * n, "copies": 20}))
        return

    if args.workload == "euler3d":
        print(json.dumps(euler3d(args)))
        return
    if args.workload == "warmstart":
        print(json.dumps(warmstart_leg(args.warmstart_train)))
        return
    if args.workload == "warmstart2d":
        print(json.dumps(warmstart_leg(args.warmstart2d_train, ntest=50, dim=2)))
        return
    if args.workload == "batch":
        print(json.dumps(batch_leg()))
        return

    mesh, order, params, coords = workloads.ns_mms_2d(args.n, args.order)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    neq = disc.neq
    u_host = workloads.mms_state_perturbed(space, neq, workloads.SEED0 + 2)
    u = torch.from_numpy(u_host).cuda()
    r = torch.empty_like(u)

    for _ in range(args.warmup):
        engine.form_residual(space, disc, u, r)
    torch.cuda.synchronize()
    # element-kernel workgroup shape chosen by timing (FP64 and mixed separately)
    tile, tile_ms = disc.autotune(u, r, reps=50)

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        engine.form_residual(space, disc, u, r)
    stop.record()
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms_per_step = total_ms / args.steps
    stage_ms, ncalls, event_pass_ms = kernel_event_pass(engine, space, disc, u, r, "fp64", args.steps)
    ndof = space.nelem * space.nbasis * neq
    gdofs = ndof / (ms_per_step * 1e-3) / 1e9
    assert bool(torch.isfinite(r).all())

    # per-kernel compulsory (unique) bytes of one launch
    info = space.info
    D, nb, nfq = space.ndim, space.nbasis, space.nfqp
    E, nfi, nfb = space.nelem, info.nface_interior, info.nface_boundary
    F = nfi + nfb
    nvert = 2 ** D
    nv = 1 + D  # trace record: value + the element's share of the DDG gradient
    geo_rec = 8 * (D + 2)  # normal, w sqrt(g), h per face point
    trace_bytes = 8 * E * 2 * D * nfq * nv * neq
    kernels = {
        # element face traces: state, vertices, face links and face-point
        # geometry read, traces written
        "trace": (stage_ms[0] / ncalls,
                  8 * E * nb * neq + 8 * E * nvert * D + 4 * 2 * D * E + geo_rec * E * 2 * D * nfq + trace_bytes),
        # face fluxes: traces of both sides, face records and geometry,
        # boundary values read, weighted normal fluxes written
        "flux": (stage_ms[1] / ncalls,
                 trace_bytes + 16 * F + geo_rec * F * nfq + 8 * nfb * nfq * neq + 8 * F * nfq * neq),
        # elements: state, vertices, face links, face fluxes and source load
        # read, residual written
        "element": (stage_ms[2] / ncalls,
                    8 * E * nb * neq + 8 * E * nvert * D + 4 * 2 * D * E + 8 * F * nfq * neq
                    + 8 * E * nb * neq + 8 * E * nb * neq),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes = kernels[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9
    nnode = (args.n + 1) ** 2
    b_alg = workloads.residual_bytes(D, order, neq, E, nnode, F)
    res_gbs = b_alg / (ms_per_step * 1e-3) / 1e9

    # mixed-precision residual on the same state: throughput and its error
    # against the FP64 residual (which matches the CPU oracle to 1e-15)
    for _ in range(args.warmup):
        engine.form_residual(space, disc, u, r, precision="mixed")
    r64 = r.clone()
    engine.form_residual(space, disc, u, r64)
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        engine.form_residual(space, disc, u, r, precision="mixed")
    stop.record()
    torch.cuda.synchronize()
    mixed_ms = start.elapsed_time(stop) / args.steps
    mstage, mcalls, _ = kernel_event_pass(engine, space, disc, u, r, "mixed", args.steps)
    mixed_err = float((r - r64).abs().max() / r64.abs().max())
    # the same comparison from the uniform free-stream state, where R is far from zero
    uf = torch.from_numpy(np.tile([1.0, 0.3, 0.0, 1.0 / 1.4 / 0.4 + 0.045], disc.size // 4)).cuda()
    rf64, rf32 = torch.empty_like(uf), torch.empty_like(uf)
    engine.form_residual(space, disc, uf, rf64)
    engine.form_residual(space, disc, uf, rf32, precision="mixed")
    torch.cuda.synchronize()
    mixed_err_fs = float((rf32 - rf64).abs().max() / rf64.abs().max())
    mixed = {
        "value": round(ndof / (mixed_ms * 1e-3) / 1e9, 4),
        "unit": "GDOF/s",
        "ms_per_step": round(mixed_ms, 6),
        "rel_err_vs_fp64": float(f"{mixed_err:.3e}"),
        "rel_err_vs_fp64_free_stream": float(f"{mixed_err_fs:.3e}"),
        "kernels_ms": {k: round(v / mcalls, 6) for k, v in zip(("trace", "flux", "element"), mstage)},
        "arithmetic": "FP32 interpolation, fluxes and face data; FP64 geometry, source load and output",
    }

    newton = None
    if not args.no_newton:
        # Newton-linesearch from a uniform free-stream state (rho 1, u 0.3,
        # v 0, p 1/gamma) to ||R||_2 < 1e-10, FP32 Jacobian blocks and
        # preconditioner with FP64 residual and refinement
        fs = np.tile([1.0, 0.3, 0.0, 1.0 / 1.4 / 0.4 + 0.045], disc.size // 4)
        un = torch.from_numpy(fs).cuda()
        solver = engine.PetscNewton(disc, tau_abs=1e-10, tau_rel=0.0, kmax=30, forcing=abi.FORCING_EW,
                                    preconditioner=abi.PC_MULTIGRID, smoother_precision=abi.PREC_FP16)
        print("newton solve ...", file=sys.stderr, flush=True)
        rep = solver.solve(un)
        print(f"newton done: {rep['iterations']} steps, {rep['seconds']:.2f} s", file=sys.stderr, flush=True)
        newton = {
            "seconds": round(rep["seconds"], 4),
            "converged": rep["converged"],
            "steps": rep["iterations"],
            "residual_norms": [float(f"{v:.3e}") for v in rep["rhist"]],
            "r0": float(f"{rep['r0']:.4e}"),
            "linear_iterations": rep["linear_iterations"],
            "jacobian_seconds": round(rep["jacobian_seconds"], 4),
            "linear_seconds": round(rep["linear_seconds"], 4),
            "initial_guess": "uniform free stream",
            "linear_solver": "GMRES(30) (classical Gram-Schmidt, one pass) with an agglomeration multigrid "
                             "V-cycle (P0 Galerkin coarse levels, block SGS smoothing on FP16 fine blocks), "
                             "FP32 operator + FP64 refinement, Eisenstat-Walker forcing",
        }
        del solver

    out = {
        "metric": "residual GDOF/s (FP64 vs mixed, 1 GPU); Newton time to ||R||<1e-10",
        "value": round(gdofs, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "ms_per_step_with_kernel_events": round(event_pass_ms, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"2D NS MMS {args.n}x{args.n} quads, 5% perturbed, P{order}, "
                        f"{order + 1}x{order + 1} GL, FP64 residual",
            "elements": E, "unknowns": ndof, "neq": neq, "order": order,
        },
        "kernels_ms": {k: round(v[0], 6) for k, v in kernels.items()},
        "kernel_bytes": {k: int(v[1]) for k, v in kernels.items()},
        "element_tile": {"best": tile, "ms": [round(t, 5) for t in tile_ms],
                         "threads": [256, 128, 64]},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": (traffic_of(dom, tile) if (args.n, order) == (128, 3) else None),
            "bytes_per_launch": int(kbytes),
        },
        "residual_roofline": {"bytes": int(b_alg), "achieved_GBs": round(res_gbs, 2),
                              "frac": round(res_gbs / HBM_PEAK_GBS, 5)},
    }
    out["mixed"] = mixed
    if newton is not None:
        out["newton"] = newton
    if not args.no_euler3d:
        del disc, space
        torch.cuda.empty_cache()
        out["euler3d"] = euler3d(args)
    if not args.no_warmstart:
        out["warmstart"] = warmstart_leg(args.warmstart_train)
        out["warmstart_2d"] = warmstart_leg(args.warmstart2d_train, ntest=50, dim=2)
    if not args.no_batch:
        out["batched"] = batch_leg()
    if not args.no_cpu_baseline:
        n, el, size = cpu_baseline(mesh, order, params, u_host, args.cpu_seconds)
        out["cpu_baseline"] = {
            "value": round(size * n / el / 1e9, 6),
            "unit": "GDOF/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{n} full residuals of the same workload on the host CPU ({el:.1f} s)",
            "gpu_speedup": round(gdofs / (size * n / el / 1e9), 1),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
