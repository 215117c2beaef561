"""Benchmark: DG residual throughput on one MI355X.

Workload (the configuration the headline metric is quoted on): 2D compressible
Navier-Stokes manufactured steady problem, 128 x 128 quads with seeded 5%
interior node perturbation, P3 Lagrange, 4 x 4 Gauss-Legendre, 4 conserved
variables, VanLeer + DDG viscous fluxes, Dirichlet boundaries, manufactured
source.  One step = one FP64 residual evaluation R(u) (form_residual) over
the whole mesh with u and R resident in HBM.  After the timed steps, the
Newton-linesearch solve of the same problem from a uniform free-stream state
to ||R||_2 < 1e-10 is timed (the metric's second part, under "newton").

Prints one JSON line:
  value     residual GDOF/s of the timed pass (K residuals back to back, no
            per-kernel events);
  roofline  the residual's algorithmic bytes B_alg = 2 s m E nb + 8 d N_vert
            + 4 2^d E + 12 F (state in, residual out, vertices, P1
            connectivity, face records) over the timed residual, against
            8 TB/s; traffic = the HBM bytes of the residual's kernels from
            rocprofv3 FETCH_SIZE / WRITE_SIZE counters taken in this run
            (child processes), scaled by the same counters on a device copy;
  kernels   per-kernel durations (HIP events in a second pass), the bytes
            each kernel moves by design and their measured traffic;
  cpu_baseline  the CPU oracle (restated reference loop) on all host cores,
            with the single-core figure beside it.
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
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure the residual kernels' HBM traffic")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2D Burgers 32x32 P2 RK3-SSP leg")
    ap.add_argument("--quick-legs", action="store_true",
                    help="few repetitions in the 3D leg (for profiler runs)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tile", type=int, default=0, help=argparse.SUPPRESS)
    return ap.parse_args()


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


KERNEL_KEYS = (("trace", "trace_kernel"), ("flux", "flux_kernel"), ("element", "element_kernel"))


def traffic_probe(args):
    """Child of measure_traffic, run under rocprofv3 --pmc: a few FP64
    residuals of the headline workload with the element tile the parent
    chose, then device copies of a known size for the counter scale."""
    import torch
    from iceicle_amd import engine, workloads
    mesh, order, params, _ = workloads.ns_mms_2d(args.n, args.order)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    disc.set_tile(args.tile)
    u = torch.from_numpy(workloads.mms_state_perturbed(space, disc.neq, workloads.SEED0 + 2)).cuda()
    r = torch.empty_like(u)
    for _ in range(5):
        engine.form_residual(space, disc, u, r)
    torch.cuda.synchronize()
    a = torch.ones(COPY_MB * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("traffic probe done", file=sys.stderr, flush=True)


COPY_MB = 256


def _counter_means(csv_dir, counter):
    """Mean counter value per dispatch, by kernel-name prefix, from a
    rocprofv3 counter_collection.csv.  For the calibration copy only the
    largest copy dispatches count (the run's small host-to-device staging
    copies use the same blit kernel)."""
    import csv
    import glob
    files = glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"no counter_collection.csv under {csv_dir}")
    rows = {}
    with open(files[0]) as fh:
        for row in csv.DictReader(fh):
            if row.get("Counter_Name") != counter:
                continue
            name = row["Kernel_Name"]
            key = None
            for k, prefix in KERNEL_KEYS:
                if prefix in name:
                    key = k
            if key is None and "copy" in name.lower():
                key = "copy"
            if key is not None:
                rows.setdefault(key, []).append((int(row.get("Grid_Size", 0) or 0), float(row["Counter_Value"])))
    if "copy" in rows:
        big = max(g for g, _ in rows["copy"])
        rows["copy"] = [(g, v) for g, v in rows["copy"] if g == big]
    return {k: sum(v for _, v in r) / len(r) for k, r in rows.items()}


def measure_traffic(args, tile, timeout=150):
    """HBM bytes per launch of the residual kernels, measured in this run:
    two child processes of this script under rocprofv3, one counting
    FETCH_SIZE and one WRITE_SIZE (KiB per dispatch), each scaled by what the
    same counter reads for a device copy of COPY_MB MiB (bytes read and
    written known).  Returns (per-kernel bytes, details) or (None, reason)."""
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    out = {}
    tmp = tempfile.mkdtemp(prefix="ice_traffic_")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = os.path.join(tmp, counter)
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "c", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", "--tile", str(tile),
               "--n", str(args.n), "--order", str(args.order)]
        print(f"traffic: {counter} child run ...", file=sys.stderr, flush=True)
        proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, start_new_session=True)
        try:
            _, err = proc.communicate(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.communicate()
            return None, f"rocprofv3 {counter} run timed out"
        try:
            out[counter] = _counter_means(d, counter)
        except Exception as e:  # noqa: BLE001
            tail = err.decode(errors="replace")[-300:] if err else ""
            return None, f"rocprofv3 {counter}: {e}; rc={proc.returncode} {tail}"
    copy_bytes = COPY_MB * (1 << 20)
    fc, wc = out["FETCH_SIZE"].get("copy"), out["WRITE_SIZE"].get("copy")
    if not fc or not wc:
        return None, "copy calibration kernel not found in the counter output"
    f_scale = copy_bytes / (fc * 1024.0)  # HBM bytes read per FETCH_SIZE byte
    w_scale = copy_bytes / (wc * 1024.0)
    per = {}
    for k, _ in KERNEL_KEYS:
        if k in out["FETCH_SIZE"] and k in out["WRITE_SIZE"]:
            per[k] = int(out["FETCH_SIZE"][k] * 1024 * f_scale + out["WRITE_SIZE"][k] * 1024 * w_scale)
    details = {"fetch_kib": {k: round(v, 1) for k, v in out["FETCH_SIZE"].items()},
               "write_kib": {k: round(v, 1) for k, v in out["WRITE_SIZE"].items()},
               "scale_fetch": round(f_scale, 4), "scale_write": round(w_scale, 4),
               "calibration": f"{COPY_MB} MiB device copy (reads and writes {copy_bytes} bytes)"}
    shutil.rmtree(tmp, ignore_errors=True)
    return per, details


def cpu_info():
    """Host CPU model string and logical core count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline_leg(mesh, order, params, u, u_free, checks, gdofs, seconds, out):
    """The CPU baseline: the oracle (FP64 restatement of the reference loop)
    on the same workload and state, timed on this host, first on one core
    for about `seconds`, then the element/face-parallel variant (bit-identical
    to the serial loop) on every thread OpenMP is given for about as long.
    Its residuals are also the checker of the device results gathered by
    the GPU legs (`checks`): the relative max-norm errors go into `out`."""
    from iceicle_amd import abi
    from oracle import oracle as O
    print("cpu baseline (oracle) ...", file=sys.stderr, flush=True)
    o = O.Oracle(mesh, abi.space_desc(order), params)

    def timed(fn):
        fn(u)  # warm
        n, t0 = 0, time.perf_counter()
        while True:
            fn(u)
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds or n >= 400:
                return n, el

    threads = O.max_threads()
    n1, s1 = timed(o.residual)
    nm, sm = timed(lambda x: o.residual_parallel(x, threads))

    def err(a, ref):
        return float(f"{np.max(np.abs(a - ref)) / np.max(np.abs(ref)):.3e}")

    ref = o.residual_parallel(u, threads)
    ref_fs = o.residual_parallel(u_free, threads)
    out["rel_err_vs_oracle"] = err(checks["fp64"], ref)
    out["rel_err_vs_oracle_free_stream"] = err(checks["fp64_free_stream"], ref_fs)
    out["mixed"]["rel_err_vs_oracle"] = err(checks["mixed"], ref)
    out["mixed"]["rel_err_vs_oracle_free_stream"] = err(checks["mixed_free_stream"], ref_fs)
    model, ncpu = cpu_info()
    multi = o.size * nm / sm / 1e9
    return {
        "value": round(multi, 6),
        "unit": "GDOF/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{nm} full FP64 residuals of the same workload and state on {threads} host threads "
                  f"({sm:.1f} s; element/face-parallel oracle loop, bit-identical to the serial one)",
        "single_core": {"value": round(o.size * n1 / s1 / 1e9, 6), "sample": f"{n1} residuals ({s1:.1f} s)"},
        "cpu_model": model, "nproc": ncpu,
        "gpu_speedup": round(gdofs / multi, 1),
    }


def c1_cpu_leg(g):
    """C1's CPU baseline: the oracle's RK3-SSP solve of the same 100 CFL
    steps on one core (the reference loop is serial), and the device's
    final state and residual checked against it."""
    from iceicle_amd import abi
    from oracle import oracle as O
    o = O.Oracle(g["mesh"], abi.space_desc(g["order"]), g["params"])
    r_ref = o.residual(g["u0"])
    t0 = time.perf_counter()
    ref, n_ref, t_ref = o.rk_solve(g["u0"], kind=0, compat=False, cfl=0.2, ntime=g["nsteps"])
    cpu_s = time.perf_counter() - t0
    return {
        "rk3_ssp_100_steps_seconds": round(cpu_s, 3), "cores": 1, "kind": "port",
        "gpu_speedup": round(cpu_s / g["gpu_seconds"], 1), "oracle_steps": n_ref,
        "final_state_rel_err_vs_oracle": float(f"{np.max(np.abs(g['u_end'] - ref)) / np.max(np.abs(ref)):.3e}"),
        "residual_rel_err_vs_oracle": float(f"{np.max(np.abs(g['r0'] - r_ref)) / np.max(np.abs(r_ref)):.3e}"),
    }


def c1_leg(nres=200, nsteps=100):
    """2D viscous Burgers (nu 0.01, b = (1, 1)) on a coupled periodic 32 x 32
    grid, P2, 3 x 3 GL, IC the L2 projection of a seeded sum of 4
    sinusoids: FP64 residual throughput, and the RK3-SSP solve of 100 steps
    at CFL 0.2 (device-side CFL step) timed on the GPU.  Returns the leg's
    record and the device results for the CPU-baseline leg."""
    import torch
    from iceicle_amd import abi, engine, workloads
    print("C1 Burgers 32x32 P2 ...", file=sys.stderr, flush=True)
    mesh, order, params, ic = workloads.burgers_2d(32, 2)
    space = engine.FESpace(mesh, order)
    disc = engine.ConservationLawDDG(space, params)
    u = torch.empty(disc.size, dtype=torch.float64, device="cuda")
    engine.project(disc, ic, u)
    u0 = u.cpu().numpy()
    r = torch.empty_like(u)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(20):
        engine.form_residual(space, disc, u, r)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(nres):
        engine.form_residual(space, disc, u, r)
    ev1.record()
    torch.cuda.synchronize()
    res_ms = ev0.elapsed_time(ev1) / nres
    r0 = r.cpu().numpy()
    # warm the stepper once, then time the full 100-step solve
    st = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_SSP, cfl=0.2, ntime=nsteps))
    st.solve(torch.from_numpy(u0).cuda())
    st = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_SSP, cfl=0.2, ntime=nsteps))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n, tend = st.solve(u)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    # the same steps replayed from one HIP graph (the CFL step size and the
    # termination flag stay on the device, so no host synchronisation)
    ug = torch.from_numpy(u0).cuda()
    sg = engine.ExplicitStepper(disc, abi.stepper_desc(abi.STEP_RK3_SSP, cfl=0.2, ntime=nsteps))
    graph = engine.StepGraph([sg], [ug], nsteps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graph.launch()
    torch.cuda.synchronize()
    graph_s = time.perf_counter() - t0
    graph_equal = bool(torch.equal(ug, u))
    del graph
    ndof = disc.size
    rec = {
        "workload": "2D viscous Burgers nu 0.01, b = (1, 1), 32x32 quads, coupled periodic, P2, 3x3 GL, "
                    "IC L2 projection of 4 seeded sinusoids, RK3-SSP CFL 0.2, 100 steps",
        "unknowns": ndof,
        "residual": {"GDOF_s": round(ndof / (res_ms * 1e-3) / 1e9, 4), "ms": round(res_ms, 5)},
        "rk3_ssp_100_steps": {"gpu_seconds": round(gpu_s, 5), "steps": n, "t_end": tend},
        "rk3_ssp_100_steps_graph": {"gpu_seconds": round(graph_s, 5), "equal_to_eager": graph_equal},
    }
    g = {"mesh": mesh, "order": order, "params": params, "u0": u0, "r0": r0, "u_end": u.cpu().numpy(),
         "nsteps": nsteps, "gpu_seconds": gpu_s}
    return rec, g


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
    # mixed-precision residual (FP32 interpolation, fluxes and face data),
    # with each element-kernel tile: 0-2 the VALU kernel in 256 / 128 / 64
    # threads, 3 the kernel whose tensor contractions run on the matrix cores
    r32 = torch.empty_like(u)
    per_tile = {}
    for t in range(4):
        disc.set_tile(t)
        for _ in range(3):
            engine.form_residual(space, disc, u, r32, precision="mixed")
        torch.cuda.synchronize()
        start.record()
        for _ in range(nres):
            engine.form_residual(space, disc, u, r32, precision="mixed")
        stop.record()
        torch.cuda.synchronize()
        per_tile[t] = (start.elapsed_time(stop) / nres, float((r32 - r).abs().max() / r.abs().max()))
    mixed_tile = min(per_tile, key=lambda t: per_tile[t][0])
    disc.set_tile(mixed_tile)
    mstage, mcalls, _ = kernel_event_pass(engine, space, disc, u, r32, "mixed", nres)
    mixed_ms, mixed_err = per_tile[mixed_tile]
    valu_best = min((0, 1, 2), key=lambda t: per_tile[t][0])
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
                  "rel_err_vs_fp64": mixed_err, "tile": mixed_tile,
                  "kernels_ms": {k: round(v / mcalls, 4) for k, v in zip(("trace", "flux", "element"), mstage)},
                  "element_kernel_matrix_cores": {
                      "ms_per_residual": round(per_tile[3][0], 4), "rel_err_vs_fp64": per_tile[3][1],
                      "instruction": "v_mfma_f32_4x4x1f32 (16 blocks of 4 x 4 per wave, one tensor line per lane)"},
                  "element_kernel_valu_best": {"tile": valu_best, "ms_per_residual": round(per_tile[valu_best][0], 4),
                                               "rel_err_vs_fp64": per_tile[valu_best][1]}},
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
    # RK3-SSP steps of the whole batch: eager launches, then one graph (the
    # order groups chained, or as concurrent branches)
    ms_eager = timed(bt.step, nsteps)
    graph = bt.graph(nsteps)
    ms_graph = timed(graph.launch, 3) / nsteps
    del graph
    graph = bt.graph(nsteps, concurrent=True)
    ms_graph_conc = timed(graph.launch, 3) / nsteps
    del graph
    finite = all(bool(torch.isfinite(g.u).all()) for g in bt.groups)
    # FP32 states: residuals, steps, and the FP64 check
    bt.reset()
    bt.reset_f32()
    ms32s = timed(lambda: bt.residual_f32(outs), reps)
    rm = [o.clone() for o in outs]
    bt.residual_f32(outs, check_fp64=True)
    torch.cuda.synchronize()
    err_check = max(float((m - c).abs().max() / c.abs().max()) for m, c in zip(rm, outs))
    for _ in range(nsteps):
        bt.step()
        bt.step_f32()
    torch.cuda.synchronize()
    drift = max(float((g.u32.double() - g.u).abs().max() / g.u.abs().max()) for g in bt.groups)
    ms_eager32 = timed(bt.step_f32, nsteps)
    graph = bt.graph_f32(nsteps)
    ms_graph32 = timed(graph.launch, 3) / nsteps
    del graph
    graph = bt.graph_f32(nsteps, concurrent=True)
    ms_graph32c = timed(graph.launch, 3) / nsteps
    del graph
    finite32 = all(bool(torch.isfinite(g.u32).all()) for g in bt.groups)
    return {
        "instances": ninst, "unknowns": int(bt.size), "groups": {g.order: len(g.insts) for g in bt.groups},
        "setup_seconds": round(t_setup, 2),
        "tiles": {str(k): {"best": v[0], "ms": [round(x, 4) for x in v[1]]} for k, v in tiles.items()},
        "residual_fp64": {"ms": round(ms64, 4), "GDOF_s": round(bt.size / (ms64 * 1e-3) / 1e9, 3)},
        "residual_mixed": {"ms": round(ms32, 4), "GDOF_s": round(bt.size / (ms32 * 1e-3) / 1e9, 3),
                           "rel_err_vs_fp64": float(f"{err32:.3e}")},
        "residual_per_instance": {"ms": round(ms_single, 3), "GDOF_s": round(bt.size / (ms_single * 1e-3) / 1e9, 3),
                                  "bit_identical_to_batched": bool(same)},
        "rk3_ssp_step_ms": {"eager": round(ms_eager, 4), "graph": round(ms_graph, 4),
                            "graph_concurrent_orders": round(ms_graph_conc, 4), "graph_steps": nsteps},
        "state_finite": finite,
        "fp32_state": {
            "residual_mixed_ms": round(ms32s, 4), "GDOF_s": round(bt.size / (ms32s * 1e-3) / 1e9, 3),
            "rel_err_vs_fp64_check": float(f"{err_check:.3e}"),
            "rk3_ssp_step_ms": {"eager": round(ms_eager32, 4), "graph": round(ms_graph32, 4),
                                "graph_concurrent_orders": round(ms_graph32c, 4)},
            f"state_drift_vs_fp64_after_{nsteps}_steps": float(f"{drift:.3e}"),
            "state_finite": finite32,
        },
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

    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.workload == "euler3d":
        print(json.dumps(euler3d(args, nres=5, nsteps=2) if args.quick_legs else euler3d(args)))
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

    # bytes each kernel moves by its design (state, geometry, face records,
    # the intermediate traces and face fluxes), for the per-kernel figures
    info = space.info
    D, nb, nfq = space.ndim, space.nbasis, space.nfqp
    E, nfi, nfb = space.nelem, info.nface_interior, info.nface_boundary
    F = nfi + nfb
    nvert = 2 ** D
    nv = 1 + D  # trace record: value + the element's share of the DDG gradient
    geo_rec = 8 * (D + 2)  # normal, w sqrt(g), h per face point
    trace_bytes = 8 * E * 2 * D * nfq * nv * neq
    design_bytes = {
        "trace": 8 * E * nb * neq + 8 * E * nvert * D + 4 * 2 * D * E + geo_rec * E * 2 * D * nfq + trace_bytes,
        "flux": trace_bytes + 16 * F + geo_rec * F * nfq + 8 * nfb * nfq * neq + 8 * F * nfq * neq,
        "element": 8 * E * nb * neq + 8 * E * nvert * D + 4 * 2 * D * E + 8 * F * nfq * neq
                   + 8 * E * nb * neq + 8 * E * nb * neq,
    }
    kms = {k: stage_ms[n] / ncalls for n, k in enumerate(("trace", "flux", "element"))}
    nnode = (args.n + 1) ** 2
    b_alg = workloads.residual_bytes(D, order, neq, E, nnode, F)
    res_gbs = b_alg / (ms_per_step * 1e-3) / 1e9

    # device results checked against the oracle in the CPU-baseline leg
    checks = {"fp64": r.cpu().numpy()}

    traffic, traffic_info = (None, "skipped (--no-traffic)")
    if not args.no_traffic:
        traffic, traffic_info = measure_traffic(args, tile)
        print(f"traffic: {traffic} {traffic_info if traffic is None else ''}", file=sys.stderr, flush=True)
    kernels = {k: {"ms": round(kms[k], 6), "design_bytes": int(design_bytes[k]),
                   "design_GBs": round(design_bytes[k] / (kms[k] * 1e-3) / 1e9, 1),
                   "hbm_traffic_bytes": (traffic or {}).get(k)} for k in kms}

    # mixed-precision residual on the same state: throughput and its error
    # against the FP64 oracle (and the device FP64 residual)
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
    checks["mixed"] = r.cpu().numpy()
    # the same residuals from the uniform free-stream state, where R is far from zero
    fs_state = np.tile([1.0, 0.3, 0.0, 1.0 / 1.4 / 0.4 + 0.045], disc.size // 4)
    uf = torch.from_numpy(fs_state).cuda()
    rf64, rf32 = torch.empty_like(uf), torch.empty_like(uf)
    engine.form_residual(space, disc, uf, rf64)
    engine.form_residual(space, disc, uf, rf32, precision="mixed")
    torch.cuda.synchronize()
    checks["fp64_free_stream"] = rf64.cpu().numpy()
    checks["mixed_free_stream"] = rf32.cpu().numpy()
    mixed = {
        "value": round(ndof / (mixed_ms * 1e-3) / 1e9, 4),
        "unit": "GDOF/s",
        "ms_per_step": round(mixed_ms, 6),
        "rel_err_vs_fp64": float(f"{mixed_err:.3e}"),
        "kernels_ms": {k: round(v / mcalls, 6) for k, v in zip(("trace", "flux", "element"), mstage)},
        "arithmetic": "FP32 interpolation, fluxes and face data; FP64 geometry, source load and output",
    }

    def newton_leg(label, **kw):
        # Newton-linesearch from a uniform free-stream state (rho 1, u 0.3,
        # v 0, p 1/gamma) to ||R||_2 < 1e-10
        un = torch.from_numpy(fs_state).cuda()
        kw.setdefault("preconditioner", abi.PC_MULTIGRID)
        solver = engine.PetscNewton(disc, tau_abs=1e-10, tau_rel=0.0, kmax=30, forcing=abi.FORCING_EW, **kw)
        print(f"newton ({label}) ...", file=sys.stderr, flush=True)
        rep = solver.solve(un)
        print(f"newton ({label}) done: {rep['iterations']} steps, {rep['seconds']:.2f} s", file=sys.stderr,
              flush=True)
        del solver
        return {
            "seconds": round(rep["seconds"], 4), "converged": rep["converged"], "steps": rep["iterations"],
            "residual_norms": [float(f"{v:.3e}") for v in rep["rhist"]],
            "alpha": [round(a, 4) for a in rep["alpha"]],
            "r0": float(f"{rep['r0']:.4e}"), "linear_iterations": rep["linear_iterations"],
            "jacobian_seconds": round(rep["jacobian_seconds"], 4), "linear_seconds": round(rep["linear_seconds"], 4),
            "linesearch_seconds": round(rep["linesearch_seconds"], 4),
        }

    newton = None
    if not args.no_newton:
        newton = newton_leg("Wolfe, FP32 blocks", linesearch=abi.LS_WOLFE, smoother_precision=abi.PREC_FP16)
        newton["initial_guess"] = "uniform free stream"
        newton["linesearch"] = "Wolfe (nonlinear_solver_utils.hpp:139-170)"
        newton["linear_solver"] = ("GMRES(30) (classical Gram-Schmidt, one pass) with an agglomeration multigrid "
                                   "V-cycle (P0 Galerkin coarse levels, block SGS smoothing on FP16 fine blocks), "
                                   "FP32 Jacobian blocks + FP64 refinement, Eisenstat-Walker forcing")
        newton["no_linesearch"] = newton_leg("full steps, FP32 blocks", linesearch=abi.LS_NONE,
                                             smoother_precision=abi.PREC_FP16)
        # FP32 Jacobian blocks + FP64 refinement against FP64 blocks, both
        # with the multicolour block-SGS preconditioner (the multigrid
        # hierarchy is built from the FP32 blocks only)
        newton["block_sgs"] = {
            "fp32_blocks_fp64_refinement": newton_leg("Wolfe, FP32 blocks, SGS", linesearch=abi.LS_WOLFE,
                                                      preconditioner=abi.PC_BLOCK_SGS,
                                                      smoother_precision=abi.PREC_FP32),
            "fp64_blocks": newton_leg("Wolfe, FP64 blocks, SGS", linesearch=abi.LS_WOLFE,
                                      preconditioner=abi.PC_BLOCK_SGS, jacobian_precision=abi.PREC_FP64,
                                      smoother_precision=abi.PREC_FP32),
        }

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
        "kernels": kernels,
        "element_tile": {"best": tile, "ms": [round(t, 5) for t in tile_ms],
                         "threads": [256, 128, 64]},
        "roofline": {
            "bound": "hbm",
            "kernel": "residual = trace_kernel + flux_kernel + element_kernel_2d (3 launches, timed pass)",
            "achieved": round(res_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(res_gbs / HBM_PEAK_GBS, 5),
            "traffic": (int(sum(traffic.values())) if traffic and len(traffic) == 3 else None),
            "bytes_per_launch": int(b_alg),
            "bytes_formula": "2*8*m*E*nb + 8*d*N_vert + 4*2^d*E + 12*(F_i+F_b)",
            "traffic_source": traffic_info if traffic is None else
            dict(traffic_info, how="rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE child runs of this script"),
        },
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
    c1 = None
    if not args.no_c1:
        c1, c1_gpu = c1_leg()
        out["c1_burgers"] = c1
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_leg(mesh, order, params, u_host, fs_state, checks, gdofs,
                                               args.cpu_seconds, out)
        if c1 is not None:
            c1["cpu_baseline"] = c1_cpu_leg(c1_gpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
