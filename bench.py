"""Benchmark: the 2048x2048 stand-alone p' solve, learned FCG versus Jacobi-PCG,
and SIMPLE iterations on one MI355X.

Workload: 2D lid-driven cavity, 2048x2048 uniform cells. A step is one
stand-alone pressure-correction solve on the 5-point p' operator (d = 1,
rho = 1, cell 0 pinned) with an RHS drawn seeded-uniform in [-1, 1]
(splitmix64, seed 20480001), its serial mean removed and b_0 = 0, solved from
x0 = 0 to an FP64 true residual |b - A x| <= 1e-8 |b|. Two methods, timed
back to back on the same resident operator and RHS:
  * learned: flexible CG (FP64 vectors) preconditioned by the U-Net (BF16
    weights, FP32 activations), the shipped weights of cfdapp_amd/weights
    (the V-cycle parametrisation with Neumann padding and the coarse tail);
  * plain: Jacobi-PCG, FP64 single-kernel iteration.
`value` is the learned solve's time. Also reported: the FP32-inner mixed
Jacobi-PCG, SIMPLE outer iterations per second at 2048x2048, Re = 1000, the
per-layer kernel times of the preconditioner, the plain iteration kernel's
HBM roofline, and the batched leg: the p' systems of the first SIMPLE
iteration of 256 cavities and channels of uneven size (seed 5005) solved
together by the batched learned FCG and by the batched mixed-precision
Jacobi-PCG (aggregate cells x iterations per second, per-case p50 / p99
completion time) and one after another by the single-case learned FCG;
the reference's own 129^2 cavity (Re 100) SIMPLE iterations per second with
stream-launched and graph-replayed p' iteration chunks; the PISO leg: time
steps of a 1024^2 cavity at Re 1000 with the reference's BiCGSTAB p' solves
and with the learned FCG; and the 3D leg: SIMPLE iterations on the 256^3
cavity at Re = 400 with FP64 and with mixed-precision p' solves, and the
mixed run's pressure / velocity error against the FP64 run.

Prints one JSON line (see the repository's bench contract).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from cfdapp_amd.workloads import batch_cases, standalone_rhs  # noqa: E402

N = 2048
SEED = 20480001
HBM_PEAK_GBS = 8000.0
# HBM bytes of one cgIterKernel<double> launch at 2048^2 from rocprofv3 counters
# (2 x FETCH_SIZE + WRITE_SIZE, calibrated on a 512 MiB copy kernel; see profiles/)
TRAFFIC_BYTES = 458.0e6


def layer_bytes(p0, ch):
    """Algorithmic HBM bytes of each preconditioner layer on the inference
    path: every stored tensor read and written once at the channels it stores
    per pixel (ch: tensor name -> channels, FP32; the shipped weights keep 2,
    the full network 16 .. 64), the FP64 residual and the FP32 scaling into
    the first layer, and the FP64 q and z of the output layer (p0 = fine
    pixels)."""
    px = [p0, p0 / 4, p0 / 16, p0 / 64]
    level = {"e0": 0, "p0": 1, "e1a": 1, "e1": 1, "p1": 2, "e2a": 2, "e2": 2, "p2": 3,
             "b3a": 3, "b3": 3, "d2a": 2, "d2": 2, "d1a": 1, "d1": 1, "d0a": 0, "d0": 0}

    def t(name):
        return 4.0 * ch[name] * px[level[name]]

    return {
        "e0b": p0 * (8 + 4) + t("e0") + t("p0"),
        "e1a": t("p0") + t("e1a"),
        "e1b": t("e1a") + t("e1") + t("p1"),
        "e2a": t("p1") + t("e2a"),
        "e2b": t("e2a") + t("e2") + t("p2"),
        "b3a": t("p2") + t("b3a"),
        "b3b": t("b3a") + t("b3"),
        "d2a": t("b3") + t("e2") + t("d2a"),
        "d2b": t("d2a") + t("d2"),
        "d1a": t("d2") + t("e1") + t("d1a"),
        "d1b": t("d1a") + t("d1"),
        "d0a": t("d1") + t("e0") + t("d0a"),
        "d0b": t("d0a") + t("d0"),
        "out": t("d0") + p0 * (4 + 8 + 8),
        "e0a": 0.0,
        # coarse tail: one channel of b3a in, one out, and its levels (H/16 ..)
        # each read and written a few times
        "tail": 4.0 * (2 * px[3] + 6 * px[3] / 3),
    }


def _lscpu():
    """Model name and physical core count of the host (lscpu)."""
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return {}
    info = {}
    for line in text.splitlines():
        key, _, val = line.partition(":")
        info[key.strip()] = val.strip()
    try:
        cores = int(info.get("Core(s) per socket", "0")) * int(info.get("Socket(s)", "1"))
    except ValueError:
        cores = 0
    return {"model": info.get("Model name", ""), "physical_cores": cores,
            "logical_cpus": int(info.get("CPU(s)", "0") or 0)}


def _ref(driver, args, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    out = subprocess.run([driver, *args], capture_output=True, text=True, env=env, timeout=600,
                         check=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def cpu_baseline(iters_gpu, pprime_its_per_outer):
    """The reference's own CPU path, compiled from its sources (oracle/Makefile.ref:
    -O3 -fopenmp -ffp-contract=off; OpenMP CSR SpMV rows, SparseMatrix.cpp:91), on
    this host, median of 3 runs each, with OMP_NUM_THREADS = the physical cores
    (at most the 16 cores this GPU's share holds) and with 1 thread:
      * C2 stand-alone p' solve: CG + Jacobi iterations on the 2048^2 system
        (10 per run at full width, 3 at one thread), scaled to the GPU solve's
        iteration count -> `value`;
      * C1 129^2 cavity, Re 100, CG + Jacobi p' to 1e-8: 5 SIMPLE iterations
        per run -> SIMPLE iter/s and p' seconds per outer iteration;
      * C2 SIMPLE outer iteration at 2048^2, Re 1000: the stages before the p'
        solve timed at 512^2 (one outer iteration stopped at its p' solve) and
        scaled by the cell count, plus the p' solve at the CG per-iteration cost
        times the GPU run's p' iterations per outer iteration (a full CPU
        iteration at 2048^2 takes several minutes).
    Falls back to the oracle restatement (one thread, 10 iterations) when the
    reference build is absent."""
    import statistics
    driver = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    cpu = _lscpu()
    threads = max(1, min(cpu.get("physical_cores") or (os.cpu_count() or 1), 16))
    if not os.path.exists(driver):
        import oracle
        c = oracle.Case.cavity(2, N, N)
        ones = np.ones(c.n_cells)
        A, _, _ = c.pressure_system(np.zeros(c.n_faces), ones, ones)
        b = standalone_rhs(c.n_cells)
        t0 = time.perf_counter()
        r = A.cg(b, max_iter=10)
        per_iter = (time.perf_counter() - t0) / max(r.iterations, 1)
        return {"value": per_iter * iters_gpu, "unit": "s", "cores": 1, "kind": "port",
                "sample": f"10 CG+Jacobi iterations of the oracle on the 2048x2048 p' system "
                          f"({per_iter * 1e3:.1f} ms/iteration), scaled to the GPU solve's "
                          f"{iters_gpu} iterations", "host": cpu}
    med = statistics.median
    per_it = {}
    for t, k in ((threads, 10), (1, 3)):
        runs = [_ref(driver, ["cgtime", "dim=2", f"nx={N}", f"ny={N}", f"seed={SEED}",
                              f"pmax={k}"], t) for _ in range(3)]
        per_it[t] = med(r["seconds"] / max(r["iterations"], 1) for r in runs)
    c1 = {}
    for t in (threads, 1):
        runs = [_ref(driver, ["simpletime", "dim=2", "nx=129", "ny=129", "re=100", "iters=5",
                              "flux=rhie_chow", "psolver=cg_jacobi", "pabs=0", "prel=1e-8"], t)
                for _ in range(3)]
        c1[f"threads_{t}"] = {
            "simple_iter_per_s": med(r["iterations"] / r["seconds"] for r in runs),
            "pprime_s_per_outer_iteration": med(r["pressure_solve"] / r["iterations"]
                                                for r in runs)}
    r = _ref(driver, ["simpletime", "dim=2", "nx=512", "ny=512", "re=1000", "iters=1",
                      "flux=rhie_chow", "psolver=cg_jacobi", "pabs=0", "prel=1e-8", "pmax=1"],
             threads)
    before_pprime = (r["setup"] + r["momentum_solve"] + r["discretization"]) * (N * N) / 512 ** 2
    c2_outer = before_pprime + per_it[threads] * pprime_its_per_outer
    return {"value": per_it[threads] * iters_gpu, "unit": "s", "cores": threads,
            "kind": "reference",
            "sample": f"median of 3 runs of 10 CG+Jacobi iterations (the reference's CG, "
                      f"OpenMP SpMV, {threads} threads) on the 2048x2048 p' system "
                      f"({per_it[threads] * 1e3:.1f} ms/iteration), scaled to the GPU solve's "
                      f"{iters_gpu} iterations",
            "build": "g++ -O3 -fopenmp -ffp-contract=off (oracle/Makefile.ref)",
            "host": cpu, "omp": "OMP_PROC_BIND=close OMP_PLACES=cores",
            "one_thread_value": per_it[1] * iters_gpu,
            "one_thread_ms_per_iteration": per_it[1] * 1e3,
            "c1_cavity129_re100": c1,
            "c2_simple_s_per_outer_iteration": c2_outer,
            "c2_simple_sample": f"stages before the p' solve of one outer iteration at 512^2 "
                                f"({r['seconds']:.2f} s) scaled by 16, plus "
                                f"{pprime_its_per_outer:.0f} p' CG iterations"}


def batch_leg(cfd, log, cases, use_graph=True, sequential=True):
    """p' systems of the first SIMPLE iteration of each case, every one to
    1e-8 FP64 true relative residual: (a) the batched flexible CG with the
    learned preconditioner (FP64 vectors), one launch per step over all active
    cases; (b) the batched mixed-precision Jacobi-PCG, one launch per step;
    (c) one case after another by the single-case learned FCG."""
    from cfdapp_amd import learned
    t0 = time.perf_counter()
    batch = cfd.BatchSolver()
    batch.set_learned()
    fcg = cfd.LinearSolverSettings(0.0, 1e-8, 2000, "FCG", "Learned", check_interval=4)
    seq_s, seq_lat = 0.0, []
    cells = 0
    sizes = []
    for kind, nx, ny, re in cases:
        d = learned.make_case(kind, nx, ny=ny, re=re, speed=1.0)
        try:
            learned.assemble_pprime(d, 0)
            batch.add_pressure(d)
            cells += nx * ny
            sizes.append(nx * ny)
            if not sequential:
                continue
            d.set_learned_weights()
            d._ctx.call("cfdhip_synchronize")
            t1 = time.perf_counter()
            r = d.solve_pressure_rhs(fcg, None, download=False)
            dt = time.perf_counter() - t1
            if r.status != 0:
                raise RuntimeError(f"sequential {kind} {nx}x{ny}: {r.status_name}")
            seq_s += dt
            seq_lat.append(dt * 1e3)
        finally:
            d.close()
    log(f"batch: {len(cases)} cases, {cells} cells, set up in {time.perf_counter() - t0:.1f} s")
    sizes = np.array(sizes, np.float64)

    def timed(settings, label):
        batch.solve(settings, use_graph=use_graph)  # warm-up (the Jacobi leg tunes its tiles)
        t1 = time.perf_counter()
        res = batch.solve(settings, use_graph=use_graph)
        wall = time.perf_counter() - t1
        bad = [k for k, r in enumerate(res) if r.status != 0 or r.fallback_used]
        if bad:
            raise RuntimeError(f"batched {label}: {len(bad)} cases did not converge on their own")
        return res, wall, dict(batch.stats)

    res, wall, st = timed(fcg, "learned FCG")
    lat = np.array([r.device_ms for r in res])
    its = np.array([r.iterations for r in res], np.float64)
    out = {"cases": len(cases), "cells": cells,
           "learned_batched_s": wall, "learned_batched_device_ms": st["device_ms"],
           "learned_cell_iter_per_s": float((its * sizes).sum()) / wall,
           "learned_p50_ms": float(np.percentile(lat, 50)),
           "learned_p99_ms": float(np.percentile(lat, 99)),
           "learned_iterations_min": int(its.min()), "learned_iterations_max": int(its.max()),
           "learned_rounds": st["steps"], "learned_graph_captures": st["graph_captures"],
           "learned_max_true_rel_res": float(max(r.true_residual / r.initial_residual
                                                 for r in res if r.initial_residual > 0))}
    bs = cfd.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi", precision="mixed",
                                  check_interval=32)
    res_j, wall_j, st_j = timed(bs, "Jacobi-PCG")
    lat_j = np.array([r.device_ms for r in res_j])
    out.update({"jacobi_batched_s": wall_j, "jacobi_batched_device_ms": st_j["device_ms"],
                "jacobi_cell_iter_per_s": st_j["cell_iterations"] / wall_j,
                "jacobi_p50_ms": float(np.percentile(lat_j, 50)),
                "jacobi_p99_ms": float(np.percentile(lat_j, 99)),
                "jacobi_tile_rows": st_j["tile_rows"], "jacobi_steps": st_j["steps"],
                "jacobi_graph_captures": st_j["graph_captures"], "tile_tuning_us": st_j["tune_us"],
                "speedup_learned_vs_jacobi_batched": wall_j / wall})
    if sequential:
        out.update({"sequential_learned_s": seq_s,
                    "sequential_learned_p50_ms": float(np.percentile(seq_lat, 50)),
                    "sequential_learned_p99_ms": float(np.percentile(seq_lat, 99)),
                    "speedup_batched_vs_sequential_learned": seq_s / wall})
    log(f"batch: learned FCG batched {wall * 1e3:.1f} ms ({out['learned_cell_iter_per_s']:.3e} "
        f"cell-it/s, {its.min():.0f}..{its.max():.0f} it, p50 {out['learned_p50_ms']:.1f} ms, "
        f"p99 {out['learned_p99_ms']:.1f} ms); Jacobi-PCG batched {wall_j * 1e3:.1f} ms "
        f"(p50 {out['jacobi_p50_ms']:.1f} ms, p99 {out['jacobi_p99_ms']:.1f} ms)" +
        (f"; learned FCG one case at a time {seq_s * 1e3:.1f} ms (p50 "
         f"{out['sequential_learned_p50_ms']:.1f} ms)" if sequential else ""))
    batch.close()
    return out


def cavity129_leg(cfd, log, iters):
    """The reference's CPU-runnable case: 129^2 cavity, Re 100 (mu = 0.01),
    Rhie-Chow, alpha_u 0.7, alpha_p 0.3, CG + Jacobi p' to 1e-8, BiCGSTAB
    momentum; `iters` SIMPLE iterations from rest with stream-launched p'
    iteration chunks and again with graph-replayed chunks."""
    out = {}
    for graph in (False, True):
        d = cfd.GpuSimpleDiscretization.cavity(2, 129, 129, viscosity=0.01)
        try:
            ps = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi", check_interval=32)
            ms = cfd.LinearSolverSettings(1e-10, 1e-8, 500, "BiCGSTAB")
            d.simple_run(2, pressure=ps, momentum=ms, use_graph=graph)  # warm-up
            d.upload("U", np.zeros(d.n_cells))
            d.upload("V", np.zeros(d.n_cells))
            d.upload("P", np.zeros(d.n_cells))
            d.upload("FLUX", np.zeros(d.n_faces))
            d._ctx.call("cfdhip_synchronize")
            t0 = time.perf_counter()
            res = d.simple_run(iters, pressure=ps, momentum=ms, use_graph=graph)
            dt = time.perf_counter() - t0
            key = "graph" if graph else "stream"
            out[f"{key}_simple_iter_per_s"] = res["iterations"] / dt
            out[f"{key}_pressure_iterations"] = res["pressure_iterations"]
        finally:
            d.close()
    out["simple_iterations"] = iters
    log(f"129^2 cavity Re 100: {iters} SIMPLE iterations, stream launches "
        f"{out['stream_simple_iter_per_s']:.1f} it/s, graph replay "
        f"{out['graph_simple_iter_per_s']:.1f} it/s ({out['graph_pressure_iterations']} p' it)")
    return out


def piso_leg(cfd, log, n, steps):
    """PISO time steps (PISO::solveTimeStep) on an n^2 cavity at Re 1000 from
    rest, dt = h / 2 (lid Courant number about 1): p' by the reference's
    BiCGSTAB, then by the learned FCG; time steps per second and the
    velocity / pressure difference between the two runs."""
    out, fields = {"cells": n * n, "steps": steps, "dt": 0.5 / n}, {}
    mom = cfd.LinearSolverSettings(1e-10, 1e-8, 500, "BiCGSTAB")
    for label, ps in (("bicgstab", cfd.LinearSolverSettings(1e-10, 1e-8, 40000, "BiCGSTAB")),
                      ("learned", cfd.LinearSolverSettings(1e-10, 1e-8, 20000, "FCG",
                                                           "Learned"))):
        d = cfd.GpuSimpleDiscretization.cavity(2, n, n, viscosity=1e-3)
        try:
            if label == "learned":
                d.set_learned_weights()
            d._ctx.call("cfdhip_mass_flux")
            d._ctx.call("cfdhip_synchronize")
            t0 = time.perf_counter()
            r = d.piso_run(steps, 0.5 / n, momentum=mom, pressure=ps)
            dt = time.perf_counter() - t0
            if r["taken"] != steps or any(s["status"] != 0 for s in r["steps"]):
                raise RuntimeError(f"PISO ({label}) failed: {r['steps'][-1]['status_name']}")
            fields[label] = (d.download("P"), d.download_velocity())
            out[f"{label}_steps_per_s"] = steps / dt
            out[f"{label}_pressure_iterations"] = sum(s["pressure_iterations"] for s in r["steps"])
            out[f"{label}_max_cfl"] = r["steps"][-1]["max_cfl"]
        finally:
            d.close()
    pb, ub = fields["bicgstab"]
    pl, ul = fields["learned"]
    out["speedup_learned_vs_bicgstab"] = out["learned_steps_per_s"] / out["bicgstab_steps_per_s"]
    out["p_rel_err_inf"] = float(np.max(np.abs(pl - pb)) / np.max(np.abs(pb)))
    umax = max(float(np.max(np.abs(c))) for c in ub)
    out["u_rel_err_inf"] = max(float(np.max(np.abs(a - b))) for a, b in zip(ul, ub)) / umax
    log(f"PISO {n}^2 Re 1000, {steps} steps: BiCGSTAB p' {out['bicgstab_steps_per_s']:.2f} "
        f"steps/s ({out['bicgstab_pressure_iterations']} p' it), learned FCG "
        f"{out['learned_steps_per_s']:.2f} steps/s ({out['learned_pressure_iterations']} p' it); "
        f"|dp|/|p| {out['p_rel_err_inf']:.1e}, |du|/|u| {out['u_rel_err_inf']:.1e}")
    return out


def mixed3d_leg(cfd, log, n, iters):
    """3D lid-driven cavity n^3 at Re = 400 (mu = 0.0025, lid ymax (1, 0, 0),
    Rhie-Chow): `iters` SIMPLE iterations with FP64 Jacobi-PCG p' solves, then
    the same with FP32 inner PCG inside FP64 refinement; the mixed run's
    pressure and velocity against the FP64 run's (max-norm relative errors)."""
    out, fields = {"cells": n ** 3, "simple_iterations": iters}, {}
    for prec in ("fp64", "mixed"):
        d = cfd.GpuSimpleDiscretization.cavity(3, n, n, n, viscosity=0.0025)
        try:
            ps = cfd.LinearSolverSettings(0.0, 1e-8, 40000, "CG", "Jacobi", precision=prec,
                                          check_interval=64)
            d._ctx.call("cfdhip_synchronize")
            t0 = time.perf_counter()
            res = d.simple_run(iters, pressure=ps)
            dt = time.perf_counter() - t0
            if res["iterations"] != iters:
                raise RuntimeError(f"3D SIMPLE ({prec}) stopped: status {res['status']}")
            fields[prec] = (d.download("P"), d.download_velocity())
            out[f"{prec}_simple_iter_per_s"] = iters / dt
            out[f"{prec}_pressure_iterations"] = res["pressure_iterations"]
        finally:
            d.close()
    p64, u64 = fields["fp64"]
    pm, um = fields["mixed"]
    out["p_rel_err_inf"] = float(np.max(np.abs(pm - p64)) / np.max(np.abs(p64)))
    umax = max(float(np.max(np.abs(c))) for c in u64)
    out["u_rel_err_inf"] = max(float(np.max(np.abs(a - b))) for a, b in zip(um, u64)) / umax
    log(f"3D {n}^3 Re 400, {iters} SIMPLE iterations: FP64 {out['fp64_simple_iter_per_s']:.3f} it/s "
        f"({out['fp64_pressure_iterations']} p' it), mixed {out['mixed_simple_iter_per_s']:.3f} "
        f"it/s ({out['mixed_pressure_iterations']} p' it); mixed vs FP64: |dp|/|p| "
        f"{out['p_rel_err_inf']:.2e}, |du|/|u| {out['u_rel_err_inf']:.2e}")
    return out


def _check(r, bnorm, label):
    """A timed solve counts only if it converged to the FP64 target."""
    rel = r.true_residual / bnorm
    if r.status != 0 or not rel <= 1.05e-8:
        raise RuntimeError(f"{label}: {r.status_name} after {r.iterations} iterations, true "
                           f"relative residual {rel:.3e}: not a valid measurement")
    return rel


def pprime_compare(cfd, log, d, label):
    """The p' system assembled on `d` solved by learned FCG and by Jacobi-PCG
    to the FP64 target (x0 = 0); their times, iterations and agreement."""
    b = d.download("PC_RHS")
    bnorm = float(np.linalg.norm(b))
    fcg = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")
    jac = cfd.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi", check_interval=64)
    out, x = {}, {}
    for name, st in (("learned_fcg", fcg), ("jacobi_pcg", jac)):
        d.solve_pressure_rhs(st, None, download=False)  # warm-up
        d._ctx.call("cfdhip_synchronize")
        t0 = time.perf_counter()
        r = d.solve_pressure_rhs(st, None, download=True)
        dt = time.perf_counter() - t0
        rel = _check(r, bnorm, f"{label} {name}")
        x[name] = r.solution
        out.update({f"{name}_s": dt, f"{name}_iterations": r.iterations,
                    f"{name}_true_relative_residual": rel, f"{name}_status": r.status_name})
        if name == "learned_fcg":
            out["learned_fcg_fallback"] = bool(r.fallback_used)
    out["speedup_learned_vs_jacobi"] = out["jacobi_pcg_s"] / out["learned_fcg_s"]
    xl, xj = x["learned_fcg"], x["jacobi_pcg"]
    out["solution_rel_diff_inf"] = float(np.max(np.abs(xl - xj)) / np.max(np.abs(xj)))
    log(f"{label}: learned FCG {out['learned_fcg_iterations']} it {out['learned_fcg_s'] * 1e3:.1f} "
        f"ms, Jacobi-PCG {out['jacobi_pcg_iterations']} it {out['jacobi_pcg_s'] * 1e3:.1f} ms "
        f"({out['speedup_learned_vs_jacobi']:.1f}x), solutions agree to "
        f"{out['solution_rel_diff_inf']:.1e}")
    return out


def c4_leg(cfd, log, n=2048, warm=3):
    """The learned preconditioner on a cavity it was not trained on: n^2, Re
    3200 (lid 1, mu = 1/3200); `warm` SIMPLE iterations from rest, then the next
    iteration's p' system solved by learned FCG and by Jacobi-PCG."""
    from cfdapp_amd import learned
    d = learned.make_case("cavity", n, re=3200.0, speed=1.0)
    try:
        d.set_learned_weights()
        fcg = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")
        mom = cfd.LinearSolverSettings(1e-12, 1e-8, 2000, "BiCGSTAB")
        res = d.simple_run(warm, pressure=fcg, momentum=mom)
        if res["iterations"] != warm:
            raise RuntimeError(f"C4 warm-up SIMPLE stopped: status {res['status']}")
        learned.assemble_pprime(d, 0)
        out = {"cells": n * n, "re": 3200.0, "warm_simple_iterations": warm,
               **pprime_compare(cfd, log, d, f"C4 {n}^2 Re 3200 p' after {warm} iterations")}
    finally:
        d.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--simple-iters", type=int, default=3,
                    help="2048^2 SIMPLE iterations from rest with each p' solver (0 skips)")
    ap.add_argument("--c2iii-outer", type=int, default=10,
                    help="the learned SIMPLE run continues to this outer iteration and solves "
                         "the next p' system with both methods (<= --simple-iters skips)")
    ap.add_argument("--no-c4", dest="c4", action="store_false",
                    help="skip the 2048^2 Re 3200 learned vs Jacobi p' leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c129-iters", type=int, default=200,
                    help="SIMPLE iterations of the 129^2 cavity leg (0 skips it)")
    ap.add_argument("--cube", type=int, default=256, help="3D mixed-precision leg: cells per side")
    ap.add_argument("--cube-iters", type=int, default=2,
                    help="SIMPLE iterations of the 3D leg (0 skips it)")
    ap.add_argument("--piso-n", type=int, default=1024, help="PISO leg: cells per side")
    ap.add_argument("--piso-steps", type=int, default=3,
                    help="PISO time steps of the PISO leg (0 skips it)")
    ap.add_argument("--batch-cases", type=int, default=256,
                    help="cases of the batched-solve leg (0 skips it)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 512 MiB device copy 20 times (counter calibration)")
    ap.add_argument("--jacobi-only", action="store_true",
                    help="only the Jacobi-PCG solves (for kernel counters)")
    ap.add_argument("--learned-only", action="store_true",
                    help="only the learned FCG solves (for kernel profiles and counters)")
    ap.add_argument("--learned-dense", action="store_true",
                    help="run the learned preconditioner through the MFMA layers over every "
                         "channel (for comparison)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched-solve leg (for kernel profiles)")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="batched leg with stream launches instead of graph replay")
    ap.add_argument("--c129-only", action="store_true",
                    help="only the 129^2 cavity SIMPLE leg (for kernel profiles)")
    ap.add_argument("--cube-only", action="store_true",
                    help="only the 3D cavity leg, FP64 and mixed (for kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU decomposition is a design, not built", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        a = torch.ones(64 * 2 ** 20, dtype=torch.float64, device="cuda")  # 512 MiB
        c = torch.empty_like(a)
        for _ in range(20):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8, "launches": 20}))
        return

    import cfdapp_amd as cfd
    if not cfd.device_available():
        sys.exit("bench.py: no gfx950 device")

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.c129_only:
        print(json.dumps(cavity129_leg(cfd, log, args.c129_iters)), flush=True)
        return
    if args.cube_only:
        print(json.dumps(mixed3d_leg(cfd, log, args.cube, args.cube_iters)), flush=True)
        return
    if args.batch_only:
        print(json.dumps(batch_leg(cfd, log, batch_cases(count=args.batch_cases),
                                   use_graph=not args.batch_no_graph, sequential=False)),
              flush=True)
        return

    # --- resident operator + RHS ------------------------------------------------
    d = cfd.GpuSimpleDiscretization.cavity(2, N, N)
    n, nf = d.n_cells, d.n_faces
    d.upload("RESPONSE", np.ones(n))
    d.upload("FLUX_STAR", np.zeros(nf))
    d.assemble_pressure_correction_resident(0)
    b = standalone_rhs(n)
    bnorm = float(np.linalg.norm(b))
    d.upload("PC_RHS", b)
    pcg = cfd.LinearSolverSettings(0.0, 1e-8, 40000, "CG", "Jacobi", check_interval=64)
    if not os.path.exists(cfd.WEIGHTS_PATH):
        sys.exit(f"bench.py: no learned weights at {cfd.WEIGHTS_PATH}")
    opts = cfd.load_learned_options()
    weights_kind = ("V-cycle parametrisation, Neumann padding, coarse tail "
                    f"({opts['tail']['levels']} levels max), model selected by the training run"
                    if opts.get("padding") == "neumann" else "zero-padded network")
    d.set_learned_weights(options=dict(opts, dense=args.learned_dense))
    fcg = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")
    log(f"setup done (learned weights: {weights_kind})")

    # --- the plain path's SpMV on its own (y = A x, device buffers) ----------------
    # x = the resident RHS, y = the pressure field (not used by the stand-alone solve)
    xs, ys = d.field_device("PC_RHS"), d.field_device("P")
    d._ctx.call("cfdhip_pressure_spmv_device", xs, ys, 5)
    d.set_timing(True)
    d._ctx.call("cfdhip_pressure_spmv_device", xs, ys, 50)
    d._ctx.call("cfdhip_synchronize")
    spmv_us = d.kernel_timing()["mean_us"]
    d.set_timing(False)
    # operator format: diag (n) + one coupling per face (nf), x (n), y (n)
    spmv_fmt_bytes = 8.0 * (3 * n + nf)
    spmv_dia_bytes = 8.0 * (5 + 2) * n  # (k + 2) n s with k = 5
    spmv_bytes = min(spmv_fmt_bytes, spmv_dia_bytes)
    spmv = {"kernel": "spmv2dKernel", "mean_us": spmv_us, "bytes": spmv_bytes,
            "gbs": spmv_bytes / (spmv_us * 1e-6) / 1e9,
            "frac": spmv_bytes / (spmv_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "dia_count_bytes": spmv_dia_bytes}
    log(f"SpMV y = A x: {spmv_us:.2f} us, {spmv['gbs']:.0f} GB/s on {spmv_bytes / 1e6:.1f} MB "
        f"({100 * spmv['frac']:.1f} % of HBM peak)")

    def timed(settings, label):
        for w in range(args.warmup):
            r = d.solve_pressure_rhs(settings, None, download=False)
            log(f"{label} warmup {w}: {r.status_name} {r.iterations} it {r.device_ms:.1f} ms")
        d._ctx.call("cfdhip_synchronize")
        rs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rs.append(d.solve_pressure_rhs(settings, None, download=False))
        d._ctx.call("cfdhip_synchronize")
        sec = (time.perf_counter() - t0) / args.steps
        for r in rs:
            _check(r, bnorm, label)
        r = rs[-1]
        log(f"{label}: {args.steps} solves, {sec * 1e3:.1f} ms/solve, {r.status_name}, "
            f"{r.iterations} it, true rel res {r.true_residual / bnorm:.3e}, "
            f"fallback {r.fallback_used}")
        return sec, rs

    if args.jacobi_only:
        timed(pcg, "Jacobi-PCG")
        return
    if args.learned_only:
        timed(fcg, "learned FCG")
        return
    fcg_s, fcg_rs = timed(fcg, "learned FCG")
    jac_s, jac_rs = timed(pcg, "Jacobi-PCG")
    rf, rj = fcg_rs[-1], jac_rs[-1]
    ms_per_step = fcg_s * 1e3

    # --- kernel timing pass: plain iteration kernel and the preconditioner layers
    d.set_timing(True)
    d.solve_pressure_rhs(pcg, None, download=False)
    kt = d.kernel_timing()
    d.solve_pressure_rhs(fcg, None, download=False)
    layer_ms, layer_calls = d.preconditioner_timing()
    d.set_timing(False)
    spmv_us = kt["mean_us"]
    # algorithmic bytes of one single-kernel PCG iteration (FP64): reads p, s,
    # r, w, x, diag (n each) and the face couplings (nf); writes p, s, r, w, x
    spmv_bytes = 8.0 * (11 * n + nf)
    achieved = spmv_bytes / (spmv_us * 1e-6) / 1e9 if spmv_us > 0 else 0.0
    log(f"PCG iteration kernel: {kt['launches']:.0f} launches, mean {spmv_us:.2f} us, "
        f"{achieved:.0f} GB/s algorithmic")
    names = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "b3a", "b3b", "d2a", "d2b", "d1a", "d1b",
             "d0a", "d0b", "out", "tail"]
    p0 = float(n)
    channels = d.learned_channels()
    lb = layer_bytes(p0, channels)
    layers = {}
    for i, nm in enumerate(names):
        if layer_calls[i] > 0:
            us = 1e3 * layer_ms[i] / layer_calls[i]
            layers[nm] = {"us": us, "bytes": lb[nm], "gbs": lb[nm] / (us * 1e-6) / 1e9}
    per_apply_us = sum(v["us"] for v in layers.values())
    log("preconditioner layers (us, GB/s): " +
        ", ".join(f"{k} {v['us']:.0f} {v['gbs']:.0f}" for k, v in layers.items()) +
        f"; total {per_apply_us:.0f} us per application")
    dom = max(layers, key=lambda k: layers[k]["us"])

    # --- mixed precision (FP32 inner PCG, FP64 refinement) on the same system ------
    mx = cfd.LinearSolverSettings(0.0, 1e-8, 40000, "CG", "Jacobi", precision="mixed",
                                  check_interval=64)
    d.solve_pressure_rhs(mx, None, download=False)
    d._ctx.call("cfdhip_synchronize")
    t2 = time.perf_counter()
    rm = d.solve_pressure_rhs(mx, None, download=False)
    dtm = time.perf_counter() - t2
    _check(rm, bnorm, "mixed Jacobi-PCG")
    mixed = {"mixed_jacobi_pcg_s": dtm, "mixed_iterations": rm.iterations,
             "mixed_refinements": rm.restarts,
             "mixed_true_relative_residual": rm.true_residual / bnorm}
    log(f"mixed Jacobi-PCG: {rm.status_name} {rm.iterations} it ({rm.restarts} cycles) "
        f"{dtm * 1e3:.1f} ms, true rel res {rm.true_residual / bnorm:.3e}")

    # --- mixed-precision learned FCG (FP32 Krylov vectors inside FP64 refinement) -------
    fcg_mx = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned", precision="mixed")
    d.solve_pressure_rhs(fcg_mx, None, download=False)
    d._ctx.call("cfdhip_synchronize")
    t2 = time.perf_counter()
    rfm = d.solve_pressure_rhs(fcg_mx, None, download=False)
    dtfm = time.perf_counter() - t2
    _check(rfm, bnorm, "mixed learned FCG")
    mixed_fcg = {"learned_fcg_mixed_s": dtfm, "learned_fcg_mixed_iterations": rfm.iterations,
                 "learned_fcg_mixed_refinements": rfm.restarts,
                 "learned_fcg_mixed_fallback": bool(rfm.fallback_used),
                 "learned_fcg_mixed_true_relative_residual": rfm.true_residual / bnorm}
    log(f"mixed learned FCG: {rfm.status_name} {rfm.iterations} it ({rfm.restarts} cycles) "
        f"{dtfm * 1e3:.1f} ms, true rel res {rfm.true_residual / bnorm:.3e}")

    # --- SIMPLE outer iterations at Re = 1000: Jacobi-PCG vs learned FCG p' solves ------
    # Both from rest for --simple-iters iterations (the same iterations, so
    # iter/s compare like for like); the learned run then continues to outer
    # iteration 10 and its next p' system (C2(iii): the operator of a developed
    # flow, not the d = 1 Laplacian) is solved by both methods.
    simple, c2iii = {}, {}
    if args.simple_iters > 0:
        runs = {}
        mom = cfd.LinearSolverSettings(1e-12, 1e-8, 2000, "BiCGSTAB")
        for label, ps in (("jacobi", cfd.LinearSolverSettings(0.0, 1e-8, 40000, "CG", "Jacobi",
                                                              check_interval=64)),
                          ("learned", fcg)):
            ds = cfd.GpuSimpleDiscretization.cavity(2, N, N, viscosity=1e-3)
            try:
                if label == "learned":
                    ds.set_learned_weights()
                ds._ctx.call("cfdhip_synchronize")
                t1 = time.perf_counter()
                res = ds.simple_run(args.simple_iters, pressure=ps, momentum=mom)
                dt = time.perf_counter() - t1
                if res["iterations"] != args.simple_iters:
                    raise RuntimeError(f"SIMPLE ({label} p') stopped: status {res['status']}")
                runs[label] = (res, dt, ds.download("P"), ds.download_velocity())
                log(f"SIMPLE ({label} p'): {res['iterations']} outer it in {dt:.2f} s, status "
                    f"{res['status']}, p' iterations {res['pressure_iterations']}")
                if label == "learned" and args.c2iii_outer > args.simple_iters:
                    more = args.c2iii_outer - args.simple_iters
                    t1 = time.perf_counter()
                    res2 = ds.simple_run(more, pressure=ps, momentum=mom)
                    dt2 = time.perf_counter() - t1
                    if res2["iterations"] != more:
                        raise RuntimeError(f"SIMPLE (learned p') stopped: status {res2['status']}")
                    from cfdapp_amd import learned as _learned
                    _learned.assemble_pprime(ds, 0)
                    c2iii = {"outer_iteration": args.c2iii_outer + 1,
                             "learned_simple_iter_per_s_to_it10": more / dt2,
                             **pprime_compare(cfd, log, ds,
                                              f"C2(iii) 2048^2 Re 1000 p' at outer iteration "
                                              f"{args.c2iii_outer + 1}")}
            finally:
                ds.close()
        (rj_, tj, pj, uj), (rl_, tl, pl, ul) = runs["jacobi"], runs["learned"]
        umax = max(float(np.max(np.abs(c))) for c in uj)
        simple = {"simple_iter_per_s": rj_["iterations"] / tj,
                  "simple_iterations": rj_["iterations"],
                  "simple_status": rj_["status"],
                  "simple_pressure_iterations": rj_["pressure_iterations"],
                  "simple_learned_iter_per_s": rl_["iterations"] / tl,
                  "simple_learned_pressure_iterations": rl_["pressure_iterations"],
                  "simple_learned_status": rl_["status"],
                  "simple_speedup_learned_vs_jacobi": tj / tl,
                  "simple_learned_vs_jacobi_p_rel_err_inf":
                      float(np.max(np.abs(pl - pj)) / np.max(np.abs(pj))),
                  "simple_learned_vs_jacobi_u_rel_err_inf":
                      max(float(np.max(np.abs(a - b))) for a, b in zip(ul, uj)) / umax}
        log(f"SIMPLE learned vs Jacobi p': {tj / tl:.1f}x iter/s; |dp|/|p| "
            f"{simple['simple_learned_vs_jacobi_p_rel_err_inf']:.2e}, |du|/|u| "
            f"{simple['simple_learned_vs_jacobi_u_rel_err_inf']:.2e}")

    c4 = {}
    if args.c4:
        c4 = c4_leg(cfd, log)

    c129 = {}
    if args.c129_iters > 0:
        c129 = cavity129_leg(cfd, log, args.c129_iters)

    piso = {}
    if args.piso_steps > 0:
        piso = piso_leg(cfd, log, args.piso_n, args.piso_steps)

    cube = {}
    if args.cube_iters > 0:
        cube = mixed3d_leg(cfd, log, args.cube, args.cube_iters)

    batched = {}
    if args.batch_cases > 0:
        batched = batch_leg(cfd, log, batch_cases(count=args.batch_cases))

    cpu = None
    if not args.no_cpu:
        per_outer = (simple["simple_pressure_iterations"] / simple["simple_iterations"]
                     if simple else float(rj.iterations))
        cpu = cpu_baseline(rj.iterations, per_outer)
        log(f"cpu baseline: {cpu['value']:.1f} s ({cpu['kind']}, {cpu['cores']} cores)")

    line = {
        "metric": "p' solve time to 1e-8 rel.res + SIMPLE iter/s: learned FCG vs Jacobi-PCG",
        "value": fcg_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D cavity 2048x2048: stand-alone p' solve to 1e-8 FP64 true "
                               "relative residual (seeded mean-free RHS), learned FCG vs "
                               "Jacobi-PCG",
                   "cells": n,
                   "learned_fcg_s": fcg_s, "learned_fcg_iterations": rf.iterations,
                   "learned_fcg_status": rf.status_name,
                   "learned_channels_stored": channels,
                   "learned_fcg_fallback": bool(rf.fallback_used),
                   "learned_fcg_true_relative_residual": rf.true_residual / bnorm,
                   "learned_weights": weights_kind,
                   "learned_precision": "FCG FP64 vectors; U-Net BF16-rounded weights, FP32 "
                                        "activations and accumulation (unused channels "
                                        "dropped: narrow FP32 kernels), FP32 coarse tail",
                   "preconditioner_us_per_application": per_apply_us,
                   "preconditioner_layers": layers,
                   "jacobi_pcg_s": jac_s, "jacobi_pcg_iterations": rj.iterations,
                   "jacobi_pcg_status": rj.status_name,
                   "jacobi_pcg_true_relative_residual": rj.true_residual / bnorm,
                   "speedup_learned_vs_jacobi": jac_s / fcg_s,
                   "plain_spmv": spmv,
                   **mixed, **mixed_fcg, **simple,
                   "c2iii_developed_flow_pprime": c2iii,
                   "c4_unseen_cavity_re3200": c4,
                   "cavity129_simple": c129,
                   "piso": piso,
                   "mixed_precision_3d": cube,
                   "batched_pprime": batched},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES,
                     "kernel": "cgIterKernel<double> (one Jacobi-PCG iteration: updates + "
                               "5-point SpMV + dots)",
                     "kernel_mean_us": spmv_us, "bytes_per_launch": spmv_bytes,
                     "learned_dominant_layer": {"layer": dom, **layers[dom],
                                                "bytes": lb[dom],
                                                "frac": layers[dom]["gbs"] / HBM_PEAK_GBS}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
