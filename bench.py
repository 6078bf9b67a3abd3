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
together by the batched mixed-precision Jacobi-PCG (aggregate cells x
iterations per second, per-case p50 / p99 completion time) and one after
another by the single-case solve; the reference's own 129^2 cavity (Re 100)
SIMPLE iterations per second with stream-launched and graph-replayed p'
iteration chunks; the PISO leg: time steps of a 1024^2 cavity at Re 1000 with
the reference's BiCGSTAB p' solves and with the learned FCG; and the 3D leg: SIMPLE iterations on the
256^3 cavity at Re = 400 with FP64 and with mixed-precision p' solves, and the
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
TRAFFIC_BYTES = 451.0e6


def layer_bytes(p0):
    """Algorithmic HBM bytes of each preconditioner layer on the solve path
    (FP32 activations, each tensor read and written once; p0 = fine pixels)."""
    p1, p2, p3 = p0 / 4, p0 / 16, p0 / 64
    f = 4.0
    return {
        "e0b": p0 * (8 + 4) + f * (16 * p0 + 16 * p1),   # r (FP64) + D^-1/2 -> e0, p0
        "e1a": f * (16 * p1 + 32 * p1),
        "e1b": f * (32 * p1 + 32 * p1 + 32 * p2),
        "e2a": f * (32 * p2 + 64 * p2),
        "e2b": f * (64 * p2 + 64 * p2 + 64 * p3),
        "b3a": f * (64 * p3 + 64 * p3),
        "b3b": f * (64 * p3 + 64 * p3),
        "d2a": f * (64 * p3 + 64 * p2 + 64 * p2),
        "d2b": f * (64 * p2 + 64 * p2),
        "d1a": f * (64 * p2 + 32 * p1 + 32 * p1),
        "d1b": f * (32 * p1 + 32 * p1),
        "d0a": f * (32 * p1 + 16 * p0 + 16 * p0),
        "d0b": f * (16 * p0 + 16 * p0),
        "out": f * 16 * p0 + p0 * (4 + 8 + 8),         # d0, D^-1/2, q -> z (FP64)
        "e0a": 0.0,
        # coarse tail: one channel of b3a in, one out, and its levels (H/16 ..)
        # each read and written a few times
        "tail": f * (2 * p3 + 6 * p3 / 3),
    }


def cpu_baseline(iters_gpu):
    """The reference's own CG + Jacobi (compiled from its sources, OpenMP SpMV),
    bounded sample of 40 iterations at 2048x2048, scaled to the GPU's iteration
    count; falls back to the oracle restatement when the reference build is absent."""
    driver = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    cores = min(os.cpu_count() or 1, 16)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    k = 40
    if os.path.exists(driver):
        out = subprocess.run([driver, "cgtime", "dim=2", f"nx={N}", f"ny={N}", f"seed={SEED}",
                              f"pmax={k}"], capture_output=True, text=True, env=env,
                             timeout=600, check=True).stdout
        rec = json.loads(out.strip().splitlines()[-1])
        per_iter = rec["seconds"] / max(rec["iterations"], 1)
        kind = "reference"
    else:
        import oracle
        c = oracle.Case.cavity(2, N, N)
        ones = np.ones(c.n_cells)
        A, _, _ = c.pressure_system(np.zeros(c.n_faces), ones, ones)
        b = standalone_rhs(c.n_cells)
        t0 = time.perf_counter()
        r = A.cg(b, max_iter=k)
        per_iter = (time.perf_counter() - t0) / max(r.iterations, 1)
        kind = "port"
    return {"value": per_iter * iters_gpu, "unit": "s", "cores": cores, "kind": kind,
            "sample": f"{k} CG+Jacobi iterations on the 2048x2048 p' system "
                      f"({per_iter * 1e3:.1f} ms/iteration), scaled to the GPU solve's "
                      f"{iters_gpu} iterations"}


def batch_leg(cfd, log, cases, use_graph=True, sequential=True):
    """p' systems of the first SIMPLE iteration of each case, solved (a) by the
    batched mixed-precision Jacobi-PCG in one launch per step and (b) one case
    after another by the single-case mixed-precision solve, both to 1e-8 FP64
    true relative residual."""
    from cfdapp_amd import learned
    t0 = time.perf_counter()
    batch = cfd.BatchSolver()
    seq_s, seq_lat, seq_iters = 0.0, [], 0
    mx = cfd.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi", precision="mixed",
                                  check_interval=64)
    cells = 0
    for kind, nx, ny, re in cases:
        d = learned.make_case(kind, nx, ny=ny, re=re, speed=1.0)
        try:
            learned.assemble_pprime(d, 0)
            batch.add_pressure(d)
            cells += nx * ny
            if not sequential:
                continue
            d._ctx.call("cfdhip_synchronize")
            t1 = time.perf_counter()
            r = d.solve_pressure_rhs(mx, None, download=False)
            dt = time.perf_counter() - t1
            if r.status != 0:
                raise RuntimeError(f"sequential {kind} {nx}x{ny}: {r.status_name}")
            seq_s += dt
            seq_lat.append(dt * 1e3)
            seq_iters += r.iterations
        finally:
            d.close()
    log(f"batch: {len(cases)} cases, {cells} cells, set up in {time.perf_counter() - t0:.1f} s")
    bs = cfd.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi", precision="mixed",
                                  check_interval=32)
    batch.solve(bs, use_graph=use_graph)  # warm-up (also tunes the tile height)
    t1 = time.perf_counter()
    res = batch.solve(bs, use_graph=use_graph)
    wall = time.perf_counter() - t1
    st = batch.stats
    bad = [k for k, r in enumerate(res) if r.status != 0]
    if bad:
        raise RuntimeError(f"batched solve: {len(bad)} cases did not converge")
    lat = np.array([r.device_ms for r in res])
    out = {"cases": len(cases), "cells": cells, "batched_s": wall,
           "batched_device_ms": st["device_ms"],
           "batched_cell_iter_per_s": st["cell_iterations"] / wall,
           "batched_p50_ms": float(np.percentile(lat, 50)),
           "batched_p99_ms": float(np.percentile(lat, 99)),
           "batched_tile_rows": st["tile_rows"], "batched_steps": st["steps"],
           "batched_graph_captures": st["graph_captures"],
           "tile_tuning_us": st["tune_us"],
           }
    if sequential:
        out.update({"sequential_s": seq_s,
                    "sequential_cell_iter_per_s": sum(r.iterations * nx * ny for r, (_, nx, ny, _)
                                                      in zip(res, cases)) / seq_s,
                    "sequential_p50_ms": float(np.percentile(seq_lat, 50)),
                    "sequential_p99_ms": float(np.percentile(seq_lat, 99)),
                    "speedup_batched_vs_sequential": seq_s / wall})
    log(f"batch: batched {wall * 1e3:.1f} ms ({out['batched_cell_iter_per_s']:.3e} cell-it/s, "
        f"p50 {out['batched_p50_ms']:.1f} ms, p99 {out['batched_p99_ms']:.1f} ms, "
        f"{st['tile_rows']}-row tiles, {st['graph_captures']} graph captures)" +
        (f"; sequential {seq_s * 1e3:.1f} ms (p50 {out['sequential_p50_ms']:.1f} ms, "
         f"p99 {out['sequential_p99_ms']:.1f} ms)" if sequential else ""))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--simple-iters", type=int, default=2)
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
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched-solve leg (for kernel profiles)")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="batched leg with stream launches instead of graph replay")
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
    d.set_learned_weights()
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
        r = rs[-1]
        log(f"{label}: {args.steps} solves, {sec * 1e3:.1f} ms/solve, {r.status_name}, "
            f"{r.iterations} it, true rel res {r.true_residual / bnorm:.3e}, "
            f"fallback {r.fallback_used}")
        return sec, rs

    if args.jacobi_only:
        timed(pcg, "Jacobi-PCG")
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
    lb = layer_bytes(p0)
    layers = {}
    for i, nm in enumerate(names):
        if layer_calls[i] > 0:
            us = 1e3 * layer_ms[i] / layer_calls[i]
            layers[nm] = {"us": us, "gbs": lb[nm] / (us * 1e-6) / 1e9}
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
    mixed_fcg = {"learned_fcg_mixed_s": dtfm, "learned_fcg_mixed_iterations": rfm.iterations,
                 "learned_fcg_mixed_refinements": rfm.restarts,
                 "learned_fcg_mixed_fallback": bool(rfm.fallback_used),
                 "learned_fcg_mixed_true_relative_residual": rfm.true_residual / bnorm}
    log(f"mixed learned FCG: {rfm.status_name} {rfm.iterations} it ({rfm.restarts} cycles) "
        f"{dtfm * 1e3:.1f} ms, true rel res {rfm.true_residual / bnorm:.3e}")

    # --- SIMPLE outer iterations at Re = 1000: Jacobi-PCG vs learned FCG p' solves ------
    simple = {}
    if args.simple_iters > 0:
        runs = {}
        for label, ps in (("jacobi", cfd.LinearSolverSettings(0.0, 1e-8, 40000, "CG", "Jacobi",
                                                              check_interval=64)),
                          ("learned", cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG",
                                                               "Learned"))):
            ds = cfd.GpuSimpleDiscretization.cavity(2, N, N, viscosity=1e-3)
            try:
                if label == "learned":
                    ds.set_learned_weights()
                ds._ctx.call("cfdhip_synchronize")
                t1 = time.perf_counter()
                res = ds.simple_run(args.simple_iters, pressure=ps)
                dt = time.perf_counter() - t1
                runs[label] = (res, dt, ds.download("P"), ds.download_velocity())
            finally:
                ds.close()
            log(f"SIMPLE ({label} p'): {res['iterations']} outer it in {dt:.2f} s, status "
                f"{res['status']}, p' iterations {res['pressure_iterations']}")
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
        cpu = cpu_baseline(rj.iterations)
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
                   "learned_fcg_fallback": bool(rf.fallback_used),
                   "learned_fcg_true_relative_residual": rf.true_residual / bnorm,
                   "learned_weights": weights_kind,
                   "learned_precision": "FCG FP64 vectors; U-Net BF16 weights, FP32 activations "
                                        "(hi/lo BF16 pairs at the MFMA), FP32 coarse tail",
                   "preconditioner_us_per_application": per_apply_us,
                   "preconditioner_layers": layers,
                   "jacobi_pcg_s": jac_s, "jacobi_pcg_iterations": rj.iterations,
                   "jacobi_pcg_true_relative_residual": rj.true_residual / bnorm,
                   "speedup_learned_vs_jacobi": jac_s / fcg_s,
                   "plain_spmv": spmv,
                   **mixed, **mixed_fcg, **simple,
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
