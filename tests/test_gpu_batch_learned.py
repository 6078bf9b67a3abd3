"""Batched flexible CG with the learned preconditioner (cfdhip_batch_set_learned
+ cfdhip_batch_solve with FCG / Learned) on the GPU.

No reference counterpart: the batch is checked against the FP64
true-residual test with A applied on the host (the oracle's p' assembly for
the small cases, the workload's host operator for the 256-case batch), the
oracle's CG solution, and the single-case learned FCG on a context, which
runs the same iteration one case at a time: the same iteration count to
within two (different reduction trees) and solutions within the tolerance
bound. Cases: cavities (cell 0 pinned) and channels (a FixedValue outlet,
the network's odd reflection on that side, found from the system) of uneven
size, down to a 1 x 9 grid, and a zero right-hand side.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [("cavity", 20, 12), ("channel", 37, 50), ("cavity", 64, 64), ("channel", 130, 33),
          ("cavity", 1, 9), ("cavity", 70, 141), ("channel", 8, 8), ("cavity", 5, 3),
          ("channel", 200, 96), ("cavity", 257, 129)]


def log(msg):
    print(msg, flush=True)


def _case(kind, nx, ny, seed):
    import oracle
    c = oracle.Case(2, nx, ny, 1, nx / max(nx, ny), ny / max(nx, ny))
    if kind == "cavity":
        for p in range(4):
            c.set_bc(p, oracle.MOVING_WALL if p == 3 else oracle.WALL, (1.0, 0, 0))
        pin = 0
    else:
        c.set_bc(0, oracle.INLET, (1.0, 0, 0), oracle.FIXED_GRADIENT)
        c.set_bc(1, oracle.OUTLET, (0, 0, 0), oracle.FIXED_VALUE, 0.0)
        c.set_bc(2, oracle.WALL)
        c.set_bc(3, oracle.WALL)
        pin = -1
    rng = np.random.default_rng(seed)
    du = 0.5 + rng.random(c.n_cells)
    dv = 0.5 + rng.random(c.n_cells)
    F = rng.uniform(-1.0, 1.0, c.n_faces)
    A, rhs, coeff = c.pressure_system(F, du, dv, ref=0)
    return A, rhs, coeff, pin


@pytest.fixture(scope="module")
def systems():
    return [(nx, ny) + _case(kind, nx, ny, 700 + k) for k, (kind, nx, ny) in enumerate(SHAPES)]


def _fcg(**kw):
    import cfdapp_amd as cfd
    return cfd.LinearSolverSettings(0.0, 1e-8, 2000, "FCG", "Learned", check_interval=4, **kw)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_learned_meets_fp64_tolerance(systems, use_graph):
    import cfdapp_amd as cfd
    b = cfd.BatchSolver()
    try:
        b.set_learned()
        zero = 7
        for k, (nx, ny, A, rhs, coeff, pin) in enumerate(systems):
            r = np.zeros_like(rhs) if k == zero else rhs
            assert b.add_system(nx, ny, pin, A.diagonal(), coeff, r) == k
        log(f"case: batched learned FCG, {len(systems)} systems, graph {use_graph}")
        res = b.solve(_fcg(), use_graph=use_graph)
        for k, (nx, ny, A, rhs, coeff, pin) in enumerate(systems):
            log(f"case {k}: {nx}x{ny} pin {pin}: {res[k].status_name} {res[k].iterations} it "
                f"fallback {res[k].fallback_used}")
            assert res[k].status == cfd.CONVERGED, (k, res[k].status_name)
            assert not res[k].fallback_used, k
            x = b.solution(k)
            if k == zero:
                assert np.all(x == 0.0)
                continue
            true = np.linalg.norm(rhs - A.spmv(x))
            assert true <= 1e-8 * np.linalg.norm(rhs) * (1 + 1e-9), (k, true)
            xo = A.cg(rhs, rel_tol=1e-12, max_iter=200000).x
            assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo), k
            # a learned preconditioner: few iterations whatever the size
            assert res[k].iterations <= 60, (k, res[k].iterations)
        assert b.stats["steps"] > 0
    finally:
        b.close()


def test_batched_learned_matches_single_case_solves():
    """The same systems solved one at a time by the context's learned FCG
    (read from assembled cavities and channels, so the Dirichlet sides come
    from the context in both)."""
    import cfdapp_amd as cfd
    from cfdapp_amd import learned
    shapes = [("cavity", 96, 64, 400.0), ("channel", 128, 40, 800.0), ("cavity", 33, 71, 150.0)]
    b = cfd.BatchSolver()
    single = []
    try:
        b.set_learned()
        for kind, nx, ny, re in shapes:
            d = learned.make_case(kind, nx, ny=ny, re=re, speed=1.0)
            try:
                learned.assemble_pprime(d, 1)
                b.add_pressure(d)
                d.set_learned_weights()
                r = d.solve_pressure_rhs(_fcg(), None)
                single.append(r)
            finally:
                d.close()
        log("case: batched vs single-case learned FCG")
        res = b.solve(_fcg())
        for k, (r1, rb) in enumerate(zip(single, res)):
            log(f"case {k}: single {r1.iterations} it, batched {rb.iterations} it")
            assert r1.status == cfd.CONVERGED and rb.status == cfd.CONVERGED
            assert abs(r1.iterations - rb.iterations) <= 2, k
            x = b.solution(k)
            scale = np.linalg.norm(r1.solution)
            assert np.linalg.norm(x - r1.solution) <= 1e-6 * scale, k
    finally:
        b.close()


def test_batched_learned_256_uneven_cases():
    """The 256-case workload (seed 5005): every case at the FP64 target with A
    applied on the host, no fallback, and the batched learned solve faster than
    the batched Jacobi-PCG on the same systems."""
    import cfdapp_amd as cfd
    from cfdapp_amd import learned
    from cfdapp_amd import workloads as wl
    log("case: batched learned FCG, 256 uneven cases, setup")
    t0 = time.perf_counter()
    b = cfd.BatchSolver()
    systems = []
    try:
        b.set_learned()
        for kind, nx, ny, re in wl.batch_cases():
            dc = learned.make_case(kind, nx, ny=ny, re=re, speed=1.0)
            try:
                learned.assemble_pprime(dc, 0)
                b.add_pressure(dc)
                systems.append((nx, ny, dc.pin, dc.download("PC_DIAG"),
                                dc.download("PC_FACE_COEFF"), dc.download("PC_RHS")))
            finally:
                dc.close()
        log(f"setup {time.perf_counter() - t0:.1f} s")
        res = b.solve(cfd.LinearSolverSettings(0.0, 1e-8, 2000, "FCG", "Learned",
                                               check_interval=4))
        st = b.stats
        lat = np.array([r.device_ms for r in res])
        its = np.array([r.iterations for r in res])
        log(f"batched learned: {st['device_ms']:.1f} ms, {st['steps']} rounds, iterations "
            f"{its.min()}..{its.max()}, p50 {np.percentile(lat, 50):.1f} ms, p99 "
            f"{np.percentile(lat, 99):.1f} ms")
        worst = 0.0
        for k, ((nx, ny, pin, diag, fc, rhs), r) in enumerate(zip(systems, res)):
            assert r.status == cfd.CONVERGED, (k, r.status_name)
            assert not r.fallback_used, k
            x = b.solution(k)
            bn = float(np.linalg.norm(rhs))
            true = float(np.linalg.norm(rhs - wl.apply_pprime_2d(nx, ny, pin, diag, fc, x)))
            worst = max(worst, true / bn if bn > 0 else true)
            assert true <= 1e-8 * bn, (k, nx, ny, true / bn)
        log(f"worst host true rel res {worst:.3e}")
        learned_ms = st["device_ms"]
        res_j = b.solve(cfd.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi",
                                                 precision="mixed", check_interval=32))
        assert all(r.status == cfd.CONVERGED for r in res_j)
        log(f"batched Jacobi-PCG (mixed): {b.stats['device_ms']:.1f} ms")
        assert learned_ms < b.stats["device_ms"]
    finally:
        b.close()
