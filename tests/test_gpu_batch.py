"""Batched p' solves on the GPU against the oracle.

Each case's operator and right-hand side come from the oracle's p' assembly
(PressureCorrectionEquation.cpp:181-339 restated) on cavities (cell 0
pinned) and channels (FixedValue outlet, no pin) of uneven size with random
response coefficients and predicted fluxes. Every batched solution must meet
the FP64 true-residual test |b - A x| <= 1e-8 |b| with A applied by the
oracle, and agree with the oracle's own CG + Jacobi solution to the
accuracy that tolerance allows (relative error bounded by cond(A) * 1e-8;
the check uses 1e-4, far above what these small systems reach).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [("cavity", 20, 12), ("channel", 37, 50), ("cavity", 64, 64), ("channel", 130, 33),
          ("cavity", 1, 9), ("cavity", 70, 141), ("channel", 8, 8), ("cavity", 5, 3)]


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
    return c, A, rhs, coeff, pin


@pytest.fixture(scope="module")
def systems():
    out = []
    for k, (kind, nx, ny) in enumerate(SHAPES):
        c, A, rhs, coeff, pin = _case(kind, nx, ny, 500 + k)
        out.append((nx, ny, pin, A, rhs, coeff))
    return out


def _batch(systems, zero_rhs=None):
    import cfdapp_amd as cfd
    b = cfd.BatchSolver()
    for k, (nx, ny, pin, A, rhs, coeff) in enumerate(systems):
        r = np.zeros_like(rhs) if k == zero_rhs else rhs
        assert b.add_system(nx, ny, pin, A.diagonal(), coeff, r) == k
    return b


def _check(systems, b, res, tol=1e-8, zero_rhs=None):
    for k, (nx, ny, pin, A, rhs, coeff) in enumerate(systems):
        print(f"case {k}: {nx}x{ny} pin {pin}: {res[k].status_name} {res[k].iterations} it",
              flush=True)
        assert res[k].status == 0, (k, res[k].status_name)
        x = b.solution(k)
        if k == zero_rhs:
            assert np.all(x == 0.0)
            continue
        r = rhs - A.spmv(x)
        assert np.linalg.norm(r) <= tol * np.linalg.norm(rhs) * (1 + 1e-9), k
        xo = A.cg(rhs, rel_tol=1e-12, max_iter=200000).x
        assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo), k


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_meets_fp64_tolerance(systems, precision):
    import cfdapp_amd as cfd
    b = _batch(systems, zero_rhs=7)
    s = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi", precision=precision,
                                 check_interval=8)
    res = b.solve(s)
    _check(systems, b, res, zero_rhs=7)
    st = b.stats
    assert st["tile_rows"] in (16, 32, 64)
    assert st["cell_iterations"] == sum(r.iterations * nx * ny
                                        for r, (nx, ny, *_) in zip(res, systems))
    lat = [r.device_ms for r in res]
    assert all(0.0 <= t <= st["device_ms"] + 1e-6 for t in lat)
    b.close()


@pytest.mark.parametrize("rows", [16, 32, 64])
def test_batch_tile_heights(systems, rows):
    import cfdapp_amd as cfd
    b = _batch(systems)
    s = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi", precision="mixed")
    res = b.solve(s, tile_rows=rows)
    assert b.stats["tile_rows"] == rows
    _check(systems, b, res)
    b.close()


def test_batch_graph_replay_matches_stream_launches(systems):
    """The captured chunks replay the same kernels: bitwise equal solutions."""
    import cfdapp_amd as cfd
    s = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi", precision="mixed",
                                 check_interval=16)
    xs = []
    for graph in (False, True):
        b = _batch(systems)
        res = b.solve(s, tile_rows=32, use_graph=graph)
        assert all(r.status == 0 for r in res)
        if graph:
            assert b.stats["graph_captures"] >= 1
        xs.append([b.solution(k) for k in range(len(systems))])
        b.close()
    for a, c in zip(*xs):
        assert np.array_equal(a, c)


def test_batch_from_resident_assembly():
    """add_pressure copies a context's assembled p' system: same answer as the
    single-case Jacobi-PCG solve on that context."""
    import cfdapp_amd as cfd
    b = cfd.BatchSolver()
    ds = []
    for nx, ny in ((48, 40), (96, 64)):
        d = cfd.GpuSimpleDiscretization.cavity(2, nx, ny)
        rng = np.random.default_rng(nx)
        d.upload("RESPONSE", 0.5 + rng.random(d.n_cells))
        d.upload("FLUX_STAR", rng.uniform(-1, 1, d.n_faces))
        d.assemble_pressure_correction_resident(0)
        b.add_pressure(d)
        ds.append(d)
    s = cfd.LinearSolverSettings(0.0, 1e-10, 20000, "CG", "Jacobi", precision="fp64")
    res = b.solve(s)
    for k, d in enumerate(ds):
        assert res[k].status == 0
        single = d.solve_pressure_rhs(s, None, download=True)
        assert single.status == 0
        x = b.solution(k)
        assert np.linalg.norm(x - single.solution) <= 1e-7 * np.linalg.norm(single.solution)
        d.close()
    b.close()
