"""HIP-graph replay of the Jacobi-PCG iteration chunks.

With settings.graph = 1 the single-kernel PCG captures one chunk of
iterations (iteration kernel + finalize, an even number so the buffer parity
returns) and replays it; iterations past the device verdict do nothing. The
replayed solve must be bitwise identical to the stream-launched one: same
kernels, same order, same reductions. The same holds for whole SIMPLE runs
whose p' solves replay their chunks (cfdhip_simple_run settings[8]).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("interval", [8, 7])
def test_graph_replay_bitwise(precision, interval):
    import cfdapp_amd as cfd
    d = cfd.GpuSimpleDiscretization.cavity(2, 96, 80)
    try:
        rng = np.random.default_rng(11)
        d.upload("RESPONSE", 0.5 + rng.random(d.n_cells))
        d.upload("FLUX_STAR", rng.uniform(-1, 1, d.n_faces))
        d.assemble_pressure_correction_resident(0)
        out = []
        for graph in (False, True):
            s = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi", precision=precision,
                                         check_interval=interval, graph=graph)
            r = d.solve_pressure_rhs(s, None, download=True)
            print(f"{precision} interval {interval} graph {graph}: {r.status_name} "
                  f"{r.iterations} it", flush=True)
            assert r.status == 0
            out.append(r)
        assert out[0].iterations == out[1].iterations
        assert np.array_equal(out[0].solution, out[1].solution)
        assert np.array_equal(out[0].residual_history, out[1].residual_history)
    finally:
        d.close()


def test_graph_replay_respects_max_iterations():
    """A cap that is not a multiple of the chunk: the replayed chunk runs past
    it on the device, where the finalize step has already ended the solve."""
    import cfdapp_amd as cfd
    d = cfd.GpuSimpleDiscretization.cavity(2, 64, 64)
    try:
        d.upload("RESPONSE", np.ones(d.n_cells))
        d.upload("FLUX_STAR", np.random.default_rng(3).uniform(-1, 1, d.n_faces))
        d.assemble_pressure_correction_resident(0)
        res = []
        for graph in (False, True):
            s = cfd.LinearSolverSettings(0.0, 1e-12, 37, "CG", "Jacobi", check_interval=16,
                                         graph=graph)
            res.append(d.solve_pressure_rhs(s, None, download=True))
        for r in res:
            assert r.status_name == "MaxIterations" and r.iterations == 37
        assert np.array_equal(res[0].solution, res[1].solution)
    finally:
        d.close()


def test_simple_run_with_graph_matches():
    import cfdapp_amd as cfd
    hs = []
    for graph in (False, True):
        d = cfd.GpuSimpleDiscretization.cavity(2, 40, 40, viscosity=0.01)
        try:
            res = d.simple_run(12, pressure=cfd.LinearSolverSettings(
                0.0, 1e-8, 20000, "CG", "Jacobi", check_interval=16), use_graph=graph)
            hs.append((res, d.download("P"), d.download("U")))
        finally:
            d.close()
    (r0, p0, u0), (r1, p1, u1) = hs
    assert r0["iterations"] == r1["iterations"] == 12
    assert r0["pressure_iterations"] == r1["pressure_iterations"]
    for k in ("u", "v", "p", "c"):
        assert np.array_equal(r0["hist"][k], r1["hist"][k])
    assert np.array_equal(p0, p1) and np.array_equal(u0, u1)
