"""The fused BiCGSTAB of the 2D momentum stencil (small grids: three kernels
per iteration, the scalar steps run inside them) against the launched
BiCGSTAB on the same system exported as CSR (cfdhip_csr_solve, itself pinned
to the reference's BiCGSTAB in test_gpu_bicgstab.py) and against the
oracle's CPU BiCGSTAB (BiCGSTAB.cpp restated): status, iterations within one
of the launched form and two of the oracle (different reduction trees),
solutions within the tolerance bound, the same history length. Grids cover
every tile height the fused form picks (4, 8, 16 and 32 rows), a one-column
grid, a grid too large for it (the launched form runs there), the Jacobi
preconditioner, a MaxIterations stop and a zero right-hand side.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (nx, ny, tile rows the fused form picks, or 0 for the launched form)
SHAPES = [(129, 129, 4), (257, 257, 8), (2048, 64, 16), (600, 600, 32), (1, 9, 4), (5, 3, 4),
          (1100, 1100, 0)]


def log(msg):
    print(msg, flush=True)


def _system(nx, ny, seed):
    import cfdapp_amd as cfd
    d = cfd.GpuSimpleDiscretization.cavity(2, nx, ny, viscosity=0.01)
    rng = np.random.default_rng(seed)
    n = d.n_cells
    d.upload("U", rng.uniform(-0.3, 0.3, n))
    d.upload("V", rng.uniform(-0.3, 0.3, n))
    d.upload("P", rng.uniform(-1.0, 1.0, n))
    d.mass_flux()
    d.assemble_momentum_resident(0.7)
    return d


def _tiles(nx, ny):
    tx = (nx + 61) // 62
    for rows in (4, 8, 16, 32):
        if tx * ((ny + rows - 3) // (rows - 2)) <= 256:
            return rows
    return 0


@pytest.mark.parametrize("nx,ny,rows", SHAPES)
def test_fused_matches_launched_and_oracle(nx, ny, rows):
    import cfdapp_amd as cfd
    import oracle
    assert _tiles(nx, ny) == rows
    log(f"case: fused BiCGSTAB {nx}x{ny} (rows {rows})")
    d = _system(nx, ny, nx + ny)
    try:
        rp, col, val = d.momentum_csr()
        A = oracle.Csr(rp, col, val)
        for pre in (("None", "Jacobi") if nx == 129 else ("None",)):
            s = cfd.LinearSolverSettings(1e-12, 1e-9, 500, "BiCGSTAB", pre)
            for k in range(2):
                b = d.download("MOM_RHS_" + "UV"[k])
                x0 = d.download("UV"[k])
                r = d.solve_momentum_resident(k, s)
                x = d.download("UV"[k] + "STAR")
                lr = cfd.make_gpu_bicgstab(s).solve(rp, col, val, b, x0)
                log(f"  {pre} {'uv'[k]}: fused {r.status_name} {r.iterations} it "
                    f"restarts {r.restarts}, launched {lr.status_name} {lr.iterations} it")
                assert r.status == lr.status == cfd.CONVERGED
                assert abs(r.iterations - lr.iterations) <= 1
                assert len(r.residual_history) == r.iterations + 1
                scale = np.linalg.norm(lr.solution)
                assert np.linalg.norm(x - lr.solution) <= 1e-6 * scale
                if nx * ny <= 70000:
                    o = A.bicgstab(b, x0, abs_tol=1e-12, rel_tol=1e-9, max_iter=500,
                                   jacobi=pre == "Jacobi")
                    assert o.converged and abs(o.iterations - r.iterations) <= 2
                    assert np.linalg.norm(x - o.x) <= 1e-6 * scale
                true = np.linalg.norm(b - A.spmv(x))
                assert true <= max(1e-12, 1e-9 * r.initial_residual) * 10
    finally:
        d.close()


def test_fused_max_iterations_and_zero_rhs():
    import cfdapp_amd as cfd
    log("case: fused BiCGSTAB MaxIterations and zero rhs, 129x129")
    d = _system(129, 129, 9)
    try:
        s = cfd.LinearSolverSettings(0.0, 1e-14, 3, "BiCGSTAB")
        r = d.solve_momentum_resident(0, s)
        rp, col, val = d.momentum_csr()
        lr = cfd.make_gpu_bicgstab(s).solve(rp, col, val, d.download("MOM_RHS_U"),
                                            d.download("U"))
        assert r.status == lr.status == cfd.MAX_ITERATIONS
        assert r.iterations == lr.iterations == 3
        assert len(r.residual_history) == 4
        n = d.n_cells
        d.upload("U", np.zeros(n))
        d.upload("MOM_RHS_U", np.zeros(n))
        r = d.solve_momentum_resident(0, cfd.LinearSolverSettings(1e-12, 1e-9, 500, "BiCGSTAB"))
        assert r.status == cfd.CONVERGED and r.iterations == 0
        assert np.all(d.download("USTAR") == 0.0)
    finally:
        d.close()
