"""The no-allocation p' solve (cfdhip_pressure_solve_device_noalloc) against
the allocating one (cfdhip_pressure_solve_rhs) on the same resident systems.

Both run the same FP64 single-kernel Jacobi-PCG kernels on the same inputs,
so the no-allocation entry point must give bit-identical solutions, the
same status, iteration count and residuals. Two calls in a row on different
contexts of one GPU share the library's fixed scratch array; the second
system is smaller than the first, so state left by the first call in the
scratch must not leak into the second. Device buffers for b and x are the
contexts' own resident fields (PC_RHS and a free cell field), so no caller
allocation is needed either.
"""
import ctypes

import numpy as np
import pytest

from cfdapp_amd import _lib

pytestmark = pytest.mark.gpu

cfd = pytest.importorskip("cfdapp_amd")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not cfd.device_available():
        pytest.fail("no gfx950 device: the HIP path must run on the GPU")


def _system(nx, ny, seed):
    d = cfd.GpuSimpleDiscretization.cavity(2, nx, ny)
    rng = np.random.default_rng(seed)
    d.upload("RESPONSE", 0.5 + rng.random(d.n_cells))
    d.upload("FLUX_STAR", rng.uniform(-1.0, 1.0, d.n_faces))
    d.assemble_pressure_correction_resident(0)
    return d


def _noalloc(d, settings):
    r = _lib.SolveResult()
    s = settings.c()
    d._ctx.call("cfdhip_pressure_solve_device_noalloc", ctypes.byref(s),
                ctypes.c_void_p(d.field_device("PC_RHS")),
                ctypes.c_void_p(d.field_device("USTAR")), ctypes.byref(r))
    return r, d.download("USTAR")


def test_noalloc_matches_the_allocating_solve_on_two_calls():
    s = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi", check_interval=16)
    for k, (nx, ny) in enumerate(((700, 520), (130, 70))):
        print(f"case: no-alloc {nx}x{ny}", flush=True)
        d = _system(nx, ny, 900 + k)
        try:
            ref = d.solve_pressure_rhs(s, None, download=True)
            r, x = _noalloc(d, s)
            print(f"{nx}x{ny}: allocating {ref.status_name} {ref.iterations} it, no-alloc "
                  f"status {r.status} {r.iterations} it", flush=True)
            assert r.status == ref.status == cfd.CONVERGED
            assert r.iterations == ref.iterations
            assert r.initial_residual == ref.initial_residual
            assert r.final_residual == ref.final_residual
            assert r.true_residual == ref.true_residual
            assert r.history_length == 0
            assert np.array_equal(x, ref.solution)
        finally:
            d.close()


def test_noalloc_rejects_what_it_cannot_hold():
    print("case: no-alloc limits", flush=True)
    d = _system(40, 30, 5)
    try:
        with pytest.raises(cfd.CfdHipError) as e:
            _noalloc(d, cfd.LinearSolverSettings(0.0, 1e-8, 70000, "CG", "Jacobi"))
        assert e.value.code == _lib.ERR_INVALID
        with pytest.raises(cfd.CfdHipError) as e:
            _noalloc(d, cfd.LinearSolverSettings(0.0, 1e-8, 100, "BiCGSTAB"))
        assert e.value.code == _lib.ERR_UNSUPPORTED
    finally:
        d.close()
    big = cfd.GpuSimpleDiscretization.cavity(2, 1100, 1000)
    try:
        big.upload("RESPONSE", np.ones(big.n_cells))
        big.upload("FLUX_STAR", np.zeros(big.n_faces))
        big.assemble_pressure_correction_resident(0)
        with pytest.raises(cfd.CfdHipError) as e:
            _noalloc(big, cfd.LinearSolverSettings(0.0, 1e-8, 100, "CG", "Jacobi"))
        assert e.value.code == _lib.ERR_INVALID
    finally:
        big.close()
