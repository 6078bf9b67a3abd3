"""The measured workloads at full size on the device, each result checked on the host.

C1  129^2 cavity, Re 100, Rhie-Chow, CG + Jacobi p' at 1e-8, BiCGSTAB
    momentum: 20 SIMPLE iterations against the compiled reference's run
    (tests/golden/simple2d_129_rc_20it.npz): histories, linear-iteration
    totals and fields.
C2  2048^2 stand-alone p' solve (d = 1, seeded mean-free RHS, cell 0
    pinned): the device operator bitwise against the oracle's CSR, then FP64
    Jacobi-PCG, mixed-precision Jacobi-PCG (its FP32 iteration runs on the
    64-row tiles at this size), learned FCG with FP64 and with FP32 Krylov
    vectors; every solution's FP64 residual is recomputed on the host with the
    oracle's operator and must meet 1e-8 |b|.
C3  256^3 cavity p' operator (d = 1): bitwise against its closed form
    (D_f = h, diagonal h x neighbour count, pinned row the identity; pinned
    against the oracle in test_workloads.py), FP64 and mixed-precision solves
    with the residual recomputed on the host from the closed form.
C4  the 2048^2 cavity at Re 3200 (mu = 3.125e-4): after three SIMPLE
    iterations, the assembled p' system (bitwise against the oracle's
    assembly of the same flux and response coefficients) solved by learned
    FCG and by Jacobi-PCG; both residuals recomputed on the host; the learned
    solve needs fewer iterations and less device time, without fallback.
C5  the 256 uneven cases (seed 5005), first SIMPLE iteration's p' systems,
    batched mixed-precision solve; every case's residual recomputed on the
    host.
"""
import time

import numpy as np
import pytest

import oracle
from cfdapp_amd import workloads as wl
from tests.conftest import load_golden
from tests.test_workloads import apply_closed_3d, closed_form_2d, closed_form_3d

pytestmark = pytest.mark.gpu

cfd = pytest.importorskip("cfdapp_amd")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not cfd.device_available():
        pytest.fail("no gfx950 device: the HIP path must run on the GPU")


def log(*a):
    print(*a, flush=True)


def rel_diff(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# ---------------------------------------------------------------------------- C1

def test_c1_cavity129_simple_matches_reference():
    log("case: C1 129^2 Re 100, 20 SIMPLE iterations")
    g = load_golden("simple2d_129_rc_20it")
    d = cfd.GpuSimpleDiscretization.cavity(2, 129, 129, viscosity=0.01)
    try:
        res = d.simple_run(20, rhie_chow=True,
                           momentum=cfd.LinearSolverSettings(1e-10, 1e-8, 500, "BiCGSTAB"),
                           pressure=cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi"))
        st, it, mom_it, p_it = (int(v) for v in g["meta"])
        assert (res["status"], res["iterations"]) == (st, it)
        assert abs(res["momentum_iterations"] - mom_it) <= 0.02 * mom_it
        assert abs(res["pressure_iterations"] - p_it) <= 0.02 * p_it
        devs = {}
        for h in ("u", "v", "p"):
            devs[h] = float(np.max(np.abs(res["hist"][h] - g["hist_" + h]) /
                                   np.maximum(np.abs(g["hist_" + h]), 1e-300)))
        # continuity RMS after correction sits at the p' solve's residual level
        devs["c"] = float(np.max(np.abs(np.log(res["hist"]["c"] / g["hist_c"]))))
        u, v = d.download_velocity()
        p = d.download_pressure()
        fields = {"u": rel_diff(u, g["u_x"]), "v": rel_diff(v, g["u_y"]), "p": rel_diff(p, g["p"]),
                  "F": rel_diff(d.download_mass_flux(), g["F"])}
        log("C1 deviations: history", devs, "fields", fields,
            f"iterations mom {res['momentum_iterations']}/{mom_it} p' "
            f"{res['pressure_iterations']}/{p_it}")
        # hist_v[0] is exactly 0 (v = 0 from rest) on both
        assert res["hist"]["v"][0] == g["hist_v"][0] == 0.0
        for h in ("u", "v", "p"):
            assert devs[h] <= 1e-6, (h, devs[h])
        assert devs["c"] <= np.log(4.0)
        for k, e in fields.items():
            assert e <= 1e-6, (k, e)
    finally:
        d.close()


# ---------------------------------------------------------------------------- C2

N2 = 2048


@pytest.fixture(scope="module")
def standalone2048():
    d = cfd.GpuSimpleDiscretization.cavity(2, N2, N2)
    d.upload("RESPONSE", np.ones(d.n_cells))
    d.upload("FLUX_STAR", np.zeros(d.n_faces))
    d.assemble_pressure_correction_resident(0)
    b = wl.standalone_rhs(d.n_cells)
    c = oracle.Case.cavity(2, N2, N2)
    A, _, coeff = c.pressure_system(np.zeros(c.n_faces), np.ones(c.n_cells), np.ones(c.n_cells))
    yield d, b, A, coeff
    d.close()


def test_c2_operator_bitwise_at_2048(standalone2048):
    log("case: C2 2048^2 p' operator")
    d, b, A, coeff = standalone2048
    rp, col, val = d.pressure_csr()
    assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col)
    assert np.array_equal(val.view(np.int64), A.val.view(np.int64))
    assert len(val) == 20963326  # p' nnz at 2048^2 (cell 0 pinned)
    diag, fc = closed_form_2d(N2, N2)
    assert np.array_equal(d.download("PC_DIAG"), diag)
    assert np.array_equal(d.download("PC_FACE_COEFF"), fc)
    assert np.array_equal(coeff, fc)


C2_METHODS = {
    "jacobi_fp64": cfd.LinearSolverSettings(0.0, 1e-8, 40000, "CG", "Jacobi", check_interval=64),
    "jacobi_mixed": cfd.LinearSolverSettings(0.0, 1e-8, 60000, "CG", "Jacobi", precision="mixed",
                                             check_interval=64),
    "learned_fp64": cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned"),
    "learned_mixed": cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned",
                                              precision="mixed"),
}


def test_c2_standalone_solves_meet_the_fp64_target(standalone2048):
    d, b, A, _ = standalone2048
    d.set_learned_weights()
    bnorm = float(np.linalg.norm(b))
    sols = {}
    for name, s in C2_METHODS.items():
        log(f"case: C2 2048^2 {name}")
        r = d.solve_pressure_rhs(s, b)
        true = float(np.linalg.norm(b - A.spmv(r.solution)))
        log(f"C2 {name}: {r.status_name} {r.iterations} it ({r.restarts} cycles) "
            f"{r.device_ms:.1f} ms, host true rel res {true / bnorm:.3e}, fallback "
            f"{r.fallback_used}")
        if name == "jacobi_mixed":
            log("C2 mixed FP64 residuals at the replacements: " +
                " ".join(f"{v / bnorm:.2e}" for v in r.residual_history[:40]))
        assert r.converged, (name, r.status_name)
        assert true <= 1e-8 * bnorm, (name, true / bnorm)
        assert r.true_residual == pytest.approx(true, rel=1e-6, abs=1e-12 * bnorm)
        if name.startswith("learned"):
            assert not r.fallback_used, name
        sols[name] = r.solution
    # |x_a - x_b| <= |A^-1| (|r_a| + |r_b|): the solutions agree to what the
    # tolerance allows on this operator (measured: far inside the bound used)
    ref = sols["jacobi_fp64"]
    for name, x in sols.items():
        e = rel_diff(x, ref)
        log(f"C2 {name} vs jacobi_fp64: max rel diff {e:.2e}")
        assert e <= 1e-4, (name, e)


# ---------------------------------------------------------------------------- C3

N3 = 256


def test_c3_cube256_fp64_and_mixed():
    log("case: C3 256^3 p' operator and solves")
    d = cfd.GpuSimpleDiscretization.cavity(3, N3, N3, N3)
    try:
        d.upload("RESPONSE", np.ones(d.n_cells))
        d.upload("FLUX_STAR", np.zeros(d.n_faces))
        d.assemble_pressure_correction_resident(0)
        diag, fc = closed_form_3d(N3, N3, N3)
        assert np.array_equal(d.download("PC_DIAG"), diag)
        assert np.array_equal(d.download("PC_FACE_COEFF"), fc)
        del diag, fc
        b = wl.standalone_rhs(d.n_cells)
        bnorm = float(np.linalg.norm(b))
        sols = {}
        for prec in ("fp64", "mixed"):
            log(f"case: C3 256^3 {prec}")
            s = cfd.LinearSolverSettings(0.0, 1e-8, 60000, "CG", "Jacobi", precision=prec,
                                         check_interval=64)
            r = d.solve_pressure_rhs(s, b)
            true = float(np.linalg.norm(b - apply_closed_3d(N3, N3, N3, r.solution)))
            log(f"C3 {prec}: {r.status_name} {r.iterations} it ({r.restarts} cycles) "
                f"{r.device_ms:.1f} ms, host true rel res {true / bnorm:.3e}, fallback "
                f"{r.fallback_used}")
            if prec == "mixed":
                log("C3 mixed FP64 residuals at the replacements: " +
                    " ".join(f"{v / bnorm:.2e}" for v in r.residual_history[:40]))
            assert r.converged and true <= 1e-8 * bnorm, (prec, r.status_name, true / bnorm)
            sols[prec] = r.solution
        e = rel_diff(sols["mixed"], sols["fp64"])
        log(f"C3 mixed vs fp64: max rel diff {e:.2e}")
        assert e <= 1e-4
    finally:
        d.close()


# ---------------------------------------------------------------------------- C4

def test_c4_learned_fcg_on_unseen_cavity_re3200():
    log("case: C4 2048^2 cavity Re 3200")
    d = cfd.GpuSimpleDiscretization.cavity(2, N2, N2, viscosity=1.0 / 3200.0)
    try:
        d.set_learned_weights()
        fcg = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "FCG", "Learned")
        jac = cfd.LinearSolverSettings(0.0, 1e-8, 60000, "CG", "Jacobi", check_interval=64)
        res = d.simple_run(3, pressure=fcg)
        assert res["iterations"] == 3 and res["status"] == 1, res["status"]
        # the p' system of the third iteration, resident: bitwise vs the oracle
        Fs, du = d.download("FLUX_STAR"), d.download("RESPONSE")
        c = oracle.Case.cavity(2, N2, N2)
        A, rhs, _ = c.pressure_system(Fs, du, du)
        rp, col, val = d.pressure_csr()
        assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col)
        assert np.array_equal(val.view(np.int64), A.val.view(np.int64))
        assert np.array_equal(d.download("PC_RHS"), rhs)
        bnorm = float(np.linalg.norm(rhs))
        out = {}
        for name, s in (("learned", fcg), ("jacobi", jac)):
            r = d.solve_pressure_rhs(s, None)
            true = float(np.linalg.norm(rhs - A.spmv(r.solution)))
            log(f"C4 {name}: {r.status_name} {r.iterations} it {r.device_ms:.1f} ms, host true "
                f"rel res {true / bnorm:.3e}, fallback {r.fallback_used}")
            assert r.converged and true <= 1e-8 * bnorm, (name, true / bnorm)
            out[name] = r
        assert not out["learned"].fallback_used
        assert out["learned"].iterations < out["jacobi"].iterations
        assert out["learned"].device_ms < out["jacobi"].device_ms
    finally:
        d.close()


# ---------------------------------------------------------------------------- C5

def test_c5_batched_256_uneven_cases():
    from cfdapp_amd import learned
    log("case: C5 256 uneven cases, setup")
    t0 = time.perf_counter()
    batch = cfd.BatchSolver()
    systems = []
    try:
        for kind, nx, ny, re in wl.batch_cases():
            dc = learned.make_case(kind, nx, ny=ny, re=re, speed=1.0)
            try:
                learned.assemble_pprime(dc, 0)
                batch.add_pressure(dc)
                systems.append((nx, ny, dc.pin, dc.download("PC_DIAG"),
                                dc.download("PC_FACE_COEFF"), dc.download("PC_RHS")))
            finally:
                dc.close()
        log(f"C5 setup {time.perf_counter() - t0:.1f} s, "
            f"{sum(nx * ny for nx, ny, *_ in systems)} cells")
        s = cfd.LinearSolverSettings(0.0, 1e-8, 100000, "CG", "Jacobi", precision="mixed",
                                     check_interval=32)
        log("case: C5 batched solve")
        res = batch.solve(s)
        st = batch.stats
        lat = np.array([r.device_ms for r in res])
        log(f"C5 batched: {st['device_ms']:.1f} ms device, {st['steps']} steps, "
            f"{st['tile_rows']}-row tiles, p50 {np.percentile(lat, 50):.1f} ms, "
            f"p99 {np.percentile(lat, 99):.1f} ms")
        worst = 0.0
        for k, ((nx, ny, pin, diag, fc, rhs), r) in enumerate(zip(systems, res)):
            assert r.status == cfd.CONVERGED, (k, r.status_name)
            x = batch.solution(k)
            bn = float(np.linalg.norm(rhs))
            true = float(np.linalg.norm(rhs - wl.apply_pprime_2d(nx, ny, pin, diag, fc, x)))
            worst = max(worst, true / bn if bn > 0 else true)
            assert true <= 1e-8 * bn, (k, nx, ny, true / bn)
        log(f"C5 worst host true rel res {worst:.3e}")
    finally:
        batch.close()
