"""Parity of the HIP path (through the C ABI) with the reference.

Bitwise (float64 bit patterns, signed zeros included) for index construction,
assembly, gradients, fluxes, corrections, continuity and SpMV; within stated
tolerances for Krylov solutions, whose reduction order differs from the
reference's serial sums. Inputs are the golden fixtures the compiled
reference wrote (tests/golden/) and seeded fields checked against the oracle.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import load_golden
from tests.test_oracle_golden import assert_bitwise, params

pytestmark = pytest.mark.gpu

cfd = pytest.importorskip("cfdapp_amd")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not cfd.device_available():
        pytest.fail("no gfx950 device: the HIP path must run on the GPU")


def gpu_case(g):
    p = params(g)
    dim = int(p["dim"])
    return cfd.GpuSimpleDiscretization.cavity(dim, int(p["nx"]), int(p["ny"]), int(p.get("nz", 1)),
                                              viscosity=1.0 / float(p.get("re", 100))), p


def print_case(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("name", ["mesh2d_5x4", "mesh3d_3x4x2"])
def test_face_order_bitwise(name):
    print_case(name)
    g = load_golden(name)
    d, _ = gpu_case(g)
    owner, neigh = d.mesh_faces()
    assert np.array_equal(owner, g["owner"])
    assert np.array_equal(neigh, g["neighbor"])


@pytest.mark.parametrize("dims", [(2, 257, 129, 1), (3, 17, 9, 12)])
def test_face_order_matches_oracle(dims):
    print_case("faces", dims)
    dim, nx, ny, nz = dims
    d = cfd.GpuSimpleDiscretization.cavity(dim, nx, ny, nz)
    owner, neigh = d.mesh_faces()
    m = oracle.Case.cavity(dim, nx, ny, nz).mesh_arrays()
    assert np.array_equal(owner, m["owner"]) and np.array_equal(neigh, m["neighbor"])


@pytest.mark.parametrize("name", ["pcorr2d_33", "pcorr2d_40x24", "pcorr3d_9x8x7"])
def test_pressure_assembly_csr_bitwise(name):
    print_case(name)
    g = load_golden(name)
    d, _ = gpu_case(g)
    d.upload("RESPONSE", np.ones(d.n_cells))
    d.upload("FLUX_STAR", g["flux"])
    d.assemble_pressure_correction_resident(0)
    rp, col, val = d.pressure_csr()
    assert np.array_equal(rp, g["A_rowptr"])
    assert np.array_equal(col, g["A_col"])
    assert_bitwise(val, g["A_val"], "p' values")
    assert_bitwise(d.download("PC_RHS"), g["rhs"], "p' rhs")
    assert_bitwise(d.download("PC_FACE_COEFF"), g["face_coeff"], "face coefficients")
    # SpMV of the resident operator == reference SparseMatrix::multiply
    assert_bitwise(d.pressure_spmv(g["cg_x"]), g["Ax"], "pressure SpMV")


@pytest.mark.parametrize("name", ["pcorr2d_33", "pcorr3d_9x8x7"])
def test_pressure_cg_solution(name):
    print_case(name)
    g = load_golden(name)
    d, _ = gpu_case(g)
    d.upload("RESPONSE", np.ones(d.n_cells))
    d.upload("FLUX_STAR", g["flux"])
    d.assemble_pressure_correction_resident(0)
    s = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi")
    r = d.solve_pressure_rhs(s, g["b"])
    ref_iters = int(g["cg_meta"][1])
    assert r.converged and abs(r.iterations - ref_iters) <= max(2, ref_iters // 50)
    b = g["b"]
    assert r.true_residual <= 1.5e-8 * np.linalg.norm(b)
    # the same converged solution as the reference CG (to the solve's accuracy)
    x_ref = g["cg_x"]
    assert np.abs(r.solution - x_ref).max() <= 1e-6 * np.abs(x_ref).max()
    # reference semantics: history[0] = |b|, recurrence residual at exit <= rel*b0
    assert r.residual_history[0] == pytest.approx(np.linalg.norm(b), rel=1e-12)
    assert r.final_residual <= 1e-8 * r.initial_residual


@pytest.mark.parametrize("shape", [(129, 129), (257, 96), (40, 36, 24)])
def test_mixed_precision_pressure_solve(shape):
    """FP32 iteration vectors with an FP64 iterate and FP64 residual
    replacement reach the FP64 true-residual target and the oracle's FP64 CG
    solution (within what that target guarantees); 2D (single-kernel
    iteration) and 3D (7-point, two kernels per iteration). The history holds
    the FP64 residuals: |b| first, then one per FP64 check."""
    print_case("mixed", shape)
    dim = len(shape)
    nx, ny = shape[:2]
    nz = shape[2] if dim == 3 else 1
    d = cfd.GpuSimpleDiscretization.cavity(dim, nx, ny, nz)
    c = oracle.Case.cavity(dim, nx, ny, nz)
    rng = np.random.default_rng(nx + ny + nz)
    D = rng.uniform(0.5, 2.0, d.n_cells)  # variable response coefficients
    Fs = rng.uniform(-1, 1, d.n_faces)
    d.upload("RESPONSE", D)
    d.upload("FLUX_STAR", Fs)
    d.assemble_pressure_correction_resident(0)
    A, rhs, _ = c.pressure_system(Fs, D, D, D if dim == 3 else None)
    ref = A.cg(rhs, rel_tol=1e-12, max_iter=50000)
    s = cfd.LinearSolverSettings(0.0, 1e-8, 50000, "CG", "Jacobi", precision="mixed")
    r = d.solve_pressure_rhs(s, rhs)
    assert r.converged, r.status_name
    true = np.linalg.norm(rhs - A.spmv(r.solution))
    assert true <= 1e-8 * np.linalg.norm(rhs)
    assert r.true_residual == pytest.approx(true, rel=1e-6)
    h = r.residual_history
    assert len(h) >= 2 and h[0] == pytest.approx(np.linalg.norm(rhs), rel=1e-12)
    assert h[-1] == r.true_residual
    assert np.abs(r.solution - ref.x).max() <= 1e-5 * np.abs(ref.x).max()
    # FP64 fused path on the same system, for the iteration comparison
    r64 = d.solve_pressure_rhs(cfd.LinearSolverSettings(0.0, 1e-8, 50000, "CG", "Jacobi"), rhs)
    assert r64.converged and r.iterations <= 2 * r64.iterations


def test_csr_spmv_and_solvers_bitwise_spmv():
    print_case("csr spmv/solve")
    g = load_golden("stages2d_33_rc")
    A = oracle.Csr(g["mu_rowptr"], g["mu_col"], g["mu_val"])
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, A.n)
    solver = cfd.make_gpu_bicgstab(cfd.LinearSolverSettings(1e-10, 1e-8, 500))
    assert_bitwise(solver.spmv(A.rowptr, A.col, A.val, x), A.spmv(x), "CSR SpMV")
    r = solver.solve(A.rowptr, A.col, A.val, g["mu_rhs"], g["in_u_x"])
    ref = g["mu_solve_x"]
    assert r.converged
    assert abs(r.iterations - int(g["mu_solve_meta"][1])) <= 2
    assert np.abs(r.solution - ref).max() <= 1e-6 * np.abs(ref).max()
    cg = cfd.make_gpu_cg(cfd.LinearSolverSettings(0.0, 1e-8, 20000, preconditioner="Jacobi"))
    P = oracle.Csr(g["pc_rowptr"], g["pc_col"], g["pc_val"])
    rc = cg.solve(P.rowptr, P.col, P.val, g["pc_rhs"])
    assert rc.converged
    pref = g["pc_solve_x"]
    assert np.abs(rc.solution - pref).max() <= 1e-6 * np.abs(pref).max()


STAGES = ["stages2d_33_rc", "stages2d_24x16_re1000", "stages2d_20_linear", "stages3d_8_rc"]


@pytest.mark.parametrize("name", STAGES)
def test_simple_stages_bitwise(name):
    """Each resident stage on the reference's own inputs, bit for bit."""
    print_case(name)
    g = load_golden(name)
    d, p = gpu_case(g)
    dim = d.dim
    comps = "xyz"[:dim]
    d.upload_initial_state([g["in_u_" + k] for k in comps], g["in_p"], g["in_F"])
    grad = d.gradient("P")
    for k in range(dim):
        assert_bitwise(grad[k], g["gradp_" + comps[k]], "grad p " + comps[k])
    d.assemble_momentum_resident(0.7)
    rp, col, val = d.momentum_csr()
    for nm in ["mu", "mv", "mw"][:dim]:
        assert np.array_equal(rp, g[nm + "_rowptr"]), nm
        assert np.array_equal(col, g[nm + "_col"]), nm
        assert_bitwise(val, g[nm + "_val"], nm + " values")
        assert_bitwise(d.download("MOM_DIAG"), g[nm + "_diag"], nm + " diagonal")
    for k, nm in enumerate(["mu", "mv", "mw"][:dim]):
        assert_bitwise(d.download("MOM_RHS_" + "UVW"[k]), g[nm + "_rhs"], nm + " rhs")
    mom = cfd.LinearSolverSettings(1e-10, 1e-8, 500, "BiCGSTAB")
    for k, nm in enumerate(["mu", "mv", "mw"][:dim]):
        r = d.solve_momentum_resident(k, mom)
        ref = g[nm + "_solve_x"]
        assert r.converged
        assert abs(r.iterations - int(g[nm + "_solve_meta"][1])) <= 2
        got = d.download("UVW"[k] + "STAR")
        assert np.abs(got - ref).max() <= 1e-6 * max(np.abs(ref).max(), 1e-30)
    # continue from the reference's own momentum solutions (bitwise chain)
    for k, nm in enumerate(["mu", "mv", "mw"][:dim]):
        d.upload("UVW"[k] + "STAR", g[nm + "_solve_x"])
    d.compute_response_coefficients()
    assert_bitwise(d.download("RESPONSE"), g["du"], "response d")
    d.compute_predicted_face_flux(p["flux"] == "rhie_chow", 0.7)
    assert_bitwise(d.download("FLUX_STAR"), g["Fstar"], "predicted face flux")
    d.assemble_pressure_correction_resident(0)
    rp, col, val = d.pressure_csr()
    assert np.array_equal(rp, g["pc_rowptr"]) and np.array_equal(col, g["pc_col"])
    assert_bitwise(val, g["pc_val"], "p' values")
    assert_bitwise(d.download("PC_RHS"), g["pc_rhs"], "p' rhs")
    assert_bitwise(d.download("PC_FACE_COEFF"), g["pc_face_coeff"], "p' face coefficients")
    if p["psolver"] == "cg_jacobi":
        ps = cfd.LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi")
    else:
        ps = cfd.LinearSolverSettings(1e-10, 1e-8, 2000, "BiCGSTAB")
    pr = d.solve_pressure_correction_resident(ps)
    pref = g["pc_solve_x"]
    assert pr.converged
    # same tolerance test as the reference run; the iterate it stops at can
    # differ in slowly converging modes (unpreconditioned BiCGSTAB), so the
    # solutions are compared at the accuracy the tolerance guarantees
    assert pr.true_residual <= 2.0 * max(ps.absolute_tolerance,
                                         ps.relative_tolerance * pr.initial_residual)
    tol = 1e-6 if p["psolver"] == "cg_jacobi" else 1e-4
    assert np.abs(d.download("PPRIME") - pref).max() <= tol * np.abs(pref).max()
    d.upload("PPRIME", pref)
    gpp = d.gradient("PPRIME", correction=True)
    for k in range(dim):
        assert_bitwise(gpp[k], g["gradpp_" + comps[k]], "grad p' " + comps[k])
    d.update_pressure(0.3)
    assert_bitwise(d.download("P"), g["p_new"], "relaxed pressure")
    d.correct()
    for k in range(dim):
        assert_bitwise(d.download("UVW"[k]), g["u_new_" + comps[k]], "corrected u" + comps[k])
    assert_bitwise(d.download("FLUX"), g["F_new"], "corrected flux")
    imb, rms, glob = d.continuity()
    assert_bitwise(imb, g["imbalance"], "continuity imbalance")
    assert glob == g["cont_scalars"][1]
    assert rms == pytest.approx(g["cont_scalars"][0], rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("dims", [(2, 129, 129, 1), (2, 64, 200, 1), (3, 24, 20, 16)])
def test_stages_bitwise_vs_oracle_random_state(dims):
    """Seeded non-trivial state at larger sizes: assembly chain vs the oracle."""
    print_case("random state", dims)
    dim, nx, ny, nz = dims
    rng = np.random.default_rng(nx * 1000 + ny)
    c = oracle.Case.cavity(dim, nx, ny, nz)
    d = cfd.GpuSimpleDiscretization.cavity(dim, nx, ny, nz, viscosity=0.004)
    n = c.n_cells
    u = [rng.uniform(-1, 1, n) for _ in range(dim)]
    pres = rng.uniform(-1, 1, n)
    F = c.mass_flux(u)
    d.upload_initial_state(u, pres, F)
    d.assemble_momentum_resident(0.7)
    A, rhs, diag = c.momentum(0, u, pres, F, 0.004, u[0], 0.7)
    rp, col, val = d.momentum_csr()
    assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col)
    assert_bitwise(val, A.val, "momentum values")
    assert_bitwise(d.download("MOM_RHS_U"), rhs, "momentum rhs")
    ustar = [rng.uniform(-1, 1, n) for _ in range(dim)]
    for k in range(dim):
        d.upload("UVW"[k] + "STAR", ustar[k])
    d.compute_response_coefficients()
    dd = c.response(diag)
    assert_bitwise(d.download("RESPONSE"), dd, "response")
    d.compute_predicted_face_flux(True, 0.7)
    Fs = c.rhie_chow(ustar, pres, dd, dd, dd if dim == 3 else None, 1.0, 0.7)
    assert_bitwise(d.download("FLUX_STAR"), Fs, "Rhie-Chow flux")
    d.assemble_pressure_correction_resident(0)
    P, prhs, coeff = c.pressure_system(Fs, dd, dd, dd if dim == 3 else None)
    rp, col, val = d.pressure_csr()
    assert np.array_equal(rp, P.rowptr) and np.array_equal(col, P.col)
    assert_bitwise(val, P.val, "p' values")
    pp = rng.uniform(-1e-3, 1e-3, n)
    assert_bitwise(d.pressure_spmv(pp), P.spmv(pp), "p' SpMV")
    d.upload("PPRIME", pp)
    d.correct()
    unew, Fnew = c.correct(ustar, dd, dd, dd if dim == 3 else None, pp, Fs, coeff)
    for k in range(dim):
        assert_bitwise(d.download("UVW"[k]), unew[k], "corrected velocity")
    assert_bitwise(d.download("FLUX"), Fnew, "corrected flux")


def test_simple_run_matches_reference_history():
    """SIMPLE 17x17 Re=100 Rhie-Chow, CG+Jacobi p', tolerance 1e-6 (reference run)."""
    print_case("simple2d_17_rc")
    g = load_golden("simple2d_17_rc")
    d, p = gpu_case(g)
    res = d.simple_run(int(p["iters"]), rhie_chow=True)
    assert res["status"] == 0
    assert abs(res["iterations"] - int(g["meta"][1])) <= 1
    k = min(res["iterations"], len(g["hist_u"]))
    np.testing.assert_allclose(res["hist"]["u"][:k], g["hist_u"][:k], rtol=1e-6)
    np.testing.assert_allclose(res["hist"]["p"][:k], g["hist_p"][:k], rtol=1e-6)
    u = d.download_velocity()
    pr = d.download_pressure()
    assert np.abs(u[0] - g["u_x"]).max() <= 1e-6 * np.abs(g["u_x"]).max()
    assert np.abs(pr - g["p"]).max() <= 1e-6 * np.abs(g["p"]).max()


@pytest.mark.parametrize("dims", [(2, 17, 17, 1, 100.0), (3, 6, 6, 6, 400.0)])
def test_steady_state_fields_within_1e10(dims):
    """Tightly converged steady state: FP64 fields within 1e-10 relative of the oracle."""
    print_case("steady", dims)
    dim, nx, ny, nz, re = dims
    tol = 1e-13
    c = oracle.Case.cavity(dim, nx, ny, nz)
    ref = c.simple(4000, re=re, vtol=tol, ptol=tol, ctol=tol, pressure="bicgstab",
                   pres=(1e-16, 1e-8, 2000), mom=(1e-16, 1e-8, 500))
    assert ref["status"] == 0
    d = cfd.GpuSimpleDiscretization.cavity(dim, nx, ny, nz, viscosity=1.0 / re)
    res = d.simple_run(4000, velocity_tol=tol, pressure_tol=tol, continuity_tol=tol,
                       momentum=cfd.LinearSolverSettings(1e-16, 1e-8, 500, "BiCGSTAB"),
                       pressure=cfd.LinearSolverSettings(1e-16, 1e-8, 2000, "BiCGSTAB"))
    assert res["status"] == 0
    u = d.download_velocity()
    pr = d.download_pressure()
    for k in range(dim):
        err = np.abs(u[k] - ref["u"][k]).max() / np.abs(ref["u"][k]).max()
        assert err <= 1e-10, (k, err)
    err = np.abs(pr - ref["p"]).max() / np.abs(ref["p"]).max()
    assert err <= 1e-10, err


def test_failure_statuses_on_device():
    """Reference numerical outcomes, forced from the host: InvalidSystem on a
    zero diagonal, NonFiniteInput on a NaN RHS, MaxIterations on a tiny budget."""
    print_case("statuses")
    g = load_golden("pcorr2d_33")
    A = oracle.Csr(g["A_rowptr"], g["A_col"], g["A_val"])
    cg = cfd.make_gpu_cg(cfd.LinearSolverSettings(0.0, 1e-8, 3, preconditioner="Jacobi"))
    r = cg.solve(A.rowptr, A.col, A.val, g["b"])
    assert r.status == cfd.MAX_ITERATIONS and r.iterations == 3
    ref = A.cg(g["b"], max_iter=3)
    assert ref.status == 1 and ref.iterations == 3
    bad = g["b"].copy()
    bad[5] = np.nan
    assert cg.solve(A.rowptr, A.col, A.val, bad).status == cfd.NONFINITE_INPUT
    val = A.val.copy()
    val[A.rowptr[7]:A.rowptr[8]][A.col[A.rowptr[7]:A.rowptr[8]] == 7] = 1e-13
    assert cg.solve(A.rowptr, A.col, val, g["b"]).status == cfd.INVALID_SYSTEM
