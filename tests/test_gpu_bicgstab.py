"""Device BiCGSTAB (cfdhip_csr_solve) on the reference's breakdown and restart cases.

The systems and the reference's own outcomes are the fixtures of
tests/golden/bicgstab_*.npz (see test_bicgstab_breakdown.py, which pins the
oracle to them). On the device the inner products are summed in a different
order, so:
  * the exactly dyadic hand-built systems (every recurrence value exact in
    FP64) must give the reference's status, iteration count, restart count
    and residuals exactly -- including the breakdown after progress that one
    restart recovers (BiCGSTAB.cpp:118-144) and the breakdowns without
    progress that are reported;
  * the stored 1024-row thermal systems must converge with no restart to the
    reference's absolute tolerance on the recomputed residual. Their
    iteration counts depend on the summation order of the inner products
    (a restatement with pairwise instead of serial sums takes 47 instead of
    the reference's 53 iterations on the rough system, 258 instead of 244 on
    the n = 200 convection-diffusion system), so the counts are bounded
    loosely (within 25 % of the reference's);
  * a system scaled by a power of two must take the same decisions as the
    unscaled one, with the history scaled exactly (the device is
    deterministic and the scaling exact), as the reference's
    ScalingTheSystemByAPowerOfTwoChangesNothing asserts.
"""
import numpy as np
import pytest

from tests.conftest import load_golden

pytestmark = pytest.mark.gpu

cfd = pytest.importorskip("cfdapp_amd")

STORED = load_golden("bicgstab_stored")
SMALL = load_golden("bicgstab_small")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not cfd.device_available():
        pytest.fail("no gfx950 device: the HIP path must run on the GPU")


def device_solve(z, key):
    a, r, m = z[key + "_settings"]
    solver = cfd.make_gpu_bicgstab(cfd.LinearSolverSettings(a if a > 0 else 0.0, r, int(m),
                                                            "BiCGSTAB"))
    return solver.solve(z[key + "_rowptr"], z[key + "_col"], z[key + "_val"], z[key + "_b"],
                        z[key + "_x0"])


def true_residual(z, key, x):
    import oracle
    A = oracle.Csr(z[key + "_rowptr"], z[key + "_col"], z[key + "_val"])
    return float(np.linalg.norm(z[key + "_b"] - A.spmv(x)))


EXACT = ["rotation", "omega", "rho_no_progress", "exact_1x1", "zero_rhs", "exact_guess"]


@pytest.mark.parametrize("key", EXACT)
def test_dyadic_systems_match_the_reference_exactly(key):
    print("case: bicgstab", key, flush=True)
    r = device_solve(SMALL, key)
    st, it, rs = SMALL[key + "_meta"]
    assert (r.status, r.iterations, r.restarts) == (st, it, rs), (r.status_name, r.iterations,
                                                                  r.restarts)
    assert [r.initial_residual, r.final_residual] == list(SMALL[key + "_res"])
    assert np.array_equal(r.residual_history, SMALL[key + "_hist"])
    assert np.array_equal(r.solution, SMALL[key + "_x"])
    if st == cfd.BREAKDOWN:
        # the reported residual is the returned iterate's (no silent acceptance)
        assert abs(r.final_residual - true_residual(SMALL, key, r.solution)) <= 1e-14


def test_breakdown_after_progress_is_recovered_by_one_restart():
    """(rHat, v) = 0 exactly in the second iteration: one restart from the
    current iterate, then convergence on the recomputed residual. The
    recurrence is exact up to the restart; after it the sums round, and the
    device sums in another order, so the last iterations may differ by one."""
    print("case: bicgstab restart", flush=True)
    r = device_solve(SMALL, "restart")
    st, it, rs = SMALL["restart_meta"]
    assert r.status == cfd.CONVERGED and r.restarts == rs == 1
    assert abs(r.iterations - it) <= 1 and r.iterations <= 50
    n_exact = 2  # b0 and the first completed iteration precede the breakdown
    assert np.array_equal(r.residual_history[:n_exact], SMALL["restart_hist"][:n_exact])
    assert true_residual(SMALL, "restart", r.solution) <= 1e-13
    again = device_solve(SMALL, "restart")  # deterministic
    assert (again.iterations, again.restarts, again.final_residual) == \
        (r.iterations, r.restarts, r.final_residual)


@pytest.mark.parametrize("key", ["smooth", "rough"])
def test_stored_thermal_systems_converge_without_restart(key):
    print("case: bicgstab stored", key, flush=True)
    r = device_solve(STORED, key)
    st, it, rs = STORED[key + "_meta"]
    assert r.status == cfd.CONVERGED and r.restarts == 0
    assert abs(r.iterations - it) <= 0.25 * it, (r.iterations, it)
    assert r.final_residual <= 1e-12
    assert true_residual(STORED, key, r.solution) <= 1e-12
    assert r.initial_residual == pytest.approx(STORED[key + "_res"][0], rel=1e-12)


def test_power_of_two_scaling_changes_nothing():
    print("case: bicgstab scaling", flush=True)
    base = device_solve(SMALL, "cd60_p0")
    assert base.status == cfd.CONVERGED
    assert abs(base.iterations - int(SMALL["cd60_p0_meta"][1])) <= 2
    for tag, e in (("cd60_m100", -100), ("cd60_p100", 100)):
        r = device_solve(SMALL, tag)
        assert r.status == cfd.CONVERGED and r.iterations == base.iterations, tag
        assert np.array_equal(r.residual_history, base.residual_history * np.ldexp(1.0, e)), tag
        assert np.array_equal(r.solution, base.solution), tag


def test_ill_conditioned_system_is_solved_at_any_scale():
    print("case: bicgstab cd200", flush=True)
    a = device_solve(SMALL, "cd200_p0")
    b = device_solve(SMALL, "cd200_m40")
    for r, tag in ((a, "cd200_p0"), (b, "cd200_m40")):
        assert r.status == cfd.CONVERGED
        assert r.final_residual <= 1e-10 * r.initial_residual
        assert true_residual(SMALL, tag, r.solution) <= 1e-9 * r.initial_residual
        assert abs(r.iterations - int(SMALL[tag + "_meta"][1])) <= 0.25 * 244
    print(f"cd200: device {a.iterations} / {b.iterations} it, reference "
          f"{int(SMALL['cd200_p0_meta'][1])}", flush=True)
    assert a.iterations == b.iterations
