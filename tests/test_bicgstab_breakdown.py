"""The oracle's BiCGSTAB against the reference's breakdown and restart cases.

Fixtures (tests/golden/bicgstab_*.npz, written by make_golden.py from the
compiled reference's ``BiCGSTAB::solve``) hold the systems of the reference's
own tests/unit/algebra/test_bicgstab_breakdown.cpp: the two stored 1024-row
thermal systems (tests/data/linear_systems/), its exactly dyadic 3x3 / 2x2 /
1x1 systems (a breakdown without progress, a breakdown after progress that a
restart recovers, exact termination, a zero right-hand side, an exact guess)
and its upwind convection-diffusion systems scaled by powers of two. The
oracle (src/algebra/BiCGSTAB.cpp:48-275 restated) must reproduce every
outcome bit for bit: status, iterations, restarts, history and solution.
The assertions of the reference test itself are checked on the fixture too,
so a fixture written from a broken build would fail here.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import load_golden
from tests.test_oracle_golden import assert_bitwise

STORED = load_golden("bicgstab_stored")
SMALL = load_golden("bicgstab_small")


def system(z, key):
    A = oracle.Csr(z[key + "_rowptr"], z[key + "_col"], z[key + "_val"])
    return A, z[key + "_b"], z[key + "_x0"], z[key + "_settings"]


def oracle_solve(z, key):
    A, b, x0, (a, r, m) = system(z, key)
    return A.bicgstab(b, x0=x0, abs_tol=a, rel_tol=r, max_iter=int(m), jacobi=False)


def test_reference_outcomes_in_the_fixtures():
    """The expectations of test_bicgstab_breakdown.cpp, on the fixture."""
    for key in ("smooth", "rough"):
        st, it, rs = STORED[key + "_meta"]
        assert (st, rs) == (0, 0) and STORED[key + "_res"][1] <= 1e-12
        A, b, _, _ = system(STORED, key)
        assert np.linalg.norm(b - A.spmv(STORED[key + "_x"])) <= 1e-12
    expect = {"rotation": (2, 0, 0), "omega": (2, 0, 0), "rho_no_progress": (2, 1, 0),
              "exact_1x1": (0, 1, 0), "zero_rhs": (0, 0, 0), "exact_guess": (0, 0, 0)}
    for key, meta in expect.items():
        assert tuple(SMALL[key + "_meta"]) == meta, key
    st, it, rs = SMALL["restart_meta"]
    assert (st, rs) == (0, 1) and it <= 50
    ref = SMALL["cd60_p0_hist"]
    for tag, e in (("cd60_m100", -100), ("cd60_p100", 100)):
        assert np.array_equal(SMALL[tag + "_hist"], ref * np.ldexp(1.0, e))
        assert np.array_equal(SMALL[tag + "_x"], SMALL["cd60_p0_x"])
    assert SMALL["cd200_p0_meta"][1] == SMALL["cd200_m40_meta"][1]


@pytest.mark.parametrize("group,key", [("stored", k) for k in STORED["cases"]] +
                         [("small", k) for k in SMALL["cases"]])
def test_oracle_bicgstab_bitwise(group, key):
    z = STORED if group == "stored" else SMALL
    r = oracle_solve(z, key)
    st, it, rs = z[key + "_meta"]
    assert (r.status, r.iterations, r.restarts) == (st, it, rs)
    assert_bitwise(r.history, z[key + "_hist"], "history")
    assert_bitwise(r.x, z[key + "_x"], "solution")
    assert_bitwise([r.initial, r.final], z[key + "_res"], "residuals")


def _bicgstab_order(z, key, dot):
    """BiCGSTAB.cpp:111-274 without restarts, every inner product summed by
    `dot`: how much the iteration count of a non-symmetric system depends on
    the order of the sums alone."""
    rp, col, val = z[key + "_rowptr"], z[key + "_col"], z[key + "_val"]
    A = oracle.Csr(rp, col, val)
    b = z[key + "_b"]
    a, rel, m = z[key + "_settings"]
    x = z[key + "_x0"].copy()
    r = b - A.spmv(x)
    rh = r.copy()
    b0 = np.sqrt(dot(r, r))
    conv = lambda q: q <= a or q / b0 <= rel  # noqa: E731
    rho_old = alpha = omega = 1.0
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    for it in range(1, int(m) + 1):
        rho = dot(rh, r)
        beta = (rho / rho_old) * (alpha / omega)
        p = r + (p - v * omega) * beta
        v = A.spmv(p)
        alpha = rho / dot(rh, v)
        s = r - v * alpha
        if conv(np.sqrt(dot(s, s))):
            return it
        t = A.spmv(s)
        omega = dot(t, s) / dot(t, t)
        x = x + p * alpha + s * omega
        r = s - t * omega
        if conv(np.sqrt(dot(r, r))):
            return it
        rho_old = rho
    return -1


def _serial(a, b):
    s = 0.0
    for u, w in zip(a.tolist(), b.tolist()):
        s += u * w
    return s


def _pairwise(a, b):
    v = a * b
    while len(v) > 1:
        if len(v) % 2:
            v = np.append(v, 0.0)
        v = v[0::2] + v[1::2]
    return float(v[0])


def test_iteration_counts_depend_on_summation_order():
    """Serial sums reproduce the reference's counts; pairwise sums (the
    device's kind of order) change them: 53 -> 47 on the rough stored system,
    244 -> 258 on the n = 200 convection-diffusion system."""
    assert _bicgstab_order(STORED, "rough", _serial) == 53
    assert _bicgstab_order(STORED, "rough", _pairwise) == 47
    assert _bicgstab_order(SMALL, "cd200_p0", _serial) == 244
    assert _bicgstab_order(SMALL, "cd200_p0", _pairwise) == 258
