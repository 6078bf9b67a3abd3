"""The oracle's PISO time step pinned against the compiled reference.

tests/golden/piso*.npz come from the reference's own PISO::solveTimeStep
(src/pressure_velocity/PISO.cpp:126-350) and assembleTransientMomentumComponent
(src/pressure_velocity/TransientMomentum.cpp), run from rest for `warm` steps on
a lid-driven cavity, then `steps` more (tests/golden/make_golden.py). The
oracle restatement (oracle/cfd_oracle.cpp: assembleMomentum with the time
term, pisoStep) must reproduce every array bit for bit: the transient momentum
CSR, right-hand side and diagonal from the dumped state, and the velocity,
pressure and face flux after every step, plus the step's CFL, continuity and
mass imbalance.
"""
import numpy as np
import pytest

import oracle

CASES = [("piso2d_17", 17, 17), ("piso2d_24x16_re400", 24, 16)]


def _params(g):
    return dict(p.split("=") for p in g["params"])


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.int64),
                          np.asarray(b, np.float64).view(np.int64))


@pytest.mark.parametrize("name,nx,ny", CASES)
def test_transient_momentum_is_bitwise(golden, name, nx, ny):
    g = golden(name)
    prm = _params(g)
    re, dt = float(prm["re"]), float(prm["dt"])
    c = oracle.Case.cavity(2, nx, ny)
    u = (g["in_u_x"], g["in_u_y"])
    for k, nm in enumerate(("mu", "mv")):
        A, rhs, diag = c.momentum(k, u, g["in_p"], g["in_F"], 1.0 / re, u[k], 1.0, dt=dt)
        assert np.array_equal(A.rowptr, g[nm + "_rowptr"]) and np.array_equal(A.col, g[nm + "_col"])
        assert _same(A.val, g[nm + "_val"]) and _same(rhs, g[nm + "_rhs"])
        assert _same(diag, g[nm + "_diag"])
    # the time term is what makes it transient: without it the diagonal shrinks by rho V / dt
    A0, _, d0 = c.momentum(0, u, g["in_p"], g["in_F"], 1.0 / re, u[0], 1.0)
    vol = 1.0 / (nx * ny)
    np.testing.assert_allclose(g["mu_diag"] - d0, vol / dt, rtol=1e-12)


@pytest.mark.parametrize("name,nx,ny", CASES)
def test_piso_steps_are_bitwise(golden, name, nx, ny):
    g = golden(name)
    prm = _params(g)
    re, dt = float(prm["re"]), float(prm["dt"])
    steps = len(g["status"])
    c = oracle.Case.cavity(2, nx, ny)
    u, p, F = (g["in_u_x"], g["in_u_y"]), g["in_p"], g["in_F"]
    for s in range(steps):
        r = c.piso(1, dt, u, p, F, re=re, pres=(1e-10, 1e-8, 20000))
        st = r["steps"][0]
        assert st["status"] == g["status"][s] == 0
        sc = g["scalars"][3 * s:3 * s + 3]
        assert _same(st["max_cfl"], sc[0]) and _same(st["continuity"], sc[1])
        assert _same(st["mass_imbalance"], sc[2])
        u, p, F = r["u"], r["p"], r["F"]
        for arr, key in ((u[0], "u_x"), (u[1], "u_y"), (p, "p"), (F, "F")):
            assert _same(arr, g[f"s{s}_{key}"]), (s, key)


def test_piso_rejects_bad_input_and_keeps_the_state():
    c = oracle.Case.cavity(2, 8, 8)
    n, nf = c.n_cells, c.n_faces
    u, p, F = (np.zeros(n), np.zeros(n)), np.zeros(n), np.zeros(nf)
    r = c.piso(1, 0.0, u, p, F)  # dt <= 0: InvalidConfiguration
    assert r["steps"][0]["status"] == 4
    bad = np.zeros(n)
    bad[3] = np.nan
    r = c.piso(1, 0.01, (bad, np.zeros(n)), p, F)  # NonFiniteState
    assert r["steps"][0]["status"] == 3 and np.isnan(r["u"][0][3])
