"""The oracle's compressible pressure-correction system against the reference.

Fixtures tests/golden/cpcorr*.npz come from the reference's own
assembleCompressiblePressureCorrection (src/compressible/
CompressiblePressureCorrection.cpp:20-80) on seeded inputs: predictor flux,
per-face density, response coefficient (both components), absolute pressure
and temperature of an ideal gas (R = 287.05); a closed cavity (cell 0 pinned)
and a cavity whose right patch is a FixedValue pressure (no pin). The oracle
must reproduce the CSR, right-hand side, face coefficients and the reference's
CG + Jacobi solution bit for bit.

Checks that do not lean on the fixtures:
  * with a huge pseudo time step the compressibility diagonal vanishes and a
    constant per-face density gives the incompressible system bit for bit;
  * the compressibility diagonal is V / dtau / (R T) per unpinned cell: the
    difference of two assemblies that differ only in T is that term's change;
  * inputs the reference rejects (non-positive pressure, temperature, gas
    constant or pseudo time step, a 3D mesh) are rejected.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import load_golden

NAMES = ["cpcorr2d_24x16", "cpcorr2d_20x12_outlet"]


def _params(g):
    return dict(str(p).split("=", 1) for p in g["params"])


def _case(g):
    p = _params(g)
    c = oracle.Case.cavity(2, int(p["nx"]), int(p["ny"]))
    if int(p.get("outlet", "0")):
        c.set_bc(1, oracle.WALL, pkind=oracle.FIXED_VALUE, pvalue=0.0)  # "right"
    return c, float(p.get("dtau", "1e-3"))


def _bits(a):
    return np.asarray(a, np.float64).view(np.int64)


def _diag(A):
    n = len(A.rowptr) - 1
    d = np.zeros(n)
    for i in range(n):
        for k in range(A.rowptr[i], A.rowptr[i + 1]):
            if A.col[k] == i:
                d[i] = A.val[k]
    return d


@pytest.mark.parametrize("name", NAMES)
def test_compressible_system_bitwise(name):
    g = load_golden(name)
    c, dtau = _case(g)
    A, rhs, coeff = c.compressible_pressure_system(g["flux"], g["d"], g["face_density"],
                                                   g["p_abs"], g["T"], 287.05, dtau)
    assert np.array_equal(A.rowptr, g["A_rowptr"]) and np.array_equal(A.col, g["A_col"])
    assert np.array_equal(_bits(A.val), _bits(g["A_val"]))
    assert np.array_equal(_bits(rhs), _bits(g["rhs"]))
    assert np.array_equal(_bits(coeff), _bits(g["face_coeff"]))
    r = A.cg(rhs, abs_tol=0.0, rel_tol=1e-12, max_iter=20000, jacobi=True)
    meta = g["cg_meta"]
    assert (r.status, r.iterations) == (int(meta[0]), int(meta[1]))
    assert np.array_equal(_bits(r.x), _bits(g["cg_x"]))


def test_vanishing_compressibility_is_the_incompressible_system():
    c = oracle.Case.cavity(2, 12, 9)
    rng = np.random.default_rng(5)
    F = rng.uniform(-1, 1, c.n_faces)
    d = rng.uniform(0.5, 1.5, c.n_cells)
    rho = 1.3
    A0, b0, f0 = c.pressure_system(F, d, d, rho=rho)
    A1, b1, f1 = c.compressible_pressure_system(F, d, np.full(c.n_faces, rho),
                                                np.full(c.n_cells, 1e5),
                                                np.full(c.n_cells, 300.0), 287.05, 1e300)
    assert np.array_equal(A0.rowptr, A1.rowptr) and np.array_equal(A0.col, A1.col)
    assert np.array_equal(b0, b1) and np.array_equal(f0, f1)
    # V / dtau / (R T) underflows to ~1e-306: adding it changes no diagonal
    assert np.array_equal(A0.val, A1.val)


def test_compressibility_diagonal_is_volume_over_dtau_rt():
    c = oracle.Case.cavity(2, 10, 7)
    rng = np.random.default_rng(9)
    F = rng.uniform(-1, 1, c.n_faces)
    d = rng.uniform(0.5, 1.5, c.n_cells)
    rf = rng.uniform(0.5, 1.5, c.n_faces)
    p = np.full(c.n_cells, 1e5)
    T1, T2 = rng.uniform(250, 350, c.n_cells), rng.uniform(250, 350, c.n_cells)
    R, dtau = 287.05, 2e-3
    A1, _, _ = c.compressible_pressure_system(F, d, rf, p, T1, R, dtau)
    A2, _, _ = c.compressible_pressure_system(F, d, rf, p, T2, R, dtau)
    vol = c.mesh_arrays()["vol"]
    expect = vol / dtau * (1.0 / (R * T1) - 1.0 / (R * T2))
    delta = _diag(A1) - _diag(A2)
    expect[0] = 0.0  # the pinned cell's row is the identity
    # each diagonal rounds once at its own magnitude (~ the face couplings)
    np.testing.assert_allclose(delta, expect, rtol=1e-12, atol=4e-16 * np.abs(_diag(A1)).max())
    assert _diag(A1)[0] == 1.0


@pytest.mark.parametrize("bad", ["p", "T", "R", "dtau"])
def test_rejected_inputs(bad):
    c = oracle.Case.cavity(2, 6, 5)
    n, nf = c.n_cells, c.n_faces
    args = dict(p=np.full(n, 1e5), T=np.full(n, 300.0), R=287.05, dtau=1e-3)
    if bad in ("p", "T"):
        args[bad] = args[bad].copy()
        args[bad][3] = 0.0
    else:
        args[bad] = -1.0
    assert c.compressible_pressure_system(np.zeros(nf), np.ones(n), np.ones(nf), args["p"],
                                          args["T"], args["R"], args["dtau"]) is None


def test_rejects_3d():
    c = oracle.Case.cavity(3, 4, 4, 4)
    n, nf = c.n_cells, c.n_faces
    assert c.compressible_pressure_system(np.zeros(nf), np.ones(n), np.ones(nf),
                                          np.full(n, 1e5), np.full(n, 300.0), 287.05,
                                          1e-3) is None
