"""Pins the CPU oracle bit-for-bit against the compiled reference's vectors.

Every array in tests/golden/*.npz was written by the reference's own sources
(oracle/Makefile.ref -> oracle/_ref/ref_driver); reference_cavity20.npz is the
reference's committed cavity result. Equality here is exact (np.array_equal on
float64), which is what makes the oracle usable as a bitwise checker for the
HIP assembly kernels.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import load_golden


def params(g):
    out = {}
    for item in g["params"]:
        k, v = str(item).split("=", 1)
        out[k] = v
    return out


def case_for(g):
    p = params(g)
    dim = int(p["dim"])
    return oracle.Case.cavity(dim, int(p["nx"]), int(p["ny"]), int(p.get("nz", 1))), p


def csr(g, name):
    return oracle.Csr(g[name + "_rowptr"], g[name + "_col"], g[name + "_val"])


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64) if a.dtype == np.float64 else a,
                                                 b.view(np.int64) if b.dtype == np.float64 else b)


def assert_bitwise(a, b, what):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.nonzero(a.view(np.int64) != b.view(np.int64))[0]
    assert len(bad) == 0, (f"{what}: {len(bad)} entries differ, first at {bad[:5]}: "
                           f"{a[bad[:3]]} vs {b[bad[:3]]}")


@pytest.mark.parametrize("name", ["mesh2d_5x4", "mesh3d_3x4x2"])
def test_mesh_ordering_bitwise(name):
    g = load_golden(name)
    c, _ = case_for(g)
    m = c.mesh_arrays()
    assert np.array_equal(m["owner"], g["owner"])
    assert np.array_equal(m["neighbor"], g["neighbor"])
    assert np.array_equal(m["cf"], g["cell_faces"])
    assert np.array_equal(m["cf_off"], g["cell_face_offsets"])
    for k, gk in [(0, "face_cx"), (1, "face_cy"), (2, "face_cz")]:
        assert_bitwise(m["fc"][:, k], g[gk], gk)
    for k, gk in [(0, "face_sx"), (1, "face_sy"), (2, "face_sz")]:
        assert_bitwise(m["sf"][:, k], g[gk], gk)
    for k, gk in [(0, "cell_cx"), (1, "cell_cy"), (2, "cell_cz")]:
        assert_bitwise(m["cc"][:, k], g[gk], gk)
    assert_bitwise(m["vol"], g["cell_vol"], "volume")
    # patch membership in patch order (left, right, bottom, top / xmin..zmax)
    patch_faces = []
    for p in range(c.n_patches):
        patch_faces.extend(np.nonzero(m["patch"] == p)[0].tolist())
    assert patch_faces == g["patch_faces"].tolist()


@pytest.mark.parametrize("name", ["pcorr2d_33", "pcorr2d_40x24", "pcorr3d_9x8x7"])
def test_pressure_system_and_cg_bitwise(name):
    g = load_golden(name)
    c, _ = case_for(g)
    ones = np.ones(c.n_cells)
    A, rhs, coeff = c.pressure_system(g["flux"], ones, ones, ones if c.dim == 3 else None)
    assert np.array_equal(A.rowptr, g["A_rowptr"])
    assert np.array_equal(A.col, g["A_col"])
    assert_bitwise(A.val, g["A_val"], "p' values")
    assert_bitwise(rhs, g["rhs"], "p' rhs")
    assert_bitwise(coeff, g["face_coeff"], "face coefficients")
    r = A.cg(g["b"], abs_tol=0.0, rel_tol=1e-8, max_iter=20000, jacobi=True)
    meta = g["cg_meta"]
    assert (r.status, r.iterations) == (int(meta[0]), int(meta[1]))
    assert_bitwise(r.x, g["cg_x"], "CG solution")
    assert_bitwise(r.history, g["cg_hist"], "CG residual history")
    assert_bitwise(A.spmv(r.x), g["Ax"], "SpMV")


STAGES = ["stages2d_33_rc", "stages2d_24x16_re1000", "stages2d_20_linear", "stages3d_8_rc"]


@pytest.mark.parametrize("name", STAGES)
def test_simple_stages_bitwise(name):
    g = load_golden(name)
    c, p = case_for(g)
    dim = c.dim
    re = float(p["re"])
    mu = 1.0 / re
    comps = "xyz"[:dim]
    u = [g["in_u_" + k] for k in comps]
    pres, F = g["in_p"], g["in_F"]
    sols, resp = [], []
    for k, nm in enumerate(["mu", "mv", "mw"][:dim]):
        A, rhs, diag = c.momentum(k, u, pres, F, mu, u[k], 0.7)
        assert np.array_equal(A.rowptr, g[nm + "_rowptr"]), nm
        assert np.array_equal(A.col, g[nm + "_col"]), nm
        assert_bitwise(A.val, g[nm + "_val"], nm + " values")
        assert_bitwise(rhs, g[nm + "_rhs"], nm + " rhs")
        assert_bitwise(diag, g[nm + "_diag"], nm + " diag")
        r = A.bicgstab(rhs, x0=u[k], abs_tol=1e-10, rel_tol=1e-8, max_iter=500)
        assert r.iterations == int(g[nm + "_solve_meta"][1])
        assert_bitwise(r.x, g[nm + "_solve_x"], nm + " BiCGSTAB solution")
        assert_bitwise(r.history, g[nm + "_solve_hist"], nm + " BiCGSTAB history")
        sols.append(r.x)
        resp.append(c.response(diag))
        assert_bitwise(resp[-1], g["d" + "uvw"[k]], "response " + "uvw"[k])
    gp = c.gradient(pres)
    for k in range(dim):
        assert_bitwise(gp[k], g["gradp_" + comps[k]], "grad p " + comps[k])
    dw = resp[2] if dim == 3 else None
    assert_bitwise(c.mass_flux(sols), g["Flinear"], "linear mass flux")
    if p["flux"] == "rhie_chow":
        Fstar = c.rhie_chow(sols, pres, resp[0], resp[1], dw, 1.0, 0.7)
    else:
        Fstar = c.mass_flux(sols)
    assert_bitwise(Fstar, g["Fstar"], "predicted face flux")
    A, rhs, coeff = c.pressure_system(Fstar, resp[0], resp[1], dw)
    assert np.array_equal(A.rowptr, g["pc_rowptr"]) and np.array_equal(A.col, g["pc_col"])
    assert_bitwise(A.val, g["pc_val"], "p' values")
    assert_bitwise(rhs, g["pc_rhs"], "p' rhs")
    assert_bitwise(coeff, g["pc_face_coeff"], "p' face coefficients")
    if p["psolver"] == "cg_jacobi":
        pr = A.cg(rhs, abs_tol=0.0, rel_tol=1e-8, max_iter=20000, jacobi=True)
    else:
        pr = A.bicgstab(rhs, abs_tol=1e-10, rel_tol=1e-8, max_iter=2000)
    assert_bitwise(pr.x, g["pc_solve_x"], "p' solution")
    assert_bitwise(pr.history, g["pc_solve_hist"], "p' history")
    gpp = c.gradient(pr.x, correction=True)
    for k in range(dim):
        assert_bitwise(gpp[k], g["gradpp_" + comps[k]], "grad p' " + comps[k])
    unew, Fnew = c.correct(sols, resp[0], resp[1], dw, pr.x, Fstar, coeff)
    for k in range(dim):
        assert_bitwise(unew[k], g["u_new_" + comps[k]], "corrected velocity " + comps[k])
    assert_bitwise(Fnew, g["F_new"], "corrected face flux")
    assert_bitwise(pres + 0.3 * pr.x, g["p_new"], "relaxed pressure")
    imb, rms, glob = c.continuity(Fnew)
    assert_bitwise(imb, g["imbalance"], "continuity imbalance")
    assert_bitwise([rms, glob], g["cont_scalars"], "continuity scalars")


@pytest.mark.parametrize("name", ["simple2d_17_rc", "simple2d_129_rc_20it", "simple3d_6_rc"])
def test_simple_run_bitwise(name):
    g = load_golden(name)
    c, p = case_for(g)
    res = c.simple(int(p["iters"]), re=float(p["re"]), rhie_chow=p["flux"] == "rhie_chow",
                   pressure=p["psolver"])
    meta = g["meta"]
    assert (res["status"], res["iterations"]) == (int(meta[0]), int(meta[1]))
    assert (res["momentum_iterations"], res["pressure_iterations"]) == (int(meta[2]), int(meta[3]))
    comps = "xyz"[:c.dim]
    for k in range(c.dim):
        assert_bitwise(res["u"][k], g["u_" + comps[k]], "u" + comps[k])
    assert_bitwise(res["p"], g["p"], "p")
    assert_bitwise(res["F"], g["F"], "face flux")
    for h in ["u", "v", "p", "c"]:
        assert_bitwise(res["hist"][h], g["hist_" + h], "history " + h)


def test_committed_reference_cavity_bitwise():
    """The reference's own cases/lid_driven_cavity result: 20x20, Re=100, linear
    face flux, unpreconditioned BiCGSTAB pressure, 3312 outer iterations."""
    g = load_golden("reference_cavity20")
    c = oracle.Case.cavity(2, 20, 20)
    res = c.simple(6000, re=100.0, rhie_chow=False, pressure="bicgstab",
                   pres=(1e-10, 1e-8, 2000))
    assert res["status"] == 0 and res["iterations"] == len(g["hist_u"]) == 3312
    assert_bitwise(res["u"][0], g["u"], "u")
    assert_bitwise(res["u"][1], g["v"], "v")
    assert_bitwise(res["p"], g["p"], "p")
    assert_bitwise(res["hist"]["u"], g["hist_u"], "u residual history")
    assert_bitwise(res["hist"]["p"], g["hist_p"], "p residual history")
    assert_bitwise(res["hist"]["c"], g["hist_c"], "continuity history")
