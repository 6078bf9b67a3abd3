"""GPU parity of the cell gradients and the p' corrector passes on general
face-list meshes (cfdhip_fmesh_gradient, cfdhip_fmesh_assemble_pressure_pass,
cfdhip_fmesh_pressure_passes).

Fixtures tests/golden/fpcorr*.npz come from the reference's own sources
(oracle/ref_driver.cpp fpcorr): distorted createStructuredQuad2D meshes and a
two-block createMultiBlock2D mesh, every face non-orthogonal or skewed. They
hold a seeded cell field and the reference's gradient() of it -- Green-Gauss
(boundary-consistent face values, skewness sweeps, oblique Neumann faces) and
least squares -- under the p' conditions of the explicit term and under
conditions with non-zero values (FixedValue 0.3 on the right patch,
FixedGradient -0.7 elsewhere); and SIMPLE's corrector passes on the p'
system: each pass's right-hand side, explicit flux and CG + Jacobi solution
(the passes after the first take the gradient of the pass before, Green-Gauss
or, in fpcorr2d_14x9_lsq, least squares), and the flux corrected after the
last. On the device:
  * every gradient bitwise;
  * each pass's system bitwise when it is assembled from the reference's p'
    of the pass before (the device computes the gradient itself);
  * the passes run end to end on the device: every solve converged, the last
    p' and the corrected flux within 1e-8 of the reference's (rel 1e-12 solves,
    reduction order differs).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["fpcorr2d_12x10", "fpcorr2d_16x8_outlet", "fpcorr2d_multiblock_outlet",
         "fpcorr2d_14x9_lsq"]


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def _params(g):
    return dict(str(p).split("=", 1) for p in g["params"])


def _mesh(g):
    import cfdapp_amd as cfd
    outlet = _params(g).get("outlet") == "1"
    nf = len(g["owner"])
    fixed = np.zeros(nf, np.int32)
    po = g["patch_offsets"]
    right = g["patch_faces"][po[1]:po[2]]  # patch order left, right, bottom, top
    if outlet:
        fixed[right] = 1
    fc = np.stack([g["face_cx"], g["face_cy"], g["face_cz"]], 1)
    sf = np.stack([g["face_sx"], g["face_sy"], g["face_sz"]], 1)
    cc = np.stack([g["cell_cx"], g["cell_cy"], g["cell_cz"]], 1)
    m = cfd.FaceMeshPressure(2, g["owner"], g["neighbor"], fc, sf, cc, g["cell_face_offsets"],
                             g["cell_faces"], fixed)
    m.set_cell_volumes(g["cell_vol"])
    return m, right


def _check(name, got, gx, gy):
    assert np.array_equal(_bits(got[:, 0]), _bits(gx)), name + " x"
    assert np.array_equal(_bits(got[:, 1]), _bits(gy)), name + " y"


@pytest.mark.parametrize("name", NAMES)
def test_gradients_bitwise(golden, name):
    g = golden(name)
    print(f"case: face-list gradients {name}", flush=True)
    m, right = _mesh(g)
    try:
        prev = g["prev"]
        _check("green-gauss, p' conditions", m.gradient(prev), g["grad_x"], g["grad_y"])
        _check("least squares, p' conditions", m.gradient(prev, scheme="least_squares"),
               g["grad_ls_x"], g["grad_ls_y"])
        nf = len(g["owner"])
        kind = np.zeros(nf, np.int32)
        val = np.full(nf, -0.7)
        kind[right] = 1
        val[right] = 0.3
        _check("green-gauss, valued conditions", m.gradient(prev, kind, val),
               g["grad_v_x"], g["grad_v_y"])
        _check("least squares, valued conditions",
               m.gradient(prev, kind, val, scheme="least_squares"), g["grad_vls_x"],
               g["grad_vls_y"])
        # the schemes and conditions differ, so every comparison above pins something
        assert np.abs(g["grad_ls_x"] - g["grad_x"]).max() > 1e-3
        assert np.abs(g["grad_v_x"] - g["grad_x"]).max() > 1e-3
    finally:
        m.close()


@pytest.mark.parametrize("name", NAMES)
def test_corrector_passes(golden, name):
    import cfdapp_amd as cfd
    g = golden(name)
    prm = _params(g)
    passes = int(prm.get("passes", "3"))
    scheme = "least_squares" if prm.get("gscheme") == "1" else "green_gauss"
    print(f"case: face-list corrector passes {name} ({passes}, {scheme})", flush=True)
    m, _ = _mesh(g)
    try:
        m.assemble(g["flux"], g["du"], g["dv"], non_orthogonal=True)
        _, rhs, _, expl = m.system()
        assert np.array_equal(_bits(rhs), _bits(g["pass1_rhs"]))
        assert not np.any(expl)
        # each later pass from the reference's p' of the pass before
        for k in range(2, passes + 1):
            m.assemble_pass(g[f"pass{k - 1}_x"], scheme=scheme)
            _, rhs, _, expl = m.system()
            assert np.array_equal(_bits(expl), _bits(g[f"pass{k}_explicit"])), k
            assert np.array_equal(_bits(rhs), _bits(g[f"pass{k}_rhs"])), k
            assert np.abs(expl).max() > 0
        # end to end on the device
        m.assemble(g["flux"], g["du"], g["dv"], non_orthogonal=True)
        x, res = m.passes(cfd.LinearSolverSettings(0.0, 1e-12, 20000, "CG", "Jacobi"), passes,
                          scheme=scheme)
        print("  iterations " + " ".join(str(r.iterations) for r in res) + " (reference " +
              " ".join(str(int(g[f"pass{k}_meta"][1])) for k in range(1, passes + 1)) + ")",
              flush=True)
        assert all(r.status == cfd.CONVERGED for r in res)
        ref = g[f"pass{passes}_x"]
        assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max()
        F = m.correct_flux()
        Fr = g["passes_flux"]
        assert np.abs(F - Fr).max() <= 1e-8 * np.abs(Fr).max()
    finally:
        m.close()


def test_gradient_refusals(golden):
    import cfdapp_amd as cfd
    g = golden("fpcorr2d_12x10")
    m, _ = _mesh(g)
    try:
        with pytest.raises(cfd.CfdHipError):  # no field
            m._ctx.call("cfdhip_fmesh_gradient", None, None, None, 0, None)
        with pytest.raises(cfd.CfdHipError):  # no first-pass assembly yet
            m.passes(cfd.LinearSolverSettings(0.0, 1e-10, 100, "CG", "Jacobi"), 2)
        with pytest.raises(ValueError):
            m.set_cell_volumes(np.ones(3))
    finally:
        m.close()
