"""GPU parity of the p' system on a general face-list mesh (cfdhip_fmesh_*).

Fixtures tests/golden/fpcorr*.npz come from the reference's own sources
(oracle/ref_driver.cpp fpcorr): a createStructuredQuad2D mesh whose interior
vertices are moved by a seeded fraction of the cell size (every internal and
boundary face non-orthogonal or skewed), and two such blocks joined by
createMultiBlock2D along a conformal interface; seeded predictor flux, response
coefficients d_u != d_v and a previous p' whose Green-Gauss gradient feeds the
explicit term; assemblePressureCorrection with the non-orthogonal correction,
its CG + Jacobi solution and correctFaceMassFlux with the explicit flux. On
the device, from the reference's mesh arrays:
  * CSR, right-hand side, face coefficients and explicit fluxes bitwise, closed
    (cell 0 pinned) and with a FixedValue outlet (no pin);
  * the corrected flux from the reference's p' bitwise;
  * the device CG + Jacobi solution within 1e-8 of the reference's (rel 1e-12
    solves; reduction order differs), and the flux corrected with it;
  * on a Cartesian mesh (the oracle's mesh arrays) the face-list assembly is
    the structured path's system bit for bit (the pcorr2d_40x24 fixture);
  * the learned preconditioner is refused on this path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["fpcorr2d_12x10", "fpcorr2d_16x8_outlet", "fpcorr2d_multiblock_outlet"]


def _bits(a):
    return np.asarray(a, np.float64).view(np.int64)


def _mesh(g, outlet):
    import cfdapp_amd as cfd
    nf = len(g["owner"])
    fixed = np.zeros(nf, np.int32)
    if outlet:  # patch order left, right, bottom, top: "right" is FixedValue
        po = g["patch_offsets"]
        fixed[g["patch_faces"][po[1]:po[2]]] = 1
    fc = np.stack([g["face_cx"], g["face_cy"], g["face_cz"]], 1)
    sf = np.stack([g["face_sx"], g["face_sy"], g["face_sz"]], 1)
    cc = np.stack([g["cell_cx"], g["cell_cy"], g["cell_cz"]], 1)
    return cfd.FaceMeshPressure(2, g["owner"], g["neighbor"], fc, sf, cc,
                                g["cell_face_offsets"], g["cell_faces"], fixed)


@pytest.mark.parametrize("name", NAMES)
def test_fmesh_system_and_flux_bitwise(golden, name):
    import cfdapp_amd as cfd
    print(f"face-list p' {name}", flush=True)
    g = golden(name)
    outlet = "outlet=1" in [str(p) for p in g["params"]]
    m = _mesh(g, outlet)
    try:
        grad = np.stack([g["grad_x"], g["grad_y"], np.zeros_like(g["grad_x"])], 1)
        m.assemble(g["flux"], g["du"], g["dv"], grad_previous=grad)
        (rp, col, val), rhs, coeff, expl = m.system()
        assert np.array_equal(rp, g["A_rowptr"]) and np.array_equal(col, g["A_col"])
        assert np.array_equal(_bits(val), _bits(g["A_val"]))
        assert np.array_equal(_bits(rhs), _bits(g["rhs"]))
        assert np.array_equal(_bits(coeff), _bits(g["face_coeff"]))
        assert np.array_equal(_bits(expl), _bits(g["explicit_flux"]))
        assert np.abs(expl).max() > 0  # the explicit term is exercised
        F = m.correct_flux(g["cg_x"])
        assert np.array_equal(_bits(F), _bits(g["flux_corrected"]))
        r = m.solve(cfd.LinearSolverSettings(0.0, 1e-12, 20000, "CG", "Jacobi"))
        assert r.status == 0
        ref = g["cg_x"]
        err = np.max(np.abs(r.solution - ref)) / np.max(np.abs(ref))
        Fe = np.max(np.abs(m.correct_flux() - g["flux_corrected"])) / np.max(
            np.abs(g["flux_corrected"]))
        print(f"{r.iterations} it (reference {int(g['cg_meta'][1])}), x err {err:.1e}, "
              f"F err {Fe:.1e}", flush=True)
        assert err < 1e-8 and Fe < 1e-8
        with pytest.raises(cfd.CfdHipError):
            m.solve(cfd.LinearSolverSettings(0.0, 1e-8, 100, "FCG", "Learned"))
    finally:
        m.close()


def test_fmesh_on_cartesian_mesh_is_the_structured_system(golden):
    import cfdapp_amd as cfd
    import oracle
    g = golden("pcorr2d_40x24")
    c = oracle.Case.cavity(2, 40, 24)
    a = c.mesh_arrays()
    m = cfd.FaceMeshPressure(2, a["owner"], a["neighbor"], a["fc"], a["sf"], a["cc"],
                             a["cf_off"], a["cf"], np.zeros(len(a["owner"]), np.int32))
    try:
        ones = np.ones(c.n_cells)
        m.assemble(g["flux"], ones, ones, non_orthogonal=True)
        (rp, col, val), rhs, coeff, expl = m.system()
        assert np.array_equal(rp, g["A_rowptr"]) and np.array_equal(col, g["A_col"])
        assert np.array_equal(_bits(val), _bits(g["A_val"]))
        assert np.array_equal(_bits(rhs), _bits(g["rhs"]))
        assert np.array_equal(_bits(coeff), _bits(g["face_coeff"]))
        assert not np.any(expl)
    finally:
        m.close()
