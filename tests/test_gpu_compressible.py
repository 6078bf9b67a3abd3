"""GPU parity of the compressible pressure-correction system.

cfdhip_assemble_compressible_pressure against the reference's
assembleCompressiblePressureCorrection (tests/golden/cpcorr*.npz, written by
the reference's own sources; the oracle reproduces them bit for bit in
tests/test_compressible_oracle.py): CSR, right-hand side and face
coefficients bitwise, on a closed cavity (cell 0 pinned) and with a FixedValue
pressure patch (no pin). The device solves of that system, Jacobi-PCG and
flexible CG with the learned preconditioner, reach the reference CG
solution to what the tolerance allows (1e-8 relative in the max norm at
rel 1e-12). Inputs the reference rejects (InvalidArgumentError there) fail with
CFDHIP_ERR_INVALID, a 3D mesh with CFDHIP_ERR_UNSUPPORTED.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["cpcorr2d_24x16", "cpcorr2d_20x12_outlet"]


def _case(g):
    import cfdapp_amd as cfd
    p = dict(str(x).split("=", 1) for x in g["params"])
    d = cfd.GpuSimpleDiscretization.cavity(2, int(p["nx"]), int(p["ny"]))
    if int(p.get("outlet", "0")):
        d.set_boundary(1, cfd.WALL, (0.0, 0.0, 0.0), cfd.FIXED_VALUE, 0.0)
    d.upload("RESPONSE", g["d"])
    d.upload("FLUX_STAR", g["flux"])
    return d, float(p.get("dtau", "1e-3"))


def _assemble(d, g, dtau):
    d.assemble_compressible_pressure_correction_resident(g["face_density"], g["p_abs"], g["T"],
                                                         287.05, dtau, 0)


def _bits(a):
    return np.asarray(a, np.float64).view(np.int64)


@pytest.mark.parametrize("name", NAMES)
def test_compressible_system_bitwise(golden, name):
    print(f"compressible p' {name}", flush=True)
    g = golden(name)
    d, dtau = _case(g)
    try:
        _assemble(d, g, dtau)
        rp, col, val = d.pressure_csr()
        assert np.array_equal(rp, g["A_rowptr"]) and np.array_equal(col, g["A_col"])
        assert np.array_equal(_bits(val), _bits(g["A_val"]))
        assert np.array_equal(_bits(d.download("PC_RHS")), _bits(g["rhs"]))
        assert np.array_equal(_bits(d.download("PC_FACE_COEFF")), _bits(g["face_coeff"]))
        # the plain assembly afterwards is the incompressible system again
        d.assemble_pressure_correction_resident(0)
        _, _, val0 = d.pressure_csr()
        assert not np.array_equal(_bits(val0), _bits(g["A_val"]))
    finally:
        d.close()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("method", ["jacobi", "learned"])
def test_compressible_solve_matches_reference(golden, name, method):
    import cfdapp_amd as cfd
    print(f"compressible p' solve {name} {method}", flush=True)
    g = golden(name)
    d, dtau = _case(g)
    try:
        if method == "learned":
            d.set_learned_weights()
            s = cfd.LinearSolverSettings(0.0, 1e-12, 2000, "FCG", "Learned")
        else:
            s = cfd.LinearSolverSettings(0.0, 1e-12, 20000, "CG", "Jacobi")
        _assemble(d, g, dtau)
        r = d.solve_pressure_correction_resident(s)
        assert r.status == 0
        x = d.download("PPRIME")
        ref = g["cg_x"]
        err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
        print(f"{r.iterations} it, |x - x_ref| / |x_ref| = {err:.1e}", flush=True)
        assert err < 1e-8
    finally:
        d.close()


@pytest.mark.parametrize("bad", ["p", "T", "R", "dtau", "ref"])
def test_compressible_rejects_invalid_inputs(golden, bad):
    import cfdapp_amd as cfd
    g = golden("cpcorr2d_24x16")
    d, dtau = _case(g)
    try:
        p, T = g["p_abs"].copy(), g["T"].copy()
        R, ref = 287.05, 0
        if bad == "p":
            p[5] = -1.0
        elif bad == "T":
            T[7] = np.nan
        elif bad == "R":
            R = 0.0
        elif bad == "dtau":
            dtau = np.inf
        else:
            ref = d.n_cells
        with pytest.raises(cfd.CfdHipError) as e:
            d.assemble_compressible_pressure_correction_resident(g["face_density"], p, T, R,
                                                                 dtau, ref)
        assert e.value.code == cfd._lib.ERR_INVALID
    finally:
        d.close()


def test_compressible_3d_is_unsupported():
    import cfdapp_amd as cfd
    d = cfd.GpuSimpleDiscretization.cavity(3, 4, 4, 4)
    try:
        with pytest.raises(cfd.CfdHipError) as e:
            d.assemble_compressible_pressure_correction_resident(
                np.ones(d.n_faces), np.full(d.n_cells, 1e5), np.full(d.n_cells, 300.0), 287.05,
                1e-3, 0)
        assert e.value.code == cfd._lib.ERR_UNSUPPORTED
    finally:
        d.close()
