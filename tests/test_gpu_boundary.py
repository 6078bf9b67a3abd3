"""The reference's GPU seams on the device.

test_reference_simple_runs_its_device_paths runs oracle/_ref/dropin_simple:
the reference's own SIMPLE::solve (src/pressure_velocity/SIMPLE.cpp), compiled
from its sources with its three GPU stubs replaced by tests/cpp/dropin's glue
over libcfdhip, on a 33 x 33 cavity. It takes the resident outer loop, the
resident p' solve with host-returned momentum systems, and the host round trip
with makeGpuCG, and each run's u, v, p and mass flux must agree with the same
solve on the reference's CPU path within 1e-10 relative (plus a converged
steady state, and the authority bookkeeping and prepare()'s refusals).

The remaining tests pin the entry points that seam uses through the Python
mirror of the C ABI: per-component momentum assembly against the
all-components pass, the host-returning corrections against the resident
ones, the all-finite guards and the workspace byte count.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.test_oracle_golden import assert_bitwise

pytestmark = pytest.mark.gpu

cfd = pytest.importorskip("cfdapp_amd")

HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN = os.path.join(HERE, "..", "oracle", "_ref", "dropin_simple")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not cfd.device_available():
        pytest.fail("no gfx950 device: the HIP path must run on the GPU")


def test_reference_simple_runs_its_device_paths():
    if not os.path.exists(DROPIN):
        pytest.skip("oracle/_ref/dropin_simple is built only where the reference sources are")
    print("case: dropin_simple 33 20", flush=True)
    p = subprocess.run([DROPIN, "33", "20"], capture_output=True, text=True, timeout=300)
    print(p.stdout[-6000:], flush=True)
    print(p.stderr[-2000:], flush=True)
    assert p.returncode == 0, "the reference's SIMPLE on libcfdhip differs from its CPU run"
    assert "PASS" in p.stdout


def _stage(d, alpha=0.7):
    """A cavity state with non-trivial velocity, pressure and flux."""
    n = d.n_cells
    rng = np.random.default_rng(5)
    u = [rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)]
    d.upload("U", u[0])
    d.upload("V", u[1])
    d.upload("P", rng.uniform(-1, 1, n))
    d.mass_flux()
    return u


def test_component_assembly_matches_all_components_pass():
    print("case: momentum per component 40x24", flush=True)
    d = cfd.GpuSimpleDiscretization.cavity(2, 40, 24)
    u = _stage(d)
    d.assemble_momentum_resident(0.7)
    rp, col, val = d.momentum_csr()
    rhs = [d.download("MOM_RHS_U"), d.download("MOM_RHS_V")]
    diag = d.download("MOM_DIAG")
    for k in range(2):
        d.assemble_momentum_component(k, 0.7)
        rp2, col2, val2 = d.momentum_csr()
        assert np.array_equal(rp, rp2) and np.array_equal(col, col2)
        assert_bitwise(val2, val, "momentum matrix")
        assert_bitwise(d.download("MOM_RHS_" + "UV"[k]), rhs[k], "momentum rhs")
        # the caller's phi_old equal to the resident component: the same bits
        d.assemble_momentum_component(k, 0.7, previous_component=u[k])
        assert_bitwise(d.download("MOM_RHS_" + "UV"[k]), rhs[k], "momentum rhs, phi_old given")
    # another phi_old moves only the relaxation source (1 - a) / a * a_P * phi_old
    shift = np.full(d.n_cells, 0.125)
    d.assemble_momentum_component(0, 0.7, previous_component=u[0] + shift)
    got = d.download("MOM_RHS_U") - rhs[0]
    want = (1.0 - 0.7) / 0.7 * (diag * 0.7) * shift
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
    with pytest.raises(cfd.CfdHipError):
        d.assemble_momentum_component(0, 0.7, convection_scheme=1)
    with pytest.raises(cfd.CfdHipError):
        d.assemble_momentum_component(0, 0.7, non_orthogonal=True)


def test_host_corrections_equal_resident_ones():
    print("case: corrections 33x33", flush=True)
    d = cfd.GpuSimpleDiscretization.cavity(2, 33, 33)
    _stage(d)
    d.assemble_momentum_resident(0.7)
    for k in range(2):
        d.solve_momentum_resident(k, cfd.LinearSolverSettings(1e-14, 1e-10, 500, "BiCGSTAB"))
    d.compute_response_coefficients()
    d.compute_predicted_face_flux(True, 0.7)
    d.assemble_pressure_correction_resident(0)
    r = d.solve_pressure_correction_resident(cfd.LinearSolverSettings(0.0, 1e-10, 5000, "CG",
                                                                      "Jacobi"))
    assert r.converged
    u0 = d.download("U")
    F0 = d.download("FLUX")
    hu = d.correct_velocity(resident=False)
    hf = d.correct_flux(resident=False)
    assert_bitwise(d.download("U"), u0, "resident velocity left as it was")
    assert_bitwise(d.download("FLUX"), F0, "resident flux left as it was")
    d.correct_velocity()
    d.correct_flux()
    assert_bitwise(d.download("U"), hu[0], "u")
    assert_bitwise(d.download("V"), hu[1], "v")
    assert_bitwise(d.download("FLUX"), hf, "flux")
    # the split corrections equal the combined one
    d.upload("U", u0)
    d.upload("FLUX", F0)
    d.correct()
    assert_bitwise(d.download("U"), hu[0], "combined u")
    assert_bitwise(d.download("FLUX"), hf, "combined flux")
    assert d.resident_solve_bytes() > 0


def test_all_finite_guards():
    print("case: all-finite 16x16", flush=True)
    d = cfd.GpuSimpleDiscretization.cavity(2, 16, 16)
    _stage(d)
    assert d.all_finite("U", "V", "W", "P", "FLUX")
    assert d.all_finite()
    p = d.download("P")
    for bad in (np.nan, np.inf, -np.inf):
        q = p.copy()
        q[37] = bad
        d.upload("P", q)
        assert not d.all_finite("P")
        assert d.all_finite("U", "V")
    F = d.download("FLUX")
    F[-1] = np.nan
    d.upload("FLUX", F)
    assert not d.all_finite("U", "FLUX")
