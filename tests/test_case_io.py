"""Case files and result writers (cfdapp_amd.case) against the reference's own.

tests/golden/cases/<name>/ holds the reference's lid_driven_cavity and
poiseuille_flow case files and their committed results (fields.csv and
residuals.csv, gzipped), copied by tests/golden/make_golden.py.

  * The writers reproduce the reference's files byte for byte: the values
    parsed from its fields.csv / residuals.csv, with the cell centroids of
    the oracle's mesh (pinned bitwise against MeshGeometry) and the velocity
    magnitude recomputed from u and v, print to the same bytes. This pins the
    columns, the number format (std::scientific, 17 digits), the magnitude
    formula and the centroid formula of the x, y columns.
  * The readers accept both cases with the values their files hold, reject
    what the reference's parsers reject (unknown keys, out-of-range values)
    with its message shape, and refuse what the device does not run.
"""
import gzip
import json
import os
import shutil

import numpy as np
import pytest

import oracle
from cfdapp_amd import case as cio

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "cases")


def _text(name, f):
    with gzip.open(os.path.join(CASES, name, f + ".gz"), "rt", newline="") as g:
        return g.read()


def _columns(text):
    rows = text.strip("\n").split("\n")
    head = rows[0].split(",")
    data = np.array([[float(x) for x in r.split(",")] for r in rows[1:]])
    return head, data


@pytest.mark.parametrize("name", ["lid_driven_cavity", "poiseuille_flow"])
def test_fields_writer_is_byte_compatible(tmp_path, name):
    cfg = cio.read_case(os.path.join(CASES, name))
    ref = _text(name, "fields.csv")
    head, data = _columns(ref)
    assert head == ["cell_id", "x", "y", "velocity_x", "velocity_y", "velocity_magnitude",
                    "pressure"]
    nx, ny = cfg["counts"]
    c = oracle.Case(2, nx, ny, 1, cfg["lengths"][0], cfg["lengths"][1])
    cent = c.mesh_arrays()["cc"]
    out = tmp_path / "fields.csv"
    cio.write_fields_csv(out, cent, [data[:, 3], data[:, 4]], data[:, 6])
    assert out.read_bytes() == ref.encode()


@pytest.mark.parametrize("name", ["lid_driven_cavity", "poiseuille_flow"])
def test_residuals_writer_is_byte_compatible(tmp_path, name):
    ref = _text(name, "residuals.csv")
    head, data = _columns(ref)
    assert head == ["iteration", "u_residual", "v_residual", "p_residual",
                    "continuity_residual", "global_mass_imbalance"]
    assert np.array_equal(data[:, 0], np.arange(1, len(data) + 1))
    assert np.all(data[:, 5] == data[-1, 5])  # the final value on every row
    hist = dict(u=data[:, 1], v=data[:, 2], p=data[:, 3], c=data[:, 4])
    out = tmp_path / "residuals.csv"
    cio.write_residuals_csv(out, hist, data[-1, 5])
    assert out.read_bytes() == ref.encode()


def test_writer_3d_layout(tmp_path):
    cent = np.array([[0.25, 0.25, 0.25], [0.75, 0.25, 0.25]])
    u = [np.array([3.0, -0.0]), np.array([4.0, 0.0]), np.array([0.0, 1e-300])]
    cio.write_fields_csv(tmp_path / "f.csv", cent, u, np.array([1.0, 2.0]))
    lines = (tmp_path / "f.csv").read_text().split("\n")
    assert lines[0] == ("cell_id,x,y,z,velocity_x,velocity_y,velocity_z,velocity_magnitude,"
                        "pressure")
    assert lines[1].split(",")[7] == "5.00000000000000000e+00"
    assert lines[2].split(",")[4] == "-0.00000000000000000e+00"
    hist = dict(u=[1.0], v=[2.0], w=[3.0], p=[4.0], c=[5.0])
    cio.write_residuals_csv(tmp_path / "r.csv", hist, 0.5)
    assert (tmp_path / "r.csv").read_text().split("\n")[0] == (
        "iteration,u_residual,v_residual,w_residual,p_residual,continuity_residual,"
        "global_mass_imbalance")
    with pytest.raises(ValueError):
        cio.write_fields_csv(tmp_path / "g.csv", cent, u, np.array([1.0, np.nan]))


def test_read_reference_cases():
    cav = cio.read_case(os.path.join(CASES, "lid_driven_cavity"))
    assert (cav["dim"], cav["counts"], cav["lengths"]) == (2, [20, 20], [1.0, 1.0])
    assert (cav["density"], cav["viscosity"]) == (1.0, 0.01)
    assert [p["name"] for p in cav["patches"]] == ["left", "right", "bottom", "top"]
    assert cav["patches"][3]["velocity"] == (1, (1.0, 0.0, 0.0))
    s = cav["solver"]
    assert (s["max_iterations"], s["velocity_relaxation"], s["pressure_relaxation"]) == \
        (6000, 0.7, 0.3)
    assert s["pressure"] == dict(type="BiCGSTAB", backend="CPU", absolute_tolerance=1e-10,
                                 relative_tolerance=1e-8, max_iterations=2000,
                                 preconditioner="none", precision="fp64")
    assert s["face_flux"] == "automatic"
    poi = cio.read_case(os.path.join(CASES, "poiseuille_flow"))
    assert poi["counts"] == [64, 8] and poi["lengths"][0] == 8.0
    assert poi["patches"][0]["velocity"] == (2, (1.0, 0.0, 0.0))
    assert poi["patches"][1]["pressure"] == (1, 0.0)


def _edited(tmp_path, edit, name="lid_driven_cavity"):
    d = tmp_path / name
    shutil.copytree(os.path.join(CASES, name), d)
    with open(d / "solver.json") as f:
        s = json.load(f)
    edit(s)
    with open(d / "solver.json", "w") as f:
        json.dump(s, f)
    return str(d)


def test_new_solver_keys(tmp_path):
    def learned(s):
        s["pressure_linear_solver"].update(type="FCG", relative_tolerance=1e-10)
        s["momentum_linear_solver"]["precision"] = "fp64"

    cfg = cio.read_case(_edited(tmp_path, learned))
    assert cfg["solver"]["pressure"]["type"] == "FCG"
    assert cfg["solver"]["pressure"]["preconditioner"] == "learned"


@pytest.mark.parametrize("edit,message", [
    (lambda s: s.update(bogus=1), 'field "bogus" must satisfy be a recognized field of '
                                  'solver.json; received unknown field'),
    (lambda s: s.update(velocity_relaxation=1.5), 'field "velocity_relaxation" must satisfy '
                                                  "satisfy 0 < value <= 1"),
    (lambda s: s.update(max_iterations=0), 'field "max_iterations" must satisfy be > 0'),
    (lambda s: s.update(type="PISO"), 'field "type" must satisfy be one of: SIMPLE'),
    (lambda s: s["pressure_linear_solver"].update(relative_tolerance=0),
     'field "pressure_linear_solver.relative_tolerance" must satisfy be > 0'),
    (lambda s: s["pressure_linear_solver"].update(preconditioner="learned"),
     "be learned exactly when type is FCG"),
    (lambda s: s["momentum_linear_solver"].update(type="FCG"),
     'field "momentum_linear_solver.type" must satisfy be one of: BiCGSTAB, CG, GMRES'),
    (lambda s: s.update(convection_scheme="quick"), "not supported on the device"),
    (lambda s: s.update(gradient_scheme="least_squares"), "not supported on the device"),
])
def test_rejected_solver_files(tmp_path, edit, message):
    with pytest.raises(cio.CaseConfigurationError) as e:
        cio.read_case(_edited(tmp_path, edit))
    assert message in str(e.value)


def test_unsupported_mesh_and_model(tmp_path):
    d = tmp_path / "c"
    shutil.copytree(os.path.join(CASES, "lid_driven_cavity"), d)
    with open(d / "mesh.json", "w") as f:
        json.dump({"type": "multiblock", "blocks": []}, f)
    with pytest.raises(cio.CaseConfigurationError, match="structured_cartesian only"):
        cio.read_case(str(d))
