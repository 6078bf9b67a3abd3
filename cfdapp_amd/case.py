"""Case directories run on the device, results written as the reference writes them.

Reads a case directory in the reference's format (case.json naming geometry,
mesh, physics, boundaries and solver files; src/io/CaseReader.cpp,
src/io/case/*ConfigParser.cpp), builds the resident SIMPLE discretization on
the GPU, runs ``cfdhip_simple_run`` and writes ``fields.csv`` and
``residuals.csv`` byte for byte in the layout of the reference's CSVWriter
(src/io/CSVWriter.cpp:16-139: ``std::scientific`` with 17 digits after the
point, classic locale, the same columns).

The accepted files are the subset this library runs: a ``rectangle`` (2D) or
``box`` (3D) geometry, a ``structured_cartesian`` mesh, the
``incompressible_laminar`` model, wall / moving_wall / inlet / outlet
velocity and fixed_gradient / fixed_value pressure conditions, and the SIMPLE
solver block with upwind convection and Green-Gauss gradients. Everything
else raises CaseConfigurationError naming what is not supported. The solver
block is validated with the reference's rules (required keys, ranges,
unknown keys rejected: SolverConfigParser.cpp:10-58,219-330) and accepts
these additional keys in a linear-solver block:

  "type": "FCG"                 flexible CG (pressure only)
  "preconditioner": "none" | "jacobi" | "learned"
  "precision": "fp64" | "mixed"

Without "preconditioner" a CG or BiCGSTAB solve is unpreconditioned, as the
reference's factory builds it from a case file; "learned" needs "FCG" and
loads the shipped network (cfdapp_amd/weights).

    python -m cfdapp_amd.case CASE_DIR [--output DIR]
"""
import json
import math
import os
import sys

import numpy as np

__all__ = ["CaseConfigurationError", "read_case", "run_case", "write_fields_csv",
           "write_residuals_csv", "format_real"]


class CaseConfigurationError(ValueError):
    """A case file the reference would reject, or one this library does not run."""


def _err(path, field, constraint, received):
    label = f'field "{field}" ' if field else ""
    raise CaseConfigurationError(f"{path}: {label}must satisfy {constraint}; received {received}")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except FileNotFoundError:
        raise CaseConfigurationError(f"{path}: file not found") from None
    except json.JSONDecodeError as e:
        raise CaseConfigurationError(f"{path}: invalid JSON ({e})") from None


def _obj(node, path, field=""):
    if not isinstance(node, dict):
        _err(path, field, "be a JSON object", json.dumps(node))
    return node


def _reject_unknown(node, path, context, allowed):
    for k in node:
        if k not in allowed:
            _err(path, k, f"be a recognized field of {context}", "unknown field")


def _real(node, path, key, label=None, default=None):
    label = label or key
    if key not in node:
        if default is not None:
            return float(default)
        _err(path, label, "be a number", "missing")
    v = node[key]
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        _err(path, label, "be a number", json.dumps(v))
    return float(v)


def _index(node, path, key, label=None, default=None):
    label = label or key
    if key not in node:
        if default is not None:
            return int(default)
        _err(path, label, "be a non-negative integer", "missing")
    v = node[key]
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        _err(path, label, "be a non-negative integer", json.dumps(v))
    return int(v)


def _string(node, path, key, label=None, default=None):
    label = label or key
    if key not in node:
        if default is not None:
            return default
        _err(path, label, "be a string", "missing")
    v = node[key]
    if not isinstance(v, str):
        _err(path, label, "be a string", json.dumps(v))
    return v


def _linear_solver(node, path, field, pressure):
    """parseLinearSolverSpec (SolverConfigParser.cpp:10-58) plus the keys above."""
    if field not in node:
        _err(path, field, "be a JSON object", "missing")
    b = _obj(node[field], path, field)
    _reject_unknown(b, path, field, ["type", "backend", "absolute_tolerance",
                                     "relative_tolerance", "max_iterations", "preconditioner",
                                     "precision"])
    kinds = ["BiCGSTAB", "CG", "GMRES"] + (["FCG"] if pressure else [])
    t = _string(b, path, "type", f"{field}.type")
    if t not in kinds:
        _err(path, f"{field}.type", "be one of: " + ", ".join(kinds), t)
    if t == "GMRES":
        raise CaseConfigurationError(f"{path}: {field}.type GMRES runs on the CPU only (the "
                                     "reference's restarted GMRES); not supported here")
    backend = _string(b, path, "backend", f"{field}.backend", "CPU")
    if backend not in ("CPU", "GPU"):
        _err(path, f"{field}.backend", "be one of: CPU, GPU", backend)
    spec = dict(type=t, backend=backend)
    for key in ("absolute_tolerance", "relative_tolerance"):
        v = _real(b, path, key, f"{field}.{key}")
        if not v > 0.0:
            _err(path, f"{field}.{key}", "be > 0", repr(v))
        spec[key] = v
    spec["max_iterations"] = _index(b, path, "max_iterations", f"{field}.max_iterations")
    if spec["max_iterations"] == 0:
        _err(path, f"{field}.max_iterations", "be > 0", "0")
    pre = _string(b, path, "preconditioner", f"{field}.preconditioner",
                  "learned" if t == "FCG" else "none")
    if pre not in ("none", "jacobi", "learned"):
        _err(path, f"{field}.preconditioner", "be one of: none, jacobi, learned", pre)
    if (pre == "learned") != (t == "FCG"):
        _err(path, f"{field}.preconditioner", "be learned exactly when type is FCG", pre)
    spec["preconditioner"] = pre
    prec = _string(b, path, "precision", f"{field}.precision", "fp64")
    if prec not in ("fp64", "mixed"):
        _err(path, f"{field}.precision", "be one of: fp64, mixed", prec)
    spec["precision"] = prec
    return spec


def _solver(node, path):
    """parseSolverConfig (SolverConfigParser.cpp:219-330)."""
    _obj(node, path)
    _reject_unknown(node, path, "solver.json",
                    ["type", "max_iterations", "velocity_relaxation", "pressure_relaxation",
                     "velocity_tolerance", "pressure_tolerance", "continuity_tolerance",
                     "momentum_linear_solver", "pressure_linear_solver", "convection_scheme",
                     "gradient_scheme", "non_orthogonal_corrections", "robustness", "face_flux"])
    t = _string(node, path, "type")
    if t != "SIMPLE":
        _err(path, "type", "be one of: SIMPLE", t)
    cfg = dict(type=t)
    cfg["max_iterations"] = _index(node, path, "max_iterations")
    if cfg["max_iterations"] == 0:
        _err(path, "max_iterations", "be > 0", "0")
    for key in ("velocity_relaxation", "pressure_relaxation"):
        v = _real(node, path, key)
        if not (0.0 < v <= 1.0):
            _err(path, key, "satisfy 0 < value <= 1", repr(v))
        cfg[key] = v
    for key in ("velocity_tolerance", "pressure_tolerance", "continuity_tolerance"):
        v = _real(node, path, key)
        if not v > 0.0:
            _err(path, key, "be > 0", repr(v))
        cfg[key] = v
    cfg["momentum"] = _linear_solver(node, path, "momentum_linear_solver", False)
    cfg["pressure"] = _linear_solver(node, path, "pressure_linear_solver", True)
    conv = _string(node, path, "convection_scheme", None, "upwind")
    if conv not in ("upwind", "central", "linear_upwind", "quick"):
        _err(path, "convection_scheme", "be one of: upwind, central, linear_upwind, quick", conv)
    grad = _string(node, path, "gradient_scheme", None, "green_gauss")
    if grad not in ("green_gauss", "least_squares"):
        _err(path, "gradient_scheme", "be one of: green_gauss, least_squares", grad)
    nonorth = node.get("non_orthogonal_corrections", 0)
    if isinstance(nonorth, bool) or not isinstance(nonorth, int) or nonorth < 0:
        _err(path, "non_orthogonal_corrections", "be >= 0", json.dumps(nonorth))
    flux = _string(node, path, "face_flux", None, "automatic")
    if flux not in ("automatic", "linear", "rhie_chow"):
        _err(path, "face_flux", "be one of: automatic, linear, rhie_chow", flux)
    if conv != "upwind":
        raise CaseConfigurationError(f"{path}: convection_scheme {conv} is not supported on the "
                                     "device (first-order upwind only)")
    if grad != "green_gauss":
        raise CaseConfigurationError(f"{path}: gradient_scheme {grad} is not supported on the "
                                     "device (Green-Gauss only)")
    if "robustness" in node:
        raise CaseConfigurationError(f"{path}: the robustness block (adaptive relaxation, "
                                     "linear-solver fallback chains) is not supported here")
    cfg.update(convection_scheme=conv, gradient_scheme=grad,
               non_orthogonal_corrections=int(nonorth), face_flux=flux)
    return cfg


_VELOCITY = {"wall": 0, "moving_wall": 1, "inlet": 2, "outlet": 3}
_PRESSURE = {"fixed_gradient": 0, "fixed_value": 1}


def read_case(case_dir):
    """The case directory's files, validated, as one dict."""
    case_path = os.path.join(case_dir, "case.json")
    case = _obj(_load(case_path), case_path)
    _reject_unknown(case, case_path, "case.json",
                    ["name", "description", "format_version", "geometry", "mesh", "physics",
                     "boundaries", "solver"])
    files = {}
    for key in ("geometry", "mesh", "physics", "boundaries", "solver"):
        name = _string(case, case_path, key)
        files[key] = os.path.join(case_dir, name)
    geo = _obj(_load(files["geometry"]), files["geometry"])
    gtype = _string(geo, files["geometry"], "type")
    if gtype not in ("rectangle", "box"):
        raise CaseConfigurationError(f"{files['geometry']}: geometry type {gtype} is not "
                                     "supported here (rectangle or box)")
    dim = 2 if gtype == "rectangle" else 3
    lengths = [_real(geo, files["geometry"], "length"), _real(geo, files["geometry"], "height")]
    if dim == 3:
        lengths.append(_real(geo, files["geometry"], "depth"))
    mesh = _obj(_load(files["mesh"]), files["mesh"])
    mtype = _string(mesh, files["mesh"], "type")
    if mtype != "structured_cartesian":
        raise CaseConfigurationError(f"{files['mesh']}: mesh type {mtype} is not supported here "
                                     "(structured_cartesian only)")
    counts = [_index(mesh, files["mesh"], "nx"), _index(mesh, files["mesh"], "ny")]
    if dim == 3:
        counts.append(_index(mesh, files["mesh"], "nz"))
    phys = _obj(_load(files["physics"]), files["physics"])
    model = _string(phys, files["physics"], "model")
    if model != "incompressible_laminar":
        raise CaseConfigurationError(f"{files['physics']}: model {model} is not supported here "
                                     "(incompressible_laminar only)")
    density = _real(phys, files["physics"], "density")
    viscosity = _real(phys, files["physics"], "dynamic_viscosity")
    bnd = _obj(_load(files["boundaries"]), files["boundaries"])
    patches = _obj(bnd.get("patches"), files["boundaries"], "patches")
    names = (["left", "right", "bottom", "top"] if dim == 2 else
             ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    conds = []
    for name in names:
        if name not in patches:
            _err(files["boundaries"], f"patches.{name}", "be present", "missing")
        p = _obj(patches[name], files["boundaries"], f"patches.{name}")
        vel = _obj(p.get("velocity"), files["boundaries"], f"patches.{name}.velocity")
        pre = _obj(p.get("pressure"), files["boundaries"], f"patches.{name}.pressure")
        vk, pk = vel.get("type"), pre.get("type")
        if vk not in _VELOCITY:
            raise CaseConfigurationError(f"{files['boundaries']}: velocity type {vk} on {name} "
                                         "is not supported here")
        if pk not in _PRESSURE:
            raise CaseConfigurationError(f"{files['boundaries']}: pressure type {pk} on {name} "
                                         "is not supported here")
        vv = [float(x) for x in vel.get("value", [0.0, 0.0, 0.0])] + [0.0, 0.0, 0.0]
        conds.append(dict(name=name, velocity=(_VELOCITY[vk], tuple(vv[:3])),
                          pressure=(_PRESSURE[pk], float(pre.get("value", 0.0)))))
    solver = _solver(_load(files["solver"]), files["solver"])
    return dict(name=case.get("name", ""), dim=dim, lengths=lengths, counts=counts,
                density=density, viscosity=viscosity, patches=conds, solver=solver)


def _settings(spec):
    from . import LinearSolverSettings
    pre = {"none": "None", "jacobi": "Jacobi", "learned": "Learned"}[spec["preconditioner"]]
    return LinearSolverSettings(spec["absolute_tolerance"], spec["relative_tolerance"],
                                spec["max_iterations"], spec["type"], pre,
                                precision=spec["precision"])


def format_real(x):
    """One value as the reference's streams print it (std::scientific, precision 17)."""
    return "%.17e" % x


def write_fields_csv(path, centroids, u, p):
    """CSVWriter::writeFields (2D) / writeFields3D (src/io/CSVWriter.cpp:16-105)."""
    dim = len(u)
    if not all(np.isfinite(c).all() for c in u) or not np.isfinite(p).all():
        raise ValueError("CSVWriter::writeFields: non-finite value")
    f = format_real
    lines = []
    if dim == 2:
        lines.append("cell_id,x,y,velocity_x,velocity_y,velocity_magnitude,pressure")
        for i in range(len(p)):
            ux, uy = float(u[0][i]), float(u[1][i])
            mag = math.sqrt((ux * ux + uy * uy) + 0.0)
            lines.append(f"{i},{f(centroids[i, 0])},{f(centroids[i, 1])},{f(ux)},{f(uy)},"
                         f"{f(mag)},{f(p[i])}")
    else:
        lines.append("cell_id,x,y,z,velocity_x,velocity_y,velocity_z,velocity_magnitude,pressure")
        for i in range(len(p)):
            ux, uy, uz = float(u[0][i]), float(u[1][i]), float(u[2][i])
            mag = math.sqrt((ux * ux + uy * uy) + uz * uz)
            lines.append(f"{i},{f(centroids[i, 0])},{f(centroids[i, 1])},{f(centroids[i, 2])},"
                         f"{f(ux)},{f(uy)},{f(uz)},{f(mag)},{f(p[i])}")
    with open(path, "w", newline="\n") as out:
        out.write("\n".join(lines) + "\n")


def write_residuals_csv(path, hist, global_imbalance):
    """CSVWriter::writeResiduals (src/io/CSVWriter.cpp:107-139): one row per outer
    iteration, the final global mass imbalance repeated on every row, a
    w_residual column for 3D."""
    f = format_real
    with_w = len(hist.get("w", [])) > 0
    n = len(hist["u"])
    lines = []
    if with_w:
        lines.append("iteration,u_residual,v_residual,w_residual,p_residual,continuity_residual,"
                     "global_mass_imbalance")
        for k in range(n):
            lines.append(f"{k + 1},{f(hist['u'][k])},{f(hist['v'][k])},{f(hist['w'][k])},"
                         f"{f(hist['p'][k])},{f(hist['c'][k])},{f(global_imbalance)}")
    else:
        lines.append("iteration,u_residual,v_residual,p_residual,continuity_residual,"
                     "global_mass_imbalance")
        for k in range(n):
            lines.append(f"{k + 1},{f(hist['u'][k])},{f(hist['v'][k])},{f(hist['p'][k])},"
                         f"{f(hist['c'][k])},{f(global_imbalance)}")
    with open(path, "w", newline="\n") as out:
        out.write("\n".join(lines) + "\n")


STATUS = ["Converged", "MaxIterations", "MomentumFailure", "PressureCorrectionFailure",
          "NonFiniteState", "InvalidConfiguration"]


def run_case(case_dir, output_dir=None, device=0):
    """Runs the case on the device from rest and writes its results; returns the
    run's summary. output_dir defaults to CASE_DIR/results."""
    from . import GpuSimpleDiscretization
    cfg = read_case(case_dir)
    s = cfg["solver"]
    d = GpuSimpleDiscretization(device)
    try:
        lx, ly = cfg["lengths"][:2]
        lz = cfg["lengths"][2] if cfg["dim"] == 3 else 1.0
        nz = cfg["counts"][2] if cfg["dim"] == 3 else 1
        ok, why = d.prepare(cfg["dim"], cfg["counts"][0], cfg["counts"][1], nz, lx, ly, lz,
                            velocity=[c["velocity"] for c in cfg["patches"]],
                            pressure=[c["pressure"] for c in cfg["patches"]],
                            density=cfg["density"], viscosity=cfg["viscosity"])
        if not ok:
            raise CaseConfigurationError(f"{case_dir}: {why}")
        if s["pressure"]["preconditioner"] == "learned":
            d.set_learned_weights()
        zeros = np.zeros(d.n_cells)
        for name in ("U", "V", "W")[:cfg["dim"]]:
            d.upload(name, zeros)
        d.upload("P", zeros)
        # face_flux automatic: linear in 2D, Rhie-Chow in 3D (SIMPLESettings.hpp)
        rc = s["face_flux"] == "rhie_chow" or (s["face_flux"] == "automatic" and cfg["dim"] == 3)
        res = d.simple_run(s["max_iterations"], s["velocity_relaxation"],
                           s["pressure_relaxation"], s["velocity_tolerance"],
                           s["pressure_tolerance"], s["continuity_tolerance"], rhie_chow=rc,
                           momentum=_settings(s["momentum"]), pressure=_settings(s["pressure"]))
        u = [d.download(n) for n in ("U", "V", "W")[:cfg["dim"]]]
        p = d.download("P")
        cent, _ = d.cell_centroids()
    finally:
        d.close()
    out = output_dir or os.path.join(case_dir, "results")
    os.makedirs(out, exist_ok=True)
    write_fields_csv(os.path.join(out, "fields.csv"), cent, u, p)
    write_residuals_csv(os.path.join(out, "residuals.csv"), res["hist"], res["global_imbalance"])
    return dict(status=STATUS[res["status"]] if 0 <= res["status"] < len(STATUS) else
                res["status"], iterations=res["iterations"], output=out,
                pressure_iterations=res["pressure_iterations"])


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m cfdapp_amd.case", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("case_dir")
    ap.add_argument("--output", default=None)
    a = ap.parse_args(argv)
    try:
        r = run_case(a.case_dir, a.output)
    except CaseConfigurationError as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
    print(json.dumps(r))
    return 0 if r["status"] in ("Converged", "MaxIterations") else 1


if __name__ == "__main__":
    sys.exit(main())
