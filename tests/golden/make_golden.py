"""Writes the golden fixtures in this directory from the compiled reference.

Runs ``oracle/_ref/ref_driver`` (the reference's own CPU sources compiled by
``oracle/Makefile.ref``; needs /root/reference, so it runs in the build
container only) and packs each command's raw arrays into ``<name>.npz``. Also
converts the reference's committed 20x20 cavity results
(``cases/lid_driven_cavity/results/{fields,residuals}.csv``) into
``reference_cavity20.npz``, and copies the case files and results of the
reference's lid_driven_cavity and poiseuille_flow cases into ``cases/``, and
writes ``bicgstab_*.npz``: the linear systems of the reference's BiCGSTAB
breakdown tests (tests/unit/algebra/test_bicgstab_breakdown.cpp -- its two
stored 1024-row systems from tests/data/linear_systems/, its hand-built 3x3
systems and its scaled convection-diffusion systems) with the outcome of the
reference's own ``BiCGSTAB::solve`` on each.

    python tests/golden/make_golden.py
"""
import csv
import glob
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")
REF = "/root/reference"

FIXTURES = {
    "mesh2d_5x4": ("mesh", dict(dim=2, nx=5, ny=4)),
    "mesh3d_3x4x2": ("mesh", dict(dim=3, nx=3, ny=4, nz=2)),
    "pcorr2d_33": ("pcorr", dict(dim=2, nx=33, ny=33, seed=20480001)),
    "pcorr2d_40x24": ("pcorr", dict(dim=2, nx=40, ny=24, seed=7)),
    "pcorr3d_9x8x7": ("pcorr", dict(dim=3, nx=9, ny=8, nz=7, seed=3003)),
    "stages2d_33_rc": ("stages", dict(dim=2, nx=33, ny=33, re=100, warm=5, flux="rhie_chow",
                                      psolver="cg_jacobi")),
    "stages2d_24x16_re1000": ("stages", dict(dim=2, nx=24, ny=16, re=1000, warm=12,
                                             flux="rhie_chow", psolver="cg_jacobi")),
    "stages2d_20_linear": ("stages", dict(dim=2, nx=20, ny=20, re=100, warm=40, flux="linear",
                                          psolver="bicgstab")),
    "stages3d_8_rc": ("stages", dict(dim=3, nx=8, ny=8, nz=8, re=400, warm=4, flux="rhie_chow",
                                     psolver="cg_jacobi")),
    "simple2d_17_rc": ("simple", dict(dim=2, nx=17, ny=17, re=100, iters=3000, flux="rhie_chow",
                                      psolver="cg_jacobi")),
    "simple2d_129_rc_20it": ("simple", dict(dim=2, nx=129, ny=129, re=100, iters=20,
                                            flux="rhie_chow", psolver="cg_jacobi")),
    "simple3d_6_rc": ("simple", dict(dim=3, nx=6, ny=6, nz=6, re=400, iters=40,
                                     flux="rhie_chow", psolver="cg_jacobi")),
    "piso2d_17": ("piso", dict(dim=2, nx=17, ny=17, re=100, dt=0.01, warm=3, steps=3,
                               psolver="bicgstab")),
    "piso2d_24x16_re400": ("piso", dict(dim=2, nx=24, ny=16, re=400, dt=0.004, warm=4, steps=2,
                                        psolver="bicgstab")),
    "cpcorr2d_24x16": ("cpcorr", dict(dim=2, nx=24, ny=16, seed=41, outlet=0)),
    "cpcorr2d_20x12_outlet": ("cpcorr", dict(dim=2, nx=20, ny=12, seed=43, outlet=1, dtau=0.05)),
    "fpcorr2d_12x10": ("fpcorr", dict(nx=12, ny=10, seed=11, outlet=0, amp=0.2)),
    "fpcorr2d_16x8_outlet": ("fpcorr", dict(nx=16, ny=8, seed=13, outlet=1, amp=0.25, lx=2.0)),
    "fpcorr2d_multiblock_outlet": ("fpcorr", dict(nx=7, ny=6, seed=17, outlet=1, amp=0.2,
                                                  mesh="multiblock")),
    # corrector passes with the least-squares gradient of the previous p'
    "fpcorr2d_14x9_lsq": ("fpcorr", dict(nx=14, ny=9, seed=19, outlet=1, amp=0.22, passes=4,
                                         gscheme=1)),
    # tight linear solves: the step's fields then differ only by reduction order
    "piso2d_32_tight": ("piso", dict(dim=2, nx=32, ny=32, re=1000, dt=0.005, warm=3, steps=2,
                                     psolver="bicgstab", mabs=0, mrel=1e-13, pabs=0, prel=1e-13)),
}


def run(cmd, params):
    with tempfile.TemporaryDirectory() as d:
        args = [DRIVER, cmd] + [f"{k}={v}" for k, v in params.items()] + [f"out={d}"]
        subprocess.run(args, check=True)
        arrays = {}
        for path in sorted(glob.glob(os.path.join(d, "*"))):
            name, ext = os.path.splitext(os.path.basename(path))
            arrays[name] = np.fromfile(path, dtype=np.float64 if ext == ".f64" else np.int64)
    arrays["params"] = np.array([f"{k}={v}" for k, v in params.items()])
    return arrays


def reference_cavity20():
    base = os.path.join(REF, "cases", "lid_driven_cavity", "results")
    rows = list(csv.DictReader(open(os.path.join(base, "fields.csv"))))
    res = list(csv.DictReader(open(os.path.join(base, "residuals.csv"))))
    col = lambda rs, k: np.array([float(r[k]) for r in rs])
    return dict(u=col(rows, "velocity_x"), v=col(rows, "velocity_y"), p=col(rows, "pressure"),
                hist_u=col(res, "u_residual"), hist_v=col(res, "v_residual"),
                hist_p=col(res, "p_residual"), hist_c=col(res, "continuity_residual"),
                hist_g=col(res, "global_mass_imbalance"))


CASES = ["lid_driven_cavity", "poiseuille_flow"]


def reference_cases():
    """The reference's case files (inputs) and committed results (fields.csv,
    residuals.csv, gzipped) of CASES, copied as data into cases/<name>/."""
    import gzip
    import shutil
    for name in CASES:
        src = os.path.join(REF, "cases", name)
        dst = os.path.join(HERE, "cases", name)
        os.makedirs(dst, exist_ok=True)
        for f in sorted(os.listdir(src)):
            if f.endswith(".json"):
                shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        for f in ("fields.csv", "residuals.csv"):
            with open(os.path.join(src, "results", f), "rb") as i, \
                    gzip.GzipFile(os.path.join(dst, f + ".gz"), "wb", mtime=0) as o:
                o.write(i.read())
        print("wrote cases/" + name)


# --- BiCGSTAB breakdown / restart systems (tests/unit/algebra/test_bicgstab_breakdown.cpp) ---

STORED = {"smooth": "p12_mesh004_thermal_smooth_a005_n32.txt",
          "rough": "p12_mesh004_thermal_rough_f03_n32.txt"}


def read_stored(name):
    """The bicg_capture text format of tests/data/linear_systems/: '#' comment
    lines, then rows nonzeros, row offsets, columns, values, rhs, guess."""
    toks = []
    with open(os.path.join(REF, "tests", "data", "linear_systems", name)) as f:
        for line in f:
            if line and line[0] != "#":
                toks.extend(line.split())
    rows, nnz = int(toks[0]), int(toks[1])
    k = 2
    rp = np.array([int(t) for t in toks[k:k + rows + 1]], np.int64)
    k += rows + 1
    col = np.array([int(t) for t in toks[k:k + nnz]], np.int64)
    k += nnz
    val = np.array([float(t) for t in toks[k:k + nnz]])
    k += nnz
    b = np.array([float(t) for t in toks[k:k + rows]])
    k += rows
    x0 = np.array([float(t) for t in toks[k:k + rows]])
    return rp, col, val, b, x0


def dense_csr(a):
    """SparseMatrixBuilder of the test's dense() helper: non-zero entries only."""
    a = np.asarray(a, dtype=np.float64)
    rp, col, val = [0], [], []
    for row in a:
        for j, v in enumerate(row):
            if v != 0.0:
                col.append(j)
                val.append(v)
        rp.append(len(col))
    return np.array(rp, np.int64), np.array(col, np.int64), np.array(val)


def convection_diffusion(n, scale):
    """The test's 1D upwind convection-diffusion system (cell Peclet 0.5)."""
    pe = 0.5
    rp, col, val = [0], [], []
    for i in range(n):
        if i > 0:
            col.append(i - 1)
            val.append(-(1.0 + pe) * scale)
        col.append(i)
        val.append((2.0 + pe) * scale)
        if i + 1 < n:
            col.append(i + 1)
            val.append(-1.0 * scale)
        rp.append(len(col))
    b = np.array([np.sin(0.1 * i) * scale for i in range(n)])
    return np.array(rp, np.int64), np.array(col, np.int64), np.array(val), b


def solve_reference(rp, col, val, b, x0, abs_tol, rel_tol, max_iter):
    with tempfile.TemporaryDirectory() as d:
        for name, a in (("rowptr.i64", rp), ("col.i64", col), ("val.f64", val), ("b.f64", b),
                        ("x0.f64", x0)):
            np.ascontiguousarray(a).tofile(os.path.join(d, name))
        out = os.path.join(d, "out")
        os.makedirs(out)
        subprocess.run([DRIVER, "bicgstab", f"in={d}", f"abs={abs_tol!r}", f"rel={rel_tol!r}",
                        f"max={max_iter}", f"out={out}"], check=True)
        read = lambda n, t: np.fromfile(os.path.join(out, n), dtype=t)
        return dict(x=read("bcg_x.f64", np.float64), hist=read("bcg_hist.f64", np.float64),
                    meta=read("bcg_meta.i64", np.int64), res=read("bcg_res.f64", np.float64))


def bicgstab_fixtures():
    """bicgstab_stored.npz: the two stored systems at the default settings
    (abs 1e-12, rel 1e-10, 1000 iterations). bicgstab_small.npz: the 3x3 /
    2x2 / 1x1 systems at the test's strict settings (abs 0, rel 1e-14, 50) and
    the convection-diffusion systems (n 60 at scales 1, 2^-100, 2^100 and
    n 200 at 1, 2^-40; abs 0, rel 1e-10, 500 / 2000). Keys: <case>_{rowptr,
    col, val, b, x0, x, hist, meta (status, iterations, restarts), res
    (initial, final), settings (abs, rel, max)}."""
    groups = {"stored": [], "small": []}
    for key, fname in STORED.items():
        rp, col, val, b, x0 = read_stored(fname)
        groups["stored"].append((key, rp, col, val, b, x0, (1e-12, 1e-10, 1000)))
    strict = (0.0, 1e-14, 50)
    hand = {
        "rotation": ([[0, 1], [-1, 0]], [1, 0]),
        "omega": ([[-2, -2, -2], [-2, -2, -1], [-2, -1, 0]], [1, 1, 0]),
        "rho_no_progress": ([[-2, -2, -2], [-2, -2, -1], [2, -2, -2]], [1, 0, 0]),
        "restart": ([[-2, -2, -2], [-2, -2, 2], [2, -2, -2]], [1, 0, 1]),
        "exact_1x1": ([[2.0]], [2.0]),
        "zero_rhs": ([[4, 1], [2, 3]], [0.0, 0.0]),
    }
    for key, (a, b) in hand.items():
        rp, col, val = dense_csr(a)
        b = np.array(b, dtype=np.float64)
        groups["small"].append((key, rp, col, val, b, np.zeros(len(b)), strict))
    rp, col, val = dense_csr([[4, 1], [2, 3]])
    groups["small"].append(("exact_guess", rp, col, val, np.array([7.0, 11.0]),
                            np.array([1.0, 3.0]), strict))
    for n, e, mx in ((60, 0, 500), (60, -100, 500), (60, 100, 500), (200, 0, 2000),
                     (200, -40, 2000)):
        rp, col, val, b = convection_diffusion(n, np.ldexp(1.0, e))
        tag = f"cd{n}_{'m' if e < 0 else 'p'}{abs(e)}"
        groups["small"].append((tag, rp, col, val, b, np.zeros(n), (0.0, 1e-10, mx)))
    for group, items in groups.items():
        arrays = {}
        for key, rp, col, val, b, x0, (a, r, m) in items:
            out = solve_reference(rp, col, val, b, x0, a, r, m)
            arrays.update({f"{key}_rowptr": rp, f"{key}_col": col, f"{key}_val": val,
                           f"{key}_b": b, f"{key}_x0": x0,
                           f"{key}_settings": np.array([a, r, float(m)])})
            arrays.update({f"{key}_{k}": v for k, v in out.items()})
        arrays["cases"] = np.array([it[0] for it in items])
        np.savez_compressed(os.path.join(HERE, f"bicgstab_{group}.npz"), **arrays)
        print(f"wrote bicgstab_{group}")


def main(names):
    if not os.path.exists(DRIVER):
        sys.exit("build oracle/_ref first: make -C oracle -f Makefile.ref")
    if names == ["bicgstab"]:
        bicgstab_fixtures()
        return
    for name in names or FIXTURES:
        cmd, params = FIXTURES[name]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(cmd, params))
        print("wrote", name)
    if not names:
        np.savez_compressed(os.path.join(HERE, "reference_cavity20.npz"), **reference_cavity20())
        print("wrote reference_cavity20")
        reference_cases()
        bicgstab_fixtures()


if __name__ == "__main__":
    main(sys.argv[1:])
