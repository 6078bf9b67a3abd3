"""Writes the golden fixtures in this directory from the compiled reference.

Runs ``oracle/_ref/ref_driver`` (the reference's own CPU sources compiled by
``oracle/Makefile.ref``; needs /root/reference, so it runs in the build
container only) and packs each command's raw arrays into ``<name>.npz``. Also
converts the reference's committed 20x20 cavity results
(``cases/lid_driven_cavity/results/{fields,residuals}.csv``) into
``reference_cavity20.npz``, and copies the case files and results of the
reference's lid_driven_cavity and poiseuille_flow cases into ``cases/``.

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


def main(names):
    if not os.path.exists(DRIVER):
        sys.exit("build oracle/_ref first: make -C oracle -f Makefile.ref")
    for name in names or FIXTURES:
        cmd, params = FIXTURES[name]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(cmd, params))
        print("wrote", name)
    if not names:
        np.savez_compressed(os.path.join(HERE, "reference_cavity20.npz"), **reference_cavity20())
        print("wrote reference_cavity20")
        reference_cases()


if __name__ == "__main__":
    main(sys.argv[1:])
