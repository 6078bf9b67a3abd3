"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/cfdhip.h declares (and nothing else under the cfdhip_ prefix), and the
host mirror validates settings like the reference. No device calls."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "cfdhip.h")
LIB = os.path.join(REPO, "cfdapp_amd", "libcfdhip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(cfdhip_[a-z0-9_]+)\s*\(", text))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_built():
    assert os.path.exists(LIB), "run __graft_entry__.build() / cfdapp_amd/build.sh"


def test_exports_match_header():
    decl = declared_symbols()
    assert len(decl) >= 30
    exported = {s for s in exported_symbols() if s.startswith("cfdhip_")}
    assert decl == exported, (decl - exported, exported - decl)


def test_ctypes_loads_every_symbol():
    from cfdapp_amd import _lib
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name)
    assert set(_lib.SIGNATURES) == declared_symbols()


def test_device_available_never_raises():
    import cfdapp_amd
    assert cfdapp_amd.device_available() in (True, False)


def test_settings_validation_mirrors_reference():
    import cfdapp_amd as cfd
    with pytest.raises(ValueError):
        cfd.LinearSolverSettings(absolute_tolerance=-1.0)
    with pytest.raises(ValueError):
        cfd.LinearSolverSettings(0.0, 0.0)
    with pytest.raises(ValueError):
        cfd.LinearSolverSettings(max_iterations=0)
    s = cfd.LinearSolverSettings(0.0, 1e-8, 100, "CG", "Jacobi").c()
    assert (s.type, s.preconditioner, s.max_iterations) == (0, 1, 100)
    assert s.inner_tolerance == 0.0
    for bad in (-0.1, 1.0, 2.0):
        with pytest.raises(ValueError):
            cfd.LinearSolverSettings(inner_tolerance=bad)
    s = cfd.LinearSolverSettings(0.0, 1e-8, 100, "CG", "Jacobi", precision="mixed",
                                 inner_tolerance=0.05).c()
    assert (s.precision, s.inner_tolerance) == (1, 0.05)


def test_settings_struct_layout_matches_header():
    """cfdhip_solver_settings as the header declares it: the ctypes mirror's
    size and the offset of the last field (a double after six int32)."""
    import ctypes

    from cfdapp_amd import _lib
    S = _lib.SolverSettings
    assert ctypes.sizeof(S) == 56
    assert S.inner_tolerance.offset == 48
    assert S.graph.offset == 44


def test_no_device_factories_return_none():
    import cfdapp_amd as cfd
    if cfd.device_available():
        pytest.skip("device present")
    assert cfd.make_gpu_cg(cfd.LinearSolverSettings()) is None
    assert cfd.make_gpu_bicgstab(cfd.LinearSolverSettings()) is None
