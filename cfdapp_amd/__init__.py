"""MI355X (gfx950) SIMPLE pressure-velocity path: host-side mirror of the
reference's linear-solver and GPU-discretization interfaces over libcfdhip.

Names follow the reference's C++ API:

* ``device_available()``                  -- ``cfd::gpu::cudaAvailable``
  (include/cfd/gpu/GPUBackend.hpp:15)
* ``LinearSolverSettings`` / ``SolverResult`` -- ``cfd::algebra`` types of the same
  name (include/cfd/algebra/LinearSolver.hpp:103-166)
* ``make_gpu_cg`` / ``make_gpu_bicgstab``  -- ``cfd::gpu::makeGpuCG/makeGpuBiCGSTAB``
  (include/cfd/gpu/GpuLinearSolver.hpp:48,52): ``None`` when no device
* ``GpuSimpleDiscretization``              -- the resident SIMPLE stage interface
  (include/cfd/gpu/GpuSimpleDiscretization.hpp:56-262) on a Cartesian case

All arithmetic happens in libcfdhip.so on the device; this module only moves
host arrays across the C ABI.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BREAKDOWN, CONVERGED, FIELDS, FIXED_GRADIENT, FIXED_VALUE, FP64, INLET,
                   INVALID_SYSTEM, MAX_ITERATIONS, MIXED, MOVING_WALL, NONFINITE_INPUT,
                   NONFINITE_RESIDUAL, OUTLET, PREC_JACOBI, PREC_LEARNED, PREC_NONE, SOLVER_BICGSTAB,
                   SOLVER_CG, SOLVER_FCG, STATUS_NAMES, WALL)

__all__ = ["device_available", "LinearSolverSettings", "SolverResult", "make_gpu_cg",
           "make_gpu_bicgstab", "GpuSimpleDiscretization", "CfdHipError", "UnetNet",
           "unet_param_count", "load_learned_weights", "load_learned_options", "WEIGHTS_PATH", "BatchSolver",
           "FaceMeshPressure"]


class CfdHipError(RuntimeError):
    """A failing libcfdhip call (code + the context's last error message)."""

    def __init__(self, code, message):
        super().__init__(f"libcfdhip error {code}: {message}")
        self.code = code


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _pf(a):
    return None if a is None else a.ctypes.data_as(_lib.P_F64)


def _pi(a):
    return None if a is None else a.ctypes.data_as(_lib.P_I64)


def device_available() -> bool:
    """True when a gfx950 device is usable (never raises)."""
    try:
        return bool(_lib.load().cfdhip_device_available())
    except OSError:
        return False


@dataclass
class LinearSolverSettings:
    """cfd::algebra::LinearSolverSettings (LinearSolver.hpp:136-166) + device knobs."""
    absolute_tolerance: float = 1e-12
    relative_tolerance: float = 1e-10
    max_iterations: int = 1000
    type: str = "BiCGSTAB"          # "CG", "BiCGSTAB", "FCG"
    preconditioner: str = "None"    # "None", "Jacobi", "Learned"
    precision: str = "fp64"         # "fp64" or "mixed"
    check_interval: int = 0
    stagnation_window: int = 0
    graph: bool = False             # replay iteration chunks from a captured HIP graph
    inner_tolerance: float = 0.0    # mixed precision: FP32 reduction between FP64 residuals

    def __post_init__(self):
        # LinearSolver's constructor validation (src/algebra/LinearSolver.cpp:9-32)
        if self.absolute_tolerance < 0.0:
            raise ValueError("LinearSolverSettings: absoluteTolerance must be >= 0")
        if self.relative_tolerance < 0.0:
            raise ValueError("LinearSolverSettings: relativeTolerance must be >= 0")
        if self.absolute_tolerance == 0.0 and self.relative_tolerance == 0.0:
            raise ValueError("LinearSolverSettings: at least one of absoluteTolerance/"
                             "relativeTolerance must be > 0")
        if self.max_iterations < 1:
            raise ValueError("LinearSolverSettings: maxIterations must be >= 1")
        if self.type not in ("CG", "BiCGSTAB", "FCG"):
            raise ValueError(f"unknown solver type {self.type}")
        if self.preconditioner not in ("None", "Jacobi", "Learned"):
            raise ValueError(f"unknown preconditioner {self.preconditioner}")
        if not 0.0 <= self.inner_tolerance < 1.0:
            raise ValueError("LinearSolverSettings: inner_tolerance must be in [0, 1)")

    def c(self) -> _lib.SolverSettings:
        s = _lib.SolverSettings()
        s.absolute_tolerance = self.absolute_tolerance
        s.relative_tolerance = self.relative_tolerance
        s.max_iterations = int(self.max_iterations)
        s.type = {"CG": SOLVER_CG, "BiCGSTAB": SOLVER_BICGSTAB, "FCG": SOLVER_FCG}[self.type]
        s.preconditioner = {"None": PREC_NONE, "Jacobi": PREC_JACOBI,
                            "Learned": PREC_LEARNED}[self.preconditioner]
        s.precision = MIXED if self.precision == "mixed" else FP64
        s.check_interval = int(self.check_interval)
        s.stagnation_window = int(self.stagnation_window)
        s.graph = 1 if self.graph else 0
        s.inner_tolerance = float(self.inner_tolerance)
        return s


@dataclass
class SolverResult:
    """cfd::algebra::SolverResult (LinearSolver.hpp:103-134); backend is always GPU."""
    solution: np.ndarray | None
    status: int
    iterations: int
    initial_residual: float
    final_residual: float
    true_residual: float
    restarts: int = 0
    device_ms: float = 0.0
    fallback_used: bool = False
    residual_history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    backend_used: str = "GPU"

    @property
    def converged(self) -> bool:
        return self.status == CONVERGED

    @property
    def status_name(self) -> str:
        return STATUS_NAMES[self.status]


class _Context:
    def __init__(self, device=0):
        self.lib = _lib.load()
        self.h = self.lib.cfdhip_create(device)
        if not self.h:
            raise CfdHipError(_lib.ERR_NO_DEVICE, "no usable gfx950 device")

    def close(self):
        if getattr(self, "h", None):
            self.lib.cfdhip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def call(self, name, *args):
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != _lib.OK:
            msg = self.lib.cfdhip_last_error(self.h)
            raise CfdHipError(rc, msg.decode() if msg else "")
        return rc

    def history(self):
        n = ctypes.c_int64(0)
        self.call("cfdhip_solve_history", None, 0, ctypes.byref(n))
        h = np.zeros(n.value)
        self.call("cfdhip_solve_history", _pf(h), n.value, ctypes.byref(n))
        return h


def _result(r: _lib.SolveResult, x, hist) -> SolverResult:
    return SolverResult(solution=x, status=r.status, iterations=r.iterations,
                        initial_residual=r.initial_residual, final_residual=r.final_residual,
                        true_residual=r.true_residual, restarts=r.restarts,
                        device_ms=r.device_ms, fallback_used=bool(r.fallback_used),
                        residual_history=hist)


class _HostCall:
    """The C callback of a host preconditioner: z = M r into the library's host
    arrays. No exception may cross the C boundary, so one is kept here and the
    library told to stop (nonzero return); solve() raises it."""

    def __init__(self, preconditioner):
        self.preconditioner = preconditioner
        self.error = None
        self.callback = _lib.HOST_PRECONDITIONER(self._apply)

    def _apply(self, r, z, n, _user):
        try:
            self.preconditioner.apply(np.ctypeslib.as_array(r, (n,)),
                                      np.ctypeslib.as_array(z, (n,)))
            return 0
        except BaseException as e:  # noqa: BLE001 - raised again by solve()
            self.error = e
            return 1


class GpuLinearSolver:
    """LinearSolver::solve(system, x0) on the device for a host CSR system.

    With a host ``preconditioner`` (an object with ``build(rowptr, col, val)``
    and ``apply(r, z)`` writing z = M r, the shape of
    cfd::algebra::Preconditioner, include/cfd/algebra/Preconditioner.hpp:11-17)
    the device CG / BiCGSTAB applies it through a host round trip per
    application (cuda/kernels/GpuLinearSolverCuda.cpp:103-118), in place of the
    settings' preconditioner. build() runs on the system before the solve; a
    ValueError from it makes the result INVALID_SYSTEM (the reference's
    try/catch of InvalidArgumentError around Preconditioner::build,
    cuda/kernels/GpuLinearSolverCuda.cpp:130-143), and an exception from apply()
    leaves solve()."""

    def __init__(self, settings: LinearSolverSettings, ctx: _Context, preconditioner=None):
        self.settings = settings
        self._ctx = ctx
        self.preconditioner = preconditioner
        self._host = None
        if preconditioner is not None:
            # the callback holds the preconditioner, not this solver: no cycle
            # keeps the context alive after the solver is dropped
            self._host = _HostCall(preconditioner)
            self._ctx.call("cfdhip_set_host_preconditioner",
                           ctypes.cast(self._host.callback, ctypes.c_void_p), None)

    def solve(self, rowptr, col, val, b, x0=None) -> SolverResult:
        rowptr, col, val, b = _i64(rowptr), _i64(col), _f64(val), _f64(b)
        n = len(rowptr) - 1
        x = np.empty(n)
        r = _lib.SolveResult()
        s = self.settings.c()
        if self.preconditioner is not None:
            try:
                self.preconditioner.build(rowptr, col, val)
            except ValueError:  # InvalidArgumentError in the reference: InvalidSystem
                x = np.array(_f64(x0)) if x0 is not None else np.zeros(n)
                return SolverResult(solution=x, status=INVALID_SYSTEM, iterations=0,
                                    initial_residual=0.0, final_residual=0.0,
                                    true_residual=float("nan"))
            s.preconditioner = _lib.PREC_HOST
            self._host.error = None
        try:
            self._ctx.call("cfdhip_csr_solve", ctypes.byref(s), n, _pi(rowptr), _pi(col),
                           _pf(val), _pf(b), _pf(None if x0 is None else _f64(x0)), _pf(x),
                           ctypes.byref(r))
        except CfdHipError:
            if self._host is not None and self._host.error is not None:
                err, self._host.error = self._host.error, None
                raise err
            raise
        return _result(r, x, self._ctx.history())

    def spmv(self, rowptr, col, val, x):
        rowptr, col, val, x = _i64(rowptr), _i64(col), _f64(val), _f64(x)
        y = np.empty(len(rowptr) - 1)
        self._ctx.call("cfdhip_csr_spmv", len(rowptr) - 1, _pi(rowptr), _pi(col), _pf(val),
                       _pf(x), _pf(y))
        return y


def _make(settings, kind, preconditioner):
    if not device_available():
        return None
    s = LinearSolverSettings(**{**settings.__dict__, "type": kind})
    return GpuLinearSolver(s, _Context(), preconditioner)


def make_gpu_cg(settings: LinearSolverSettings, preconditioner=None):
    """cfd::gpu::makeGpuCG: a device CG solver, or None when no device. A host
    preconditioner object is applied through the host (GpuLinearSolver)."""
    return _make(settings, "CG", preconditioner)


def make_gpu_bicgstab(settings: LinearSolverSettings, preconditioner=None):
    """cfd::gpu::makeGpuBiCGSTAB: a device BiCGSTAB solver, or None when no
    device. A host preconditioner object is applied through the host."""
    return _make(settings, "BiCGSTAB", preconditioner)


class GpuSimpleDiscretization:
    """The resident SIMPLE stages on a uniform Cartesian case.

    ``prepare`` mirrors GpuSimpleDiscretization::prepare (all or nothing):
    it returns (False, reason) for configurations outside the structured path.
    Patches: 2D left, right, bottom, top; 3D xmin, xmax, ymin, ymax, zmin, zmax.
    """

    def __init__(self, device=0):
        self._ctx = _Context(device)
        self.dim = 0
        self.n_cells = self.n_faces = self.n_patches = 0
        self._fixed_value = set()  # patches with a FixedValue pressure condition
        self.pin = -1              # p' reference cell of the last assembly (-1: none)

    def close(self):
        self._ctx.close()

    # -- plan ---------------------------------------------------------------
    def prepare(self, dim, nx, ny, nz=1, lx=1.0, ly=1.0, lz=1.0, velocity=None, pressure=None,
                density=1.0, viscosity=0.01):
        """velocity: per patch (kind, (vx, vy, vz)); pressure: per patch (kind, value)."""
        try:
            self._ctx.call("cfdhip_mesh_cartesian", dim, nx, ny, nz, lx, ly, lz)
        except CfdHipError as e:
            if e.code == _lib.ERR_UNSUPPORTED:
                return False, str(e)
            raise
        sizes = np.zeros(4, np.int64)
        self._ctx.call("cfdhip_mesh_sizes", _pi(sizes))
        self.n_cells, self.n_faces, self.n_patches, self.dim = (int(v) for v in sizes)
        self.shape = (nx, ny, nz if dim == 3 else 1)
        for p in range(self.n_patches):
            vk, vv = (velocity[p] if velocity else (WALL, (0.0, 0.0, 0.0)))
            pk, pv = (pressure[p] if pressure else (FIXED_GRADIENT, 0.0))
            self.set_boundary(p, vk, vv, pk, pv)
        self._ctx.call("cfdhip_set_fluid", density, viscosity)
        return True, ""

    @classmethod
    def cavity(cls, dim, nx, ny, nz=1, lid=1.0, viscosity=0.01, device=0):
        """Lid-driven cavity: lid on 'top' (2D) / 'ymax' (3D), other walls no-slip,
        zero-gradient pressure everywhere (cases/lid_driven_cavity/boundaries.json)."""
        d = cls(device)
        npatch = 2 * dim
        vel = [(MOVING_WALL, (lid, 0.0, 0.0)) if p == 3 else (WALL, (0.0, 0.0, 0.0))
               for p in range(npatch)]
        ok, why = d.prepare(dim, nx, ny, nz, velocity=vel, viscosity=viscosity)
        if not ok:
            raise CfdHipError(_lib.ERR_UNSUPPORTED, why)
        return d

    def cell_centroids(self):
        """(n_cells, 3) centroids and (n_cells,) volumes in cell order."""
        c = np.zeros(3 * self.n_cells)
        vol = np.zeros(self.n_cells)
        self._ctx.call("cfdhip_mesh_cells", _pf(c), _pf(vol))
        return c.reshape(-1, 3), vol

    def set_boundary(self, patch, vkind, vvalue, pkind, pvalue):
        v = _f64(list(vvalue) + [0.0] * (3 - len(vvalue)))
        self._ctx.call("cfdhip_set_boundary", patch, vkind, _pf(v), pkind, float(pvalue))
        if pkind == FIXED_VALUE:
            self._fixed_value.add(patch)
        else:
            self._fixed_value.discard(patch)

    def mesh_faces(self):
        o = np.zeros(self.n_faces, np.int64)
        nb = np.zeros(self.n_faces, np.int64)
        self._ctx.call("cfdhip_mesh_faces", _pi(o), _pi(nb))
        return o, nb

    # -- fields -------------------------------------------------------------
    def _count(self, name):
        return self.n_faces if name in FACE_FIELDS else self.n_cells

    def upload(self, name, values):
        a = _f64(values)
        self._ctx.call("cfdhip_field_upload", FIELDS[name], _pf(a), len(a))

    def download(self, name):
        a = np.empty(self._count(name))
        self._ctx.call("cfdhip_field_download", FIELDS[name], _pf(a), len(a))
        return a

    def piso_run(self, steps, dt, momentum: LinearSolverSettings | None = None,
                 pressure: LinearSolverSettings | None = None, reference_cell=0):
        """PISO::solveTimeStep `steps` times on the resident state (cfdhip_piso_run).
        Defaults are the reference's: BiCGSTAB for both systems. Returns
        {steps: [{status, status_name, max_cfl, continuity, mass_imbalance,
        momentum_iterations, pressure_iterations}], taken}."""
        momentum = momentum or LinearSolverSettings(1e-10, 1e-8, 500, "BiCGSTAB")
        pressure = pressure or LinearSolverSettings(1e-10, 1e-8, 20000, "BiCGSTAB")
        meta = np.zeros(6 * int(steps))
        taken = ctypes.c_int64(0)
        ms, ps = momentum.c(), pressure.c()
        self._ctx.call("cfdhip_piso_run", float(dt), int(steps), int(reference_cell),
                       ctypes.byref(ms), ctypes.byref(ps), _pf(meta), ctypes.byref(taken))
        names = ["Converged", "MomentumFailure", "PressureCorrectionFailure", "NonFiniteState",
                 "InvalidConfiguration"]
        rows = meta.reshape(-1, 6)[:taken.value]
        return dict(taken=taken.value, steps=[
            dict(status=int(r[0]), status_name=names[int(r[0])], max_cfl=float(r[1]),
                 continuity=float(r[2]), mass_imbalance=float(r[3]),
                 momentum_iterations=int(r[4]), pressure_iterations=int(r[5])) for r in rows])

    def field_device(self, name):
        """Device address of a resident field (borrowed from the context)."""
        ptr = ctypes.c_void_p()
        self._ctx.call("cfdhip_field_device", FIELDS[name], ctypes.byref(ptr))
        return ptr.value

    def upload_initial_state(self, velocity, pressure, mass_flux=None, viscosity=None):
        for k, name in enumerate("UVW"[:self.dim]):
            self.upload(name, velocity[k])
        self.upload("P", pressure)
        if mass_flux is not None:
            self.upload("FLUX", mass_flux)
        if viscosity is not None:
            self.upload("VISCOSITY", viscosity)

    def download_velocity(self):
        return [self.download(n) for n in "UVW"[:self.dim]]

    def download_pressure(self):
        return self.download("P")

    def download_mass_flux(self):
        return self.download("FLUX")

    # -- stages -------------------------------------------------------------
    def gradient(self, name="P", correction=False):
        self._ctx.call("cfdhip_gradient", FIELDS[name], 1 if correction else 0)
        return [self.download(n) for n in ("GRAD_X", "GRAD_Y", "GRAD_Z")[:self.dim]]

    def mass_flux(self):
        self._ctx.call("cfdhip_mass_flux")
        return self.download("FLUX")

    def assemble_momentum_resident(self, alpha):
        self._ctx.call("cfdhip_assemble_momentum", float(alpha))

    def assemble_momentum_component(self, component, alpha, convection_scheme=0,
                                    non_orthogonal=False, previous_component=None):
        """assembleMomentum / assembleMomentumResident for one component
        (include/cfd/gpu/GpuSimpleDiscretization.hpp:162-164,181-182); with
        previous_component the relaxation's phi_old is the caller's field."""
        prev = None if previous_component is None else _f64(previous_component)
        if prev is not None and prev.size != self.n_cells:
            raise ValueError("assembleMomentum: previousComponent size does not match the mesh")
        self._ctx.call("cfdhip_assemble_momentum_component", int(component), float(alpha),
                       int(convection_scheme), 1 if non_orthogonal else 0, _pf(prev))

    def assemble_transient_momentum(self, dt):
        """The PISO predictor's momentum system (TransientMomentum.cpp)."""
        self._ctx.call("cfdhip_assemble_transient_momentum", float(dt))

    def solve_momentum_resident(self, component, settings: LinearSolverSettings) -> SolverResult:
        r = _lib.SolveResult()
        s = settings.c()
        self._ctx.call("cfdhip_solve_momentum", component, ctypes.byref(s), ctypes.byref(r))
        return _result(r, None, self._ctx.history())

    def compute_response_coefficients(self):
        self._ctx.call("cfdhip_response")

    def compute_predicted_face_flux(self, rhie_chow, alpha):
        self._ctx.call("cfdhip_predicted_flux", 1 if rhie_chow else 0, float(alpha))

    def assemble_pressure_correction_resident(self, reference_cell=0):
        self._ctx.call("cfdhip_assemble_pressure", int(reference_cell))
        # the reference row is pinned only without a FixedValue pressure patch
        # (PressureCorrectionEquation.cpp:221-228)
        self.pin = -1 if self._fixed_value else int(reference_cell)

    def assemble_compressible_pressure_correction_resident(self, face_density, pressure_absolute,
                                                           temperature, gas_constant,
                                                           pseudo_time_step, reference_cell=0):
        """assembleCompressiblePressureCorrection
        (src/compressible/CompressiblePressureCorrection.cpp:20-80) from the
        resident FLUX_STAR and RESPONSE: per-face density, ideal-gas
        compressibility diagonal V / dtau / (R T)."""
        rf, pa, t = _f64(face_density), _f64(pressure_absolute), _f64(temperature)
        if rf.size != self.n_faces or pa.size != self.n_cells or t.size != self.n_cells:
            raise ValueError("assembleCompressiblePressureCorrection: field size does not match "
                             "the mesh")
        P = ctypes.POINTER(ctypes.c_double)
        self._ctx.call("cfdhip_assemble_compressible_pressure", int(reference_cell),
                       rf.ctypes.data_as(P), pa.ctypes.data_as(P), t.ctypes.data_as(P),
                       float(gas_constant), float(pseudo_time_step))
        self.pin = -1 if self._fixed_value else int(reference_cell)

    def solve_pressure_correction_resident(self, settings: LinearSolverSettings) -> SolverResult:
        r = _lib.SolveResult()
        s = settings.c()
        self._ctx.call("cfdhip_solve_pressure", ctypes.byref(s), ctypes.byref(r))
        return _result(r, None, self._ctx.history())

    def solve_pressure_rhs(self, settings: LinearSolverSettings, b=None, download=True):
        """Stand-alone p' solve on the resident operator with a caller RHS (x0 = 0)."""
        r = _lib.SolveResult()
        s = settings.c()
        x = np.empty(self.n_cells) if download else None
        self._ctx.call("cfdhip_pressure_solve_rhs", ctypes.byref(s),
                       _pf(None if b is None else _f64(b)), _pf(x), ctypes.byref(r))
        return _result(r, x, self._ctx.history())

    def update_pressure(self, alpha):
        self._ctx.call("cfdhip_update_pressure", float(alpha))

    def correct(self):
        self._ctx.call("cfdhip_correct")

    def correct_velocity(self, resident=True):
        """correctVelocityResident (resident=True: into U/V/W) or
        correctVelocity (returns the corrected components, U/V/W unchanged)."""
        if resident:
            self._ctx.call("cfdhip_correct_velocity", None)
            return None
        out = np.empty(3 * self.n_cells)
        self._ctx.call("cfdhip_correct_velocity", _pf(out))
        return [out[k * self.n_cells:(k + 1) * self.n_cells] for k in range(self.dim)]

    def correct_flux(self, resident=True):
        """correctFaceMassFluxResident (into FLUX) or correctFaceMassFlux
        (returns the corrected flux, FLUX unchanged)."""
        if resident:
            self._ctx.call("cfdhip_correct_flux", None)
            return None
        out = np.empty(self.n_faces)
        self._ctx.call("cfdhip_correct_flux", _pf(out))
        return out

    def all_finite(self, *names):
        """SIMPLE's resident finiteness guards: True when no listed field
        holds a NaN or an infinity."""
        ids = (ctypes.c_int * max(1, len(names)))(*[FIELDS[n] for n in names])
        ok = ctypes.c_int(0)
        self._ctx.call("cfdhip_fields_all_finite", ids, len(names), ctypes.byref(ok))
        return bool(ok.value)

    def set_density(self, density):
        self._ctx.call("cfdhip_set_density", float(density))

    def resident_solve_bytes(self):
        b = ctypes.c_int64(0)
        self._ctx.call("cfdhip_resident_bytes", ctypes.byref(b))
        return int(b.value)

    def continuity(self):
        out = np.zeros(2)
        self._ctx.call("cfdhip_continuity", _pf(out))
        return self.download("IMBALANCE"), float(out[0]), float(out[1])

    def pressure_spmv(self, x):
        x = _f64(x)
        y = np.empty(self.n_cells)
        self._ctx.call("cfdhip_pressure_spmv", _pf(x), _pf(y))
        return y

    def _csr(self, name):
        n = self.n_cells
        cap = (2 * self.dim + 1) * n
        rp = np.zeros(n + 1, np.int64)
        col = np.zeros(cap, np.int64)
        val = np.zeros(cap)
        nnz = ctypes.c_int64(0)
        self._ctx.call(name, _pi(rp), _pi(col), _pf(val), cap, ctypes.byref(nnz))
        return rp, col[:nnz.value], val[:nnz.value]

    def pressure_csr(self):
        return self._csr("cfdhip_pressure_csr")

    def momentum_csr(self):
        return self._csr("cfdhip_momentum_csr")

    def set_timing(self, enabled):
        self._ctx.call("cfdhip_set_timing", 1 if enabled else 0)

    def preconditioner_timing(self):
        """Per-layer (ms, launches) of the learned preconditioner while timing is on."""
        ms = np.zeros(16)
        calls = np.zeros(16, np.int64)
        self._ctx.call("cfdhip_preconditioner_timing", _pf(ms), _pi(calls))
        return ms, calls

    def kernel_timing(self):
        out = np.zeros(3)
        self._ctx.call("cfdhip_kernel_timing", _pf(out))
        return dict(launches=out[0], mean_us=out[1], iterations=out[2])

    def set_learned_weights(self, weights=None, options=None):
        """Loads the learned p' preconditioner (FP32 master weights and how they
        are applied; the trained set shipped with the package by default)."""
        if weights is None:
            weights = load_learned_weights()
            if options is None:
                options = load_learned_options()
        self.set_learned_options(options)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        self._ctx.call("cfdhip_set_learned_weights", w.ctypes.data_as(_lib.P_F32), w.size)

    def learned_channels(self):
        """Channels stored per pixel by each forward tensor of the learned
        preconditioner's inference path (dict by tensor name; 0: not stored)."""
        out = (ctypes.c_int32 * 17)()
        self._ctx.call("cfdhip_learned_channels", out, 17)
        return dict(zip(UnetNet.TENSORS[:17], [int(v) for v in out]))

    def set_learned_options(self, options=None):
        """Padding and coarse tail of the learned preconditioner (a dict, see
        _unet_options; None: zero padding, no tail)."""
        o = _unet_options(options)
        self._ctx.call("cfdhip_set_learned_options", ctypes.byref(o))

    def harvest(self, settings: LinearSolverSettings, stride, max_pairs, h, w, r_out, e_out,
                ce_out=None, cn_out=None, b=None):
        """Training pairs from the assembled p' system (see cfdhip_pressure_harvest).
        r_out / e_out / ce_out / cn_out are device buffers (torch tensors or raw
        addresses) of max_pairs*h*w and h*w floats. Returns (pairs, den, result)."""
        den = np.zeros(max(1, int(max_pairs)), dtype=np.float32)
        pairs = ctypes.c_int64(0)
        _sync_producer()
        r = _lib.SolveResult()
        cs = settings.c()
        self._ctx.call("cfdhip_pressure_harvest", ctypes.byref(cs),
                       _pf(None if b is None else _f64(b)), int(stride), int(max_pairs), int(h),
                       int(w), _dev(r_out), _dev(e_out), _dev(ce_out), _dev(cn_out),
                       den.ctypes.data_as(_lib.P_F32), ctypes.byref(pairs), ctypes.byref(r))
        return pairs.value, den[:pairs.value].copy(), _result(r, None, self._ctx.history())

    def harvest_noise(self, seed, count, h, w, r_out, e_out):
        """White-noise pairs on the assembled p' operator (see
        cfdhip_pressure_harvest_noise). Returns (pairs, den)."""
        den = np.zeros(max(1, int(count)), dtype=np.float32)
        pairs = ctypes.c_int64(0)
        _sync_producer()
        self._ctx.call("cfdhip_pressure_harvest_noise", int(seed), int(count), int(h), int(w),
                       _dev(r_out), _dev(e_out), den.ctypes.data_as(_lib.P_F32),
                       ctypes.byref(pairs))
        return pairs.value, den[:pairs.value].copy()

    def simple_run(self, iterations, alpha_u=0.7, alpha_p=0.3, velocity_tol=1e-6,
                   pressure_tol=1e-6, continuity_tol=1e-6, rhie_chow=True, reference_cell=0,
                   momentum: LinearSolverSettings | None = None,
                   pressure: LinearSolverSettings | None = None, use_graph=False):
        """SIMPLE::solve on the device from the uploaded state (u, p)."""
        momentum = momentum or LinearSolverSettings(1e-10, 1e-8, 500, "BiCGSTAB")
        pressure = pressure or LinearSolverSettings(0.0, 1e-8, 20000, "CG", "Jacobi")
        s = _f64([iterations, alpha_u, alpha_p, velocity_tol, pressure_tol, continuity_tol,
                  1.0 if rhie_chow else 0.0, reference_cell, 1.0 if use_graph else 0.0, 0, 0])
        meta = np.zeros(6)
        cap = int(iterations)
        hist = np.zeros(5 * cap)
        ms, ps = momentum.c(), pressure.c()
        self._ctx.call("cfdhip_simple_run", _pf(s), ctypes.byref(ms), ctypes.byref(ps),
                       _pf(meta), _pf(hist), cap)
        it = int(meta[1])
        h = hist.reshape(5, cap)[:, :it]
        return dict(status=int(meta[0]), iterations=it, momentum_iterations=int(meta[2]),
                    pressure_iterations=int(meta[3]), final_continuity=meta[4],
                    global_imbalance=meta[5],
                    hist=dict(u=h[0], v=h[1], w=h[2] if self.dim == 3 else h[2][:0], p=h[3],
                              c=h[4]))


class FaceMeshPressure:
    """The p' system on a general face-list mesh (cfdhip_fmesh_*): non-orthogonal
    and multiblock meshes, with the over-relaxed coefficient and the explicit
    non-orthogonal flux (assembleGeometricPressureCorrection,
    PressureCorrectionEquation.cpp:181-339), solved by CG / BiCGSTAB on the CSR
    operator, and correctFaceMassFlux. The mesh arrays are those of the
    reference's Mesh: owner, neighbour (-1 on boundary faces), face centroids
    and area vectors (n_faces x 3), cell centroids (n_cells x 3), each cell's
    face list (offsets + faces) and a FixedValue-pressure flag per face."""

    def __init__(self, dim, owner, neighbor, face_centroid, area_vector, cell_centroid,
                 cell_face_offsets, cell_faces, fixed_value, device=0):
        self._ctx = _Context(device)
        self.n_cells = len(cell_face_offsets) - 1
        self.n_faces = len(owner)
        self._keep = [_i64(owner), _i64(neighbor), _f64(face_centroid).reshape(-1),
                      _f64(area_vector).reshape(-1), _f64(cell_centroid).reshape(-1),
                      _i64(cell_face_offsets), _i64(cell_faces),
                      np.ascontiguousarray(fixed_value, dtype=np.int32)]
        k = self._keep
        self._ctx.call("cfdhip_fmesh_set", int(dim), self.n_cells, self.n_faces, _pi(k[0]),
                       _pi(k[1]), _pf(k[2]), _pf(k[3]), _pf(k[4]), _pi(k[5]), _pi(k[6]),
                       k[7].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))

    def close(self):
        self._ctx.close()

    def assemble(self, flux_star, du, dv, dw=None, density=1.0, face_density=None,
                 non_orthogonal=True, grad_previous=None, extra_diagonal=None,
                 reference_cell=0):
        """grad_previous: (n_cells, 3) gradient of the previous p' (explicit term)."""
        held = []

        def ptr(a):
            if a is None:
                return None
            a = _f64(a).reshape(-1)
            held.append(a)
            return _pf(a)

        self._ctx.call("cfdhip_fmesh_assemble_pressure", ptr(flux_star), ptr(face_density),
                       float(density), ptr(du), ptr(dv), ptr(dw), 1 if non_orthogonal else 0,
                       ptr(grad_previous), ptr(extra_diagonal), int(reference_cell))

    def system(self):
        nnz = ctypes.c_int64(0)
        self._ctx.call("cfdhip_fmesh_pressure_system", None, None, None, 0, ctypes.byref(nnz),
                       None, None, None)
        rp = np.zeros(self.n_cells + 1, np.int64)
        col = np.zeros(max(nnz.value, 1), np.int64)
        val = np.zeros(max(nnz.value, 1))
        rhs, coeff, expl = np.zeros(self.n_cells), np.zeros(self.n_faces), np.zeros(self.n_faces)
        self._ctx.call("cfdhip_fmesh_pressure_system", _pi(rp), _pi(col), _pf(val), nnz.value,
                       ctypes.byref(nnz), _pf(rhs), _pf(coeff), _pf(expl))
        return (rp, col[:nnz.value], val[:nnz.value]), rhs, coeff, expl

    def solve(self, settings: "LinearSolverSettings"):
        x = np.zeros(self.n_cells)
        r = _lib.SolveResult()
        s = settings.c()
        self._ctx.call("cfdhip_fmesh_solve_pressure", ctypes.byref(s), _pf(x), ctypes.byref(r))
        return _result(r, x, self._ctx.history())

    def correct_flux(self, pprime=None):
        F = np.zeros(self.n_faces)
        p = None if pprime is None else _f64(pprime)
        self._ctx.call("cfdhip_fmesh_correct_flux", None if p is None else _pf(p), _pf(F))
        return F

    def set_cell_volumes(self, volume):
        """Cell::volume() of every cell: the gradients need it."""
        v = _f64(volume)
        if v.size != self.n_cells:
            raise ValueError("set_cell_volumes: one volume per cell")
        self._ctx.call("cfdhip_fmesh_set_cell_volumes", _pf(v))

    GRADIENT_SCHEMES = {"green_gauss": 0, "least_squares": 1}

    def gradient(self, field, bc_kind=None, bc_value=None, scheme="green_gauss"):
        """cfd::discretization::gradient of a cell field (n_cells, 3): bc_kind per
        face (1 FixedValue, 0 FixedGradient; boundary faces) with bc_value (the
        value / the gradient); None: the p' conditions of the explicit term."""
        phi = _f64(field)
        g = np.zeros((self.n_cells, 3))
        kind = None if bc_kind is None else np.ascontiguousarray(bc_kind, dtype=np.int32)
        val = None if bc_value is None else _f64(bc_value)
        self._ctx.call("cfdhip_fmesh_gradient", _pf(phi),
                       None if kind is None else kind.ctypes.data_as(
                           ctypes.POINTER(ctypes.c_int32)),
                       None if val is None else _pf(val), self.GRADIENT_SCHEMES[scheme],
                       _pf(g))
        return g

    def assemble_pass(self, previous=None, scheme="green_gauss"):
        """A corrector pass's assembly: the last assemble()'s inputs with the
        explicit flux of the device gradient of `previous` (None: the last
        solve's p')."""
        p = None if previous is None else _f64(previous)
        self._ctx.call("cfdhip_fmesh_assemble_pressure_pass", None if p is None else _pf(p),
                       self.GRADIENT_SCHEMES[scheme])

    def passes(self, settings: "LinearSolverSettings", passes, scheme="green_gauss"):
        """The p' corrector passes (SIMPLE.cpp:876-924) from the first-pass
        assembly: the last p' and one SolverResult per pass."""
        x = np.zeros(self.n_cells)
        rs = (_lib.SolveResult * passes)()
        s = settings.c()
        self._ctx.call("cfdhip_fmesh_pressure_passes", ctypes.byref(s), int(passes),
                       self.GRADIENT_SCHEMES[scheme], _pf(x), rs)
        return x, [_result(r, None, np.zeros(0)) for r in rs]


FACE_FIELDS = _lib.FACE_FIELDS


# ---------------------------------------------------------------------------
# Learned preconditioner

WEIGHTS_PATH = __import__("os").path.join(_lib.HERE, "weights", "pprime_unet.f32")


class BatchSolver:
    """Many independent 2D p' systems of uneven size solved together
    (cfdhip_batch_*): Jacobi-PCG, one launch per step for the whole batch,
    FP32 inner iterations inside FP64 refinement (precision "mixed") or FP64;
    or, after set_learned(), flexible CG with the learned preconditioner
    (settings type "FCG", preconditioner "Learned", FP64 vectors), each step
    one launch over all active cases. No reference counterpart; each case
    meets the same FP64 true-residual test as a single solve
    (src/algebra/CG.cpp:14-148)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.cfdhip_batch_create(device, ctypes.byref(h))
        if rc != _lib.OK:
            raise CfdHipError(rc, "cannot create a batch")
        self.h = h
        self.shapes = []
        self.stats = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.cfdhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != _lib.OK:
            msg = self.lib.cfdhip_batch_last_error(self.h)
            raise CfdHipError(rc, msg.decode() if msg else "")

    def add_system(self, nx, ny, pin, diag, face_coeff, rhs) -> int:
        """Host arrays: diagonal (nx*ny), face couplings (x-faces then y-faces),
        rhs; pin = pinned cell or -1."""
        diag, face_coeff, rhs = _f64(diag), _f64(face_coeff), _f64(rhs)
        if diag.size != nx * ny or rhs.size != nx * ny or \
                face_coeff.size != (nx + 1) * ny + nx * (ny + 1):
            raise ValueError("batch system: array sizes do not match nx, ny")
        idx = ctypes.c_int64(-1)
        self._call("cfdhip_batch_add_system", nx, ny, pin, _pf(diag), _pf(face_coeff), _pf(rhs),
                   ctypes.byref(idx))
        self.shapes.append((nx, ny))
        return idx.value

    def add_pressure(self, d: "GpuSimpleDiscretization") -> int:
        """The p' system the discretization has assembled (copied on the device)."""
        idx = ctypes.c_int64(-1)
        self._call("cfdhip_batch_add_pressure", d._ctx.h, ctypes.byref(idx))
        self.shapes.append(tuple(d.shape[:2]))
        return idx.value

    def set_learned(self, weights=None, options=None):
        """The learned preconditioner of the batched FCG solve (the shipped
        weights and their options by default)."""
        if weights is None:
            weights = load_learned_weights()
            if options is None:
                options = load_learned_options()
        w = np.ascontiguousarray(weights, dtype=np.float32)
        o = _unet_options(options)
        self._call("cfdhip_batch_set_learned", w.ctypes.data_as(_lib.P_F32), w.size,
                   ctypes.byref(o))

    def solve(self, settings: LinearSolverSettings, tile_rows=0, use_graph=True):
        """Returns a list of SolverResult (solution None; device_ms = the case's
        completion time from the start of the batch); self.stats holds the
        batch figures."""
        n = len(self.shapes)
        res = (_lib.SolveResult * max(n, 1))()
        lat = np.zeros(max(n, 1))
        st = np.zeros(8)
        s = settings.c()
        self._call("cfdhip_batch_solve", ctypes.byref(s), int(tile_rows), int(bool(use_graph)),
                   res, _pf(lat), _pf(st))
        self.stats = {"device_ms": st[0], "steps": int(st[1]), "cell_iterations": st[2],
                      "tile_rows": int(st[3]), "graph_captures": int(st[4]),
                      "tune_us": {16: st[5], 32: st[6], 64: st[7]}}
        return [_result(res[k], None, None) for k in range(n)]

    def solution(self, index) -> np.ndarray:
        nx, ny = self.shapes[index]
        x = np.empty(nx * ny)
        self._call("cfdhip_batch_solution", int(index), _pf(x), x.size)
        return x


def unet_param_count() -> int:
    return int(_lib.load().cfdhip_unet_param_count())


def load_learned_weights(path: str | None = None) -> np.ndarray:
    """The trained network's FP32 master weights (raw little-endian float32,
    layer order e0a .. out, each W[ky][kx][cin][cout])."""
    w = np.fromfile(path or WEIGHTS_PATH, dtype="<f4")
    if w.size != unet_param_count():
        raise ValueError(f"{path or WEIGHTS_PATH}: {w.size} weights, expected "
                         f"{unet_param_count()}")
    return w


def load_learned_options(path: str | None = None) -> dict:
    """How the shipped weights are applied (the "options" record of the JSON
    beside the weight file; zero padding and no tail when it has none)."""
    import json
    import os
    meta = os.path.splitext(path or WEIGHTS_PATH)[0] + ".json"
    if os.path.exists(meta):
        with open(meta) as f:
            opts = json.load(f).get("options")
        if opts:
            return opts
    return {"padding": "zero"}


SIDES = {"left": 1, "right": 2, "bottom": 4, "top": 8}


def _unet_options(opts) -> "_lib.UnetOptions":
    """cfdhip_unet_options from a dict: padding "zero" / "neumann", an
    optional tail {levels, omega, a, k}, dirichlet_sides (names of the
    sides with odd reflection: left, right, bottom, top) and dense (True: the
    MFMA layers over every channel, never the narrow kernels)."""
    o = _lib.UnetOptions()
    opts = opts or {}
    o.padding = 1 if opts.get("padding", "zero") == "neumann" else 0
    o.dirichlet_sides = sum(SIDES[s] for s in (opts.get("dirichlet_sides") or ()))
    o.dense = 1 if opts.get("dense") else 0
    tail = opts.get("tail") or {}
    o.tail_levels = int(tail.get("levels", 0))
    o.tail_omega = float(tail.get("omega", 0.0))
    for q in range(3):
        o.tail_a[q] = float((tail.get("a") or (0.0, 0.0, 0.0))[q])
        o.tail_k[q] = float((tail.get("k") or (0.0, 0.0, 0.0))[q])
    return o


def _sync_producer():
    """Device buffers handed to the library are usually filled by torch on its
    own stream; the library works on its own stream, so wait for torch first."""
    t = __import__("sys").modules.get("torch")
    if t is not None and t.cuda.is_initialized():
        t.cuda.synchronize()


def _dev(t):
    """Device address of a torch tensor (or an int address, or None)."""
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    return ctypes.c_void_p(t.data_ptr())


class UnetNet:
    """The learned preconditioner's network on the device (cfdhip_unet_*).
    Inputs and outputs are caller-owned FP32 device buffers (torch tensors);
    every call is synchronous with respect to the library's own stream, so
    the caller synchronises its producer stream before calling."""

    def __init__(self, weights=None, options=None):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.cfdhip_unet_create(ctypes.byref(h))
        if rc != _lib.OK:
            raise CfdHipError(rc, "cfdhip_unet_create failed")
        self.h = h
        if options is not None:
            self.set_options(options)
        if weights is not None:
            self.set_weights(weights)

    def set_options(self, options):
        o = _unet_options(options)
        self._call("cfdhip_unet_set_options", ctypes.byref(o))

    def close(self):
        if getattr(self, "h", None):
            self.lib.cfdhip_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _call(self, name, *args):
        _sync_producer()
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != _lib.OK:
            msg = self.lib.cfdhip_unet_last_error(self.h)
            raise CfdHipError(rc, msg.decode() if msg else "")

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._call("cfdhip_unet_set_weights", w.ctypes.data_as(_lib.P_F32), w.size)

    def get_weights(self) -> np.ndarray:
        w = np.zeros(unet_param_count(), dtype=np.float32)
        self._call("cfdhip_unet_get_weights", w.ctypes.data_as(_lib.P_F32), w.size)
        return w

    def forward(self, x, y, batch, h, w):
        self._call("cfdhip_unet_forward", _dev(x), _dev(y), int(batch), int(h), int(w))

    def apply(self, x, y, batch, h, w):
        """The inference path (first layer fused into the second)."""
        self._call("cfdhip_unet_apply", _dev(x), _dev(y), int(batch), int(h), int(w))

    def backward(self, x, gy, batch, h, w) -> np.ndarray:
        g = np.zeros(unet_param_count(), dtype=np.float32)
        self._call("cfdhip_unet_backward", _dev(x), _dev(gy), int(batch), int(h), int(w),
                   g.ctypes.data_as(_lib.P_F32))
        return g

    TENSORS = ["e0a", "e0", "p0", "e1a", "e1", "p1", "e2a", "e2", "p2", "b3a", "b3", "d2a", "d2",
               "d1a", "d1", "d0a", "d0", "gd0", "gd0a", "gd1", "ge0s", "gd1a", "gd2", "ge1s",
               "gd2a", "gb3", "ge2s", "gb3a", "gp2", "ge2a", "gp1", "ge1a", "gp0", "ge0a"]

    def tensor(self, name) -> np.ndarray:
        """An activation (or gradient) tensor of the last pass, flat NHWC."""
        t = self.TENSORS.index(name)
        m = int(self.lib.cfdhip_unet_tensor_size(self.h, t))
        if m < 0:
            raise CfdHipError(_lib.ERR_INVALID, f"tensor {name} is not available")
        out = np.zeros(m, dtype=np.float32)
        self._call("cfdhip_unet_tensor", t, out.ctypes.data_as(_lib.P_F32), m)
        return out

    def loss_grad(self, x, e, ce, cn, sys_offset, den, batch, h, w, cd=None):
        """A-norm loss and its gradient in the master weights (no update);
        cd: the operator's diagonal (None: unit diagonal)."""
        loss = ctypes.c_double(0.0)
        g = np.zeros(unet_param_count(), dtype=np.float32)
        self._call("cfdhip_unet_loss_grad", _dev(x), _dev(e), _dev(ce), _dev(cn), _dev(cd),
                   _dev(sys_offset), _dev(den), int(batch), int(h), int(w), ctypes.byref(loss),
                   g.ctypes.data_as(_lib.P_F32))
        return loss.value, g

    def spectral_loss_grad(self, z, ce, cn, sys_offset, batch, h, w, steps, cd=None,
                           grad=True):
        """Multi-step spectral loss of x <- x + N(b - A x) on the error images z
        (cfdhip_unet_spectral_loss_grad) and, with grad, its gradient in the
        master weights; returns (loss, gradient or None)."""
        loss = ctypes.c_double(0.0)
        g = np.zeros(unet_param_count(), dtype=np.float32) if grad else None
        self._call("cfdhip_unet_spectral_loss_grad", _dev(z), _dev(ce), _dev(cn), _dev(cd),
                   _dev(sys_offset), int(batch), int(h), int(w), int(steps), ctypes.byref(loss),
                   g.ctypes.data_as(_lib.P_F32) if grad else None)
        return loss.value, g

    def train_step(self, x, e, ce, cn, sys_offset, den, batch, h, w, lr, update=True,
                   cd=None) -> float:
        loss = ctypes.c_double(0.0)
        self._call("cfdhip_unet_train_step", _dev(x), _dev(e), _dev(ce), _dev(cn), _dev(cd),
                   _dev(sys_offset), _dev(den), int(batch), int(h), int(w), float(lr),
                   1 if update else 0, ctypes.byref(loss))
        return loss.value
