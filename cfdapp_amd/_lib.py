"""ctypes binding of libcfdhip.so (include/cfdhip.h).

The library is built in-tree by ``cfdapp_amd/build.sh`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback on the
product path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfdhip.so")

P_F64 = ctypes.POINTER(ctypes.c_double)
P_I64 = ctypes.POINTER(ctypes.c_int64)
P_F32 = ctypes.POINTER(ctypes.c_float)
VP = ctypes.c_void_p

# include/cfdhip.h enums
OK, ERR_INVALID, ERR_RUNTIME, ERR_NO_DEVICE, ERR_UNSUPPORTED, ERR_NONFINITE = range(6)
CONVERGED, MAX_ITERATIONS, BREAKDOWN, NONFINITE_INPUT, NONFINITE_RESIDUAL, INVALID_SYSTEM = range(6)
SOLVER_CG, SOLVER_BICGSTAB, SOLVER_FCG = 0, 1, 3
PREC_NONE, PREC_JACOBI, PREC_LEARNED, PREC_HOST = 0, 1, 2, 3
FP64, MIXED = 0, 1
WALL, MOVING_WALL, INLET, OUTLET = 0, 1, 2, 3
FIXED_GRADIENT, FIXED_VALUE = 0, 1

FIELDS = dict(U=0, V=1, W=2, P=3, PPRIME=4, FLUX=5, FLUX_STAR=6, VISCOSITY=7, RESPONSE=8,
              USTAR=9, VSTAR=10, WSTAR=11, GRAD_X=12, GRAD_Y=13, GRAD_Z=14, IMBALANCE=15,
              MOM_RHS_U=16, MOM_RHS_V=17, MOM_RHS_W=18, MOM_DIAG=19, PC_RHS=20, PC_DIAG=21,
              PC_FACE_COEFF=22)
FACE_FIELDS = {"FLUX", "FLUX_STAR", "PC_FACE_COEFF"}

STATUS_NAMES = ["Converged", "MaxIterations", "Breakdown", "NonFiniteInput",
                "NonFiniteResidual", "InvalidSystem"]


class SolverSettings(ctypes.Structure):
    _fields_ = [("absolute_tolerance", ctypes.c_double),
                ("relative_tolerance", ctypes.c_double),
                ("max_iterations", ctypes.c_int64),
                ("type", ctypes.c_int32),
                ("preconditioner", ctypes.c_int32),
                ("precision", ctypes.c_int32),
                ("check_interval", ctypes.c_int32),
                ("stagnation_window", ctypes.c_int32),
                ("graph", ctypes.c_int32),
                ("inner_tolerance", ctypes.c_double)]


class SolveResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32),
                ("fallback_used", ctypes.c_int32),
                ("iterations", ctypes.c_int64),
                ("restarts", ctypes.c_int64),
                ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double),
                ("true_residual", ctypes.c_double),
                ("device_ms", ctypes.c_double),
                ("history_length", ctypes.c_int64)]


class UnetOptions(ctypes.Structure):
    _fields_ = [("padding", ctypes.c_int32),
                ("tail_levels", ctypes.c_int32),
                ("tail_omega", ctypes.c_float),
                ("tail_a", ctypes.c_float * 3),
                ("tail_k", ctypes.c_float * 3),
                ("dirichlet_sides", ctypes.c_int32),
                ("dense", ctypes.c_int32)]


# Every exported symbol with its signature; the not-gpu tests check that the
# built library exports exactly these names.
# int (*)(const double* r, double* z, int64_t n, void* user)
HOST_PRECONDITIONER = ctypes.CFUNCTYPE(ctypes.c_int, P_F64, P_F64, ctypes.c_int64, ctypes.c_void_p)

SIGNATURES = {
    "cfdhip_device_available": (ctypes.c_int, []),
    "cfdhip_create": (ctypes.c_void_p, [ctypes.c_int]),
    "cfdhip_destroy": (None, [ctypes.c_void_p]),
    "cfdhip_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "cfdhip_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "cfdhip_mesh_cartesian": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                             ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double]),
    "cfdhip_mesh_sizes": (ctypes.c_int, [ctypes.c_void_p, P_I64]),
    "cfdhip_mesh_faces": (ctypes.c_int, [ctypes.c_void_p, P_I64, P_I64]),
    "cfdhip_set_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P_F64,
                                           ctypes.c_int, ctypes.c_double]),
    "cfdhip_set_fluid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "cfdhip_field_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, P_F64, ctypes.c_int64]),
    "cfdhip_field_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, P_F64,
                                             ctypes.c_int64]),
    "cfdhip_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "cfdhip_mass_flux": (ctypes.c_int, [ctypes.c_void_p]),
    "cfdhip_assemble_momentum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "cfdhip_solve_momentum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(SolverSettings),
                                             ctypes.POINTER(SolveResult)]),
    "cfdhip_response": (ctypes.c_int, [ctypes.c_void_p]),
    "cfdhip_predicted_flux": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "cfdhip_assemble_pressure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "cfdhip_mesh_cells": (ctypes.c_int, [ctypes.c_void_p, P_F64, P_F64]),
    "cfdhip_assemble_compressible_pressure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64,
                                                             P_F64, P_F64, P_F64, ctypes.c_double,
                                                             ctypes.c_double]),
    "cfdhip_solve_pressure": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverSettings),
                                             ctypes.POINTER(SolveResult)]),
    "cfdhip_update_pressure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "cfdhip_correct": (ctypes.c_int, [ctypes.c_void_p]),
    "cfdhip_correct_velocity": (ctypes.c_int, [ctypes.c_void_p, P_F64]),
    "cfdhip_correct_flux": (ctypes.c_int, [ctypes.c_void_p, P_F64]),
    "cfdhip_fields_all_finite": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "cfdhip_set_density": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "cfdhip_resident_bytes": (ctypes.c_int, [ctypes.c_void_p, P_I64]),
    "cfdhip_assemble_momentum_component": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                          ctypes.c_double, ctypes.c_int,
                                                          ctypes.c_int, P_F64]),
    "cfdhip_continuity": (ctypes.c_int, [ctypes.c_void_p, P_F64]),
    "cfdhip_simple_run": (ctypes.c_int, [ctypes.c_void_p, P_F64, ctypes.POINTER(SolverSettings),
                                         ctypes.POINTER(SolverSettings), P_F64, P_F64,
                                         ctypes.c_int64]),
    "cfdhip_pressure_csr": (ctypes.c_int, [ctypes.c_void_p, P_I64, P_I64, P_F64, ctypes.c_int64,
                                           P_I64]),
    "cfdhip_momentum_csr": (ctypes.c_int, [ctypes.c_void_p, P_I64, P_I64, P_F64, ctypes.c_int64,
                                           P_I64]),
    "cfdhip_csr_spmv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, P_I64, P_I64, P_F64,
                                       P_F64, P_F64]),
    "cfdhip_fmesh_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                        ctypes.c_int64, P_I64, P_I64, P_F64, P_F64, P_F64,
                                        P_I64, P_I64, ctypes.POINTER(ctypes.c_int32)]),
    "cfdhip_fmesh_assemble_pressure": (ctypes.c_int, [ctypes.c_void_p, P_F64, P_F64,
                                                      ctypes.c_double, P_F64, P_F64, P_F64,
                                                      ctypes.c_int, P_F64, P_F64,
                                                      ctypes.c_int64]),
    "cfdhip_fmesh_pressure_system": (ctypes.c_int, [ctypes.c_void_p, P_I64, P_I64, P_F64,
                                                    ctypes.c_int64, P_I64, P_F64, P_F64, P_F64]),
    "cfdhip_fmesh_solve_pressure": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverSettings),
                                                   P_F64, ctypes.POINTER(SolveResult)]),
    "cfdhip_fmesh_correct_flux": (ctypes.c_int, [ctypes.c_void_p, P_F64, P_F64]),
    "cfdhip_fmesh_set_cell_volumes": (ctypes.c_int, [ctypes.c_void_p, P_F64]),
    "cfdhip_fmesh_gradient": (ctypes.c_int, [ctypes.c_void_p, P_F64,
                                             ctypes.POINTER(ctypes.c_int32), P_F64, ctypes.c_int,
                                             P_F64]),
    "cfdhip_fmesh_assemble_pressure_pass": (ctypes.c_int, [ctypes.c_void_p, P_F64, ctypes.c_int]),
    "cfdhip_fmesh_pressure_passes": (ctypes.c_int, [ctypes.c_void_p,
                                                    ctypes.POINTER(SolverSettings), ctypes.c_int,
                                                    ctypes.c_int, P_F64,
                                                    ctypes.POINTER(SolveResult)]),
    "cfdhip_csr_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverSettings),
                                        ctypes.c_int64, P_I64, P_I64, P_F64, P_F64, P_F64, P_F64,
                                        ctypes.POINTER(SolveResult)]),
    "cfdhip_pressure_spmv": (ctypes.c_int, [ctypes.c_void_p, P_F64, P_F64]),
    "cfdhip_pressure_solve_rhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverSettings),
                                                 P_F64, P_F64, ctypes.POINTER(SolveResult)]),
    "cfdhip_pressure_solve_device_noalloc": (ctypes.c_int, [ctypes.c_void_p,
                                                            ctypes.POINTER(SolverSettings),
                                                            ctypes.c_void_p, ctypes.c_void_p,
                                                            ctypes.POINTER(SolveResult)]),
    "cfdhip_solve_history": (ctypes.c_int, [ctypes.c_void_p, P_F64, ctypes.c_int64, P_I64]),
    "cfdhip_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "cfdhip_set_host_preconditioner": (ctypes.c_int, [ctypes.c_void_p, VP, VP]),
    "cfdhip_kernel_timing": (ctypes.c_int, [ctypes.c_void_p, P_F64]),
    "cfdhip_preconditioner_timing": (ctypes.c_int, [ctypes.c_void_p, P_F64, P_I64]),
    # learned preconditioner (device buffers are passed as void*)
    "cfdhip_unet_param_count": (ctypes.c_int64, []),
    "cfdhip_set_learned_weights": (ctypes.c_int, [VP, P_F32, ctypes.c_int64]),
    "cfdhip_set_learned_options": (ctypes.c_int, [VP, ctypes.POINTER(UnetOptions)]),
    "cfdhip_learned_channels": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_int32),
                                               ctypes.c_int32]),
    "cfdhip_pressure_harvest": (ctypes.c_int, [VP, ctypes.POINTER(SolverSettings), P_F64,
                                               ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_int64, VP, VP, VP, VP, P_F32, P_I64,
                                               ctypes.POINTER(SolveResult)]),
    "cfdhip_pressure_harvest_noise": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.c_int64,
                                                     ctypes.c_int64, ctypes.c_int64, VP, VP,
                                                     P_F32, P_I64]),
    "cfdhip_unet_create": (ctypes.c_int, [ctypes.POINTER(VP)]),
    "cfdhip_unet_destroy": (None, [VP]),
    "cfdhip_unet_last_error": (ctypes.c_char_p, [VP]),
    "cfdhip_unet_set_weights": (ctypes.c_int, [VP, P_F32, ctypes.c_int64]),
    "cfdhip_unet_set_options": (ctypes.c_int, [VP, ctypes.POINTER(UnetOptions)]),
    "cfdhip_unet_get_weights": (ctypes.c_int, [VP, P_F32, ctypes.c_int64]),
    "cfdhip_unet_forward": (ctypes.c_int, [VP, VP, VP, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int64]),
    "cfdhip_unet_apply": (ctypes.c_int, [VP, VP, VP, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int64]),
    "cfdhip_unet_backward": (ctypes.c_int, [VP, VP, VP, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, P_F32]),
    "cfdhip_unet_tensor_size": (ctypes.c_int64, [VP, ctypes.c_int]),
    "cfdhip_unet_tensor": (ctypes.c_int, [VP, ctypes.c_int, P_F32, ctypes.c_int64]),
    "cfdhip_unet_loss_grad": (ctypes.c_int, [VP, VP, VP, VP, VP, VP, VP, VP, ctypes.c_int64,
                                             ctypes.c_int64, ctypes.c_int64, P_F64, P_F32]),
    "cfdhip_unet_train_step": (ctypes.c_int, [VP, VP, VP, VP, VP, VP, VP, VP, ctypes.c_int64,
                                              ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                              ctypes.c_int, P_F64]),
    "cfdhip_unet_spectral_loss_grad": (ctypes.c_int, [VP, VP, VP, VP, VP, VP, ctypes.c_int64,
                                                      ctypes.c_int64, ctypes.c_int64,
                                                      ctypes.c_int32, P_F64, P_F32]),
    "cfdhip_pressure_spmv_device": (ctypes.c_int, [VP, VP, VP, ctypes.c_int]),
    "cfdhip_field_device": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(VP)]),
    "cfdhip_assemble_transient_momentum": (ctypes.c_int, [VP, ctypes.c_double]),
    "cfdhip_piso_run": (ctypes.c_int, [VP, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.POINTER(SolverSettings),
                                       ctypes.POINTER(SolverSettings), P_F64, P_I64]),
    "cfdhip_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "cfdhip_batch_destroy": (None, [VP]),
    "cfdhip_batch_last_error": (ctypes.c_char_p, [VP]),
    "cfdhip_batch_size": (ctypes.c_int64, [VP]),
    "cfdhip_batch_add_system": (ctypes.c_int, [VP, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                               P_F64, P_F64, P_F64, P_I64]),
    "cfdhip_batch_add_pressure": (ctypes.c_int, [VP, VP, P_I64]),
    "cfdhip_batch_solve": (ctypes.c_int, [VP, ctypes.POINTER(SolverSettings), ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(SolveResult), P_F64,
                                          P_F64]),
    "cfdhip_batch_solution": (ctypes.c_int, [VP, ctypes.c_int64, P_F64, ctypes.c_int64]),
    "cfdhip_batch_set_learned": (ctypes.c_int, [VP, P_F32, ctypes.c_int64,
                                                ctypes.POINTER(UnetOptions)]),
}

_lib = None


def load():
    """Loads libcfdhip.so (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with cfdapp_amd/build.sh "
                "(or __graft_entry__.build()); there is no CPU fallback")
        # One HIP runtime per process. torch ships its own libamdhip64 /
        # libhsa-runtime64 under the same sonames as ROCm's: imported first,
        # its copies also serve libcfdhip.so; loaded after libcfdhip.so has
        # pulled in ROCm's, torch maps a second runtime that finds no device.
        # So torch, when it is installed, is loaded first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
