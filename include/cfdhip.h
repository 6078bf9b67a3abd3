/*
 * cfdhip.h -- C ABI of libcfdhip, the MI355X (gfx950) pressure-velocity path.
 *
 * Plain C: opaque context, caller-owned host buffers, int status codes
 * (0 = OK, see cfdhip_last_error), no exceptions and no C++ types crossing the
 * boundary. Device memory belongs to the context; one HIP stream per context.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository). The reference reaches these through
 *   cfd::gpu::makeGpuCG / makeGpuBiCGSTAB    (include/cfd/gpu/GpuLinearSolver.hpp:48,52)
 *   cfd::gpu::GpuSimpleDiscretization        (include/cfd/gpu/GpuSimpleDiscretization.hpp:56-262)
 *   cfd::gpu::cudaAvailable                   (include/cfd/gpu/GPUBackend.hpp:15)
 * and a maintainer binds them behind those interfaces (a thin pimpl over the
 * C++ adapter in cfdhip.hpp).
 *
 * Arithmetic: everything on the reference path is FP64. Assembly, gradient,
 * flux, correction and SpMV results are bitwise equal to the reference CPU
 * path on the same inputs; Krylov iterates differ only by reduction order.
 * The mixed-precision (CFDHIP_MIXED) and learned-preconditioner solves are
 * additions with no reference counterpart; they are judged by the same FP64
 * true-residual tolerance.
 *
 * Threading: a context is not re-entrant; calls on one context must not
 * overlap. Distinct contexts are independent.
 */
#ifndef CFDHIP_H
#define CFDHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cfdhip_ctx cfdhip_ctx;

/* Return codes. */
enum {
  CFDHIP_OK = 0,
  CFDHIP_ERR_INVALID = 1,     /* bad argument or state (reference: InvalidArgumentError) */
  CFDHIP_ERR_RUNTIME = 2,     /* HIP runtime failure (reference: NumericalError via checkCuda) */
  CFDHIP_ERR_NO_DEVICE = 3,   /* no usable gfx950 device */
  CFDHIP_ERR_UNSUPPORTED = 4, /* configuration outside the structured path (prepare() == false) */
  CFDHIP_ERR_NONFINITE = 5    /* assembled system or state not finite (SIMPLE NonFiniteState) */
};

/* cfd::algebra::SolverStatus (include/cfd/algebra/LinearSolver.hpp:16-23), same order. */
enum {
  CFDHIP_CONVERGED = 0,
  CFDHIP_MAX_ITERATIONS = 1,
  CFDHIP_BREAKDOWN = 2,
  CFDHIP_NONFINITE_INPUT = 3,
  CFDHIP_NONFINITE_RESIDUAL = 4,
  CFDHIP_INVALID_SYSTEM = 5
};

/* cfd::algebra::LinearSolverType (LinearSolver.hpp:31-39) plus flexible CG. */
enum { CFDHIP_SOLVER_CG = 0, CFDHIP_SOLVER_BICGSTAB = 1, CFDHIP_SOLVER_FCG = 3 };
/* cfd::algebra::PreconditionerType (LinearSolver.hpp:62-65) plus the learned one, and
 * CFDHIP_PREC_HOST: the caller's own host preconditioner (cfdhip_set_host_preconditioner),
 * the Preconditioner object makeGpuCG / makeGpuBiCGSTAB take (GpuLinearSolver.hpp:48-52). */
enum { CFDHIP_PREC_NONE = 0, CFDHIP_PREC_JACOBI = 1, CFDHIP_PREC_LEARNED = 2,
       CFDHIP_PREC_HOST = 3 };
/* Arithmetic of the inner iteration. */
enum { CFDHIP_FP64 = 0, CFDHIP_MIXED = 1 };

/* Velocity / pressure boundary-condition kinds (src/boundary/Wall.cpp etc.). */
enum { CFDHIP_WALL = 0, CFDHIP_MOVING_WALL = 1, CFDHIP_INLET = 2, CFDHIP_OUTLET = 3 };
enum { CFDHIP_FIXED_GRADIENT = 0, CFDHIP_FIXED_VALUE = 1 };

/* cfd::algebra::LinearSolverSettings (LinearSolver.hpp:136-166) as POD. */
typedef struct {
  double absolute_tolerance;
  double relative_tolerance;
  int64_t max_iterations;
  int32_t type;           /* CFDHIP_SOLVER_* */
  int32_t preconditioner; /* CFDHIP_PREC_* */
  int32_t precision;      /* CFDHIP_FP64 or CFDHIP_MIXED */
  int32_t check_interval; /* iterations between host reads of the device verdict (0 = 32) */
  int32_t stagnation_window; /* FCG: fall back to Jacobi-PCG after this many non-improving iterations (0 = 200) */
  int32_t graph;          /* 1: replay each chunk of check_interval iterations from a captured
                             HIP graph (2D Jacobi-PCG and the p' solves of cfdhip_simple_run) */
  double inner_tolerance; /* precision CFDHIP_MIXED: the FP32 reduction between FP64 residual
                             recomputations (0 = default). Jacobi-PCG replaces its residual by
                             the FP64 b - A x (its iterate is FP64) and restarts its direction
                             once the FP32 recurrence residual is below this fraction of its
                             largest value since the last replacement (default 1e-5);
                             flexible CG and the batched solver restart from the FP64
                             residual after this much reduction (defaults 1e-4 and 1e-5).
                             Must be in [0, 1). */
} cfdhip_solver_settings;

/* cfd::algebra::SolverResult (LinearSolver.hpp:103-134) without the solution
 * vector (copied separately into caller memory). */
typedef struct {
  int32_t status;          /* CFDHIP_CONVERGED ... */
  int32_t fallback_used;   /* 1 if FP64 Jacobi-PCG finished the solve (FCG stagnation, or a
                              mixed-precision refinement cycle that did not halve |r|) */
  int64_t iterations;
  int64_t restarts;        /* BiCGSTAB restarts / refinement cycles (mixed) */
  double initial_residual; /* |b - A x0| */
  double final_residual;   /* recurrence residual at exit (reference semantics) */
  double true_residual;    /* |b - A x| recomputed in FP64 at exit */
  double device_ms;        /* device time of the solve (HIP events on the context stream) */
  int64_t history_length;  /* residual history entries available (cfdhip_solve_history) */
} cfdhip_solve_result;

/* Persistent device fields of a case (GpuSimpleDiscretization::PersistentField
 * and the intermediate stages). Cell fields hold n doubles, face fields nf. */
enum {
  CFDHIP_FIELD_U = 0, CFDHIP_FIELD_V = 1, CFDHIP_FIELD_W = 2,
  CFDHIP_FIELD_P = 3, CFDHIP_FIELD_PPRIME = 4,
  CFDHIP_FIELD_FLUX = 5,       /* face mass flux F (faces) */
  CFDHIP_FIELD_FLUX_STAR = 6,  /* predicted face flux F* (faces) */
  CFDHIP_FIELD_VISCOSITY = 7,  /* effective viscosity (cells) */
  CFDHIP_FIELD_RESPONSE = 8,   /* d = V / a_P (cells; equal for u, v, w) */
  CFDHIP_FIELD_USTAR = 9, CFDHIP_FIELD_VSTAR = 10, CFDHIP_FIELD_WSTAR = 11,
  CFDHIP_FIELD_GRAD_X = 12, CFDHIP_FIELD_GRAD_Y = 13, CFDHIP_FIELD_GRAD_Z = 14,
  CFDHIP_FIELD_IMBALANCE = 15, /* continuity imbalance per cell */
  CFDHIP_FIELD_MOM_RHS_U = 16, CFDHIP_FIELD_MOM_RHS_V = 17, CFDHIP_FIELD_MOM_RHS_W = 18,
  CFDHIP_FIELD_MOM_DIAG = 19,  /* relaxed momentum diagonal (cells) */
  CFDHIP_FIELD_PC_RHS = 20, CFDHIP_FIELD_PC_DIAG = 21,
  CFDHIP_FIELD_PC_FACE_COEFF = 22, /* p' face coupling D_f (faces) */
  CFDHIP_FIELD_COUNT = 23
};

/* ---- device ---------------------------------------------------------------- */

/* cfd::gpu::cudaAvailable (include/cfd/gpu/GPUBackend.hpp:15; src/gpu/GPUBackend.cpp:10):
 * 1 if a gfx950 device is usable, else 0. Never fails. */
int cfdhip_device_available(void);

/* Context = device + stream + resident case state. NULL when no device
 * (the reference's makeGpu* then returns nullptr and the factory falls back,
 * src/algebra/LinearSolverFactory.cpp:12-31). */
cfdhip_ctx* cfdhip_create(int device);
void cfdhip_destroy(cfdhip_ctx* ctx);
/* Message of the last failing call on this context ("" if none). */
const char* cfdhip_last_error(const cfdhip_ctx* ctx);
/* Blocks until all work queued on the context stream has finished. */
int cfdhip_synchronize(cfdhip_ctx* ctx);

/* ---- mesh plan (GpuSimpleDiscretization::prepare,
 *      include/cfd/gpu/GpuSimpleDiscretization.hpp:74) ------------------------- */

/* Uniform Cartesian mesh in the order of MeshGeometry::createCartesian2D/3D
 * (src/mesh/MeshGeometry.cpp:428-553, 1047-1160). dim 2 ignores nz/lz.
 * Returns CFDHIP_ERR_UNSUPPORTED (with the reason) for geometry the structured
 * kernels do not cover. */
int cfdhip_mesh_cartesian(cfdhip_ctx* ctx, int dim, int64_t nx, int64_t ny, int64_t nz,
                          double lx, double ly, double lz);
/* out[0..3] = {cells, faces, patches, dim}. */
int cfdhip_mesh_sizes(cfdhip_ctx* ctx, int64_t* out);
/* Face owner/neighbour (neighbour = -1 on a boundary face), generated on the
 * device in the reference face order (Mesh face list / DeviceMeshCuda upload). */
int cfdhip_mesh_faces(cfdhip_ctx* ctx, int64_t* owner, int64_t* neighbor);
/* Cell centroids (x, y, z per cell, z = 0 in 2D) and volumes (may be NULL) in
 * cell order, as Mesh::cell(id).centroid() / volume() of
 * MeshGeometry::createCartesian2D/3D (src/mesh/MeshGeometry.cpp:428-553,
 * 1047-1160); what CSVWriter::writeFields prints in its x, y (, z) columns. */
int cfdhip_mesh_cells(cfdhip_ctx* ctx, double* centroid, double* volume);
/* Boundary conditions of patch p (2D: left, right, bottom, top; 3D: xmin..zmax),
 * as the BoundaryConditionSet the reference passes to prepare(). */
int cfdhip_set_boundary(cfdhip_ctx* ctx, int patch, int velocity_kind, const double* velocity3,
                        int pressure_kind, double pressure_value);
/* FluidProperties density + uniform laminar viscosity (LaminarModel:
 * mu_eff = mu + 0, src/turbulence/TurbulenceModel.cpp:13-24). */
int cfdhip_set_fluid(cfdhip_ctx* ctx, double density, double viscosity);

/* ---- fields (uploadInitialState / download*, GpuSimpleDiscretization.hpp) ---- */
int cfdhip_field_upload(cfdhip_ctx* ctx, int field, const double* host, int64_t count);
int cfdhip_field_download(cfdhip_ctx* ctx, int field, double* host, int64_t count);
/* Device address of a resident field (borrowed: owned by the context, valid
 * until it is destroyed or its mesh changes), for callers that hand device
 * buffers to the *_device entry points, as DeviceCsrMatrix::adoptDevice
 * borrows the assembly buffers (include/cfd/gpu/DeviceCsrMatrix.hpp:75-88). */
int cfdhip_field_device(cfdhip_ctx* ctx, int field, void** out);

/* ---- SIMPLE stages (GpuSimpleDiscretization resident methods) ---------------- */

/* Green-Gauss gradient of a cell field into GRAD_X/Y/Z; bc = 0 pressure
 * conditions, 1 p' conditions (makeGradientBoundaries). gradient(),
 * src/discretization/Gradient.cpp:534-605. */
int cfdhip_gradient(cfdhip_ctx* ctx, int field, int bc);
/* calculateMassFlux (src/physics/MassFlux.cpp:15) of U/V/W into FLUX. */
int cfdhip_mass_flux(cfdhip_ctx* ctx);
/* assembleRelaxedMomentumComponent for all components at once
 * (src/pressure_velocity/RelaxedMomentum.cpp:30; the u, v, w matrices are
 * identical, the right-hand sides differ). assembleMomentumResident. */
int cfdhip_assemble_momentum(cfdhip_ctx* ctx, double alpha);
/* GpuSimpleDiscretization::assembleMomentum / assembleMomentumResident for one
 * component (include/cfd/gpu/GpuSimpleDiscretization.hpp:162-164,181-182;
 * cuda/kernels/GpuSimpleDiscretizationCuda.cpp:321-392): the relaxed momentum
 * system of `component` from the resident U/V/W, P, FLUX and VISCOSITY. The
 * matrix, the same for every component, is exported by cfdhip_momentum_csr;
 * the right-hand side lands in MOM_RHS_U + component and the relaxed diagonal
 * in MOM_DIAG. previous_component (host, n doubles, may be NULL) is the
 * under-relaxation's phi_old; NULL takes the resident component, as
 * assembleMomentumResident does. convection_scheme is
 * cfd::discretization::ConvectionScheme (Convection.hpp:22-27): only 0
 * (Upwind) is on the structured path, and non_orthogonal must be 0; anything
 * else returns CFDHIP_ERR_UNSUPPORTED. CFDHIP_ERR_NONFINITE when the assembled
 * system is not finite. */
int cfdhip_assemble_momentum_component(cfdhip_ctx* ctx, int component, double alpha,
                                       int convection_scheme, int non_orthogonal,
                                       const double* previous_component);
/* assembleTransientMomentumComponent (src/pressure_velocity/TransientMomentum.cpp,
 * effective-viscosity overload) for both components of a 2D case: the same
 * terms from (U, P, FLUX) plus the implicit-Euler time term
 * (TimeDerivative.cpp:13-34) of the current velocity, no relaxation. The
 * PISO predictor's system; exported by cfdhip_momentum_csr. */
int cfdhip_assemble_transient_momentum(cfdhip_ctx* ctx, double dt);
/* solveMomentumResident: BiCGSTAB (or the type in settings) for one
 * component, warm-started from that component; solution into USTAR+comp. */
int cfdhip_solve_momentum(cfdhip_ctx* ctx, int component, const cfdhip_solver_settings* s,
                          cfdhip_solve_result* r);
/* computeResponseCoefficients: RESPONSE = V / a_P
 * (PressureCorrectionEquation.cpp:103-114). */
int cfdhip_response(cfdhip_ctx* ctx);
/* computePredictedFaceFlux: FLUX_STAR from USTAR/VSTAR/WSTAR, linear or
 * Rhie-Chow (src/pressure_velocity/RhieChow.cpp:53-65). */
int cfdhip_predicted_flux(cfdhip_ctx* ctx, int rhie_chow, double alpha);
/* assemblePressureCorrectionResident (PressureCorrectionEquation.cpp:181-367):
 * PC_DIAG, PC_FACE_COEFF, PC_RHS; reference_cell pinned when no patch is FixedValue. */
int cfdhip_assemble_pressure(cfdhip_ctx* ctx, int64_t reference_cell);
/* assembleCompressiblePressureCorrection (src/compressible/
 * CompressiblePressureCorrection.cpp:20-80, 2D): the same system with the
 * per-face density face_density[n_faces] in the face coupling and the
 * compressibility diagonal V / pseudo_time_step * drho/dp added to each
 * unpinned row after the face terms, drho/dp = 1 / (gas_constant * T) of the
 * ideal gas (IdealGasEOS.cpp:35-39) at pressure_absolute[n_cells] and
 * temperature[n_cells]. Uses FLUX_STAR and RESPONSE (one response coefficient
 * for both components, as on these meshes); host or device pointers.
 * CFDHIP_ERR_INVALID where the reference throws InvalidArgumentError (a
 * pressure, temperature, gas constant or pseudo time step not finite and > 0,
 * reference_cell out of range); CFDHIP_ERR_UNSUPPORTED on a 3D mesh. */
int cfdhip_assemble_compressible_pressure(cfdhip_ctx* ctx, int64_t reference_cell,
                                          const double* face_density,
                                          const double* pressure_absolute,
                                          const double* temperature, double gas_constant,
                                          double pseudo_time_step);
/* --- p' on a general face-list mesh (non-orthogonal, multiblock) ---------
 * The reference's CSR route (cuda/kernels/DevicePressureCorrectionKernel.cu:
 * 176-216) for meshes the structured path does not take: the mesh as its
 * Mesh holds it -- face owner / neighbour (-1 on a boundary face), face
 * centroid and area vector (x, y, z per face), cell centroid (x, y, z per
 * cell), each cell's face list (Cell::faceIds() order, CSR offsets) and a
 * FixedValue-pressure flag per face (boundary faces only). One mesh per
 * context, replaced by the next call. */
int cfdhip_fmesh_set(cfdhip_ctx* ctx, int dim, int64_t n_cells, int64_t n_faces,
                     const int64_t* owner, const int64_t* neighbor, const double* face_centroid,
                     const double* area_vector, const double* cell_centroid,
                     const int64_t* cell_face_offsets, const int64_t* cell_faces,
                     const int32_t* fixed_value);
/* assembleGeometricPressureCorrection (PressureCorrectionEquation.cpp:181-339)
 * on that mesh from host arrays: predictor flux, per-face density (NULL: the
 * constant density), response coefficients du, dv (dw on a 3D mesh, else
 * NULL); non_orthogonal = the over-relaxed implicit coefficient; with
 * grad_previous (x, y, z per cell: the gradient of the previous p' under
 * the p' conditions) the explicit -rho T . grad(p') flux moves to the RHS;
 * extra_diagonal (NULL: none) is added to each unpinned row after the face
 * terms (the compressible equation). Pins reference_cell when no face is
 * FixedValue. CFDHIP_ERR_NONFINITE for degenerate geometry or a non-finite
 * system, as the reference's NumericalError. */
int cfdhip_fmesh_assemble_pressure(cfdhip_ctx* ctx, const double* flux_star,
                                   const double* face_density, double density, const double* du,
                                   const double* dv, const double* dw, int non_orthogonal,
                                   const double* grad_previous, const double* extra_diagonal,
                                   int64_t reference_cell);
/* The assembled system as SparseMatrixBuilder builds it (columns ascending,
 * exact-zero sums dropped): rowptr[n+1], col/val[capacity] (NULL rowptr: only
 * *nnz), rhs[n], face_coeff[n_faces], explicit_flux[n_faces]; any may be NULL. */
int cfdhip_fmesh_pressure_system(cfdhip_ctx* ctx, int64_t* rowptr, int64_t* col, double* val,
                                 int64_t capacity, int64_t* nnz, double* rhs, double* face_coeff,
                                 double* explicit_flux);
/* p' from x0 = 0 by CG / BiCGSTAB (Jacobi or none) on the CSR operator; x[n]
 * (host, may be NULL) receives it. The learned preconditioner needs the
 * structured path (CFDHIP_ERR_UNSUPPORTED). */
int cfdhip_fmesh_solve_pressure(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, double* x,
                                cfdhip_solve_result* r);
/* correctFaceMassFlux (PressureCorrectionEquation.cpp:369-402): flux[n_faces]
 * = F* + D_f (p'_P - p'_N) (+ the explicit flux when the assembly had one),
 * from pprime[n] (host) or, when NULL, the last solve's p'. */
int cfdhip_fmesh_correct_flux(cfdhip_ctx* ctx, const double* pprime, double* flux);
/* Cell volumes (Cell::volume(), n_cells values) of the face-list mesh: the
 * gradients below need them. */
int cfdhip_fmesh_set_cell_volumes(cfdhip_ctx* ctx, const double* cell_volume);
/* cfd::discretization::gradient (src/discretization/Gradient.cpp:598-607) of
 * a cell field on the face-list mesh (2D): scheme 0 Green-Gauss with the
 * boundary-consistent face values, skewness correction and oblique Neumann
 * faces (greenGaussGradient, :534-596), 1 weighted least squares with the
 * Green-Gauss fallback (leastSquaresGradient, :361-447). field[n] and
 * grad[3 n] (x, y, z per cell) are host arrays; bc_kind / bc_value give each
 * boundary face's scalar condition (1 FixedValue: the value, 0 FixedGradient:
 * the gradient; internal faces ignored); both NULL: the p' conditions of the
 * explicit term (FixedValue 0 on the mesh's FixedValue faces, FixedGradient 0
 * elsewhere, makeGradientBoundaries, PressureCorrectionEquation.cpp:43-54).
 * CFDHIP_ERR_UNSUPPORTED on a 3D mesh. */
int cfdhip_fmesh_gradient(cfdhip_ctx* ctx, const double* field, const int32_t* bc_kind,
                          const double* bc_value, int scheme, double* grad);
/* A corrector pass's assembly (SIMPLE.cpp:880-891): the last
 * cfdhip_fmesh_assemble_pressure's inputs again, with the explicit
 * non-orthogonal flux from the gradient (gradient_scheme as above, the p'
 * conditions) of previous_pprime (host, n values; NULL: the last solve's p'),
 * computed on the device. Needs non_orthogonal = 1. */
int cfdhip_fmesh_assemble_pressure_pass(cfdhip_ctx* ctx, const double* previous_pprime,
                                        int gradient_scheme);
/* The p' corrector passes of SIMPLE.cpp:876-924 (nonOrthogonalCorrections =
 * passes) on the device: the first-pass system (the last
 * cfdhip_fmesh_assemble_pressure, without an explicit term) solved from
 * x0 = 0, then passes 2.. each reassembled with the explicit flux of the
 * previous pass's p' and solved from x0 = 0. results[passes] (may be NULL)
 * receive each solve; a pass that does not converge ends the loop (later
 * results: status -1). x[n] (host, may be NULL) receives the last p';
 * cfdhip_fmesh_correct_flux(NULL) then applies the last pass's explicit flux. */
int cfdhip_fmesh_pressure_passes(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, int passes,
                                 int gradient_scheme, double* x, cfdhip_solve_result* results);
/* solvePressureCorrectionResident: p' from x0 = 0 into PPRIME. */
int cfdhip_solve_pressure(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, cfdhip_solve_result* r);
/* updatePressure: P += alpha * PPRIME (SIMPLE.cpp:939-942). */
int cfdhip_update_pressure(cfdhip_ctx* ctx, double alpha);
/* correctVelocityResident + correctFaceMassFluxResident
 * (PressureCorrectionEquation.cpp:369-438): U/V/W and FLUX. */
int cfdhip_correct(cfdhip_ctx* ctx);
/* correctVelocityResident / correctVelocity (GpuSimpleDiscretization.hpp:136,
 * 252-253; correctVelocity, PressureCorrectionEquation.cpp:404-438): u = u* -
 * d grad(p') from USTAR.., RESPONSE and PPRIME with the p' conditions (Green-
 * Gauss). out NULL: the result replaces U/V/W (the resident form). Otherwise
 * out (host, 3 x n doubles: x, then y, then z; z = 0 in 2D) receives it and
 * U/V/W are left as they were. */
int cfdhip_correct_velocity(cfdhip_ctx* ctx, double* out);
/* correctFaceMassFluxResident / correctFaceMassFlux (GpuSimpleDiscretization.hpp:
 * 137,257; PressureCorrectionEquation.cpp:369-402): F = F* + D_f (p'_P - p'_N)
 * from FLUX_STAR, PC_FACE_COEFF and PPRIME. out NULL: into FLUX; otherwise into
 * out (host, n_faces doubles) with FLUX unchanged. */
int cfdhip_correct_flux(cfdhip_ctx* ctx, double* out);
/* SIMPLE's finiteness guards on resident fields
 * (residentMomentumPredictorAllFinite, residentVelocityAllFinite,
 * residentPressureAllFinite, residentPressureCorrectionAllFinite:
 * GpuSimpleDiscretization.hpp:143,200-201,241; SIMPLE.cpp:645-648,866-869,
 * 1006-1012): *all_finite = 1 when no entry of the `count` listed fields
 * (CFDHIP_FIELD_*) is a NaN or an infinity, else 0. One integer crosses the
 * boundary. */
int cfdhip_fields_all_finite(cfdhip_ctx* ctx, const int* fields, int count, int* all_finite);
/* FluidProperties::density for the flux and p' stages, which the reference
 * passes per call (computePredictedFaceFlux / assemblePressureCorrection*,
 * GpuSimpleDiscretization.hpp:209-210,213-215,230-231); the viscosity field
 * is left as it is. */
int cfdhip_set_density(cfdhip_ctx* ctx, double density);
/* residentSolveBytes (GpuSimpleDiscretization.hpp:245): device bytes held by
 * this context's Krylov work vectors and reduction buffers. */
int cfdhip_resident_bytes(cfdhip_ctx* ctx, int64_t* bytes);
/* evaluateContinuity (src/physics/ContinuityEquation.cpp:14-55): IMBALANCE per
 * cell; out[0] = RMS imbalance, out[1] = global net boundary flux. */
int cfdhip_continuity(cfdhip_ctx* ctx, double* out);

/* One whole SIMPLE run on the device (SIMPLE::solve, SIMPLE.cpp:148-1124,
 * resident loop, laminar, default absolute convergence gate).
 * settings[0..10] = {max_iterations, alpha_u, alpha_p, velocity_tol,
 *   pressure_tol, continuity_tol, rhie_chow (0/1), reference_cell,
 *   use_graph (0/1: the p' solves replay their iteration chunks from a
 *   captured HIP graph, as settings.graph = 1), 0, 0}; momentum/pressure are
 *   the two linear solvers.
 * meta[0..5] out = {status (SIMPLEStatus: 0 converged, 1 max iterations,
 *   2 momentum failure, 3 pressure failure, 4 non-finite), iterations,
 *   momentum linear iterations, pressure linear iterations, last continuity
 *   RMS, last |global imbalance|}; history (5 x cap, may be NULL) =
 *   u, v, w, p initial residuals and continuity RMS per outer iteration. */
int cfdhip_simple_run(cfdhip_ctx* ctx, const double* settings,
                      const cfdhip_solver_settings* momentum,
                      const cfdhip_solver_settings* pressure, double* meta, double* history,
                      int64_t cap);

/* ---- assembled systems as CSR (the host-round-trip assemble* variants,
 *      GpuSimpleDiscretizationCuda.cpp:44-57: exact zeros dropped) ------------ */
/* Returns nnz through *nnz; buffers sized n+1 / capacity. */
/* PISO time steps on the resident state (U, V, P, FLUX): `steps` calls of
 * cfd::pressure_velocity::PISO::solveTimeStep (include/cfd/pressure_velocity/
 * PISO.hpp:89-90; src/pressure_velocity/PISO.cpp:126-350), laminar, 2D. Per
 * step: transient momentum predictor (TransientMomentum.cpp), linear predictor
 * flux, two pressure corrections (p += p', unrelaxed), final continuity. The
 * reference solves both systems with BiCGSTAB whatever the settings' type
 * (PISO.cpp:142-162); here the settings choose (BiCGSTAB reproduces it; CG +
 * Jacobi or FCG + learned use the resident p' solvers). meta: 6 doubles per
 * step {status (cfd::solver::TransientStepStatus: 0 Converged, 1
 * MomentumFailure, 2 PressureCorrectionFailure, 3 NonFiniteState, 4
 * InvalidConfiguration), maxCFL (pre-step), continuityResidual (max cell
 * imbalance), massImbalance (|global net flux|), momentum iterations, pressure
 * iterations}. *taken = steps run; a failed step ends the run and leaves the
 * state of the step before it. */
int cfdhip_piso_run(cfdhip_ctx* ctx, double dt, int64_t steps, int64_t reference_cell,
                    const cfdhip_solver_settings* momentum, const cfdhip_solver_settings* pressure,
                    double* meta, int64_t* taken);
int cfdhip_pressure_csr(cfdhip_ctx* ctx, int64_t* rowptr, int64_t* col, double* val,
                        int64_t capacity, int64_t* nnz);
int cfdhip_momentum_csr(cfdhip_ctx* ctx, int64_t* rowptr, int64_t* col, double* val,
                        int64_t capacity, int64_t* nnz);

/* ---- stand-alone linear algebra (GpuCG / GpuBiCGSTAB,
 *      cuda/kernels/GpuLinearSolverCuda.cpp:164-650) ----------------------------- */

/* y = A x for a host CSR matrix (SparseMatrix::multiply,
 * src/algebra/SparseMatrix.cpp:72-100; csrSpmvKernel). Row sums in ascending
 * column order from 0.0: bitwise equal to the reference. */
int cfdhip_csr_spmv(cfdhip_ctx* ctx, int64_t n, const int64_t* rowptr, const int64_t* col,
                    const double* val, const double* x, double* y);
/* LinearSolver::solve(system, x0) for a host CSR system
 * (include/cfd/algebra/LinearSolver.hpp:171-186): CG / BiCGSTAB with the
 * reference's convergence and breakdown rules. A 5-/7-point structured
 * pattern on the current mesh runs on the stencil kernels; anything else
 * on the CSR kernels. x receives the solution. */
int cfdhip_csr_solve(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, int64_t n,
                     const int64_t* rowptr, const int64_t* col, const double* val,
                     const double* b, const double* x0, double* x, cfdhip_solve_result* r);
/* y = A x with the resident p' operator (PC_DIAG / PC_FACE_COEFF). */
int cfdhip_pressure_spmv(cfdhip_ctx* ctx, const double* x, double* y);
/* y = A x on device buffers (n doubles each), `repeat` times back to back on
 * the context stream (asynchronous; cfdhip_synchronize waits). With timing on
 * (cfdhip_set_timing) the mean kernel time is reported by
 * cfdhip_kernel_timing. SparseMatrix::multiply (SparseMatrix.cpp:72-100) on
 * the resident p' operator, without the host copies of cfdhip_pressure_spmv. */
int cfdhip_pressure_spmv_device(cfdhip_ctx* ctx, const double* x, double* y, int repeat);
/* Solve the resident p' operator for a caller RHS (the stand-alone p' solve):
 * x0 = 0, solution into x (may be NULL to keep it on the device in PPRIME). */
int cfdhip_pressure_solve_rhs(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, const double* b,
                              double* x, cfdhip_solve_result* r);
/* No counterpart in the reference: cfdhip_pressure_solve_rhs for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation). FP64 Jacobi-PCG (type
 * CFDHIP_SOLVER_CG, preconditioner CFDHIP_PREC_JACOBI, precision CFDHIP_FP64)
 * on the resident 2D p' operator from x0 = 0, with the right-hand side b and
 * the solution x in the caller's device memory (n doubles each); the same
 * kernels and iterates as cfdhip_pressure_solve_rhs. Every work buffer is a
 * fixed device array of the library (67.6 MB per GPU, held from the first
 * use of the library on that GPU until the process ends); the call makes no
 * device allocation, creates no HIP event or graph (settings.graph is
 * ignored) and keeps no residual history (history_length 0; device_ms is the
 * host wall time of the solve). Limits: at most 1048576 cells and 65536
 * iterations (CFDHIP_ERR_INVALID beyond); CFDHIP_ERR_UNSUPPORTED for other
 * solver settings or a 3D mesh. All contexts on one GPU share the array:
 * calls that overlap on one GPU race. */
int cfdhip_pressure_solve_device_noalloc(cfdhip_ctx* ctx, const cfdhip_solver_settings* s,
                                         const double* b, double* x, cfdhip_solve_result* r);
/* Residual history of the last solve on this context (up to cap entries). */
int cfdhip_solve_history(cfdhip_ctx* ctx, double* history, int64_t cap, int64_t* length);

/* A caller's preconditioner z = M r applied on the host (cfd::algebra::Preconditioner::apply,
 * include/cfd/algebra/Preconditioner.hpp:16), for CG and BiCGSTAB solves with preconditioner
 * CFDHIP_PREC_HOST on this context (cfdhip_solve_pressure*, cfdhip_solve_momentum,
 * cfdhip_csr_solve): each application copies the one vector to the host and z back, as the
 * reference's GPU solvers do for such a preconditioner
 * (cuda/kernels/GpuLinearSolverCuda.cpp:103-118); the recurrence stays on the device. r and z
 * are host arrays of n doubles, valid during the call. The callback returns 0, or nonzero when
 * it failed (the reference's Preconditioner::apply throwing): the solve then stops and returns
 * CFDHIP_ERR_RUNTIME. fn null clears it. */
typedef int (*cfdhip_host_preconditioner)(const double* r, double* z, int64_t n, void* user);
int cfdhip_set_host_preconditioner(cfdhip_ctx* ctx, cfdhip_host_preconditioner fn, void* user);

/* ---- timing -------------------------------------------------------------------- */
/* Kernel timing of the last solve: out[0] = SpMV kernel launches,
 * out[1] = their mean duration in microseconds (HIP events on the context
 * stream), out[2] = iterations timed. Filled when timing is enabled. */
int cfdhip_set_timing(cfdhip_ctx* ctx, int enabled);
int cfdhip_kernel_timing(cfdhip_ctx* ctx, double* out);
/* Learned preconditioner kernels timed while timing is enabled: 16 entries,
 * the 15 layers (e0a .. out; e0a runs inside e0b's kernel on the solve path)
 * and the coarse tail, total milliseconds and launches since timing was
 * enabled. */
int cfdhip_preconditioner_timing(cfdhip_ctx* ctx, double* ms, int64_t* calls);

/* ---- learned preconditioner ------------------------------------------------------ */
/* No counterpart in the reference: the p' solve with preconditioner
 * CFDHIP_PREC_LEARNED is flexible CG (type CFDHIP_SOLVER_FCG) preconditioned by
 * a linear 4-level convolutional U-Net (3x3 convolutions, 16-64 channels,
 * BF16 weights, FP32 activations split into two BF16 halves at the matrix
 * cores, FP32 accumulation) applied to the Jacobi-scaled residual,
 * z = D^-1/2 N(D^-1/2 r). The Krylov vectors are FP64 (precision CFDHIP_FP64)
 * or FP32 inside FP64 refinement (CFDHIP_MIXED); the solve ends at the
 * requested FP64 true residual, and an inner solve that stops improving for
 * stagnation_window iterations hands the solve to FP64 Jacobi-PCG
 * (fallback_used = 1). 2D p' operator only. */

/* Number of FP32 master weights: layers e0a e0b e1a e1b e2a e2b b3a b3b d2a
 * d2b d1a d1b d0a d0b out, each W[ky][kx][cin][cout]. */
int64_t cfdhip_unet_param_count(void);
/* Loads the network weights used by CFDHIP_PREC_LEARNED on this context. */
int cfdhip_set_learned_weights(cfdhip_ctx* ctx, const float* weights, int64_t count);

/* How the network is applied (no counterpart in the reference).
 * padding 0: every convolution zero-padded, z = D^-1/2 N(D^-1/2 r).
 * padding 1: Neumann (replicate) padding, the padding that matches the p'
 *   operator's zero-gradient walls, and z = S N(S r) with S^-2 a smoothed
 *   interior-equivalent diagonal (the sum of a cell's internal-face couplings
 *   times 2 dim / their count, geometric-mean filtered over a 9 x 9 window;
 *   the pinned cell keeps its diagonal).
 *   The backward pass then needs kernels mirror-symmetric in both axes.
 * tail_levels > 0 (padding 1 only): a one-channel symmetric V(1,1) multigrid
 *   cycle below the bottom level (H/16 .. the coarsest grid with even sides,
 *   at most tail_levels levels), Neumann padding, damped-Jacobi weight
 *   tail_omega, level operator tail_a and transfer kernel tail_k (centre,
 *   edge, corner); it restricts channel 1 of the bottom tensor b3a and writes
 *   the prolonged correction into its channel 2, before b3b.
 * dirichlet_sides (padding 1): sides where the padding is the odd reflection
 *   (zero value at the face) instead of the even one, bits 1 left, 2 right,
 *   4 bottom, 8 top. On a solver context the sides are the case's FixedValue
 *   pressure patches and this field is ignored.
 * dense 0 (default): inference drops the channels no output depends on, and
 *   a network whose every tensor keeps at most 4 live channels (the shipped
 *   V-cycle weights keep 2) runs on narrow FP32 kernels over only those;
 *   dense 1: inference runs every channel through the BF16 MFMA layers.
 *   Both compute the same linear map (they differ in rounding only). */
typedef struct {
  int32_t padding;
  int32_t tail_levels;
  float tail_omega;
  float tail_a[3];
  float tail_k[3];
  int32_t dirichlet_sides;
  int32_t dense;
} cfdhip_unet_options;
int cfdhip_set_learned_options(cfdhip_ctx* ctx, const cfdhip_unet_options* options);
/* Channels stored per pixel by each of the network's 17 forward tensors on
 * this context's inference path (e0a e0 p0 e1a e1 p1 e2a e2 p2 b3a b3 d2a d2
 * d1a d1 d0a d0; the full widths 16 .. 64 on the MFMA layers, 2 or 4 with the
 * unused channels dropped, 0 for a tensor that is not stored), after
 * cfdhip_set_learned_weights. No counterpart in the reference. */
int cfdhip_learned_channels(cfdhip_ctx* ctx, int32_t* stored, int32_t count);

/* Training pairs from the assembled p' system (b: host RHS of n doubles, or
 * NULL for the assembled one): an FP64 Jacobi-PCG solve (settings s) keeps its
 * iterate every `stride` iterations; with x* the converged solution (left in
 * CFDHIP_FIELD_PPRIME) each kept iterate gives e = x* - x_k and r = A e, written
 * in the Jacobi-scaled system (e^ = D^1/2 e, r^ = D^-1/2 r), scaled to unit RMS
 * of r^ and zero-padded to h x w. Device buffers: r_out, e_out (max_pairs x h x
 * w floats), ce_out, cn_out (h x w: scaled couplings to x + 1 and y + 1, may be
 * NULL); host: den_out (max_pairs floats: (e^, r^) per pair). */
int cfdhip_pressure_harvest(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, const double* b,
                            int64_t stride, int64_t max_pairs, int64_t h, int64_t w, float* r_out,
                            float* e_out, float* ce_out, float* cn_out, float* den_out,
                            int64_t* pairs, cfdhip_solve_result* result);

/* White-noise training pairs on the assembled p' operator: e uniform in [-1, 1]
 * per cell (splitmix64 of seed, pair and cell), r = A e, written like the pairs
 * of cfdhip_pressure_harvest (scaled system, unit RMS of r^, padded to h x w;
 * r_out, e_out device, den_out host). They carry the rough error modes a
 * converging solve has already removed. */
int cfdhip_pressure_harvest_noise(cfdhip_ctx* ctx, uint64_t seed, int64_t count, int64_t h,
                                  int64_t w, float* r_out, float* e_out, float* den_out,
                                  int64_t* pairs);

/* A network handle for training and evaluation on caller-owned device buffers
 * (FP32, batch x h x w, h and w multiples of 8). Calls are synchronous. */
typedef struct cfdhip_unet cfdhip_unet;
int cfdhip_unet_create(cfdhip_unet** out);
void cfdhip_unet_destroy(cfdhip_unet* net);
const char* cfdhip_unet_last_error(const cfdhip_unet* net);
int cfdhip_unet_set_weights(cfdhip_unet* net, const float* weights, int64_t count);
/* Padding and coarse tail of this network (cfdhip_unet_options above). */
int cfdhip_unet_set_options(cfdhip_unet* net, const cfdhip_unet_options* options);
int cfdhip_unet_get_weights(cfdhip_unet* net, float* weights, int64_t count);
/* y = N(x). */
int cfdhip_unet_forward(cfdhip_unet* net, const float* x, float* y, int64_t batch, int64_t h,
                        int64_t w);
/* y = N(x) along the preconditioner's inference path (the first layer fused
 * into the second; its output is not kept). Same values as cfdhip_unet_forward.
 * Any h, w >= 1: the network runs on the grid rounded up to multiples of 8 with
 * the image in its corner; each level's padding reflects at that level's
 * edges (rounded-up halves), and a 2x2 pool that straddles an odd edge takes
 * the padding's ghost of the pixel inside (what the learned p' solve does on
 * grids whose sides are not multiples of 8). */
int cfdhip_unet_apply(cfdhip_unet* net, const float* x, float* y, int64_t batch, int64_t h,
                      int64_t w);
/* grad (host, param_count floats) = d/dW of sum(gy * N(x)). */
int cfdhip_unet_backward(cfdhip_unet* net, const float* x, const float* gy, int64_t batch, int64_t h,
                         int64_t w, float* grad);
/* Activation / gradient tensor `tensor` of the last pass (BF16 widened to
 * FP32, NHWC): forward e0a e0 p0 e1a e1 p1 e2a e2 p2 b3a b3 d2a d2 d1a d1 d0a d0
 * (ids 0-16), after a backward pass also the gradients (ids 17-33). */
int64_t cfdhip_unet_tensor_size(cfdhip_unet* net, int tensor);
int cfdhip_unet_tensor(cfdhip_unet* net, int tensor, float* out, int64_t count);
/* A-norm loss L = mean_s (N(x_s) - e_s, A_s (N(x_s) - e_s)) / den_s with A_s the
 * operator of couplings ce, cn (to x + 1, y + 1) and diagonal cd (NULL: unit
 * diagonal), all + sys_offset[s] (device int64 per sample), and its gradient
 * in the master weights (grad: host, param_count floats), no update: for
 * training a parametrisation of the weights on the host. */
int cfdhip_unet_loss_grad(cfdhip_unet* net, const float* x, const float* e, const float* ce,
                          const float* cn, const float* cd, const int64_t* sys_offset,
                          const float* den, int64_t batch, int64_t h, int64_t w, double* loss,
                          float* grad);
/* Multi-step spectral loss of the stationary iteration x <- x + N(b - A x) (no
 * counterpart in the reference): per sample s, with error image z_s and the
 * operator A_s of cfdhip_unet_loss_grad, v_0 = z_s / |z_s|_A,
 * v_j = v_{j-1} - N(A_s v_{j-1}) for j = 1..steps, and
 * L = mean_s log(|v_steps|_A^2) / (2 steps): by Gelfand's formula an estimate of
 * the log of the iteration's spectral radius, which bounds the convergence of
 * the preconditioned Krylov solve from above, where the one-application
 * loss only weighs the average error. grad (host, param_count floats; NULL:
 * the loss only) is dL/dW, accumulated backwards through the steps with
 * N^T = N (exact for a self-adjoint network such as the symmetric V-cycle
 * weights; the Neumann-padded backward pass requires mirror-symmetric
 * kernels). steps in 1..32. */
int cfdhip_unet_spectral_loss_grad(cfdhip_unet* net, const float* z, const float* ce,
                                   const float* cn, const float* cd, const int64_t* sys_offset,
                                   int64_t batch, int64_t h, int64_t w, int32_t steps,
                                   double* loss, float* grad);
/* The same loss; with update != 0 one Adam step (lr, betas 0.9 / 0.999). */
int cfdhip_unet_train_step(cfdhip_unet* net, const float* x, const float* e, const float* ce,
                           const float* cn, const float* cd, const int64_t* sys_offset,
                           const float* den, int64_t batch, int64_t h, int64_t w, float lr,
                           int update, double* loss);

/* ---- batched p' solves ------------------------------------------------------------ */
/* No counterpart in the reference, which solves one case at a time
 * (src/algebra/CG.cpp:14-148 per SIMPLE iteration): many independent 2D p'
 * systems of uneven size, solved together by Jacobi-PCG with one launch per
 * step for the whole batch. Each case reaches |b - A x| <= max(abs, rel |b|)
 * in FP64 from x0 = 0; precision CFDHIP_MIXED runs the inner iterations in
 * FP32 inside FP64 refinement, CFDHIP_FP64 runs them in FP64. Cases that are
 * done leave the launch grid at the next host check (every check_interval
 * steps). The batch owns its stream and its device memory. */
typedef struct cfdhip_batch cfdhip_batch;
int cfdhip_batch_create(int device, cfdhip_batch** out);
void cfdhip_batch_destroy(cfdhip_batch* batch);
const char* cfdhip_batch_last_error(const cfdhip_batch* batch);
int64_t cfdhip_batch_size(const cfdhip_batch* batch);
/* Adds the 2D system of an nx x ny case from host arrays: diagonal (n), face
 * couplings D_f in face order (x-faces j-outer, then y-faces; (nx+1) ny +
 * nx (ny+1) values; boundary faces carry no coupling) and right-hand side
 * (n); pin = the cell whose row is the identity-scaled diagonal, or -1. */
int cfdhip_batch_add_system(cfdhip_batch* batch, int64_t nx, int64_t ny, int64_t pin,
                            const double* diag, const double* face_coeff, const double* rhs,
                            int64_t* index);
/* Adds the p' system a context has assembled (copied on the device). */
int cfdhip_batch_add_pressure(cfdhip_batch* batch, cfdhip_ctx* ctx, int64_t* index);
/* Solves every case (type CFDHIP_SOLVER_CG, preconditioner
 * CFDHIP_PREC_JACOBI). tile_rows: 16, 32 or 64 rows per 64-column tile, or 0
 * to time the three on the batch and take the fastest; use_graph != 0 replays
 * each chunk of steps from a captured HIP graph. results: one per case
 * (device_ms = the case's completion time from the start of the batch);
 * latency_ms: the same times, or NULL; stats (NULL or 8 doubles): device ms of
 * the batch, steps, sum of cells x iterations, tile rows used, graph captures,
 * tuning times of the 16/32/64-row steps (us). */
int cfdhip_batch_solve(cfdhip_batch* batch, const cfdhip_solver_settings* s, int tile_rows,
                       int use_graph, cfdhip_solve_result* results, double* latency_ms,
                       double* stats);
/* The learned preconditioner of the batch's flexible-CG solve (no counterpart
 * in the reference): weights (cfdhip_unet_param_count() floats) and options
 * as for cfdhip_set_learned_weights / cfdhip_set_learned_options, Neumann
 * padding required; each case's Dirichlet sides are its FixedValue pressure
 * patches (from the context for cfdhip_batch_add_pressure; for
 * cfdhip_batch_add_system the sides whose every cell has a diagonal larger
 * than the sum of its internal-face couplings). The weights must leave at
 * most 4 live channels per tensor (the narrow kernels; the shipped V-cycle
 * weights keep 2). With them, cfdhip_batch_solve with type
 * CFDHIP_SOLVER_FCG, preconditioner CFDHIP_PREC_LEARNED and precision
 * CFDHIP_FP64 solves every case (sides up to 1024 cells) by flexible CG with
 * FP64 Krylov vectors, each step one launch over all active cases: the
 * network's layers over (case, tile) tables, the coarse tail one workgroup
 * per case, the vector kernels over (case, block) tables, one-block-per-case
 * scalar steps. A case whose recurrence residual drifted from its FP64 true
 * residual restarts from b - A x; one that stagnated (stagnation_window) or
 * broke down is solved by the batched FP64 Jacobi-PCG (fallback_used = 1).
 * tile_rows is ignored; stats: device ms, rounds, sum of cells x iterations,
 * 0, graph captures. */
int cfdhip_batch_set_learned(cfdhip_batch* batch, const float* weights, int64_t count,
                             const cfdhip_unet_options* options);
/* Copies the solution of case `index` of the last solve to the host. */
int cfdhip_batch_solution(cfdhip_batch* batch, int64_t index, double* host, int64_t count);

#ifdef __cplusplus
}
#endif

#endif /* CFDHIP_H */
