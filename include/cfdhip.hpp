// cfdhip.hpp -- C++20 host adapter over the C ABI of libcfdhip (cfdhip.h).
//
// Presents the shapes the reference's SIMPLE loop already calls, so a
// maintainer can bind them behind the reference's own interfaces with a few
// lines of glue per seam:
//   LinearSolverSettings / SolverResult / SolverStatus
//       include/cfd/algebra/LinearSolver.hpp:16-166
//   makeGpuCG / makeGpuBiCGSTAB (nullptr when unusable, never throw)
//       include/cfd/gpu/GpuLinearSolver.hpp:48,52
//   deviceAvailable            include/cfd/gpu/GPUBackend.hpp:15 (cudaAvailable)
//   GpuSimpleDiscretization    include/cfd/gpu/GpuSimpleDiscretization.hpp:56-262
// Conventions follow the reference: numerical outcomes are SolverStatus
// values; invalid arguments throw InvalidArgumentError; a device/runtime
// failure throws NumericalError (include/cfd/gpu/CudaCheck.hpp:24-28).
#pragma once

#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cfdhip.h"

namespace cfdhip {

struct InvalidArgumentError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct NumericalError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum class SolverStatus {
  Converged,
  MaxIterations,
  Breakdown,
  NonFiniteInput,
  NonFiniteResidual,
  InvalidSystem
};
enum class LinearSolverType { CG, BiCGSTAB, GMRES, FCG };
enum class LinearSolverBackend { CPU, GPU };
enum class PreconditionerType { None, Jacobi, Learned };
enum class Precision { FP64, Mixed };

struct LinearSolverSettings {
  double absoluteTolerance{1e-12};
  double relativeTolerance{1e-10};
  std::size_t maxIterations{1000};
  LinearSolverType type{LinearSolverType::BiCGSTAB};
  LinearSolverBackend backend{LinearSolverBackend::GPU};
  PreconditionerType preconditioner{PreconditionerType::None};
  Precision precision{Precision::FP64};
  bool graph{false};  // replay iteration chunks from a captured HIP graph
  // mixed precision: FP32 reduction between FP64 residual recomputations (0: default)
  double innerTolerance{0.0};

  // LinearSolver's constructor validation (src/algebra/LinearSolver.cpp:9-32).
  void validate() const {
    if (absoluteTolerance < 0.0)
      throw InvalidArgumentError("LinearSolverSettings: absoluteTolerance must be >= 0");
    if (relativeTolerance < 0.0)
      throw InvalidArgumentError("LinearSolverSettings: relativeTolerance must be >= 0");
    if (absoluteTolerance == 0.0 && relativeTolerance == 0.0)
      throw InvalidArgumentError(
          "LinearSolverSettings: at least one of absoluteTolerance/relativeTolerance must be > 0");
    if (maxIterations == 0)
      throw InvalidArgumentError("LinearSolverSettings: maxIterations must be >= 1");
    if (!(innerTolerance >= 0.0 && innerTolerance < 1.0))
      throw InvalidArgumentError("LinearSolverSettings: innerTolerance must be in [0, 1)");
  }
  cfdhip_solver_settings c() const {
    cfdhip_solver_settings s{};
    s.absolute_tolerance = absoluteTolerance;
    s.relative_tolerance = relativeTolerance;
    s.max_iterations = static_cast<int64_t>(maxIterations);
    s.type = type == LinearSolverType::CG    ? CFDHIP_SOLVER_CG
             : type == LinearSolverType::FCG ? CFDHIP_SOLVER_FCG
                                             : CFDHIP_SOLVER_BICGSTAB;
    s.preconditioner = preconditioner == PreconditionerType::Jacobi    ? CFDHIP_PREC_JACOBI
                       : preconditioner == PreconditionerType::Learned ? CFDHIP_PREC_LEARNED
                                                                      : CFDHIP_PREC_NONE;
    s.precision = precision == Precision::Mixed ? CFDHIP_MIXED : CFDHIP_FP64;
    s.graph = graph ? 1 : 0;
    s.inner_tolerance = innerTolerance;
    return s;
  }
};

struct SolverResult {
  std::vector<double> solution;
  SolverStatus status{SolverStatus::MaxIterations};
  std::size_t iterations{0};
  double initialResidual{};
  double finalResidual{};
  double trueResidual{};
  std::vector<double> residualHistory;
  std::size_t restarts{0};
  LinearSolverBackend backendUsed{LinearSolverBackend::GPU};
  [[nodiscard]] bool converged() const noexcept { return status == SolverStatus::Converged; }
};

// A CSR system as the reference stores it (SparseMatrix: size_t indices,
// strictly ascending columns, rowOffsets of length rows + 1).
struct CsrView {
  std::size_t rows{0};
  const std::size_t* rowOffsets{nullptr};
  const std::size_t* columnIndices{nullptr};
  const double* values{nullptr};
};

[[nodiscard]] inline bool deviceAvailable() noexcept { return cfdhip_device_available() != 0; }

namespace detail {
struct CtxDeleter {
  void operator()(cfdhip_ctx* c) const noexcept { cfdhip_destroy(c); }
};
using CtxPtr = std::unique_ptr<cfdhip_ctx, CtxDeleter>;

inline void check(cfdhip_ctx* ctx, int rc) {
  if (rc == CFDHIP_OK) return;
  const std::string msg = cfdhip_last_error(ctx);
  if (rc == CFDHIP_ERR_INVALID) throw InvalidArgumentError(msg);
  throw NumericalError(msg);
}

inline SolverResult fromC(const cfdhip_solve_result& r, cfdhip_ctx* ctx) {
  SolverResult out;
  out.status = static_cast<SolverStatus>(r.status);
  out.iterations = static_cast<std::size_t>(r.iterations);
  out.initialResidual = r.initial_residual;
  out.finalResidual = r.final_residual;
  out.trueResidual = r.true_residual;
  out.restarts = static_cast<std::size_t>(r.restarts);
  int64_t len = 0;
  check(ctx, cfdhip_solve_history(ctx, nullptr, 0, &len));
  out.residualHistory.resize(static_cast<std::size_t>(len));
  check(ctx, cfdhip_solve_history(ctx, out.residualHistory.data(), len, &len));
  return out;
}
}  // namespace detail

// A caller's host preconditioner, z = M r on n values
// (cfd::algebra::Preconditioner::apply, include/cfd/algebra/Preconditioner.hpp:16).
using HostPreconditioner = std::function<void(const double* r, double* z, std::size_t n)>;

// LinearSolver::solve(system, x0) on the device.
class GpuLinearSolver {
 public:
  GpuLinearSolver(LinearSolverSettings settings, detail::CtxPtr ctx)
      : settings_(settings), ctx_(std::move(ctx)) {}
  // With a host preconditioner (the reference's makeGpuCG(settings, preconditioner) with a
  // Preconditioner it has no device form of): every application crosses to the host and back.
  GpuLinearSolver(LinearSolverSettings settings, detail::CtxPtr ctx, HostPreconditioner pc)
      : settings_(settings), ctx_(std::move(ctx)),
        pc_(std::make_unique<HostCall>(HostCall{std::move(pc), nullptr})) {
    detail::check(ctx_.get(), cfdhip_set_host_preconditioner(ctx_.get(), &trampoline,
                                                             pc_.get()));
  }

  [[nodiscard]] const LinearSolverSettings& settings() const noexcept { return settings_; }

  [[nodiscard]] SolverResult solve(const CsrView& A, std::span<const double> b,
                                   std::span<const double> x0) const {
    if (b.size() != A.rows || x0.size() != A.rows)
      throw InvalidArgumentError("GpuLinearSolver::solve: size mismatch");
    static_assert(sizeof(std::size_t) == sizeof(int64_t));
    std::vector<double> x(A.rows);
    cfdhip_solver_settings s = settings_.c();
    if (pc_) s.preconditioner = CFDHIP_PREC_HOST;
    cfdhip_solve_result r{};
    if (pc_) pc_->error = nullptr;
    const int rc = cfdhip_csr_solve(ctx_.get(), &s, static_cast<int64_t>(A.rows),
                                    reinterpret_cast<const int64_t*>(A.rowOffsets),
                                    reinterpret_cast<const int64_t*>(A.columnIndices), A.values,
                                    b.data(), x0.data(), x.data(), &r);
    // what the host preconditioner threw leaves solve() as it would the reference's
    if (pc_ && pc_->error) std::rethrow_exception(std::exchange(pc_->error, nullptr));
    detail::check(ctx_.get(), rc);
    SolverResult out = detail::fromC(r, ctx_.get());
    out.solution = std::move(x);
    return out;
  }

  // Convenience overload: x0 = 0 (LinearSolver::solve(system)).
  [[nodiscard]] SolverResult solve(const CsrView& A, std::span<const double> b) const {
    const std::vector<double> zero(A.rows, 0.0);
    return solve(A, b, zero);
  }

 private:
  struct HostCall {
    HostPreconditioner fn;
    std::exception_ptr error;  // what fn threw, rethrown by solve()
  };
  // No exception crosses the C boundary: it is kept and the library told to stop.
  static int trampoline(const double* r, double* z, int64_t n, void* user) {
    auto* call = static_cast<HostCall*>(user);
    try {
      call->fn(r, z, static_cast<std::size_t>(n));
      return 0;
    } catch (...) {
      call->error = std::current_exception();
      return 1;
    }
  }
  LinearSolverSettings settings_;
  detail::CtxPtr ctx_;
  std::unique_ptr<HostCall> pc_;  // stable address for the C callback
};

namespace detail {
inline std::unique_ptr<GpuLinearSolver> make(LinearSolverSettings s, LinearSolverType type,
                                             HostPreconditioner pc = nullptr) {
  s.type = type;
  s.validate();  // the reference's constructor validation throws the same way
  if (!deviceAvailable()) return nullptr;
  CtxPtr ctx(cfdhip_create(0));
  if (!ctx) return nullptr;
  if (pc) return std::make_unique<GpuLinearSolver>(s, std::move(ctx), std::move(pc));
  return std::make_unique<GpuLinearSolver>(s, std::move(ctx));
}
}  // namespace detail

// makeGpuCG / makeGpuBiCGSTAB: nullptr iff the device is unusable. The overloads with
// a host preconditioner apply it in place of the settings' preconditioner.
[[nodiscard]] inline std::unique_ptr<GpuLinearSolver> makeGpuCG(LinearSolverSettings s) {
  return detail::make(s, LinearSolverType::CG);
}
[[nodiscard]] inline std::unique_ptr<GpuLinearSolver> makeGpuBiCGSTAB(LinearSolverSettings s) {
  return detail::make(s, LinearSolverType::BiCGSTAB);
}
[[nodiscard]] inline std::unique_ptr<GpuLinearSolver> makeGpuCG(LinearSolverSettings s,
                                                                HostPreconditioner pc) {
  return detail::make(s, LinearSolverType::CG, std::move(pc));
}
[[nodiscard]] inline std::unique_ptr<GpuLinearSolver> makeGpuBiCGSTAB(LinearSolverSettings s,
                                                                      HostPreconditioner pc) {
  return detail::make(s, LinearSolverType::BiCGSTAB, std::move(pc));
}

// Boundary condition of one patch as prepare() receives it.
struct PatchCondition {
  int velocityKind{CFDHIP_WALL};
  double velocity[3]{0.0, 0.0, 0.0};
  int pressureKind{CFDHIP_FIXED_GRADIENT};
  double pressureValue{0.0};
};

// Cartesian description a maintainer extracts from Mesh::structuredGrid().
struct CartesianMesh {
  int dim{2};
  std::size_t nx{0}, ny{0}, nz{1};
  double lx{1.0}, ly{1.0}, lz{1.0};
};

// The resident SIMPLE stages (GpuSimpleDiscretization): pimpl over one context,
// non-copyable, movable; prepare() is all-or-nothing with a reason.
// cfd::solver::TransientStepStatus / TransientStepResult
// (include/cfd/solver/TransientSolver.hpp:29-53) without the state, which
// stays resident (download it with GpuSimpleDiscretization::download).
enum class TransientStepStatus {
  Converged,
  MomentumFailure,
  PressureCorrectionFailure,
  NonFiniteState,
  InvalidConfiguration,
};
struct TransientStepResult {
  TransientStepStatus status{TransientStepStatus::InvalidConfiguration};
  double maxCFL{};
  double continuityResidual{};
  double massImbalance{};
  std::size_t momentumIterations{}, pressureIterations{};
};

// A system assembled on the device and brought back for a host solver: the
// CSR arrays SparseMatrix takes (size_t indices, ascending columns, exact zeros
// dropped, GpuSimpleDiscretizationCuda.cpp:44-57) and the right-hand side.
struct HostSystem {
  std::size_t rows{0};
  std::vector<std::size_t> rowOffsets, columnIndices;
  std::vector<double> values, rhs;
};

class GpuSimpleDiscretization {
 public:
  // GpuSimpleDiscretization::FieldAuthority / PersistentField
  // (include/cfd/gpu/GpuSimpleDiscretization.hpp:89-98): which side holds the
  // authoritative copy of a persistent field. Host bookkeeping only.
  enum class FieldAuthority { HostOnly, DeviceOwned, Synchronized };
  enum class PersistentField { Velocity, Pressure, MassFlux, Viscosity };

  GpuSimpleDiscretization() = default;
  GpuSimpleDiscretization(const GpuSimpleDiscretization&) = delete;
  GpuSimpleDiscretization& operator=(const GpuSimpleDiscretization&) = delete;
  GpuSimpleDiscretization(GpuSimpleDiscretization&&) noexcept = default;
  GpuSimpleDiscretization& operator=(GpuSimpleDiscretization&&) noexcept = default;

  [[nodiscard]] bool active() const noexcept { return ctx_ != nullptr; }

  // prepare(): all or nothing, with the reason when the mesh or a condition
  // is outside the structured path. This overload also sets a uniform
  // laminar viscosity and the density.
  [[nodiscard]] bool prepare(const CartesianMesh& mesh, std::span<const PatchCondition> patches,
                             double density, double viscosity, std::string& reason) {
    if (!prepare(mesh, patches, reason)) return false;
    const int rc = cfdhip_set_fluid(ctx_.get(), density, viscosity);
    if (rc != CFDHIP_OK) {
      reason = cfdhip_last_error(ctx_.get());
      ctx_.reset();
      return false;
    }
    viscosityAuthority_ = FieldAuthority::Synchronized;
    return true;
  }
  // The reference's prepare(mesh, velocityBoundaries, pressureBoundaries,
  // reason): the fluid arrives later, the viscosity with uploadInitialState
  // and the density with the flux and p' stages.
  [[nodiscard]] bool prepare(const CartesianMesh& mesh, std::span<const PatchCondition> patches,
                             std::string& reason) {
    ctx_.reset();
    resetAuthority();
    if (!deviceAvailable()) {
      reason = "no usable gfx950 device";
      return false;
    }
    detail::CtxPtr ctx(cfdhip_create(0));
    if (!ctx) {
      reason = "device context creation failed";
      return false;
    }
    int rc = cfdhip_mesh_cartesian(ctx.get(), mesh.dim, static_cast<int64_t>(mesh.nx),
                                   static_cast<int64_t>(mesh.ny), static_cast<int64_t>(mesh.nz),
                                   mesh.lx, mesh.ly, mesh.lz);
    int64_t sizes[4] = {0, 0, 0, 0};
    if (rc == CFDHIP_OK) rc = cfdhip_mesh_sizes(ctx.get(), sizes);
    if (rc == CFDHIP_OK && static_cast<int64_t>(patches.size()) != sizes[2]) {
      reason = "one boundary condition per patch is required";
      return false;
    }
    for (std::size_t p = 0; rc == CFDHIP_OK && p < patches.size(); ++p)
      rc = cfdhip_set_boundary(ctx.get(), static_cast<int>(p), patches[p].velocityKind,
                               patches[p].velocity, patches[p].pressureKind,
                               patches[p].pressureValue);
    if (rc != CFDHIP_OK) {
      reason = cfdhip_last_error(ctx.get());
      return false;
    }
    cells_ = static_cast<std::size_t>(sizes[0]);
    faces_ = static_cast<std::size_t>(sizes[1]);
    dim_ = static_cast<int>(sizes[3]);
    ctx_ = std::move(ctx);
    return true;
  }

  [[nodiscard]] std::size_t cells() const noexcept { return cells_; }
  [[nodiscard]] std::size_t faces() const noexcept { return faces_; }
  [[nodiscard]] int dimension() const noexcept { return dim_; }
  // The device's mesh: cell centroids (x, y, z per cell) and volumes, face
  // owner / neighbour (-1 on a boundary face), in the reference's order.
  void meshCells(std::span<double> centroid, std::span<double> volume) {
    if (centroid.size() != 3 * cells_ || volume.size() != cells_)
      throw InvalidArgumentError("meshCells: output sizes do not match the mesh");
    detail::check(ctx(), cfdhip_mesh_cells(ctx(), centroid.data(), volume.data()));
  }
  void meshFaces(std::span<int64_t> owner, std::span<int64_t> neighbor) {
    if (owner.size() != faces_ || neighbor.size() != faces_)
      throw InvalidArgumentError("meshFaces: output sizes do not match the mesh");
    detail::check(ctx(), cfdhip_mesh_faces(ctx(), owner.data(), neighbor.data()));
  }

  [[nodiscard]] FieldAuthority authority(PersistentField field) const noexcept {
    switch (field) {
      case PersistentField::Velocity: return velocityAuthority_;
      case PersistentField::Pressure: return pressureAuthority_;
      case PersistentField::MassFlux: return massFluxAuthority_;
      case PersistentField::Viscosity: return viscosityAuthority_;
    }
    return FieldAuthority::HostOnly;
  }

  // Velocity components and pressure; the mass flux is computed on the device
  // from the velocity (calculateMassFlux, as SIMPLE.cpp:229 does before the
  // loop). An empty `w` leaves W at zero.
  void uploadInitialState(std::span<const double> u, std::span<const double> v,
                          std::span<const double> w, std::span<const double> p) {
    upload(CFDHIP_FIELD_U, u);
    upload(CFDHIP_FIELD_V, v);
    if (!w.empty()) upload(CFDHIP_FIELD_W, w);
    upload(CFDHIP_FIELD_P, p);
    detail::check(ctx(), cfdhip_mass_flux(ctx()));
    markUploaded(false);
  }
  // uploadInitialState(velocity, pressure, massFlux, effectiveViscosity)
  // (GpuSimpleDiscretization.hpp:104-107): the host state becomes
  // authoritative for all four fields.
  void uploadInitialState(std::span<const double> u, std::span<const double> v,
                          std::span<const double> w, std::span<const double> p,
                          std::span<const double> massFlux,
                          std::span<const double> effectiveViscosity) {
    upload(CFDHIP_FIELD_U, u);
    upload(CFDHIP_FIELD_V, v);
    if (!w.empty()) upload(CFDHIP_FIELD_W, w);
    upload(CFDHIP_FIELD_P, p);
    upload(CFDHIP_FIELD_FLUX, massFlux);
    upload(CFDHIP_FIELD_VISCOSITY, effectiveViscosity);
    markUploaded(true);
  }
  // beginIteration (GpuSimpleDiscretization.hpp:155-158) = uploadInitialState.
  void beginIteration(std::span<const double> u, std::span<const double> v,
                      std::span<const double> w, std::span<const double> p,
                      std::span<const double> massFlux,
                      std::span<const double> effectiveViscosity) {
    uploadInitialState(u, v, w, p, massFlux, effectiveViscosity);
  }
  // beginIterationResident: no transfer; the device continues from its own
  // corrected velocity and flux and updated pressure.
  void beginIterationResident() {
    if (!active()) return;
    if (!stateUploaded_)
      throw InvalidArgumentError(
          "GpuSimpleDiscretization::beginIterationResident: uploadInitialState has not run, so "
          "there is no resident state to continue from");
  }
  void setViscosity(std::span<const double> effectiveViscosity) {
    if (!active()) return;
    upload(CFDHIP_FIELD_VISCOSITY, effectiveViscosity);
    viscosityAuthority_ = FieldAuthority::Synchronized;
  }

  // Relaxed momentum of all components in one pass (the matrix is shared).
  void assembleMomentumResident(double alpha) {
    detail::check(ctx(), cfdhip_assemble_momentum(ctx(), alpha));
  }
  // assembleMomentumResident(component, alpha, convectionScheme,
  // applyNonOrthogonalCorrection): phi_old from the resident component.
  void assembleMomentumResident(int component, double alpha, int convectionScheme,
                                bool applyNonOrthogonalCorrection) {
    detail::check(ctx(), cfdhip_assemble_momentum_component(ctx(), component, alpha,
                                                            convectionScheme,
                                                            applyNonOrthogonalCorrection ? 1 : 0,
                                                            nullptr));
  }
  // assembleMomentum: the same assembly with the caller's phi_old, and the
  // system brought back for a host solver.
  [[nodiscard]] HostSystem assembleMomentum(int component,
                                            std::span<const double> previousComponent,
                                            double alpha, int convectionScheme,
                                            bool applyNonOrthogonalCorrection) {
    if (previousComponent.size() != cells_)
      throw InvalidArgumentError("assembleMomentum: previousComponent size does not match the mesh");
    detail::check(ctx(), cfdhip_assemble_momentum_component(ctx(), component, alpha,
                                                            convectionScheme,
                                                            applyNonOrthogonalCorrection ? 1 : 0,
                                                            previousComponent.data()));
    HostSystem sys = exportCsr(cfdhip_momentum_csr);
    sys.rhs.resize(cells_);
    download(CFDHIP_FIELD_MOM_RHS_U + component, sys.rhs);
    return sys;
  }
  // setMomentumSolution: a host solver's u* back into the device predictor.
  void setMomentumSolution(int component, std::span<const double> solution) {
    if (!active()) return;
    if (component < 0 || component >= dim_)
      throw InvalidArgumentError("setMomentumSolution: component out of range");
    upload(CFDHIP_FIELD_USTAR + component, solution);
  }
  [[nodiscard]] SolverResult solveMomentumResident(int component, const LinearSolverSettings& s) {
    const cfdhip_solver_settings c = s.c();
    cfdhip_solve_result r{};
    detail::check(ctx(), cfdhip_solve_momentum(ctx(), component, &c, &r));
    return detail::fromC(r, ctx());  // empty solution: it stays resident
  }
  [[nodiscard]] bool residentMomentumPredictorAllFinite(bool threeDimensional) {
    const int f[3] = {CFDHIP_FIELD_USTAR, CFDHIP_FIELD_VSTAR, CFDHIP_FIELD_WSTAR};
    return allFinite(f, threeDimensional ? 3 : 2);
  }
  // U, V and W whatever the dimension (W holds zeros in 2D), as the host
  // allFinite(VectorField) tests all three.
  [[nodiscard]] bool residentVelocityAllFinite() {
    const int f[3] = {CFDHIP_FIELD_U, CFDHIP_FIELD_V, CFDHIP_FIELD_W};
    return allFinite(f, 3);
  }
  [[nodiscard]] bool residentPressureAllFinite() {
    if (!active()) return true;
    const int f[1] = {CFDHIP_FIELD_P};
    return allFinite(f, 1);
  }
  [[nodiscard]] bool residentPressureCorrectionAllFinite() {
    const int f[1] = {CFDHIP_FIELD_PPRIME};
    return allFinite(f, 1);
  }

  void computeResponseCoefficients() { detail::check(ctx(), cfdhip_response(ctx())); }
  void computeResponseCoefficients(bool /*threeDimensional: the mesh's own dimension rules*/) {
    computeResponseCoefficients();
  }
  void computePredictedFaceFlux(bool rhieChow, double alpha) {
    detail::check(ctx(), cfdhip_predicted_flux(ctx(), rhieChow ? 1 : 0, alpha));
  }
  // computePredictedFaceFlux(rhieChow, density, alpha, gradientScheme, 3D):
  // the Rhie-Chow pressure gradient is Green-Gauss, as the reference's device
  // path computes it whatever gradientScheme says
  // (GpuSimpleDiscretizationCuda.cpp:493-499).
  void computePredictedFaceFlux(bool rhieChow, double density, double alpha,
                                int /*gradientScheme*/, bool /*threeDimensional*/) {
    detail::check(ctx(), cfdhip_set_density(ctx(), density));
    computePredictedFaceFlux(rhieChow, alpha);
  }

  void assemblePressureCorrectionResident(std::size_t referenceCell) {
    detail::check(ctx(), cfdhip_assemble_pressure(ctx(), static_cast<int64_t>(referenceCell)));
  }
  void assemblePressureCorrectionResident(bool nonOrthogonal, std::size_t referenceCell,
                                          double density, bool /*threeDimensional*/) {
    requireOrthogonal(nonOrthogonal, "assemblePressureCorrectionResident");
    detail::check(ctx(), cfdhip_set_density(ctx(), density));
    assemblePressureCorrectionResident(referenceCell);
  }
  // assemblePressureCorrection: pass 1 only (previousPressureCorrection must be
  // empty; SIMPLE keeps corrector passes on the CPU), brought back to the host.
  [[nodiscard]] HostSystem assemblePressureCorrection(
      bool nonOrthogonal, std::size_t referenceCell, double density,
      std::span<const double> previousPressureCorrection, bool threeDimensional) {
    if (!previousPressureCorrection.empty())
      throw InvalidArgumentError(
          "assemblePressureCorrection: corrector passes (a previous p') are not on the "
          "structured path");
    assemblePressureCorrectionResident(nonOrthogonal, referenceCell, density, threeDimensional);
    HostSystem sys = exportCsr(cfdhip_pressure_csr);
    sys.rhs.resize(cells_);
    download(CFDHIP_FIELD_PC_RHS, sys.rhs);
    return sys;
  }
  // assembleCompressiblePressureCorrection (2D): per-face density (faces),
  // absolute pressure and temperature (cells) of an ideal gas with constant R.
  void assembleCompressiblePressureCorrectionResident(std::span<const double> faceDensity,
                                                      std::span<const double> pressureAbsolute,
                                                      std::span<const double> temperature,
                                                      double gasConstant, double pseudoTimeStep,
                                                      std::size_t referenceCell) {
    if (faceDensity.size() != faces_ || pressureAbsolute.size() != cells_ ||
        temperature.size() != cells_)
      throw InvalidArgumentError(
          "assembleCompressiblePressureCorrection: field size does not match the mesh");
    detail::check(ctx(), cfdhip_assemble_compressible_pressure(
                             ctx(), static_cast<int64_t>(referenceCell), faceDensity.data(),
                             pressureAbsolute.data(), temperature.data(), gasConstant,
                             pseudoTimeStep));
  }
  [[nodiscard]] SolverResult solvePressureCorrectionResident(const LinearSolverSettings& s) {
    const cfdhip_solver_settings c = s.c();
    cfdhip_solve_result r{};
    detail::check(ctx(), cfdhip_solve_pressure(ctx(), &c, &r));
    return detail::fromC(r, ctx());
  }
  // setPressureCorrection: a host solver's p' back to the device.
  void setPressureCorrection(std::span<const double> pressureCorrection) {
    if (!active()) return;
    upload(CFDHIP_FIELD_PPRIME, pressureCorrection);
  }
  [[nodiscard]] std::size_t residentSolveBytes() const noexcept {
    if (!ctx_) return 0;
    int64_t b = 0;
    return cfdhip_resident_bytes(ctx_.get(), &b) == CFDHIP_OK ? static_cast<std::size_t>(b) : 0;
  }
  void updatePressure(double alpha) {
    detail::check(ctx(), cfdhip_update_pressure(ctx(), alpha));
    pressureAuthority_ = FieldAuthority::DeviceOwned;
  }
  // PISO::solveTimeStep (include/cfd/pressure_velocity/PISO.hpp:89-90) on the
  // resident state (U, V, P, FLUX), 2D; the solvers' types are honoured (the
  // reference always uses BiCGSTAB). A failed step leaves the previous state.
  [[nodiscard]] TransientStepResult solveTimeStepResident(double dt,
                                                          const LinearSolverSettings& momentum,
                                                          const LinearSolverSettings& pressure,
                                                          std::size_t referenceCell = 0) {
    const cfdhip_solver_settings m = momentum.c(), p = pressure.c();
    double meta[6] = {};
    int64_t taken = 0;
    detail::check(ctx(), cfdhip_piso_run(ctx(), dt, 1, static_cast<int64_t>(referenceCell), &m,
                                         &p, meta, &taken));
    TransientStepResult r;
    r.status = static_cast<TransientStepStatus>(static_cast<int>(meta[0]));
    r.maxCFL = meta[1];
    r.continuityResidual = meta[2];
    r.massImbalance = meta[3];
    r.momentumIterations = static_cast<std::size_t>(meta[4]);
    r.pressureIterations = static_cast<std::size_t>(meta[5]);
    return r;
  }
  // correctVelocityResident / correctFaceMassFluxResident: the corrections
  // replace the persistent velocity / mass flux on the device. Green-Gauss is
  // the structured path's gradient; the explicit non-orthogonal flux term
  // belongs to corrector passes, which SIMPLE keeps on the CPU.
  void correctVelocityResident(int gradientScheme, bool /*threeDimensional*/) {
    requireGreenGauss(gradientScheme, "correctVelocityResident");
    detail::check(ctx(), cfdhip_correct_velocity(ctx(), nullptr));
    velocityAuthority_ = FieldAuthority::DeviceOwned;
  }
  void correctFaceMassFluxResident(bool withExplicitTerm) {
    requireOrthogonal(withExplicitTerm, "correctFaceMassFluxResident");
    detail::check(ctx(), cfdhip_correct_flux(ctx(), nullptr));
    massFluxAuthority_ = FieldAuthority::DeviceOwned;
  }
  void correctVelocityAndFluxResident() {
    detail::check(ctx(), cfdhip_correct(ctx()));
    velocityAuthority_ = massFluxAuthority_ = FieldAuthority::DeviceOwned;
  }
  // correctVelocity / correctFaceMassFlux: the corrected fields come back to
  // the host (velocity: x, then y, then z, each of cells()); the persistent
  // fields are left as they were.
  void correctVelocity(int gradientScheme, bool /*threeDimensional*/, std::span<double> corrected) {
    requireGreenGauss(gradientScheme, "correctVelocity");
    if (corrected.size() != 3 * cells_)
      throw InvalidArgumentError("correctVelocity: output must hold 3 x cells values");
    detail::check(ctx(), cfdhip_correct_velocity(ctx(), corrected.data()));
  }
  void correctFaceMassFlux(bool withExplicitTerm, std::span<double> corrected) {
    requireOrthogonal(withExplicitTerm, "correctFaceMassFlux");
    if (corrected.size() != faces_)
      throw InvalidArgumentError("correctFaceMassFlux: output must hold one value per face");
    detail::check(ctx(), cfdhip_correct_flux(ctx(), corrected.data()));
  }
  // out[0] = RMS imbalance, out[1] = global net boundary flux
  std::pair<double, double> continuity() {
    double out[2] = {0.0, 0.0};
    detail::check(ctx(), cfdhip_continuity(ctx(), out));
    return {out[0], out[1]};
  }
  // downloadVelocity / downloadPressure / downloadMassFlux: the host copy
  // becomes current (Synchronized).
  void downloadVelocity(std::span<double> u, std::span<double> v, std::span<double> w) {
    download(CFDHIP_FIELD_U, u);
    download(CFDHIP_FIELD_V, v);
    download(CFDHIP_FIELD_W, w);
    velocityAuthority_ = FieldAuthority::Synchronized;
  }
  void downloadPressure(std::span<double> p) {
    download(CFDHIP_FIELD_P, p);
    pressureAuthority_ = FieldAuthority::Synchronized;
  }
  void downloadMassFlux(std::span<double> flux) {
    download(CFDHIP_FIELD_FLUX, flux);
    massFluxAuthority_ = FieldAuthority::Synchronized;
  }
  void download(int field, std::span<double> out) {
    detail::check(ctx(), cfdhip_field_download(ctx(), field, out.data(),
                                               static_cast<int64_t>(out.size())));
  }
  void upload(int field, std::span<const double> in) {
    detail::check(ctx(), cfdhip_field_upload(ctx(), field, in.data(),
                                             static_cast<int64_t>(in.size())));
  }

 private:
  cfdhip_ctx* ctx() const {
    if (!ctx_) throw InvalidArgumentError("GpuSimpleDiscretization: prepare() first");
    return ctx_.get();
  }
  bool allFinite(const int* fields, int count) {
    int ok = 0;
    detail::check(ctx(), cfdhip_fields_all_finite(ctx(), fields, count, &ok));
    return ok != 0;
  }
  static void requireGreenGauss(int gradientScheme, const char* what) {
    if (gradientScheme != 0)  // cfd::discretization::GradientScheme::GreenGauss
      throw InvalidArgumentError(std::string(what) +
                                 ": the structured path's gradient is Green-Gauss");
  }
  static void requireOrthogonal(bool nonOrthogonal, const char* what) {
    if (nonOrthogonal)
      throw InvalidArgumentError(std::string(what) +
                                 ": non-orthogonal correction is not on the structured path");
  }
  template <class Export>
  HostSystem exportCsr(Export fn) {
    HostSystem sys;
    sys.rows = cells_;
    sys.rowOffsets.resize(cells_ + 1);
    const std::size_t cap = (2 * static_cast<std::size_t>(dim_) + 1) * cells_;
    std::vector<int64_t> col(cap);
    sys.values.resize(cap);
    int64_t nnz = 0;
    static_assert(sizeof(std::size_t) == sizeof(int64_t));
    detail::check(ctx(), fn(ctx(), reinterpret_cast<int64_t*>(sys.rowOffsets.data()), col.data(),
                            sys.values.data(), static_cast<int64_t>(cap), &nnz));
    sys.values.resize(static_cast<std::size_t>(nnz));
    sys.columnIndices.assign(col.begin(), col.begin() + nnz);
    return sys;
  }
  void markUploaded(bool viscosity) {
    velocityAuthority_ = pressureAuthority_ = massFluxAuthority_ = FieldAuthority::Synchronized;
    if (viscosity) viscosityAuthority_ = FieldAuthority::Synchronized;
    stateUploaded_ = true;
  }
  void resetAuthority() {
    velocityAuthority_ = pressureAuthority_ = massFluxAuthority_ = viscosityAuthority_ =
        FieldAuthority::HostOnly;
    stateUploaded_ = false;
  }
  detail::CtxPtr ctx_;
  std::size_t cells_{0}, faces_{0};
  int dim_{2};
  FieldAuthority velocityAuthority_{FieldAuthority::HostOnly};
  FieldAuthority pressureAuthority_{FieldAuthority::HostOnly};
  FieldAuthority massFluxAuthority_{FieldAuthority::HostOnly};
  FieldAuthority viscosityAuthority_{FieldAuthority::HostOnly};
  bool stateUploaded_{false};
};

}  // namespace cfdhip
