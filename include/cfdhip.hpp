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

// LinearSolver::solve(system, x0) on the device.
class GpuLinearSolver {
 public:
  GpuLinearSolver(LinearSolverSettings settings, detail::CtxPtr ctx)
      : settings_(settings), ctx_(std::move(ctx)) {}

  [[nodiscard]] const LinearSolverSettings& settings() const noexcept { return settings_; }

  [[nodiscard]] SolverResult solve(const CsrView& A, std::span<const double> b,
                                   std::span<const double> x0) const {
    if (b.size() != A.rows || x0.size() != A.rows)
      throw InvalidArgumentError("GpuLinearSolver::solve: size mismatch");
    static_assert(sizeof(std::size_t) == sizeof(int64_t));
    std::vector<double> x(A.rows);
    const cfdhip_solver_settings s = settings_.c();
    cfdhip_solve_result r{};
    detail::check(ctx_.get(),
                  cfdhip_csr_solve(ctx_.get(), &s, static_cast<int64_t>(A.rows),
                                   reinterpret_cast<const int64_t*>(A.rowOffsets),
                                   reinterpret_cast<const int64_t*>(A.columnIndices), A.values,
                                   b.data(), x0.data(), x.data(), &r));
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
  LinearSolverSettings settings_;
  detail::CtxPtr ctx_;
};

namespace detail {
inline std::unique_ptr<GpuLinearSolver> make(LinearSolverSettings s, LinearSolverType type) {
  s.type = type;
  s.validate();  // the reference's constructor validation throws the same way
  if (!deviceAvailable()) return nullptr;
  CtxPtr ctx(cfdhip_create(0));
  if (!ctx) return nullptr;
  return std::make_unique<GpuLinearSolver>(s, std::move(ctx));
}
}  // namespace detail

// makeGpuCG / makeGpuBiCGSTAB: nullptr iff the device is unusable.
[[nodiscard]] inline std::unique_ptr<GpuLinearSolver> makeGpuCG(LinearSolverSettings s) {
  return detail::make(s, LinearSolverType::CG);
}
[[nodiscard]] inline std::unique_ptr<GpuLinearSolver> makeGpuBiCGSTAB(LinearSolverSettings s) {
  return detail::make(s, LinearSolverType::BiCGSTAB);
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

class GpuSimpleDiscretization {
 public:
  GpuSimpleDiscretization() = default;
  GpuSimpleDiscretization(const GpuSimpleDiscretization&) = delete;
  GpuSimpleDiscretization& operator=(const GpuSimpleDiscretization&) = delete;
  GpuSimpleDiscretization(GpuSimpleDiscretization&&) noexcept = default;
  GpuSimpleDiscretization& operator=(GpuSimpleDiscretization&&) noexcept = default;

  [[nodiscard]] bool active() const noexcept { return ctx_ != nullptr; }

  [[nodiscard]] bool prepare(const CartesianMesh& mesh, std::span<const PatchCondition> patches,
                             double density, double viscosity, std::string& reason) {
    ctx_.reset();
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
    if (rc == CFDHIP_OK) rc = cfdhip_set_fluid(ctx.get(), density, viscosity);
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

  void uploadInitialState(std::span<const double> u, std::span<const double> v,
                          std::span<const double> w, std::span<const double> p) {
    upload(CFDHIP_FIELD_U, u);
    upload(CFDHIP_FIELD_V, v);
    if (dim_ == 3) upload(CFDHIP_FIELD_W, w);
    upload(CFDHIP_FIELD_P, p);
    detail::check(ctx(), cfdhip_mass_flux(ctx()));
  }
  void assembleMomentumResident(double alpha) {
    detail::check(ctx(), cfdhip_assemble_momentum(ctx(), alpha));
  }
  [[nodiscard]] SolverResult solveMomentumResident(int component, const LinearSolverSettings& s) {
    const cfdhip_solver_settings c = s.c();
    cfdhip_solve_result r{};
    detail::check(ctx(), cfdhip_solve_momentum(ctx(), component, &c, &r));
    return detail::fromC(r, ctx());  // empty solution: it stays resident
  }
  void computeResponseCoefficients() { detail::check(ctx(), cfdhip_response(ctx())); }
  void computePredictedFaceFlux(bool rhieChow, double alpha) {
    detail::check(ctx(), cfdhip_predicted_flux(ctx(), rhieChow ? 1 : 0, alpha));
  }
  void assemblePressureCorrectionResident(std::size_t referenceCell) {
    detail::check(ctx(), cfdhip_assemble_pressure(ctx(), static_cast<int64_t>(referenceCell)));
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
  void updatePressure(double alpha) { detail::check(ctx(), cfdhip_update_pressure(ctx(), alpha)); }
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
  void correctVelocityAndFluxResident() { detail::check(ctx(), cfdhip_correct(ctx())); }
  // out[0] = RMS imbalance, out[1] = global net boundary flux
  std::pair<double, double> continuity() {
    double out[2] = {0.0, 0.0};
    detail::check(ctx(), cfdhip_continuity(ctx(), out));
    return {out[0], out[1]};
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
  detail::CtxPtr ctx_;
  std::size_t cells_{0}, faces_{0};
  int dim_{2};
};

}  // namespace cfdhip
