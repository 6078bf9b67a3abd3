// The reference's GPU factory seams bound to libcfdhip, compiled against the
// reference's own headers and linked in place of its CPU stubs
// src/gpu/GPUBackend.cpp and src/gpu/GpuLinearSolver.cpp:
//   cfd::gpu::cudaAvailable             include/cfd/gpu/GPUBackend.hpp:15
//   cfd::gpu::makeGpuCG / makeGpuBiCGSTAB include/cfd/gpu/GpuLinearSolver.hpp:48-54
// The solvers run cfdhip_csr_solve on the device through the C++ adapter in
// include/cfdhip.hpp; a SolverResult comes back with the reference's fields.
#include <memory>
#include <span>
#include <utility>

#include "cfd/algebra/LinearSolver.hpp"
#include "cfd/algebra/LinearSystem.hpp"
#include "cfd/algebra/Preconditioner.hpp"
#include "cfd/algebra/SparseMatrix.hpp"
#include "cfd/core/Exception.hpp"
#include "cfd/gpu/GPUBackend.hpp"
#include "cfd/gpu/GpuLinearSolver.hpp"
#include "cfdhip.hpp"

namespace cfd::gpu {

bool cudaAvailable() { return cfdhip::deviceAvailable(); }

namespace {

using cfd::algebra::LinearSolver;
using cfd::algebra::LinearSolverSettings;
using cfd::algebra::LinearSystem;
using cfd::algebra::Preconditioner;
using cfd::algebra::PreconditionerType;
using cfd::algebra::SolverResult;
using cfd::algebra::Vector;

class HipKrylovSolver final : public LinearSolver {
 public:
  HipKrylovSolver(LinearSolverSettings settings, std::unique_ptr<cfdhip::GpuLinearSolver> device,
                  std::shared_ptr<Preconditioner> host = nullptr)
      : LinearSolver(settings), device_(std::move(device)), host_(std::move(host)) {}

  SolverResult solve(const LinearSystem& system, const Vector& initialGuess) const override {
    const auto& A = system.matrix();
    if (initialGuess.size() != system.size())
      throw cfd::InvalidArgumentError("LinearSolver::solve: initial guess size mismatch");
    // a host preconditioner is built on this system's matrix, as the reference's
    // solvers do before they iterate; one that rejects it makes the result
    // InvalidSystem (cuda/kernels/GpuLinearSolverCuda.cpp:130-143)
    if (host_) {
      try {
        host_->build(A);
      } catch (const cfd::InvalidArgumentError&) {
        SolverResult bad;
        bad.solution = initialGuess;
        bad.status = cfd::algebra::SolverStatus::InvalidSystem;
        bad.backendUsed = cfd::algebra::LinearSolverBackend::GPU;
        return bad;
      }
    }
    const cfdhip::CsrView view{A.rows(), A.rowOffsetsData(), A.columnIndicesData(),
                               A.valuesData()};
    cfdhip::SolverResult r;
    try {
      r = device_->solve(view, std::span<const double>(system.rhs().data(), system.size()),
                         std::span<const double>(initialGuess.data(), initialGuess.size()));
    } catch (const cfdhip::InvalidArgumentError& e) {
      throw cfd::InvalidArgumentError(e.what());
    } catch (const cfdhip::NumericalError& e) {
      throw cfd::NumericalError(e.what());
    }
    SolverResult out;
    out.solution = Vector(r.solution.size());
    for (std::size_t i = 0; i < r.solution.size(); ++i) out.solution[i] = r.solution[i];
    out.status = static_cast<cfd::algebra::SolverStatus>(static_cast<int>(r.status));
    out.iterations = r.iterations;
    out.initialResidual = r.initialResidual;
    out.finalResidual = r.finalResidual;
    out.residualHistory = std::move(r.residualHistory);
    out.restarts = r.restarts;
    out.backendUsed = cfd::algebra::LinearSolverBackend::GPU;
    return out;
  }

 private:
  std::unique_ptr<cfdhip::GpuLinearSolver> device_;
  std::shared_ptr<Preconditioner> host_;  // applied by the device solver's host callback
};

cfdhip::LinearSolverSettings toDevice(const LinearSolverSettings& s, bool jacobi) {
  cfdhip::LinearSolverSettings d;
  d.absoluteTolerance = s.absoluteTolerance;
  d.relativeTolerance = s.relativeTolerance;
  d.maxIterations = s.maxIterations;
  d.backend = cfdhip::LinearSolverBackend::GPU;
  d.preconditioner = jacobi ? cfdhip::PreconditionerType::Jacobi : cfdhip::PreconditionerType::None;
  return d;
}

// The caller's preconditioner: a JacobiPreconditioner is diag(A)^-1, which
// the device builds from the same matrix (the settings' Jacobi path);
// IdentityPreconditioner is no preconditioner. Any other host preconditioner
// has no device form: the device solver applies it through a host round trip
// per application, as the reference's GPU solvers do
// (cuda/kernels/GpuLinearSolverCuda.cpp:103-118).
bool deviceForm(const LinearSolverSettings& s, const std::shared_ptr<Preconditioner>& pc,
                bool& jacobi) {
  jacobi = s.preconditioner == PreconditionerType::Jacobi;
  if (pc == nullptr) return true;
  if (dynamic_cast<const cfd::algebra::JacobiPreconditioner*>(pc.get()) != nullptr) {
    jacobi = true;
    return true;
  }
  if (dynamic_cast<const cfd::algebra::IdentityPreconditioner*>(pc.get()) != nullptr) {
    jacobi = false;
    return true;
  }
  return false;
}

std::unique_ptr<LinearSolver> make(LinearSolverSettings settings,
                                   const std::shared_ptr<Preconditioner>& pc, bool cg) {
  bool jacobi = false;
  if (!deviceForm(settings, pc, jacobi)) {
    const cfdhip::LinearSolverSettings d = toDevice(settings, false);
    Preconditioner* raw = pc.get();  // kept alive by the solver's host_ member
    cfdhip::HostPreconditioner apply = [raw](const double* r, double* z, std::size_t n) {
      Vector in(n), out(n);
      for (std::size_t i = 0; i < n; ++i) in[i] = r[i];
      raw->apply(in, out);
      for (std::size_t i = 0; i < n; ++i) z[i] = out[i];
    };
    std::unique_ptr<cfdhip::GpuLinearSolver> device =
        cg ? cfdhip::makeGpuCG(d, std::move(apply)) : cfdhip::makeGpuBiCGSTAB(d, std::move(apply));
    if (device == nullptr) return nullptr;
    return std::make_unique<HipKrylovSolver>(settings, std::move(device), pc);
  }
  const cfdhip::LinearSolverSettings d = toDevice(settings, jacobi);
  std::unique_ptr<cfdhip::GpuLinearSolver> device =
      cg ? cfdhip::makeGpuCG(d) : cfdhip::makeGpuBiCGSTAB(d);
  if (device == nullptr) return nullptr;
  return std::make_unique<HipKrylovSolver>(settings, std::move(device));
}

}  // namespace

std::unique_ptr<cfd::algebra::LinearSolver> makeGpuCG(
    cfd::algebra::LinearSolverSettings settings,
    std::shared_ptr<cfd::algebra::Preconditioner> preconditioner) {
  return make(settings, preconditioner, true);
}

std::unique_ptr<cfd::algebra::LinearSolver> makeGpuBiCGSTAB(
    cfd::algebra::LinearSolverSettings settings,
    std::shared_ptr<cfd::algebra::Preconditioner> preconditioner) {
  return make(settings, preconditioner, false);
}

}  // namespace cfd::gpu
