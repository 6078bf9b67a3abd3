// The reference's own SIMPLE::solve (src/pressure_velocity/SIMPLE.cpp:148-1124)
// driven into libcfdhip through the reference's GPU seams, against the same
// solve on the reference's CPU path. Built from the reference's sources minus
// its three GPU stubs (src/gpu/GPUBackend.cpp, GpuLinearSolver.cpp,
// GpuSimpleDiscretization.cpp), whose interfaces gpu_backend_hip.cpp and
// simple_discretization_hip.cpp implement over libcfdhip.
//
// Lid-driven cavity (unit square, lid u = 1, Re = 100, Rhie-Chow flux):
//   trajectory  a fixed number of SIMPLE iterations with tightly converged
//               linear solves (abs 0, rel 1e-12), run four ways:
//     cpu_bicgstab     CPU discretization, BiCGSTAB / BiCGSTAB on the CPU
//     gpu_resident     device discretization, both solvers BiCGSTAB on the
//                      device: SIMPLE takes its resident outer loop
//                      (SIMPLE.cpp:355-380: assembleMomentumResident,
//                      solveMomentumResident, the all-finite guards, ...)
//     gpu_resident_p   device discretization, momentum BiCGSTAB on the CPU
//                      (assembleMomentum / setMomentumSolution host round
//                      trip), p' BiCGSTAB resident on the device
//     cpu_cg, gpu_roundtrip
//                      p' CG + Jacobi: on the CPU, and through makeGpuCG on
//                      the device discretization's host round trip
//                      (assemblePressureCorrection / setPressureCorrection)
//   steady      the resident loop and the CPU run converged to outer
//               tolerances of 1e-12.
// Host preconditioners: makeGpuCG / makeGpuBiCGSTAB with a Preconditioner
// object the device has no form of (symmetric Gauss-Seidel, a Jacobi the
// factory cannot recognise, one that throws) on a 2D Poisson system, against
// the reference's CG / BiCGSTAB with the same object: status, iterations
// within one, solutions within the tolerance bound, one apply() per
// iteration, and the exception of a failing apply() leaving solve().
// Every device run must agree with its CPU run within 1e-10 relative (max
// over cells / faces, relative to the field's max) on u, v, p and the mass
// flux, with matching status and iteration count. A lifecycle section checks
// the authority bookkeeping and prepare()'s refusals directly.
//
// usage: dropin_simple [n] [iterations]   (defaults 33 and 20); exit 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "cfd/algebra/BiCGSTAB.hpp"
#include "cfd/algebra/CG.hpp"
#include "cfd/algebra/LinearSystem.hpp"
#include "cfd/algebra/Preconditioner.hpp"
#include "cfd/algebra/SparseMatrix.hpp"
#include "cfd/boundary/FixedGradient.hpp"
#include "cfd/core/Exception.hpp"
#include "cfd/boundary/MovingWall.hpp"
#include "cfd/boundary/Symmetry.hpp"
#include "cfd/boundary/Wall.hpp"
#include "cfd/gpu/GPUBackend.hpp"
#include "cfd/gpu/GpuLinearSolver.hpp"
#include "cfd/gpu/GpuSimpleDiscretization.hpp"
#include "cfd/mesh/MeshGeometry.hpp"
#include "cfd/mesh/MeshGrading.hpp"
#include "cfd/physics/FluidProperties.hpp"
#include "cfd/physics/MassFlux.hpp"
#include "cfd/pressure_velocity/SIMPLE.hpp"

using namespace cfd;
using cfd::algebra::LinearSolverBackend;
using cfd::algebra::LinearSolverSettings;
using cfd::algebra::LinearSolverType;
using cfd::algebra::PreconditionerType;
using cfd::fields::ScalarField;
using cfd::fields::SurfaceField;
using cfd::fields::VectorField;
using cfd::gpu::GpuSimpleDiscretization;
using cfd::mesh::Mesh;
using cfd::mesh::MeshGeometry;
using cfd::pressure_velocity::SIMPLE;
using cfd::pressure_velocity::SIMPLEResult;
using cfd::pressure_velocity::SIMPLESettings;

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok) ++failures;
}

struct Cavity {
  Mesh mesh;
  boundary::BoundaryConditionSet vbc, pbc;
};

Cavity cavity(Mesh mesh) {
  Cavity c{std::move(mesh), {}, {}};
  for (const auto& patch : c.mesh.boundaryPatches()) {
    const std::string& name = patch.name();
    if (name == "top")
      c.vbc.set(c.mesh, name, std::make_unique<boundary::MovingWall>(Vector3{1.0, 0.0, 0.0}));
    else
      c.vbc.set(c.mesh, name, std::make_unique<boundary::Wall>());
    c.pbc.set(c.mesh, name, std::make_unique<boundary::FixedGradient>(0.0));
  }
  return c;
}

LinearSolverSettings linearSolver(LinearSolverType type, LinearSolverBackend backend,
                            PreconditionerType pc, double abs, double rel, std::size_t maxIt) {
  LinearSolverSettings s;
  s.type = type;
  s.backend = backend;
  s.preconditioner = pc;
  s.absoluteTolerance = abs;
  s.relativeTolerance = rel;
  s.maxIterations = maxIt;
  return s;
}

SIMPLEResult run(const Cavity& c, SIMPLESettings s) {
  const physics::FluidProperties fluid(1.0, 1.0 / 100.0);
  const std::size_t n = c.mesh.numberOfCells();
  const SIMPLE simple(s, 0);
  return simple.solve(c.mesh, fluid, c.vbc, c.pbc, VectorField(n, Vector3{}), ScalarField(n, 0.0));
}

template <class F, class G>
double maxRel(std::size_t n, F a, G b) {
  double diff = 0.0, scale = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    diff = std::fmax(diff, std::fabs(a(i) - b(i)));
    scale = std::fmax(scale, std::fabs(b(i)));
  }
  return scale > 0.0 ? diff / scale : diff;
}

void compare(const std::string& name, const SIMPLEResult& gpu, const SIMPLEResult& cpu,
             double tol) {
  expect(gpu.status == cpu.status, name + ": status " + std::to_string(int(gpu.status)) +
                                       " (cpu " + std::to_string(int(cpu.status)) + ")");
  expect(gpu.iterations == cpu.iterations,
         name + ": iterations " + std::to_string(gpu.iterations) + " (cpu " +
             std::to_string(cpu.iterations) + ")");
  const std::size_t n = cpu.velocity.size(), nf = cpu.massFlux.size();
  if (gpu.velocity.size() != n || gpu.massFlux.size() != nf || gpu.pressure.size() != n) {
    expect(false, name + ": field sizes");
    return;
  }
  const double eu = maxRel(n, [&](std::size_t i) { return gpu.velocity[i].x; },
                           [&](std::size_t i) { return cpu.velocity[i].x; });
  const double ev = maxRel(n, [&](std::size_t i) { return gpu.velocity[i].y; },
                           [&](std::size_t i) { return cpu.velocity[i].y; });
  const double ep = maxRel(n, [&](std::size_t i) { return gpu.pressure[i]; },
                           [&](std::size_t i) { return cpu.pressure[i]; });
  const double ef = maxRel(nf, [&](std::size_t i) { return gpu.massFlux[i]; },
                           [&](std::size_t i) { return cpu.massFlux[i]; });
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: max rel u %.3e v %.3e p %.3e flux %.3e (bound %.0e)",
                name.c_str(), eu, ev, ep, ef, tol);
  expect(eu <= tol && ev <= tol && ep <= tol && ef <= tol, buf);
  const std::size_t k = std::min(gpu.uResidualHistory.size(), cpu.uResidualHistory.size());
  double eh = 0.0;
  for (std::size_t i = 0; i < k; ++i) {
    eh = std::fmax(eh, std::fabs(gpu.uResidualHistory[i] - cpu.uResidualHistory[i]) /
                           std::fabs(cpu.uResidualHistory[i]));
    eh = std::fmax(eh, std::fabs(gpu.pressureResidualHistory[i] - cpu.pressureResidualHistory[i]) /
                           std::fabs(cpu.pressureResidualHistory[i]));
  }
  std::snprintf(buf, sizeof(buf), "%s: u / p' residual histories within %.3e relative", name.c_str(),
                eh);
  expect(eh <= 1e-6, buf);
}

void report(const std::string& name, const SIMPLEResult& r) {
  std::printf("run %-15s status %d iterations %zu gpuDiscretization %d residentPressureSolve %d "
              "residentSimpleLoop %d momentum its %zu pressure its %zu%s%s\n",
              name.c_str(), int(r.status), r.iterations, int(r.gpuDiscretization),
              int(r.residentPressureSolve), int(r.residentSimpleLoop), r.momentumLinearIterations,
              r.pressureLinearIterations, r.gpuDiscretizationFallbackReason.empty() ? "" : " reason: ",
              r.gpuDiscretizationFallbackReason.c_str());
  std::fflush(stdout);
}

void lifecycle(const Cavity& c) {
  using A = GpuSimpleDiscretization::FieldAuthority;
  using F = GpuSimpleDiscretization::PersistentField;
  GpuSimpleDiscretization d;
  expect(d.active(), "lifecycle: active() with a device");
  std::string reason;
  expect(d.prepare(c.mesh, c.vbc, c.pbc, reason), "lifecycle: prepare on the cavity " + reason);
  expect(d.authority(F::Velocity) == A::HostOnly, "lifecycle: HostOnly before the upload");
  bool threw = false;
  try {
    d.beginIterationResident();
  } catch (const InvalidArgumentError&) {
    threw = true;
  }
  expect(threw, "lifecycle: beginIterationResident before uploadInitialState throws");
  const physics::FluidProperties fluid(1.0, 0.01);
  const std::size_t n = c.mesh.numberOfCells();
  VectorField u(n, Vector3{});
  for (std::size_t i = 0; i < n; ++i) u[i] = Vector3{0.01 * double(i % 7), -0.02, 0.0};
  const ScalarField p(n, 0.5), mu(n, 0.01);
  const SurfaceField F0 = physics::calculateMassFlux(c.mesh, u, fluid, c.vbc);
  d.uploadInitialState(u, p, F0, mu);
  expect(d.authority(F::Velocity) == A::Synchronized && d.authority(F::Pressure) ==
             A::Synchronized && d.authority(F::MassFlux) == A::Synchronized &&
             d.authority(F::Viscosity) == A::Synchronized,
         "lifecycle: Synchronized after uploadInitialState");
  d.beginIterationResident();
  d.assembleMomentumResident(0, 0.7, 0, false);
  d.assembleMomentumResident(1, 0.7, 0, false);
  const auto s = linearSolver(LinearSolverType::BiCGSTAB, LinearSolverBackend::GPU,
                        PreconditionerType::None, 1e-14, 1e-10, 500);
  const auto r0 = d.solveMomentumResident(0, s);
  const auto r1 = d.solveMomentumResident(1, s);
  expect(r0.converged() && r1.converged() && r0.solution.size() == 0,
         "lifecycle: resident momentum solves converge and return no solution");
  expect(d.residentMomentumPredictorAllFinite(false), "lifecycle: predictor finite");
  d.computeResponseCoefficients(false);
  d.computePredictedFaceFlux(true, 1.0, 0.7, 0, false);
  d.assemblePressureCorrectionResident(false, 0, 1.0, false);
  const auto rp = d.solvePressureCorrectionResident(
      linearSolver(LinearSolverType::BiCGSTAB, LinearSolverBackend::GPU, PreconditionerType::None,
             1e-14, 1e-10, 2000));
  expect(rp.converged(), "lifecycle: resident p' solve converges");
  expect(d.residentPressureCorrectionAllFinite(), "lifecycle: p' finite");
  expect(d.residentSolveBytes() > 0, "lifecycle: residentSolveBytes " +
                                         std::to_string(d.residentSolveBytes()));
  d.updatePressure(0.3);
  expect(d.authority(F::Pressure) == A::DeviceOwned, "lifecycle: DeviceOwned after updatePressure");
  // the host-returning corrections leave the persistent fields untouched
  VectorField corrected;
  SurfaceField fluxCorrected;
  d.correctVelocity(0, false, corrected);
  d.correctFaceMassFlux(false, fluxCorrected);
  VectorField before;
  d.downloadVelocity(before);
  double du = 0.0;
  for (std::size_t i = 0; i < n; ++i) du = std::fmax(du, std::fabs(before[i].y - u[i].y));
  expect(du == 0.0, "lifecycle: correctVelocity leaves the resident velocity as it was");
  d.correctVelocityResident(0, false);
  d.correctFaceMassFluxResident(false);
  expect(d.authority(F::Velocity) == A::DeviceOwned && d.authority(F::MassFlux) ==
             A::DeviceOwned,
         "lifecycle: DeviceOwned after the resident corrections");
  VectorField after;
  SurfaceField flux;
  d.downloadVelocity(after);
  d.downloadMassFlux(flux);
  double e = 0.0;
  for (std::size_t i = 0; i < n; ++i)
    e = std::fmax(e, std::fmax(std::fabs(after[i].x - corrected[i].x),
                               std::fabs(after[i].y - corrected[i].y)));
  for (std::size_t f = 0; f < flux.size(); ++f)
    e = std::fmax(e, std::fabs(flux[f] - fluxCorrected[f]));
  expect(e == 0.0, "lifecycle: resident and host-returning corrections agree bit for bit");
  expect(d.authority(F::Velocity) == A::Synchronized && d.authority(F::MassFlux) ==
             A::Synchronized,
         "lifecycle: Synchronized after the downloads");
  expect(d.residentVelocityAllFinite() && d.residentPressureAllFinite(),
         "lifecycle: resident velocity and pressure finite");

  // refusals: a graded (non-uniform) grid and a condition off the structured path
  cfd::mesh::AxisGrading graded;
  graded.type = cfd::mesh::GradingType::Geometric;
  graded.ratio = 2.0;
  const Cavity g = cavity(MeshGeometry::createGraded2D(16, 16, 1.0, 1.0, graded, graded));
  GpuSimpleDiscretization dg;
  reason.clear();
  const bool gradedDeclined = !dg.prepare(g.mesh, g.vbc, g.pbc, reason);
  expect(gradedDeclined && !reason.empty(), "lifecycle: a graded grid is declined (" + reason + ")");
  Cavity sym = cavity(MeshGeometry::createCartesian2D(16, 16, 1.0, 1.0));
  sym.vbc.replace(sym.mesh, "left", std::make_unique<boundary::Symmetry>());
  GpuSimpleDiscretization ds;
  reason.clear();
  const bool symmetryDeclined = !ds.prepare(sym.mesh, sym.vbc, sym.pbc, reason);
  expect(symmetryDeclined && !reason.empty(),
         "lifecycle: a Symmetry condition is declined (" + reason + ")");
}

// z = (D + U)^-1 D (D + L)^-1 r: symmetric Gauss-Seidel, SPD for an SPD A
class SymmetricGaussSeidel final : public cfd::algebra::Preconditioner {
 public:
  void build(const cfd::algebra::SparseMatrix& m) override {
    A_ = std::make_unique<cfd::algebra::SparseMatrix>(m);
  }
  void apply(const cfd::algebra::Vector& r, cfd::algebra::Vector& z) const override {
    ++calls;
    const std::size_t n = A_->rows();
    const double* v = A_->valuesData();
    const std::size_t* col = A_->columnIndicesData();
    const std::size_t* rp = A_->rowOffsetsData();
    std::vector<double> y(n);
    for (std::size_t i = 0; i < n; ++i) {  // (D + L) y = r
      double acc = r[i];
      for (std::size_t k = rp[i]; k < rp[i + 1]; ++k)
        if (col[k] < i) acc -= v[k] * y[col[k]];
      y[i] = acc / A_->diagonal(i);
    }
    z = cfd::algebra::Vector(n);
    for (std::size_t i = n; i-- > 0;) {  // (D + U) z = D y
      double acc = A_->diagonal(i) * y[i];
      for (std::size_t k = rp[i]; k < rp[i + 1]; ++k)
        if (col[k] > i) acc -= v[k] * z[col[k]];
      z[i] = acc / A_->diagonal(i);
    }
  }
  mutable int calls = 0;

 private:
  std::unique_ptr<cfd::algebra::SparseMatrix> A_;
};

// diag(A)^-1 behind a type the factory does not know
class WrappedJacobi final : public cfd::algebra::Preconditioner {
 public:
  void build(const cfd::algebra::SparseMatrix& m) override { inner_.build(m); }
  void apply(const cfd::algebra::Vector& r, cfd::algebra::Vector& z) const override {
    ++calls;
    inner_.apply(r, z);
  }
  mutable int calls = 0;

 private:
  cfd::algebra::JacobiPreconditioner inner_;
};

// throws on its fifth application
class FailingPreconditioner final : public cfd::algebra::Preconditioner {
 public:
  void build(const cfd::algebra::SparseMatrix&) override {}
  void apply(const cfd::algebra::Vector& r, cfd::algebra::Vector& z) const override {
    if (++calls == 5) throw NumericalError("FailingPreconditioner: fifth application");
    z = r;
  }
  mutable int calls = 0;
};

// 5-point Poisson on an m x m grid with Dirichlet sides (SPD), sorted columns
cfd::algebra::LinearSystem poisson(std::size_t m) {
  std::vector<double> val;
  std::vector<std::size_t> col, rp{0};
  const std::size_t n = m * m;
  for (std::size_t j = 0; j < m; ++j)
    for (std::size_t i = 0; i < m; ++i) {
      const std::size_t c = j * m + i;
      auto put = [&](std::size_t k, double a) {
        col.push_back(k);
        val.push_back(a);
      };
      if (j > 0) put(c - m, -1.0);
      if (i > 0) put(c - 1, -1.0);
      put(c, 4.0 + 0.01 * static_cast<double>((c * 7) % 5));
      if (i + 1 < m) put(c + 1, -1.0);
      if (j + 1 < m) put(c + m, -1.0);
      rp.push_back(col.size());
    }
  cfd::algebra::Vector b(n);
  for (std::size_t c = 0; c < n; ++c) b[c] = std::sin(0.37 * static_cast<double>(c)) + 0.25;
  return {cfd::algebra::SparseMatrix(n, n, std::move(val), std::move(col), std::move(rp)),
          std::move(b)};
}

void hostPreconditioners() {
  using cfd::algebra::SolverStatus;
  const cfd::algebra::LinearSystem sys = poisson(48);
  const std::size_t n = sys.size();
  const cfd::algebra::Vector x0(n, 0.0);
  auto set = [](LinearSolverBackend b) {
    return linearSolver(LinearSolverType::CG, b, PreconditionerType::None, 0.0, 1e-10, 5000);
  };
  auto same = [&](const std::string& name, const cfd::algebra::SolverResult& g,
                  const cfd::algebra::SolverResult& c) {
    const double e = maxRel(n, [&](std::size_t i) { return g.solution[i]; },
                            [&](std::size_t i) { return c.solution[i]; });
    char buf[240];
    std::snprintf(buf, sizeof(buf),
                  "%s: status %d (cpu %d), %zu it (cpu %zu), backend GPU, max rel x %.2e", name.c_str(),
                  int(g.status), int(c.status), std::size_t(g.iterations),
                  std::size_t(c.iterations), e);
    const long d = long(g.iterations) - long(c.iterations);
    expect(g.status == SolverStatus::Converged && c.status == SolverStatus::Converged &&
               d >= -1 && d <= 1 && e <= 1e-7 &&
               g.backendUsed == LinearSolverBackend::GPU,
           buf);
  };

  auto sgsC = std::make_shared<SymmetricGaussSeidel>();
  auto sgsG = std::make_shared<SymmetricGaussSeidel>();
  const auto cpuCg = cfd::algebra::CG(set(LinearSolverBackend::CPU), sgsC).solve(sys, x0);
  const auto gpuCgSolver = cfd::gpu::makeGpuCG(set(LinearSolverBackend::GPU), sgsG);
  expect(gpuCgSolver != nullptr, "host pc: makeGpuCG with symmetric Gauss-Seidel is a device solver");
  if (gpuCgSolver == nullptr) return;
  const auto gpuCg = gpuCgSolver->solve(sys, x0);
  same("host pc cg+sgs", gpuCg, cpuCg);
  expect(sgsG->calls >= int(gpuCg.iterations) && sgsG->calls <= int(gpuCg.iterations) + 1,
         "host pc cg+sgs: " + std::to_string(sgsG->calls) + " host applications for " +
             std::to_string(gpuCg.iterations) + " iterations");

  auto wjC = std::make_shared<WrappedJacobi>();
  auto wjG = std::make_shared<WrappedJacobi>();
  const auto cpuJ = cfd::algebra::CG(set(LinearSolverBackend::CPU), wjC).solve(sys, x0);
  const auto gpuJ = cfd::gpu::makeGpuCG(set(LinearSolverBackend::GPU), wjG)->solve(sys, x0);
  same("host pc cg+wrapped jacobi", gpuJ, cpuJ);
  expect(gpuCg.iterations < gpuJ.iterations,
         "host pc: symmetric Gauss-Seidel (" + std::to_string(gpuCg.iterations) +
             " it) beats Jacobi (" + std::to_string(gpuJ.iterations) + " it)");
  LinearSolverSettings jac = set(LinearSolverBackend::GPU);
  jac.preconditioner = PreconditionerType::Jacobi;
  const auto devJ = cfd::gpu::makeGpuCG(jac)->solve(sys, x0);
  same("host pc: wrapped jacobi vs the device's own jacobi", gpuJ, devJ);

  LinearSolverSettings bs = set(LinearSolverBackend::CPU);
  bs.type = LinearSolverType::BiCGSTAB;
  const auto cpuB =
      cfd::algebra::BiCGSTAB(bs, std::make_shared<SymmetricGaussSeidel>()).solve(sys, x0);
  bs.backend = LinearSolverBackend::GPU;
  const auto gpuB =
      cfd::gpu::makeGpuBiCGSTAB(bs, std::make_shared<SymmetricGaussSeidel>())->solve(sys, x0);
  same("host pc bicgstab+sgs", gpuB, cpuB);

  const auto bad = cfd::gpu::makeGpuCG(set(LinearSolverBackend::GPU),
                                       std::make_shared<FailingPreconditioner>());
  std::string what = "no exception";
  try {
    (void)bad->solve(sys, x0);
  } catch (const NumericalError& e) {
    what = e.what();
  } catch (const std::exception& e) {
    what = std::string("other exception: ") + e.what();
  }
  expect(what == "FailingPreconditioner: fifth application",
         "host pc: a throwing apply() leaves solve() with its exception (" + what + ")");
}

}  // namespace

int main(int argc, char** argv) {
  const std::size_t n = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 33;
  const std::size_t iters = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 20;
  if (!cfd::gpu::cudaAvailable()) {
    std::printf("FAIL no usable gfx950 device\n");
    return 2;
  }
  const Cavity c = cavity(MeshGeometry::createCartesian2D(n, n, 1.0, 1.0));
  std::printf("case: host preconditioners, 48 x 48 Poisson\n");
  hostPreconditioners();
  std::printf("case: %zu x %zu cavity, Re 100, Rhie-Chow\n", n, n);
  lifecycle(c);

  using T = LinearSolverType;
  using B = LinearSolverBackend;
  using P = PreconditionerType;
  SIMPLESettings base;
  base.maxIterations = iters;
  base.velocityRelaxation = 0.7;
  base.pressureRelaxation = 0.3;
  base.velocityTolerance = base.pressureTolerance = base.continuityTolerance = 1e-14;
  base.faceFlux = pressure_velocity::FaceFluxScheme::RhieChow;

  // At a relative 1e-12, CG + Jacobi fails a p' solve at outer iteration 13
  // of this cavity on the reference's CPU path (PressureCorrectionFailure),
  // so the CG runs use 1e-10.
  auto settings = [&](B mb, T pt, B pb, P pp, bool disc) {
    SIMPLESettings s = base;
    s.momentumSolver = linearSolver(T::BiCGSTAB, mb, P::None, 0.0, 1e-12, 1000);
    s.pressureSolver = linearSolver(pt, pb, pp, 0.0, pt == T::CG ? 1e-10 : 1e-12, 20000);
    s.enableGpuDiscretization = disc;
    return s;
  };

  const SIMPLEResult cpuB = run(c, settings(B::CPU, T::BiCGSTAB, B::CPU, P::None, false));
  report("cpu_bicgstab", cpuB);
  const SIMPLEResult gpuR = run(c, settings(B::GPU, T::BiCGSTAB, B::GPU, P::None, true));
  report("gpu_resident", gpuR);
  expect(gpuR.gpuDiscretization && gpuR.residentPressureSolve && gpuR.residentSimpleLoop,
         "gpu_resident: SIMPLE took the resident outer loop");
  compare("gpu_resident", gpuR, cpuB, 1e-10);
  const SIMPLEResult gpuP = run(c, settings(B::CPU, T::BiCGSTAB, B::GPU, P::None, true));
  report("gpu_resident_p", gpuP);
  expect(gpuP.gpuDiscretization && gpuP.residentPressureSolve && !gpuP.residentSimpleLoop,
         "gpu_resident_p: device discretization, resident p' solve, host momentum");
  compare("gpu_resident_p", gpuP, cpuB, 1e-10);

  const SIMPLEResult cpuC = run(c, settings(B::CPU, T::CG, B::CPU, P::Jacobi, false));
  report("cpu_cg", cpuC);
  const SIMPLEResult gpuC = run(c, settings(B::CPU, T::CG, B::GPU, P::Jacobi, true));
  report("gpu_roundtrip", gpuC);
  expect(gpuC.gpuDiscretization && !gpuC.residentPressureSolve,
         "gpu_roundtrip: device discretization with host-returned systems, p' by makeGpuCG");
  compare("gpu_roundtrip", gpuC, cpuC, 1e-10);

  // converged steady state
  SIMPLESettings steady = base;
  steady.maxIterations = 5000;
  steady.velocityTolerance = steady.pressureTolerance = steady.continuityTolerance = 1e-12;
  steady.momentumSolver = linearSolver(T::BiCGSTAB, B::CPU, P::None, 1e-16, 1e-8, 500);
  steady.pressureSolver = linearSolver(T::BiCGSTAB, B::CPU, P::None, 1e-16, 1e-8, 2000);
  const SIMPLEResult cpuS = run(c, steady);
  report("cpu_steady", cpuS);
  steady.enableGpuDiscretization = true;
  steady.momentumSolver.backend = steady.pressureSolver.backend = B::GPU;
  const SIMPLEResult gpuS = run(c, steady);
  report("gpu_steady", gpuS);
  expect(cpuS.status == pressure_velocity::SIMPLEStatus::Converged &&
             gpuS.status == pressure_velocity::SIMPLEStatus::Converged && gpuS.residentSimpleLoop,
         "steady: both converged, the device run on the resident loop");
  const std::size_t nc = c.mesh.numberOfCells();
  const double eu = maxRel(nc, [&](std::size_t i) { return gpuS.velocity[i].x; },
                           [&](std::size_t i) { return cpuS.velocity[i].x; });
  const double ev = maxRel(nc, [&](std::size_t i) { return gpuS.velocity[i].y; },
                           [&](std::size_t i) { return cpuS.velocity[i].y; });
  const double ep = maxRel(nc, [&](std::size_t i) { return gpuS.pressure[i]; },
                           [&](std::size_t i) { return cpuS.pressure[i]; });
  char buf[200];
  std::snprintf(buf, sizeof(buf), "steady: max rel u %.3e v %.3e p %.3e (bound 1e-10)", eu, ev, ep);
  expect(eu <= 1e-10 && ev <= 1e-10 && ep <= 1e-10, buf);

  std::printf("%s: %d failure(s)\n", failures == 0 ? "PASS" : "FAIL", failures);
  return failures == 0 ? 0 : 1;
}
