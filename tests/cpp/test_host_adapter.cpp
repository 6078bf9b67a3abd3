// C++20 host adapter checks (run by tests/test_host_adapter.py).
//   ./test_host_adapter nodevice  -- factory contract without a device
//   ./test_host_adapter device    -- solves and SIMPLE stages on the GPU
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cfdhip.hpp"

namespace {

int failures = 0;
#define EXPECT(cond)                                                      \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                         \
    }                                                                     \
  } while (0)

// The hand-solved 2x2 SPD system of the reference's test_cg.cpp shape:
// [4 1; 1 3] x = [1; 2]  ->  x = [1/11, 7/11].
void smallSystem(std::vector<std::size_t>& rp, std::vector<std::size_t>& col,
                 std::vector<double>& val) {
  rp = {0, 2, 4};
  col = {0, 1, 0, 1};
  val = {4.0, 1.0, 1.0, 3.0};
}

void noDevice() {
  cfdhip::LinearSolverSettings s;
  bool threw = false;
  try {
    s.maxIterations = 0;
    (void)cfdhip::makeGpuCG(s);
  } catch (const cfdhip::InvalidArgumentError&) {
    threw = true;
  }
  EXPECT(threw);
  s.maxIterations = 10;
  if (!cfdhip::deviceAvailable()) {
    EXPECT(cfdhip::makeGpuCG(s) == nullptr);
    EXPECT(cfdhip::makeGpuBiCGSTAB(s) == nullptr);
    cfdhip::GpuSimpleDiscretization d;
    std::string reason;
    const std::vector<cfdhip::PatchCondition> bc(4);
    EXPECT(!d.prepare({2, 8, 8, 1}, bc, 1.0, 0.01, reason));
    EXPECT(!reason.empty());
  }
}

void device() {
  EXPECT(cfdhip::deviceAvailable());
  std::vector<std::size_t> rp, col;
  std::vector<double> val;
  smallSystem(rp, col, val);
  const cfdhip::CsrView A{2, rp.data(), col.data(), val.data()};
  const std::vector<double> b = {1.0, 2.0};
  cfdhip::LinearSolverSettings s;
  s.absoluteTolerance = 1e-14;
  s.relativeTolerance = 1e-12;
  s.maxIterations = 50;
  s.preconditioner = cfdhip::PreconditionerType::Jacobi;
  std::vector<std::unique_ptr<cfdhip::GpuLinearSolver>> solvers;
  solvers.push_back(cfdhip::makeGpuCG(s));
  solvers.push_back(cfdhip::makeGpuBiCGSTAB(s));
  for (const auto& solver : solvers) {
    EXPECT(solver != nullptr);
    if (!solver) continue;
    const auto r = solver->solve(A, b);
    EXPECT(r.converged());
    EXPECT(r.solution.size() == 2);
    EXPECT(std::fabs(r.solution[0] - 1.0 / 11.0) < 1e-12);
    EXPECT(std::fabs(r.solution[1] - 7.0 / 11.0) < 1e-12);
    EXPECT(r.residualHistory.size() == r.iterations + 1);
    EXPECT(r.backendUsed == cfdhip::LinearSolverBackend::GPU);
  }
  // MaxIterations is a status, not an exception (test_cg.cpp shape)
  s.maxIterations = 1;
  s.relativeTolerance = 1e-15;
  s.absoluteTolerance = 0.0;
  const auto cg = cfdhip::makeGpuCG(s);
  std::vector<std::size_t> rp3 = {0, 2, 5, 7}, col3 = {0, 1, 0, 1, 2, 1, 2};
  std::vector<double> val3 = {4.0, -1.0, -1.0, 4.0, -1.0, -1.0, 4.0};
  const cfdhip::CsrView A3{3, rp3.data(), col3.data(), val3.data()};
  const auto r3 = cg->solve(A3, std::vector<double>{1.0, 2.0, 3.0});
  EXPECT(r3.status == cfdhip::SolverStatus::MaxIterations);
  EXPECT(r3.iterations == 1);

  // A 16x16 cavity through the resident stages.
  cfdhip::GpuSimpleDiscretization d;
  std::vector<cfdhip::PatchCondition> bc(4);
  bc[3].velocityKind = CFDHIP_MOVING_WALL;
  bc[3].velocity[0] = 1.0;
  std::string reason;
  EXPECT(d.prepare({2, 16, 16, 1}, bc, 1.0, 0.01, reason));
  if (!d.active()) return;
  const std::size_t n = d.cells();
  const std::vector<double> zero(n, 0.0);
  d.uploadInitialState(zero, zero, zero, zero);
  cfdhip::LinearSolverSettings mom;
  mom.absoluteTolerance = 1e-10;
  mom.relativeTolerance = 1e-8;
  mom.maxIterations = 500;
  cfdhip::LinearSolverSettings pres;
  pres.type = cfdhip::LinearSolverType::CG;
  pres.preconditioner = cfdhip::PreconditionerType::Jacobi;
  pres.absoluteTolerance = 0.0;
  pres.relativeTolerance = 1e-8;
  pres.maxIterations = 2000;
  double lastCont = 1e300;
  for (int it = 0; it < 5; ++it) {
    d.assembleMomentumResident(0.7);
    EXPECT(d.solveMomentumResident(0, mom).converged());
    EXPECT(d.solveMomentumResident(1, mom).converged());
    d.computeResponseCoefficients();
    d.computePredictedFaceFlux(true, 0.7);
    d.assemblePressureCorrectionResident(0);
    const auto pr = d.solvePressureCorrectionResident(pres);
    EXPECT(pr.converged());
    d.updatePressure(0.3);
    d.correctVelocityAndFluxResident();
    lastCont = d.continuity().first;
  }
  EXPECT(std::isfinite(lastCont) && lastCont < 1e-6);
  std::vector<double> u(n);
  d.download(CFDHIP_FIELD_U, u);
  double umax = 0.0;
  for (double x : u) umax = std::fmax(umax, std::fabs(x));
  EXPECT(umax > 0.01 && umax < 1.0);

  // PISO time steps from this state (the reference's BiCGSTAB for both systems)
  cfdhip::LinearSolverSettings bicg;
  bicg.absoluteTolerance = 1e-10;
  bicg.relativeTolerance = 1e-8;
  bicg.maxIterations = 2000;
  for (int s = 0; s < 3; ++s) {
    const auto t = d.solveTimeStepResident(0.01, mom, bicg);
    EXPECT(t.status == cfdhip::TransientStepStatus::Converged);
    EXPECT(t.maxCFL > 0.0 && t.continuityResidual < 1e-6 && t.pressureIterations > 0);
  }
  const auto bad = d.solveTimeStepResident(-1.0, mom, bicg);
  EXPECT(bad.status == cfdhip::TransientStepStatus::InvalidConfiguration);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "device") == 0) device();
  else noDevice();
  if (failures == 0) std::printf("ok\n");
  return failures == 0 ? 0 : 1;
}
