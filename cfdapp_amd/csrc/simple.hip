// The resident SIMPLE outer iteration (SIMPLE::solve, src/pressure_velocity/
// SIMPLE.cpp:148-1124, on its GPU-resident path: laminar model, both linear
// solvers on the device, no linear-solver fallback, default absolute
// OuterIterationMonitor gate, src/solver/SolverRobustness.cpp:354-363).
// Fields never leave the device; per iteration the host reads only the
// solver verdicts and the residual scalars.
#include <cmath>

#include "context.hpp"

namespace cfdhip {

namespace {

enum SimpleStatus { kConverged = 0, kMaxIterations = 1, kMomentumFailure = 2,
                    kPressureFailure = 3, kNonFinite = 4 };

bool finite(Context& ctx, std::initializer_list<std::pair<const double*, int64_t>> arrays) {
  std::vector<const double*> p;
  std::vector<int64_t> n;
  for (const auto& a : arrays) {
    p.push_back(a.first);
    n.push_back(a.second);
  }
  return countNonFinite(ctx, p.data(), n.data(), static_cast<int>(p.size())) == 0;
}

}  // namespace

void simpleRun(Context& ctx, const double* s, const cfdhip_solver_settings& mom,
               const cfdhip_solver_settings& pres, double* meta, double* history, int64_t cap) {
  ctx.requireMesh();
  require(ctx.fluidSet, "fluid properties not set");
  const View& v = ctx.view;
  const int D = v.dim;
  const int64_t n = v.n, nf = v.nf;
  const int64_t maxIter = static_cast<int64_t>(s[0]);
  const double au = s[1], ap = s[2], vtol = s[3], ptol = s[4], ctol = s[5];
  const bool rhieChow = s[6] != 0.0;
  const int64_t ref = static_cast<int64_t>(s[7]);
  require(maxIter > 0, "validateSIMPLESettings: maxIterations must be > 0");
  require(std::isfinite(au) && au > 0.0 && au <= 1.0 && std::isfinite(ap) && ap > 0.0 && ap <= 1.0,
          "validateSIMPLESettings: relaxation must be finite and in (0, 1]");
  require(vtol > 0.0 && ptol > 0.0 && ctol > 0.0, "validateSIMPLESettings: tolerances must be > 0");
  require(ref >= 0 && ref < n, "SIMPLE::solve: referenceCell out of range");
  // settings[8]: replay the p' solve's iteration chunks from a captured HIP graph
  cfdhip_solver_settings presRun = pres;
  if (s[8] != 0.0) presRun.graph = 1;

  for (int k = 0; k < 6; ++k) meta[k] = 0.0;
  // the loop's own checks cover its solves' inputs (see Context::inLoop)
  struct InLoop {
    Context& c;
    explicit InLoop(Context& cc) : c(cc) { c.inLoop = true; }
    ~InLoop() { c.inLoop = false; }
  } inLoop(ctx);
  int status = kMaxIterations;
  int64_t it = 0, momIters = 0, presIters = 0;
  double lastCont = 0.0, lastGlobal = 0.0;

  // massFlux = calculateMassFlux(velocity) before the loop (SIMPLE.cpp:229).
  launchMassFlux(ctx, ctx.f(CFDHIP_FIELD_U), ctx.f(CFDHIP_FIELD_V), ctx.f(CFDHIP_FIELD_W),
                 ctx.f(CFDHIP_FIELD_FLUX));
  if (!finite(ctx, {{ctx.f(CFDHIP_FIELD_U), n}, {ctx.f(CFDHIP_FIELD_V), n},
                    {D == 3 ? ctx.f(CFDHIP_FIELD_W) : nullptr, D == 3 ? n : 0},
                    {ctx.f(CFDHIP_FIELD_P), n}, {ctx.f(CFDHIP_FIELD_FLUX), nf}})) {
    meta[0] = kNonFinite;
    return;
  }

  for (; it < maxIter; ++it) {
    try {
      launchMomentum(ctx, au);
    } catch (const Error& e) {
      if (e.code != CFDHIP_ERR_NONFINITE) throw;
      status = kNonFinite;
      break;
    }
    double initial[3] = {0.0, 0.0, 0.0};
    bool failed = false;
    for (int k = 0; k < D; ++k) {
      double* x = ctx.f(CFDHIP_FIELD_USTAR + k);
      CFDHIP_CHECK(hipMemcpyAsync(x, ctx.f(CFDHIP_FIELD_U + k), n * sizeof(double),
                                  hipMemcpyDeviceToDevice, ctx.stream));
      const cfdhip_solve_result r = solveStencil(ctx, mom, ctx.f(CFDHIP_FIELD_MOM_RHS_U + k), x);
      if (r.status != CFDHIP_CONVERGED) {
        failed = true;
        break;
      }
      initial[k] = r.initial_residual;
      momIters += r.iterations;
    }
    if (failed) {
      status = kMomentumFailure;
      break;
    }
    if (!finite(ctx, {{ctx.f(CFDHIP_FIELD_USTAR), n}, {ctx.f(CFDHIP_FIELD_VSTAR), n},
                      {D == 3 ? ctx.f(CFDHIP_FIELD_WSTAR) : nullptr, D == 3 ? n : 0}})) {
      status = kNonFinite;
      break;
    }
    launchResponse(ctx);
    launchPredictedFlux(ctx, rhieChow, au);
    if (!finite(ctx, {{ctx.f(CFDHIP_FIELD_FLUX_STAR), nf}})) {
      status = kNonFinite;
      break;
    }
    try {
      launchPressureAssembly(ctx, ref);
    } catch (const Error& e) {
      if (e.code != CFDHIP_ERR_NONFINITE) throw;
      status = kNonFinite;
      break;
    }
    const cfdhip_solve_result pr =
        solvePressure(ctx, presRun, ctx.f(CFDHIP_FIELD_PC_RHS), ctx.f(CFDHIP_FIELD_PPRIME), true);
    if (pr.status != CFDHIP_CONVERGED) {
      status = kPressureFailure;
      break;
    }
    presIters += pr.iterations;
    if (!finite(ctx, {{ctx.f(CFDHIP_FIELD_PPRIME), n}})) {
      status = kNonFinite;
      break;
    }
    launchUpdatePressure(ctx, ap);
    launchCorrect(ctx);
    if (!finite(ctx, {{ctx.f(CFDHIP_FIELD_U), n}, {ctx.f(CFDHIP_FIELD_V), n},
                      {D == 3 ? ctx.f(CFDHIP_FIELD_W) : nullptr, D == 3 ? n : 0},
                      {ctx.f(CFDHIP_FIELD_P), n}, {ctx.f(CFDHIP_FIELD_FLUX), nf}})) {
      status = kNonFinite;
      break;
    }
    double cont[2];
    launchContinuity(ctx, cont);
    lastCont = cont[0];
    lastGlobal = std::fabs(cont[1]);
    if (history != nullptr && it < cap) {
      history[0 * cap + it] = initial[0];
      history[1 * cap + it] = initial[1];
      history[2 * cap + it] = initial[2];
      history[3 * cap + it] = pr.initial_residual;
      history[4 * cap + it] = lastCont;
    }
    const bool wOk = D != 3 || initial[2] <= vtol;
    if (initial[0] <= vtol && initial[1] <= vtol && pr.initial_residual <= ptol &&
        lastCont <= ctol && lastGlobal <= ctol && wOk) {
      status = kConverged;
      ++it;
      break;
    }
  }
  meta[0] = status;
  meta[1] = static_cast<double>(it);
  meta[2] = static_cast<double>(momIters);
  meta[3] = static_cast<double>(presIters);
  meta[4] = lastCont;
  meta[5] = lastGlobal;
}

}  // namespace cfdhip
