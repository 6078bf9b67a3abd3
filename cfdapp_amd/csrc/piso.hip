// The PISO time step on the resident state (PISO::solveTimeStep,
// include/cfd/pressure_velocity/PISO.hpp:89-90; src/pressure_velocity/PISO.cpp:
// 126-350, detail::solvePisoStep without ALE), laminar, 2D, as the reference
// restricts it (assembleTransientMomentumComponent requires a 2D mesh):
//   transient momentum predictor from (U, P, FLUX) with the implicit-Euler
//   time term, both components solved from the previous velocity;
//   predictor flux F* = calculateMassFlux(u*) (linear, no Rhie-Chow);
//   d = V / a_P of the transient diagonal, kept for both corrections;
//   correction 1: p' system from F*, solve, p += p', u = u* - d grad p',
//     F = F* + D_f (p'_P - p'_N);
//   correction 2: the same from the corrected (u, F);
//   final continuity from the last flux (max cell imbalance, |net flux|);
//   the Courant number is the pre-step one of the previous flux.
// The reference constructs BiCGSTAB for both solves whatever the settings'
// type (PISO.cpp:142-162); here the settings choose (the LinearSolverFactory
// route): BiCGSTAB reproduces the reference, CG / FCG + learned use the
// resident p' solvers. A failed step restores the previous state and reports
// the reference's TransientStepStatus.
#include <cmath>
#include <vector>

#include "context.hpp"

namespace cfdhip {

namespace {

enum PisoStatus { kPisoConverged = 0, kPisoMomentumFailure = 1, kPisoPressureFailure = 2,
                  kPisoNonFinite = 3, kPisoInvalid = 4 };

bool finiteFields(Context& ctx, std::initializer_list<std::pair<const double*, int64_t>> arrays) {
  std::vector<const double*> p;
  std::vector<int64_t> n;
  for (const auto& a : arrays) {
    p.push_back(a.first);
    n.push_back(a.second);
  }
  return countNonFinite(ctx, p.data(), n.data(), static_cast<int>(p.size())) == 0;
}

void copyD(Context& ctx, double* dst, const double* src, int64_t n) {
  CFDHIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx.stream));
}

}  // namespace

// One step; meta = {status, maxCFL, continuity, massImbalance, momentum
// iterations, pressure iterations}. The state fields are U, V, P, FLUX.
void pisoStep(Context& ctx, double dt, int64_t ref, const cfdhip_solver_settings& mom,
              const cfdhip_solver_settings& pres, double* meta,
              DeviceArray<double> (&save)[4]) {
  const View& v = ctx.view;
  const int64_t n = v.n, nf = v.nf;
  for (int k = 0; k < 6; ++k) meta[k] = 0.0;
  meta[0] = kPisoInvalid;
  if (v.dim != 2 || !std::isfinite(dt) || !(dt > 0.0) || ref < 0 || ref >= n) return;
  double* fields[4] = {ctx.f(CFDHIP_FIELD_U), ctx.f(CFDHIP_FIELD_V), ctx.f(CFDHIP_FIELD_P),
                       ctx.f(CFDHIP_FIELD_FLUX)};
  const int64_t counts[4] = {n, n, n, nf};
  if (!finiteFields(ctx, {{fields[0], n}, {fields[1], n}, {fields[2], n}, {fields[3], nf}})) {
    meta[0] = kPisoNonFinite;
    return;
  }
  for (int q = 0; q < 4; ++q) {
    save[q].ensure(counts[q]);
    copyD(ctx, save[q].get(), fields[q], counts[q]);
  }
  int status = kPisoConverged;
  int64_t momIters = 0, presIters = 0;
  double cfl = 0.0, cont[2] = {0.0, 0.0}, maxImb = 0.0;
  do {
    cfl = launchMaxCell(ctx, true, dt);
    try {
      launchMomentum(ctx, 1.0, dt);
    } catch (const Error& e) {
      if (e.code != CFDHIP_ERR_NONFINITE) throw;
      status = kPisoNonFinite;
      break;
    }
    bool failed = false;
    for (int k = 0; k < 2; ++k) {
      double* x = ctx.f(CFDHIP_FIELD_USTAR + k);
      copyD(ctx, x, fields[k], n);
      const cfdhip_solve_result r = solveStencil(ctx, mom, ctx.f(CFDHIP_FIELD_MOM_RHS_U + k), x);
      if (r.status != CFDHIP_CONVERGED) {
        failed = true;
        break;
      }
      momIters += r.iterations;
    }
    if (failed) {
      status = kPisoMomentumFailure;
      break;
    }
    if (!finiteFields(ctx, {{ctx.f(CFDHIP_FIELD_USTAR), n}, {ctx.f(CFDHIP_FIELD_VSTAR), n}})) {
      status = kPisoNonFinite;
      break;
    }
    launchPredictedFlux(ctx, false, 1.0);
    if (!finiteFields(ctx, {{ctx.f(CFDHIP_FIELD_FLUX_STAR), nf}})) {
      status = kPisoNonFinite;
      break;
    }
    launchResponse(ctx);
    for (int corr = 0; corr < 2 && status == kPisoConverged; ++corr) {
      if (corr == 1) {  // the second correction starts from the first's (u, F)
        copyD(ctx, ctx.f(CFDHIP_FIELD_USTAR), fields[0], n);
        copyD(ctx, ctx.f(CFDHIP_FIELD_VSTAR), fields[1], n);
        copyD(ctx, ctx.f(CFDHIP_FIELD_FLUX_STAR), fields[3], nf);
      }
      try {
        launchPressureAssembly(ctx, ref);
      } catch (const Error& e) {
        if (e.code != CFDHIP_ERR_NONFINITE) throw;
        status = kPisoNonFinite;
        break;
      }
      const cfdhip_solve_result pr =
          solvePressure(ctx, pres, ctx.f(CFDHIP_FIELD_PC_RHS), ctx.f(CFDHIP_FIELD_PPRIME), true);
      if (pr.status != CFDHIP_CONVERGED) {
        status = kPisoPressureFailure;
        break;
      }
      presIters += pr.iterations;
      if (!finiteFields(ctx, {{ctx.f(CFDHIP_FIELD_PPRIME), n}})) {
        status = kPisoNonFinite;
        break;
      }
      launchUpdatePressure(ctx, 1.0);  // p + 1 * p' == p + p' exactly
      launchCorrect(ctx);
      if (!finiteFields(ctx, {{fields[0], n}, {fields[1], n}, {fields[2], n}, {fields[3], nf}})) {
        status = kPisoNonFinite;
        break;
      }
    }
    if (status != kPisoConverged) break;
    launchContinuity(ctx, cont);
    maxImb = launchMaxCell(ctx, false, 0.0);
  } while (false);
  if (status != kPisoConverged) {
    for (int q = 0; q < 4; ++q) copyD(ctx, fields[q], save[q].get(), counts[q]);
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    meta[0] = status;
    meta[4] = static_cast<double>(momIters);
    meta[5] = static_cast<double>(presIters);
    return;
  }
  meta[0] = kPisoConverged;
  meta[1] = cfl;
  meta[2] = maxImb;
  meta[3] = std::fabs(cont[1]);
  meta[4] = static_cast<double>(momIters);
  meta[5] = static_cast<double>(presIters);
}

int64_t pisoRun(Context& ctx, double dt, int64_t steps, int64_t ref,
                const cfdhip_solver_settings& mom, const cfdhip_solver_settings& pres,
                double* meta) {
  ctx.requireMesh();
  require(ctx.fluidSet, "fluid properties not set");
  require(steps >= 1, "PISO: steps must be >= 1");
  require(ctx.view.dim == 2, "PISO: the transient momentum predictor is 2D only "
                             "(assembleTransientMomentumComponent)", CFDHIP_ERR_UNSUPPORTED);
  DeviceArray<double> save[4];
  int64_t k = 0;
  for (; k < steps; ++k) {
    pisoStep(ctx, dt, ref, mom, pres, meta + 6 * k, save);
    if (meta[6 * k] != static_cast<double>(kPisoConverged)) {
      ++k;
      break;
    }
  }
  return k;
}

}  // namespace cfdhip
