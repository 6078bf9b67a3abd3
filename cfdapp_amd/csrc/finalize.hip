// The finalize kernel: reduces block partials in a fixed order and runs one
// scalar step of a Krylov recurrence on the device (CG.cpp:45-147,
// BiCGSTAB.cpp:88-274), so the host never waits on an inner product.
#include <cfloat>

#include "reduce.hpp"

namespace cfdhip {

namespace {

__device__ void step(KState& s, double* hist, double* out, int stepId, const double* t) {
  switch (stepId) {
    case kStepStartNorm: {  // r0 = b - A x0 (CG.cpp:45-57, BiCGSTAB.cpp:82-93)
      const double norm = sqrt(t[0]);
      s.b0 = norm;
      s.rnorm = norm;
      s.rHatNorm = norm;
      s.restartResidual = norm;
      s.best = norm;
      s.stagnant = 0;
      pushHist(s, hist, norm);
      if (norm == 0.0 || convergedTest(s, norm)) finish(s, CFDHIP_CONVERGED);
      return;
    }
    case kStepCgStartRz: {  // rz = (r, z) (CG.cpp:59-69)
      s.rz = t[0];
      if (!isfinite(s.rz)) finish(s, CFDHIP_BREAKDOWN);
      return;
    }
    case kStepCgPAp: {  // CG.cpp:77-95
      const double pAp = t[0];
      s.pAp = pAp;
      if (!isfinite(pAp) || fabs(pAp) < 1e-30) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double alpha = s.rz / pAp;
      if (!isfinite(alpha)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.alpha = alpha;
      return;
    }
    case kStepCgUpdate: {  // CG.cpp:97-140
      const double rnorm = sqrt(t[0]);
      s.iter += 1;
      pushHist(s, hist, rnorm);
      s.rnorm = rnorm;
      if (!isfinite(rnorm)) {
        finish(s, CFDHIP_NONFINITE_RESIDUAL);
        return;
      }
      if (convergedTest(s, rnorm)) {
        finish(s, CFDHIP_CONVERGED);
        return;
      }
      const double rzNew = t[1];
      if (!isfinite(rzNew)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double beta = rzNew / s.rz;
      if (!isfinite(beta)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.beta = beta;
      s.rz = rzNew;
      if (s.iter >= s.maxIter) finish(s, CFDHIP_MAX_ITERATIONS);
      return;
    }
    case kStepCg1Start: {  // gamma0 = (r0,u0), delta0 = (w0,u0) = (p, Ap) of iteration 1
      const double gamma = t[0], delta = t[1];
      s.rz = gamma;
      if (!isfinite(gamma)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      if (!isfinite(delta) || fabs(delta) < 1e-30) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double a = gamma / delta;
      if (!isfinite(a)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.alpha = a;
      s.beta = 0.0;
      return;
    }
    case kStepCg1Iter:
      cg1IterStep(s, hist, t);
      return;
    case kStepBcgRho:
    case kStepBcgV:
    case kStepBcgS:
    case kStepBcgT:
    case kStepBcgX:
      bcgStep(s, hist, stepId, t);
      return;
    case kStepBcgRestart: {  // BiCGSTAB.cpp:128-144, after the rNorm < restartResidual test
      const double norm = sqrt(t[0]);
      s.rnorm = norm;
      s.restart = 0;
      if (!isfinite(norm)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      if (convergedTest(s, norm)) {
        finish(s, CFDHIP_CONVERGED);
        return;
      }
      s.rHatNorm = norm;
      s.restartResidual = norm;
      s.rhoOld = 1.0;
      s.alpha = 1.0;
      s.omega = 1.0;
      s.restarts += 1;
      s.restart = 2;  // the host zeroes v and p
      return;
    }
    case kStepCgRobustStart: {  // p = u: alpha = (r, u) / (u, A u)
      const double gamma = t[0], delta = t[1];
      s.rz = gamma;
      if (!isfinite(gamma) || !isfinite(delta) || fabs(delta) < 1e-30) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.alpha = gamma / delta;
      s.beta = 0.0;
      if (!isfinite(s.alpha)) finish(s, CFDHIP_BREAKDOWN);
      return;
    }
    case kStepCgRobust:
    case kStepCgReplace: {
      // Explicit conjugation against the last direction p (s = A p by
      // recurrence): beta = -(u, A p) / (p, A p), so the next direction is
      // A-conjugate to p whatever rounding or an FP64 residual replacement
      // did to r, (p', A p') = (u, A u) + beta (u, A p), and the exact line
      // search alpha = (r, p') / (p', A p') with (r, p') = (r, u) + beta (r, p).
      // In exact arithmetic these are CG's scalars (CG.cpp:77-142).
      const double gamma = t[0], delta = t[1], us = t[3], ps = t[4], rp = t[5];
      const double rnorm = sqrt(t[2]);
      s.rnorm = rnorm;
      if (stepId == kStepCgReplace) {  // a replacement completes no iteration
        s.done = 0;
        s.replace = 0;
        s.status = -1;
        s.rmax = rnorm;
      } else {
        s.iter += 1;
        pushHist(s, hist, rnorm);
        if (!isfinite(rnorm)) {
          finish(s, CFDHIP_NONFINITE_RESIDUAL);
          return;
        }
        // the FP32 recurrence only proposes; the host checks in FP64
        s.rmax = fmax(s.rmax, rnorm);
        if (convergedTest(s, rnorm) || rnorm <= s.replaceRatio * s.rmax ||
            s.iter >= s.maxIter) {
          s.replace = 1;
          s.done = 1;
          return;
        }
      }
      if (stepId == kStepCgReplace && s.restartOnReplace) {
        // restart from the replaced residual: p = u, (p, A p) = (u, A u)
        if (!isfinite(gamma) || !isfinite(delta) || fabs(delta) < 1e-30) {
          finish(s, CFDHIP_BREAKDOWN);
          return;
        }
        s.alpha = gamma / delta;
        s.beta = 0.0;
        s.rz = gamma;
        s.dirStart = s.iter;
        if (!isfinite(s.alpha)) finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double beta = -us / ps;
      const double pAp = delta + beta * us;
      if (!isfinite(gamma) || !isfinite(beta) || !isfinite(pAp) || fabs(pAp) < 1e-30) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double alpha = (gamma + beta * rp) / pAp;
      if (!isfinite(alpha)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.beta = beta;
      s.alpha = alpha;
      s.rz = gamma;
      return;
    }
    case kStepNorm:
      out[0] = sqrt(t[0]);
      return;
    case kStepFcgAlpha: {  // flexible CG: alpha = (p, r) / (p, q)
      const double pq = t[0], pr = t[1];
      s.pAp = pq;
      // scale-free: with A and M positive definite (p, A p) > 0 for any p != 0,
      // however small the residual
      if (!isfinite(pq) || !isfinite(pr) || !(pq > 0.0)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double a = pr / pq;
      if (!isfinite(a)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.alpha = a;
      return;
    }
    case kStepFcgUpdate: {  // |r| -> convergence; stagnation window
      const double rnorm = sqrt(t[0]);
      s.iter += 1;
      pushHist(s, hist, rnorm);
      s.rnorm = rnorm;
      if (!isfinite(rnorm)) {
        finish(s, CFDHIP_NONFINITE_RESIDUAL);
        return;
      }
      if (convergedTest(s, rnorm)) {
        finish(s, CFDHIP_CONVERGED);
        return;
      }
      if (rnorm < s.best) {
        s.best = rnorm;
        s.stagnant = 0;
      } else if (++s.stagnant >= s.window && s.window > 0) {
        s.stagnated = 1;
        finish(s, CFDHIP_MAX_ITERATIONS);
        return;
      }
      if (s.iter >= s.maxIter) finish(s, CFDHIP_MAX_ITERATIONS);
      return;
    }
    case kStepFcgBeta: {  // beta = -(z, q_old) / (p_old, q_old)
      const double b = -t[0] / s.pAp;
      if (!isfinite(b)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.beta = b;
      return;
    }
    default:
      return;
  }
}

__global__ __launch_bounds__(kFinThreads) void finalizeKernel(FinalArgs a) {
  KState& s = *a.st;
  const bool bcg = a.step >= kStepBcgRho && a.step <= kStepBcgX;
  const bool skip =
      a.step != kStepNorm && a.step != kStepCgReplace && (s.done || (bcg && s.restart));
  if (skip) return;
  double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (a.step == kStepCgReplace || a.step == kStepCgRobust || a.step == kStepCgRobustStart) {
    sumPartials<6>(a.partial, a.nParts, t);
  } else if (a.step == kStepCgUpdate || a.step == kStepBcgRho || a.step == kStepFcgAlpha) {
    double v[2];
    sumPartials<2>(a.partial, a.nParts, v);
    t[0] = v[0], t[1] = v[1];
  } else if (a.step == kStepCg1Start || a.step == kStepCg1Iter || a.step == kStepBcgV ||
             a.step == kStepBcgT) {
    double v[3];
    sumPartials<3>(a.partial, a.nParts, v);
    t[0] = v[0], t[1] = v[1], t[2] = v[2];
  } else {
    double v[1];
    sumPartials<1>(a.partial, a.nParts, v);
    t[0] = v[0];
  }
  if (threadIdx.x == 0) step(s, a.hist, a.out, a.step, t);
}

}  // namespace

// The scalar steps of many independent solves at once (a batch of cases):
// block b runs case active[b]: it sums that case's W partial streams over its
// slots slot0 .. slot0 + slots - 1 in slot order and steps its state (no
// history). kStepNorm writes out[case].
__global__ __launch_bounds__(256) void caseFinalizeKernel(const double* __restrict__ partial,
                                                          const int64_t* __restrict__ slot0,
                                                          const int64_t* __restrict__ slots,
                                                          const int32_t* __restrict__ active,
                                                          KState* states, int W, int stepId,
                                                          double* out) {
  __shared__ double lds[6][256 / 64];
  const int c = active[blockIdx.x];
  KState& s = states[c];
  if (stepId != kStepNorm && s.done) return;  // uniform: one case per block
  double x[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t b0 = slot0[c], nb = slots[c];
  for (int64_t k = threadIdx.x; k < nb; k += blockDim.x)
    for (int w = 0; w < W; ++w) x[w] += partial[(b0 + k) * W + w];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int w = 0; w < W; ++w) {
    double y = x[w];
    for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o, 64);
    if (lane == 0) lds[w][wid] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < W; ++w)
      for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t[w] += lds[w][i];
    step(s, nullptr, out != nullptr ? out + c : nullptr, stepId, t);
  }
}

void launchFinalize(Context& ctx, const FinalArgs& a) {
  hipLaunchKernelGGL(finalizeKernel, dim3(1), dim3(kFinThreads), 0, ctx.stream, a);
}

void launchCaseFinalize(hipStream_t stream, const double* partial, const int64_t* slot0,
                        const int64_t* slots, const int32_t* active, int64_t nActive,
                        KState* states, int W, int stepId, double* out) {
  require(W >= 1 && W <= 6, "case finalize: 1 to 6 sums");
  if (nActive <= 0) return;
  hipLaunchKernelGGL(caseFinalizeKernel, dim3(static_cast<unsigned>(nActive)), dim3(256), 0,
                     stream, partial, slot0, slots, active, states, W, stepId, out);
}

}  // namespace cfdhip
