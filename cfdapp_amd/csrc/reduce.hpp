// Deterministic two-stage reductions without in-kernel fences.
//
// Stage 1 (inside a worker kernel): each block reduces its values with wave
// shuffles and LDS and writes one partial per sum. Stage 2 (finalizeKernel,
// one block, launched right after on the same stream): sums the partials in a
// fixed order and runs the solver's scalar step on the device. The kernel
// boundary makes the partials visible; no device-scope fence is needed
// inside the worker, which on gfx950 would write back the L2 of the XCD once
// per block while the kernel is streaming its stores.
#pragma once

#include <cfloat>

#include "context.hpp"

namespace cfdhip {

constexpr int kRedThreads = 256;

// XCD-aware tile order: workgroups are dispatched to the 8 XCDs round-robin
// by workgroup id, so workgroup b runs on XCD b % 8. Giving each XCD one
// contiguous range of tiles (row-major) keeps the halo a tile shares with its
// left/right and upper/lower neighbours in the same L2 instead of fetching it
// once per XCD. A bijection of [0, T) for any tile count T.
constexpr int kXcds = 8;
__device__ inline int64_t xcdTileOrder(int64_t b, int64_t T) {
  const int64_t q = T / kXcds, r = T % kXcds;
  const int64_t x = b % kXcds, k = b / kXcds;
  return x * q + (x < r ? x : r) + k;
}

template <int W>
__device__ inline void blockPartials(double (&val)[W], double* partial) {
  __shared__ double lds[W][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < W; ++w) {
    double x = val[w];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) lds[w][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < W) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += lds[threadIdx.x][i];
    partial[static_cast<int64_t>(blockIdx.x) * W + threadIdx.x] = t;
  }
}

// The finalize kernel is one block: it is latency-bound, so it uses the
// widest block and issues four independent loads per stream before adding.
constexpr int kFinThreads = 1024;

// Sums W partial streams of nParts blocks in a fixed order (one block of
// kFinThreads).
template <int W>
__device__ inline void sumPartials(const double* __restrict__ partial, int64_t nParts,
                                   double (&tot)[W]) {
  __shared__ double lds[W][kFinThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double x[W];
  for (int w = 0; w < W; ++w) x[w] = 0.0;
  int64_t b = threadIdx.x;
  for (; b + 3 * kFinThreads < nParts; b += 4 * kFinThreads) {
    double v[4][W];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int w = 0; w < W; ++w) v[k][w] = partial[(b + k * kFinThreads) * W + w];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] += v[k][w];
  }
  for (; b < nParts; b += kFinThreads)
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] += partial[b * W + w];
  for (int w = 0; w < W; ++w) {
    double y = x[w];
    for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o, 64);
    if (lane == 0) lds[w][wid] = y;
  }
  __syncthreads();
  for (int w = 0; w < W; ++w) {
    double t = 0.0;
    for (int i = 0; i < kFinThreads / 64; ++i) t += lds[w][i];
    tot[w] = t;
  }
}

// Scalar steps run by the finalize kernel (thread 0) on the reduced sums.
enum FinalStep : int {
  kStepStartNorm = 0,   // b0 = |r0|, converged at start?          (1 sum)
  kStepCgStartRz,       // rz = (r, z)                             (1 sum)
  kStepCgPAp,           // pAp -> alpha                            (1 sum)
  kStepCgUpdate,        // |r|, (r, z) -> convergence, beta        (2 sums)
  kStepCg1Start,        // gamma, delta -> alpha (single-kernel CG) (3 sums)
  kStepCg1Iter,         // gamma, delta, |r| -> alpha, beta        (3 sums)
  kStepBcgRho,          // rho, sum|.| -> beta                     (2 sums)
  kStepBcgV,            // (rHat, v), sum|.|, |v|^2 -> alpha       (3 sums)
  kStepBcgS,            // |s| -> early convergence                (1 sum)
  kStepBcgT,            // (t,t), (t,s), sum|.| -> omega           (3 sums)
  kStepBcgX,            // |r| -> convergence                      (1 sum)
  kStepBcgRestart,      // |r| after the true-residual restart     (1 sum)
  kStepNorm,            // sqrt of the sum into out                (1 sum)
  kStepFcgAlpha,        // (p, q), (p, r) -> alpha                 (2 sums)
  kStepFcgUpdate,       // |r| -> convergence, stagnation          (1 sum)
  kStepFcgBeta,         // (z, q) -> beta                          (1 sum)
  // the mixed-precision iteration with explicit conjugation (6 sums each:
  // (r,u), (w,u), |r|^2, (u,s), (p,s), (r,p) with p, s = A p the last direction)
  kStepCgRobustStart,   // first direction p = u -> alpha
  kStepCgRobust,        // |r| -> convergence / replacement; beta, alpha
  kStepCgReplace        // the same scalars after an FP64 residual replacement
};

__device__ inline bool convergedTest(const KState& s, double a) {
  return a <= s.absTol || ((s.b0 > 0.0) && (a / s.b0 <= s.relTol));
}
__device__ inline void finish(KState& s, int status) {
  s.status = status;
  s.done = 1;
}
// Appends to the residual history (hist null: counts the entry only).
__device__ inline void pushHist(KState& s, double* hist, double value) {
  if (hist != nullptr) hist[s.histLen] = value;
  s.histLen += 1;
}

// kStepCg1Iter (CG.cpp:97-140 in the single-reduction form): t = (r', u'),
// (w', u'), |r'|^2 of the iteration just done -> the recurrence residual, the
// convergence test, and beta, alpha of the next iteration. Run by the
// finalize kernel and, redundantly in every block, by the fused small-grid
// iteration kernel.
__device__ inline void cg1IterStep(KState& s, double* hist, const double* t) {
  const double gamma = t[0], delta = t[1];
  const double rnorm = sqrt(t[2]);
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
  if (!isfinite(gamma)) {
    finish(s, CFDHIP_BREAKDOWN);
    return;
  }
  const double beta = gamma / s.rz;
  const double pAp = delta - beta * gamma / s.alpha;  // (p', A p') of the next step
  if (!isfinite(beta) || !isfinite(pAp) || fabs(pAp) < 1e-30) {
    finish(s, CFDHIP_BREAKDOWN);
    return;
  }
  const double alpha = gamma / pAp;
  if (!isfinite(alpha)) {
    finish(s, CFDHIP_BREAKDOWN);
    return;
  }
  s.beta = beta;
  s.alpha = alpha;
  s.rz = gamma;
  if (s.iter >= s.maxIter) finish(s, CFDHIP_MAX_ITERATIONS);
}

// Mixed-precision Jacobi-PCG: whether an FP64 residual replacement restarts
// the search direction (1) or continues it with explicit conjugation (0).
// Measured on the d = 1 p' operators (2048^2, 256^3): continuing the
// direction across a replacement is stable at 512^2 but not at 2048^2 or
// 256^3, where the residual jumps after the replacement; restarting every
// 1e-5 of FP32 reduction is the fastest setting in 2D (0.82 x the FP64 time).
constexpr int kRestartOnReplace = 1;

struct FinalArgs {
  const double* partial;
  int64_t nParts;
  KState* st;
  double* hist;
  double* out;  // kStepNorm target
  int step;
};

void launchFinalize(Context& ctx, const FinalArgs& a);
// The finalize of many independent solves (one block per active case, its
// partials in slots slot0[case] .. slot0[case] + slots[case] - 1).
void launchCaseFinalize(hipStream_t stream, const double* partial, const int64_t* slot0,
                        const int64_t* slots, const int32_t* active, int64_t nActive,
                        KState* states, int W, int stepId, double* out);


// cancelledToRoundingLevel (BiCGSTAB.cpp:34-41) with sum_i |x_i y_i| at hand.
__device__ inline bool cancelled(double ip, double absSum, double nx, double ny) {
  if (fabs(ip) > DBL_EPSILON * nx * ny) return false;
  return fabs(ip) <= DBL_EPSILON * absSum;
}

// The scalar steps of BiCGSTAB (BiCGSTAB.cpp:149-268) on the sums t of the
// worker that precedes each: run by the finalize kernel and, redundantly in
// every block, by the fused 2D stencil BiCGSTAB kernels.
__device__ inline void bcgStep(KState& s, double* hist, int stepId, const double* t) {
  switch (stepId) {
    case kStepBcgRho: {  // BiCGSTAB.cpp:149-163
      const double rho = t[0];
      if (!isfinite(rho)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      if (cancelled(rho, t[1], s.rHatNorm, s.rnorm)) {
        s.restart = 1;
        return;
      }
      const double beta = (rho / s.rhoOld) * (s.alpha / s.omega);
      if (!isfinite(beta)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.rho = rho;
      s.beta = beta;
      return;
    }
    case kStepBcgV: {  // BiCGSTAB.cpp:171-185
      const double rHatDotV = t[0];
      if (!isfinite(rHatDotV)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      if (cancelled(rHatDotV, t[1], s.rHatNorm, sqrt(t[2]))) {
        s.restart = 1;
        return;
      }
      const double alpha = s.rho / rHatDotV;
      if (!isfinite(alpha)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.alpha = alpha;
      return;
    }
    case kStepBcgS: {  // BiCGSTAB.cpp:187-207
      const double sNorm = sqrt(t[0]);
      s.snorm = sNorm;
      const int64_t iter = s.completed + 1;
      if (!isfinite(sNorm)) {
        s.iter = iter;
        pushHist(s, hist, sNorm);
        s.rnorm = sNorm;
        finish(s, CFDHIP_NONFINITE_RESIDUAL);
        return;
      }
      if (convergedTest(s, sNorm)) {
        s.iter = iter;
        pushHist(s, hist, sNorm);
        s.rnorm = sNorm;
        s.sConverged = 1;  // x += pHat alpha in the next worker
      }
      return;
    }
    case kStepBcgT: {  // BiCGSTAB.cpp:209-238
      if (s.sConverged) {
        finish(s, CFDHIP_CONVERGED);
        return;
      }
      const double tDotT = t[0];
      if (!isfinite(tDotT) || tDotT < DBL_MIN) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      const double tDotS = t[1];
      if (!isfinite(tDotS)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      if (cancelled(tDotS, t[2], sqrt(tDotT), s.snorm)) {
        s.restart = 1;
        return;
      }
      const double omega = tDotS / tDotT;
      if (!isfinite(omega)) {
        finish(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.omega = omega;
      return;
    }
    case kStepBcgX: {  // BiCGSTAB.cpp:240-268
      const double rn = sqrt(t[0]);
      const int64_t iter = s.completed + 1;
      s.rnorm = rn;
      s.iter = iter;
      pushHist(s, hist, rn);
      if (!isfinite(rn)) {
        finish(s, CFDHIP_NONFINITE_RESIDUAL);
        return;
      }
      s.completed = iter;
      if (convergedTest(s, rn)) {
        finish(s, CFDHIP_CONVERGED);
        return;
      }
      s.rhoOld = s.rho;
      if (s.completed >= s.maxIter) finish(s, CFDHIP_MAX_ITERATIONS);
      return;
    }
    default:
      return;
  }
}

}  // namespace cfdhip
