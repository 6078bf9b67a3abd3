// Flexible CG with the learned preconditioner on the 2D p' operator.
//
// Inner solve (vectors in FP64, or FP32 with precision CFDHIP_MIXED; FP64
// reductions), truncated flexible CG with one stored direction (Notay's
// FCG(1)):
//   z = M r;  beta = -(z, q_old) / (p_old, q_old);  p = z + beta p_old
//   q = A p;  alpha = (p, r) / (p, q);  d += alpha p;  r -= alpha q
// with M r = D^-1/2 N(D^-1/2 r), N the network (FP32 activations, BF16
// weights), and z = r / a_P at the pinned reference cell. Outer loop (FP64):
// r = b - A x, inner solve for the correction, x += d, until the FP64 true
// residual meets the tolerance relative to |b - A x0|. FP32 inner vectors lose
// A-orthogonality fast enough to cost more iterations than their bandwidth
// saves (measured: 1024^2 p' system, 1035 FP32 iterations with restarts every
// 1e-4 versus 3497 without), so they restart every 1e-4 of reduction; FP64
// vectors run the inner solve to the final tolerance in one cycle. An inner solve whose residual has not reached a new
// minimum for `stagnation_window` iterations hands the solve to FP64
// Jacobi-PCG from the current x (the fallback, reported in fallback_used).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <functional>
#include <vector>

#include "reduce.hpp"
#include "unet_net.hpp"

namespace cfdhip {

namespace {

constexpr int kMaxParts = 2048;

template <class T>
struct OpT {
  int64_t nx, ny, pin, offY;
  const T* diag;
  const T* coeff;
};

inline dim3 stride(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > kMaxParts) b = kMaxParts;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : b));
}

template <class T>
__device__ inline T pAt(const T* __restrict__ z, const T* __restrict__ pOld, T beta, bool first,
                        int64_t c) {
  return first ? z[c] : z[c] + beta * pOld[c];
}

// p = z + beta pOld; q = A p; partials (p, q), (p, r).
template <class T>
__global__ __launch_bounds__(256) void fcgDirSpmvKernel(OpT<T> op, const T* __restrict__ z,
                                                        const T* __restrict__ pOld,
                                                        T* __restrict__ pNew, T* __restrict__ q,
                                                        const T* __restrict__ r,
                                                        const KState* __restrict__ st, int first,
                                                        double* partial) {
  double acc[2] = {0.0, 0.0};
  if (!st->done) {
    const T beta = static_cast<T>(st->beta);
    const bool f = first != 0;
    const int64_t n = op.nx * op.ny;
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const int64_t i = c % op.nx, j = c / op.nx;
      const T pc = pAt(z, pOld, beta, f, c);
      T y;
      if (c == op.pin) {
        y = T(0) + op.diag[c] * pc;
      } else {
        y = T(0);
        if (j > 0) y += (-op.coeff[op.offY + j * op.nx + i]) * pAt(z, pOld, beta, f, c - op.nx);
        if (i > 0) y += (-op.coeff[j * (op.nx + 1) + i]) * pAt(z, pOld, beta, f, c - 1);
        y += op.diag[c] * pc;
        if (i < op.nx - 1) y += (-op.coeff[j * (op.nx + 1) + i + 1]) * pAt(z, pOld, beta, f, c + 1);
        if (j < op.ny - 1)
          y += (-op.coeff[op.offY + (j + 1) * op.nx + i]) * pAt(z, pOld, beta, f, c + op.nx);
      }
      pNew[c] = pc;
      q[c] = y;
      acc[0] += static_cast<double>(pc) * y;
      acc[1] += static_cast<double>(pc) * r[c];
    }
  }
  blockPartials<2>(acc, partial);
}

// d += alpha p; r -= alpha q; partial |r|^2.
template <class T>
__global__ __launch_bounds__(256) void fcgUpdateKernel(int64_t n, const T* __restrict__ p,
                                                       const T* __restrict__ q, T* __restrict__ d,
                                                       T* __restrict__ r,
                                                       const KState* __restrict__ st,
                                                       double* partial) {
  double acc[1] = {0.0};
  if (!st->done) {
    const T alpha = static_cast<T>(st->alpha);
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      d[c] += alpha * p[c];
      const T rc = r[c] - alpha * q[c];
      r[c] = rc;
      acc[0] += static_cast<double>(rc) * rc;
    }
  }
  blockPartials<1>(acc, partial);
}

template <class T>
__global__ void normKernelT(int64_t n, const T* __restrict__ r, double* partial) {
  double acc[1] = {0.0};
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    acc[0] += static_cast<double>(r[c]) * r[c];
  blockPartials<1>(acc, partial);
}

// D^-1/2 in FP32 (diagonal entries are positive on a valid p' system).
__global__ void invSqrtKernel(int64_t n, const double* __restrict__ diag, float* __restrict__ dis) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dis[c] = static_cast<float>(1.0 / std::sqrt(diag[c]));
}

// Scaling of a Neumann-padded network, z = S N(S r). S^-2 is a smoothed
// interior-equivalent diagonal: per cell the sum of its internal-face
// couplings times 4 / their count (the diagonal the cell would have with all
// four neighbours, so that a constant field scales like the interior),
// geometric-mean filtered over a (2 kScaleRadius + 1)^2 window (indices
// clamped to the grid). A cell-by-cell scaling would map the operator's
// smooth near-null modes to rough ones wherever the coefficients are rough
// (a one-cell wall layer, noisy response coefficients), which the network's
// fixed stencils cannot follow; the filtered scaling keeps them smooth. The
// pinned cell keeps D^-1/2 (z = r / a_P there). 2D face order: x-faces
// j (nx + 1) + i, y-faces offY + j nx + i.
constexpr int kScaleRadius = 4;

__global__ void logInteriorDiagKernel(int64_t nx, int64_t ny, int64_t offY,
                                      const double* __restrict__ diag,
                                      const double* __restrict__ coeff, float* __restrict__ out) {
  const int64_t n = nx * ny;
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % nx, j = c / nx;
    double s = 0.0;
    int m = 0;
    if (i > 0) { s += coeff[j * (nx + 1) + i]; ++m; }
    if (i < nx - 1) { s += coeff[j * (nx + 1) + i + 1]; ++m; }
    if (j > 0) { s += coeff[offY + j * nx + i]; ++m; }
    if (j < ny - 1) { s += coeff[offY + (j + 1) * nx + i]; ++m; }
    const double dt = (m == 0 || !(s > 0.0)) ? diag[c] : s * 4.0 / m;
    out[c] = static_cast<float>(std::log(dt));
  }
}

// Window mean along x (dir 0) or y (dir 1), indices clamped to the grid; the
// y pass also forms S = exp(-mean / 2) (final != 0), D^-1/2 at the pin.
__global__ void boxMeanKernel(int64_t nx, int64_t ny, int dir, const float* __restrict__ in,
                              float* __restrict__ out, int final, int64_t pin,
                              const double* __restrict__ diag) {
  const int64_t n = nx * ny;
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % nx, j = c / nx;
    float s = 0.f;
    for (int k = -kScaleRadius; k <= kScaleRadius; ++k) {
      int64_t ii = i, jj = j;
      if (dir == 0) ii = i + k < 0 ? 0 : (i + k >= nx ? nx - 1 : i + k);
      else jj = j + k < 0 ? 0 : (j + k >= ny ? ny - 1 : j + k);
      s += in[jj * nx + ii];
    }
    const float mean = s / static_cast<float>(2 * kScaleRadius + 1);
    if (!final) out[c] = mean;
    else out[c] = c == pin ? static_cast<float>(1.0 / std::sqrt(diag[c])) : std::exp(-0.5f * mean);
  }
}

__global__ void zeroKernelD64(int64_t n, double* a) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[c] = 0.0;
}

template <class T>
__global__ void zeroT(int64_t n, T* a) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[c] = T(0);
}

template <class T>
__global__ void addCorrection(int64_t n, const T* __restrict__ d, double* __restrict__ x) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[c] += static_cast<double>(d[c]);
}

// r64 = b - A x (FP64) -> rOut (FP32 or FP64), |r64|^2 partials.
template <class T>
__global__ __launch_bounds__(256) void outerResidualKernel(int64_t nx, int64_t ny, int64_t pin,
                                                           int64_t offY,
                                                           const double* __restrict__ diag,
                                                           const double* __restrict__ coeff,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ x,
                                                           T* __restrict__ r32,
                                                           double* partial) {
  const int64_t n = nx * ny;
  double acc[1] = {0.0};
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % nx, j = c / nx;
    double y;
    if (c == pin) {
      y = 0.0 + diag[c] * x[c];
    } else {
      y = 0.0;
      if (j > 0) y += (-coeff[offY + j * nx + i]) * x[c - nx];
      if (i > 0) y += (-coeff[j * (nx + 1) + i]) * x[c - 1];
      y += diag[c] * x[c];
      if (i < nx - 1) y += (-coeff[j * (nx + 1) + i + 1]) * x[c + 1];
      if (j < ny - 1) y += (-coeff[offY + (j + 1) * nx + i]) * x[c + nx];
    }
    const double rc = b[c] - y;
    r32[c] = static_cast<T>(rc);
    acc[0] += rc * rc;
  }
  blockPartials<1>(acc, partial);
}

struct InnerResult {
  int64_t iterations = 0;
  int status = CFDHIP_MAX_ITERATIONS;
  bool stagnated = false;
  int64_t replacements = 0;  // reliable updates made
  bool replacedOut = false;  // ended by a reliable update (the caller's FP64 test decides)
};

// The preconditioner: z = M r on the solver grid (ny x nx), network grid h x w.
template <class T>
void applyM(Context& ctx, const T* r, const T* q, T* z, double* partial, int parts, int h, int w) {
  const View& v = ctx.view;
  Unet& net = *ctx.net;
  net.forwardFrom(r, std::is_same_v<T, double>, ctx.kw.dis32.get(), static_cast<int>(v.ny),
                  static_cast<int>(v.nx), 0, 1, h, w, true);
  net.outputApply<T>(ctx.kw.dis32.get(), r, q, z, static_cast<int>(v.ny), static_cast<int>(v.nx),
                     ctx.pin, h, w, partial, parts);
}

// FCG(1) on r (T vectors) into the correction d, to relTol of |r| at the
// start. With `replace` (FP32 vectors), a reliable update whenever the
// recurrence residual has fallen by replaceRatio since the last one: replace()
// moves d into the FP64 solution, recomputes r = b - A x in FP64 and returns
// its norm, and the iteration continues with the same search direction (a
// flexible method takes the perturbed residual as it comes), instead of
// restarting the Krylov space at every FP64 correction.
template <class T>
InnerResult innerFcg(Context& ctx, const OpT<T>& op, double relTol, int64_t maxIter,
                     int64_t window, T* r, T* d, T* const work[4], std::vector<double>& history,
                     const std::function<double()>* replace = nullptr,
                     double replaceRatio = 0.0) {
  const View& v = ctx.view;
  const int64_t n = v.n;
  KrylovWork& w = ctx.kw;
  const int h = static_cast<int>((v.ny + 7) / 8 * 8), wd = static_cast<int>((v.nx + 7) / 8 * 8);
  ctx.net->ensureShape(1, h, wd, false);
  T* z = work[0];
  T* q = work[1];
  T* p[2] = {work[2], work[3]};
  const dim3 g = stride(n);
  const int parts = static_cast<int>(g.x);
  KState k{};
  k.absTol = 0.0;
  k.relTol = relTol;
  k.maxIter = maxIter;
  k.window = window;
  k.status = -1;
  *w.hostState = k;
  CFDHIP_CHECK(hipMemcpyAsync(w.state.get(), w.hostState, sizeof(KState), hipMemcpyHostToDevice,
                              ctx.stream));
  auto fin = [&](int step, int64_t nParts) {
    launchFinalize(ctx, FinalArgs{w.partial.get(), nParts, w.state.get(), w.hist.get(), nullptr,
                                  step});
  };
  hipLaunchKernelGGL((zeroT<T>), g, dim3(256), 0, ctx.stream, n, d);
  hipLaunchKernelGGL((normKernelT<T>), g, dim3(256), 0, ctx.stream, n, r, w.partial.get());
  fin(kStepStartNorm, parts);
  CFDHIP_CHECK(hipGetLastError());
  const int chunk = 4;
  int64_t launched = 0;
  int cur = 0;
  InnerResult res;
  double lastTrue = 0.0;  // the FP64 residual norm at the last reliable update
  while (true) {
    for (int c = 0; c < chunk && launched < maxIter; ++c, ++launched) {
      const bool first = launched == 0;
      // z = M r and (z, q_old) -> beta
      applyM<T>(ctx, r, first ? nullptr : q, z, first ? nullptr : w.partial.get(), parts, h, wd);
      if (!first) fin(kStepFcgBeta, parts);
      hipLaunchKernelGGL((fcgDirSpmvKernel<T>), g, dim3(256), 0, ctx.stream, op, z, p[1 - cur],
                         p[cur], q, r, w.state.get(), first ? 1 : 0, w.partial.get());
      fin(kStepFcgAlpha, parts);
      hipLaunchKernelGGL((fcgUpdateKernel<T>), g, dim3(256), 0, ctx.stream, n, p[cur], q, d, r,
                         w.state.get(), w.partial.get());
      fin(kStepFcgUpdate, parts);
      cur = 1 - cur;
    }
    CFDHIP_CHECK(hipGetLastError());
    CFDHIP_CHECK(hipMemcpyAsync(w.hostState, w.state.get(), sizeof(KState), hipMemcpyDeviceToHost,
                                ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    if (w.hostState->done || launched >= maxIter) break;
    KState& hs = *w.hostState;
    if (replace != nullptr && hs.rnorm <= replaceRatio * (lastTrue > 0.0 ? lastTrue : hs.b0)) {
      const double prev = lastTrue > 0.0 ? lastTrue : hs.b0;
      lastTrue = (*replace)();
      ++res.replacements;
      // the FP64 residual meets the target, or no longer falls (the FP32
      // floor): the caller's FP64 test decides
      if (lastTrue <= relTol * hs.b0 || !(lastTrue < 0.5 * prev)) {
        res.replacedOut = true;
        break;
      }
    }
  }
  const KState& s = *w.hostState;
  res.iterations = s.iter;
  res.status = s.done ? s.status : CFDHIP_MAX_ITERATIONS;
  res.stagnated = s.stagnated != 0;
  std::vector<double> hist(static_cast<size_t>(s.histLen));
  if (s.histLen > 0)
    CFDHIP_CHECK(hipMemcpy(hist.data(), w.hist.get(), sizeof(double) * hist.size(),
                           hipMemcpyDeviceToHost));
  history.insert(history.end(), hist.begin(), hist.end());
  return res;
}

}  // namespace

void learnedScaling(hipStream_t stream, int64_t nx, int64_t ny, int64_t offY, int64_t pin,
                    const double* diag, const double* coeff, float* tmp0, float* tmp1,
                    float* out) {
  const dim3 g = stride(nx * ny);
  hipLaunchKernelGGL(logInteriorDiagKernel, g, dim3(256), 0, stream, nx, ny, offY, diag, coeff,
                     tmp0);
  hipLaunchKernelGGL(boxMeanKernel, g, dim3(256), 0, stream, nx, ny, 0, tmp0, tmp1, 0, pin, diag);
  hipLaunchKernelGGL(boxMeanKernel, g, dim3(256), 0, stream, nx, ny, 1, tmp1, out, 1, pin, diag);
  CFDHIP_CHECK(hipGetLastError());
}

void loadLearnedWeights(Context& ctx, const float* weights) {
  if (!ctx.net) ctx.net = std::make_unique<Unet>(ctx.stream);
  ctx.net->setWeights(weights);
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
}

cfdhip_solve_result learnedFcg2d(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                                 double* x) {
  require(ctx.net && ctx.net->ready(),
          "the learned preconditioner has no weights: call cfdhip_set_learned_weights first");
  const View& v = ctx.view;
  const int64_t n = v.n, nf = v.nf;
  const bool mixed = set.precision == CFDHIP_MIXED;
  KrylovWork& w = ctx.kw;
  w.dis32.ensure(n);
  float* r32 = nullptr;
  float* d32 = nullptr;
  float* work32[4] = {nullptr, nullptr, nullptr, nullptr};
  double* r64 = nullptr;
  double* d64 = nullptr;
  double* work64[4] = {nullptr, nullptr, nullptr, nullptr};
  if (mixed) {
    w.r32.ensure(n);
    w.x32.ensure(n);
    for (int q = 0; q < 4; ++q) w.f8[q].ensure(n);
    w.dinv32.ensure(n);
    w.w32.ensure(nf);
    castToFloat(ctx, n, ctx.f(CFDHIP_FIELD_PC_DIAG), w.dinv32.get());
    castToFloat(ctx, nf, ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), w.w32.get());
    r32 = w.r32.get();
    d32 = w.x32.get();
    for (int q = 0; q < 4; ++q) work32[q] = w.f8[q].get();
  } else {
    for (int q = 0; q < 6; ++q) w.d8[q].ensure(n);
    r64 = w.d8[4].get();
    d64 = w.d8[5].get();
    for (int q = 0; q < 4; ++q) work64[q] = w.d8[q].get();
  }
  const OpT<float> op32{v.nx, v.ny, ctx.pin, v.offY, w.dinv32.get(), w.w32.get()};
  const OpT<double> op64{v.nx, v.ny, ctx.pin, v.offY, ctx.f(CFDHIP_FIELD_PC_DIAG),
                         ctx.f(CFDHIP_FIELD_PC_FACE_COEFF)};
  const dim3 g = stride(n);
  if (ctx.net->padding() == 1) {
    // odd reflection on the sides with a FixedValue pressure patch (2D patch
    // order left, right, bottom, top)
    int sides = 0;
    for (int p = 0; p < 4 && p < v.npatch; ++p)
      if (v.patch[p].pkind == CFDHIP_FIXED_VALUE) sides |= 1 << p;
    ctx.net->setSides(sides);
    w.scaleTmp[0].ensure(n);
    w.scaleTmp[1].ensure(n);
    learnedScaling(ctx.stream, v.nx, v.ny, v.offY, ctx.pin, ctx.f(CFDHIP_FIELD_PC_DIAG),
                   ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), w.scaleTmp[0].get(), w.scaleTmp[1].get(),
                   w.dis32.get());
  } else
    hipLaunchKernelGGL(invSqrtKernel, g, dim3(256), 0, ctx.stream, n, ctx.f(CFDHIP_FIELD_PC_DIAG),
                       w.dis32.get());
  w.partial.ensure(kMaxParts * 4 + 8);
  w.state.ensure(1);
  w.scal.ensure(4);
  w.hist.ensure(set.max_iterations + 8);
  if (!w.hostState) CFDHIP_CHECK(hipHostMalloc(&w.hostState, sizeof(KState)));
  double* normOut = w.scal.get();
  hipEvent_t t0, t1;
  CFDHIP_CHECK(hipEventCreate(&t0));
  CFDHIP_CHECK(hipEventCreate(&t1));
  CFDHIP_CHECK(hipEventRecord(t0, ctx.stream));
  hipLaunchKernelGGL(zeroKernelD64, g, dim3(256), 0, ctx.stream, n, x);
  // FP32 vectors: recurrence reduction between two reliable updates (the
  // settings' inner tolerance). At 2048^2 the iteration count does not depend
  // on it from 0.5 to 0.1 (20 against 16 with FP64 vectors) and grows below
  // (22 at 0.05, 27 at 1e-4); the updates cost an FP64 residual each.
  const double innerTol = set.inner_tolerance > 0.0 ? set.inner_tolerance : 0.1;
  auto outerResidual = [&]() {
    if (mixed)
      hipLaunchKernelGGL((outerResidualKernel<float>), g, dim3(256), 0, ctx.stream, v.nx, v.ny,
                         ctx.pin, v.offY, ctx.f(CFDHIP_FIELD_PC_DIAG),
                         ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b, x, r32, w.partial.get());
    else
      hipLaunchKernelGGL((outerResidualKernel<double>), g, dim3(256), 0, ctx.stream, v.nx, v.ny,
                         ctx.pin, v.offY, ctx.f(CFDHIP_FIELD_PC_DIAG),
                         ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b, x, r64, w.partial.get());
    launchFinalize(ctx, FinalArgs{w.partial.get(), static_cast<int64_t>(g.x), w.state.get(),
                                  nullptr, normOut, kStepNorm});
    CFDHIP_CHECK(hipGetLastError());
    double v0 = 0.0;
    CFDHIP_CHECK(hipMemcpyAsync(&v0, normOut, sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    return v0;
  };
  // a reliable update of the FP32 inner solve: d into x, r = b - A x in FP64
  const std::function<double()> replace = [&]() {
    hipLaunchKernelGGL((addCorrection<float>), g, dim3(256), 0, ctx.stream, n, d32, x);
    hipLaunchKernelGGL((zeroT<float>), g, dim3(256), 0, ctx.stream, n, d32);
    CFDHIP_CHECK(hipGetLastError());
    return outerResidual();
  };
  std::vector<double> history;
  double b0 = 0.0, rTrue = 0.0;
  int64_t total = 0, cycles = 0;
  int status = CFDHIP_MAX_ITERATIONS;
  bool fallback = false;
  // a cycle whose inner solve and fallback add no iteration cannot change x:
  // without this the loop would repeat it forever
  bool noProgress = false;
  int lastStatus = CFDHIP_MAX_ITERATIONS;
  while (true) {
    if (cycles > 0 && noProgress) {  // the last cycle moved nothing: stop with its verdict
      status = lastStatus == CFDHIP_CONVERGED ? CFDHIP_BREAKDOWN : lastStatus;
      break;
    }
    rTrue = outerResidual();
    if (cycles == 0) b0 = rTrue;
    if (!std::isfinite(rTrue)) {
      status = CFDHIP_NONFINITE_RESIDUAL;
      break;
    }
    if (rTrue == 0.0 || rTrue <= set.absolute_tolerance ||
        (b0 > 0.0 && rTrue / b0 <= set.relative_tolerance)) {
      status = CFDHIP_CONVERGED;
      break;
    }
    if (total >= set.max_iterations) break;
    const double need = std::max(set.relative_tolerance * b0, set.absolute_tolerance) / rTrue;
    const int64_t window = set.stagnation_window == 0 ? 200 : std::max(0, set.stagnation_window);
    // FP64 vectors aim at the final target (with a small margin for the
    // recurrence drift); FP32 vectors too, with a reliable update every
    // innerTol of reduction
    const double tol = std::max(0.5 * need, 1e-15);
    InnerResult ir;
    if (mixed)
      ir = innerFcg<float>(ctx, op32, tol, set.max_iterations - total, window, r32, d32, work32,
                           history, &replace, innerTol);
    else
      ir = innerFcg<double>(ctx, op64, tol, set.max_iterations - total, window, r64, d64, work64,
                            history);
    total += ir.iterations;
    ++cycles;
    noProgress = ir.iterations == 0;
    lastStatus = ir.status;
    if (ir.iterations > 0) {
      if (mixed)
        hipLaunchKernelGGL((addCorrection<float>), g, dim3(256), 0, ctx.stream, n, d32, x);
      else
        hipLaunchKernelGGL((addCorrection<double>), g, dim3(256), 0, ctx.stream, n, d64, x);
      CFDHIP_CHECK(hipGetLastError());
    }
    if (ir.stagnated || (ir.iterations == 0 && ir.status != CFDHIP_CONVERGED)) {
      // FP64 Jacobi-PCG from the current x, to the same FP64 target
      fallback = true;
      cfdhip_solver_settings fb = set;
      fb.type = CFDHIP_SOLVER_CG;
      fb.preconditioner = CFDHIP_PREC_JACOBI;
      fb.precision = CFDHIP_FP64;
      fb.max_iterations = std::max<int64_t>(1, set.max_iterations - total);
      // its relative test is against |b - A x| at its start; rescale to |b - A x0|
      const double rNow = outerResidual();
      fb.relative_tolerance = rNow > 0.0 ? std::min(1.0, set.relative_tolerance * b0 / rNow) : 0.0;
      double* work[8];
      for (int q = 0; q < 8; ++q) {
        w.d8[q].ensure(n);
        work[q] = w.d8[q].get();
      }
      const cfdhip_solve_result fr = singleKernelPcg2d<double>(
          ctx, fb, ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b, x, work);
      total += fr.iterations;
      history.insert(history.end(), ctx.lastHistory.begin(), ctx.lastHistory.end());
      if (fr.iterations > 0) noProgress = false;
      lastStatus = fr.status;
      continue;  // the outer test decides, in FP64
    }
  }
  CFDHIP_CHECK(hipEventRecord(t1, ctx.stream));
  CFDHIP_CHECK(hipEventSynchronize(t1));
  float ms = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  cfdhip_solve_result res{};
  res.status = status;
  res.fallback_used = fallback ? 1 : 0;
  res.iterations = total;
  res.restarts = cycles;
  res.initial_residual = b0;
  res.final_residual = rTrue;
  res.true_residual = rTrue;
  res.device_ms = ms;
  ctx.lastHistory = history;
  ctx.lastFallback = fallback ? 1 : 0;
  res.history_length = static_cast<int64_t>(history.size());
  return res;
}

}  // namespace cfdhip
