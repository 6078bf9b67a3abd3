// Training pairs for the learned preconditioner, harvested from the solver's
// own p' solves: an FP64 Jacobi-PCG solve of the assembled system keeps its
// iterate every `stride` iterations; with x* the converged solution, each kept
// x_k gives the error e = x* - x_k and its residual r = A e. The pairs are
// written in the symmetrically scaled system (unit diagonal):
//   e^ = D^1/2 e,  r^ = D^-1/2 r = D^-1/2 A D^-1/2 e^,
// normalised to unit RMS of r^, zero-padded to the network grid h x w, with
// den = (e^, r^) (the squared A-norm of the error, for the loss). White-noise
// errors (harvestNoisePairs) add the rough modes a converging solve no longer
// carries.
#include <cmath>
#include <vector>

#include "reduce.hpp"

namespace cfdhip {

namespace {

constexpr int kParts = 1024;

// Over the padded h x w grid: e^, r^ of one snapshot (unnormalised) and the
// partials (r^, r^), (e^, r^).
__global__ __launch_bounds__(256) void pairKernel(int64_t nx, int64_t ny, int64_t pin, int64_t offY,
                                                  const double* __restrict__ diag,
                                                  const double* __restrict__ coeff,
                                                  const double* __restrict__ xs,
                                                  const double* __restrict__ xk, int64_t h,
                                                  int64_t w, float* __restrict__ rOut,
                                                  float* __restrict__ eOut, double* partial) {
  double acc[2] = {0.0, 0.0};
  const int64_t total = h * w;
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = p % w, j = p / w;
    if (i >= nx || j >= ny) {
      rOut[p] = 0.f;
      eOut[p] = 0.f;
      continue;
    }
    const int64_t c = j * nx + i;
    auto e = [&](int64_t k) { return xs[k] - xk[k]; };
    double y;
    if (c == pin) {
      y = diag[c] * e(c);
    } else {
      y = 0.0;
      if (j > 0) y -= coeff[offY + j * nx + i] * e(c - nx);
      if (i > 0) y -= coeff[j * (nx + 1) + i] * e(c - 1);
      y += diag[c] * e(c);
      if (i < nx - 1) y -= coeff[j * (nx + 1) + i + 1] * e(c + 1);
      if (j < ny - 1) y -= coeff[offY + (j + 1) * nx + i] * e(c + nx);
    }
    const double sq = std::sqrt(diag[c]);
    const double eh = sq * e(c), rh = y / sq;
    eOut[p] = static_cast<float>(eh);
    rOut[p] = static_cast<float>(rh);
    acc[0] += rh * rh;
    acc[1] += eh * rh;
  }
  blockPartials<2>(acc, partial);
}

__global__ void scalePairKernel(int64_t total, float s, float* __restrict__ r, float* __restrict__ e) {
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    r[p] *= s;
    e[p] *= s;
  }
}

__global__ void sumPairKernel(const double* __restrict__ partial, int parts, double* out) {
  double v[2];
  sumPartials<2>(partial, parts, v);
  if (threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
  }
}

// Scaled couplings of D^-1/2 A D^-1/2 on the padded grid: ce to (i + 1, j),
// cn to (i, j + 1); zero at the domain edge, in the padding and around the
// pinned cell (its row is the identity).
__global__ void couplingKernel(int64_t nx, int64_t ny, int64_t pin, int64_t offY,
                               const double* __restrict__ diag, const double* __restrict__ coeff,
                               int64_t h, int64_t w, float* __restrict__ ce,
                               float* __restrict__ cn) {
  const int64_t total = h * w;
  for (int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; p < total;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = p % w, j = p / w;
    float e = 0.f, nn = 0.f;
    if (i < nx && j < ny) {
      const int64_t c = j * nx + i;
      if (i + 1 < nx && c != pin && c + 1 != pin)
        e = static_cast<float>(coeff[j * (nx + 1) + i + 1] / std::sqrt(diag[c] * diag[c + 1]));
      if (j + 1 < ny && c != pin && c + nx != pin)
        nn = static_cast<float>(coeff[offY + (j + 1) * nx + i] / std::sqrt(diag[c] * diag[c + nx]));
    }
    ce[p] = e;
    cn[p] = nn;
  }
}

// White-noise error vectors: e_i uniform in [-1, 1] from splitmix64 of
// (seed, pair, cell), zero outside the mesh (x* - x_k with x_k = 0 and x* = e
// in pairKernel).
__global__ void noiseKernel(int64_t n, uint64_t seed, double* __restrict__ e) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(c + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    e[c] = 2.0 * (static_cast<double>(z >> 11) * 0x1.0p-53) - 1.0;
  }
}

__global__ void zeroKernelH(int64_t n, double* __restrict__ a) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[c] = 0.0;
}

// Writes pair k from the error x* - x_k (device vectors of n): scaled pair
// into rOut / eOut, den into denOut[k]; false when the error is zero.
bool emitPair(Context& ctx, const double* xs, const double* xk, int64_t h, int64_t w, float* r,
              float* e, float* den, DeviceArray<double>& part, DeviceArray<double>& sums) {
  const View& v = ctx.view;
  const dim3 g(kParts);
  hipLaunchKernelGGL(pairKernel, g, dim3(256), 0, ctx.stream, v.nx, v.ny, ctx.pin, v.offY,
                     ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), xs, xk, h, w,
                     r, e, part.get());
  hipLaunchKernelGGL(sumPairKernel, dim3(1), dim3(kFinThreads), 0, ctx.stream, part.get(), kParts,
                     sums.get());
  double hs[2];
  CFDHIP_CHECK(hipMemcpyAsync(hs, sums.get(), sizeof(hs), hipMemcpyDeviceToHost, ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  const double ms = hs[0] / static_cast<double>(v.n);
  if (!(ms > 0.0) || !std::isfinite(ms) || !(hs[1] > 0.0)) return false;  // x_k == x*
  const double sc = 1.0 / std::sqrt(ms);
  hipLaunchKernelGGL(scalePairKernel, g, dim3(256), 0, ctx.stream, h * w, static_cast<float>(sc),
                     r, e);
  *den = static_cast<float>(hs[1] * sc * sc);
  return true;
}

}  // namespace

int64_t harvestNoisePairs(Context& ctx, uint64_t seed, int64_t count, int64_t h, int64_t w,
                          float* rOut, float* eOut, float* denOut) {
  const View& v = ctx.view;
  require(v.dim == 2, "harvesting needs a 2D mesh");
  require(h >= v.ny && w >= v.nx, "the network grid must cover the mesh");
  require(count >= 0, "count must be >= 0");
  const int64_t n = v.n;
  DeviceArray<double> e, zero, part, sums;
  e.ensure(n);
  zero.ensure(n);
  part.ensure(kParts * 2 + 8);
  sums.ensure(2);
  const dim3 g(kParts);
  hipLaunchKernelGGL(zeroKernelH, g, dim3(256), 0, ctx.stream, n, zero.get());
  int64_t kept = 0;
  for (int64_t k = 0; k < count; ++k) {
    hipLaunchKernelGGL(noiseKernel, g, dim3(256), 0, ctx.stream, n,
                       seed * 0x2545F4914F6CDD1Dull + static_cast<uint64_t>(k) * 0xD1B54A32D192ED03ull,
                       e.get());
    if (emitPair(ctx, e.get(), zero.get(), h, w, rOut + kept * h * w, eOut + kept * h * w,
                 denOut + kept, part, sums))
      ++kept;
  }
  CFDHIP_CHECK(hipGetLastError());
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return kept;
}

int64_t harvestPairs(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                     int64_t stride, int64_t maxPairs, int64_t h, int64_t w, float* rOut,
                     float* eOut, float* ceOut, float* cnOut, float* denOut,
                     cfdhip_solve_result* result) {
  const View& v = ctx.view;
  require(v.dim == 2, "harvesting needs a 2D mesh");
  require(h >= v.ny && w >= v.nx, "the network grid must cover the mesh");
  require(stride > 0 && maxPairs >= 0, "stride must be positive");
  const int64_t n = v.n;
  KrylovWork& kw = ctx.kw;
  DeviceArray<double> snaps;
  snaps.ensure(std::max<int64_t>(1, maxPairs) * n);
  double* x = ctx.f(CFDHIP_FIELD_PPRIME);
  CFDHIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n, ctx.stream));
  double* work[8];
  for (int q = 0; q < 8; ++q) {
    kw.d8[q].ensure(n);
    work[q] = kw.d8[q].get();
  }
  Snapshots sn;
  sn.buf = snaps.get();
  sn.stride = stride;
  sn.cap = maxPairs;
  cfdhip_solver_settings s = set;
  s.type = CFDHIP_SOLVER_CG;
  s.preconditioner = CFDHIP_PREC_JACOBI;
  s.precision = CFDHIP_FP64;
  const cfdhip_solve_result res =
      singleKernelPcg2d<double>(ctx, s, ctx.f(CFDHIP_FIELD_PC_DIAG),
                                ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b, x, work, &sn);
  if (result != nullptr) *result = res;
  const int64_t usable = std::min<int64_t>(sn.taken, res.iterations / stride);
  DeviceArray<double> part, sums;
  part.ensure(kParts * 2 + 8);
  sums.ensure(2);
  int64_t kept = 0;
  const dim3 g(kParts);
  for (int64_t k = 0; k < usable; ++k) {
    if (emitPair(ctx, x, snaps.get() + k * n, h, w, rOut + kept * h * w, eOut + kept * h * w,
                 denOut + kept, part, sums))
      ++kept;
  }
  if (ceOut != nullptr && cnOut != nullptr)
    hipLaunchKernelGGL(couplingKernel, g, dim3(256), 0, ctx.stream, v.nx, v.ny, ctx.pin, v.offY,
                       ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), h, w, ceOut,
                       cnOut);
  CFDHIP_CHECK(hipGetLastError());
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return kept;
}

}  // namespace cfdhip
