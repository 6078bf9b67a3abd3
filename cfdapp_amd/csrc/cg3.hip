// Jacobi-PCG for the 3D p' operator (7-point), in FP64 or FP32, and the
// mixed-precision solve (FP32 iteration with FP64 residual replacement).
//
// Same Chronopoulos-Gear recurrence and device-side scalar steps as the 2D
// single-kernel iteration (cg1.hip; src/algebra/CG.cpp:77-142 in exact
// arithmetic). FP64 runs one kernel per iteration on z-marching tiles
// (cg3IterKernel below); the FP32 iteration of the mixed solve keeps two:
//   update: s = w + beta s, r -= alpha s, p = M r_old + beta p, x += alpha p,
//           u = M r                                  (pointwise, in place)
//   apply:  w = A u; (r, u), (w, u), |r|^2           (7-point stencil)
// Per iteration: update reads p, s, r, w, x, diag and writes p, s, r, x, u;
// apply reads u (+ neighbours, from cache), r, diag, the three face arrays and
// writes w: 8 * 19 n bytes in FP64.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <utility>
#include <vector>

#include "reduce.hpp"

namespace cfdhip {

namespace {

constexpr int kMaxBlocks3 = 2048;

template <class T>
struct Op3 {
  int64_t nx, ny, nz, pin, offY, offZ;
  const T* diag;
  const T* coeff;
};

// Row c of A applied to u, ascending column order from 0 (PressureOp::apply).
template <class T>
__device__ inline T applyRow3(const Op3<T>& g, int64_t i, int64_t j, int64_t k, int64_t c,
                              const T* __restrict__ u) {
  if (c == g.pin) return T(0) + g.diag[c] * u[c];
  const int64_t sxy = g.nx * g.ny;
  T y = T(0);
  if (k > 0) y += (-g.coeff[g.offZ + (k * g.ny + j) * g.nx + i]) * u[c - sxy];
  if (j > 0) y += (-g.coeff[g.offY + (k * (g.ny + 1) + j) * g.nx + i]) * u[c - g.nx];
  if (i > 0) y += (-g.coeff[(k * g.ny + j) * (g.nx + 1) + i]) * u[c - 1];
  y += g.diag[c] * u[c];
  if (i < g.nx - 1) y += (-g.coeff[(k * g.ny + j) * (g.nx + 1) + i + 1]) * u[c + 1];
  if (j < g.ny - 1) y += (-g.coeff[g.offY + (k * (g.ny + 1) + j + 1) * g.nx + i]) * u[c + g.nx];
  if (k < g.nz - 1) y += (-g.coeff[g.offZ + ((k + 1) * g.ny + j) * g.nx + i]) * u[c + sxy];
  return y;
}

inline dim3 grid3(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > kMaxBlocks3) b = kMaxBlocks3;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : b));
}

// r0 = b - A x0, |r0|^2.
template <class T>
__global__ __launch_bounds__(256) void start3Kernel(Op3<T> g, const T* __restrict__ b,
                                                    const T* __restrict__ x, T* __restrict__ r,
                                                    double* partial) {
  const int64_t n = g.nx * g.ny * g.nz;
  double acc[1] = {0.0};
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % g.nx, t = c / g.nx, j = t % g.ny, k = t / g.ny;
    const T ri = b[c] - applyRow3<T>(g, i, j, k, c, x);
    r[c] = ri;
    acc[0] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  blockPartials<1>(acc, partial);
}

// The iterate x has its own type TX (FP64 under the mixed solve's FP32
// iteration vectors, as in the 2D iteration).
template <class T, class TX = T>
__global__ __launch_bounds__(256) void update3Kernel(int64_t n, const T* __restrict__ diag,
                                                     T* __restrict__ p, T* __restrict__ s,
                                                     T* __restrict__ r, const T* __restrict__ w,
                                                     TX* __restrict__ x, T* __restrict__ u,
                                                     const KState* __restrict__ st, int start) {
  if (st->done) return;
  const T alpha = static_cast<T>(st->alpha);
  const T beta = static_cast<T>(st->beta);
  const TX alphaX = static_cast<TX>(st->alpha);
  const bool first = st->iter == st->dirStart;
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const T d = diag[c];
    T rn = r[c];
    if (!start) {
      const T w0 = w[c];
      const T sn = first ? w0 : w0 + (s[c] * beta);
      const T uOld = (T(1) / d) * rn;
      const T pn = first ? uOld : uOld + (p[c] * beta);
      rn = rn - (sn * alpha);
      s[c] = sn;
      p[c] = pn;
      r[c] = rn;
      x[c] = x[c] + (static_cast<TX>(pn) * alphaX);
    }
    u[c] = (T(1) / d) * rn;
  }
}

// NSUM = 6 adds the explicit-conjugation sums of the mixed solve, (u, s),
// (p, s), (r, p), with p, s the direction the update kernel just formed.
template <class T, int NSUM = 3>
__global__ __launch_bounds__(256) void apply3Kernel(Op3<T> g, const T* __restrict__ u,
                                                    const T* __restrict__ r, T* __restrict__ w,
                                                    const T* __restrict__ p,
                                                    const T* __restrict__ s,
                                                    const KState* __restrict__ st,
                                                    double* partial) {
  double acc[NSUM] = {};
  if (!st->done) {
    const int64_t n = g.nx * g.ny * g.nz;
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const int64_t i = c % g.nx, t = c / g.nx, j = t % g.ny, k = t / g.ny;
      const T wc = applyRow3<T>(g, i, j, k, c, u);
      w[c] = wc;
      const T uc = u[c], rc = r[c];
      acc[0] += static_cast<double>(rc) * static_cast<double>(uc);
      acc[1] += static_cast<double>(wc) * static_cast<double>(uc);
      acc[2] += static_cast<double>(rc) * static_cast<double>(rc);
      if constexpr (NSUM == 6) {
        const double sc = s[c], pc = p[c];
        acc[3] += static_cast<double>(uc) * sc;
        acc[4] += pc * sc;
        acc[5] += static_cast<double>(rc) * pc;
      }
    }
  }
  blockPartials<NSUM>(acc, partial);
}

// r64 = b - A x (FP64) -> r32, |r64|^2 partials (refinement step).
__global__ __launch_bounds__(256) void refine3Kernel(Op3<double> g, const double* __restrict__ b,
                                                     const double* __restrict__ x,
                                                     float* __restrict__ r32, double* partial) {
  const int64_t n = g.nx * g.ny * g.nz;
  double acc[1] = {0.0};
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % g.nx, t = c / g.nx, j = t % g.ny, k = t / g.ny;
    const double rc = b[c] - applyRow3<double>(g, i, j, k, c, x);
    r32[c] = static_cast<float>(rc);
    acc[0] += rc * rc;
  }
  blockPartials<1>(acc, partial);
}

__global__ void zero3d(int64_t n, double* __restrict__ x) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[c] = 0.0;
}

// u = M r with the iteration's rounding (update3Kernel), whatever the verdict.
__global__ void invDiag3(int64_t n, const float* __restrict__ diag, const float* __restrict__ r,
                         float* __restrict__ u) {
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    u[c] = (1.0f / diag[c]) * r[c];
}

// After a residual replacement: w = A u and the six sums of kStepCgReplace,
// (r, u), (w, u), |r|^2, (u, s), (p, s), (r, p).
__global__ __launch_bounds__(256) void replace3Kernel(Op3<float> g, const float* __restrict__ u,
                                                      const float* __restrict__ r,
                                                      const float* __restrict__ p,
                                                      const float* __restrict__ s,
                                                      float* __restrict__ w, double* partial) {
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t n = g.nx * g.ny * g.nz;
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % g.nx, t = c / g.nx, j = t % g.ny, k = t / g.ny;
    const float wc = applyRow3<float>(g, i, j, k, c, u);
    w[c] = wc;
    const double uc = u[c], rc = r[c], sc = s[c], pc = p[c];
    acc[0] += rc * uc;
    acc[1] += static_cast<double>(wc) * uc;
    acc[2] += rc * rc;
    acc[3] += uc * sc;
    acc[4] += pc * sc;
    acc[5] += rc * pc;
  }
  blockPartials<6>(acc, partial);
}

// One FP64 Jacobi-PCG iteration in a single kernel (the Chronopoulos-Gear
// step of cg1.hip's 2D iteration) on z-marching tiles: a block owns 62 x
// (ROWS - 2) interior cells of an x-y tile (one lane per column, +1 halo on
// every side) over KZ planes, and walks its planes upwards. The walk is
// sequential (two barriers per plane), so short walks over many blocks win. For each plane it
// forms the new residual and u = M r of every tile point (the halo ring and
// the planes just below and above its range too, from the previous iterate)
// into a three-plane LDS ring, writes p, s, r, x of its own cells, and applies
// the 7-point stencil to the plane below from the ring. Per iteration a cell
// is read once (d, r, w, s, p, x, its three lower faces) and written once
// (p, s, r, x, w): 14 values against the 19 of the two-kernel form. r, s, w
// alternate between two buffers (a halo reads the previous iterate); p and x
// are only read by their own block and stay in place. start == 1: the start
// step w0 = A M r0 (rOld = r0, only wNew written).
struct Grid3 {
  int64_t nx, ny, nz, pin, offY, offZ;
};
constexpr int kOut3 = 62;

template <int ROWS, int KZ>
__global__ __launch_bounds__(256) void cg3IterKernel(
    Grid3 g, const double* __restrict__ diag, const double* __restrict__ coeff,
    double* __restrict__ p, const double* __restrict__ sOld, double* __restrict__ sNew,
    const double* __restrict__ rOld, double* __restrict__ rNew, const double* __restrict__ wOld,
    double* __restrict__ wNew, double* __restrict__ x, int64_t tilesX, int64_t tilesY,
    const KState* st, double* partial, int start) {
  constexpr int kRW = ROWS / 4;  // rows per wave
  constexpr int kIR = ROWS - 2;  // interior rows
  __shared__ double U[3][ROWS][64];
  const bool skip = st->done;
  const double alpha = st->alpha, beta = st->beta;
  const bool first = st->iter == st->dirStart;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t tile = xcdTileOrder(blockIdx.x, gridDim.x);
  const int64_t perPlane = tilesX * tilesY;
  const int64_t tz = tile / perPlane, txy = tile % perPlane;
  const int64_t tx = txy % tilesX, ty = txy / tilesX;
  const int64_t i = tx * kOut3 - 1 + lane;
  const int64_t jBase = ty * kIR - 1;
  const int64_t kz0 = tz * KZ;
  const bool col = i >= 0 && i < g.nx;
  const bool outCol = col && lane >= 1 && lane <= kOut3;
  const int64_t sxy = g.nx * g.ny;
  double acc[3] = {0.0, 0.0, 0.0};
  auto slot = [](int kk) { return (kk + 3) % 3; };
  for (int kk = -1; kk <= KZ && !skip; ++kk) {
    const int64_t k = kz0 + kk;
    const bool planeValid = k >= 0 && k < g.nz;
    const bool planeOwn = kk >= 0 && kk < KZ && planeValid;
    const int sl = slot(kk);
#pragma unroll
    for (int q = 0; q < kRW; ++q) {
      const int tr = wave * kRW + q;
      const int64_t j = jBase + tr;
      const bool valid = planeValid && col && j >= 0 && j < g.ny;
      const bool center = valid && planeOwn && outCol && tr >= 1 && tr <= kIR;
      double u = 0.0;
      if (valid) {
        const int64_t c = k * sxy + j * g.nx + i;
        const double d = diag[c];
        const double r0 = rOld[c];
        double sn = 0.0, rn = r0;
        if (!start) {
          const double w0 = wOld[c];
          sn = first ? w0 : w0 + (sOld[c] * beta);
          rn = r0 - (sn * alpha);
        }
        u = (1.0 / d) * rn;
        if (center) {
          if (!start) {
            const double uOld = (1.0 / d) * r0;
            const double pn = first ? uOld : uOld + (p[c] * beta);
            p[c] = pn;
            sNew[c] = sn;
            rNew[c] = rn;
            x[c] = x[c] + (pn * alpha);
          }
          acc[0] += rn * u;
          acc[2] += rn * rn;
        }
      }
      U[sl][tr][lane] = u;
    }
    __syncthreads();
    // the stencil on the plane below (its neighbours above are now in the ring)
    const int kb = kk - 1;
    const int64_t k1 = kz0 + kb;
    if (kb >= 0 && k1 < g.nz) {
      const int sm = slot(kb), sd = slot(kb - 1), su = slot(kk);
#pragma unroll
      for (int q = 0; q < kRW; ++q) {
        const int tr = wave * kRW + q;
        const int64_t j = jBase + tr;
        const bool center = outCol && j >= 0 && j < g.ny && tr >= 1 && tr <= kIR;
        const bool inGrid = col && j >= 0 && j < g.ny;
        const double uC = U[sm][tr][lane];
        const double uW = __shfl_up(uC, 1, 64), uE = __shfl_down(uC, 1, 64);
        const int64_t c = k1 * sxy + j * g.nx + i;
        // every tile point loads its west face, so the east face of the last
        // interior column comes from the halo lane
        const double fxW = inGrid ? coeff[(k1 * g.ny + j) * (g.nx + 1) + i] : 0.0;
        const double fxE = __shfl_down(fxW, 1, 64);
        if (center) {
          const double d = diag[c];
          double y;
          if (c == g.pin) {
            y = 0.0 + d * uC;
          } else {
            y = 0.0;
            if (k1 > 0) y += (-coeff[g.offZ + (k1 * g.ny + j) * g.nx + i]) * U[sd][tr][lane];
            if (j > 0) y += (-coeff[g.offY + (k1 * (g.ny + 1) + j) * g.nx + i]) * U[sm][tr - 1][lane];
            if (i > 0) y += (-fxW) * uW;
            y += d * uC;
            if (i < g.nx - 1) y += (-fxE) * uE;
            if (j < g.ny - 1)
              y += (-coeff[g.offY + (k1 * (g.ny + 1) + j + 1) * g.nx + i]) * U[sm][tr + 1][lane];
            if (k1 < g.nz - 1)
              y += (-coeff[g.offZ + ((k1 + 1) * g.ny + j) * g.nx + i]) * U[su][tr][lane];
          }
          wNew[c] = y;
          acc[1] += y * uC;
        }
      }
    }
    __syncthreads();  // the next plane overwrites the ring slot just read
  }
  blockPartials<3>(acc, partial);
}

}  // namespace

// work: r s w p u (n each)
template <class T>
cfdhip_solve_result pcg3d(Context& ctx, const cfdhip_solver_settings& set, const T* diag,
                          const T* coeff, const T* b, T* x, T* const work[5]) {
  const View& v = ctx.view;
  require(v.dim == 3, "the 3D PCG needs a 3D mesh");
  KrylovWork& w = ctx.kw;
  const int64_t n = v.n;
  const dim3 g = grid3(n);
  w.partial.ensure(std::max<int64_t>(w.partial.count, kMaxBlocks3 * 3 + 8));
  w.state.ensure(1);
  if (!w.hostState) CFDHIP_CHECK(hipHostMalloc(&w.hostState, sizeof(KState)));
  w.hist.ensure(set.max_iterations + 2);
  KState k{};
  k.absTol = set.absolute_tolerance;
  k.relTol = set.relative_tolerance;
  k.maxIter = set.max_iterations;
  k.status = -1;
  *w.hostState = k;
  CFDHIP_CHECK(hipMemcpyAsync(w.state.get(), w.hostState, sizeof(KState), hipMemcpyHostToDevice,
                              ctx.stream));
  const Op3<T> op{v.nx, v.ny, v.nz, ctx.pin, v.offY, v.offZ, diag, coeff};
  // the single-kernel iteration's buffers: r, s, w twice (work 0-2 and two
  // more of the context's), p in place (work 3)
  for (int q = 5; q < 8; ++q) w.d8[q].ensure(n);
  static_assert(std::is_same_v<T, double>, "the single-kernel 3D iteration runs in FP64");
  T* Rv[2] = {work[0], w.d8[5].get()};
  T* Sv[2] = {work[1], w.d8[6].get()};
  T* Wv[2] = {work[2], w.d8[7].get()};
  T* p = work[3];
  const Grid3 g3{v.nx, v.ny, v.nz, ctx.pin, v.offY, v.offZ};
  // tile shape measured at 256^3 (ms per iteration): 16 x 4 planes 0.538, 12 x 6 0.540,
  // 12 x 8 0.550, 8 x 8 0.579, 16 x 32 0.746, 32 x 8 1.65; the two-kernel form 0.633
  constexpr int kRows3 = 16, kPlanes3 = 4;
  const int64_t tilesX = (v.nx + kOut3 - 1) / kOut3;
  const int64_t tilesY = (v.ny + (kRows3 - 2) - 1) / (kRows3 - 2);
  const int64_t zChunks = (v.nz + kPlanes3 - 1) / kPlanes3;
  const dim3 gt(static_cast<unsigned>(tilesX * tilesY * zChunks));
  w.partial.ensure(std::max<int64_t>(w.partial.count, static_cast<int64_t>(gt.x) * 3 + 8));
  auto fin = [&](int step, dim3 grid) {
    launchFinalize(ctx, FinalArgs{w.partial.get(), static_cast<int64_t>(grid.x), w.state.get(),
                                  w.hist.get(), nullptr, step});
  };
  auto iterate = [&](int cur, int start) {
    hipLaunchKernelGGL((cg3IterKernel<kRows3, kPlanes3>), gt, dim3(256), 0, ctx.stream, g3, diag,
                       coeff, p, Sv[cur], Sv[1 - cur], Rv[cur], Rv[1 - cur], Wv[cur], Wv[1 - cur], x,
                       tilesX, tilesY, w.state.get(), w.partial.get(), start);
  };
  hipEvent_t t0, t1;
  CFDHIP_CHECK(hipEventCreate(&t0));
  CFDHIP_CHECK(hipEventCreate(&t1));
  CFDHIP_CHECK(hipEventRecord(t0, ctx.stream));
  hipLaunchKernelGGL((start3Kernel<T>), g, dim3(256), 0, ctx.stream, op, b, x, Rv[0],
                     w.partial.get());
  fin(kStepStartNorm, g);
  // start step: w0 = A M r0 from R[0] into W[1]; r0 then moves to R[1], so
  // the first iteration reads index 1 (R, W; S is not read on a first step)
  iterate(0, 1);
  fin(kStepCg1Start, gt);
  std::swap(Rv[0], Rv[1]);
  CFDHIP_CHECK(hipGetLastError());
  int cur = 1;
  const int chunk = set.check_interval > 0 ? set.check_interval : 32;
  int64_t launched = 0;
  while (true) {
    for (int q = 0; q < chunk && launched < set.max_iterations; ++q, ++launched) {
      iterate(cur, 0);
      fin(kStepCg1Iter, gt);
      cur = 1 - cur;
    }
    CFDHIP_CHECK(hipGetLastError());
    CFDHIP_CHECK(hipMemcpyAsync(w.hostState, w.state.get(), sizeof(KState), hipMemcpyDeviceToHost,
                                ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    if (w.hostState->done || launched >= set.max_iterations) break;
  }
  CFDHIP_CHECK(hipEventRecord(t1, ctx.stream));
  CFDHIP_CHECK(hipEventSynchronize(t1));
  float ms = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  const KState& S = *w.hostState;
  cfdhip_solve_result res{};
  res.status = S.done ? S.status : CFDHIP_MAX_ITERATIONS;
  res.iterations = S.iter;
  res.initial_residual = S.b0;
  res.final_residual = S.rnorm;
  res.true_residual = S.rnorm;
  res.device_ms = ms;
  ctx.lastHistory.resize(S.histLen);
  if (S.histLen > 0)
    CFDHIP_CHECK(hipMemcpy(ctx.lastHistory.data(), w.hist.get(), S.histLen * sizeof(double),
                           hipMemcpyDeviceToHost));
  res.history_length = S.histLen;
  return res;
}

template cfdhip_solve_result pcg3d<double>(Context&, const cfdhip_solver_settings&, const double*,
                                           const double*, const double*, double*,
                                           double* const[5]);

// Mixed precision in 3D with FP64 residual replacement: the scheme of the 2D
// solve (cg1.hip, mixedPcg2d) on the two-kernel 3D iteration. FP32 p, s, r,
// w, u on an FP32 copy of the operator, accumulating into the FP64 x; when
// the FP32 recurrence residual is below `ratio` times its largest value
// since the last replacement (or below the target) the host recomputes
// r = b - A x in FP64 and continues with the same search
// direction (kStepCgReplace). A replacement that does not halve the FP64
// residual, or an FP32 breakdown, hands the rest to FP64 Jacobi-PCG from x.
cfdhip_solve_result mixedPcg3d(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                               double* x) {
  const View& v = ctx.view;
  const int64_t n = v.n, nf = v.nf;
  KrylovWork& w = ctx.kw;
  for (int q = 0; q < 5; ++q) w.f8[q].ensure(n);
  w.dinv32.ensure(n);
  w.w32.ensure(nf);
  castToFloat(ctx, n, ctx.f(CFDHIP_FIELD_PC_DIAG), w.dinv32.get());
  castToFloat(ctx, nf, ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), w.w32.get());
  w.partial.ensure(std::max<int64_t>(w.partial.count, kMaxBlocks3 * 6 + 8));
  w.state.ensure(1);
  w.scal.ensure(4);
  if (!w.hostState) CFDHIP_CHECK(hipHostMalloc(&w.hostState, sizeof(KState)));
  w.hist.ensure(set.max_iterations + 2);
  const dim3 g = grid3(n);
  const double* diag64 = ctx.f(CFDHIP_FIELD_PC_DIAG);
  const double* coeff64 = ctx.f(CFDHIP_FIELD_PC_FACE_COEFF);
  const Op3<double> op64{v.nx, v.ny, v.nz, ctx.pin, v.offY, v.offZ, diag64, coeff64};
  const Op3<float> op32{v.nx, v.ny, v.nz, ctx.pin, v.offY, v.offZ, w.dinv32.get(), w.w32.get()};
  const float* diag32 = w.dinv32.get();
  float* r = w.f8[0].get();
  float* s = w.f8[1].get();
  float* wv = w.f8[2].get();
  float* p = w.f8[3].get();
  float* u = w.f8[4].get();
  const double ratio = set.inner_tolerance > 0.0 ? set.inner_tolerance : 1e-5;
  auto fin = [&](int step) {
    launchFinalize(ctx, FinalArgs{w.partial.get(), static_cast<int64_t>(g.x), w.state.get(),
                                  w.hist.get(), w.scal.get(), step});
  };
  auto trueResidual = [&]() {
    hipLaunchKernelGGL(refine3Kernel, g, dim3(256), 0, ctx.stream, op64, b, x, r,
                       w.partial.get());
    fin(kStepNorm);
    CFDHIP_CHECK(hipGetLastError());
    double out = 0.0;
    CFDHIP_CHECK(hipMemcpyAsync(&out, w.scal.get(), sizeof(double), hipMemcpyDeviceToHost,
                                ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    return out;
  };
  hipEvent_t t0, t1;
  CFDHIP_CHECK(hipEventCreate(&t0));
  CFDHIP_CHECK(hipEventCreate(&t1));
  CFDHIP_CHECK(hipEventRecord(t0, ctx.stream));
  hipLaunchKernelGGL(zero3d, g, dim3(256), 0, ctx.stream, n, x);
  const double b0 = trueResidual();
  double rTrue = b0;
  auto converged = [&](double a) {
    return a == 0.0 || a <= set.absolute_tolerance ||
           (b0 > 0.0 && a / b0 <= set.relative_tolerance);
  };
  int status = CFDHIP_MAX_ITERATIONS;
  int64_t replacements = 0, fallbackIters = 0;
  bool fallback = false;
  std::vector<double> history{b0};
  KState k{};
  k.b0 = b0;
  k.absTol = set.absolute_tolerance;
  k.relTol = set.relative_tolerance;
  k.maxIter = set.max_iterations;
  k.status = -1;
  k.replaceRatio = ratio;
  k.restartOnReplace = kRestartOnReplace;
  k.rmax = b0;
  *w.hostState = k;
  if (!std::isfinite(b0)) {
    status = CFDHIP_NONFINITE_RESIDUAL;
  } else if (converged(b0)) {
    status = CFDHIP_CONVERGED;
  } else {
    CFDHIP_CHECK(hipMemcpyAsync(w.state.get(), w.hostState, sizeof(KState),
                                hipMemcpyHostToDevice, ctx.stream));
    hipLaunchKernelGGL((update3Kernel<float, double>), g, dim3(256), 0, ctx.stream, n, diag32, p,
                       s, r, wv, x, u, w.state.get(), 1);
    hipLaunchKernelGGL((apply3Kernel<float, 6>), g, dim3(256), 0, ctx.stream, op32, u, r, wv, p,
                       s, w.state.get(), w.partial.get());
    fin(kStepCgRobustStart);
    CFDHIP_CHECK(hipGetLastError());
    const int chunk = set.check_interval > 0 ? set.check_interval : 32;
    double lastTrue = b0;
    while (true) {
      const int64_t left = set.max_iterations - w.hostState->iter;
      for (int q = 0; q < chunk && q < left; ++q) {
        hipLaunchKernelGGL((update3Kernel<float, double>), g, dim3(256), 0, ctx.stream, n,
                           diag32, p, s, r, wv, x, u, w.state.get(), 0);
        hipLaunchKernelGGL((apply3Kernel<float, 6>), g, dim3(256), 0, ctx.stream, op32, u, r, wv,
                           p, s, w.state.get(), w.partial.get());
        fin(kStepCgRobust);
      }
      CFDHIP_CHECK(hipGetLastError());
      CFDHIP_CHECK(hipMemcpyAsync(w.hostState, w.state.get(), sizeof(KState),
                                  hipMemcpyDeviceToHost, ctx.stream));
      CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
      KState& S = *w.hostState;
      if (!S.done) continue;
      rTrue = trueResidual();
      history.push_back(rTrue);
      if (!std::isfinite(rTrue)) {
        status = CFDHIP_NONFINITE_RESIDUAL;
        break;
      }
      if (converged(rTrue)) {
        status = CFDHIP_CONVERGED;
        break;
      }
      if (S.iter >= set.max_iterations) {
        status = CFDHIP_MAX_ITERATIONS;
        break;
      }
      if (!S.replace || (replacements > 0 && rTrue > 0.5 * lastTrue)) {
        cfdhip_solver_settings rest = set;
        rest.precision = CFDHIP_FP64;
        rest.relative_tolerance = set.relative_tolerance * b0 / rTrue;
        rest.max_iterations = std::max<int64_t>(1, set.max_iterations - S.iter);
        double* work64[5];
        for (int q = 0; q < 5; ++q) {
          w.d8[q].ensure(n);
          work64[q] = w.d8[q].get();
        }
        const int64_t done = S.iter;
        const cfdhip_solve_result fr = pcg3d<double>(ctx, rest, diag64, coeff64, b, x, work64);
        fallback = true;
        fallbackIters = fr.iterations;
        w.hostState->iter = done;
        history.insert(history.end(), ctx.lastHistory.begin(), ctx.lastHistory.end());
        // pcg3d reports its recurrence residual: measure the FP64 one
        rTrue = trueResidual();
        status = converged(rTrue) ? CFDHIP_CONVERGED
                                  : (fr.status == CFDHIP_CONVERGED ? CFDHIP_MAX_ITERATIONS
                                                                   : fr.status);
        break;
      }
      lastTrue = rTrue;
      ++replacements;
      hipLaunchKernelGGL(invDiag3, g, dim3(256), 0, ctx.stream, n, diag32, r, u);
      hipLaunchKernelGGL(replace3Kernel, g, dim3(256), 0, ctx.stream, op32, u, r, p, s, wv,
                         w.partial.get());
      fin(kStepCgReplace);
      CFDHIP_CHECK(hipGetLastError());
      CFDHIP_CHECK(hipMemcpyAsync(w.hostState, w.state.get(), sizeof(KState),
                                  hipMemcpyDeviceToHost, ctx.stream));
      CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    }
  }
  const int64_t inner = w.hostState->iter;
  CFDHIP_CHECK(hipEventRecord(t1, ctx.stream));
  CFDHIP_CHECK(hipEventSynchronize(t1));
  float ms = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  cfdhip_solve_result res{};
  res.status = status;
  res.fallback_used = fallback ? 1 : 0;
  res.iterations = inner + fallbackIters;
  res.restarts = replacements;
  res.initial_residual = b0;
  res.final_residual = rTrue;
  res.true_residual = rTrue;
  res.device_ms = ms;
  ctx.lastHistory = history;
  res.history_length = static_cast<int64_t>(history.size());
  return res;
}

}  // namespace cfdhip
