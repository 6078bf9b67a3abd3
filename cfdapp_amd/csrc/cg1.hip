// Single-kernel Jacobi-PCG iteration for the 2D p' operator.
//
// Chronopoulos-Gear form of preconditioned CG: the same iterates as the
// two-reduction CG of src/algebra/CG.cpp:77-142 in exact arithmetic, with
// both inner products of an iteration taken at once, so one kernel per
// iteration does all of it:
//   p = u + beta p,  s = w + beta s          (s = A p by recurrence)
//   x += alpha p,    r -= alpha s,  u = M r  (M = Jacobi)
//   w = A u,         gamma = (r, u), delta = (w, u), |r|^2
//   beta' = gamma'/gamma, alpha' = gamma' / (delta' - beta' gamma' / alpha)
// delta' - beta' gamma'/alpha equals (p', A p') of the two-reduction form,
// so the reference's |pAp| < 1e-30 breakdown test applies to it unchanged.
// The convergence test is the reference's: recurrence residual relative to
// |b - A x0|.
//
// Memory: one block per 62 x 30 (FP64) or 62 x 62 (FP32) tile (+1 halo): the tile's r, s, w, diag
// are read in one burst, u = M r goes to LDS, the stencil reads LDS. r, s, w,
// p are double-buffered (neighbouring tiles read the previous iterate), x is
// updated in place. Per iteration the kernel reads p, s, x, r, w, diag and the
// face couplings and writes p, s, x, r, w: 13 n values plus the tile halo.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "reduce.hpp"

namespace cfdhip {

namespace {

constexpr int kOut = 62;
constexpr int kWaves = 4;
constexpr int kMaxBlocks = 2048;

struct Grid2 {
  int64_t nx, ny, pin, offY;
};

// Row of A applied to u at (i, j): ascending column order from 0.
template <class T>
__device__ inline T applyRow(const Grid2& g, int64_t i, int64_t j, int64_t c, T d, T uS, T uW,
                             T uC, T uE, T uN, T fyS, T fxW, T fxE, T fyN) {
  if (c == g.pin) return T(0) + d * uC;
  T y = T(0);
  if (j > 0) y += (-fyS) * uS;
  if (i > 0) y += (-fxW) * uW;
  y += d * uC;
  if (i < g.nx - 1) y += (-fxE) * uE;
  if (j < g.ny - 1) y += (-fyN) * uN;
  return y;
}

// XCD-aware tile order: workgroups are dispatched to the 8 XCDs round-robin
// by workgroup id, so workgroup b runs on XCD b % 8. Giving each XCD one
// contiguous range of tiles (row-major) keeps the halo a tile shares with its
// left/right and upper/lower neighbours in the same L2 instead of fetching it
// once per XCD. A bijection of [0, T) for any tile count T.
constexpr int kXcds = 8;
__device__ inline int64_t xcdTile(int64_t b, int64_t T) {
  const int64_t q = T / kXcds, r = T % kXcds;
  const int64_t x = b % kXcds, k = b / kXcds;
  return x * q + (x < r ? x : r) + k;
}

// Tile geometry: a block covers 62 interior columns (+1 halo each side, one
// lane per column) and 30 (FP64) or 62 (FP32) interior rows (+1 halo each
// side); each of its 4 waves owns a quarter of the tile rows. On large grids
// FP32 tiles are twice as tall (half the halo share); at 64 rows the FP64
// kernel would drop to 3 waves per SIMD, which costs more than the halo
// saves, and on small grids the taller tiles leave CUs idle.
template <int ROWS>
struct TileShape {
  static constexpr int kTileRows = ROWS;
  static constexpr int kRowsPerWave = kTileRows / kWaves;
  static constexpr int kInteriorRows = kTileRows - 2;
};
// 32-row tiles of a grid with at least this many use 64-row ones in FP32
constexpr int64_t kTallTileMin = 2048;

// One CG iteration (start == 0) or the start step w0 = A M r0 (start == 1, r0
// in rOld). Phase 1 loads the tile in one burst, forms the new residual and
// u = M r for every tile point (into LDS) and writes p, s, r, x of the
// interior; phase 2 applies the 5-point stencil to u from LDS.
template <class T, int ROWS>
__global__ __launch_bounds__(256) void cgIterKernel(
    Grid2 g, const T* __restrict__ diag, const T* __restrict__ coeff,
    const T* __restrict__ pOld, T* __restrict__ pNew, const T* __restrict__ sOld,
    T* __restrict__ sNew, const T* __restrict__ rOld, T* __restrict__ rNew,
    const T* __restrict__ wOld, T* __restrict__ wNew, T* __restrict__ x, int64_t tilesX,
    KState* st, double* partial, int start) {
  constexpr int kTileRows = TileShape<ROWS>::kTileRows;
  constexpr int kRowsPerWave = TileShape<ROWS>::kRowsPerWave;
  constexpr int kInteriorRows = TileShape<ROWS>::kInteriorRows;
  __shared__ T uTile[kTileRows][64];
  const bool skip = st->done;
  const T alpha = static_cast<T>(st->alpha);
  const T beta = static_cast<T>(st->beta);
  const bool first = st->iter == 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t tile = xcdTile(blockIdx.x, gridDim.x);
  const int64_t tx = tile % tilesX, ty = tile / tilesX;
  const int64_t i = tx * kOut - 1 + lane;
  const int64_t jBase = ty * kInteriorRows - 1;  // tile row 0
  const bool col = i >= 0 && i < g.nx;
  const bool outCol = col && lane >= 1 && lane <= kOut;
  double acc[3] = {0.0, 0.0, 0.0};
  T dK[kRowsPerWave], uK[kRowsPerWave], fxW[kRowsPerWave];
  if (!skip) {
#pragma unroll
    for (int k = 0; k < kRowsPerWave; ++k) {
      const int tr = wave * kRowsPerWave + k;
      const int64_t j = jBase + tr;
      const bool valid = col && j >= 0 && j < g.ny;
      const bool center = valid && outCol && tr >= 1 && tr <= kInteriorRows;
      T u = T(0), d = T(1), fx = T(0);
      if (valid) {
        const int64_t c = j * g.nx + i;
        d = diag[c];
        const T r0 = rOld[c];
        T sn = T(0), rn = r0;
        if (!start) {
          const T w0 = wOld[c];
          sn = first ? w0 : w0 + (sOld[c] * beta);
          rn = r0 - (sn * alpha);
        }
        u = (T(1) / d) * rn;
        fx = coeff[j * (g.nx + 1) + i];
        if (center) {
          if (!start) {
            const T uOld = (T(1) / d) * r0;
            const T p = first ? uOld : uOld + (pOld[c] * beta);
            pNew[c] = p;
            sNew[c] = sn;
            rNew[c] = rn;
            x[c] = x[c] + (p * alpha);
          }
          acc[0] += static_cast<double>(rn) * static_cast<double>(u);
          acc[2] += static_cast<double>(rn) * static_cast<double>(rn);
        }
      }
      uTile[tr][lane] = u;
      dK[k] = d;
      uK[k] = u;
      fxW[k] = fx;
    }
  }
  __syncthreads();
  if (!skip) {
#pragma unroll
    for (int k = 0; k < kRowsPerWave; ++k) {
      const int tr = wave * kRowsPerWave + k;
      const int64_t j = jBase + tr;
      const T fxE = __shfl_down(fxW[k], 1, 64);
      const T uW = __shfl_up(uK[k], 1, 64), uE = __shfl_down(uK[k], 1, 64);
      const bool center = outCol && j >= 0 && j < g.ny && tr >= 1 && tr <= kInteriorRows;
      if (center) {
        const int64_t c = j * g.nx + i;
        const T fyS = coeff[g.offY + j * g.nx + i];
        const T fyN = coeff[g.offY + (j + 1) * g.nx + i];
        const T w = applyRow<T>(g, i, j, c, dK[k], uTile[tr - 1][lane], uW, uK[k], uE,
                                uTile[tr + 1][lane], fyS, fxW[k], fxE, fyN);
        wNew[c] = w;
        acc[1] += static_cast<double>(w) * static_cast<double>(uK[k]);
      }
    }
  }
  blockPartials<3>(acc, partial);
}

// r0 = b - A x0 and |r0| (CG.cpp:45-57), one cell per thread.
template <class T>
__global__ __launch_bounds__(256) void cgStartKernel(Grid2 g, const T* __restrict__ diag,
                                                     const T* __restrict__ coeff,
                                                     const T* __restrict__ b,
                                                     const T* __restrict__ x, T* __restrict__ r,
                                                     double* partial) {
  const int64_t n = g.nx * g.ny;
  double acc[1] = {0.0};
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < n) {
    const int64_t i = c % g.nx, j = c / g.nx;
    const T xS = j > 0 ? x[c - g.nx] : T(0);
    const T xW = i > 0 ? x[c - 1] : T(0);
    const T xE = i < g.nx - 1 ? x[c + 1] : T(0);
    const T xN = j < g.ny - 1 ? x[c + g.nx] : T(0);
    const T y = applyRow<T>(g, i, j, c, diag[c], xS, xW, x[c], xE, xN,
                            coeff[g.offY + j * g.nx + i], coeff[j * (g.nx + 1) + i],
                            coeff[j * (g.nx + 1) + i + 1], coeff[g.offY + (j + 1) * g.nx + i]);
    const T ri = b[c] - y;
    r[c] = ri;
    acc[0] += static_cast<double>(ri) * static_cast<double>(ri);
  }
  blockPartials<1>(acc, partial);
}

}  // namespace

template <class T>
cfdhip_solve_result singleKernelPcg2d(Context& ctx, const cfdhip_solver_settings& set,
                                      const T* diag, const T* coeff, const T* b, T* x,
                                      T* const work[8], Snapshots* snap) {
  const View& v = ctx.view;
  require(v.dim == 2, "the single-kernel PCG needs a 2D mesh");
  KrylovWork& w = ctx.kw;
  const int64_t tilesX = (v.nx + kOut - 1) / kOut;
  const int64_t tiles32 = tilesX * ((v.ny + TileShape<32>::kInteriorRows - 1) /
                                    TileShape<32>::kInteriorRows);
  const bool tall = sizeof(T) == 4 && tiles32 >= kTallTileMin;
  const int64_t tiles = tall ? tilesX * ((v.ny + TileShape<64>::kInteriorRows - 1) /
                                         TileShape<64>::kInteriorRows)
                             : tiles32;
  const dim3 grid(static_cast<unsigned>(tiles));
  w.partial.ensure(std::max<int64_t>(kMaxBlocks * 4, static_cast<int64_t>(grid.x) * 3 + 8));
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
  const Grid2 g{v.nx, v.ny, ctx.pin, v.offY};
  // work: r0 r1 w0 w1 p0 p1 s0 s1
  T* r[2] = {work[0], work[1]};
  T* wv[2] = {work[2], work[3]};
  T* p[2] = {work[4], work[5]};
  T* s[2] = {work[6], work[7]};
  auto finalize = [&](int step, int64_t parts) {
    launchFinalize(ctx, FinalArgs{w.partial.get(), parts, w.state.get(), w.hist.get(), nullptr,
                                  step});
  };
  // the iteration kernel; its finalize step is launched by the caller
  auto launch = [&](int cur, int start) {
    if (tall)
      hipLaunchKernelGGL((cgIterKernel<T, 64>), grid, dim3(256), 0, ctx.stream, g, diag, coeff,
                         p[cur], p[1 - cur], s[cur], s[1 - cur], r[cur], r[1 - cur], wv[cur],
                         wv[1 - cur], x, tilesX, w.state.get(), w.partial.get(), start);
    else
      hipLaunchKernelGGL((cgIterKernel<T, 32>), grid, dim3(256), 0, ctx.stream, g, diag, coeff,
                         p[cur], p[1 - cur], s[cur], s[1 - cur], r[cur], r[1 - cur], wv[cur],
                         wv[1 - cur], x, tilesX, w.state.get(), w.partial.get(), start);
  };
  hipEvent_t t0, t1;
  CFDHIP_CHECK(hipEventCreate(&t0));
  CFDHIP_CHECK(hipEventCreate(&t1));
  CFDHIP_CHECK(hipEventRecord(t0, ctx.stream));
  // start: r0 = b - A x0 into r[0], then w0 = A M r0 into w[1] (cur = 0 reads
  // r[0] as rOld and writes wNew = w[1]); the first iteration therefore runs
  // with cur = 1 reading w[1] and r0 from r[1] once r[0] and r[1] are swapped.
  const int64_t nb = (v.n + 255) / 256;
  w.partial.ensure(std::max<int64_t>(w.partial.count, nb + 8));
  hipLaunchKernelGGL((cgStartKernel<T>), dim3(static_cast<unsigned>(nb)), dim3(256), 0,
                     ctx.stream, g, diag, coeff, b, x, r[0], w.partial.get());
  finalize(kStepStartNorm, nb);
  launch(0, 1);
  finalize(kStepCg1Start, grid.x);
  CFDHIP_CHECK(hipGetLastError());
  std::swap(r[0], r[1]);  // r0 now in r[1]: the first iteration reads index 1
  const int chunk = set.check_interval > 0 ? set.check_interval : 32;
  std::vector<hipEvent_t> ev;
  if (ctx.timing) {
    ev.resize(2 * chunk);
    for (hipEvent_t& e : ev) CFDHIP_CHECK(hipEventCreate(&e));
  }
  double kMs = 0.0, kLaunches = 0.0;
  int64_t launched = 0;
  int cur = 1;
  // graph mode: one chunk (an even number of iterations, so the buffer
  // parity returns to where it started) is captured once and replayed; the
  // kernels of iterations past the device verdict do nothing
  const bool graphed = set.graph != 0 && snap == nullptr && !ctx.timing;
  const int gChunk = chunk + (chunk & 1);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  while (true) {
    int q = 0;
    if (graphed) {
      if (exec == nullptr) {
        CFDHIP_CHECK(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < gChunk; ++k) {
          launch(cur, 0);
          finalize(kStepCg1Iter, grid.x);
          cur = 1 - cur;
        }
        CFDHIP_CHECK(hipStreamEndCapture(ctx.stream, &graph));
        CFDHIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      }
      CFDHIP_CHECK(hipGraphLaunch(exec, ctx.stream));
      launched += gChunk;
      q = gChunk;
    }
    for (; !graphed && q < chunk && launched < set.max_iterations; ++q, ++launched) {
      if (ctx.timing) CFDHIP_CHECK(hipEventRecord(ev[2 * q], ctx.stream));
      launch(cur, 0);
      if (ctx.timing) CFDHIP_CHECK(hipEventRecord(ev[2 * q + 1], ctx.stream));
      finalize(kStepCg1Iter, grid.x);
      cur = 1 - cur;
      if constexpr (std::is_same_v<T, double>) {
        if (snap != nullptr && (launched + 1) % snap->stride == 0 && snap->taken < snap->cap) {
          CFDHIP_CHECK(hipMemcpyAsync(snap->buf + snap->taken * v.n, x, sizeof(T) * v.n,
                                      hipMemcpyDeviceToDevice, ctx.stream));
          ++snap->taken;
        }
      }
    }
    CFDHIP_CHECK(hipGetLastError());
    const int64_t before = w.hostState->iter;
    CFDHIP_CHECK(hipMemcpyAsync(w.hostState, w.state.get(), sizeof(KState), hipMemcpyDeviceToHost,
                                ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    if (ctx.timing) {
      const int64_t worked = std::min<int64_t>(q, w.hostState->iter - before);
      for (int e = 0; e < worked; ++e) {
        float ms = 0.f;
        CFDHIP_CHECK(hipEventElapsedTime(&ms, ev[2 * e], ev[2 * e + 1]));
        kMs += ms;
        kLaunches += 1.0;
      }
    }
    if (w.hostState->done || launched >= set.max_iterations) break;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  CFDHIP_CHECK(hipEventRecord(t1, ctx.stream));
  CFDHIP_CHECK(hipEventSynchronize(t1));
  float ms = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  if (ctx.timing) {
    ctx.spmvCount = kLaunches;
    ctx.spmvMeanUs = kLaunches > 0 ? 1000.0 * kMs / kLaunches : 0.0;
    ctx.timedIters = static_cast<double>(w.hostState->iter);
  }
  KState& S = *w.hostState;
  if (!S.done) S.status = CFDHIP_MAX_ITERATIONS;
  cfdhip_solve_result res{};
  res.status = S.status;
  res.iterations = S.iter;
  res.initial_residual = S.b0;
  res.final_residual = S.rnorm;
  res.true_residual = NAN;
  res.device_ms = ms;
  res.history_length = S.histLen;
  ctx.lastHistory.resize(S.histLen);
  if (S.histLen > 0)
    CFDHIP_CHECK(hipMemcpy(ctx.lastHistory.data(), w.hist.get(), S.histLen * sizeof(double),
                           hipMemcpyDeviceToHost));
  return res;
}

template cfdhip_solve_result singleKernelPcg2d<double>(Context&, const cfdhip_solver_settings&,
                                                       const double*, const double*,
                                                       const double*, double*,
                                                       double* const[8], Snapshots*);
template cfdhip_solve_result singleKernelPcg2d<float>(Context&, const cfdhip_solver_settings&,
                                                      const float*, const float*, const float*,
                                                      float*, float* const[8], Snapshots*);

}  // namespace cfdhip
