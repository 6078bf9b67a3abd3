// Fused Jacobi-PCG for the 2D p' operator (5-point, diagonal + face couplings).
//
// One iteration is two kernels instead of three:
//   A  p_k = z + beta p_{k-1} (z = r / diag) evaluated on the fly at the
//      5-point stencil, Ap = A p_k, p . Ap          -> alpha
//   B  x += p alpha, r -= Ap alpha, r . r, r . z   -> convergence, beta
// Kernel A marches each 64-lane wave down a strip of rows: the lanes hold one
// column each, rows j-1, j, j+1 of the new direction live in registers, the
// west/east neighbours come from adjacent lanes (the wave computes 64 columns
// and writes the inner 62), so every operator entry and every vector element
// is read once per iteration. The new direction goes to a second buffer
// (neighbouring waves still read the old one).
//
// FP64 arithmetic is the reference's (CG.cpp:77-142; row sums in ascending
// column order from 0.0); T = float gives the inner iteration of the
// mixed-precision solve, with FP64 inner products.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>

#include "context.hpp"

namespace cfdhip {

namespace {

constexpr int kWave = 64;
constexpr int kOut = kWave - 2;  // output columns per wave
constexpr int kWavesPerBlock = 4;
constexpr int kMaxBlocks = 2048;

struct Op2D {
  int64_t nx, ny, pin;
  int64_t offY;  // first y-face id
};

template <int W>
__device__ bool reduceLastBlock(double (&val)[W], double* partial, unsigned* counter,
                                double (&tot)[W]) {
  __shared__ double lds[W][kWavesPerBlock];
  __shared__ int amLast;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int w = 0; w < W; ++w) {
    double x = val[w];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) lds[w][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < W; ++w) {
      double t = 0.0;
      for (int i = 0; i < kWavesPerBlock; ++i) t += lds[w][i];
      partial[blockIdx.x * W + w] = t;
    }
    __threadfence();
    amLast = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!amLast) return false;
  __threadfence();
  for (int w = 0; w < W; ++w) {
    double x = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) x += partial[b * W + w];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    __syncthreads();
    if (lane == 0) lds[w][wid] = x;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kWavesPerBlock; ++i) t += lds[w][i];
    tot[w] = t;
  }
  if (threadIdx.x == 0) *counter = 0;
  return true;
}

// Kernel A. Tiles: tilesX = ceil(nx / 62) wave-columns, strips of H rows.
template <class T>
__global__ __launch_bounds__(256) void fusedSpmvKernel(Op2D op, const T* __restrict__ diag,
                                                       const T* __restrict__ coeff,
                                                       const T* __restrict__ r,
                                                       const T* __restrict__ pOld, T* pNew,
                                                       T* Ap, int H, int64_t tilesX,
                                                       int64_t tiles, KState* st, double* partial,
                                                       unsigned* counter) {
  const bool skip = st->done;
  const T beta = static_cast<T>(st->beta);
  const bool first = st->iter == 0;  // p_1 = z
  const int lane = threadIdx.x & 63;
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
  double dotAcc[1] = {0.0};
  if (!skip && tile < tiles) {
    const int64_t tx = tile % tilesX, ty = tile / tilesX;
    const int64_t i = tx * kOut - 1 + lane;
    const int64_t j0 = ty * H;
    const int64_t j1 = j0 + H < op.ny ? j0 + H : op.ny;
    const bool col = i >= 0 && i < op.nx;
    const bool out = col && lane >= 1 && lane <= kOut;
    const int64_t nx = op.nx, ny = op.ny;
    const int64_t sx = nx + 1;  // x-faces per row
    // direction value of (i, j), zero outside the mesh
    auto dir = [&](int64_t j) -> T {
      if (!col || j < 0 || j >= ny) return T(0);
      const int64_t c = j * nx + i;
      const T z = (T(1) / diag[c]) * r[c];
      return first ? z : z + (pOld[c] * beta);
    };
    T pPrev = dir(j0 - 1);
    T pCur = dir(j0);
    for (int64_t j = j0; j < j1; ++j) {
      const T pNext = dir(j + 1);
      const int64_t c = j * nx + i;
      // west face of (i, j) is x-face (i, j); the east face is the next lane's west face
      const T fxW = col ? coeff[j * sx + i] : T(0);
      const T fxE = __shfl_down(fxW, 1, 64);
      const T pW = __shfl_up(pCur, 1, 64);
      const T pE = __shfl_down(pCur, 1, 64);
      if (out) {
        const T d = diag[c];
        T y;
        if (c == op.pin) {
          y = T(0) + d * pCur;
        } else {
          y = T(0);
          if (j > 0) y += (-coeff[op.offY + j * nx + i]) * pPrev;
          if (i > 0) y += (-fxW) * pW;
          y += d * pCur;
          if (i < nx - 1) y += (-fxE) * pE;
          if (j < ny - 1) y += (-coeff[op.offY + (j + 1) * nx + i]) * pNext;
        }
        Ap[c] = y;
        pNew[c] = pCur;
        dotAcc[0] += static_cast<double>(pCur) * static_cast<double>(y);
      }
      pPrev = pCur;
      pCur = pNext;
    }
  }
  double tot[1];
  if (!reduceLastBlock<1>(dotAcc, partial, counter, tot) || threadIdx.x != 0 || skip) return;
  KState& s = *st;
  const double pAp = tot[0];
  s.pAp = pAp;
  if (!isfinite(pAp) || fabs(pAp) < 1e-30) {
    s.status = CFDHIP_BREAKDOWN;
    s.done = 1;
    return;
  }
  const double alpha = s.rz / pAp;
  if (!isfinite(alpha)) {
    s.status = CFDHIP_BREAKDOWN;
    s.done = 1;
    return;
  }
  s.alpha = alpha;
}

// Kernel B: x += p alpha; r = r - Ap alpha; |r|, r . z.
template <class T>
__global__ __launch_bounds__(256) void updateKernel(int64_t n, T* x, T* r,
                                                    const T* __restrict__ p,
                                                    const T* __restrict__ Ap,
                                                    const T* __restrict__ diag, KState* st,
                                                    double* hist, double* partial,
                                                    unsigned* counter) {
  const bool skip = st->done;
  double acc[2] = {0.0, 0.0};
  if (!skip) {
    const T alpha = static_cast<T>(st->alpha);
    for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      x[k] += p[k] * alpha;
      const T rn = r[k] - (Ap[k] * alpha);
      r[k] = rn;
      const T z = (T(1) / diag[k]) * rn;
      acc[0] += static_cast<double>(rn) * static_cast<double>(rn);
      acc[1] += static_cast<double>(rn) * static_cast<double>(z);
    }
  }
  double tot[2];
  if (!reduceLastBlock<2>(acc, partial, counter, tot) || threadIdx.x != 0 || skip) return;
  KState& s = *st;
  const double rnorm = sqrt(tot[0]);
  s.iter += 1;
  hist[s.histLen] = rnorm;
  s.histLen += 1;
  s.rnorm = rnorm;
  const bool conv = rnorm <= s.absTol || ((s.b0 > 0.0) && (rnorm / s.b0 <= s.relTol));
  if (!isfinite(rnorm)) {
    s.status = CFDHIP_NONFINITE_RESIDUAL;
    s.done = 1;
    return;
  }
  if (conv) {
    s.status = CFDHIP_CONVERGED;
    s.done = 1;
    return;
  }
  const double rzNew = tot[1];
  const double beta = rzNew / s.rz;
  if (!isfinite(rzNew) || !isfinite(beta)) {
    s.status = CFDHIP_BREAKDOWN;
    s.done = 1;
    return;
  }
  s.beta = beta;
  s.rz = rzNew;
  if (s.iter >= s.maxIter) {
    s.status = CFDHIP_MAX_ITERATIONS;
    s.done = 1;
  }
}

// r = b - A x (x == 0 allowed), |r|, r . z -> b0, rz  (CG.cpp:45-75)
template <class T>
__global__ __launch_bounds__(256) void startKernel(Op2D op, const T* __restrict__ diag,
                                                   const T* __restrict__ coeff,
                                                   const T* __restrict__ b,
                                                   const T* __restrict__ x, T* r, KState* st,
                                                   double* hist, double* partial,
                                                   unsigned* counter) {
  const int64_t n = op.nx * op.ny;
  double acc[2] = {0.0, 0.0};
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % op.nx, j = c / op.nx;
    T y;
    if (c == op.pin) {
      y = T(0) + diag[c] * x[c];
    } else {
      y = T(0);
      if (j > 0) y += (-coeff[op.offY + j * op.nx + i]) * x[c - op.nx];
      if (i > 0) y += (-coeff[j * (op.nx + 1) + i]) * x[c - 1];
      y += diag[c] * x[c];
      if (i < op.nx - 1) y += (-coeff[j * (op.nx + 1) + i + 1]) * x[c + 1];
      if (j < op.ny - 1) y += (-coeff[op.offY + (j + 1) * op.nx + i]) * x[c + op.nx];
    }
    const T ri = b[c] - y;
    r[c] = ri;
    const T z = (T(1) / diag[c]) * ri;
    acc[0] += static_cast<double>(ri) * static_cast<double>(ri);
    acc[1] += static_cast<double>(ri) * static_cast<double>(z);
  }
  double tot[2];
  if (!reduceLastBlock<2>(acc, partial, counter, tot) || threadIdx.x != 0) return;
  KState& s = *st;
  const double b0 = sqrt(tot[0]);
  s.b0 = b0;
  s.rnorm = b0;
  hist[s.histLen] = b0;
  s.histLen += 1;
  if (b0 == 0.0 || b0 <= s.absTol || ((b0 > 0.0) && (b0 / b0 <= s.relTol))) {
    s.status = CFDHIP_CONVERGED;
    s.done = 1;
    return;
  }
  s.rz = tot[1];
  if (!isfinite(s.rz)) {
    s.status = CFDHIP_BREAKDOWN;
    s.done = 1;
  }
}

// r64 = b - A x (FP64), r32 = (float) r64, |r64|^2 -> out[0] (iterative refinement).
__global__ __launch_bounds__(256) void refineResidualKernel(Op2D op, const double* __restrict__ diag,
                                                            const double* __restrict__ coeff,
                                                            const double* __restrict__ b,
                                                            const double* __restrict__ x,
                                                            float* r32, double* partial,
                                                            unsigned* counter, double* out) {
  const int64_t n = op.nx * op.ny;
  double acc[1] = {0.0};
  for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = c % op.nx, j = c / op.nx;
    double y;
    if (c == op.pin) {
      y = 0.0 + diag[c] * x[c];
    } else {
      y = 0.0;
      if (j > 0) y += (-coeff[op.offY + j * op.nx + i]) * x[c - op.nx];
      if (i > 0) y += (-coeff[j * (op.nx + 1) + i]) * x[c - 1];
      y += diag[c] * x[c];
      if (i < op.nx - 1) y += (-coeff[j * (op.nx + 1) + i + 1]) * x[c + 1];
      if (j < op.ny - 1) y += (-coeff[op.offY + (j + 1) * op.nx + i]) * x[c + op.nx];
    }
    const double ri = b[c] - y;
    r32[c] = static_cast<float>(ri);
    acc[0] += ri * ri;
  }
  double tot[1];
  if (!reduceLastBlock<1>(acc, partial, counter, tot) || threadIdx.x != 0) return;
  out[0] = sqrt(tot[0]);
}

__global__ void addCorrectionKernel(int64_t n, const float* __restrict__ e, double* x) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[k] += static_cast<double>(e[k]);
}

__global__ void zeroKernelD(int64_t n, double* a) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[k] = 0.0;
}

__global__ void zeroFloatKernel(int64_t n, float* a) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[k] = 0.f;
}

template <class T>
__global__ void castKernel(int64_t n, const double* __restrict__ src, T* dst) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[k] = static_cast<T>(src[k]);
}

}  // namespace

// Rows per wave strip of kernel A; CFDHIP_PCG2D_STRIP overrides it (tuning).
int pcg2dStripHeight = [] {
  const char* e = std::getenv("CFDHIP_PCG2D_STRIP");
  const int h = e ? std::atoi(e) : 16;
  return h >= 1 && h <= 4096 ? h : 16;
}();

// Jacobi-PCG on the resident 2D p' operator with the fused kernels.
// diag/coeff/b/x are T arrays on the device; x holds x0 on entry.
template <class T>
cfdhip_solve_result fusedPcg2d(Context& ctx, const cfdhip_solver_settings& set, const T* diag,
                               const T* coeff, const T* b, T* x, T* r, T* p0, T* p1, T* Ap) {
  const View& v = ctx.view;
  require(v.dim == 2, "fused 2D PCG needs a 2D mesh");
  const int64_t n = v.n;
  KrylovWork& w = ctx.kw;
  w.partial.ensure(kMaxBlocks * 4);
  if (!w.counter.get()) {
    w.counter.ensure(2);
    CFDHIP_CHECK(hipMemsetAsync(w.counter.get(), 0, 2 * sizeof(unsigned), ctx.stream));
  }
  w.state.ensure(1);
  if (!w.hostState) CFDHIP_CHECK(hipHostMalloc(&w.hostState, sizeof(KState)));
  w.hist.ensure(set.max_iterations + 2);
  KState k{};
  k.absTol = set.absolute_tolerance;
  k.relTol = set.relative_tolerance;
  k.maxIter = set.max_iterations;
  k.status = -1;
  *w.hostState = k;
  CFDHIP_CHECK(
      hipMemcpyAsync(w.state.get(), w.hostState, sizeof(KState), hipMemcpyHostToDevice, ctx.stream));
  Op2D op{v.nx, v.ny, ctx.pin, v.offY};
  const int H = pcg2dStripHeight;
  const int64_t tilesX = (v.nx + kOut - 1) / kOut;
  const int64_t tilesY = (v.ny + H - 1) / H;
  const int64_t tiles = tilesX * tilesY;
  const dim3 gA(static_cast<unsigned>((tiles + kWavesPerBlock - 1) / kWavesPerBlock));
  int64_t gb = (n + 255) / 256;
  if (gb > kMaxBlocks) gb = kMaxBlocks;
  const dim3 gB(static_cast<unsigned>(gb));
  // kernel A's grid can exceed the partial buffer: size it for the larger grid
  w.partial.ensure(std::max<int64_t>(kMaxBlocks * 4, gA.x * 2 + 8));

  hipEvent_t t0, t1;
  CFDHIP_CHECK(hipEventCreate(&t0));
  CFDHIP_CHECK(hipEventCreate(&t1));
  CFDHIP_CHECK(hipEventRecord(t0, ctx.stream));
  hipLaunchKernelGGL((startKernel<T>), gB, dim3(256), 0, ctx.stream, op, diag, coeff, b, x, r,
                     w.state.get(), w.hist.get(), w.partial.get(), w.counter.get());
  CFDHIP_CHECK(hipGetLastError());
  const int chunk = set.check_interval > 0 ? set.check_interval : 32;
  std::vector<hipEvent_t> ev;
  if (ctx.timing) {
    ev.resize(2 * chunk);
    for (hipEvent_t& e : ev) CFDHIP_CHECK(hipEventCreate(&e));
  }
  double spmvMs = 0.0, spmvLaunches = 0.0;
  int64_t launched = 0;
  T* pa = p0;
  T* pb = p1;
  while (true) {
    int q = 0;
    T* pStart = pa;
    for (; q < chunk && launched < set.max_iterations; ++q, ++launched) {
      if (ctx.timing) CFDHIP_CHECK(hipEventRecord(ev[2 * q], ctx.stream));
      hipLaunchKernelGGL((fusedSpmvKernel<T>), gA, dim3(256), 0, ctx.stream, op, diag, coeff, r,
                         pa, pb, Ap, H, tilesX, tiles, w.state.get(), w.partial.get(),
                         w.counter.get());
      if (ctx.timing) CFDHIP_CHECK(hipEventRecord(ev[2 * q + 1], ctx.stream));
      hipLaunchKernelGGL((updateKernel<T>), gB, dim3(256), 0, ctx.stream, n, x, r, pb, Ap, diag,
                         w.state.get(), w.hist.get(), w.partial.get(), w.counter.get());
      std::swap(pa, pb);
    }
    (void)pStart;
    CFDHIP_CHECK(hipGetLastError());
    const int64_t before = w.hostState->iter;
    CFDHIP_CHECK(hipMemcpyAsync(w.hostState, w.state.get(), sizeof(KState), hipMemcpyDeviceToHost,
                                ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    if (ctx.timing) {
      const int64_t worked =
          std::min<int64_t>(q, w.hostState->iter - before + (w.hostState->done ? 1 : 0));
      for (int e = 0; e < worked; ++e) {
        float ms = 0.f;
        CFDHIP_CHECK(hipEventElapsedTime(&ms, ev[2 * e], ev[2 * e + 1]));
        spmvMs += ms;
        spmvLaunches += 1.0;
      }
    }
    if (w.hostState->done || launched >= set.max_iterations) break;
  }
  CFDHIP_CHECK(hipEventRecord(t1, ctx.stream));
  CFDHIP_CHECK(hipEventSynchronize(t1));
  float ms = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  if (ctx.timing) {
    ctx.spmvCount = spmvLaunches;
    ctx.spmvMeanUs = spmvLaunches > 0 ? 1000.0 * spmvMs / spmvLaunches : 0.0;
    ctx.timedIters = static_cast<double>(w.hostState->iter);
  }
  KState& s = *w.hostState;
  if (!s.done) {
    s.status = CFDHIP_MAX_ITERATIONS;
    s.done = 1;
  }
  cfdhip_solve_result res{};
  res.status = s.status;
  res.iterations = s.iter;
  res.initial_residual = s.b0;
  res.final_residual = s.rnorm;
  res.true_residual = NAN;
  res.device_ms = ms;
  res.history_length = s.histLen;
  ctx.lastHistory.resize(s.histLen);
  if (s.histLen > 0)
    CFDHIP_CHECK(hipMemcpy(ctx.lastHistory.data(), w.hist.get(), s.histLen * sizeof(double),
                           hipMemcpyDeviceToHost));
  return res;
}

template cfdhip_solve_result fusedPcg2d<double>(Context&, const cfdhip_solver_settings&,
                                                const double*, const double*, const double*,
                                                double*, double*, double*, double*, double*);
template cfdhip_solve_result fusedPcg2d<float>(Context&, const cfdhip_solver_settings&,
                                               const float*, const float*, const float*, float*,
                                               float*, float*, float*, float*);

void castToFloat(Context& ctx, int64_t n, const double* src, float* dst);

// Mixed precision: FP32 fused Jacobi-PCG inner solves on an FP32 copy of the
// operator, FP64 true residual and correction (iterative refinement) until
// |b - A x| <= max(abs, rel |b - A x0|) in FP64 (x0 = 0).
cfdhip_solve_result mixedPcg2d(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                               double* x) {
  const View& v = ctx.view;
  const int64_t n = v.n, nf = v.nf;
  KrylovWork& w = ctx.kw;
  w.r32.ensure(n);
  w.p32.ensure(n);
  w.z32.ensure(n);
  w.Ap32.ensure(n);
  w.x32.ensure(n);
  w.q32.ensure(n);
  w.dinv32.ensure(n);  // FP32 diagonal
  w.w32.ensure(nf);    // FP32 face couplings
  castToFloat(ctx, n, ctx.f(CFDHIP_FIELD_PC_DIAG), w.dinv32.get());
  castToFloat(ctx, nf, ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), w.w32.get());
  Op2D op{v.nx, v.ny, ctx.pin, v.offY};
  int64_t gb = (n + 255) / 256;
  if (gb > kMaxBlocks) gb = kMaxBlocks;
  const dim3 g(static_cast<unsigned>(gb));
  w.partial.ensure(kMaxBlocks * 4);
  if (!w.counter.get()) {
    w.counter.ensure(2);
    CFDHIP_CHECK(hipMemsetAsync(w.counter.get(), 0, 2 * sizeof(unsigned), ctx.stream));
  }
  double* normOut = w.partial.get() + kMaxBlocks * 4 - 2;
  hipEvent_t t0, t1;
  CFDHIP_CHECK(hipEventCreate(&t0));
  CFDHIP_CHECK(hipEventCreate(&t1));
  CFDHIP_CHECK(hipEventRecord(t0, ctx.stream));
  hipLaunchKernelGGL(zeroKernelD, g, dim3(256), 0, ctx.stream, n, x);
  std::vector<double> history;
  double b0 = 0.0, rTrue = 0.0;
  int64_t total = 0, cycles = 0;
  int status = CFDHIP_MAX_ITERATIONS;
  const double innerTol = 1e-4;
  double spmvMs = 0.0, spmvLaunches = 0.0;
  while (true) {
    hipLaunchKernelGGL(refineResidualKernel, g, dim3(256), 0, ctx.stream, op,
                       ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b, x,
                       w.r32.get(), w.partial.get(), w.counter.get() + 1, normOut);
    CFDHIP_CHECK(hipGetLastError());
    CFDHIP_CHECK(hipMemcpyAsync(&rTrue, normOut, sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
    if (cycles == 0) b0 = rTrue;
    history.push_back(rTrue);
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
    cfdhip_solver_settings inner = set;
    inner.absolute_tolerance = 0.0;
    // the inner solve aims at the outer target when that is within FP32 reach
    const double need = set.relative_tolerance * b0 / rTrue;
    inner.relative_tolerance = std::max(innerTol, need);
    inner.max_iterations = set.max_iterations - total;
    hipLaunchKernelGGL(zeroFloatKernel, g, dim3(256), 0, ctx.stream, n, w.x32.get());
    const cfdhip_solve_result ir =
        fusedPcg2d<float>(ctx, inner, w.dinv32.get(), w.w32.get(), w.r32.get(), w.x32.get(),
                          w.z32.get(), w.p32.get(), w.q32.get(), w.Ap32.get());
    total += ir.iterations;
    ++cycles;
    if (ctx.timing) {
      spmvMs += ctx.spmvMeanUs * ctx.spmvCount / 1000.0;
      spmvLaunches += ctx.spmvCount;
    }
    if (ir.iterations == 0 && ir.status != CFDHIP_CONVERGED) {
      status = ir.status;
      break;
    }
    hipLaunchKernelGGL(addCorrectionKernel, g, dim3(256), 0, ctx.stream, n, w.x32.get(), x);
    CFDHIP_CHECK(hipGetLastError());
  }
  CFDHIP_CHECK(hipEventRecord(t1, ctx.stream));
  CFDHIP_CHECK(hipEventSynchronize(t1));
  float ms = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (ctx.timing) {
    ctx.spmvCount = spmvLaunches;
    ctx.spmvMeanUs = spmvLaunches > 0 ? 1000.0 * spmvMs / spmvLaunches : 0.0;
    ctx.timedIters = static_cast<double>(total);
  }
  cfdhip_solve_result res{};
  res.status = status;
  res.iterations = total;
  res.restarts = cycles;
  res.initial_residual = b0;
  res.final_residual = rTrue;
  res.true_residual = rTrue;
  res.device_ms = ms;
  ctx.lastHistory = history;
  res.history_length = static_cast<int64_t>(history.size());
  return res;
}

void castToFloat(Context& ctx, int64_t n, const double* src, float* dst) {
  int64_t gb = (n + 255) / 256;
  if (gb > kMaxBlocks) gb = kMaxBlocks;
  hipLaunchKernelGGL((castKernel<float>), dim3(static_cast<unsigned>(gb)), dim3(256), 0,
                     ctx.stream, n, src, dst);
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace cfdhip
