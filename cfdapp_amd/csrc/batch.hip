// Batched p' solves: many independent 2D cases of uneven size advanced
// together, one launch per step for the whole batch.
//
// Each case runs Jacobi-PCG in the single-reduction (Chronopoulos-Gear) form
// of cg1.hip on its own p' operator (PressureCorrectionEquation.cpp:181-339:
// diagonal + face couplings, optional pinned row), inside FP64 iterative
// refinement: the inner iteration works in T (float: mixed precision, double:
// plain FP64), the true residual b - A x, the correction x += d and the
// convergence test |b - A x| <= max(abs, rel |b|) are FP64. The batch has no
// counterpart in the reference, which solves one case at a time
// (src/algebra/CG.cpp:14-148).
//
// Per case a small state machine lives on the device:
//   RESID   r = b - A x (FP64), r_T = (T) r, |r|    -> outer test, inner target
//   START   u = M r_T, w = A u, (r,u), (w,u), |r|^2  -> alpha
//   ITER    one PCG iteration (cg1.hip)             -> inner test, alpha, beta
//   CORRECT x += x_T, x_T = 0                        -> RESID
// Every block of a step belongs to one case and does that case's phase; a
// one-block-per-case finalize kernel then sums the case's block partials in
// tile order and moves its state. Cases that are done drop out of the block
// table the host rebuilds between chunks of steps, so the grid shrinks as
// cases converge. A chunk (steps x {step kernel, finalize}) can be captured
// once into a HIP graph and replayed while the set of active cases is
// unchanged. The tile height (rows of a 64-lane-wide tile) is picked by timing
// the START step, which is idempotent, for each candidate on the batch itself.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>

#include "batch.hpp"

using namespace cfdhip;

struct cfdhip_ctx {
  Context c;
};

namespace cfdhip {
namespace {

enum Phase : int32_t { kResid = 0, kStart = 1, kIter = 2, kCorrect = 3 };

constexpr int kLanes = 64;            // tile width incl. one halo column each side
constexpr int kOutCols = kLanes - 2;  // interior columns of a tile
constexpr int kWaves = 4;

template <class T>
struct CaseT {
  int64_t nx, ny, pin, offY;
  int64_t tilesX, tile0, tiles;
  const double *diag, *coeff, *b;
  double* x;
  const T *diagT, *coeffT;
  T *r[2], *w[2], *p[2], *s[2], *xT;
};

// Row of A applied at (i, j), ascending column order from 0.
template <class T>
__device__ inline T rowApply(int64_t nx, int64_t ny, int64_t pin, int64_t i, int64_t j,
                             int64_t c, T d, T uS, T uW, T uC, T uE, T uN, T fyS, T fxW, T fxE,
                             T fyN) {
  if (c == pin) return T(0) + d * uC;
  T y = T(0);
  if (j > 0) y += (-fyS) * uS;
  if (i > 0) y += (-fxW) * uW;
  y += d * uC;
  if (i < nx - 1) y += (-fxE) * uE;
  if (j < ny - 1) y += (-fyN) * uN;
  return y;
}

// Block partials of W sums into slot `slot` (fixed order within the block).
template <int W>
__device__ inline void partialsAt(double (&val)[W], double* partial, int64_t slot) {
  __shared__ double lds[W][kWaves];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < W; ++q) {
    double x = val[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) lds[q][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < W) {
    double t = 0.0;
    for (int i = 0; i < kWaves; ++i) t += lds[threadIdx.x][i];
    partial[slot * W + threadIdx.x] = t;
  }
}

// One step of every active case. blockCase / blockTile map a block to its
// case and to its tile within the case; ROWS is the tile height (incl. halo).
template <class T, int ROWS>
__global__ __launch_bounds__(256) void batchStepKernel(const CaseT<T>* __restrict__ cases,
                                                       const CaseState* __restrict__ states,
                                                       const int32_t* __restrict__ blockCase,
                                                       const int32_t* __restrict__ blockTile,
                                                       double* partial) {
  constexpr int kRowsPerWave = ROWS / kWaves;
  constexpr int kInner = ROWS - 2;
  __shared__ T uTile[ROWS][kLanes];
  const int32_t cid = blockCase[blockIdx.x];
  const int64_t tile = blockTile[blockIdx.x];
  const CaseT<T>& cs = cases[cid];
  const CaseState& st = states[cid];
  const int phase = st.phase;
  const int cur = st.cur;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t nx = cs.nx, ny = cs.ny;
  const int64_t tx = tile % cs.tilesX, ty = tile / cs.tilesX;
  const int64_t i = tx * kOutCols - 1 + lane;
  const int64_t jBase = ty * kInner - 1;
  const bool col = i >= 0 && i < nx;
  const bool outCol = col && lane >= 1 && lane <= kOutCols;
  const int64_t slot = cs.tile0 + tile;
  double acc[3] = {0.0, 0.0, 0.0};
  if (st.done) return;  // uniform per block: every thread reads the same state

  if (phase == kCorrect || phase == kResid) {
    // pointwise over the tile's interior cells (global loads; rare steps)
#pragma unroll
    for (int k = 0; k < kRowsPerWave; ++k) {
      const int tr = wave * kRowsPerWave + k;
      const int64_t j = jBase + tr;
      if (!(outCol && tr >= 1 && tr <= kInner && j >= 0 && j < ny)) continue;
      const int64_t c = j * nx + i;
      if (phase == kCorrect) {
        cs.x[c] = cs.x[c] + static_cast<double>(cs.xT[c]);
        cs.xT[c] = T(0);
      } else {
        const double* x = cs.x;
        const double* f = cs.coeff;
        const double xS = j > 0 ? x[c - nx] : 0.0;
        const double xW = i > 0 ? x[c - 1] : 0.0;
        const double xE = i < nx - 1 ? x[c + 1] : 0.0;
        const double xN = j < ny - 1 ? x[c + nx] : 0.0;
        const double y = rowApply<double>(nx, ny, cs.pin, i, j, c, cs.diag[c], xS, xW, x[c], xE,
                                          xN, f[cs.offY + j * nx + i], f[j * (nx + 1) + i],
                                          f[j * (nx + 1) + i + 1], f[cs.offY + (j + 1) * nx + i]);
        const double ri = cs.b[c] - y;
        cs.r[cur][c] = static_cast<T>(ri);
        acc[0] += ri * ri;
      }
    }
    partialsAt<3>(acc, partial, slot);
    return;
  }

  // START (w = A M r) or one PCG iteration, as cgIterKernel in cg1.hip.
  const bool start = phase == kStart;
  const T alpha = static_cast<T>(st.alpha);
  const T beta = static_cast<T>(st.beta);
  const bool first = st.iter == 0;
  const T* rOld = cs.r[cur];
  const T* wOld = cs.w[cur];
  const T* pOld = cs.p[cur];
  const T* sOld = cs.s[cur];
  T* rNew = cs.r[1 - cur];
  T* pNew = cs.p[1 - cur];
  T* sNew = cs.s[1 - cur];
  T* wNew = start ? cs.w[cur] : cs.w[1 - cur];
  T dK[kRowsPerWave], uK[kRowsPerWave], fxW[kRowsPerWave];
#pragma unroll
  for (int k = 0; k < kRowsPerWave; ++k) {
    const int tr = wave * kRowsPerWave + k;
    const int64_t j = jBase + tr;
    const bool valid = col && j >= 0 && j < ny;
    const bool center = valid && outCol && tr >= 1 && tr <= kInner;
    T u = T(0), d = T(1), fx = T(0);
    if (valid) {
      const int64_t c = j * nx + i;
      d = cs.diagT[c];
      const T r0 = rOld[c];
      T sn = T(0), rn = r0;
      if (!start) {
        const T w0 = wOld[c];
        sn = first ? w0 : w0 + (sOld[c] * beta);
        rn = r0 - (sn * alpha);
      }
      u = (T(1) / d) * rn;
      fx = cs.coeffT[j * (nx + 1) + i];
      if (center) {
        if (!start) {
          const T uOld = (T(1) / d) * r0;
          const T p = first ? uOld : uOld + (pOld[c] * beta);
          pNew[c] = p;
          sNew[c] = sn;
          rNew[c] = rn;
          cs.xT[c] = cs.xT[c] + (p * alpha);
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
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRowsPerWave; ++k) {
    const int tr = wave * kRowsPerWave + k;
    const int64_t j = jBase + tr;
    const T fxE = __shfl_down(fxW[k], 1, 64);
    const T uW = __shfl_up(uK[k], 1, 64), uE = __shfl_down(uK[k], 1, 64);
    const bool center = outCol && j >= 0 && j < ny && tr >= 1 && tr <= kInner;
    if (center) {
      const int64_t c = j * nx + i;
      const T fyS = cs.coeffT[cs.offY + j * nx + i];
      const T fyN = cs.coeffT[cs.offY + (j + 1) * nx + i];
      const T w = rowApply<T>(nx, ny, cs.pin, i, j, c, dK[k], uTile[tr - 1][lane], uW, uK[k], uE,
                              uTile[tr + 1][lane], fyS, fxW[k], fxE, fyN);
      wNew[c] = w;
      acc[1] += static_cast<double>(w) * static_cast<double>(uK[k]);
    }
  }
  partialsAt<3>(acc, partial, slot);
}

__device__ inline void finishCase(CaseState& s, int status) {
  s.status = status;
  s.done = 1;
  s.doneStep = s.steps;
}

// One block per active case: sums the case's partials in tile order and runs
// the scalar step of its phase (thread 0).
__global__ __launch_bounds__(256) void batchFinalizeKernel(const int32_t* __restrict__ active,
                                                           const int64_t* __restrict__ tile0,
                                                           const int64_t* __restrict__ tiles,
                                                           CaseState* states,
                                                           const double* __restrict__ partial) {
  __shared__ double lds[3][4];
  const int32_t cid = active[blockIdx.x];
  CaseState& s = states[cid];
  if (s.done) return;
  const int64_t t0 = tile0[cid], nt = tiles[cid];
  double x[3] = {0.0, 0.0, 0.0};
  for (int64_t b = threadIdx.x; b < nt; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < 3; ++q) x[q] += partial[(t0 + b) * 3 + q];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < 3; ++q) {
    double y = x[q];
    for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o, 64);
    if (lane == 0) lds[q][wid] = y;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[3];
  for (int q = 0; q < 3; ++q) t[q] = lds[q][0] + lds[q][1] + lds[q][2] + lds[q][3];
  s.steps += 1;
  switch (s.phase) {
    case kCorrect:
      s.phase = kResid;
      return;
    case kResid: {
      const double r = sqrt(t[0]);
      if (s.cycles == 0) s.b0 = r;
      s.rTrue = r;
      if (!isfinite(r)) {
        finishCase(s, CFDHIP_NONFINITE_RESIDUAL);
      } else if (r == 0.0 || r <= s.absTol || (s.b0 > 0.0 && r / s.b0 <= s.relTol)) {
        finishCase(s, CFDHIP_CONVERGED);
      } else if (s.total >= s.maxIter) {
        finishCase(s, CFDHIP_MAX_ITERATIONS);
      } else {
        s.iter = 0;
        s.phase = kStart;
      }
      return;
    }
    case kStart: {
      const double gamma = t[0], delta = t[1];
      s.ib0 = sqrt(t[2]);
      // the inner solve aims at the outer target when that is within reach
      const double need = s.b0 > 0.0 ? s.relTol * s.b0 / s.rTrue : 0.0;
      s.innerTol = fmax(s.innerFloor, need);
      s.rz = gamma;
      const double a = gamma / delta;
      if (!isfinite(gamma) || !isfinite(delta) || fabs(delta) < 1e-30 || !isfinite(a)) {
        finishCase(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.alpha = a;
      s.beta = 0.0;
      s.phase = kIter;
      return;
    }
    case kIter: {
      const double gamma = t[0], delta = t[1], rn = sqrt(t[2]);
      s.iter += 1;
      s.total += 1;
      s.cur = 1 - s.cur;
      s.rnorm = rn;
      if (!isfinite(rn)) {
        finishCase(s, CFDHIP_NONFINITE_RESIDUAL);
        return;
      }
      if (rn <= s.innerTol * s.ib0 || s.total >= s.maxIter) {
        s.cycles += 1;
        s.phase = kCorrect;
        return;
      }
      const double beta = gamma / s.rz;
      const double pAp = delta - beta * gamma / s.alpha;
      const double alpha = gamma / pAp;
      if (!isfinite(gamma) || !isfinite(beta) || !isfinite(pAp) || fabs(pAp) < 1e-30 ||
          !isfinite(alpha)) {
        // keep what the inner solve gained and restart from the true residual
        s.cycles += 1;
        s.phase = kCorrect;
        if (s.iter <= 1) finishCase(s, CFDHIP_BREAKDOWN);
        return;
      }
      s.beta = beta;
      s.rz = gamma;
      s.alpha = alpha;
      return;
    }
    default:
      return;
  }
}

template <class T>
__global__ void fillKernel(int64_t n, T* a) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[k] = T(0);
}

template <class T>
__global__ void castKernelT(int64_t n, const double* __restrict__ src, T* dst) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[k] = static_cast<T>(src[k]);
}

constexpr int kRowChoices[3] = {16, 32, 64};

}  // namespace
}  // namespace cfdhip

namespace cfdhip {
namespace {

int64_t faceCount2d(int64_t nx, int64_t ny) { return (nx + 1) * ny + nx * (ny + 1); }

inline dim3 strideGrid(int64_t n) {
  int64_t b = (n + 255) / 256;
  return dim3(static_cast<unsigned>(std::clamp<int64_t>(b, 1, 2048)));
}

struct Layout {
  std::vector<int64_t> tilesX, tile0, tiles;
  int64_t totalTiles = 0;
};

Layout tileLayout(const cfdhip_batch& B, int rows) {
  Layout L;
  const int inner = rows - 2;
  for (const auto& s : B.sys) {
    const int64_t tx = (s->nx + kOutCols - 1) / kOutCols;
    const int64_t ty = (s->ny + inner - 1) / inner;
    L.tilesX.push_back(tx);
    L.tile0.push_back(L.totalTiles);
    L.tiles.push_back(tx * ty);
    L.totalTiles += tx * ty;
  }
  return L;
}

template <class T>
void launchStep(cfdhip_batch& B, int rows, const CaseT<T>* cases, unsigned blocks) {
  switch (rows) {
    case 16:
      hipLaunchKernelGGL((batchStepKernel<T, 16>), dim3(blocks), dim3(256), 0, B.stream, cases,
                         B.states.get(), B.blockCase.get(), B.blockTile.get(), B.partial.get());
      break;
    case 32:
      hipLaunchKernelGGL((batchStepKernel<T, 32>), dim3(blocks), dim3(256), 0, B.stream, cases,
                         B.states.get(), B.blockCase.get(), B.blockTile.get(), B.partial.get());
      break;
    default:
      hipLaunchKernelGGL((batchStepKernel<T, 64>), dim3(blocks), dim3(256), 0, B.stream, cases,
                         B.states.get(), B.blockCase.get(), B.blockTile.get(), B.partial.get());
      break;
  }
}

// Uploads the block table (block -> case, tile) and the active-case list for
// the cases not yet done; returns (blocks, active cases).
std::pair<int64_t, int64_t> uploadTable(cfdhip_batch& B, const Layout& L,
                                        const std::vector<char>& done) {
  int64_t blocks = 0, act = 0;
  const int64_t nc = static_cast<int64_t>(B.sys.size());
  for (int64_t c = 0; c < nc; ++c)
    if (!done[c]) {
      blocks += L.tiles[c];
      ++act;
    }
  const int64_t need = 2 * blocks + act;
  if (need > B.hostTableCap) {
    if (B.hostTable) CFDHIP_CHECK(hipHostFree(B.hostTable));
    B.hostTable = nullptr;
    CFDHIP_CHECK(hipHostMalloc(&B.hostTable, sizeof(int32_t) * std::max<int64_t>(need, 1)));
    B.hostTableCap = need;
  }
  // the previous upload may still be in flight from this staging buffer
  CFDHIP_CHECK(hipStreamSynchronize(B.stream));
  int32_t* bc = B.hostTable;
  int32_t* bt = bc + blocks;
  int32_t* ac = bt + blocks;
  int64_t k = 0, a = 0;
  for (int64_t c = 0; c < nc; ++c) {
    if (done[c]) continue;
    for (int64_t t = 0; t < L.tiles[c]; ++t, ++k) {
      bc[k] = static_cast<int32_t>(c);
      bt[k] = static_cast<int32_t>(t);
    }
    ac[a++] = static_cast<int32_t>(c);
  }
  if (blocks > 0) {
    CFDHIP_CHECK(hipMemcpyAsync(B.blockCase.get(), bc, sizeof(int32_t) * blocks,
                                hipMemcpyHostToDevice, B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(B.blockTile.get(), bt, sizeof(int32_t) * blocks,
                                hipMemcpyHostToDevice, B.stream));
  }
  if (act > 0)
    CFDHIP_CHECK(hipMemcpyAsync(B.active.get(), ac, sizeof(int32_t) * act,
                                hipMemcpyHostToDevice, B.stream));
  return {blocks, act};
}

template <class T>
void solveBatch(cfdhip_batch& B, const cfdhip_solver_settings& set, int rowsReq, int useGraph,
                cfdhip_solve_result* results, double* latencyMs, double* stats) {
  const int64_t nc = static_cast<int64_t>(B.sys.size());
  require(nc > 0, "the batch holds no systems");
  require(nc < (int64_t{1} << 31), "too many systems in the batch");
  require(rowsReq == 0 || rowsReq == 16 || rowsReq == 32 || rowsReq == 64,
          "tile rows must be 0 (autotune), 16, 32 or 64");
  require(set.max_iterations > 0, "max_iterations must be positive");
  constexpr bool kF32 = std::is_same_v<T, float>;
  // arenas: x (FP64, per case) and the T work vectors r0 r1 w0 w1 p0 p1 s0 s1 xT
  // (+ diag, coeff copies for float)
  int64_t used64 = 0, usedT = 0;
  std::vector<int64_t>& offX = B.xOff;
  offX.assign(nc, 0);
  std::vector<int64_t> offT(nc);
  auto pad = [](int64_t n) { return (n + 63) / 64 * 64; };
  for (int64_t c = 0; c < nc; ++c) {
    const auto& s = *B.sys[c];
    const int64_t n = s.nx * s.ny, nf = faceCount2d(s.nx, s.ny);
    offX[c] = used64;
    used64 += pad(n);
    offT[c] = usedT;
    usedT += 9 * pad(n) + (kF32 ? pad(n) + pad(nf) : 0);
  }
  require(used64 + (kF32 ? 0 : usedT) < (int64_t{1} << 40), "the batch is too large");
  B.arena64.ensure(used64 + (kF32 ? 0 : usedT));
  if (kF32) B.arena32.ensure(usedT);
  T* baseT = kF32 ? reinterpret_cast<T*>(B.arena32.get())
                  : reinterpret_cast<T*>(B.arena64.get() + used64);
  std::vector<CaseT<T>> hc(nc);
  for (int64_t c = 0; c < nc; ++c) {
    auto& s = *B.sys[c];
    const int64_t n = s.nx * s.ny, nf = faceCount2d(s.nx, s.ny);
    CaseT<T>& h = hc[c];
    h.nx = s.nx;
    h.ny = s.ny;
    h.pin = s.pin;
    h.offY = (s.nx + 1) * s.ny;
    h.diag = s.diag.get();
    h.coeff = s.coeff.get();
    h.b = s.b.get();
    h.x = B.arena64.get() + offX[c];
    T* q = baseT + offT[c];
    for (int k = 0; k < 2; ++k) {
      h.r[k] = q + (0 + k) * pad(n);
      h.w[k] = q + (2 + k) * pad(n);
      h.p[k] = q + (4 + k) * pad(n);
      h.s[k] = q + (6 + k) * pad(n);
    }
    h.xT = q + 8 * pad(n);
    if constexpr (kF32) {
      T* dT = q + 9 * pad(n);
      T* cT = dT + pad(n);
      hipLaunchKernelGGL((castKernelT<T>), strideGrid(n), dim3(256), 0, B.stream, n,
                         s.diag.get(), dT);
      hipLaunchKernelGGL((castKernelT<T>), strideGrid(nf), dim3(256), 0, B.stream, nf,
                         s.coeff.get(), cT);
      h.diagT = dT;
      h.coeffT = cT;
    } else {
      h.diagT = reinterpret_cast<const T*>(s.diag.get());
      h.coeffT = reinterpret_cast<const T*>(s.coeff.get());
    }
  }
  // x = 0, x_T = 0 (fill kernels; the rest is written before it is read)
  hipLaunchKernelGGL(fillKernel<double>, strideGrid(used64), dim3(256), 0, B.stream, used64,
                     B.arena64.get());
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t n = B.sys[c]->nx * B.sys[c]->ny;
    hipLaunchKernelGGL(fillKernel<T>, strideGrid(n), dim3(256), 0, B.stream, n, hc[c].xT);
  }
  CFDHIP_CHECK(hipGetLastError());

  // states
  // FP32 reduction per refinement cycle (the settings' inner tolerance)
  const double innerTol = set.inner_tolerance > 0.0 ? set.inner_tolerance : 1e-5;
  B.states.ensure(nc);
  if (B.hostStates) CFDHIP_CHECK(hipHostFree(B.hostStates));
  B.hostStates = nullptr;
  CFDHIP_CHECK(hipHostMalloc(&B.hostStates, sizeof(CaseState) * nc));
  for (int64_t c = 0; c < nc; ++c) {
    CaseState z{};
    z.absTol = set.absolute_tolerance;
    z.relTol = set.relative_tolerance;
    z.innerFloor = kF32 ? innerTol : 0.0;  // FP64: the inner solve aims at the outer target
    z.maxIter = set.max_iterations;
    z.status = -1;
    z.phase = kResid;
    B.hostStates[c] = z;
  }
  CFDHIP_CHECK(hipMemcpyAsync(B.states.get(), B.hostStates, sizeof(CaseState) * nc,
                              hipMemcpyHostToDevice, B.stream));

  // tile layouts for the candidates; partial slots sized for the largest
  Layout lay[3];
  int64_t maxTiles = 0;
  for (int q = 0; q < 3; ++q) {
    lay[q] = tileLayout(B, kRowChoices[q]);
    maxTiles = std::max(maxTiles, lay[q].totalTiles);
  }
  require(maxTiles < (int64_t{1} << 31), "too many tiles in the batch");
  B.partial.ensure(3 * maxTiles + 8);
  B.blockCase.ensure(maxTiles);
  B.blockTile.ensure(maxTiles);
  B.active.ensure(nc);
  B.tile0.ensure(nc);
  B.tiles.ensure(nc);
  B.caseBuf.ensure(static_cast<int64_t>(sizeof(CaseT<T>)) * nc);
  CaseT<T>* dCases = reinterpret_cast<CaseT<T>*>(B.caseBuf.get());

  auto setLayout = [&](int q) {
    for (int64_t c = 0; c < nc; ++c) {
      hc[c].tilesX = lay[q].tilesX[c];
      hc[c].tile0 = lay[q].tile0[c];
      hc[c].tiles = lay[q].tiles[c];
    }
    CFDHIP_CHECK(hipMemcpyAsync(dCases, hc.data(), sizeof(CaseT<T>) * nc, hipMemcpyHostToDevice,
                                B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(B.tile0.get(), lay[q].tile0.data(), sizeof(int64_t) * nc,
                                hipMemcpyHostToDevice, B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(B.tiles.get(), lay[q].tiles.data(), sizeof(int64_t) * nc,
                                hipMemcpyHostToDevice, B.stream));
    CFDHIP_CHECK(hipStreamSynchronize(B.stream));  // hc is host pageable memory
  };
  std::vector<char> done(nc, 0);

  hipEvent_t e0, e1;
  CFDHIP_CHECK(hipEventCreate(&e0));
  CFDHIP_CHECK(hipEventCreate(&e1));
  // the first step (RESID from x = 0) with any layout, then the START step is
  // timed for every candidate tile height (it is idempotent: it reads r, writes
  // w and the partials, and the state moves only in the finalize)
  int q = rowsReq == 16 ? 0 : rowsReq == 64 ? 2 : 1;
  setLayout(q);
  CFDHIP_CHECK(hipEventRecord(e0, B.stream));
  auto tb = uploadTable(B, lay[q], done);
  launchStep<T>(B, kRowChoices[q], dCases, static_cast<unsigned>(tb.first));
  hipLaunchKernelGGL(batchFinalizeKernel, dim3(static_cast<unsigned>(tb.second)), dim3(256), 0,
                     B.stream, B.active.get(), B.tile0.get(), B.tiles.get(), B.states.get(),
                     B.partial.get());
  CFDHIP_CHECK(hipGetLastError());
  double tuneUs[3] = {0.0, 0.0, 0.0};
  if (rowsReq == 0) {
    hipEvent_t a, b;
    CFDHIP_CHECK(hipEventCreate(&a));
    CFDHIP_CHECK(hipEventCreate(&b));
    float best = 0.f;
    int bestQ = 1;
    for (int cand = 0; cand < 3; ++cand) {
      setLayout(cand);
      auto t2 = uploadTable(B, lay[cand], done);
      launchStep<T>(B, kRowChoices[cand], dCases, static_cast<unsigned>(t2.first));  // warm
      float ms = 0.f;
      for (int rep = 0; rep < 3; ++rep) {
        CFDHIP_CHECK(hipEventRecord(a, B.stream));
        launchStep<T>(B, kRowChoices[cand], dCases, static_cast<unsigned>(t2.first));
        CFDHIP_CHECK(hipEventRecord(b, B.stream));
        CFDHIP_CHECK(hipEventSynchronize(b));
        float m = 0.f;
        CFDHIP_CHECK(hipEventElapsedTime(&m, a, b));
        ms += m;
      }
      tuneUs[cand] = 1000.0 * ms / 3.0;
      if (cand == 0 || ms < best) {
        best = ms;
        bestQ = cand;
      }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    q = bestQ;
    setLayout(q);
  }
  const int rows = kRowChoices[q];
  B.lastRows = rows;

  // chunks of steps; the host reads the case states between chunks
  const int chunk = set.check_interval > 0 ? set.check_interval : 32;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int captures = 0;
  int64_t steps = 1;
  std::vector<double> chunkEndMs;    // device time at the end of each chunk
  std::vector<int64_t> chunkEndStep; // global step count at the end of each chunk
  chunkEndMs.push_back(0.0);
  chunkEndStep.push_back(0);
  bool tableDirty = true;
  int64_t blocks = 0, act = 0;
  hipEvent_t ce;
  CFDHIP_CHECK(hipEventCreate(&ce));
  const int64_t stepCap = 4 * set.max_iterations + 64;
  while (true) {
    CFDHIP_CHECK(hipMemcpyAsync(B.hostStates, B.states.get(), sizeof(CaseState) * nc,
                                hipMemcpyDeviceToHost, B.stream));
    CFDHIP_CHECK(hipEventRecord(ce, B.stream));
    CFDHIP_CHECK(hipEventSynchronize(ce));
    float at = 0.f;
    CFDHIP_CHECK(hipEventElapsedTime(&at, e0, ce));
    chunkEndMs.push_back(at);
    chunkEndStep.push_back(steps);
    bool any = false;
    for (int64_t c = 0; c < nc; ++c) {
      if (!done[c] && B.hostStates[c].done) {
        done[c] = 1;
        tableDirty = true;
      }
      any = any || !done[c];
    }
    if (!any || steps >= stepCap) break;
    if (tableDirty) {
      auto t = uploadTable(B, lay[q], done);
      blocks = t.first;
      act = t.second;
      tableDirty = false;
      if (exec) {
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
      }
      if (graph) {
        (void)hipGraphDestroy(graph);
        graph = nullptr;
      }
    }
    auto enqueue = [&]() {
      for (int k = 0; k < chunk; ++k) {
        launchStep<T>(B, rows, dCases, static_cast<unsigned>(blocks));
        hipLaunchKernelGGL(batchFinalizeKernel, dim3(static_cast<unsigned>(act)), dim3(256), 0,
                           B.stream, B.active.get(), B.tile0.get(), B.tiles.get(),
                           B.states.get(), B.partial.get());
      }
    };
    if (useGraph) {
      if (!exec) {
        CFDHIP_CHECK(hipStreamSynchronize(B.stream));  // table upload done before capture
        CFDHIP_CHECK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        CFDHIP_CHECK(hipStreamEndCapture(B.stream, &graph));
        CFDHIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        ++captures;
      }
      CFDHIP_CHECK(hipGraphLaunch(exec, B.stream));
    } else {
      enqueue();
    }
    CFDHIP_CHECK(hipGetLastError());
    steps += chunk;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  CFDHIP_CHECK(hipEventRecord(e1, B.stream));
  CFDHIP_CHECK(hipEventSynchronize(e1));
  float total = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&total, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(ce);

  // per-case latency: the device time at the case's last step, interpolated
  // between the host checks around it (the first check follows the first
  // step and the tile tuning)
  int64_t cellIters = 0;
  for (int64_t c = 0; c < nc; ++c) {
    const CaseState& s = B.hostStates[c];
    cfdhip_solve_result& r = results[c];
    r = cfdhip_solve_result{};
    r.status = s.done ? s.status : CFDHIP_MAX_ITERATIONS;
    r.iterations = s.total;
    r.restarts = s.cycles;
    r.initial_residual = s.b0;
    r.final_residual = s.rTrue;
    r.true_residual = s.rTrue;
    double lat = total;
    if (s.done) {
      const int64_t st = s.doneStep;
      size_t k = 1;
      while (k < chunkEndStep.size() && chunkEndStep[k] < st) ++k;
      if (k >= chunkEndStep.size()) {
        lat = total;
      } else {
        const double s0 = static_cast<double>(chunkEndStep[k - 1]);
        const double s1 = static_cast<double>(chunkEndStep[k]);
        const double f = s1 > s0 ? (static_cast<double>(st) - s0) / (s1 - s0) : 1.0;
        lat = chunkEndMs[k - 1] + std::clamp(f, 0.0, 1.0) * (chunkEndMs[k] - chunkEndMs[k - 1]);
      }
    }
    r.device_ms = lat;
    if (latencyMs) latencyMs[c] = lat;
    cellIters += s.total * B.sys[c]->nx * B.sys[c]->ny;
  }
  if (stats) {
    stats[0] = total;
    stats[1] = static_cast<double>(steps);
    stats[2] = static_cast<double>(cellIters);
    stats[3] = rows;
    stats[4] = captures;
    stats[5] = tuneUs[0];
    stats[6] = tuneUs[1];
    stats[7] = tuneUs[2];
  }
}

template <class F>
int guardedBatch(cfdhip_batch* b, F&& fn) {
  if (b == nullptr) return CFDHIP_ERR_INVALID;
  try {
    CFDHIP_CHECK(hipSetDevice(b->device));
    fn(*b);
    b->lastError.clear();
    return CFDHIP_OK;
  } catch (const Error& e) {
    b->lastError = e.what();
    return e.code;
  } catch (const std::exception& e) {
    b->lastError = e.what();
    return CFDHIP_ERR_RUNTIME;
  }
}

// Sides whose every cell (but a pinned one) has a diagonal larger than the sum
// of its internal-face couplings: a FixedValue pressure patch there adds its
// face coupling to the diagonal, a zero-gradient wall adds nothing (bits 1
// left, 2 right, 4 bottom, 8 top, the 2D patch order).
int inferFixedValueSides(int64_t nx, int64_t ny, int64_t pin, const double* diag,
                         const double* coeff) {
  const int64_t offY = (nx + 1) * ny;
  auto excess = [&](int64_t i, int64_t j) {
    const int64_t c = j * nx + i;
    double s = 0.0;
    if (i > 0) s += coeff[j * (nx + 1) + i];
    if (i < nx - 1) s += coeff[j * (nx + 1) + i + 1];
    if (j > 0) s += coeff[offY + j * nx + i];
    if (j < ny - 1) s += coeff[offY + (j + 1) * nx + i];
    return c == pin || diag[c] - s > 1e-12 * std::fabs(diag[c]);
  };
  int sides = 0;
  bool all[4] = {true, true, true, true};
  for (int64_t j = 0; j < ny; ++j) {
    all[0] = all[0] && excess(0, j);
    all[1] = all[1] && excess(nx - 1, j);
  }
  for (int64_t i = 0; i < nx; ++i) {
    all[2] = all[2] && excess(i, 0);
    all[3] = all[3] && excess(i, ny - 1);
  }
  for (int q = 0; q < 4; ++q)
    if (all[q]) sides |= 1 << q;
  return sides;
}

}  // namespace

void solveBatchJacobi(cfdhip_batch& B, const cfdhip_solver_settings& set, int rows, int useGraph,
                      cfdhip_solve_result* results, double* latencyMs, double* stats) {
  if (set.precision == CFDHIP_MIXED)
    solveBatch<float>(B, set, rows, useGraph, results, latencyMs, stats);
  else
    solveBatch<double>(B, set, rows, useGraph, results, latencyMs, stats);
}

}  // namespace cfdhip

extern "C" {

int cfdhip_batch_create(int device, cfdhip_batch** out) {
  if (out == nullptr) return CFDHIP_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return CFDHIP_ERR_NO_DEVICE;
  auto* b = new (std::nothrow) cfdhip_batch();
  if (b == nullptr) return CFDHIP_ERR_RUNTIME;
  b->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
    delete b;
    return CFDHIP_ERR_RUNTIME;
  }
  *out = b;
  return CFDHIP_OK;
}

void cfdhip_batch_destroy(cfdhip_batch* b) {
  if (b == nullptr) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  delete b;
}

const char* cfdhip_batch_last_error(const cfdhip_batch* b) {
  return b == nullptr ? "null batch" : b->lastError.c_str();
}

int64_t cfdhip_batch_size(const cfdhip_batch* b) {
  return b == nullptr ? -1 : static_cast<int64_t>(b->sys.size());
}

int cfdhip_batch_add_system(cfdhip_batch* b, int64_t nx, int64_t ny, int64_t pin,
                            const double* diag, const double* face_coeff, const double* rhs,
                            int64_t* index) {
  return guardedBatch(b, [&](cfdhip_batch& B) {
    require(nx >= 1 && ny >= 1 && nx <= (int64_t{1} << 20) && ny <= (int64_t{1} << 20),
            "batch system: nx and ny must be in [1, 2^20]");
    require(nx * ny < (int64_t{1} << 31), "batch system: too many cells");
    require(pin >= -1 && pin < nx * ny, "batch system: pinned cell out of range");
    require(diag != nullptr && face_coeff != nullptr && rhs != nullptr,
            "batch system: null array");
    const int64_t n = nx * ny, nf = faceCount2d(nx, ny);
    auto s = std::make_unique<cfdhip_batch::Sys>();
    s->nx = nx;
    s->ny = ny;
    s->pin = pin;
    s->diag.ensure(n);
    s->coeff.ensure(nf);
    s->b.ensure(n);
    CFDHIP_CHECK(hipMemcpyAsync(s->diag.get(), diag, n * sizeof(double), hipMemcpyHostToDevice,
                                B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(s->coeff.get(), face_coeff, nf * sizeof(double),
                                hipMemcpyHostToDevice, B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(s->b.get(), rhs, n * sizeof(double), hipMemcpyHostToDevice,
                                B.stream));
    s->sides = inferFixedValueSides(nx, ny, pin, diag, face_coeff);
    CFDHIP_CHECK(hipStreamSynchronize(B.stream));
    B.sys.push_back(std::move(s));
    if (index) *index = static_cast<int64_t>(B.sys.size()) - 1;
  });
}

int cfdhip_batch_add_pressure(cfdhip_batch* b, cfdhip_ctx* ctx, int64_t* index) {
  if (ctx == nullptr) return CFDHIP_ERR_INVALID;
  return guardedBatch(b, [&](cfdhip_batch& B) {
    Context& c = ctx->c;
    require(c.device == B.device, "batch system: the context is on another device");
    require(c.hasMesh && c.view.dim == 2, "batch system: needs a 2D mesh");
    require(c.pressureAssembled, "batch system: assemble the pressure correction first");
    const int64_t nx = c.view.nx, ny = c.view.ny;
    const int64_t n = nx * ny, nf = faceCount2d(nx, ny);
    require(nf == c.view.nf, "batch system: unexpected face count");
    auto s = std::make_unique<cfdhip_batch::Sys>();
    s->nx = nx;
    s->ny = ny;
    s->pin = c.pin;
    for (int p = 0; p < 4 && p < c.view.npatch; ++p)
      if (c.view.patch[p].pkind == CFDHIP_FIXED_VALUE) s->sides |= 1 << p;
    s->diag.ensure(n);
    s->coeff.ensure(nf);
    s->b.ensure(n);
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
    CFDHIP_CHECK(hipMemcpyAsync(s->diag.get(), c.f(CFDHIP_FIELD_PC_DIAG), n * sizeof(double),
                                hipMemcpyDeviceToDevice, B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(s->coeff.get(), c.f(CFDHIP_FIELD_PC_FACE_COEFF),
                                nf * sizeof(double), hipMemcpyDeviceToDevice, B.stream));
    CFDHIP_CHECK(hipMemcpyAsync(s->b.get(), c.f(CFDHIP_FIELD_PC_RHS), n * sizeof(double),
                                hipMemcpyDeviceToDevice, B.stream));
    CFDHIP_CHECK(hipStreamSynchronize(B.stream));
    B.sys.push_back(std::move(s));
    if (index) *index = static_cast<int64_t>(B.sys.size()) - 1;
  });
}

int cfdhip_batch_solve(cfdhip_batch* b, const cfdhip_solver_settings* s, int tile_rows,
                       int use_graph, cfdhip_solve_result* results, double* latency_ms,
                       double* stats) {
  if (s == nullptr || results == nullptr) return CFDHIP_ERR_INVALID;
  return guardedBatch(b, [&](cfdhip_batch& B) {
    const bool learned = s->type == CFDHIP_SOLVER_FCG && s->preconditioner == CFDHIP_PREC_LEARNED;
    require(learned || (s->type == CFDHIP_SOLVER_CG && s->preconditioner == CFDHIP_PREC_JACOBI),
            "batched solves run CG with the Jacobi preconditioner or flexible CG with the "
            "learned one");
    require(s->absolute_tolerance >= 0.0 && s->relative_tolerance >= 0.0,
            "tolerances must be >= 0");
    require(s->inner_tolerance >= 0.0 && s->inner_tolerance < 1.0,
            "LinearSolverSettings: inner_tolerance must be in [0, 1)");
    if (learned)
      solveBatchLearned(B, *s, use_graph, results, latency_ms, stats);
    else
      solveBatchJacobi(B, *s, tile_rows, use_graph, results, latency_ms, stats);
  });
}

int cfdhip_batch_set_learned(cfdhip_batch* b, const float* weights, int64_t count,
                             const cfdhip_unet_options* options) {
  return guardedBatch(b, [&](cfdhip_batch& B) {
    require(weights != nullptr && count == kUnetParams,
            "learned weights: expected cfdhip_unet_param_count() floats");
    require(options != nullptr && options->padding == 1,
            "the batched learned solve needs options with Neumann padding (padding 1)");
    if (!B.net) B.net = std::make_unique<Unet>(B.stream);
    B.net->setDense(false);
    B.net->setOptions(options->padding, 0, tailFrom(*options));
    B.net->setWeights(weights);
    CFDHIP_CHECK(hipStreamSynchronize(B.stream));
    require(B.net->narrowActive(),
            "the batched learned solve runs the narrow kernels: the weights must leave at most "
            "4 live channels per tensor");
  });
}

int cfdhip_batch_solution(cfdhip_batch* b, int64_t index, double* host, int64_t count) {
  return guardedBatch(b, [&](cfdhip_batch& B) {
    require(index >= 0 && index < static_cast<int64_t>(B.sys.size()), "batch index out of range");
    require(host != nullptr, "null output");
    require(static_cast<int64_t>(B.xOff.size()) == static_cast<int64_t>(B.sys.size()),
            "solve the batch first");
    const int64_t n = B.sys[index]->nx * B.sys[index]->ny;
    require(count >= n, "output buffer too small");
    CFDHIP_CHECK(hipMemcpyAsync(host, B.arena64.get() + B.xOff[index], n * sizeof(double),
                                hipMemcpyDeviceToHost, B.stream));
    CFDHIP_CHECK(hipStreamSynchronize(B.stream));
  });
}

}  // extern "C"
