// Batched flexible CG with the learned preconditioner: cfdhip_batch_solve
// with type CFDHIP_SOLVER_FCG and preconditioner CFDHIP_PREC_LEARNED.
//
// Every active case of a batch of 2D p' systems of uneven size advances one
// FCG(1) iteration per round, and each step of a round is one launch over all
// active cases: the network's layers over a table of (case, tile) entries
// (Unet::forwardNarrowBatch), the coarse tail one workgroup per case, the
// Krylov vector kernels over (case, block) entries, and a one-block-per-case
// finalize that moves every case's scalars. The host reads the cases' states
// every check_interval rounds and drops the finished cases from the tables;
// a round chunk is captured into a HIP graph and replayed while the active
// set is unchanged. No counterpart in the reference, which solves one case
// at a time (src/algebra/CG.cpp:14-148).
//
// The iteration is the single-case learned solve's (fcg.hip) with FP64
// Krylov vectors: z = S N(S r) (z = r / a_P at the pinned cell), beta =
// -(z, q_old) / (p_old, q_old), p = z + beta p_old, q = A p, alpha = (p, r) /
// (p, q), x += alpha p, r -= alpha q. At the end the FP64 true residual of
// every case is recomputed; a case whose recurrence residual drifted from it
// restarts from r = b - A x (at most kMaxRestarts times), and a case that
// stagnated for stagnation_window iterations or broke down is solved again
// by the batched FP64 Jacobi-PCG from x = 0 (fallback_used).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "reduce.hpp"

namespace cfdhip {
namespace {

constexpr int64_t kVecCells = 2048;  // cells per block of the vector kernels (256 x 8)
constexpr int kMaxRestarts = 3;

// One case as the vector kernels see it: its operator (diagonal, face
// couplings, right-hand side), its offset in every vector region of the
// arena, and its index in the batch (its state).
struct LCase {
  int64_t nx, ny, pin, offY, cells;
  const double* diag;
  const double* coeff;
  const double* b;
  int32_t id;
};

// Row of A applied at (i, j) of case L, the p' operator in ascending column
// order from 0 (fcg.hip's fcgDirSpmvKernel, cell c case-local).
template <class F>
__device__ inline double applyRowAt(const LCase& L, int64_t c, F at) {
  const int64_t i = c % L.nx, j = c / L.nx;
  if (c == L.pin) return 0.0 + L.diag[c] * at(c);
  double y = 0.0;
  if (j > 0) y += (-L.coeff[L.offY + j * L.nx + i]) * at(c - L.nx);
  if (i > 0) y += (-L.coeff[j * (L.nx + 1) + i]) * at(c - 1);
  y += L.diag[c] * at(c);
  if (i < L.nx - 1) y += (-L.coeff[j * (L.nx + 1) + i + 1]) * at(c + 1);
  if (j < L.ny - 1) y += (-L.coeff[L.offY + (j + 1) * L.nx + i]) * at(c + L.nx);
  return y;
}

// The block's W sums into partial[slot * W + w] (fixed order in the block).
template <int W>
__device__ inline void slotPartials(double (&v)[W], double* partial, int64_t slot) {
  __shared__ double lds[W][256 / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int w = 0; w < W; ++w) {
    double x = v[w];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) lds[w][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < W) {
    double t = 0.0;
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) t += lds[threadIdx.x][i];
    partial[slot * W + threadIdx.x] = t;
  }
}

// Block entries e = (case, first cell, cells, slot) of the vector kernels.
// x = 0, r = b, q = 0, |b|^2.
__global__ __launch_bounds__(256) void fcgStartBatchKernel(const LCase* __restrict__ cases,
                                                           const int4* __restrict__ blocks,
                                                           double* __restrict__ x,
                                                           double* __restrict__ r,
                                                           double* __restrict__ q,
                                                           double* partial) {
  const int4 e = blocks[blockIdx.x];
  const LCase& L = cases[e.x];
  double acc[1] = {0.0};
  for (int64_t c = e.y + threadIdx.x; c < e.y + e.z; c += blockDim.x) {
    const int64_t k = L.cells + c;
    const double bc = L.b[c];
    x[k] = 0.0;
    r[k] = bc;
    q[k] = 0.0;
    acc[0] += bc * bc;
  }
  slotPartials<1>(acc, partial, e.w);
}

// p = z + beta p_old (the first direction of a case: p = z); q = A p;
// (p, q), (p, r).
__global__ __launch_bounds__(256) void fcgDirBatchKernel(const LCase* __restrict__ cases,
                                                         const int4* __restrict__ blocks,
                                                         const KState* __restrict__ states,
                                                         const double* __restrict__ z,
                                                         const double* __restrict__ pOld,
                                                         double* __restrict__ pNew,
                                                         double* __restrict__ q,
                                                         const double* __restrict__ r,
                                                         double* partial) {
  const int4 e = blocks[blockIdx.x];
  const LCase& L = cases[e.x];
  const KState& st = states[L.id];
  if (st.done) return;  // one case per block: uniform
  const bool first = st.iter == st.dirStart;
  const double beta = st.beta;
  const double* zc = z + L.cells;
  const double* pc = pOld + L.cells;
  auto pAt = [&](int64_t c) { return first ? zc[c] : zc[c] + beta * pc[c]; };
  double acc[2] = {0.0, 0.0};
  for (int64_t c = e.y + threadIdx.x; c < e.y + e.z; c += blockDim.x) {
    const double pv = pAt(c);
    const double y = applyRowAt(L, c, pAt);
    pNew[L.cells + c] = pv;
    q[L.cells + c] = y;
    acc[0] += pv * y;
    acc[1] += pv * r[L.cells + c];
  }
  slotPartials<2>(acc, partial, e.w);
}

// x += alpha p; r -= alpha q; |r|^2.
__global__ __launch_bounds__(256) void fcgUpdateBatchKernel(const LCase* __restrict__ cases,
                                                            const int4* __restrict__ blocks,
                                                            const KState* __restrict__ states,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ q,
                                                            double* __restrict__ x,
                                                            double* __restrict__ r,
                                                            double* partial) {
  const int4 e = blocks[blockIdx.x];
  const LCase& L = cases[e.x];
  const KState& st = states[L.id];
  if (st.done) return;
  const double alpha = st.alpha;
  double acc[1] = {0.0};
  for (int64_t c = e.y + threadIdx.x; c < e.y + e.z; c += blockDim.x) {
    const int64_t k = L.cells + c;
    x[k] += alpha * p[k];
    const double rc = r[k] - alpha * q[k];
    r[k] = rc;
    acc[0] += rc * rc;
  }
  slotPartials<1>(acc, partial, e.w);
}

// The FP64 true residual r = b - A x and |r|^2.
__global__ __launch_bounds__(256) void trueResidualBatchKernel(const LCase* __restrict__ cases,
                                                               const int4* __restrict__ blocks,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ r,
                                                               double* partial) {
  const int4 e = blocks[blockIdx.x];
  const LCase& L = cases[e.x];
  const double* xc = x + L.cells;
  double acc[1] = {0.0};
  for (int64_t c = e.y + threadIdx.x; c < e.y + e.z; c += blockDim.x) {
    const double rc = L.b[c] - applyRowAt(L, c, [&](int64_t k) { return xc[k]; });
    r[L.cells + c] = rc;
    acc[0] += rc * rc;
  }
  slotPartials<1>(acc, partial, e.w);
}

inline int64_t pad64(int64_t n) { return (n + 63) / 64 * 64; }
inline dim3 grid1(int64_t n) { return dim3(static_cast<unsigned>(std::max<int64_t>(1, n))); }

// Host-built tables of the active cases, uploaded in one copy.
struct Tables {
  NarrowBatch nb;
  const LCase* lcases = nullptr;
  const int4* vec = nullptr;
  int64_t vecCount = 0;
  const int32_t* active = nullptr;
  int64_t nActive = 0;
};

}  // namespace

void solveBatchLearned(cfdhip_batch& B, const cfdhip_solver_settings& set, int useGraph,
                       cfdhip_solve_result* results, double* latencyMs, double* stats) {
  const int64_t nc = static_cast<int64_t>(B.sys.size());
  require(nc > 0, "the batch holds no systems");
  require(nc < (int64_t{1} << 31), "too many systems in the batch");
  require(B.net && B.net->ready(),
          "the batch has no learned weights: call cfdhip_batch_set_learned_weights first");
  require(B.net->padding() == 1,
          "the batched learned solve needs the Neumann-padded network (padding 1)");
  require(set.precision == CFDHIP_FP64,
          "the batched learned solve keeps FP64 Krylov vectors (precision CFDHIP_FP64)");
  require(set.max_iterations > 0, "max_iterations must be positive");
  Unet& net = *B.net;
  hipStream_t st = B.stream;

  // ---- layout: vector regions x r z q p0 p1, the scaling, the partial slots
  std::vector<LCase> lc(nc);
  std::vector<NarrowCase> ncs(nc);
  std::vector<int64_t> vSlot0(nc), vSlots(nc), lSlot0(nc), lSlots(nc);
  int64_t N = 0, nVecSlots = 0, nLastSlots = 0;
  for (int64_t c = 0; c < nc; ++c) {
    const auto& s = *B.sys[c];
    require(s.nx <= 1024 && s.ny <= 1024,
            "the batched learned solve takes cases with sides up to 1024 cells");
    const int64_t n = s.nx * s.ny;
    lc[c] = LCase{s.nx, s.ny, s.pin, (s.nx + 1) * s.ny, N, s.diag.get(), s.coeff.get(),
                  s.b.get(), static_cast<int32_t>(c)};
    NarrowCase& k = ncs[c];
    k = NarrowCase{};
    k.ny = static_cast<int>(s.ny);
    k.nx = static_cast<int>(s.nx);
    k.pad = kPadNeumann | (s.sides << 1);
    k.cells = N;
    k.pin = s.pin;
    vSlot0[c] = nVecSlots;
    vSlots[c] = (n + kVecCells - 1) / kVecCells;
    nVecSlots += vSlots[c];
    lSlot0[c] = nLastSlots;
    lSlots[c] = (s.ny + kNarrowLastRows - 1) / kNarrowLastRows;
    nLastSlots += lSlots[c];
    N += pad64(n);
  }
  const int64_t actFloats = net.narrowBatchLayout(ncs);
  B.arena64.ensure(6 * N);
  double* X = B.arena64.get();
  double* R = X + N;
  double* Z = R + N;
  double* Q = Z + N;
  double* P[2] = {Q + N, Q + 2 * N};
  B.scale32.ensure(3 * N);
  float* S = B.scale32.get();
  B.act.ensure(actFloats);
  B.xOff.assign(nc, 0);
  for (int64_t c = 0; c < nc; ++c) B.xOff[c] = lc[c].cells;
  // partials: the vector kernels' slots (2 sums each at most), then the
  // output layer's; the slot ranges per case (static: indexed by case id)
  B.partial.ensure(2 * nVecSlots + nLastSlots + 8);
  double* vPart = B.partial.get();
  double* lPart = vPart + 2 * nVecSlots;
  B.kstates.ensure(nc);
  B.norms.ensure(nc);

  // ---- the scaling S of every case
  for (int64_t c = 0; c < nc; ++c)
    learnedScaling(st, lc[c].nx, lc[c].ny, lc[c].offY, lc[c].pin, lc[c].diag, lc[c].coeff,
                   S + N + lc[c].cells, S + 2 * N + lc[c].cells, S + lc[c].cells);

  // ---- device tables: static slot ranges, then the active set's entries
  int64_t maxTab[4] = {0, 0, 0, 0}, maxLast = 0, maxVec = 0;
  for (int64_t c = 0; c < nc; ++c) {
    for (int l = 0; l < 4; ++l)
      maxTab[l] += static_cast<int64_t>(ncs[c].tilesX[l]) *
                   ((ncs[c].h[l] + kNarrowTileH - 1) / kNarrowTileH);
    maxLast += lSlots[c];
    maxVec += vSlots[c];
  }
  const int64_t shapeBytes = Unet::tailShapeBytes();
  auto al = [](int64_t b) { return (b + 255) / 256 * 256; };
  const int64_t bStatic = al(4 * nc * static_cast<int64_t>(sizeof(int64_t)));
  const int64_t bCases = al(nc * static_cast<int64_t>(sizeof(NarrowCase)));
  const int64_t bL = al(nc * static_cast<int64_t>(sizeof(LCase)));
  const int64_t bShapes = al(nc * shapeBytes);
  int64_t bTab[4];
  for (int l = 0; l < 4; ++l) bTab[l] = al(maxTab[l] * static_cast<int64_t>(sizeof(int2)));
  const int64_t bLast = al(maxLast * static_cast<int64_t>(sizeof(int4)));
  const int64_t bVec = al(maxVec * static_cast<int64_t>(sizeof(int4)));
  const int64_t bAct = al(nc * static_cast<int64_t>(sizeof(int32_t)));
  const int64_t total = bStatic + bCases + bL + bShapes + bTab[0] + bTab[1] + bTab[2] + bTab[3] +
                        bLast + bVec + bAct;
  B.table.ensure(total);
  if (B.hostStagingBytes < total) {
    if (B.hostStaging) CFDHIP_CHECK(hipHostFree(B.hostStaging));
    B.hostStaging = nullptr;
    CFDHIP_CHECK(hipHostMalloc(&B.hostStaging, total));
    B.hostStagingBytes = total;
  }
  unsigned char* dev = B.table.get();
  unsigned char* host = static_cast<unsigned char*>(B.hostStaging);
  // static part: slot ranges (vector slot0, slots, output-layer slot0, slots)
  CFDHIP_CHECK(hipStreamSynchronize(st));
  {
    int64_t* h = reinterpret_cast<int64_t*>(host);
    std::memcpy(h, vSlot0.data(), nc * sizeof(int64_t));
    std::memcpy(h + nc, vSlots.data(), nc * sizeof(int64_t));
    std::memcpy(h + 2 * nc, lSlot0.data(), nc * sizeof(int64_t));
    std::memcpy(h + 3 * nc, lSlots.data(), nc * sizeof(int64_t));
  }
  const int64_t* dVSlot0 = reinterpret_cast<const int64_t*>(dev);
  const int64_t* dVSlots = dVSlot0 + nc;
  const int64_t* dLSlot0 = dVSlot0 + 2 * nc;
  const int64_t* dLSlots = dVSlot0 + 3 * nc;
  Tables T;
  // the active set's tables (compact case indices; host staging, one copy)
  auto build = [&](const std::vector<int32_t>& ids) {
    CFDHIP_CHECK(hipStreamSynchronize(st));  // the staging buffer may still be in flight
    int64_t o = bStatic;
    NarrowCase* hCases = reinterpret_cast<NarrowCase*>(host + o);
    T.nb.cases = reinterpret_cast<const NarrowCase*>(dev + o);
    o += bCases;
    LCase* hL = reinterpret_cast<LCase*>(host + o);
    T.lcases = reinterpret_cast<const LCase*>(dev + o);
    o += bL;
    T.nb.tailShapes = dev + o;
    const int64_t oShapes = o;
    o += bShapes;
    int2* hTab[4];
    for (int l = 0; l < 4; ++l) {
      hTab[l] = reinterpret_cast<int2*>(host + o);
      T.nb.tab[l] = reinterpret_cast<const int2*>(dev + o);
      T.nb.tabCount[l] = 0;
      o += bTab[l];
    }
    int4* hLast = reinterpret_cast<int4*>(host + o);
    T.nb.last = reinterpret_cast<const int4*>(dev + o);
    o += bLast;
    int4* hVec = reinterpret_cast<int4*>(host + o);
    T.vec = reinterpret_cast<const int4*>(dev + o);
    o += bVec;
    int32_t* hAct = reinterpret_cast<int32_t*>(host + o);
    T.active = reinterpret_cast<const int32_t*>(dev + o);
    std::vector<NarrowCase> actCases;
    int64_t nLast = 0, nVec = 0;
    for (size_t k = 0; k < ids.size(); ++k) {
      const int32_t c = ids[k];
      hCases[k] = ncs[c];
      actCases.push_back(ncs[c]);
      hL[k] = lc[c];
      hAct[k] = c;
      for (int l = 0; l < 4; ++l) {
        const int tiles = ncs[c].tilesX[l] * ((ncs[c].h[l] + kNarrowTileH - 1) / kNarrowTileH);
        for (int t = 0; t < tiles; ++t)
          hTab[l][T.nb.tabCount[l]++] = make_int2(static_cast<int>(k), t);
      }
      for (int64_t q = 0; q < lSlots[c]; ++q)
        hLast[nLast++] = make_int4(static_cast<int>(k), static_cast<int>(q * kNarrowLastRows),
                                   kNarrowLastRows, static_cast<int>(lSlot0[c] + q));
      const int64_t n = lc[c].nx * lc[c].ny;
      for (int64_t q = 0; q < vSlots[c]; ++q) {
        const int64_t c0 = q * kVecCells;
        hVec[nVec++] = make_int4(static_cast<int>(k), static_cast<int>(c0),
                                 static_cast<int>(std::min(kVecCells, n - c0)),
                                 static_cast<int>(vSlot0[c] + q));
      }
    }
    T.nb.nCases = static_cast<int64_t>(ids.size());
    T.nb.lastCount = nLast;
    T.nb.act = B.act.get();
    T.vecCount = nVec;
    T.nActive = static_cast<int64_t>(ids.size());
    // the tail shapes go through the network's own staging copy
    net.narrowBatchTail(actCases, host + oShapes, dev + oShapes);
    CFDHIP_CHECK(hipMemcpyAsync(dev, host, o + bAct, hipMemcpyHostToDevice, st));
  };

  // ---- states
  std::vector<KState> hs(nc);
  for (int64_t c = 0; c < nc; ++c) {
    KState k{};
    k.absTol = set.absolute_tolerance;
    k.relTol = set.relative_tolerance;
    k.maxIter = set.max_iterations;
    k.window = set.stagnation_window > 0 ? set.stagnation_window : 200;
    k.status = -1;
    k.pAp = 1.0;  // the first beta step divides by it; that direction takes p = z
    hs[c] = k;
  }
  CFDHIP_CHECK(hipMemcpyAsync(B.kstates.get(), hs.data(), sizeof(KState) * nc,
                              hipMemcpyHostToDevice, st));
  CFDHIP_CHECK(hipStreamSynchronize(st));

  hipEvent_t e0, e1, ce;
  CFDHIP_CHECK(hipEventCreate(&e0));
  CFDHIP_CHECK(hipEventCreate(&e1));
  CFDHIP_CHECK(hipEventCreate(&ce));
  CFDHIP_CHECK(hipEventRecord(e0, st));
  std::vector<int32_t> ids(nc);
  for (int64_t c = 0; c < nc; ++c) ids[c] = static_cast<int32_t>(c);
  build(ids);
  KState* K = B.kstates.get();
  hipLaunchKernelGGL(fcgStartBatchKernel, grid1(T.vecCount), dim3(256), 0, st, T.lcases, T.vec,
                     X, R, Q, vPart);
  launchCaseFinalize(st, vPart, dVSlot0, dVSlots, T.active, T.nActive, K, 1, kStepStartNorm,
                     nullptr);
  CFDHIP_CHECK(hipGetLastError());

  int cur = 0;
  auto round = [&]() {
    net.forwardNarrowBatch(T.nb, R, S, Q, Z, lPart);
    launchCaseFinalize(st, lPart, dLSlot0, dLSlots, T.active, T.nActive, K, 1, kStepFcgBeta,
                       nullptr);
    hipLaunchKernelGGL(fcgDirBatchKernel, grid1(T.vecCount), dim3(256), 0, st, T.lcases, T.vec,
                       K, Z, P[1 - cur], P[cur], Q, R, vPart);
    launchCaseFinalize(st, vPart, dVSlot0, dVSlots, T.active, T.nActive, K, 2, kStepFcgAlpha,
                       nullptr);
    hipLaunchKernelGGL(fcgUpdateBatchKernel, grid1(T.vecCount), dim3(256), 0, st, T.lcases,
                       T.vec, K, P[cur], Q, X, R, vPart);
    launchCaseFinalize(st, vPart, dVSlot0, dVSlots, T.active, T.nActive, K, 1, kStepFcgUpdate,
                       nullptr);
    cur = 1 - cur;
  };
  // chunks of rounds (an even count, so the direction buffers' parity
  // returns); the host reads the states after each
  const int chunk = std::max(2, (set.check_interval > 0 ? set.check_interval : 4) & ~1);
  std::vector<char> done(nc, 0);
  std::vector<double> doneMs(nc, 0.0);
  std::vector<int> restarts(nc, 0);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int captures = 0;
  int64_t rounds = 0, cellIters = 0;
  bool dirty = false;
  auto dropGraph = [&]() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  };
  while (true) {
    // every case done: the FP64 true residuals; drifted cases restart
    auto readStates = [&]() {
      CFDHIP_CHECK(hipMemcpyAsync(hs.data(), K, sizeof(KState) * nc, hipMemcpyDeviceToHost, st));
      CFDHIP_CHECK(hipEventRecord(ce, st));
      CFDHIP_CHECK(hipEventSynchronize(ce));
      float at = 0.f;
      CFDHIP_CHECK(hipEventElapsedTime(&at, e0, ce));
      return static_cast<double>(at);
    };
    const double at = readStates();
    std::vector<int32_t> act;
    for (int64_t c = 0; c < nc; ++c) {
      if (!done[c] && hs[c].done) {
        done[c] = 1;
        doneMs[c] = at;
        dirty = true;
      }
      if (!done[c]) act.push_back(static_cast<int32_t>(c));
    }
    if (act.empty()) {
      // true residuals of every case, then restarts of the converged ones
      // whose recurrence residual drifted from the true one
      build(ids);
      hipLaunchKernelGGL(trueResidualBatchKernel, grid1(T.vecCount), dim3(256), 0, st, T.lcases,
                         T.vec, X, R, vPart);
      launchCaseFinalize(st, vPart, dVSlot0, dVSlots, T.active, T.nActive, K, 1, kStepNorm,
                         B.norms.get());
      std::vector<double> tr(nc);
      CFDHIP_CHECK(hipMemcpyAsync(tr.data(), B.norms.get(), sizeof(double) * nc,
                                  hipMemcpyDeviceToHost, st));
      CFDHIP_CHECK(hipStreamSynchronize(st));
      bool again = false;
      for (int64_t c = 0; c < nc; ++c) {
        KState& k = hs[c];
        const double target = std::max(k.absTol, k.relTol * k.b0);
        if (k.status == CFDHIP_CONVERGED && tr[c] > target && restarts[c] < kMaxRestarts &&
            k.iter < k.maxIter) {
          ++restarts[c];
          k.done = 0;
          k.status = -1;
          k.dirStart = k.iter;
          k.best = tr[c];
          k.stagnant = 0;
          done[c] = 0;
          again = true;
        }
        k.rnorm = k.done ? tr[c] : k.rnorm;
      }
      if (!again) break;
      CFDHIP_CHECK(hipMemcpyAsync(K, hs.data(), sizeof(KState) * nc, hipMemcpyHostToDevice, st));
      dirty = true;
      continue;
    }
    if (rounds >= 4 * set.max_iterations + 64) break;
    if (dirty) {
      build(act);
      dropGraph();
      dirty = false;
    }
    int64_t actCells = 0;
    for (int32_t c : act) actCells += lc[c].nx * lc[c].ny;
    if (useGraph) {
      if (!exec) {
        CFDHIP_CHECK(hipStreamSynchronize(st));
        CFDHIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < chunk; ++k) round();
        CFDHIP_CHECK(hipStreamEndCapture(st, &graph));
        CFDHIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        ++captures;
      }
      CFDHIP_CHECK(hipGraphLaunch(exec, st));
    } else {
      for (int k = 0; k < chunk; ++k) round();
    }
    CFDHIP_CHECK(hipGetLastError());
    rounds += chunk;
    cellIters += actCells * chunk;
  }
  dropGraph();
  CFDHIP_CHECK(hipEventRecord(e1, st));
  CFDHIP_CHECK(hipEventSynchronize(e1));
  float totalMs = 0.f;
  CFDHIP_CHECK(hipEventElapsedTime(&totalMs, e0, e1));

  // ---- fallback: the cases that did not converge, by the batched FP64 Jacobi-PCG
  std::vector<int64_t> fb;
  for (int64_t c = 0; c < nc; ++c)
    if (hs[c].status != CFDHIP_CONVERGED) fb.push_back(c);
  std::vector<cfdhip_solve_result> fr(fb.size());
  std::vector<double> fLat(fb.size(), 0.0);
  if (!fb.empty()) {
    cfdhip_batch sub;
    sub.device = B.device;
    sub.stream = st;
    for (int64_t c : fb) sub.sys.push_back(std::move(B.sys[c]));
    cfdhip_solver_settings js = set;
    js.type = CFDHIP_SOLVER_CG;
    js.preconditioner = CFDHIP_PREC_JACOBI;
    js.precision = CFDHIP_FP64;
    try {
      solveBatchJacobi(sub, js, 0, useGraph, fr.data(), fLat.data(), nullptr);
      for (size_t k = 0; k < fb.size(); ++k) {
        const int64_t c = fb[k];
        const int64_t n = lc[c].nx * lc[c].ny;
        CFDHIP_CHECK(hipMemcpyAsync(X + lc[c].cells, sub.arena64.get() + sub.xOff[k],
                                    n * sizeof(double), hipMemcpyDeviceToDevice, st));
      }
      CFDHIP_CHECK(hipStreamSynchronize(st));
    } catch (...) {
      for (size_t k = 0; k < fb.size(); ++k) B.sys[fb[k]] = std::move(sub.sys[k]);
      sub.stream = nullptr;
      throw;
    }
    for (size_t k = 0; k < fb.size(); ++k) B.sys[fb[k]] = std::move(sub.sys[k]);
    sub.stream = nullptr;  // the batch's own stream
    B.xOff.assign(nc, 0);
    for (int64_t c = 0; c < nc; ++c) B.xOff[c] = lc[c].cells;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(ce);

  for (int64_t c = 0; c < nc; ++c) {
    const KState& k = hs[c];
    cfdhip_solve_result& r = results[c];
    r = cfdhip_solve_result{};
    r.status = k.done ? k.status : CFDHIP_MAX_ITERATIONS;
    r.iterations = k.iter;
    r.restarts = restarts[c];
    r.initial_residual = k.b0;
    r.final_residual = k.rnorm;
    r.true_residual = k.rnorm;
    r.device_ms = done[c] ? doneMs[c] : totalMs;
  }
  for (size_t q = 0; q < fb.size(); ++q) {
    cfdhip_solve_result& r = results[fb[q]];
    r.status = fr[q].status;
    r.iterations += fr[q].iterations;
    r.true_residual = fr[q].true_residual;
    r.final_residual = fr[q].final_residual;
    r.fallback_used = 1;
    r.device_ms = totalMs + fLat[q];
  }
  if (latencyMs)
    for (int64_t c = 0; c < nc; ++c) latencyMs[c] = results[c].device_ms;
  if (stats) {
    double fbMs = 0.0;
    for (double v : fLat) fbMs = std::max(fbMs, v);
    stats[0] = totalMs + fbMs;
    stats[1] = static_cast<double>(rounds);
    stats[2] = static_cast<double>(cellIters);
    stats[3] = 0.0;
    stats[4] = static_cast<double>(captures);
    stats[5] = stats[6] = stats[7] = 0.0;
  }
}

}  // namespace cfdhip
