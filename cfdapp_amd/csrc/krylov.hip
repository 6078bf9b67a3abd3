// Krylov solvers on the device: CG (src/algebra/CG.cpp:14-148) and BiCGSTAB
// (src/algebra/BiCGSTAB.cpp:48-275) with Jacobi (Preconditioner.cpp:11-45),
// over three operators: the resident p' operator (diagonal + face couplings),
// a 5-/7-point stencil (the momentum matrix) and a general CSR matrix.
//
// Every inner product is a two-stage deterministic reduction: wave shuffles
// give block partials, then a one-block finalize kernel sums the partials in
// a fixed order and evaluates the solver's scalar recurrence and its
// convergence / breakdown tests on the device. The host reads the verdict
// only every check_interval iterations, so an iteration costs no host round
// trip (the reference pays 5 host syncs per BiCGSTAB iteration,
// cuda/kernels/DeviceVectorOpsKernel.cu:226-260).
#include <cfloat>
#include <algorithm>
#include <cmath>
#include <memory>

#include "reduce.hpp"

namespace cfdhip {

namespace {

constexpr int kMaxBlocks = 2048;

inline dim3 gridStride(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : b));
}

// ---------------------------------------------------------------------------
// Operators. apply(c, x) = row c of A times x, summed in ascending column
// order from 0.0 (SparseMatrix::multiply); a stored zero contributes a signed
// zero, which leaves a sum that started at +0.0 unchanged, so the stencil
// forms equal the zero-dropped CSR bit for bit on finite x.

struct PressureOp {
  View v;
  const double* diag;
  const double* coeff;
  int64_t pin;
  __device__ double apply(int64_t c, const double* __restrict__ x) const {
    if (c == pin) return 0.0 + diag[c] * x[c];
    int64_t idx[3];
    cellIndex(v, c, idx);
    const int D = v.dim;
    double y = 0.0;
    // low neighbours, highest axis first (ascending column)
    for (int a = D - 1; a >= 0; --a) {
      if (idx[a] == 0) continue;
      int64_t f[3] = {idx[0], idx[1], idx[2]};
      y += (-coeff[faceId(v, a, f)]) * x[c - v.stride[a]];
    }
    y += diag[c] * x[c];
    for (int a = 0; a < D; ++a) {
      if (idx[a] == v.ncell[a] - 1) continue;
      int64_t f[3] = {idx[0], idx[1], idx[2]};
      f[a] += 1;
      y += (-coeff[faceId(v, a, f)]) * x[c + v.stride[a]];
    }
    return y;
  }
  __device__ double diagonal(int64_t c) const { return diag[c]; }
};

struct StencilOp {
  View v;
  const double* coef;  // (2D+1) * n in column order
  __device__ double apply(int64_t c, const double* __restrict__ x) const {
    int64_t idx[3];
    cellIndex(v, c, idx);
    const int D = v.dim;
    double y = 0.0;
    for (int a = D - 1; a >= 0; --a) {
      if (idx[a] == 0) continue;
      y += coef[(D - 1 - a) * v.n + c] * x[c - v.stride[a]];
    }
    y += coef[D * v.n + c] * x[c];
    for (int a = 0; a < D; ++a) {
      if (idx[a] == v.ncell[a] - 1) continue;
      y += coef[(D + 1 + a) * v.n + c] * x[c + v.stride[a]];
    }
    return y;
  }
  __device__ double diagonal(int64_t c) const { return coef[v.dim * v.n + c]; }
};

struct CsrOp {
  int64_t n;
  const int64_t* rp;
  const int64_t* col;
  const double* val;
  __device__ double apply(int64_t r, const double* __restrict__ x) const {
    double sum = 0.0;
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) sum += val[k] * x[col[k]];
    return sum;
  }
  // SparseMatrix::diagonal: the stored diagonal, 0 when absent (rejected).
  __device__ double diagonal(int64_t r) const {
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
      if (col[k] == r) return val[k];
    return 0.0;
  }
};

struct Red {
  double* partial;
  KState* st;
};

// ---------------------------------------------------------------------------
// Shared kernels.

template <class Op>
__global__ void residualKernel(Op A, int64_t n, const double* __restrict__ b,
                               const double* __restrict__ x, double* r, double* copy, Red R,
                               int mode) {
  // mode 0: CG / BiCGSTAB start, 1: BiCGSTAB restart (finalized by the host's step)
  (void)mode;
  double rr[1] = {0.0};
  KState& s = *R.st;
  const bool skip = s.done;
  if (!skip)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double ri = b[i] - A.apply(i, x);
      r[i] = ri;
      if (copy) copy[i] = ri;
      rr[0] += ri * ri;
    }
  blockPartials<1>(rr, R.partial);
}

// Jacobi inverse diagonal with the reference's rejection rules; counts rejects.
template <class Op>
__global__ void jacobiKernel(Op A, int64_t n, double* dinv, unsigned* bad) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double d = A.diagonal(i);
    if (!isfinite(d) || d == 0.0 || fabs(d) < 1e-12) {
      atomicAdd(bad, 1u);
      dinv[i] = 0.0;
    } else {
      dinv[i] = 1.0 / d;
    }
  }
}

// ---------------------------------------------------------------------------
// CG (CG.cpp:59-147).

// z = M r; p = z; rz = r . z
__global__ void cgStartKernel(int64_t n, const double* __restrict__ r,
                              const double* __restrict__ dinv, double* p, Red R) {
  double acc[1] = {0.0};
  KState& s = *R.st;
  const bool skip = s.done;
  if (!skip)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double z = dinv ? dinv[i] * r[i] : r[i];
      p[i] = z;
      acc[0] += r[i] * z;
    }
  blockPartials<1>(acc, R.partial);
}

// Ap = A p; pAp -> alpha
template <class Op>
__global__ void cgSpmvKernel(Op A, int64_t n, const double* __restrict__ p, double* Ap, Red R) {
  double acc[1] = {0.0};
  KState& s = *R.st;
  const bool skip = s.done;
  if (!skip)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double y = A.apply(i, p);
      Ap[i] = y;
      acc[0] += p[i] * y;
    }
  blockPartials<1>(acc, R.partial);
}

// x += p alpha; r = r - Ap alpha; |r|, r . z -> convergence, beta
__global__ void cgUpdateKernel(int64_t n, double* x, double* r, const double* __restrict__ p,
                               const double* __restrict__ Ap, const double* __restrict__ dinv,
                               Red R) {
  double acc[2] = {0.0, 0.0};
  KState& s = *R.st;
  const bool skip = s.done;
  if (!skip) {
    const double alpha = s.alpha;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      x[i] += p[i] * alpha;
      const double rn = r[i] - (Ap[i] * alpha);
      r[i] = rn;
      const double z = dinv ? dinv[i] * rn : rn;
      acc[0] += rn * rn;
      acc[1] += rn * z;
    }
  }
  blockPartials<2>(acc, R.partial);
}

// p = z + p beta
__global__ void cgDirectionKernel(int64_t n, const double* __restrict__ r,
                                  const double* __restrict__ dinv, double* p, const KState* st) {
  if (st->done) return;
  const double beta = st->beta;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double z = dinv ? dinv[i] * r[i] : r[i];
    p[i] = z + (p[i] * beta);
  }
}

// The caller's host preconditioner (CFDHIP_PREC_HOST): z arrives from the
// host, so the kernels that form z = M r inline take it as a vector.
// p = z; r . z
__global__ void cgStartZKernel(int64_t n, const double* __restrict__ r,
                               const double* __restrict__ z, double* p, Red R) {
  double acc[1] = {0.0};
  if (!R.st->done)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      p[i] = z[i];
      acc[0] += r[i] * z[i];
    }
  blockPartials<1>(acc, R.partial);
}

// x += p alpha; r = r - Ap alpha (no sums: z = M r comes from the host next)
__global__ void cgUpdateXRKernel(int64_t n, double* x, double* r, const double* __restrict__ p,
                                 const double* __restrict__ Ap, const KState* st) {
  if (st->done) return;
  const double alpha = st->alpha;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    x[i] += p[i] * alpha;
    r[i] = r[i] - (Ap[i] * alpha);
  }
}

// |r|^2, r . z (the sums of kStepCgUpdate)
__global__ void rzKernel(int64_t n, const double* __restrict__ r, const double* __restrict__ z,
                         Red R) {
  double acc[2] = {0.0, 0.0};
  if (!R.st->done)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      acc[0] += r[i] * r[i];
      acc[1] += r[i] * z[i];
    }
  blockPartials<2>(acc, R.partial);
}

// p = z + p beta
__global__ void cgDirectionZKernel(int64_t n, const double* __restrict__ z, double* p,
                                   const KState* st) {
  if (st->done) return;
  const double beta = st->beta;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = z[i] + (p[i] * beta);
}

// ---------------------------------------------------------------------------
// BiCGSTAB (BiCGSTAB.cpp:111-274).

// rho = rHat . r -> beta
__global__ void bcgRhoKernel(int64_t n, const double* __restrict__ rhat,
                             const double* __restrict__ r, Red R) {
  double acc[2] = {0.0, 0.0};
  KState& s = *R.st;
  const bool skip = s.done || s.restart;
  if (!skip)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double t = rhat[i] * r[i];
      acc[0] += t;
      acc[1] += fabs(t);
    }
  blockPartials<2>(acc, R.partial);
}

// p = r + (p - v omega) beta; pHat = M p
__global__ void bcgDirectionKernel(int64_t n, const double* __restrict__ r,
                                   const double* __restrict__ v, const double* __restrict__ dinv,
                                   double* p, double* phat, const KState* st) {
  if (st->done || st->restart) return;
  const double beta = st->beta, omega = st->omega;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double pn = r[i] + ((p[i] - (v[i] * omega)) * beta);
    p[i] = pn;
    phat[i] = dinv ? dinv[i] * pn : pn;
  }
}

// v = A pHat; rHat . v -> alpha
template <class Op>
__global__ void bcgVKernel(Op A, int64_t n, const double* __restrict__ phat,
                           const double* __restrict__ rhat, double* v, Red R) {
  double acc[3] = {0.0, 0.0, 0.0};
  KState& s = *R.st;
  const bool skip = s.done || s.restart;
  if (!skip)
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double y = A.apply(i, phat);
      v[i] = y;
      const double t = rhat[i] * y;
      acc[0] += t;
      acc[1] += fabs(t);
      acc[2] += y * y;
    }
  blockPartials<3>(acc, R.partial);
}

// s = r - v alpha; |s| -> early convergence; sHat = M s
__global__ void bcgSKernel(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                           const double* __restrict__ dinv, double* sv, double* shat, Red R) {
  double acc[1] = {0.0};
  KState& s = *R.st;
  const bool skip = s.done || s.restart;
  if (!skip) {
    const double alpha = s.alpha;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double si = r[i] - (v[i] * alpha);
      sv[i] = si;
      shat[i] = dinv ? dinv[i] * si : si;
      acc[0] += si * si;
    }
  }
  blockPartials<1>(acc, R.partial);
}

// t = A sHat; (t,t), (t,s) -> omega   (or the early-exit x update)
template <class Op>
__global__ void bcgTKernel(Op A, int64_t n, const double* __restrict__ shat,
                           const double* __restrict__ sv, const double* __restrict__ phat,
                           double* t, double* x, Red R) {
  double acc[3] = {0.0, 0.0, 0.0};
  KState& s = *R.st;
  const bool skip = s.done || s.restart;
  const bool early = s.sConverged;
  if (!skip) {
    const double alpha = s.alpha;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      if (early) {
        x[i] += phat[i] * alpha;
        continue;
      }
      const double y = A.apply(i, shat);
      t[i] = y;
      const double ts = y * sv[i];
      acc[0] += y * y;
      acc[1] += ts;
      acc[2] += fabs(ts);
    }
  }
  blockPartials<3>(acc, R.partial);
}

// x += pHat alpha; x += sHat omega; r = s - t omega; |r|
__global__ void bcgXKernel(int64_t n, double* x, double* r, const double* __restrict__ phat,
                           const double* __restrict__ shat, const double* __restrict__ sv,
                           const double* __restrict__ t, Red R) {
  double acc[1] = {0.0};
  KState& s = *R.st;
  const bool skip = s.done || s.restart;
  if (!skip) {
    const double alpha = s.alpha, omega = s.omega;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      x[i] += phat[i] * alpha;
      x[i] += shat[i] * omega;
      const double ri = sv[i] - (t[i] * omega);
      r[i] = ri;
      acc[0] += ri * ri;
    }
  }
  blockPartials<1>(acc, R.partial);
}

// ---------------------------------------------------------------------------
// Fused BiCGSTAB for the 2D momentum stencil on small grids. The worker
// kernels above and their five finalize launches make eleven dependent
// launches per iteration; here three kernels do an iteration, each running
// the scalar steps that precede its work itself (bcgStep, redundantly in
// every block, on the previous kernel's partials summed in one fixed order):
//   dirV   [X, Rho]  p = r + (p - v omega) beta, pHat = M p, v = A pHat
//   sT     [V]       s = r - v alpha, sHat = M s, t = A sHat
//   xr     [S, T]    x += pHat alpha (+ sHat omega), r = s - t omega
// dirV and sT form the preconditioned vector of their tile's halo ring
// themselves (the same expression on the same inputs, so the same bits as
// the ring's own tile) and apply the stencil to it from LDS. p and v are
// double-buffered, since a halo reads the previous iterate. Tiles as in the
// single-kernel PCG: 62 interior columns (one lane each, +1 halo each side)
// by ROWS - 2 interior rows, four waves.
constexpr int kBcgOut = 62;
constexpr int64_t kBcgMaxTiles = 256;

struct Bcg2 {
  int64_t nx, ny, n, tilesX;
  const double* coef;  // 5 n: the S, W, C, E, N coefficients of each row
  const double* dinv;  // Jacobi inverse diagonal, or null
  const double* rhat;
  double* r;
  double* x;
  double* p[2];
  double* v[2];
  double* phat;
  double* s;
  double* shat;
  double* t;
  double* hist;
};

struct Bcg2Launch {
  int cur;               // p[cur], v[cur]: the previous iterate
  const KState* stIn;    // the state the steps start from
  KState* stOut;         // block 0 writes the stepped state here
  const double* partIn;  // the previous kernel's partials (one W-tuple per block)
  double* partOut;
  int pending;           // dirV: run the X and Rho steps first (0 after a chunk boundary)
};

// Sums of W partial streams over n blocks, the same order in every block;
// valid in thread 0.
template <int W>
__device__ inline void bcgSumParts(const double* __restrict__ part, int64_t n, double (&t)[W]) {
  __shared__ double lds[W][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double a[W];
#pragma unroll
  for (int q = 0; q < W; ++q) a[q] = 0.0;
  for (int64_t b = threadIdx.x; b < n; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < W; ++q) a[q] += part[b * W + q];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    double y = a[q];
    for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o, 64);
    if (lane == 0) lds[q][wid] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < W; ++q) {
      double y = 0.0;
      for (int i = 0; i < kRedThreads / 64; ++i) y += lds[q][i];
      t[q] = y;
    }
}

// a BiCGSTAB step unless the solve is done or waits for a restart (the
// finalize kernel's rule)
__device__ inline void bcgStepIfRunning(KState& s, double* hist, int stepId, const double* t) {
  if (!s.done && !s.restart) bcgStep(s, hist, stepId, t);
}

template <int ROWS>
struct BcgTile {
  static constexpr int kRowsPerWave = ROWS / 4;
  static constexpr int kInterior = ROWS - 2;
  int64_t i, jBase;
  bool col, outCol;
  int lane, wave;
  __device__ BcgTile(const Bcg2& a) {
    lane = threadIdx.x & 63;
    wave = threadIdx.x >> 6;
    const int64_t tile = xcdTileOrder(blockIdx.x, gridDim.x);
    const int64_t tx = tile % a.tilesX, ty = tile / a.tilesX;
    i = tx * kBcgOut - 1 + lane;
    jBase = ty * kInterior - 1;
    col = i >= 0 && i < a.nx;
    outCol = col && lane >= 1 && lane <= kBcgOut;
  }
  __device__ int row(int k) const { return wave * kRowsPerWave + k; }
  __device__ bool valid(const Bcg2& a, int tr) const {
    const int64_t j = jBase + tr;
    return col && j >= 0 && j < a.ny;
  }
  __device__ bool center(const Bcg2& a, int tr) const {
    return valid(a, tr) && outCol && tr >= 1 && tr <= kInterior;
  }
};

// Row c of the momentum stencil (StencilOp::apply in 2D: S, W, C, E, N, the
// order of SparseMatrix::multiply) on the tile values h (LDS) and the row's
// own value hc and its west/east neighbours.
template <int ROWS>
__device__ inline double bcgApply(const Bcg2& a, int64_t i, int64_t j, int64_t c,
                                  const double (&h)[ROWS][64], int tr, int lane, double hW,
                                  double hC, double hE) {
  double y = 0.0;
  if (j > 0) y += a.coef[c] * h[tr - 1][lane];
  if (i > 0) y += a.coef[a.n + c] * hW;
  y += a.coef[2 * a.n + c] * hC;
  if (i < a.nx - 1) y += a.coef[3 * a.n + c] * hE;
  if (j < a.ny - 1) y += a.coef[4 * a.n + c] * h[tr + 1][lane];
  return y;
}

// the scalar steps of a launch, run by thread 0 of every block on its own
// copy; block 0 publishes the result
template <int W, class F>
__device__ inline void bcgSteps(const Bcg2& a, const Bcg2Launch& L, KState& ks, bool sum, F run) {
  double t[W > 0 ? W : 1];
  if constexpr (W > 0) {
    if (sum) bcgSumParts<W>(L.partIn, gridDim.x, t);
  }
  if (threadIdx.x == 0) {
    ks = *L.stIn;
    if (sum) run(ks, blockIdx.x == 0 ? a.hist : nullptr, t);
    if (blockIdx.x == 0) *L.stOut = ks;
  }
  __syncthreads();
}

template <int ROWS>
__global__ __launch_bounds__(256) void bcgDirVKernel(Bcg2 a, Bcg2Launch L) {
  __shared__ KState ks;
  __shared__ double h[ROWS][64];
  bcgSteps<3>(a, L, ks, L.pending != 0, [](KState& s, double* hist, const double* t) {
    bcgStepIfRunning(s, hist, kStepBcgX, t);
    bcgStepIfRunning(s, hist, kStepBcgRho, t + 1);
  });
  double acc[3] = {0.0, 0.0, 0.0};
  const BcgTile<ROWS> T(a);
  if (!ks.done && !ks.restart) {
    const double beta = ks.beta, omega = ks.omega;
    const double* pOld = a.p[L.cur];
    const double* vOld = a.v[L.cur];
    double* pNew = a.p[1 - L.cur];
    double* vNew = a.v[1 - L.cur];
    double hK[BcgTile<ROWS>::kRowsPerWave];
#pragma unroll
    for (int k = 0; k < BcgTile<ROWS>::kRowsPerWave; ++k) {
      const int tr = T.row(k);
      double ph = 0.0;
      if (T.valid(a, tr)) {
        const int64_t c = (T.jBase + tr) * a.nx + T.i;
        const double pn = a.r[c] + ((pOld[c] - (vOld[c] * omega)) * beta);
        ph = a.dinv ? a.dinv[c] * pn : pn;
        if (T.center(a, tr)) {
          pNew[c] = pn;
          a.phat[c] = ph;
        }
      }
      h[tr][T.lane] = ph;
      hK[k] = ph;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BcgTile<ROWS>::kRowsPerWave; ++k) {
      const int tr = T.row(k);
      const double hW = __shfl_up(hK[k], 1, 64), hE = __shfl_down(hK[k], 1, 64);
      if (T.center(a, tr)) {
        const int64_t j = T.jBase + tr, c = j * a.nx + T.i;
        const double y = bcgApply<ROWS>(a, T.i, j, c, h, tr, T.lane, hW, hK[k], hE);
        vNew[c] = y;
        const double q = a.rhat[c] * y;
        acc[0] += q;
        acc[1] += fabs(q);
        acc[2] += y * y;
      }
    }
  }
  blockPartials<3>(acc, L.partOut);
}

template <int ROWS>
__global__ __launch_bounds__(256) void bcgSTKernel(Bcg2 a, Bcg2Launch L) {
  __shared__ KState ks;
  __shared__ double h[ROWS][64];
  bcgSteps<3>(a, L, ks, true, [](KState& s, double* hist, const double* t) {
    bcgStepIfRunning(s, hist, kStepBcgV, t);
  });
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const BcgTile<ROWS> T(a);
  if (!ks.done && !ks.restart) {
    const double alpha = ks.alpha;
    const double* v = a.v[1 - L.cur];
    double hK[BcgTile<ROWS>::kRowsPerWave], sK[BcgTile<ROWS>::kRowsPerWave];
#pragma unroll
    for (int k = 0; k < BcgTile<ROWS>::kRowsPerWave; ++k) {
      const int tr = T.row(k);
      double sh = 0.0, si = 0.0;
      if (T.valid(a, tr)) {
        const int64_t c = (T.jBase + tr) * a.nx + T.i;
        si = a.r[c] - (v[c] * alpha);
        sh = a.dinv ? a.dinv[c] * si : si;
        if (T.center(a, tr)) {
          a.s[c] = si;
          a.shat[c] = sh;
          acc[0] += si * si;
        }
      }
      h[tr][T.lane] = sh;
      hK[k] = sh;
      sK[k] = si;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BcgTile<ROWS>::kRowsPerWave; ++k) {
      const int tr = T.row(k);
      const double hW = __shfl_up(hK[k], 1, 64), hE = __shfl_down(hK[k], 1, 64);
      if (T.center(a, tr)) {
        const int64_t j = T.jBase + tr, c = j * a.nx + T.i;
        const double y = bcgApply<ROWS>(a, T.i, j, c, h, tr, T.lane, hW, hK[k], hE);
        a.t[c] = y;
        const double ts = y * sK[k];
        acc[1] += y * y;
        acc[2] += ts;
        acc[3] += fabs(ts);
      }
    }
  }
  blockPartials<4>(acc, L.partOut);
}

template <int ROWS>
__global__ __launch_bounds__(256) void bcgXRKernel(Bcg2 a, Bcg2Launch L) {
  __shared__ KState ks;
  __shared__ int mode;  // 0: nothing, 1: the early x update (s converged), 2: full update
  double t[4];
  bcgSumParts<4>(L.partIn, gridDim.x, t);
  if (threadIdx.x == 0) {
    ks = *L.stIn;
    int m = 0;
    if (!ks.done && !ks.restart) {
      double* hist = blockIdx.x == 0 ? a.hist : nullptr;
      bcgStep(ks, hist, kStepBcgS, t);
      if (!ks.done) {
        // the early update needs only the S step; the T step then finishes
        m = ks.sConverged ? 1 : 2;
        bcgStep(ks, hist, kStepBcgT, t + 1);
        if (m == 2 && (ks.done || ks.restart)) m = 0;
      }
    }
    mode = m;
    if (blockIdx.x == 0) *L.stOut = ks;
  }
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
  const BcgTile<ROWS> T(a);
  if (mode != 0) {
    const double alpha = ks.alpha, omega = ks.omega;
#pragma unroll
    for (int k = 0; k < BcgTile<ROWS>::kRowsPerWave; ++k) {
      const int tr = T.row(k);
      if (!T.center(a, tr)) continue;
      const int64_t c = (T.jBase + tr) * a.nx + T.i;
      a.x[c] += a.phat[c] * alpha;
      if (mode == 1) continue;
      a.x[c] += a.shat[c] * omega;
      const double ri = a.s[c] - (a.t[c] * omega);
      a.r[c] = ri;
      acc[0] += ri * ri;
      const double q = a.rhat[c] * ri;
      acc[1] += q;
      acc[2] += fabs(q);
    }
  }
  blockPartials<3>(acc, L.partOut);
}

// The X and Rho steps of the last xr kernel of a chunk (one block).
__global__ __launch_bounds__(256) void bcgPendingKernel(const double* part, int64_t nParts,
                                                        KState* st, double* hist) {
  double t[3];
  bcgSumParts<3>(part, nParts, t);
  if (threadIdx.x == 0) {
    KState& s = *st;
    bcgStepIfRunning(s, hist, kStepBcgX, t);
    bcgStepIfRunning(s, hist, kStepBcgRho, t + 1);
  }
}

__global__ void zeroKernel(int64_t n, double* a, double* b) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    a[i] = 0.0;
    if (b) b[i] = 0.0;
  }
}

__global__ void copyKernel(int64_t n, const double* __restrict__ src, double* dst) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void nonFiniteKernel(int64_t n, const double* __restrict__ a, unsigned* out) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    if (!isfinite(a[i])) atomicAdd(out, 1u);
}

template <class Op>
__global__ void spmvKernel(Op A, int64_t n, const double* __restrict__ x, double* y) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[i] = A.apply(i, x);
}

// True residual |b - A x| (FP64), written to state.tDotT slot via a reduction.
template <class Op>
__global__ void trueResidualKernel(Op A, int64_t n, const double* __restrict__ b,
                                   const double* __restrict__ x, double* partial) {
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double ri = b[i] - A.apply(i, x);
    acc[0] += ri * ri;
  }
  blockPartials<1>(acc, partial);
}

// ---------------------------------------------------------------------------
// Host drivers.

struct Solve {
  Context& ctx;
  int64_t n;
  Red R;
  KState* hs;
  explicit Solve(Context& c, int64_t nn) : ctx(c), n(nn) {
    KrylovWork& w = ctx.kw;
    w.partial.ensure(kMaxBlocks * 4);
    w.state.ensure(1);
    if (!w.hostState) CFDHIP_CHECK(hipHostMalloc(&w.hostState, sizeof(KState)));
    R = Red{w.partial.get(), w.state.get()};
    hs = w.hostState;
  }
  void begin(const cfdhip_solver_settings& s) {
    require(s.absolute_tolerance >= 0.0 && s.relative_tolerance >= 0.0,
            "LinearSolverSettings: tolerances must be >= 0");
    require(!(s.absolute_tolerance == 0.0 && s.relative_tolerance == 0.0),
            "LinearSolverSettings: at least one of absoluteTolerance/relativeTolerance must be > 0");
    require(s.max_iterations >= 1, "LinearSolverSettings: maxIterations must be >= 1");
    KState k{};
    k.absTol = s.absolute_tolerance;
    k.relTol = s.relative_tolerance;
    k.maxIter = s.max_iterations;
    k.status = -1;
    k.rhoOld = 1.0;
    k.alpha = 1.0;
    k.omega = 1.0;
    *hs = k;
    CFDHIP_CHECK(hipMemcpyAsync(R.st, hs, sizeof(KState), hipMemcpyHostToDevice, ctx.stream));
    ctx.kw.hist.ensure(s.max_iterations + 2);
  }
  // the finalize step after a worker kernel of `grid` blocks
  void fin(int step, dim3 grid) {
    launchFinalize(ctx, FinalArgs{R.partial, static_cast<int64_t>(grid.x), R.st,
                                  ctx.kw.hist.get(), nullptr, step});
  }
  void pull() {
    CFDHIP_CHECK(hipMemcpyAsync(hs, R.st, sizeof(KState), hipMemcpyDeviceToHost, ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  }
  void push() {
    CFDHIP_CHECK(hipMemcpyAsync(R.st, hs, sizeof(KState), hipMemcpyHostToDevice, ctx.stream));
  }
  dim3 grid() const { return gridStride(n); }
  cfdhip_solve_result result(double trueRes, double ms) {
    cfdhip_solve_result r{};
    r.status = hs->status < 0 ? CFDHIP_MAX_ITERATIONS : hs->status;
    r.iterations = hs->iter;
    r.restarts = hs->restarts;
    r.initial_residual = hs->b0;
    r.final_residual = hs->rnorm;
    r.true_residual = trueRes;
    r.device_ms = ms;
    r.history_length = hs->histLen;
    if (ctx.inLoop) {
      ctx.lastHistory.clear();
      return r;
    }
    ctx.lastHistory.resize(hs->histLen);
    if (hs->histLen > 0)
      CFDHIP_CHECK(hipMemcpy(ctx.lastHistory.data(), ctx.kw.hist.get(),
                             hs->histLen * sizeof(double), hipMemcpyDeviceToHost));
    return r;
  }
};

template <class Op>
bool inputsFinite(Context& ctx, Op&, const std::vector<std::pair<const double*, int64_t>>& arrays) {
  if (ctx.inLoop) return true;
  for (const auto& a : arrays) {
    if (a.second <= 0 || a.first == nullptr) continue;
    hipLaunchKernelGGL(nonFiniteKernel, gridStride(a.second), dim3(kBlock), 0, ctx.stream,
                       a.second, a.first, ctx.flags.get() + 1);
  }
  unsigned h = 0;
  CFDHIP_CHECK(hipMemcpyAsync(&h, ctx.flags.get() + 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  CFDHIP_CHECK(hipMemsetAsync(ctx.flags.get() + 1, 0, sizeof(unsigned), ctx.stream));
  return h == 0;
}

template <class Op>
bool buildJacobi(Context& ctx, Op A, int64_t n, double* dinv) {
  hipLaunchKernelGGL((jacobiKernel<Op>), gridStride(n), dim3(kBlock), 0, ctx.stream, A, n, dinv,
                     ctx.flags.get() + 1);
  CFDHIP_CHECK(hipGetLastError());
  unsigned h = 0;
  CFDHIP_CHECK(hipMemcpyAsync(&h, ctx.flags.get() + 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  CFDHIP_CHECK(hipMemsetAsync(ctx.flags.get() + 1, 0, sizeof(unsigned), ctx.stream));
  return h == 0;
}

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    CFDHIP_CHECK(hipEventCreate(&a));
    CFDHIP_CHECK(hipEventCreate(&b));
    CFDHIP_CHECK(hipEventRecord(a, s));
  }
  double stop() {
    CFDHIP_CHECK(hipEventRecord(b, s));
    CFDHIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CFDHIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

template <class Op>
double trueResidual(Context& ctx, Op A, int64_t n, const double* b, const double* x) {
  if (ctx.inLoop) return NAN;
  KrylovWork& w = ctx.kw;
  w.scal.ensure(4);
  w.partial.ensure(kMaxBlocks * 4);
  w.state.ensure(1);
  double* out = w.scal.get() + 1;
  const dim3 g = gridStride(n);
  hipLaunchKernelGGL((trueResidualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x,
                     w.partial.get());
  CFDHIP_CHECK(hipGetLastError());
  launchFinalize(ctx, FinalArgs{w.partial.get(), static_cast<int64_t>(g.x), w.state.get(),
                                nullptr, out, kStepNorm});
  double h = 0.0;
  CFDHIP_CHECK(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return h;
}

// CG with optional Jacobi; b, x device pointers (x holds x0 on entry).
template <class Op>
cfdhip_solve_result runCG(Context& ctx, Op A, int64_t n, const cfdhip_solver_settings& set,
                          const double* b, double* x,
                          const std::vector<std::pair<const double*, int64_t>>& opArrays) {
  KrylovWork& w = ctx.kw;
  w.r.ensure(n);
  w.p.ensure(n);
  w.Ap.ensure(n);
  w.dinv.ensure(n);
  Solve S(ctx, n);
  S.begin(set);
  auto arrays = opArrays;
  arrays.push_back({b, n});
  arrays.push_back({x, n});
  if (!inputsFinite(ctx, A, arrays)) {
    S.hs->status = CFDHIP_NONFINITE_INPUT;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  const bool jacobi = set.preconditioner == CFDHIP_PREC_JACOBI;
  if (jacobi && !buildJacobi(ctx, A, n, w.dinv.get())) {
    S.hs->status = CFDHIP_INVALID_SYSTEM;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  const double* dinv = jacobi ? w.dinv.get() : nullptr;
  Timer timer(ctx.stream);
  const dim3 g = S.grid();
  hipLaunchKernelGGL((residualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x, w.r.get(),
                     nullptr, S.R, 0);
  S.fin(kStepStartNorm, g);
  hipLaunchKernelGGL(cgStartKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(), dinv, w.p.get(),
                     S.R);
  S.fin(kStepCgStartRz, g);
  CFDHIP_CHECK(hipGetLastError());
  const int chunk = set.check_interval > 0 ? set.check_interval : 32;
  int64_t launched = 0;
  while (true) {
    for (int k = 0; k < chunk && launched < set.max_iterations; ++k, ++launched) {
      hipLaunchKernelGGL((cgSpmvKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, w.p.get(),
                         w.Ap.get(), S.R);
      S.fin(kStepCgPAp, g);
      hipLaunchKernelGGL(cgUpdateKernel, g, dim3(kBlock), 0, ctx.stream, n, x, w.r.get(),
                         w.p.get(), w.Ap.get(), dinv, S.R);
      S.fin(kStepCgUpdate, g);
      hipLaunchKernelGGL(cgDirectionKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(), dinv,
                         w.p.get(), S.R.st);
    }
    CFDHIP_CHECK(hipGetLastError());
    S.pull();
    if (S.hs->done || launched >= set.max_iterations) break;
  }
  const double ms = timer.stop();
  if (!S.hs->done) {
    S.hs->status = CFDHIP_MAX_ITERATIONS;
    S.hs->done = 1;
  }
  const double tr = trueResidual(ctx, A, n, b, x);
  return S.result(tr, ms);
}

template <class Op>
cfdhip_solve_result runBiCGSTAB(Context& ctx, Op A, int64_t n, const cfdhip_solver_settings& set,
                                const double* b, double* x,
                                const std::vector<std::pair<const double*, int64_t>>& opArrays) {
  KrylovWork& w = ctx.kw;
  for (DeviceArray<double>* a : {&w.r, &w.p, &w.rhat, &w.v, &w.s, &w.t, &w.phat, &w.shat, &w.dinv})
    a->ensure(n);
  Solve S(ctx, n);
  S.begin(set);
  auto arrays = opArrays;
  arrays.push_back({b, n});
  arrays.push_back({x, n});
  if (!inputsFinite(ctx, A, arrays)) {
    S.hs->status = CFDHIP_NONFINITE_INPUT;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  const bool jacobi = set.preconditioner == CFDHIP_PREC_JACOBI;
  if (jacobi && !buildJacobi(ctx, A, n, w.dinv.get())) {
    S.hs->status = CFDHIP_INVALID_SYSTEM;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  const double* dinv = jacobi ? w.dinv.get() : nullptr;
  Timer timer(ctx.stream);
  const dim3 g = S.grid();
  hipLaunchKernelGGL((residualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x, w.r.get(),
                     w.rhat.get(), S.R, 0);
  S.fin(kStepStartNorm, g);
  hipLaunchKernelGGL(zeroKernel, g, dim3(kBlock), 0, ctx.stream, n, w.v.get(), w.p.get());
  CFDHIP_CHECK(hipGetLastError());
  const int chunk = set.check_interval > 0 ? set.check_interval : 8;
  int64_t launched = 0;
  // A restart does not complete an iteration but still consumes launch
  // budget; bound the total launches generously.
  const int64_t launchCap = 4 * set.max_iterations + 64;
  while (true) {
    for (int k = 0; k < chunk && launched < launchCap; ++k, ++launched) {
      hipLaunchKernelGGL(bcgRhoKernel, g, dim3(kBlock), 0, ctx.stream, n, w.rhat.get(), w.r.get(),
                         S.R);
      S.fin(kStepBcgRho, g);
      hipLaunchKernelGGL(bcgDirectionKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(),
                         w.v.get(), dinv, w.p.get(), w.phat.get(), S.R.st);
      hipLaunchKernelGGL((bcgVKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, w.phat.get(),
                         w.rhat.get(), w.v.get(), S.R);
      S.fin(kStepBcgV, g);
      hipLaunchKernelGGL(bcgSKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(), w.v.get(), dinv,
                         w.s.get(), w.shat.get(), S.R);
      S.fin(kStepBcgS, g);
      hipLaunchKernelGGL((bcgTKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, w.shat.get(),
                         w.s.get(), w.phat.get(), w.t.get(), x, S.R);
      S.fin(kStepBcgT, g);
      hipLaunchKernelGGL(bcgXKernel, g, dim3(kBlock), 0, ctx.stream, n, x, w.r.get(), w.phat.get(),
                         w.shat.get(), w.s.get(), w.t.get(), S.R);
      S.fin(kStepBcgX, g);
    }
    CFDHIP_CHECK(hipGetLastError());
    S.pull();
    if (S.hs->done) break;
    if (S.hs->restart == 1) {
      // BiCGSTAB.cpp:128-130: restart only after progress in this sequence.
      if (!(S.hs->rnorm < S.hs->restartResidual)) {
        S.hs->status = CFDHIP_BREAKDOWN;
        S.hs->done = 1;
        S.hs->iter = S.hs->completed;
        break;
      }
      hipLaunchKernelGGL((residualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x,
                         w.r.get(), w.rhat.get(), S.R, 1);
      S.fin(kStepBcgRestart, g);
      hipLaunchKernelGGL(zeroKernel, g, dim3(kBlock), 0, ctx.stream, n, w.v.get(), w.p.get());
      CFDHIP_CHECK(hipGetLastError());
      S.pull();
      if (S.hs->done) {
        S.hs->iter = S.hs->completed;
        break;
      }
      S.hs->restart = 0;
      S.push();
    }
    if (launched >= launchCap) break;
  }
  const double ms = timer.stop();
  if (!S.hs->done) {
    S.hs->status = CFDHIP_MAX_ITERATIONS;
    S.hs->done = 1;
  }
  if (S.hs->status == CFDHIP_BREAKDOWN) S.hs->iter = S.hs->completed;
  const double tr = trueResidual(ctx, A, n, b, x);
  return S.result(tr, ms);
}

// BiCGSTAB (BiCGSTAB.cpp:48-275) on the 2D momentum stencil of a small grid
// with the fused kernels: the same recurrence, statuses, restarts and
// history as runBiCGSTAB, three launches per iteration. A chunk of
// check_interval iterations starts from a state whose Rho step is done
// (bcgRhoKernel + finalize at the start and after a restart), runs its
// iterations, and ends with the pending X and Rho steps in one block; the
// state then lives in S[0] and the host reads it once per chunk.
cfdhip_solve_result runBiCGSTABFused2d(Context& ctx, const StencilOp& A, int rows, int64_t tiles,
                                       const cfdhip_solver_settings& set, const double* b,
                                       double* x,
                                       const std::vector<std::pair<const double*, int64_t>>& opArrays) {
  const View& vw = ctx.view;
  const int64_t n = vw.n;
  KrylovWork& w = ctx.kw;
  for (DeviceArray<double>* a : {&w.r, &w.p, &w.rhat, &w.v, &w.s, &w.t, &w.phat, &w.shat,
                                 &w.dinv, &w.Ap, &w.z})
    a->ensure(n);
  w.state.ensure(2);  // before Solve takes the address of state 0
  w.partial.ensure(kMaxBlocks * 4);
  Solve S(ctx, n);
  S.begin(set);
  auto arrays = opArrays;
  arrays.push_back({b, n});
  arrays.push_back({x, n});
  if (!inputsFinite(ctx, A, arrays)) {
    S.hs->status = CFDHIP_NONFINITE_INPUT;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  const bool jacobi = set.preconditioner == CFDHIP_PREC_JACOBI;
  if (jacobi && !buildJacobi(ctx, A, n, w.dinv.get())) {
    S.hs->status = CFDHIP_INVALID_SYSTEM;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  KState* St[2] = {S.R.st, S.R.st + 1};
  double* P[2] = {w.partial.get(), w.partial.get() + 4 * kBcgMaxTiles};
  // p and v alternate between (p, Ap) and (v, z)
  Bcg2 a{vw.nx,     vw.ny,  n,         (vw.nx + kBcgOut - 1) / kBcgOut,
         A.coef,    jacobi ? w.dinv.get() : nullptr,   w.rhat.get(),
         w.r.get(), x,      {w.p.get(), w.Ap.get()},   {w.v.get(), w.z.get()},
         w.phat.get(), w.s.get(), w.shat.get(), w.t.get(), ctx.kw.hist.get()};
  Timer timer(ctx.stream);
  const dim3 g = S.grid();
  const dim3 gt(static_cast<unsigned>(tiles));
  int cur = 0;
  auto zeroPV = [&]() {
    hipLaunchKernelGGL(zeroKernel, g, dim3(kBlock), 0, ctx.stream, n, a.v[cur], a.p[cur]);
  };
  auto rho = [&]() {
    hipLaunchKernelGGL(bcgRhoKernel, g, dim3(kBlock), 0, ctx.stream, n, w.rhat.get(), w.r.get(),
                       S.R);
    S.fin(kStepBcgRho, g);
  };
  hipLaunchKernelGGL((residualKernel<StencilOp>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x,
                     w.r.get(), w.rhat.get(), S.R, 0);
  S.fin(kStepStartNorm, g);
  zeroPV();
  rho();
  CFDHIP_CHECK(hipGetLastError());
  const int chunk = set.check_interval > 0 ? set.check_interval : 8;
  int64_t launched = 0;
  const int64_t launchCap = 4 * set.max_iterations + 64;
#define CFDHIP_BCG_LAUNCH(KERNEL, LA)                                                        \
  switch (rows) {                                                                             \
    case 4: hipLaunchKernelGGL((KERNEL<4>), gt, dim3(256), 0, ctx.stream, a, LA); break;      \
    case 8: hipLaunchKernelGGL((KERNEL<8>), gt, dim3(256), 0, ctx.stream, a, LA); break;      \
    case 16: hipLaunchKernelGGL((KERNEL<16>), gt, dim3(256), 0, ctx.stream, a, LA); break;    \
    default: hipLaunchKernelGGL((KERNEL<32>), gt, dim3(256), 0, ctx.stream, a, LA); break;    \
  }
  while (true) {
    // launch q of the chunk reads state St[q & 1] and writes St[(q + 1) & 1],
    // reads partials P[(q - 1) & 1] and writes P[q & 1]
    int q = 0;
    for (int k = 0; k < chunk && launched < launchCap; ++k, ++launched) {
      const Bcg2Launch l1{cur, St[q & 1], St[(q + 1) & 1], P[(q + 1) & 1], P[q & 1], k > 0};
      CFDHIP_BCG_LAUNCH(bcgDirVKernel, l1);
      ++q;
      const Bcg2Launch l2{cur, St[q & 1], St[(q + 1) & 1], P[(q + 1) & 1], P[q & 1], 0};
      CFDHIP_BCG_LAUNCH(bcgSTKernel, l2);
      ++q;
      const Bcg2Launch l3{cur, St[q & 1], St[(q + 1) & 1], P[(q + 1) & 1], P[q & 1], 0};
      CFDHIP_BCG_LAUNCH(bcgXRKernel, l3);
      ++q;
      cur = 1 - cur;
    }
    // the last xr's steps, into St[0]
    if (q > 0) {
      if ((q & 1) != 0)
        CFDHIP_CHECK(hipMemcpyAsync(St[0], St[1], sizeof(KState), hipMemcpyDeviceToDevice,
                                    ctx.stream));
      hipLaunchKernelGGL(bcgPendingKernel, dim3(1), dim3(256), 0, ctx.stream, P[(q - 1) & 1],
                         static_cast<int64_t>(tiles), St[0], ctx.kw.hist.get());
    }
    CFDHIP_CHECK(hipGetLastError());
    S.pull();
    if (S.hs->done) break;
    if (S.hs->restart == 1) {
      // BiCGSTAB.cpp:128-130: restart only after progress in this sequence.
      if (!(S.hs->rnorm < S.hs->restartResidual)) {
        S.hs->status = CFDHIP_BREAKDOWN;
        S.hs->done = 1;
        S.hs->iter = S.hs->completed;
        break;
      }
      hipLaunchKernelGGL((residualKernel<StencilOp>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x,
                         w.r.get(), w.rhat.get(), S.R, 1);
      S.fin(kStepBcgRestart, g);
      zeroPV();
      CFDHIP_CHECK(hipGetLastError());
      S.pull();
      if (S.hs->done) {
        S.hs->iter = S.hs->completed;
        break;
      }
      S.hs->restart = 0;
      S.push();
      rho();
    }
    if (launched >= launchCap) break;
  }
#undef CFDHIP_BCG_LAUNCH
  const double ms = timer.stop();
  if (!S.hs->done) {
    S.hs->status = CFDHIP_MAX_ITERATIONS;
    S.hs->done = 1;
  }
  if (S.hs->status == CFDHIP_BREAKDOWN) S.hs->iter = S.hs->completed;
  const double tr = trueResidual(ctx, A, n, b, x);
  return S.result(tr, ms);
}

// The fused form's tiles for a 2D grid: the shortest rows (4, 8, 16 or 32)
// whose tile count fits kBcgMaxTiles; rows = 0 when none does.
void bcgFusedTiles(const View& v, int& rows, int64_t& tiles) {
  rows = 0;
  tiles = 0;
  if (v.dim != 2) return;
  const int64_t tilesX = (v.nx + kBcgOut - 1) / kBcgOut;
  for (int r : {4, 8, 16, 32}) {
    const int64_t t = tilesX * ((v.ny + r - 3) / (r - 2));
    if (t <= kBcgMaxTiles) {
      rows = r;
      tiles = t;
      return;
    }
  }
}

// out = M in through the caller's host preconditioner: the one vector crosses
// to the host and back per application, as in the reference's GPU solvers
// (cuda/kernels/GpuLinearSolverCuda.cpp:103-118).
void hostApply(Context& ctx, const double* in, double* out, int64_t n) {
  ctx.hostPcIn.resize(n);
  ctx.hostPcOut.assign(n, 0.0);
  CFDHIP_CHECK(hipMemcpyAsync(ctx.hostPcIn.data(), in, n * sizeof(double), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  if (ctx.hostPc(ctx.hostPcIn.data(), ctx.hostPcOut.data(), n, ctx.hostPcUser) != 0)
    throw Error(CFDHIP_ERR_RUNTIME, "the host preconditioner failed");
  CFDHIP_CHECK(hipMemcpyAsync(out, ctx.hostPcOut.data(), n * sizeof(double),
                              hipMemcpyHostToDevice, ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));  // the host buffer is reused
}

// CG (CG.cpp:59-147) with the caller's host preconditioner: the state is read
// after every iteration (the host is in the loop anyway), so M is not applied
// once the solve is done.
template <class Op>
cfdhip_solve_result runCGHost(Context& ctx, Op A, int64_t n, const cfdhip_solver_settings& set,
                              const double* b, double* x,
                              const std::vector<std::pair<const double*, int64_t>>& opArrays) {
  require(ctx.hostPc != nullptr,
          "CFDHIP_PREC_HOST: no host preconditioner set (cfdhip_set_host_preconditioner)");
  KrylovWork& w = ctx.kw;
  w.r.ensure(n);
  w.p.ensure(n);
  w.Ap.ensure(n);
  w.z.ensure(n);
  Solve S(ctx, n);
  S.begin(set);
  auto arrays = opArrays;
  arrays.push_back({b, n});
  arrays.push_back({x, n});
  if (!inputsFinite(ctx, A, arrays)) {
    S.hs->status = CFDHIP_NONFINITE_INPUT;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  Timer timer(ctx.stream);
  const dim3 g = S.grid();
  hipLaunchKernelGGL((residualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x, w.r.get(),
                     nullptr, S.R, 0);
  S.fin(kStepStartNorm, g);
  S.pull();
  if (!S.hs->done) {
    hostApply(ctx, w.r.get(), w.z.get(), n);
    hipLaunchKernelGGL(cgStartZKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(), w.z.get(),
                       w.p.get(), S.R);
    S.fin(kStepCgStartRz, g);
    for (int64_t it = 0; it < set.max_iterations; ++it) {
      hipLaunchKernelGGL((cgSpmvKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, w.p.get(),
                         w.Ap.get(), S.R);
      S.fin(kStepCgPAp, g);
      hipLaunchKernelGGL(cgUpdateXRKernel, g, dim3(kBlock), 0, ctx.stream, n, x, w.r.get(),
                         w.p.get(), w.Ap.get(), S.R.st);
      CFDHIP_CHECK(hipGetLastError());
      S.pull();
      if (S.hs->done) break;
      hostApply(ctx, w.r.get(), w.z.get(), n);
      hipLaunchKernelGGL(rzKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(), w.z.get(), S.R);
      S.fin(kStepCgUpdate, g);
      hipLaunchKernelGGL(cgDirectionZKernel, g, dim3(kBlock), 0, ctx.stream, n, w.z.get(),
                         w.p.get(), S.R.st);
      CFDHIP_CHECK(hipGetLastError());
    }
    S.pull();
  }
  const double ms = timer.stop();
  if (!S.hs->done) {
    S.hs->status = CFDHIP_MAX_ITERATIONS;
    S.hs->done = 1;
  }
  const double tr = trueResidual(ctx, A, n, b, x);
  return S.result(tr, ms);
}

// BiCGSTAB (BiCGSTAB.cpp:48-275) with the caller's host preconditioner:
// phat = M p and shat = M s cross to the host, the rest is the device
// recurrence of runBiCGSTAB (the kernels form phat = p, shat = s, which the
// host result then replaces).
template <class Op>
cfdhip_solve_result runBiCGSTABHost(Context& ctx, Op A, int64_t n,
                                    const cfdhip_solver_settings& set, const double* b, double* x,
                                    const std::vector<std::pair<const double*, int64_t>>& opArrays) {
  require(ctx.hostPc != nullptr,
          "CFDHIP_PREC_HOST: no host preconditioner set (cfdhip_set_host_preconditioner)");
  KrylovWork& w = ctx.kw;
  for (DeviceArray<double>* a : {&w.r, &w.p, &w.rhat, &w.v, &w.s, &w.t, &w.phat, &w.shat})
    a->ensure(n);
  Solve S(ctx, n);
  S.begin(set);
  auto arrays = opArrays;
  arrays.push_back({b, n});
  arrays.push_back({x, n});
  if (!inputsFinite(ctx, A, arrays)) {
    S.hs->status = CFDHIP_NONFINITE_INPUT;
    S.hs->done = 1;
    return S.result(NAN, 0.0);
  }
  Timer timer(ctx.stream);
  const dim3 g = S.grid();
  hipLaunchKernelGGL((residualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x, w.r.get(),
                     w.rhat.get(), S.R, 0);
  S.fin(kStepStartNorm, g);
  hipLaunchKernelGGL(zeroKernel, g, dim3(kBlock), 0, ctx.stream, n, w.v.get(), w.p.get());
  CFDHIP_CHECK(hipGetLastError());
  const int64_t launchCap = 4 * set.max_iterations + 64;
  for (int64_t launched = 0; launched < launchCap; ++launched) {
    S.pull();
    if (S.hs->done) break;
    if (S.hs->restart == 1) {
      // BiCGSTAB.cpp:128-130: restart only after progress in this sequence.
      if (!(S.hs->rnorm < S.hs->restartResidual)) {
        S.hs->status = CFDHIP_BREAKDOWN;
        S.hs->done = 1;
        S.hs->iter = S.hs->completed;
        break;
      }
      hipLaunchKernelGGL((residualKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, b, x,
                         w.r.get(), w.rhat.get(), S.R, 1);
      S.fin(kStepBcgRestart, g);
      hipLaunchKernelGGL(zeroKernel, g, dim3(kBlock), 0, ctx.stream, n, w.v.get(), w.p.get());
      CFDHIP_CHECK(hipGetLastError());
      S.pull();
      if (S.hs->done) {
        S.hs->iter = S.hs->completed;
        break;
      }
      S.hs->restart = 0;
      S.push();
    }
    hipLaunchKernelGGL(bcgRhoKernel, g, dim3(kBlock), 0, ctx.stream, n, w.rhat.get(), w.r.get(),
                       S.R);
    S.fin(kStepBcgRho, g);
    hipLaunchKernelGGL(bcgDirectionKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(),
                       w.v.get(), nullptr, w.p.get(), w.phat.get(), S.R.st);
    S.pull();
    if (S.hs->done || S.hs->restart == 1) continue;
    hostApply(ctx, w.p.get(), w.phat.get(), n);
    hipLaunchKernelGGL((bcgVKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, w.phat.get(),
                       w.rhat.get(), w.v.get(), S.R);
    S.fin(kStepBcgV, g);
    hipLaunchKernelGGL(bcgSKernel, g, dim3(kBlock), 0, ctx.stream, n, w.r.get(), w.v.get(),
                       nullptr, w.s.get(), w.shat.get(), S.R);
    S.fin(kStepBcgS, g);
    S.pull();
    if (S.hs->done || S.hs->restart == 1) continue;
    hostApply(ctx, w.s.get(), w.shat.get(), n);
    hipLaunchKernelGGL((bcgTKernel<Op>), g, dim3(kBlock), 0, ctx.stream, A, n, w.shat.get(),
                       w.s.get(), w.phat.get(), w.t.get(), x, S.R);
    S.fin(kStepBcgT, g);
    hipLaunchKernelGGL(bcgXKernel, g, dim3(kBlock), 0, ctx.stream, n, x, w.r.get(), w.phat.get(),
                       w.shat.get(), w.s.get(), w.t.get(), S.R);
    S.fin(kStepBcgX, g);
    CFDHIP_CHECK(hipGetLastError());
  }
  S.pull();
  const double ms = timer.stop();
  if (!S.hs->done) {
    S.hs->status = CFDHIP_MAX_ITERATIONS;
    S.hs->done = 1;
  }
  if (S.hs->status == CFDHIP_BREAKDOWN) S.hs->iter = S.hs->completed;
  const double tr = trueResidual(ctx, A, n, b, x);
  return S.result(tr, ms);
}

template <class Op>
cfdhip_solve_result dispatch(Context& ctx, Op A, int64_t n, const cfdhip_solver_settings& s,
                             const double* b, double* x,
                             const std::vector<std::pair<const double*, int64_t>>& arrays) {
  const bool host = s.preconditioner == CFDHIP_PREC_HOST;
  if (s.type == CFDHIP_SOLVER_CG)
    return host ? runCGHost(ctx, A, n, s, b, x, arrays) : runCG(ctx, A, n, s, b, x, arrays);
  if (s.type == CFDHIP_SOLVER_BICGSTAB)
    return host ? runBiCGSTABHost(ctx, A, n, s, b, x, arrays)
                : runBiCGSTAB(ctx, A, n, s, b, x, arrays);
  throw Error(CFDHIP_ERR_INVALID, "unsupported linear solver type for this operator");
}

PressureOp pressureOp(Context& ctx) {
  return PressureOp{ctx.view, ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF),
                    ctx.pin};
}

}  // namespace

cfdhip_solve_result solvePressure(Context& ctx, const cfdhip_solver_settings& s, const double* b,
                                  double* x, bool zeroGuess) {
  ctx.requireMesh();
  require(ctx.pressureAssembled, "no p' system assembled");
  const int64_t n = ctx.view.n;
  if (zeroGuess) {
    hipLaunchKernelGGL(zeroKernel, gridStride(n), dim3(kBlock), 0, ctx.stream, n, x, nullptr);
    CFDHIP_CHECK(hipGetLastError());
  }
  const std::vector<std::pair<const double*, int64_t>> opArrays = {
      {ctx.f(CFDHIP_FIELD_PC_DIAG), n}, {ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), ctx.view.nf}};
  if (s.preconditioner == CFDHIP_PREC_LEARNED) {
    require(ctx.view.dim == 2 && zeroGuess &&
                (s.type == CFDHIP_SOLVER_FCG || s.type == CFDHIP_SOLVER_CG),
            "the learned preconditioner runs flexible CG on the 2D p' operator from x0 = 0",
            CFDHIP_ERR_UNSUPPORTED);
    PressureOp A = pressureOp(ctx);
    KrylovWork& w = ctx.kw;
    auto arrays = opArrays;
    arrays.push_back({b, n});
    cfdhip_solve_result res{};
    w.dinv.ensure(n);
    if (!inputsFinite(ctx, A, arrays)) {
      res.status = CFDHIP_NONFINITE_INPUT;
      ctx.lastHistory.clear();
      return res;
    }
    if (!buildJacobi(ctx, A, n, w.dinv.get())) {  // the scaling needs a positive diagonal
      res.status = CFDHIP_INVALID_SYSTEM;
      ctx.lastHistory.clear();
      return res;
    }
    return learnedFcg2d(ctx, s, b, x);
  }
  const bool fused = s.type == CFDHIP_SOLVER_CG && s.preconditioner == CFDHIP_PREC_JACOBI;
  if (s.precision == CFDHIP_MIXED)
    require(fused && zeroGuess,
            "mixed precision is implemented for the p' operator with CG + Jacobi from x0 = 0",
            CFDHIP_ERR_UNSUPPORTED);
  if (!fused) return dispatch(ctx, pressureOp(ctx), n, s, b, x, opArrays);
  // fused 2D Jacobi-PCG: same input checks and Jacobi rejection as runCG
  PressureOp A = pressureOp(ctx);
  KrylovWork& w = ctx.kw;
  auto arrays = opArrays;
  arrays.push_back({b, n});
  arrays.push_back({x, n});
  cfdhip_solve_result res{};
  w.dinv.ensure(n);
  if (!inputsFinite(ctx, A, arrays)) {
    res.status = CFDHIP_NONFINITE_INPUT;
    ctx.lastHistory.clear();
    return res;
  }
  if (!buildJacobi(ctx, A, n, w.dinv.get())) {
    res.status = CFDHIP_INVALID_SYSTEM;
    ctx.lastHistory.clear();
    return res;
  }
  if (s.precision == CFDHIP_MIXED)
    return ctx.view.dim == 2 ? mixedPcg2d(ctx, s, b, x) : mixedPcg3d(ctx, s, b, x);
  double* work[8];
  for (int q = 0; q < 8; ++q) {
    w.d8[q].ensure(n);
    work[q] = w.d8[q].get();
  }
  if (ctx.view.dim == 2)
    res = singleKernelPcg2d<double>(ctx, s, ctx.f(CFDHIP_FIELD_PC_DIAG),
                                    ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b, x, work);
  else
    res = pcg3d<double>(ctx, s, ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), b,
                        x, work);
  res.true_residual = trueResidual(ctx, A, n, b, x);
  return res;
}

cfdhip_solve_result solveStencil(Context& ctx, const cfdhip_solver_settings& s, const double* b,
                                 double* x) {
  ctx.requireMesh();
  require(ctx.momentumAssembled, "no momentum system assembled");
  const int64_t n = ctx.view.n;
  StencilOp A{ctx.view, ctx.momCoef.get()};
  const std::vector<std::pair<const double*, int64_t>> arrays = {
      {ctx.momCoef.get(), (2 * ctx.view.dim + 1) * n}};
  // small 2D grids: the fused BiCGSTAB (three launches per iteration)
  int rows = 0;
  int64_t tiles = 0;
  bcgFusedTiles(ctx.view, rows, tiles);
  if (s.type == CFDHIP_SOLVER_BICGSTAB && rows > 0 &&
      (s.preconditioner == CFDHIP_PREC_NONE || s.preconditioner == CFDHIP_PREC_JACOBI))
    return runBiCGSTABFused2d(ctx, A, rows, tiles, s, b, x, arrays);
  return dispatch(ctx, A, n, s, b, x, arrays);
}

cfdhip_solve_result solveCsr(Context& ctx, const cfdhip_solver_settings& s, const CsrDevice& M,
                             const double* b, double* x) {
  CsrOp A{M.n, M.rp.get(), M.col.get(), M.val.get()};
  return dispatch(ctx, A, M.n, s, b, x, {{M.val.get(), M.val.count}});
}

namespace {

// y = A x for the 2D p' operator (SparseMatrix::multiply, SparseMatrix.cpp:72-100,
// on the operator of PressureCorrectionEquation.cpp:181-339): one lane per
// column, 64 columns x 16 rows per block (4 rows per thread), ascending column
// order from 0 (south, west, centre, east, north) as the CSR row holds them.
// Streams diag, the two face-coefficient arrays, x and y once each.
constexpr int kSpmvRows = 16;

__global__ __launch_bounds__(256) void spmv2dKernel(int64_t nx, int64_t ny, int64_t pin,
                                                    int64_t offY, const double* __restrict__ diag,
                                                    const double* __restrict__ coeff,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 64 + (threadIdx.x & 63);
  const int64_t jb = static_cast<int64_t>(blockIdx.y) * kSpmvRows + (threadIdx.x >> 6);
  if (i >= nx) return;
#pragma unroll
  for (int k = 0; k < kSpmvRows / 4; ++k) {
    const int64_t j = jb + 4 * k;
    if (j >= ny) break;
    const int64_t c = j * nx + i;
    double v;
    if (c == pin) {
      v = 0.0 + diag[c] * x[c];
    } else {
      v = 0.0;
      if (j > 0) v += (-coeff[offY + j * nx + i]) * x[c - nx];
      if (i > 0) v += (-coeff[j * (nx + 1) + i]) * x[c - 1];
      v += diag[c] * x[c];
      if (i < nx - 1) v += (-coeff[j * (nx + 1) + i + 1]) * x[c + 1];
      if (j < ny - 1) v += (-coeff[offY + (j + 1) * nx + i]) * x[c + nx];
    }
    y[c] = v;
  }
}

}  // namespace

void pressureSpmv(Context& ctx, const double* x, double* y) {
  ctx.requireMesh();
  require(ctx.pressureAssembled, "no p' system assembled");
  const View& v = ctx.view;
  if (v.dim == 2) {
    const dim3 grid(static_cast<unsigned>((v.nx + 63) / 64),
                    static_cast<unsigned>((v.ny + kSpmvRows - 1) / kSpmvRows));
    hipLaunchKernelGGL(spmv2dKernel, grid, dim3(256), 0, ctx.stream, v.nx, v.ny, ctx.pin, v.offY,
                       ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), x, y);
  } else {
    hipLaunchKernelGGL((spmvKernel<PressureOp>), gridStride(v.n), dim3(kBlock), 0, ctx.stream,
                       pressureOp(ctx), v.n, x, y);
  }
  CFDHIP_CHECK(hipGetLastError());
}

void csrSpmv(Context& ctx, const CsrDevice& M, const double* x, double* y) {
  CsrOp A{M.n, M.rp.get(), M.col.get(), M.val.get()};
  hipLaunchKernelGGL((spmvKernel<CsrOp>), gridStride(M.n), dim3(kBlock), 0, ctx.stream, A, M.n, x,
                     y);
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace cfdhip
