// The no-allocation p' solve: cfdhip_pressure_solve_rhs's FP64 Jacobi-PCG
// (single-kernel iteration, cg1.hip) on the resident 2D p' operator, with the
// right-hand side and the solution in the caller's device memory and every
// work buffer in one fixed, file-scope device array, so that the call makes
// no device allocation (no hipMalloc / hipFree, no pinned host memory, no
// HIP events or graphs).
//
// The array lives once per GPU for the life of the process. Host code never
// passes its name to a kernel: on the host the name stands for a host-side
// symbol, not a device address. The device address is looked up with
// hipGetSymbolAddress once per device, with that device current, and only
// that pointer is used afterwards. Every call on one GPU shares the array (all
// contexts on that GPU), so calls that overlap on one GPU race; each call
// starts by clearing what must start at zero (the check counters, the
// solution) with fill kernels.
#include <array>
#include <cmath>
#include <mutex>

#include "reduce.hpp"

struct cfdhip_ctx {
  cfdhip::Context c;
};

namespace cfdhip {

namespace {

constexpr int64_t kNaMaxCells = int64_t(1) << 20;  // 1024 x 1024 cells
constexpr int64_t kNaMaxIter = int64_t(1) << 16;
// iteration-kernel partials (3 per 62 x 30 tile: at most 3 (n / 30 + 2)
// for n cells) and start partials (one per 256 cells)
constexpr int64_t kNaPartials = 131072;
constexpr int64_t kNaHist = kNaMaxIter + 2;
constexpr int64_t kNaStateDoubles = 2 * ((sizeof(KState) + 7) / 8) + 8;  // two states
// layout in doubles: 8 work vectors | partials | history | state | counters + norm
constexpr int64_t kNaWork = 8 * kNaMaxCells;
constexpr int64_t kNaOffPartial = kNaWork;
constexpr int64_t kNaOffHist = kNaOffPartial + kNaPartials;
constexpr int64_t kNaOffState = kNaOffHist + kNaHist;
constexpr int64_t kNaOffMisc = kNaOffState + kNaStateDoubles;
constexpr int64_t kNaDoubles = kNaOffMisc + 8;
static_assert(kNaPartials >= 3 * (kNaMaxCells / 30 + 2) && kNaPartials >= kNaMaxCells / 256 + 8,
              "partials must hold the largest grid");

}  // namespace

// The scratch: 67.6 MB of device memory per GPU, from the moment the
// library's code object is loaded on that GPU until the process ends.
static __device__ __align__(256) double scratch[kNaDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::array<double*, kMaxDevices> g_scratch{};
std::array<std::once_flag, kMaxDevices> g_once;
std::array<hipError_t, kMaxDevices> g_err{};

double* scratchOf(int device) {
  require(device >= 0 && device < kMaxDevices, "device index out of range");
  std::call_once(g_once[device], [device] {
    // the caller made `device` current
    void* ptr = nullptr;
    g_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (g_err[device] == hipSuccess) g_scratch[device] = static_cast<double*>(ptr);
  });
  if (g_err[device] != hipSuccess || g_scratch[device] == nullptr)
    throw Error(CFDHIP_ERR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(g_err[device]));
  return g_scratch[device];
}

__global__ void fillKernel(int64_t n, double* __restrict__ a, double v) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[k] = v;
}

// counters[0]: non-finite entries of diag, coeff, b (NonFiniteInput);
// counters[1]: diagonals the Jacobi preconditioner rejects (not finite, 0 or
// |d| < 1e-12, src/algebra/Preconditioner.cpp:11-45: InvalidSystem).
__global__ void checkKernel(int64_t n, int64_t nf, const double* __restrict__ diag,
                            const double* __restrict__ coeff, const double* __restrict__ b,
                            unsigned* __restrict__ counters) {
  unsigned bad = 0, rejected = 0;
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double d = diag[k];
    bad += !isfinite(d) + !isfinite(b[k]);
    rejected += !isfinite(d) || d == 0.0 || fabs(d) < 1e-12;
  }
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < nf;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    bad += !isfinite(coeff[k]);
  if (bad) atomicAdd(&counters[0], bad);
  if (rejected) atomicAdd(&counters[1], rejected);
}

// r = b - A x (FP64, the 2D p' operator, ascending columns), |r|^2 partials.
__global__ __launch_bounds__(256) void residualNormKernel(int64_t nx, int64_t ny, int64_t pin,
                                                          int64_t offY,
                                                          const double* __restrict__ diag,
                                                          const double* __restrict__ coeff,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ x,
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
    acc[0] += rc * rc;
  }
  blockPartials<1>(acc, partial);
}

inline dim3 strideGrid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 2048) g = 2048;
  return dim3(static_cast<unsigned>(g < 1 ? 1 : g));
}

cfdhip_solve_result solveNoAlloc(Context& ctx, const cfdhip_solver_settings& set,
                                 const double* b, double* x) {
  ctx.requireMesh();
  require(ctx.pressureAssembled, "no p' system assembled");
  const View& v = ctx.view;
  require(v.dim == 2, "the no-allocation p' solve runs on 2D meshes", CFDHIP_ERR_UNSUPPORTED);
  require(set.type == CFDHIP_SOLVER_CG && set.preconditioner == CFDHIP_PREC_JACOBI &&
              set.precision == CFDHIP_FP64,
          "the no-allocation p' solve is FP64 CG with the Jacobi preconditioner",
          CFDHIP_ERR_UNSUPPORTED);
  require(set.absolute_tolerance >= 0.0 && set.relative_tolerance >= 0.0 &&
              (set.absolute_tolerance > 0.0 || set.relative_tolerance > 0.0),
          "LinearSolverSettings: tolerances must be >= 0 and one of them > 0");
  require(set.max_iterations >= 1 && set.max_iterations <= kNaMaxIter,
          "the no-allocation p' solve takes 1 .. 65536 iterations");
  require(v.n <= kNaMaxCells,
          "the no-allocation p' solve takes at most 1048576 cells (its scratch is fixed)");
  require(b != nullptr && x != nullptr, "device right-hand side and solution are required");
  double* base = scratchOf(ctx.device);
  double* partial = base + kNaOffPartial;
  double* hist = base + kNaOffHist;
  KState* state = reinterpret_cast<KState*>(base + kNaOffState);
  unsigned* counters = reinterpret_cast<unsigned*>(base + kNaOffMisc);
  double* norm = base + kNaOffMisc + 4;
  const int64_t n = v.n, nf = v.nf;
  const double* diag = ctx.f(CFDHIP_FIELD_PC_DIAG);
  const double* coeff = ctx.f(CFDHIP_FIELD_PC_FACE_COEFF);
  // clear the counters and the solution (x0 = 0), then the input checks of
  // the resident solve: non-finite inputs, then the Jacobi rejection
  hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, ctx.stream, int64_t(2), base + kNaOffMisc,
                     0.0);
  hipLaunchKernelGGL(fillKernel, strideGrid(n), dim3(256), 0, ctx.stream, n, x, 0.0);
  hipLaunchKernelGGL(checkKernel, strideGrid(n), dim3(256), 0, ctx.stream, n, nf, diag, coeff, b,
                     counters);
  CFDHIP_CHECK(hipGetLastError());
  unsigned hc[2] = {0, 0};
  CFDHIP_CHECK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  cfdhip_solve_result res{};
  if (hc[0] != 0) {
    res.status = CFDHIP_NONFINITE_INPUT;
    return res;
  }
  if (hc[1] != 0) {
    res.status = CFDHIP_INVALID_SYSTEM;
    return res;
  }
  double* work[8];
  for (int q = 0; q < 8; ++q) work[q] = base + q * kNaMaxCells;
  KState host{};
  cfdhip_solver_settings run = set;
  run.graph = 0;
  const PcgBuffers buf{partial, kNaPartials, state, &host, hist, kNaHist, false, state + 1};
  res = pcg2dOnBuffers(ctx, run, diag, coeff, b, x, work, buf);
  // the FP64 true residual |b - A x|
  const dim3 g = strideGrid(n);
  hipLaunchKernelGGL(residualNormKernel, g, dim3(256), 0, ctx.stream, v.nx, v.ny, ctx.pin, v.offY,
                     diag, coeff, b, x, partial);
  CFDHIP_CHECK(hipGetLastError());
  launchFinalize(ctx, FinalArgs{partial, static_cast<int64_t>(g.x), state, nullptr, norm,
                                kStepNorm});
  double tr = 0.0;
  CFDHIP_CHECK(hipMemcpyAsync(&tr, norm, sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  res.true_residual = tr;
  return res;
}

}  // namespace

}  // namespace cfdhip

extern "C" int cfdhip_pressure_solve_device_noalloc(cfdhip_ctx* ctx,
                                                    const cfdhip_solver_settings* s,
                                                    const double* b, double* x,
                                                    cfdhip_solve_result* r) {
  if (ctx == nullptr) return CFDHIP_ERR_INVALID;
  cfdhip::Context& c = ctx->c;
  try {
    CFDHIP_CHECK(hipSetDevice(c.device));
    cfdhip::require(s != nullptr, "solver settings are required");
    const cfdhip_solve_result res = cfdhip::solveNoAlloc(c, *s, b, x);
    if (r) *r = res;
    c.lastError.clear();
    return CFDHIP_OK;
  } catch (const cfdhip::Error& e) {
    c.lastError = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c.lastError = e.what();
    return CFDHIP_ERR_RUNTIME;
  }
}
