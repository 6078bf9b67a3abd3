// The p' system on a general face-list mesh: non-orthogonal and multiblock
// meshes that the structured path does not take (the reference's CSR route,
// cuda/kernels/DevicePressureCorrectionKernel.cu with the explicit-flux terms,
// :176-216).
//
// The caller hands over the mesh as the reference's Mesh holds it (face
// owner / neighbour, face centroid and area vector, cell centroid, each
// cell's face list, which boundary faces carry a FixedValue pressure), as
// DeviceMesh uploads it. Per assembly:
//   pressureCorrectionFaceCoupling (PressureCorrectionEquation.cpp:142-179)
//   with the over-relaxed decomposition of the response vector
//   (MeshGeometry::decomposeAreaVector, MeshGeometry.cpp:144-174) and the
//   explicit non-orthogonal flux -T . grad(p'_prev) of a caller-supplied
//   gradient (:235-290), one thread per face;
//   the CSR values in SparseMatrixBuilder's order: a host plan lists, for each
//   (row, column) slot, the contributions in insertion order (face loop, then
//   the additional diagonal, then the pin), so one thread per slot sums them
//   as the builder does (first value, then +=);
//   the right-hand side -imbalance (cell face-list order) with the explicit
//   terms added in face order (:291-303);
// then Jacobi CG / BiCGSTAB on the CSR operator (krylov.hip) and
// correctFaceMassFlux with the explicit term (:369-402).
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.hpp"

namespace cfdhip {

namespace {

constexpr int kFmBlock = 256;

dim3 fmBlocks(int64_t n) { return dim3(static_cast<unsigned>((n + kFmBlock - 1) / kFmBlock)); }

struct V3 {
  double x, y, z;
};

__device__ inline V3 ld3(const double* a, int64_t i) { return V3{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
__device__ inline V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 scale3(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
// Vector3 dot: (a.x b.x + a.y b.y) + a.z b.z
__device__ inline double dot3(V3 a, V3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__device__ inline double mag3(V3 a) { return sqrt(dot3(a, a)); }
__device__ inline bool crossZero(V3 a, V3 b) {
  return (a.y * b.z) - (a.z * b.y) == 0.0 && (a.z * b.x) - (a.x * b.z) == 0.0 &&
         (a.x * b.y) - (a.y * b.x) == 0.0;
}
__device__ inline bool axisAligned(V3 s) {
  return (s.x != 0.0 ? 1 : 0) + (s.y != 0.0 ? 1 : 0) + (s.z != 0.0 ? 1 : 0) == 1;
}

// coupling (PressureCorrectionEquation.cpp:73-98): the coefficient and the
// non-orthogonal vector of one face.
__device__ inline void coupling(V3 sf, V3 d, double rho, double du, double dv, double dw,
                                bool nonOrth, double& coef, V3& nvec) {
  const double distance = mag3(d);
  nvec = V3{0.0, 0.0, 0.0};
  if (axisAligned(sf) && crossZero(d, sf)) {
    const double dc = sf.x != 0.0 ? du : (sf.y != 0.0 ? dv : dw);
    coef = rho * mag3(sf) * dc / distance;
    return;
  }
  const V3 rv = sf.z == 0.0 ? V3{du * sf.x, dv * sf.y, 0.0} : V3{du * sf.x, dv * sf.y, dw * sf.z};
  if (nonOrth) {
    // decomposeAreaVector: isFaceGeometryWellPosed, the exact-parallel
    // short cut, then the over-relaxed split
    const double dDotS = dot3(d, rv);
    const double dm = mag3(d), sm = mag3(rv);
    if (dm > 0.0 && sm > 0.0 && dDotS > (1e-6 * dm * sm)) {
      V3 orth = rv, non{0.0, 0.0, 0.0};
      if (!crossZero(d, rv)) {
        const double sDotS = dot3(rv, rv);
        orth = scale3(d, sDotS / dDotS);
        non = sub3(rv, orth);
      }
      coef = rho * mag3(orth) / distance;
      nvec = scale3(non, rho);
      return;
    }
  }
  coef = rho * mag3(rv) / distance;
}

struct FaceArgs {
  int64_t nf;
  const int64_t* owner;
  const int64_t* neigh;
  const double* fc;
  const double* sf;
  const double* cc;
  const int* fixedValue;
  const double* rhoF;  // nullptr: rho
  double rho;
  const double* du;
  const double* dv;
  const double* dw;  // nullptr in 2D
  int nonOrth;
  const double* grad;  // 3 per cell, nullptr: no explicit term
  double* coeff;
  double* expl;
  unsigned* nonFinite;
};

__global__ void faceCouplingKernel(FaceArgs a) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= a.nf) return;
  const int64_t o = a.owner[f], nb = a.neigh[f];
  if (nb < 0 && !a.fixedValue[f]) {
    a.coeff[f] = 0.0;
    a.expl[f] = 0.0;
    return;
  }
  const V3 sf = ld3(a.sf, f), fc = ld3(a.fc, f), co = ld3(a.cc, o);
  const double rho = a.rhoF != nullptr ? a.rhoF[f] : a.rho;
  double coef;
  V3 nvec;
  double dPf = 0.0, dNf = 0.0;
  if (nb < 0) {
    coupling(sf, sub3(fc, co), rho, a.du[o], a.dv[o], a.dw ? a.dw[o] : 0.0, a.nonOrth, coef, nvec);
  } else {
    const V3 cn = ld3(a.cc, nb);
    const V3 d = sub3(cn, co);
    dPf = mag3(sub3(fc, co));  // MeshGeometry::distance(owner, face)
    dNf = mag3(sub3(fc, cn));
    auto interp = [&](const double* r) { return ((dNf * r[o]) + (dPf * r[nb])) / (dPf + dNf); };
    if (axisAligned(sf) && crossZero(d, sf)) {
      const double* r = sf.x != 0.0 ? a.du : (sf.y != 0.0 ? a.dv : a.dw);
      const double df = interp(r);
      coupling(sf, d, rho, df, df, df, a.nonOrth, coef, nvec);
    } else {
      coupling(sf, d, rho, interp(a.du), interp(a.dv), a.dw ? interp(a.dw) : 0.0, a.nonOrth, coef,
               nvec);
    }
  }
  if (!isfinite(coef) || !isfinite(nvec.x) || !isfinite(nvec.y) || !isfinite(nvec.z))
    atomicAdd(a.nonFinite, 1u);
  a.coeff[f] = coef;
  double e = 0.0;
  if (a.grad != nullptr) {
    V3 g;
    if (nb < 0) {
      g = ld3(a.grad, o);
    } else {  // interpolateInternalFace (vector): times the reciprocal
      const V3 gp = ld3(a.grad, o), gn = ld3(a.grad, nb);
      const double inv = 1.0 / (dPf + dNf);
      g = V3{((gp.x * dNf) + (gn.x * dPf)) * inv, ((gp.y * dNf) + (gn.y * dPf)) * inv,
             ((gp.z * dNf) + (gn.z * dPf)) * inv};
    }
    e = -dot3(nvec, g);
  }
  a.expl[f] = e;
}

// One CSR slot: its contributions in insertion order. Code >= 0: face
// code >> 1 with sign (code & 1 ? -1 : +1); -1: the pin's 1.0; <= -2: the
// additional diagonal of row -code - 2.
__global__ void slotKernel(int64_t nslots, const int64_t* __restrict__ off,
                           const int64_t* __restrict__ code, const double* __restrict__ coeff,
                           const double* __restrict__ extra, double* val) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nslots) return;
  double v = 0.0;
  for (int64_t q = off[k]; q < off[k + 1]; ++q) {
    const int64_t c = code[q];
    double t;
    if (c >= 0) {
      const double cf = coeff[c >> 1];
      t = (c & 1) ? -cf : cf;
    } else if (c == -1) {
      t = 1.0;
    } else {
      t = extra[-c - 2];
    }
    v = q == off[k] ? t : v + t;
  }
  val[k] = v;
}

__global__ void fmRhsKernel(int64_t n, const int64_t* __restrict__ cfOff,
                            const int64_t* __restrict__ cf, const int64_t* __restrict__ cfSorted,
                            const int64_t* __restrict__ owner, const double* __restrict__ flux,
                            const double* __restrict__ expl, int hasExplicit, int64_t pin,
                            double* rhs) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double s = 0.0;  // evaluateContinuity: the cell's face list order
  for (int64_t k = cfOff[c]; k < cfOff[c + 1]; ++k) {
    const int64_t f = cf[k];
    s += owner[f] == c ? flux[f] : -flux[f];
  }
  double b = -s;
  if (hasExplicit)  // the explicit terms, faces in id order
    for (int64_t k = cfOff[c]; k < cfOff[c + 1]; ++k) {
      const int64_t f = cfSorted[k];
      if (owner[f] == c)
        b -= expl[f];
      else
        b += expl[f];
    }
  rhs[c] = c == pin ? 0.0 : b;
}

__global__ void fmCorrectFluxKernel(int64_t nf, const int64_t* __restrict__ owner,
                                    const int64_t* __restrict__ neigh,
                                    const double* __restrict__ fstar,
                                    const double* __restrict__ coeff,
                                    const double* __restrict__ pp, const double* __restrict__ expl,
                                    double* out) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double po = pp[owner[f]];
  const double pn = neigh[f] < 0 ? 0.0 : pp[neigh[f]];
  double F = fstar[f] + (coeff[f] * (po - pn));
  if (expl != nullptr) F += expl[f];
  out[f] = F;
}

template <class T>
void upload(Context& ctx, DeviceArray<T>& d, const T* h, int64_t n) {
  d.ensure(n > 0 ? n : 1);
  if (n > 0)
    CFDHIP_CHECK(hipMemcpyAsync(d.get(), h, n * sizeof(T), hipMemcpyHostToDevice, ctx.stream));
}

// The CSR pattern and slot contributions for the pin and the presence of an
// additional diagonal (PressureCorrectionEquation.cpp:255-318).
void buildPlan(Context& ctx, FaceMesh& m, int64_t pin, bool extra) {
  if (m.planPin == pin && m.planExtra == extra) return;
  std::vector<std::vector<std::pair<int64_t, int64_t>>> rows(m.n);  // (col, code) in order
  for (int64_t f = 0; f < m.nf; ++f) {
    const int64_t o = m.hOwner[f], nb = m.hNeigh[f];
    if (nb < 0) {
      if (m.hFixed[f] && (pin < 0 || o != pin)) rows[o].push_back({o, 2 * f});
      continue;
    }
    if (pin < 0 || o != pin) {
      rows[o].push_back({o, 2 * f});
      rows[o].push_back({nb, 2 * f + 1});
    }
    if (pin < 0 || nb != pin) {
      rows[nb].push_back({nb, 2 * f});
      rows[nb].push_back({o, 2 * f + 1});
    }
  }
  if (extra)
    for (int64_t c = 0; c < m.n; ++c)
      if (pin < 0 || c != pin) rows[c].push_back({c, -c - 2});
  if (pin >= 0) rows[pin].push_back({pin, -1});
  std::vector<int64_t> rp{0}, col, off{0}, code;
  for (int64_t r = 0; r < m.n; ++r) {
    auto& e = rows[r];
    std::stable_sort(e.begin(), e.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t i = 0; i < e.size(); ++i) {
      if (i == 0 || e[i].first != e[i - 1].first) {
        if (i > 0) off.push_back(static_cast<int64_t>(code.size()));
        col.push_back(e[i].first);
      }
      code.push_back(e[i].second);
    }
    if (!e.empty()) off.push_back(static_cast<int64_t>(code.size()));
    rp.push_back(static_cast<int64_t>(col.size()));
  }
  m.hRp = rp;
  m.hCol = col;
  m.A.n = m.n;
  upload(ctx, m.A.rp, rp.data(), static_cast<int64_t>(rp.size()));
  upload(ctx, m.A.col, col.data(), static_cast<int64_t>(col.size()));
  m.A.val.ensure(col.empty() ? 1 : static_cast<int64_t>(col.size()));
  m.A.val.count = static_cast<int64_t>(col.size());
  upload(ctx, m.slotOff, off.data(), static_cast<int64_t>(off.size()));
  upload(ctx, m.slotCode, code.data(), static_cast<int64_t>(code.size()));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  m.planPin = pin;
  m.planExtra = extra;
}

// the non-finite counter (flags[0]) read and cleared
unsigned takeNonFinite(Context& ctx) {
  unsigned h = 0;
  CFDHIP_CHECK(hipMemcpyAsync(&h, ctx.flags.get(), sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  CFDHIP_CHECK(hipMemsetAsync(ctx.flags.get(), 0, sizeof(unsigned), ctx.stream));
  return h;
}

FaceMesh& need(Context& ctx) {
  require(ctx.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
  return *ctx.fmesh;
}

}  // namespace

void fmeshSet(Context& ctx, int dim, int64_t n, int64_t nf, const int64_t* owner,
              const int64_t* neighbor, const double* faceCentroid, const double* areaVector,
              const double* cellCentroid, const int64_t* cellFaceOffsets,
              const int64_t* cellFaces, const int32_t* fixedValue) {
  require(dim == 2 || dim == 3, "face-list mesh: dim must be 2 or 3");
  require(n > 0 && nf > 0, "face-list mesh: empty mesh");
  require(owner && neighbor && faceCentroid && areaVector && cellCentroid && cellFaceOffsets &&
              cellFaces && fixedValue,
          "face-list mesh: every array is required");
  require(cellFaceOffsets[0] == 0, "face-list mesh: cell face offsets must start at 0");
  for (int64_t c = 0; c < n; ++c)
    require(cellFaceOffsets[c + 1] >= cellFaceOffsets[c],
            "face-list mesh: cell face offsets must be non-decreasing");
  const int64_t ncf = cellFaceOffsets[n];
  auto m = std::make_unique<FaceMesh>();
  m->dim = dim;
  m->n = n;
  m->nf = nf;
  m->hOwner.assign(owner, owner + nf);
  m->hNeigh.assign(neighbor, neighbor + nf);
  m->hFixed.assign(fixedValue, fixedValue + nf);
  m->hFc.assign(faceCentroid, faceCentroid + 3 * nf);
  m->hSf.assign(areaVector, areaVector + 3 * nf);
  m->hCc.assign(cellCentroid, cellCentroid + 3 * n);
  m->hCfOff.assign(cellFaceOffsets, cellFaceOffsets + n + 1);
  m->open = false;
  for (int64_t f = 0; f < nf; ++f) {
    require(owner[f] >= 0 && owner[f] < n && neighbor[f] >= -1 && neighbor[f] < n &&
                neighbor[f] != owner[f],
            "face-list mesh: owner/neighbour out of range");
    require(!fixedValue[f] || neighbor[f] < 0,
            "face-list mesh: only boundary faces carry a FixedValue pressure");
    m->open = m->open || fixedValue[f] != 0;
  }
  m->hCf.assign(cellFaces, cellFaces + ncf);
  std::vector<int64_t> sorted(cellFaces, cellFaces + ncf);
  for (int64_t c = 0; c < n; ++c) {
    for (int64_t k = cellFaceOffsets[c]; k < cellFaceOffsets[c + 1]; ++k) {
      const int64_t f = cellFaces[k];
      require(f >= 0 && f < nf && (owner[f] == c || neighbor[f] == c),
              "face-list mesh: a cell's face list names a face not on that cell");
    }
    std::sort(sorted.begin() + cellFaceOffsets[c], sorted.begin() + cellFaceOffsets[c + 1]);
  }
  upload(ctx, m->owner, owner, nf);
  upload(ctx, m->neigh, neighbor, nf);
  upload(ctx, m->fc, faceCentroid, 3 * nf);
  upload(ctx, m->sf, areaVector, 3 * nf);
  upload(ctx, m->cc, cellCentroid, 3 * n);
  upload(ctx, m->cfOff, cellFaceOffsets, n + 1);
  upload(ctx, m->cf, cellFaces, ncf);
  upload(ctx, m->cfSorted, sorted.data(), ncf);
  upload(ctx, m->fixed, fixedValue, nf);
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  ctx.fmesh = std::move(m);
}

namespace {

// The assembly on the inputs held in m; gradDev: the gradient of the previous
// p' (3 per cell, device) for the explicit term, or null.
void assembleOnDevice(Context& ctx, FaceMesh& m, const double* gradDev) {
  const int64_t pin = m.open ? -1 : m.ref;
  m.assembled = false;
  m.coeff.ensure(m.nf);
  m.expl.ensure(m.nf);
  m.rhs.ensure(m.n);
  buildPlan(ctx, m, pin, m.hasExtra);
  FaceArgs a{m.nf, m.owner.get(), m.neigh.get(), m.fc.get(), m.sf.get(), m.cc.get(), m.fixed.get(),
             m.hasRhoF ? m.rhoF.get() : nullptr, m.rho, m.du.get(), m.dv.get(),
             m.hasDw ? m.dw.get() : nullptr, m.nonOrth, gradDev, m.coeff.get(), m.expl.get(),
             ctx.flags.get()};
  hipLaunchKernelGGL(faceCouplingKernel, fmBlocks(m.nf), dim3(kFmBlock), 0, ctx.stream, a);
  CFDHIP_CHECK(hipGetLastError());
  require(takeNonFinite(ctx) == 0,
          "pressureCorrectionFaceCoupling: non-finite face coupling (degenerate face geometry)",
          CFDHIP_ERR_NONFINITE);
  const int64_t nslots = static_cast<int64_t>(m.hCol.size());
  hipLaunchKernelGGL(slotKernel, fmBlocks(nslots), dim3(kFmBlock), 0, ctx.stream, nslots,
                     m.slotOff.get(), m.slotCode.get(), m.coeff.get(),
                     m.hasExtra ? m.extra.get() : nullptr, m.A.val.get());
  hipLaunchKernelGGL(fmRhsKernel, fmBlocks(m.n), dim3(kFmBlock), 0, ctx.stream, m.n, m.cfOff.get(),
                     m.cf.get(), m.cfSorted.get(), m.owner.get(), m.flux.get(), m.expl.get(),
                     gradDev ? 1 : 0, pin, m.rhs.get());
  CFDHIP_CHECK(hipGetLastError());
  const double* arrays[2] = {m.A.val.get(), m.rhs.get()};
  const int64_t counts[2] = {nslots, m.n};
  require(countNonFinite(ctx, arrays, counts, 2) == 0,
          "assembleGeometricPressureCorrection: assembled system contains a non-finite value",
          CFDHIP_ERR_NONFINITE);
  m.hasExplicit = gradDev != nullptr;
  m.assembled = true;
}

}  // namespace

void fmeshAssemblePressure(Context& ctx, const double* fluxStar, const double* faceDensity,
                           double rho, const double* du, const double* dv, const double* dw,
                           int nonOrthogonal, const double* gradPrev, const double* extraDiag,
                           int64_t ref) {
  FaceMesh& m = need(ctx);
  require(fluxStar && du && dv, "face-list p': predictor flux and response coefficients are "
                                "required");
  require(m.dim == 2 || dw != nullptr, "face-list p': a 3D mesh needs the w response coefficient",
          CFDHIP_ERR_INVALID);
  require(ref >= 0 && ref < m.n, "assembleGeometricPressureCorrection: referenceCell out of range");
  m.assembled = false;
  upload(ctx, m.flux, fluxStar, m.nf);
  upload(ctx, m.du, du, m.n);
  upload(ctx, m.dv, dv, m.n);
  if (dw) upload(ctx, m.dw, dw, m.n);
  if (faceDensity) upload(ctx, m.rhoF, faceDensity, m.nf);
  if (gradPrev) upload(ctx, m.grad, gradPrev, 3 * m.n);
  if (extraDiag) upload(ctx, m.extra, extraDiag, m.n);
  m.rho = rho;
  m.nonOrth = nonOrthogonal;
  m.hasRhoF = faceDensity != nullptr;
  m.hasDw = dw != nullptr;
  m.hasExtra = extraDiag != nullptr;
  m.ref = ref;
  assembleOnDevice(ctx, m, gradPrev ? m.grad.get() : nullptr);
}

void fmeshReassemble(Context& ctx, const double* gradDev) {
  FaceMesh& m = need(ctx);
  require(m.flux.get() != nullptr && m.du.get() != nullptr,
          "face-list p': no assembly to repeat (call cfdhip_fmesh_assemble_pressure first)");
  assembleOnDevice(ctx, m, gradDev);
}

int64_t fmeshSystem(Context& ctx, int64_t* rowptr, int64_t* col, double* val, int64_t cap,
                    double* rhs, double* faceCoeff, double* explicitFlux) {
  FaceMesh& m = need(ctx);
  require(m.assembled, "face-list p': no assembled system");
  const int64_t slots = static_cast<int64_t>(m.hCol.size());
  std::vector<double> v(slots > 0 ? slots : 1);
  CFDHIP_CHECK(hipMemcpy(v.data(), m.A.val.get(), slots * sizeof(double), hipMemcpyDeviceToHost));
  // SparseMatrixBuilder drops entries that sum to an exact zero
  int64_t nnz = 0;
  for (int64_t k = 0; k < slots; ++k) nnz += v[k] != 0.0 ? 1 : 0;
  if (rowptr != nullptr) {
    require(col && val && cap >= nnz, "face-list p': CSR buffers too small");
    int64_t w = 0;
    rowptr[0] = 0;
    for (int64_t r = 0; r < m.n; ++r) {
      for (int64_t k = m.hRp[r]; k < m.hRp[r + 1]; ++k)
        if (v[k] != 0.0) {
          col[w] = m.hCol[k];
          val[w] = v[k];
          ++w;
        }
      rowptr[r + 1] = w;
    }
  }
  if (rhs) CFDHIP_CHECK(hipMemcpy(rhs, m.rhs.get(), m.n * sizeof(double), hipMemcpyDeviceToHost));
  if (faceCoeff)
    CFDHIP_CHECK(
        hipMemcpy(faceCoeff, m.coeff.get(), m.nf * sizeof(double), hipMemcpyDeviceToHost));
  if (explicitFlux)
    CFDHIP_CHECK(
        hipMemcpy(explicitFlux, m.expl.get(), m.nf * sizeof(double), hipMemcpyDeviceToHost));
  return nnz;
}

cfdhip_solve_result fmeshSolve(Context& ctx, const cfdhip_solver_settings& s, double* x) {
  FaceMesh& m = need(ctx);
  require(m.assembled, "face-list p': no assembled system");
  require(s.preconditioner != CFDHIP_PREC_LEARNED && s.type != CFDHIP_SOLVER_FCG,
          "face-list p': the learned preconditioner needs the structured path",
          CFDHIP_ERR_UNSUPPORTED);
  m.x.ensure(m.n);
  CFDHIP_CHECK(hipMemsetAsync(m.x.get(), 0, m.n * sizeof(double), ctx.stream));
  const cfdhip_solve_result r = solveCsr(ctx, s, m.A, m.rhs.get(), m.x.get());
  if (x) {
    CFDHIP_CHECK(hipMemcpyAsync(x, m.x.get(), m.n * sizeof(double), hipMemcpyDeviceToHost,
                                ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  }
  m.solved = true;
  return r;
}

void fmeshCorrectFlux(Context& ctx, const double* pprime, double* flux) {
  FaceMesh& m = need(ctx);
  require(m.assembled, "face-list p': no assembled system");
  require(flux != nullptr, "face-list p': output flux buffer is required");
  const double* pp = m.x.get();
  DeviceArray<double> tmp;
  if (pprime != nullptr) {
    upload(ctx, tmp, pprime, m.n);
    pp = tmp.get();
  } else {
    require(m.solved, "face-list p': no p' given and none solved");
  }
  DeviceArray<double> out;
  out.ensure(m.nf);
  hipLaunchKernelGGL(fmCorrectFluxKernel, fmBlocks(m.nf), dim3(kFmBlock), 0, ctx.stream, m.nf,
                     m.owner.get(), m.neigh.get(), m.flux.get(), m.coeff.get(), pp,
                     m.hasExplicit ? m.expl.get() : nullptr, out.get());
  CFDHIP_CHECK(hipGetLastError());
  CFDHIP_CHECK(
      hipMemcpyAsync(flux, out.get(), m.nf * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
}

}  // namespace cfdhip
