// SIMPLE discretization kernels on the structured Cartesian plan.
//
// Each kernel is a gather over one cell (or one face) that evaluates the
// reference's expressions in the reference's operand order, with
// -ffp-contract=off, so results are bitwise equal to the CPU path:
//   gradient          src/discretization/Gradient.cpp:104-162, 511-596
//   velocity gradient src/discretization/VectorGradient.cpp:19-80
//   mass flux         src/physics/MassFlux.cpp:15-36
//   momentum          src/pressure_velocity/RelaxedMomentum.cpp:30-104,
//                     src/physics/MomentumEquation.cpp:187-430,
//                     src/discretization/NonOrthogonalDiffusion.cpp:51-113,
//                     src/pressure_velocity/UnderRelaxation.cpp:13-36
//   response          src/pressure_velocity/PressureCorrectionEquation.cpp:103-114
//   Rhie-Chow         src/pressure_velocity/RhieChow.cpp:17-65
//   p' assembly       src/pressure_velocity/PressureCorrectionEquation.cpp:146-339
//   corrections       PressureCorrectionEquation.cpp:369-438, SIMPLE.cpp:939-942
//   continuity        src/physics/ContinuityEquation.cpp:14-55
//
// A matrix row's entries arrive in the order SparseMatrixBuilder would sum
// them: the face loop visits a cell's faces in creation order (W, E, S, N,
// B, T), diffusion for all faces before convection, relaxation last; an
// entry's sum starts from its first triplet.
#include <hipcub/hipcub.hpp>

#include "context.hpp"

namespace cfdhip {

namespace {

struct Slot {
  bool boundary;
  int64_t face;
  int64_t cell;  // neighbour cell across the face (-1 on a boundary)
  int64_t fi;    // face index along the axis
};

__device__ inline Slot slotOf(const View& v, int64_t c, const int64_t idx[3], int a, int side) {
  Slot s;
  int64_t fidx[3] = {idx[0], idx[1], idx[2]};
  s.fi = idx[a] + side;
  fidx[a] = s.fi;
  s.face = faceId(v, a, fidx);
  s.boundary = side ? (idx[a] == v.ncell[a] - 1) : (idx[a] == 0);
  s.cell = s.boundary ? -1 : (side ? c + v.stride[a] : c - v.stride[a]);
  return s;
}

// interpolateInternalFace, scalar (Interpolation.cpp:21-29): division by dPf+dNf.
__device__ inline double interpScalar(const View& v, int a, int64_t fi, double phiO, double phiN) {
  return ((tabv(v, a, kDNf, fi) * phiO) + (tabv(v, a, kDPf, fi) * phiN)) / tabv(v, a, kDSum, fi);
}
// interpolateInternalFace, vector (Interpolation.cpp:31-41): times 1/(dPf+dNf).
__device__ inline double interpVector(const View& v, int a, int64_t fi, double phiO, double phiN) {
  return ((phiO * tabv(v, a, kDNf, fi)) + (phiN * tabv(v, a, kDPf, fi))) *
         (1.0 / tabv(v, a, kDSum, fi));
}

__device__ inline bool prescribes(int vkind) { return vkind != CFDHIP_OUTLET; }

// Boundary velocity component (Wall / MovingWall / Inlet / Outlet::boundaryValue).
__device__ inline double boundaryVelocity(const PatchGeom& P, int k, double ownerValue) {
  if (P.vkind == CFDHIP_OUTLET) return ownerValue;
  if (P.vkind == CFDHIP_WALL) return 0.0;
  return P.vval[k];
}

// Scalar boundary value: bcMode 0 = the pressure conditions, 1 = the p'
// conditions of makeGradientBoundaries (PressureCorrectionEquation.cpp:43-58).
__device__ inline double scalarBoundary(const PatchGeom& P, int bcMode, double ownerValue) {
  const bool fixedValue = P.pkind == CFDHIP_FIXED_VALUE;
  const double value = bcMode == 0 ? P.pval : 0.0;
  if (fixedValue) return value;
  return ownerValue + (value * P.bdist);  // FixedGradient::boundaryValue
}

// Green-Gauss gradient of a scalar at cell c with the boundary-consistent
// opposite-face correction and, when the plan needs it, the four transfer
// sweeps (each sweep only reads this cell's previous gradient).
__device__ void scalarGradient(const View& v, const double* __restrict__ phi, int bcMode,
                               int64_t c, double g[3]) {
  int64_t idx[3];
  cellIndex(v, c, idx);
  const int D = v.dim;
  double fv[6];
  int64_t nb[6];
  bool bnd[6];
  const double pc = phi[c];
  for (int a = 0; a < D; ++a)
    for (int side = 0; side < 2; ++side) {
      const int s = 2 * a + side;
      const Slot sl = slotOf(v, c, idx, a, side);
      bnd[s] = sl.boundary;
      nb[s] = sl.cell;
      if (sl.boundary) {
        fv[s] = scalarBoundary(v.patch[s], bcMode, pc);
      } else {
        const double po = side ? pc : phi[sl.cell];
        const double pn = side ? phi[sl.cell] : pc;
        fv[s] = interpScalar(v, a, sl.fi, po, pn);
      }
    }
  double og[3] = {0.0, 0.0, 0.0};
  const int sweeps = v.transfer ? 5 : 1;
  for (int sw = 0; sw < sweeps; ++sw) {
    double val[6];
    for (int s = 0; s < 2 * D; ++s) val[s] = fv[s];
    for (int s = 0; s < 2 * D; ++s) {
      if (!bnd[s]) continue;
      const PatchGeom& P = v.patch[s];
      const int opp = s ^ 1;
      if (!P.consistent || bnd[opp]) continue;
      const double phiB = fv[s] + dot3(og, P.offset);
      const double phiF = phi[nb[opp]];
      const double back = P.back, L = P.farDist;
      const double d2 =
          2.0 * (((phiB - pc) / back) + ((phiF - pc) / L)) / (back + L);
      const double corr = -0.5 * P.w * (1.0 - P.w) * L * L * d2;
      val[opp] = fv[opp] + corr;
    }
    const double inv = 1.0 / v.vol;
    for (int a = 0; a < 3; ++a) {
      if (a >= D) {
        g[a] = 0.0;
        continue;
      }
      const double sLow = -v.S[a][a];  // cell-outward area component of the low face
      const double sHigh = v.S[a][a];
      const double sum = (0.0 + (sLow * val[2 * a])) + (sHigh * val[2 * a + 1]);
      g[a] = sum * inv;
    }
    og[0] = g[0];
    og[1] = g[1];
    og[2] = g[2];
  }
}

__global__ void gradientKernel(View v, const double* __restrict__ phi, int bcMode, double* gx,
                               double* gy, double* gz) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  double g[3];
  scalarGradient(v, phi, bcMode, c, g);
  gx[c] = g[0];
  gy[c] = g[1];
  if (gz) gz[c] = g[2];
}

// Plain Green-Gauss gradient of velocity component k at cell c
// (greenGaussVelocityGradient: interpolated interior values, boundary values
// from the velocity condition, no boundary-consistent correction).
__device__ void velocityGradient(const View& v, const double* const* u, int k, int64_t c,
                                 double g[3]) {
  int64_t idx[3];
  cellIndex(v, c, idx);
  const double inv = 1.0 / v.vol;
  const double* uk = u[k];
  for (int a = 0; a < 3; ++a) {
    if (a >= v.dim) {
      g[a] = 0.0;
      continue;
    }
    double val[2];
    for (int side = 0; side < 2; ++side) {
      const Slot sl = slotOf(v, c, idx, a, side);
      if (sl.boundary) {
        val[side] = boundaryVelocity(v.patch[2 * a + side], k, uk ? uk[c] : 0.0);
      } else {
        const double uo = uk ? (side ? uk[c] : uk[sl.cell]) : 0.0;
        const double un = uk ? (side ? uk[sl.cell] : uk[c]) : 0.0;
        val[side] = interpVector(v, a, sl.fi, uo, un);
      }
    }
    const double sum = (0.0 + ((-v.S[a][a]) * val[0])) + (v.S[a][a] * val[1]);
    g[a] = sum * inv;
  }
}

struct Acc {
  double s;
  bool has;
  __device__ void add(double x) {
    if (has) {
      s = s + x;
    } else {
      s = x;
      has = true;
    }
  }
};

// CSR column slot of the neighbour across (axis a, side): low neighbours in
// descending axis order, then the diagonal, then high neighbours.
__device__ __host__ inline int colSlot(int D, int a, int side) {
  return side ? D + 1 + a : D - 1 - a;
}

struct MomentumArgs {
  const double* u[3];
  const double* gp[3];  // pressure gradient
  const double* F;
  const double* mu;
  double* coef;  // (2D+1) * n
  double* rhs[3];
  double* diag;
  unsigned* nonFinite;
  double alpha;
  // transient momentum (dt > 0, alpha 1): the implicit-Euler time term of the
  // previous velocity uOld with density rho
  double rho, dt;
  const double* uOld[3];
  const double* uRelax[3];  // the under-relaxation's phi_old (steady: the resident velocity)
};

__global__ void momentumKernel(View v, MomentumArgs m) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  const int D = v.dim;
  int64_t idx[3];
  cellIndex(v, c, idx);
  Acc acc[7];
  for (int q = 0; q < 7; ++q) acc[q] = Acc{0.0, false};
  double rhs[3] = {0.0, 0.0, 0.0};
  const int P = D;  // diagonal slot

  // Diffusion (MomentumEquation.cpp:187-262, effective-viscosity overload).
  for (int a = 0; a < D; ++a)
    for (int side = 0; side < 2; ++side) {
      const Slot sl = slotOf(v, c, idx, a, side);
      if (!sl.boundary) {
        const double muO = side ? m.mu[c] : m.mu[sl.cell];
        const double muN = side ? m.mu[sl.cell] : m.mu[c];
        const double muf = interpScalar(v, a, sl.fi, muO, muN);
        const double d = muf * v.area[a] / tabv(v, a, kDPN, sl.fi);
        acc[P].add(d);
        acc[colSlot(D, a, side)].add(-d);
        continue;
      }
      const PatchGeom& PG = v.patch[2 * a + side];
      const double gamma = m.mu[c];
      double coeff, bv, farCoeff = 0.0;
      double expl[3] = {0.0, 0.0, 0.0};
      if (prescribes(PG.vkind) && PG.stencil) {
        const double gA = gamma * PG.area;
        coeff = gA * PG.cP;
        farCoeff = gA * PG.cF;
        bv = gA * PG.cB;
        if (!PG.deltaZero) {
          const int64_t far = side ? c - v.stride[a] : c + v.stride[a];
          for (int k = 0; k < D; ++k) {
            double gc[3], gf[3];
            velocityGradient(v, m.u, k, c, gc);
            velocityGradient(v, m.u, k, far, gf);
            expl[k] = gA * ((PG.cP * dot3(gc, PG.deltaP)) - (PG.cF * dot3(gf, PG.deltaF)));
          }
        }
      } else {
        coeff = gamma * PG.area / PG.bdist;
        bv = coeff;
      }
      for (int k = 0; k < D; ++k) {
        const double phiB = boundaryVelocity(PG, k, m.u[k][c]);
        rhs[k] += expl[k];
        rhs[k] += bv * phiB;
      }
      acc[P].add(coeff);
      if (farCoeff != 0.0) acc[colSlot(D, a, 1 - side)].add(-farCoeff);
    }

  // Upwind convection (MomentumEquation.cpp:264-327).
  for (int a = 0; a < D; ++a)
    for (int side = 0; side < 2; ++side) {
      const Slot sl = slotOf(v, c, idx, a, side);
      const double F = m.F[sl.face];
      if (sl.boundary) {
        if (F >= 0.0) {
          acc[P].add(F);
        } else {
          const PatchGeom& PG = v.patch[2 * a + side];
          for (int k = 0; k < D; ++k) rhs[k] -= F * boundaryVelocity(PG, k, m.u[k][c]);
        }
        continue;
      }
      if (side == 0) {  // this cell is the face's neighbour
        if (F >= 0.0) acc[colSlot(D, a, 0)].add(-F);
        else acc[P].add(-F);
      } else {  // this cell is the face's owner
        if (F >= 0.0) acc[P].add(F);
        else acc[colSlot(D, a, 1)].add(F);
      }
    }

  // Pressure source: rhs += -V * dp/dx_k (MomentumEquation.cpp:415-430).
  for (int k = 0; k < D; ++k) rhs[k] += -v.vol * m.gp[k][c];

  // Implicit-Euler time term (TimeDerivative.cpp:13-34), added after the
  // pressure source as applyTransientTerm does (TransientMomentum.cpp).
  if (m.dt > 0.0) {
    const double aPTime = (m.rho * v.vol) / m.dt;
    acc[P].add(aPTime);
    for (int k = 0; k < D; ++k) rhs[k] += aPTime * m.uOld[k][c];
  }

  // Implicit under-relaxation (UnderRelaxation.cpp:13-36): the unrelaxed
  // diagonal must exist (a zero sum is dropped by the CSR build).
  bool ok = acc[P].has && acc[P].s != 0.0;
  if (m.alpha != 1.0) {
    const double factor = (1.0 / m.alpha) - 1.0;
    const double extra = acc[P].s * factor;
    acc[P].add(extra);
    for (int k = 0; k < D; ++k) rhs[k] += extra * m.uRelax[k][c];
  }

  for (int q = 0; q < 2 * D + 1; ++q) {
    const double x = acc[q].has ? acc[q].s : 0.0;
    m.coef[q * v.n + c] = x;
    ok = ok && isfinite(x);
  }
  m.diag[c] = acc[P].s;
  for (int k = 0; k < D; ++k) {
    m.rhs[k][c] = rhs[k];
    ok = ok && isfinite(rhs[k]);
  }
  if (!ok) atomicAdd(m.nonFinite, 1u);
}

__global__ void responseKernel(View v, const double* __restrict__ diag, double* d) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  d[c] = v.vol / diag[c];
}

struct FluxArgs {
  const double* u[3];
  const double* p;
  const double* gp[3];
  const double* D;  // response
  double* F;
  int rhieChow;
  double alpha;
};

// calculateMassFlux + rhieChowFaceCorrection for one face.
__global__ void fluxKernel(View v, FluxArgs m) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= v.nf) return;
  int64_t idx[3];
  const int a = decodeFace(v, f, idx);
  const int64_t fi = idx[a];
  const bool boundary = fi == 0 || fi == v.ncell[a];
  double uf[3] = {0.0, 0.0, 0.0};
  double S[3];
  int64_t o = 0, nb = 0;
  if (boundary) {
    const PatchGeom& PG = v.patch[2 * a + (fi == 0 ? 0 : 1)];
    idx[a] = fi == 0 ? 0 : fi - 1;
    o = cellOf(v, idx);
    for (int k = 0; k < 3; ++k) {
      uf[k] = boundaryVelocity(PG, k, m.u[k] ? m.u[k][o] : 0.0);
      S[k] = PG.S[k];
    }
    if (v.dim == 2) uf[2] = 0.0;
  } else {
    idx[a] = fi - 1;
    o = cellOf(v, idx);
    nb = o + v.stride[a];
    for (int k = 0; k < 3; ++k) {
      uf[k] = m.u[k] ? interpVector(v, a, fi, m.u[k][o], m.u[k][nb])
                     : interpVector(v, a, fi, 0.0, 0.0);
      S[k] = v.S[a][k];
    }
  }
  double F = v.rho * dot3(uf, S);
  if (m.rhieChow) {
    double corr = 0.0;
    if (!boundary) {
      const double dF = interpScalar(v, a, fi, m.D[o], m.D[nb]);
      const double coupling = v.rho * v.area[a] * dF / tabv(v, a, kDPN, fi);
      double gf[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < 3; ++k)
        gf[k] = m.gp[k] ? interpVector(v, a, fi, m.gp[k][o], m.gp[k][nb])
                        : interpVector(v, a, fi, 0.0, 0.0);
      double d[3] = {0.0, 0.0, 0.0};
      d[a] = tabv(v, a, kDC, fi);
      const double cmi = (m.p[nb] - m.p[o]) - dot3(gf, d);
      corr = -(coupling / m.alpha) * cmi;
    }
    F += corr;
  }
  m.F[f] = F;
}

// p' face coupling D_f (pressureCorrectionFaceCoupling, axis-aligned branch),
// with the constant density or a per-face one (rhoF, the compressible equation).
__global__ void pressureFaceKernel(View v, const double* __restrict__ D,
                                   const double* __restrict__ rhoF, double* coeff) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= v.nf) return;
  int64_t idx[3];
  const int a = decodeFace(v, f, idx);
  const int64_t fi = idx[a];
  if (fi == 0 || fi == v.ncell[a]) {
    const PatchGeom& PG = v.patch[2 * a + (fi == 0 ? 0 : 1)];
    if (PG.pkind != CFDHIP_FIXED_VALUE) {
      coeff[f] = 0.0;
      return;
    }
    idx[a] = fi == 0 ? 0 : fi - 1;
    const int64_t o = cellOf(v, idx);
    coeff[f] = (rhoF != nullptr ? rhoF[f] : v.rho) * PG.area * D[o] / PG.bdist;
    return;
  }
  idx[a] = fi - 1;
  const int64_t o = cellOf(v, idx);
  const double dF = interpScalar(v, a, fi, D[o], D[o + v.stride[a]]);
  coeff[f] = (rhoF != nullptr ? rhoF[f] : v.rho) * v.area[a] * dF / tabv(v, a, kDPN, fi);
}

// Signed face-flux sum of a cell in face-list order (ContinuityEquation.cpp:21-29).
__device__ inline double cellImbalance(const View& v, const double* __restrict__ F, int64_t c,
                                       const int64_t idx[3]) {
  double sum = 0.0;
  for (int a = 0; a < v.dim; ++a)
    for (int side = 0; side < 2; ++side) {
      const Slot sl = slotOf(v, c, idx, a, side);
      const double flux = F[sl.face];
      // owner on the high side always; on the low side only at a boundary
      sum += (side || sl.boundary) ? flux : -flux;
    }
  return sum;
}

// Row of the p' system: diagonal and RHS (assembleGeometricPressureCorrection);
// an additional diagonal (extra, the compressibility term) is added after the
// face terms, as the reference's builder sums it (PressureCorrectionEquation.cpp:
// 304-310), except on the pinned row.
__global__ void pressureRowKernel(View v, const double* __restrict__ coeff,
                                  const double* __restrict__ Fstar,
                                  const double* __restrict__ extra, int64_t pin, double* diag,
                                  double* rhs, unsigned* nonFinite) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  int64_t idx[3];
  cellIndex(v, c, idx);
  double d = 1.0;
  double b = 0.0;
  if (c != pin) {
    Acc acc{0.0, false};
    for (int a = 0; a < v.dim; ++a)
      for (int side = 0; side < 2; ++side) {
        const Slot sl = slotOf(v, c, idx, a, side);
        if (sl.boundary && v.patch[2 * a + side].pkind != CFDHIP_FIXED_VALUE) continue;
        acc.add(coeff[sl.face]);
      }
    if (extra != nullptr) acc.add(extra[c]);
    d = acc.has ? acc.s : 0.0;
    b = -cellImbalance(v, Fstar, c, idx);
  }
  diag[c] = d;
  rhs[c] = b;
  if (!isfinite(d) || !isfinite(b)) atomicAdd(nonFinite, 1u);
}

__global__ void faceCoeffFiniteKernel(int64_t nf, const double* __restrict__ coeff,
                                      unsigned* nonFinite) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f < nf && !isfinite(coeff[f])) atomicAdd(nonFinite, 1u);
}

__global__ void updatePressureKernel(int64_t n, double alpha, const double* __restrict__ pp,
                                     double* p) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < n) p[c] = p[c] + (alpha * pp[c]);
}

struct CorrectArgs {
  const double* ustar[3];
  double* u[3];
  const double* D;
  const double* pp;
};

// correctVelocity: u = u* - d * grad(p') with the p' conditions.
__global__ void correctVelocityKernel(View v, CorrectArgs m) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  double g[3];
  scalarGradient(v, m.pp, 1, c, g);
  const double d = m.D[c];
  for (int k = 0; k < v.dim; ++k) m.u[k][c] = m.ustar[k][c] - (d * g[k]);
}

// correctFaceMassFlux: F = F* + D_f (p'_P - p'_N) (boundary p'_N = 0).
__global__ void correctFluxKernel(View v, const double* __restrict__ Fstar,
                                  const double* __restrict__ coeff, const double* __restrict__ pp,
                                  double* F) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= v.nf) return;
  int64_t idx[3];
  const int a = decodeFace(v, f, idx);
  const int64_t fi = idx[a];
  double po, pn;
  if (fi == 0 || fi == v.ncell[a]) {
    idx[a] = fi == 0 ? 0 : fi - 1;
    po = pp[cellOf(v, idx)];
    pn = 0.0;
  } else {
    idx[a] = fi - 1;
    const int64_t o = cellOf(v, idx);
    po = pp[o];
    pn = pp[o + v.stride[a]];
  }
  F[f] = Fstar[f] + (coeff[f] * (po - pn));
}

__global__ void imbalanceKernel(View v, const double* __restrict__ F, double* imb,
                                double* partial) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (c < v.n) {
    int64_t idx[3];
    cellIndex(v, c, idx);
    const double s = cellImbalance(v, F, c, idx);
    imb[c] = s;
    sq = s * s;
  }
  // block sum of squares for the RMS (reduction order differs from the
  // reference's serial sum; the per-cell imbalance itself is exact)
  __shared__ double lds[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    partial[blockIdx.x] = t;
  }
}

// Global net boundary flux in patch order, face order within a patch, and
// the sum of the imbalance partials in block order: the reference's serial
// summation order, kept by thread 0 adding in sequence. The block first
// loads each chunk of values into LDS (in parallel), so the additions wait on
// LDS instead of on one memory round trip each (256^3: 94 ms per call
// before).
constexpr int kFluxThreads = 256, kFluxChunk = 4096;
__global__ __launch_bounds__(kFluxThreads) void globalFluxKernel(View v,
                                                                const double* __restrict__ F,
                                                                const double* partial,
                                                                int64_t nPartial, double* out) {
  __shared__ double buf[kFluxChunk];
  double g = 0.0, ss = 0.0;  // thread 0's sums
  // chunk [c0, c0 + len) of a sequence whose element l is at(l)
  auto addChunks = [&](int64_t m, double& sum, auto at) {
    for (int64_t c0 = 0; c0 < m; c0 += kFluxChunk) {
      const int len = static_cast<int>(m - c0 < kFluxChunk ? m - c0 : kFluxChunk);
      for (int t = threadIdx.x; t < len; t += blockDim.x) buf[t] = at(c0 + t);
      __syncthreads();
      if (threadIdx.x == 0) {
        // eight LDS reads in flight ahead of the (in-order) additions
        int t = 0;
        for (; t + 8 <= len; t += 8) {
          double x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) x[u] = buf[t + u];
#pragma unroll
          for (int u = 0; u < 8; ++u) sum += x[u];
        }
        for (; t < len; ++t) sum += buf[t];
      }
      __syncthreads();
    }
  };
  for (int p = 0; p < 2 * v.dim; ++p) {
    const int a = p / 2;
    const bool high = p % 2;
    int64_t lim[3] = {v.nx, v.ny, v.nz};
    lim[a] = 1;
    // element l of the patch: k outer, j, i inner (l = (k lim1 + j) lim0 + i)
    addChunks(lim[0] * lim[1] * lim[2], g, [&](int64_t l) {
      int64_t idx[3] = {l % lim[0], (l / lim[0]) % lim[1], l / (lim[0] * lim[1])};
      idx[a] = high ? v.ncell[a] : 0;
      return F[faceId(v, a, idx)];
    });
  }
  addChunks(nPartial, ss, [&](int64_t b) { return partial[b]; });
  if (threadIdx.x == 0) {
    out[0] = sqrt(ss / static_cast<double>(v.n));
    out[1] = g;
  }
}

// CSR row counts / fills with exact zeros dropped (SparseMatrixBuilder::build).
__device__ inline double pressureEntry(const View& v, const double* coeff, const double* diag,
                                       int64_t pin, int64_t c, const int64_t idx[3], int q,
                                       int64_t& col) {
  const int D = v.dim;
  if (q == D) {
    col = c;
    return diag[c];
  }
  const int side = q > D ? 1 : 0;
  const int a = side ? q - D - 1 : D - 1 - q;
  const Slot sl = slotOf(v, c, idx, a, side);
  if (sl.boundary || c == pin) return 0.0;
  col = sl.cell;
  return -coeff[sl.face];
}

__global__ void pressureCsrCount(View v, const double* coeff, const double* diag, int64_t pin,
                                 int64_t* count) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  int64_t idx[3];
  cellIndex(v, c, idx);
  int64_t k = 0, col;
  for (int q = 0; q < 2 * v.dim + 1; ++q)
    if (pressureEntry(v, coeff, diag, pin, c, idx, q, col) != 0.0) ++k;
  count[c] = k;
}

__global__ void pressureCsrFill(View v, const double* coeff, const double* diag, int64_t pin,
                                const int64_t* rp, int64_t* colOut, double* valOut) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  int64_t idx[3];
  cellIndex(v, c, idx);
  int64_t k = rp[c], col = 0;
  for (int q = 0; q < 2 * v.dim + 1; ++q) {
    const double x = pressureEntry(v, coeff, diag, pin, c, idx, q, col);
    if (x != 0.0) {
      colOut[k] = col;
      valOut[k] = x;
      ++k;
    }
  }
}

__device__ inline int64_t stencilColumn(const View& v, int64_t c, const int64_t idx[3], int q,
                                        bool& exists) {
  const int D = v.dim;
  exists = true;
  if (q == D) return c;
  const int side = q > D ? 1 : 0;
  const int a = side ? q - D - 1 : D - 1 - q;
  const Slot sl = slotOf(v, c, idx, a, side);
  exists = !sl.boundary;
  return sl.cell;
}

__global__ void stencilCsrCount(View v, const double* coef, int64_t* count) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  int64_t idx[3];
  cellIndex(v, c, idx);
  int64_t k = 0;
  for (int q = 0; q < 2 * v.dim + 1; ++q) {
    bool ex;
    stencilColumn(v, c, idx, q, ex);
    if (ex && coef[q * v.n + c] != 0.0) ++k;
  }
  count[c] = k;
}

__global__ void stencilCsrFill(View v, const double* coef, const int64_t* rp, int64_t* colOut,
                               double* valOut) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= v.n) return;
  int64_t idx[3];
  cellIndex(v, c, idx);
  int64_t k = rp[c];
  for (int q = 0; q < 2 * v.dim + 1; ++q) {
    bool ex;
    const int64_t col = stencilColumn(v, c, idx, q, ex);
    const double x = coef[q * v.n + c];
    if (ex && x != 0.0) {
      colOut[k] = col;
      valOut[k] = x;
      ++k;
    }
  }
}

__global__ void countNonFiniteKernel(int64_t n, const double* __restrict__ a, unsigned* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && !isfinite(a[i])) atomicAdd(out, 1u);
}

unsigned readAndClear(Context& ctx) {
  unsigned h = 0;
  CFDHIP_CHECK(hipMemcpyAsync(&h, ctx.flags.get(), sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  CFDHIP_CHECK(hipMemsetAsync(ctx.flags.get(), 0, sizeof(unsigned), ctx.stream));
  return h;
}

// Inclusive-scan the counts into rowptr[1..n] (rowptr[0] = 0) on the device.
int64_t scanRows(Context& ctx, int64_t* dCount, int64_t* dRowptr, int64_t n) {
  CFDHIP_CHECK(hipMemsetAsync(dRowptr, 0, sizeof(int64_t), ctx.stream));
  size_t bytes = 0;
  CFDHIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, dCount, dRowptr + 1,
                                                static_cast<int>(n), ctx.stream));
  DeviceArray<unsigned char> tmp;
  tmp.ensure(static_cast<int64_t>(bytes) + 1);
  CFDHIP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp.get(), bytes, dCount, dRowptr + 1,
                                                static_cast<int>(n), ctx.stream));
  int64_t nnz = 0;
  CFDHIP_CHECK(hipMemcpyAsync(&nnz, dRowptr + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return nnz;
}

}  // namespace

void launchGradient(Context& ctx, const double* phi, int bcKind, double* gx, double* gy,
                    double* gz) {
  ctx.requireMesh();
  hipLaunchKernelGGL(gradientKernel, blocksFor(ctx.view.n), dim3(kBlock), 0, ctx.stream,
                     ctx.view, phi, bcKind, gx, gy, ctx.view.dim == 3 ? gz : nullptr);
  CFDHIP_CHECK(hipGetLastError());
}

void launchMassFlux(Context& ctx, const double* u, const double* v, const double* w, double* F) {
  ctx.requireMesh();
  FluxArgs a{};
  a.u[0] = u;
  a.u[1] = v;
  a.u[2] = ctx.view.dim == 3 ? w : nullptr;
  a.F = F;
  a.rhieChow = 0;
  a.alpha = 1.0;
  hipLaunchKernelGGL(fluxKernel, blocksFor(ctx.view.nf), dim3(kBlock), 0, ctx.stream, ctx.view, a);
  CFDHIP_CHECK(hipGetLastError());
}

void launchMomentum(Context& ctx, double alpha, double dt, int oldComponent,
                    const double* oldValues) {
  ctx.requireMesh();
  require(std::isfinite(alpha) && alpha > 0.0 && alpha <= 1.0,
          "applyImplicitUnderRelaxation: alpha must be finite and in (0, 1]");
  require(dt == 0.0 || (std::isfinite(dt) && dt > 0.0 && alpha == 1.0),
          "assembleTransientMomentumComponent: dt must be finite and > 0 (no relaxation)");
  require(ctx.fluidSet, "fluid properties not set");
  const View& v = ctx.view;
  const int D = v.dim;
  ctx.momCoef.ensure((2 * D + 1) * v.n);
  // pressure gradient of the current p (also used by the Rhie-Chow flux)
  launchGradient(ctx, ctx.f(CFDHIP_FIELD_P), 0, ctx.f(CFDHIP_FIELD_GRAD_X),
                 ctx.f(CFDHIP_FIELD_GRAD_Y), ctx.f(CFDHIP_FIELD_GRAD_Z));
  MomentumArgs m{};
  for (int k = 0; k < 3; ++k) {
    m.u[k] = k < D ? ctx.f(CFDHIP_FIELD_U + k) : nullptr;
    m.gp[k] = ctx.f(CFDHIP_FIELD_GRAD_X + k);
    m.rhs[k] = ctx.f(CFDHIP_FIELD_MOM_RHS_U + k);
  }
  m.F = ctx.f(CFDHIP_FIELD_FLUX);
  m.mu = ctx.f(CFDHIP_FIELD_VISCOSITY);
  m.coef = ctx.momCoef.get();
  m.diag = ctx.f(CFDHIP_FIELD_MOM_DIAG);
  m.nonFinite = ctx.flags.get();
  m.alpha = alpha;
  m.rho = ctx.rho;
  m.dt = dt;
  for (int k = 0; k < 3; ++k) m.uOld[k] = m.uRelax[k] = m.u[k];
  if (oldComponent >= 0 && oldComponent < D && oldValues != nullptr)
    m.uRelax[oldComponent] = oldValues;
  hipLaunchKernelGGL(momentumKernel, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v, m);
  CFDHIP_CHECK(hipGetLastError());
  const unsigned bad = readAndClear(ctx);
  ctx.momentumAssembled = bad == 0;
  require(bad == 0, "assembleRelaxedMomentumComponent: assembled system contains a non-finite value",
          CFDHIP_ERR_NONFINITE);
}

void launchResponse(Context& ctx) {
  ctx.requireMesh();
  require(ctx.momentumAssembled, "no momentum system assembled");
  hipLaunchKernelGGL(responseKernel, blocksFor(ctx.view.n), dim3(kBlock), 0, ctx.stream,
                     ctx.view, ctx.f(CFDHIP_FIELD_MOM_DIAG), ctx.f(CFDHIP_FIELD_RESPONSE));
  CFDHIP_CHECK(hipGetLastError());
}

void launchPredictedFlux(Context& ctx, bool rhieChow, double alpha) {
  ctx.requireMesh();
  require(std::isfinite(alpha) && alpha > 0.0 && alpha <= 1.0,
          "rhieChowFaceCorrection: alpha must be finite and in (0, 1]");
  const View& v = ctx.view;
  FluxArgs a{};
  for (int k = 0; k < 3; ++k) {
    a.u[k] = k < v.dim ? ctx.f(CFDHIP_FIELD_USTAR + k) : nullptr;
    a.gp[k] = k < v.dim ? ctx.f(CFDHIP_FIELD_GRAD_X + k) : nullptr;
  }
  a.p = ctx.f(CFDHIP_FIELD_P);
  a.D = ctx.f(CFDHIP_FIELD_RESPONSE);
  a.F = ctx.f(CFDHIP_FIELD_FLUX_STAR);
  a.rhieChow = rhieChow ? 1 : 0;
  a.alpha = alpha;
  if (rhieChow)  // gradient of the current pressure (SIMPLE.cpp:751-753)
    launchGradient(ctx, ctx.f(CFDHIP_FIELD_P), 0, ctx.f(CFDHIP_FIELD_GRAD_X),
                   ctx.f(CFDHIP_FIELD_GRAD_Y), ctx.f(CFDHIP_FIELD_GRAD_Z));
  hipLaunchKernelGGL(fluxKernel, blocksFor(v.nf), dim3(kBlock), 0, ctx.stream, v, a);
  CFDHIP_CHECK(hipGetLastError());
}

void launchPressureAssembly(Context& ctx, int64_t ref, const double* rhoF, const double* extra) {
  ctx.requireMesh();
  const View& v = ctx.view;
  require(ref >= 0 && ref < v.n, "assemblePressureCorrection: referenceCell out of range");
  bool open = false;
  for (int p = 0; p < v.npatch; ++p) open = open || v.patch[p].pkind == CFDHIP_FIXED_VALUE;
  ctx.pin = open ? -1 : ref;
  hipLaunchKernelGGL(pressureFaceKernel, blocksFor(v.nf), dim3(kBlock), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_RESPONSE), rhoF, ctx.f(CFDHIP_FIELD_PC_FACE_COEFF));
  CFDHIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(faceCoeffFiniteKernel, blocksFor(v.nf), dim3(kBlock), 0, ctx.stream, v.nf,
                     ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), ctx.flags.get());
  hipLaunchKernelGGL(pressureRowKernel, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), ctx.f(CFDHIP_FIELD_FLUX_STAR), extra,
                     ctx.pin, ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.f(CFDHIP_FIELD_PC_RHS),
                     ctx.flags.get());
  CFDHIP_CHECK(hipGetLastError());
  const unsigned bad = readAndClear(ctx);
  ctx.pressureAssembled = bad == 0;
  require(bad == 0,
          "assembleGeometricPressureCorrection: assembled system contains a non-finite value",
          CFDHIP_ERR_NONFINITE);
}

// Compressibility diagonal of the compressible p' equation
// (CompressiblePressureCorrection.cpp:52-62): V / dtau * drho/dp per cell, with
// the ideal gas drho/dp = 1 / (R T) (IdealGasEOS.cpp:35-39); counts the cells
// whose absolute pressure or temperature is not finite and > 0 (the EOS
// rejects them).
__global__ void compressibilityKernel(int64_t n, double vol, double dtau, double R,
                                      const double* __restrict__ pAbs,
                                      const double* __restrict__ T, double* extra,
                                      unsigned* invalid) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double p = pAbs[c], t = T[c];
  if (!(isfinite(p) && p > 0.0 && isfinite(t) && t > 0.0)) {
    atomicAdd(invalid, 1u);
    extra[c] = 0.0;
    return;
  }
  const double dRhoDp = 1.0 / (R * t);
  extra[c] = vol / dtau * dRhoDp;
}

void launchCompressiblePressureAssembly(Context& ctx, int64_t ref, const double* rhoF,
                                        const double* pAbs, const double* T, double R,
                                        double dtau) {
  ctx.requireMesh();
  const View& v = ctx.view;
  require(v.dim == 2, "assembleCompressiblePressureCorrection: the reference's equation is 2D only",
          CFDHIP_ERR_UNSUPPORTED);
  require(std::isfinite(R) && R > 0.0, "IdealGasEOS: gasConstant must be finite and > 0");
  require(std::isfinite(dtau) && dtau > 0.0,
          "assembleCompressiblePressureCorrection: pseudoTimeStep must be finite and > 0");
  require(ref >= 0 && ref < v.n, "assembleCompressiblePressureCorrection: referenceCell out of range");
  CompressibleWork& w = ctx.comp;
  w.rhoF.ensure(v.nf);
  w.pAbs.ensure(v.n);
  w.T.ensure(v.n);
  w.extra.ensure(v.n);
  CFDHIP_CHECK(hipMemcpyAsync(w.rhoF.get(), rhoF, v.nf * sizeof(double), hipMemcpyDefault,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(w.pAbs.get(), pAbs, v.n * sizeof(double), hipMemcpyDefault,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(w.T.get(), T, v.n * sizeof(double), hipMemcpyDefault, ctx.stream));
  hipLaunchKernelGGL(compressibilityKernel, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v.n,
                     v.vol, dtau, R, w.pAbs.get(), w.T.get(), w.extra.get(), ctx.flags.get());
  CFDHIP_CHECK(hipGetLastError());
  const unsigned bad = readAndClear(ctx);
  require(bad == 0, "IdealGasEOS: pressure and temperature must be finite and > 0");
  launchPressureAssembly(ctx, ref, w.rhoF.get(), w.extra.get());
}

void launchUpdatePressure(Context& ctx, double alpha) {
  ctx.requireMesh();
  hipLaunchKernelGGL(updatePressureKernel, blocksFor(ctx.view.n), dim3(kBlock), 0, ctx.stream,
                     ctx.view.n, alpha, ctx.f(CFDHIP_FIELD_PPRIME), ctx.f(CFDHIP_FIELD_P));
  CFDHIP_CHECK(hipGetLastError());
}

void launchCorrectVelocity(Context& ctx, double* const dest[3]) {
  ctx.requireMesh();
  const View& v = ctx.view;
  CorrectArgs m{};
  for (int k = 0; k < 3; ++k) {
    m.ustar[k] = k < v.dim ? ctx.f(CFDHIP_FIELD_USTAR + k) : nullptr;
    m.u[k] = k < v.dim ? (dest != nullptr ? dest[k] : ctx.f(CFDHIP_FIELD_U + k)) : nullptr;
  }
  m.D = ctx.f(CFDHIP_FIELD_RESPONSE);
  m.pp = ctx.f(CFDHIP_FIELD_PPRIME);
  hipLaunchKernelGGL(correctVelocityKernel, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v, m);
  CFDHIP_CHECK(hipGetLastError());
}

void launchCorrectFlux(Context& ctx, double* dest) {
  ctx.requireMesh();
  const View& v = ctx.view;
  hipLaunchKernelGGL(correctFluxKernel, blocksFor(v.nf), dim3(kBlock), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_FLUX_STAR), ctx.f(CFDHIP_FIELD_PC_FACE_COEFF),
                     ctx.f(CFDHIP_FIELD_PPRIME), dest != nullptr ? dest : ctx.f(CFDHIP_FIELD_FLUX));
  CFDHIP_CHECK(hipGetLastError());
}

void launchCorrect(Context& ctx) {
  launchCorrectVelocity(ctx);
  launchCorrectFlux(ctx);
}

// Block maxima of v over a grid-stride range (0 when empty).
__device__ inline void blockMax(double v, double* partial) {
  __shared__ double lds[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t = fmax(t, lds[w]);
    partial[blockIdx.x] = t;
  }
}

// Courant number per cell (CFL.cpp:28-42): (dt / (2 V)) (sum over the cell's
// faces, in face-list order, of |F|, / rho); block maxima.
__global__ void cflKernel(View v, const double* __restrict__ F, double rho, double dt,
                          double* partial) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double cfl = 0.0;
  if (c < v.n) {
    int64_t idx[3];
    cellIndex(v, c, idx);
    double s = 0.0;
    for (int a = 0; a < v.dim; ++a)
      for (int side = 0; side < 2; ++side) s += fabs(F[slotOf(v, c, idx, a, side).face]);
    cfl = (dt / (2.0 * v.vol)) * (s / rho);
  }
  blockMax(cfl, partial);
}

__global__ void absMaxKernel(int64_t n, const double* __restrict__ x, double* partial) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  blockMax(c < n ? fabs(x[c]) : 0.0, partial);
}

// max(0, partials) by one block of kBlock threads (a max is exact in any
// order; fmax drops a NaN operand as the serial form does).
__global__ __launch_bounds__(kBlock) void maxFinalKernel(const double* __restrict__ partial,
                                                         int64_t parts, double* out) {
  __shared__ double lds[kBlock / 64];
  double t = 0.0;
  for (int64_t b = threadIdx.x; b < parts; b += blockDim.x) t = fmax(t, partial[b]);
  for (int o = 32; o > 0; o >>= 1) t = fmax(t, __shfl_xor(t, o, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) m = fmax(m, lds[w]);
    out[0] = m;
  }
}

// max over cells of the Courant number (FLUX) or of |IMBALANCE| (exact: a max
// does not depend on the order).
double launchMaxCell(Context& ctx, bool cfl, double dt) {
  ctx.requireMesh();
  const View& v = ctx.view;
  const dim3 grid = blocksFor(v.n);
  DeviceArray<double>& part = ctx.kw.partial;
  part.ensure(static_cast<int64_t>(grid.x) + 2);
  if (cfl)
    hipLaunchKernelGGL(cflKernel, grid, dim3(kBlock), 0, ctx.stream, v, ctx.f(CFDHIP_FIELD_FLUX),
                       ctx.rho, dt, part.get());
  else
    hipLaunchKernelGGL(absMaxKernel, grid, dim3(kBlock), 0, ctx.stream, v.n,
                       ctx.f(CFDHIP_FIELD_IMBALANCE), part.get());
  hipLaunchKernelGGL(maxFinalKernel, dim3(1), dim3(kBlock), 0, ctx.stream, part.get(),
                     static_cast<int64_t>(grid.x), part.get() + grid.x);
  CFDHIP_CHECK(hipGetLastError());
  double out = 0.0;
  CFDHIP_CHECK(hipMemcpyAsync(&out, part.get() + grid.x, sizeof(double), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return out;
}

void launchContinuity(Context& ctx, double* out) {
  ctx.requireMesh();
  const View& v = ctx.view;
  const dim3 grid = blocksFor(v.n);
  DeviceArray<double>& part = ctx.kw.partial;
  part.ensure(static_cast<int64_t>(grid.x) + 2);
  hipLaunchKernelGGL(imbalanceKernel, grid, dim3(kBlock), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_FLUX), ctx.f(CFDHIP_FIELD_IMBALANCE), part.get());
  CFDHIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(globalFluxKernel, dim3(1), dim3(kFluxThreads), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_FLUX), part.get(), static_cast<int64_t>(grid.x),
                     part.get() + grid.x);
  CFDHIP_CHECK(hipGetLastError());
  if (out != nullptr) {
    CFDHIP_CHECK(hipMemcpyAsync(out, part.get() + grid.x, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, ctx.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  }
}

int64_t exportPressureCsr(Context& ctx, int64_t* rp, int64_t* col, double* val, int64_t cap) {
  ctx.requireMesh();
  require(ctx.pressureAssembled, "no p' system assembled");
  const View& v = ctx.view;
  DeviceArray<int64_t> cnt, drp, dcol;
  DeviceArray<double> dval;
  cnt.ensure(v.n);
  drp.ensure(v.n + 1);
  hipLaunchKernelGGL(pressureCsrCount, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.pin,
                     cnt.get());
  CFDHIP_CHECK(hipGetLastError());
  const int64_t nnz = scanRows(ctx, cnt.get(), drp.get(), v.n);
  require(nnz <= cap, "CSR capacity too small");
  dcol.ensure(nnz);
  dval.ensure(nnz);
  hipLaunchKernelGGL(pressureCsrFill, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v,
                     ctx.f(CFDHIP_FIELD_PC_FACE_COEFF), ctx.f(CFDHIP_FIELD_PC_DIAG), ctx.pin,
                     drp.get(), dcol.get(), dval.get());
  CFDHIP_CHECK(hipGetLastError());
  CFDHIP_CHECK(hipMemcpyAsync(rp, drp.get(), (v.n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(col, dcol.get(), nnz * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(val, dval.get(), nnz * sizeof(double), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return nnz;
}

int64_t exportMomentumCsr(Context& ctx, int64_t* rp, int64_t* col, double* val, int64_t cap) {
  ctx.requireMesh();
  require(ctx.momentumAssembled, "no momentum system assembled");
  const View& v = ctx.view;
  DeviceArray<int64_t> cnt, drp, dcol;
  DeviceArray<double> dval;
  cnt.ensure(v.n);
  drp.ensure(v.n + 1);
  hipLaunchKernelGGL(stencilCsrCount, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v,
                     ctx.momCoef.get(), cnt.get());
  CFDHIP_CHECK(hipGetLastError());
  const int64_t nnz = scanRows(ctx, cnt.get(), drp.get(), v.n);
  require(nnz <= cap, "CSR capacity too small");
  dcol.ensure(nnz);
  dval.ensure(nnz);
  hipLaunchKernelGGL(stencilCsrFill, blocksFor(v.n), dim3(kBlock), 0, ctx.stream, v,
                     ctx.momCoef.get(), drp.get(), dcol.get(), dval.get());
  CFDHIP_CHECK(hipGetLastError());
  CFDHIP_CHECK(hipMemcpyAsync(rp, drp.get(), (v.n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(col, dcol.get(), nnz * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(val, dval.get(), nnz * sizeof(double), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  return nnz;
}

unsigned countNonFinite(Context& ctx, const double* const* arrays, const int64_t* counts, int k) {
  for (int i = 0; i < k; ++i) {
    if (counts[i] <= 0) continue;
    hipLaunchKernelGGL(countNonFiniteKernel, blocksFor(counts[i]), dim3(kBlock), 0, ctx.stream,
                       counts[i], arrays[i], ctx.flags.get());
    CFDHIP_CHECK(hipGetLastError());
  }
  return readAndClear(ctx);
}

}  // namespace cfdhip
