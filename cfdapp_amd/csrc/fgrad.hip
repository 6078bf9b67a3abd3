// Cell gradients on a general face-list mesh (2D): the Green-Gauss gradient
// with the reference's boundary-consistent face values, skewness correction
// and oblique Neumann faces (src/discretization/Gradient.cpp:104-218,
// 449-596, greenGaussGradient), and the weighted least-squares gradient with
// its Green-Gauss fallback (Gradient.cpp:297-447), for the p' corrector passes
// on non-orthogonal and multiblock meshes (SIMPLE.cpp:876-924).
//
// Everything that depends only on the geometry -- interpolation distances,
// owner-neighbour crossings and skew vectors (MeshGeometry.cpp:197-227), the
// opposite interior face, far cell, crossing weight and boundary-line
// intersection of every boundary-consistent face value and which boundary
// face's value claims an opposite face (MeshGeometry.cpp:45-123, 229-251),
// the oblique-face split (Gradient.cpp:190-217) -- is computed once on the
// host with the reference's expressions and operand order (Vector3 dot with
// its zero-z rule, magnitude = sqrt(dot)). The device does the arithmetic on
// the field, one thread per face or per cell, in the reference's order: the
// per-cell Green-Gauss sum runs over the cell's faces in Cell::faceIds()
// order, so every gradient is bitwise the reference's.
#include <cmath>
#include <vector>

#include "context.hpp"

namespace cfdhip {

namespace {

constexpr int kGBlock = 256;
dim3 gBlocks(int64_t n) { return dim3(static_cast<unsigned>((n + kGBlock - 1) / kGBlock)); }

template <class T>
void put(Context& ctx, DeviceArray<T>& d, const std::vector<T>& h) {
  d.ensure(h.empty() ? 1 : static_cast<int64_t>(h.size()));
  if (!h.empty())
    CFDHIP_CHECK(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice,
                                ctx.stream));
}

// --- the reference's Vector3 arithmetic (include/cfd/core/Vector3.hpp) -------

struct R3 {
  double x = 0.0, y = 0.0, z = 0.0;
};
__host__ __device__ inline R3 operator+(R3 a, R3 b) { return R3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline R3 operator-(R3 a, R3 b) { return R3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline R3 operator*(R3 a, double s) { return R3{a.x * s, a.y * s, a.z * s}; }
// a.x b.x + a.y b.y, plus a.z b.z only when that product is not zero
__host__ __device__ inline double dotR(R3 a, R3 b) {
  const double inPlane = (a.x * b.x) + (a.y * b.y);
  const double normal = a.z * b.z;
  return normal == 0.0 ? inPlane : inPlane + normal;
}
__host__ __device__ inline double magR(R3 a) { return sqrt(dotR(a, a)); }
inline R3 crossR(R3 a, R3 b) {
  return R3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
inline bool isZero(R3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }
__host__ __device__ inline R3 at3(const double* a, int64_t i) {
  return R3{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
}

// --- host geometry (MeshGeometry.cpp) ----------------------------------------

struct Geo {
  const FaceMesh& m;
  R3 fc(int64_t f) const { return at3(m.hFc.data(), f); }
  R3 sf(int64_t f) const { return at3(m.hSf.data(), f); }
  R3 cc(int64_t c) const { return at3(m.hCc.data(), c); }
  bool boundary(int64_t f) const { return m.hNeigh[f] < 0; }
  R3 unitNormal(int64_t f) const { return sf(f) * (1.0 / magR(sf(f))); }
  // isFaceGeometryWellPosed (MeshGeometry.cpp:106-117)
  static bool wellPosed(R3 d, R3 s) {
    const double dDotSf = dotR(d, s);
    const double dMag = magR(d), sfMag = magR(s);
    if (!(dMag > 0.0) || !(sfMag > 0.0)) return false;
    return dDotSf > (1e-6 * dMag * sfMag);
  }
  // ownerNeighborCrossing (MeshGeometry.cpp:197-227): t and the skew vector
  bool crossing(int64_t f, double& t, R3& skew) const {
    const R3 co = cc(m.hOwner[f]);
    const R3 d = cc(m.hNeigh[f]) - co;
    const R3 s = sf(f);
    if (!wellPosed(d, s)) return false;
    const R3 toFace = fc(f) - co;
    if (isZero(crossR(toFace, d))) {
      t = dotR(toFace, d) / dotR(d, d);
      skew = R3{};
      return true;
    }
    t = dotR(fc(f) - co, s) / dotR(d, s);
    const R3 point = co + (d * t);
    skew = fc(f) - point;
    return true;
  }
  // boundaryLineIntersection (MeshGeometry.cpp:229-251)
  bool intersection(int64_t bf, R3 direction, double& distance, R3& point) const {
    const R3 co = cc(m.hOwner[bf]);
    const R3 s = sf(bf);
    if (!wellPosed(direction * -1.0, s)) return false;
    distance = dotR(co - fc(bf), s) / dotR(direction, s);
    if (!(distance > 0.0)) return false;
    point = co - (direction * distance);
    return true;
  }
  // oppositeInteriorFace (MeshGeometry.cpp:45-123) of cell c across face ref
  int64_t opposite(int64_t c, int64_t ref) const {
    const R3 raw = unitNormal(ref);
    const R3 outRef = m.hOwner[ref] == c ? raw : raw * -1.0;
    int64_t best = -1;
    double bestDot = 0.0;
    for (int64_t k = m.hCfOff[c]; k < m.hCfOff[c + 1]; ++k) {
      const int64_t f = m.hCf[k];
      if (f == ref || boundary(f)) continue;
      const R3 rn = unitNormal(f);
      const R3 out = m.hOwner[f] == c ? rn : rn * -1.0;
      const double alignment = dotR(outRef, out);
      if (alignment < 0.0 && (best < 0 || alignment < bestDot)) {
        bestDot = alignment;
        best = f;
      }
    }
    return best;
  }
  int64_t other(int64_t f, int64_t c) const {
    return m.hOwner[f] == c ? m.hNeigh[f] : m.hOwner[f];
  }
};

void buildGradPlan(Context& ctx, FaceMesh& m) {
  FaceGradPlan& p = m.gp;
  if (p.built) return;
  require(m.dim == 2, "face-list gradient: 2D meshes only", CFDHIP_ERR_UNSUPPORTED);
  for (int64_t c = 0; c < m.n; ++c)
    require(m.hCc[3 * c + 2] == 0.0, "face-list gradient: 2D meshes only (cell centroid z != 0)",
            CFDHIP_ERR_UNSUPPORTED);
  for (int64_t f = 0; f < m.nf; ++f)
    require(m.hFc[3 * f + 2] == 0.0 && m.hSf[3 * f + 2] == 0.0,
            "face-list gradient: 2D meshes only (face z != 0)", CFDHIP_ERR_UNSUPPORTED);
  const Geo G{m};
  // interpolation distances (Interpolation.cpp:21-30, 106-113)
  std::vector<double> wts(2 * m.nf, 0.0);
  std::vector<int64_t> skewFace;
  std::vector<double> skewGeo;
  for (int64_t f = 0; f < m.nf; ++f) {
    wts[2 * f] = magR(G.fc(f) - G.cc(m.hOwner[f]));
    if (G.boundary(f)) continue;
    wts[2 * f + 1] = magR(G.fc(f) - G.cc(m.hNeigh[f]));
    double t = 0.0;
    R3 skew;
    if (G.crossing(f, t, skew) && !isZero(skew)) {
      skewFace.push_back(f);
      skewGeo.insert(skewGeo.end(), {t, skew.x, skew.y, skew.z});
    }
  }
  // boundary-consistent face values (Gradient.cpp:104-170) and their claims
  // (Gradient.cpp:491-517: per cell and opposite face, the most
  // anti-parallel claim wins, ties to the lower boundary face id)
  const int64_t ncf = m.hCfOff[m.n];
  std::vector<int32_t> slotClaim(ncf, -1);
  std::vector<int64_t> claimFaces;
  std::vector<double> claimGeo;
  bool transfer = false;
  for (int64_t c = 0; c < m.n; ++c) {
    struct Claim {
      int64_t opp, bf, far;
      double alignment, farDist, w, back;
      R3 offset;
    };
    std::vector<Claim> claims;
    for (int64_t k = m.hCfOff[c]; k < m.hCfOff[c + 1]; ++k) {
      const int64_t bf = m.hCf[k];
      if (!G.boundary(bf)) continue;
      const int64_t opp = G.opposite(c, bf);
      if (opp < 0) continue;
      const int64_t far = G.other(opp, c);
      const R3 d = G.cc(far) - G.cc(c);
      const double farDist = magR(d);
      if (!(farDist > 0.0)) continue;
      const R3 inward = d * (1.0 / farDist);
      double back = 0.0;
      R3 point;
      const bool hit = G.intersection(bf, inward, back, point);
      // boundaryTransferNeeded (Gradient.cpp:449-476)
      if (hit && !isZero(point - G.fc(bf))) transfer = true;
      double w = 0.0;
      R3 skew;
      if (!G.crossing(opp, w, skew) || !hit) continue;
      const R3 refN = G.unitNormal(bf);
      const R3 oppN = m.hOwner[opp] == c ? G.unitNormal(opp) : G.unitNormal(opp) * -1.0;
      const Claim cl{opp, bf, far, dotR(refN, oppN), farDist, w, back, point - G.fc(bf)};
      bool placed = false;
      for (Claim& e : claims)
        if (e.opp == opp) {
          placed = true;
          if (cl.alignment < e.alignment || (cl.alignment == e.alignment && cl.bf < e.bf)) e = cl;
        }
      if (!placed) claims.push_back(cl);
    }
    for (const Claim& cl : claims) {
      for (int64_t k = m.hCfOff[c]; k < m.hCfOff[c + 1]; ++k)
        if (m.hCf[k] == cl.opp) slotClaim[k] = static_cast<int32_t>(claimFaces.size() / 3);
      claimFaces.insert(claimFaces.end(), {cl.bf, cl.opp, cl.far});
      claimGeo.insert(claimGeo.end(),
                      {cl.farDist, cl.w, cl.back, cl.offset.x, cl.offset.y, cl.offset.z});
    }
  }
  put(ctx, p.wts, wts);
  put(ctx, p.skewFace, skewFace);
  put(ctx, p.skewGeo, skewGeo);
  put(ctx, p.slotClaim, slotClaim);
  put(ctx, p.claimFaces, claimFaces);
  put(ctx, p.claimGeo, claimGeo);
  p.nSkew = static_cast<int64_t>(skewFace.size());
  p.nClaims = static_cast<int64_t>(claimFaces.size() / 3);
  p.transfer = transfer;
  p.oblKinds.clear();
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  p.built = true;
}

// ObliqueNeumannFace (Gradient.cpp:190-217) of every boundary face whose
// condition does not prescribe the value (kind 0, FixedGradient)
void buildOblique(Context& ctx, FaceMesh& m, const std::vector<int32_t>& kinds) {
  FaceGradPlan& p = m.gp;
  if (p.oblKinds == kinds) return;
  const Geo G{m};
  std::vector<double> geo(8 * m.nf, 0.0);
  std::vector<int64_t> faces;
  for (int64_t f = 0; f < m.nf; ++f) {
    if (!G.boundary(f) || kinds[f] != 0) continue;
    const R3 d = G.fc(f) - G.cc(m.hOwner[f]);
    const R3 nrm = G.unitNormal(f);
    const double dn = dotR(d, nrm);
    if (!(dn > 0.0)) continue;
    const R3 tang = d - (nrm * dn);
    if (isZero(tang)) continue;
    faces.push_back(f);
    double* g = &geo[8 * f];
    g[0] = 1.0;
    g[1] = dn;
    g[2] = nrm.x, g[3] = nrm.y, g[4] = nrm.z;
    g[5] = tang.x, g[6] = tang.y, g[7] = tang.z;
  }
  put(ctx, p.oblGeo, geo);
  put(ctx, p.oblFaces, faces);
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));  // the host vectors end here
  p.nObl = static_cast<int64_t>(faces.size());
  p.oblKinds = kinds;
}

// --- device ------------------------------------------------------------------

// ScalarBoundaryCondition::boundaryValue: FixedValue (kind 1) its value,
// FixedGradient (kind 0) phi_P + g d (FixedValue.cpp:22-24, FixedGradient.cpp:22-27)
__device__ inline double boundaryValue(int kind, double value, double phiP, double dist) {
  return kind == 1 ? value : phiP + (value * dist);
}

// interpolate() (Interpolation.cpp:21-30, 106-113, 157-166)
__global__ void faceValueKernel(int64_t nf, const int64_t* __restrict__ owner,
                                const int64_t* __restrict__ neigh, const double* __restrict__ wts,
                                const double* __restrict__ phi, const int32_t* __restrict__ kind,
                                const double* __restrict__ value, double* fv) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double phiP = phi[owner[f]];
  const double dPf = wts[2 * f];
  if (neigh[f] < 0) {
    fv[f] = boundaryValue(kind[f], value[f], phiP, dPf);
    return;
  }
  const double dNf = wts[2 * f + 1];
  fv[f] = ((dNf * phiP) + (dPf * phi[neigh[f]])) / (dPf + dNf);
}

// interpolateInternalFaceSkewCorrected (Interpolation.cpp:43-62) with the
// last sweep's gradient, on the skewed faces
__global__ void skewFaceKernel(int64_t nSkew, const int64_t* __restrict__ faces,
                               const double* __restrict__ geo, const int64_t* __restrict__ owner,
                               const int64_t* __restrict__ neigh, const double* __restrict__ phi,
                               const double* __restrict__ grad, double* fv) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nSkew) return;
  const int64_t f = faces[k], o = owner[f], nb = neigh[f];
  const double t = geo[4 * k];
  const R3 skew{geo[4 * k + 1], geo[4 * k + 2], geo[4 * k + 3]};
  const double phiCross = phi[o] + (t * (phi[nb] - phi[o]));
  const R3 gCross = at3(grad, o) + ((at3(grad, nb) - at3(grad, o)) * t);
  fv[f] = phiCross + dotR(gCross, skew);
}

// the oblique faces re-evaluated with the last sweep's owner gradient
// (Gradient.cpp:577-583)
__global__ void obliqueFaceKernel(int64_t nObl, const int64_t* __restrict__ faces,
                                  const double* __restrict__ geo,
                                  const int64_t* __restrict__ owner, const double* __restrict__ phi,
                                  const int32_t* __restrict__ kind,
                                  const double* __restrict__ value,
                                  const double* __restrict__ grad, double* fv) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nObl) return;
  const int64_t f = faces[k], o = owner[f];
  const double* g = geo + 8 * f;
  const R3 tang{g[5], g[6], g[7]};
  fv[f] = boundaryValue(kind[f], value[f], phi[o], g[1]) + dotR(at3(grad, o), tang);
}

struct SweepArgs {
  int64_t n;
  const int64_t* cfOff;
  const int64_t* cf;
  const int64_t* owner;
  const double* sf;
  const double* vol;
  const int32_t* slotClaim;
  const int64_t* claimFaces;
  const double* claimGeo;
  const double* phi;
  const double* fv;
  const double* prev;  // the last sweep's gradient, or null on the first sweep
  double* out;
};

// greenGaussSweep (Gradient.cpp:478-532): per cell, the boundary-consistent
// value of every claimed opposite face (boundaryConsistentFaceValue,
// Gradient.cpp:104-170), then sum S_f phi_f over the cell's faces in order,
// divided by the volume
__global__ void sweepKernel(SweepArgs a) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  const double phiP = a.phi[c];
  const R3 og = a.prev != nullptr ? at3(a.prev, c) : R3{};
  R3 sum{};
  for (int64_t k = a.cfOff[c]; k < a.cfOff[c + 1]; ++k) {
    const int64_t f = a.cf[k];
    const R3 s = at3(a.sf, f);
    const R3 sfCell = a.owner[f] == c ? s : s * -1.0;
    double value = a.fv[f];
    const int32_t q = a.slotClaim[k];
    if (q >= 0) {
      const int64_t bf = a.claimFaces[3 * q], far = a.claimFaces[3 * q + 2];
      const double* g = a.claimGeo + 6 * q;
      const double farDist = g[0], w = g[1], back = g[2];
      const R3 offset{g[3], g[4], g[5]};
      const double phiB = a.fv[bf] + dotR(og, offset);
      const double second =
          2.0 * (((phiB - phiP) / back) + ((a.phi[far] - phiP) / farDist)) / (back + farDist);
      const double correction = -0.5 * w * (1.0 - w) * farDist * farDist * second;
      value = value + correction;
    }
    sum = sum + (sfCell * value);
  }
  const R3 r = sum * (1.0 / a.vol[c]);
  a.out[3 * c] = r.x;
  a.out[3 * c + 1] = r.y;
  a.out[3 * c + 2] = r.z;
}

struct LsqArgs {
  int64_t n;
  const int64_t* cfOff;
  const int64_t* cf;
  const int64_t* owner;
  const int64_t* neigh;
  const double* cc;
  const double* fc;
  const double* wts;
  const double* obl;
  const int32_t* kind;
  const double* value;
  const double* phi;
  const double* gg;  // the Green-Gauss fallback
  double* out;
};

// leastSquaresGradient (Gradient.cpp:361-447) with the 2D normal equations
// of solveLeastSquaresGradient (Gradient.cpp:297-359)
__global__ void lsqKernel(LsqArgs a) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= a.n) return;
  const double phiP = a.phi[c];
  const R3 cP = at3(a.cc, c);
  double Sxx = 0.0, Sxy = 0.0, Syy = 0.0, bx = 0.0, by = 0.0;
  for (int64_t k = a.cfOff[c]; k < a.cfOff[c + 1]; ++k) {
    const int64_t f = a.cf[k];
    R3 d;
    double dphi;
    if (a.neigh[f] >= 0) {
      const int64_t nb = a.owner[f] == c ? a.neigh[f] : a.owner[f];
      d = at3(a.cc, nb) - cP;
      dphi = a.phi[nb] - phiP;
    } else if (a.obl[8 * f] != 0.0) {
      const double* g = a.obl + 8 * f;
      d = R3{g[2], g[3], g[4]} * g[1];
      dphi = boundaryValue(a.kind[f], a.value[f], phiP, g[1]) - phiP;
    } else {
      d = at3(a.fc, f) - cP;
      dphi = boundaryValue(a.kind[f], a.value[f], phiP, a.wts[2 * f]) - phiP;
    }
    const double dx = d.x, dy = d.y;
    const double d2 = (dx * dx) + (dy * dy);
    if (!(d2 > 0.0)) continue;
    const double weight = 1.0 / d2;
    Sxx += weight * dx * dx;
    Sxy += weight * dx * dy;
    Syy += weight * dy * dy;
    bx += weight * dx * dphi;
    by += weight * dy * dphi;
  }
  const double det = (Sxx * Syy) - (Sxy * Sxy);
  const double scale = Sxx + Syy;
  double gx = 0.0, gy = 0.0;
  bool ok = scale > 0.0 && !(det < (1e-10 * scale * scale));
  if (ok) {
    gx = ((bx * Syy) - (by * Sxy)) / det;
    gy = ((Sxx * by) - (Sxy * bx)) / det;
    ok = isfinite(gx) && isfinite(gy);
  }
  if (ok) {
    a.out[3 * c] = gx;
    a.out[3 * c + 1] = gy;
    a.out[3 * c + 2] = 0.0;
  } else {
    a.out[3 * c] = a.gg[3 * c];
    a.out[3 * c + 1] = a.gg[3 * c + 1];
    a.out[3 * c + 2] = a.gg[3 * c + 2];
  }
}

constexpr int kSkewSweeps = 4;  // kGreenGaussSkewCorrectionSweeps (Gradient.hpp:80)

// greenGaussGradient (Gradient.cpp:534-596) into out
void greenGauss(Context& ctx, FaceMesh& m, const double* phi, const int32_t* kind,
                const double* value, double* out) {
  FaceGradPlan& p = m.gp;
  m.fv.ensure(m.nf);
  m.gradB.ensure(3 * m.n);
  hipLaunchKernelGGL(faceValueKernel, gBlocks(m.nf), dim3(kGBlock), 0, ctx.stream, m.nf,
                     m.owner.get(), m.neigh.get(), p.wts.get(), phi, kind, value, m.fv.get());
  SweepArgs a{m.n,         m.cfOff.get(),     m.cf.get(),       m.owner.get(), m.sf.get(),
              m.vol.get(), p.slotClaim.get(), p.claimFaces.get(), p.claimGeo.get(), phi,
              m.fv.get(),  nullptr,           out};
  const bool sweeps = p.nSkew > 0 || p.nObl > 0 || p.transfer;
  // with sweeps, the first one goes to the buffer that makes the last land in out
  double* bufs[2] = {out, m.gradB.get()};
  int cur = sweeps ? (kSkewSweeps & 1) : 0;
  a.out = bufs[cur];
  hipLaunchKernelGGL(sweepKernel, gBlocks(m.n), dim3(kGBlock), 0, ctx.stream, a);
  for (int s = 0; sweeps && s < kSkewSweeps; ++s) {
    const double* g = bufs[cur];
    if (p.nSkew > 0)
      hipLaunchKernelGGL(skewFaceKernel, gBlocks(p.nSkew), dim3(kGBlock), 0, ctx.stream, p.nSkew,
                         p.skewFace.get(), p.skewGeo.get(), m.owner.get(), m.neigh.get(), phi, g,
                         m.fv.get());
    if (p.nObl > 0)
      hipLaunchKernelGGL(obliqueFaceKernel, gBlocks(p.nObl), dim3(kGBlock), 0, ctx.stream,
                         p.nObl, p.oblFaces.get(), p.oblGeo.get(), m.owner.get(), phi, kind, value,
                         g, m.fv.get());
    a.prev = g;
    cur = 1 - cur;
    a.out = bufs[cur];
    hipLaunchKernelGGL(sweepKernel, gBlocks(m.n), dim3(kGBlock), 0, ctx.stream, a);
  }
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace

void fmeshSetVolumes(Context& ctx, const double* cellVolume) {
  require(ctx.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
  FaceMesh& m = *ctx.fmesh;
  require(cellVolume != nullptr, "face-list mesh: cell volumes are required");
  std::vector<double> v(cellVolume, cellVolume + m.n);
  for (double x : v)
    require(std::isfinite(x) && x > 0.0, "face-list mesh: cell volumes must be finite and > 0");
  put(ctx, m.vol, v);
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  m.hasVolumes = true;
}

void fmeshGradient(Context& ctx, const double* field, const int32_t* kind, const double* value,
                   int scheme, double* out) {
  require(ctx.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
  FaceMesh& m = *ctx.fmesh;
  require(m.hasVolumes, "face-list gradient: call cfdhip_fmesh_set_cell_volumes first");
  require(scheme == 0 || scheme == 1, "gradient: scheme must be 0 (Green-Gauss) or 1 (least "
                                      "squares)");
  buildGradPlan(ctx, m);
  std::vector<int32_t> kinds(m.nf);
  CFDHIP_CHECK(hipMemcpyAsync(kinds.data(), kind, m.nf * sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
  for (int64_t f = 0; f < m.nf; ++f)
    if (m.hNeigh[f] >= 0) kinds[f] = -1;  // internal faces carry no condition
  buildOblique(ctx, m, kinds);
  if (scheme == 0) {
    greenGauss(ctx, m, field, kind, value, out);
    return;
  }
  m.gg.ensure(3 * m.n);
  greenGauss(ctx, m, field, kind, value, m.gg.get());
  FaceGradPlan& p = m.gp;
  LsqArgs a{m.n,         m.cfOff.get(), m.cf.get(), m.owner.get(), m.neigh.get(), m.cc.get(),
            m.fc.get(),  p.wts.get(),   p.oblGeo.get(), kind,      value,          field,
            m.gg.get(),  out};
  hipLaunchKernelGGL(lsqKernel, gBlocks(m.n), dim3(kGBlock), 0, ctx.stream, a);
  CFDHIP_CHECK(hipGetLastError());
}

// The p' conditions of the explicit term's gradient (makeGradientBoundaries,
// PressureCorrectionEquation.cpp:43-54): FixedValue 0 where the pressure is
// FixedValue, FixedGradient 0 elsewhere.
void fmeshPprimeConditions(Context& ctx, FaceMesh& m) {
  std::vector<int32_t> kinds(m.hFixed.begin(), m.hFixed.end());
  std::vector<double> values(m.nf, 0.0);
  put(ctx, m.bcKind, kinds);
  put(ctx, m.bcVal, values);
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
}

void fmeshAssemblePass(Context& ctx, const double* previousDev, int scheme) {
  require(ctx.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
  FaceMesh& m = *ctx.fmesh;
  require(m.nonOrth != 0, "face-list p': the corrector passes need the non-orthogonal "
                          "correction (non_orthogonal = 1)");
  fmeshPprimeConditions(ctx, m);
  m.grad.ensure(3 * m.n);
  fmeshGradient(ctx, previousDev, m.bcKind.get(), m.bcVal.get(), scheme, m.grad.get());
  fmeshReassemble(ctx, m.grad.get());
}

}  // namespace cfdhip
