// Mesh plan: the structured Cartesian case in the reference's order, with the
// geometric quantities the kernels need evaluated once on the host in the
// reference's arithmetic (the analogue of the reference's device plan
// builders, e.g. cuda/kernels/DevicePressureCorrectionPlan.cpp:69-185).
#include <cmath>
#include <cstring>

#include "context.hpp"

namespace cfdhip {

namespace {

// Host 3-vector with the operations of include/cfd/core/Vector3.hpp.
struct G3 {
  double x = 0.0, y = 0.0, z = 0.0;
};
G3 operator+(G3 a, G3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
G3 operator-(G3 a, G3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
G3 operator*(G3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(G3 a, G3 b) {
  const double inPlane = (a.x * b.x) + (a.y * b.y);
  const double normal = a.z * b.z;
  return normal == 0.0 ? inPlane : inPlane + normal;
}
G3 cross(G3 a, G3 b) {
  return {(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
double mag(G3 v) { return std::sqrt(dot(v, v)); }
bool zero(G3 v) { return v.x == 0.0 && v.y == 0.0 && v.z == 0.0; }
bool wellPosed(G3 d, G3 s) {  // MeshGeometry.cpp:99-107
  const double dDotSf = dot(d, s);
  const double dm = mag(d), sm = mag(s);
  if (!(dm > 0.0) || !(sm > 0.0)) return false;
  return dDotSf > (1e-6 * dm * sm);
}
double comp(G3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
G3 unit(int a, double s) {
  G3 v{0.0, 0.0, 0.0};
  if (a == 0) v.x = s;
  if (a == 1) v.y = s;
  if (a == 2) v.z = s;
  return v;
}

// The structured mesh with explicit coordinates (AxisSpacing::uniform,
// src/mesh/MeshGrading.cpp:125-172).
struct Mesh {
  int dim;
  int64_t n[3];
  double delta[3];
  double areaPos[3];  // positive area of an x/y/z face
  double node(int a, int64_t i) const { return static_cast<double>(i) * delta[a]; }
  double center(int a, int64_t i) const { return (static_cast<double>(i) + 0.5) * delta[a]; }
  // 3D coordinates are origin(0) + value; 0.0 + v == v bitwise for v >= 0.
  G3 cell(int64_t i, int64_t j, int64_t k) const {
    return dim == 3 ? G3{0.0 + center(0, i), 0.0 + center(1, j), 0.0 + center(2, k)}
                    : G3{center(0, i), center(1, j), 0.0};
  }
  // Face on axis a with face index fi along a and cell indices elsewhere.
  G3 faceCentroid(int a, int64_t idx[3]) const {
    double c[3];
    for (int b = 0; b < 3; ++b) c[b] = b == a ? node(b, idx[b]) : center(b, idx[b]);
    if (dim == 2) return G3{c[0], c[1], 0.0};
    return G3{0.0 + c[0], 0.0 + c[1], 0.0 + c[2]};
  }
  // Stored area vector: owner -> neighbour (+axis) internally, outward on a
  // boundary; the 3D low side is the negated vector (-A, -0.0, -0.0), the 2D
  // one is built as (-h, 0.0) / (0.0, -w).
  G3 areaVector(int a, int64_t fi) const {
    const bool low = fi == 0;
    if (dim == 3) {
      G3 pos = unit(a, areaPos[a]);
      return low ? G3{-pos.x, -pos.y, -pos.z} : pos;
    }
    return unit(a, low ? -areaPos[a] : areaPos[a]);
  }
};

// Geometry of one cell face as the reference evaluates it.
struct FaceGeo {
  int axis;
  int64_t fi;       // face index along the axis
  bool boundary;
  G3 centroid, S;
  int64_t owner[3], neigh[3];
};

FaceGeo faceOf(const Mesh& m, int a, const int64_t cellIdx[3], bool high) {
  FaceGeo g;
  g.axis = a;
  int64_t idx[3] = {cellIdx[0], cellIdx[1], cellIdx[2]};
  g.fi = cellIdx[a] + (high ? 1 : 0);
  idx[a] = g.fi;
  g.boundary = g.fi == 0 || g.fi == m.n[a];
  g.centroid = m.faceCentroid(a, idx);
  g.S = m.areaVector(a, g.fi);
  for (int b = 0; b < 3; ++b) g.owner[b] = g.neigh[b] = cellIdx[b];
  if (g.boundary) {
    g.owner[a] = g.fi == 0 ? 0 : m.n[a] - 1;
  } else {
    g.owner[a] = g.fi - 1;
    g.neigh[a] = g.fi;
  }
  return g;
}

G3 cellCentroid(const Mesh& m, const int64_t idx[3]) { return m.cell(idx[0], idx[1], idx[2]); }
bool sameCell(const int64_t a[3], const int64_t b[3]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}
G3 unitNormal(const FaceGeo& f) { return f.S * (1.0 / mag(f.S)); }

// MeshGeometry::oppositeInteriorFace (MeshGeometry.cpp:47-96) over the cell's
// faces in creation order (W, E, S, N, B, T).
bool oppositeFace(const Mesh& m, const int64_t cell[3], const FaceGeo& ref, FaceGeo& out) {
  const G3 raw = unitNormal(ref);
  const G3 outRef = sameCell(ref.owner, cell) ? raw : raw * -1.0;
  bool found = false;
  double best = 0.0;
  for (int a = 0; a < m.dim; ++a)
    for (int side = 0; side < 2; ++side) {
      const FaceGeo cand = faceOf(m, a, cell, side == 1);
      if (cand.axis == ref.axis && cand.fi == ref.fi) continue;
      if (cand.boundary) continue;
      const G3 rn = unitNormal(cand);
      const G3 o = sameCell(cand.owner, cell) ? rn : rn * -1.0;
      const double al = dot(outRef, o);
      if (al < 0.0 && (!found || al < best)) {
        best = al;
        out = cand;
        found = true;
      }
    }
  return found;
}

bool samePatch(const PatchGeom& a, const PatchGeom& b) {
  return std::memcmp(&a, &b, sizeof(PatchGeom)) == 0;
}

// Per-face reference geometry of boundary face `f` (owner cell idx), written
// into a PatchGeom (conditions left zero).
PatchGeom boundaryGeometry(const Mesh& m, const FaceGeo& f, bool& transfer, bool& oblique,
                           bool& alignedCoupling) {
  PatchGeom g;
  std::memset(&g, 0, sizeof(g));
  const G3 cc = cellCentroid(m, f.owner);
  g.S[0] = f.S.x;
  g.S[1] = f.S.y;
  g.S[2] = f.S.z;
  g.area = mag(f.S);
  g.bdist = mag(f.centroid - cc);  // MeshGeometry::distance(owner, face)
  // p' boundary coupling takes the axis-aligned branch (PressureCorrectionEquation.cpp:79-87)
  {
    const G3 d = f.centroid - cc;
    const int nonZero = (f.S.x != 0.0) + (f.S.y != 0.0) + (f.S.z != 0.0);
    alignedCoupling = nonZero == 1 && zero(cross(d, f.S));
  }
  // obliqueNeumannFace geometry (Gradient.cpp:186-216)
  {
    const G3 d = f.centroid - cc;
    const G3 n = unitNormal(f);
    const double dn = dot(d, n);
    oblique = dn > 0.0 && !zero(d - (n * dn));
  }
  FaceGeo opp;
  if (!oppositeFace(m, f.owner, f, opp)) return g;
  const int64_t* far = sameCell(opp.owner, f.owner) ? opp.neigh : opp.owner;
  const G3 cf = cellCentroid(m, far);
  // boundaryInwardStencil (MeshGeometry.cpp:226-257)
  {
    const G3 n = unitNormal(f);
    const double h1 = dot(f.centroid - cc, n);
    const double h2 = dot(f.centroid - cf, n);
    if (h1 > 0.0 && h2 > h1) {
      g.stencil = 1;
      const G3 dP = cc - (f.centroid - (n * h1));
      const G3 dF = cf - (f.centroid - (n * h2));
      g.deltaP[0] = dP.x, g.deltaP[1] = dP.y, g.deltaP[2] = dP.z;
      g.deltaF[0] = dF.x, g.deltaF[1] = dF.y, g.deltaF[2] = dF.z;
      g.deltaZero = zero(dP) && zero(dF);
      // NonOrthogonalDiffusion.cpp:86-102 (geometry part)
      g.cP = h2 / (h1 * (h2 - h1));
      g.cF = h1 / (h2 * (h2 - h1));
      g.cB = (1.0 / h1) + (1.0 / h2);
    }
  }
  // boundaryConsistentFaceValue (Gradient.cpp:104-162), geometry part
  {
    const G3 d = cf - cc;
    const double farDistance = mag(d);
    if (!(farDistance > 0.0)) return g;
    const G3 inward = d * (1.0 / farDistance);
    // ownerNeighborCrossing of the opposite face (MeshGeometry.cpp:169-199)
    const G3 co = cellCentroid(m, opp.owner);
    const G3 dd = cellCentroid(m, opp.neigh) - co;
    if (!wellPosed(dd, opp.S)) return g;
    const G3 toFace = opp.centroid - co;
    double t;
    if (zero(cross(toFace, dd))) {
      t = dot(toFace, dd) / dot(dd, dd);
    } else {
      t = dot(opp.centroid - co, opp.S) / dot(dd, opp.S);
    }
    // boundaryLineIntersection (MeshGeometry.cpp:201-224)
    if (!wellPosed(inward * -1.0, f.S)) return g;
    const double dist = dot(cc - f.centroid, f.S) / dot(inward, f.S);
    if (!(dist > 0.0)) return g;
    const G3 point = cc - (inward * dist);
    const G3 off = point - f.centroid;
    if (!zero(off)) transfer = true;
    g.consistent = 1;
    g.back = dist;
    g.farDist = farDistance;
    g.w = t;
    g.offset[0] = off.x, g.offset[1] = off.y, g.offset[2] = off.z;
  }
  return g;
}

}  // namespace

void buildCartesianPlan(Context& ctx, int dim, int64_t nx, int64_t ny, int64_t nz, double lx,
                        double ly, double lz) {
  require(dim == 2 || dim == 3, "dim must be 2 or 3");
  if (dim == 2) {
    nz = 1;
    lz = 1.0;
  }
  require(nx > 0 && ny > 0 && nz > 0, "Cartesian mesh requires nx, ny, nz > 0");
  require(std::isfinite(lx) && lx > 0.0 && std::isfinite(ly) && ly > 0.0 &&
              std::isfinite(lz) && lz > 0.0,
          "Cartesian mesh requires finite, positive lengths");
  Mesh m;
  m.dim = dim;
  m.n[0] = nx;
  m.n[1] = ny;
  m.n[2] = nz;
  m.delta[0] = lx / static_cast<double>(nx);
  m.delta[1] = ly / static_cast<double>(ny);
  m.delta[2] = lz / static_cast<double>(nz);
  if (dim == 2) {
    // createRectilinear2D: vertical faces carry the cell height, horizontal the width.
    m.areaPos[0] = m.delta[1];
    m.areaPos[1] = m.delta[0];
    m.areaPos[2] = 0.0;
  } else {
    m.areaPos[0] = m.delta[1] * m.delta[2];  // createCartesian3D: areaX = dy*dz
    m.areaPos[1] = m.delta[0] * m.delta[2];
    m.areaPos[2] = m.delta[0] * m.delta[1];
  }

  View v{};
  v.dim = dim;
  v.nx = nx;
  v.ny = ny;
  v.nz = nz;
  v.n = nx * ny * nz;
  v.nfx = (nx + 1) * ny * nz;
  v.offY = v.nfx;
  v.offZ = v.offY + nx * (ny + 1) * nz;
  v.nf = v.offZ + (dim == 3 ? nx * ny * (nz + 1) : 0);
  v.ncell[0] = nx;
  v.ncell[1] = ny;
  v.ncell[2] = nz;
  v.stride[0] = 1;
  v.stride[1] = nx;
  v.stride[2] = nx * ny;
  v.vol = dim == 3 ? (m.delta[0] * m.delta[1]) * m.delta[2] : m.delta[0] * m.delta[1];
  v.rho = ctx.rho;

  // Per-axis tables from the reference expressions. Internal face fi of axis
  // a lies between cells fi-1 and fi; every other coordinate difference is
  // the exact zero (x - x), so the quantities depend on fi alone.
  for (int a = 0; a < 3; ++a) {
    const int64_t na = m.n[a];
    std::vector<double> t(5 * (na + 1), 0.0);
    if (a < dim) {
      for (int64_t fi = 0; fi <= na; ++fi) {
        const G3 fc = unit(a, m.node(a, fi));
        if (fi == 0 || fi == na) {
          const G3 cc = unit(a, m.center(a, fi == 0 ? 0 : na - 1));
          t[kDPf * (na + 1) + fi] = mag(fc - cc);
          continue;
        }
        const G3 co = unit(a, m.center(a, fi - 1)), cn = unit(a, m.center(a, fi));
        const double dPf = mag(fc - co);  // MeshGeometry::distance(owner, face)
        const double dNf = mag(fc - cn);
        t[kDPf * (na + 1) + fi] = dPf;
        t[kDNf * (na + 1) + fi] = dNf;
        t[kDSum * (na + 1) + fi] = dPf + dNf;
        const G3 d = cn - co;
        t[kDPN * (na + 1) + fi] = mag(d);
        t[kDC * (na + 1) + fi] = comp(d, a);
        // Internal faces are axis-aligned, parallel to d and unskewed
        // (MeshGeometry::ownerNeighborCrossing returns a zero skew vector).
        const G3 S = m.areaVector(a, fi);
        require(zero(cross(d, S)) && zero(cross(fc - co, d)),
                "internal face not parallel to the owner-neighbour line", CFDHIP_ERR_UNSUPPORTED);
      }
      const G3 S = m.areaVector(a, 1);
      v.area[a] = mag(S);
      v.S[a][0] = S.x;
      v.S[a][1] = S.y;
      v.S[a][2] = S.z;
    }
    ctx.tables[a].ensure(static_cast<int64_t>(t.size()));
    CFDHIP_CHECK(hipMemcpy(ctx.tables[a].get(), t.data(), t.size() * sizeof(double),
                           hipMemcpyHostToDevice));
    v.tab[a] = ctx.tables[a].get();
  }

  // Boundary patches: evaluate every face, keep one copy per patch, refuse a
  // patch whose faces differ in any bit.
  v.npatch = 2 * dim;
  bool transfer = false;
  for (int p = 0; p < v.npatch; ++p) {
    const int a = p / 2;
    const bool high = (p % 2) == 1;
    bool first = true;
    PatchGeom ref{};
    int64_t lim[3] = {m.n[0], m.n[1], m.n[2]};
    lim[a] = 1;
    for (int64_t k = 0; k < lim[2]; ++k)
      for (int64_t j = 0; j < lim[1]; ++j)
        for (int64_t i = 0; i < lim[0]; ++i) {
          int64_t cell[3] = {i, j, k};
          cell[a] = high ? m.n[a] - 1 : 0;
          const FaceGeo f = faceOf(m, a, cell, high);
          bool oblique = false, aligned = false;
          const PatchGeom g = boundaryGeometry(m, f, transfer, oblique, aligned);
          require(!oblique, "oblique boundary face", CFDHIP_ERR_UNSUPPORTED);
          require(aligned, "boundary face not aligned with its owner", CFDHIP_ERR_UNSUPPORTED);
          if (first) {
            ref = g;
            first = false;
          } else {
            require(samePatch(ref, g), "boundary patch geometry is not uniform",
                    CFDHIP_ERR_UNSUPPORTED);
          }
        }
    v.patch[p] = ref;
    v.patch[p].vkind = CFDHIP_WALL;
    v.patch[p].pkind = CFDHIP_FIXED_GRADIENT;
  }
  v.transfer = transfer ? 1 : 0;
  ctx.view = v;
  for (int a = 0; a < 3; ++a) ctx.delta[a] = m.delta[a];
  ctx.hasMesh = true;
  ctx.pressureAssembled = ctx.momentumAssembled = false;
  for (bool& b : ctx.bcSet) b = false;
  ctx.ensureFields();
}

void setBoundary(Context& ctx, int patch, int vkind, const double* v3, int pkind, double pvalue) {
  ctx.requireMesh();
  require(patch >= 0 && patch < ctx.view.npatch, "patch index out of range");
  require(vkind >= CFDHIP_WALL && vkind <= CFDHIP_OUTLET, "unknown velocity condition");
  require(pkind == CFDHIP_FIXED_GRADIENT || pkind == CFDHIP_FIXED_VALUE,
          "unknown pressure condition");
  require(std::isfinite(pvalue), "pressure condition value must be finite");
  PatchGeom& g = ctx.view.patch[patch];
  g.vkind = vkind;
  for (int c = 0; c < 3; ++c) {
    const double x = (v3 != nullptr && (vkind == CFDHIP_MOVING_WALL || vkind == CFDHIP_INLET))
                         ? v3[c]
                         : 0.0;
    require(std::isfinite(x), "velocity condition must be finite");
    g.vval[c] = x;
  }
  if (ctx.view.dim == 2) g.vval[2] = 0.0;  // Vector2 conditions carry z = 0
  g.pkind = pkind;
  g.pval = pvalue;
  ctx.bcSet[patch] = true;
}

__global__ void facesKernel(View v, int64_t* owner, int64_t* neighbor) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= v.nf) return;
  int64_t idx[3];
  const int a = decodeFace(v, f, idx);
  const int64_t fi = idx[a];
  if (fi == 0) {
    owner[f] = cellOf(v, idx);
    neighbor[f] = -1;
  } else if (fi == v.ncell[a]) {
    idx[a] = fi - 1;
    owner[f] = cellOf(v, idx);
    neighbor[f] = -1;
  } else {
    idx[a] = fi - 1;
    owner[f] = cellOf(v, idx);
    idx[a] = fi;
    neighbor[f] = cellOf(v, idx);
  }
}

void exportFaces(Context& ctx, int64_t* owner, int64_t* neighbor) {
  ctx.requireMesh();
  const int64_t nf = ctx.view.nf;
  DeviceArray<int64_t> o, nb;
  o.ensure(nf);
  nb.ensure(nf);
  hipLaunchKernelGGL(facesKernel, blocksFor(nf), dim3(kBlock), 0, ctx.stream, ctx.view, o.get(),
                     nb.get());
  CFDHIP_CHECK(hipGetLastError());
  CFDHIP_CHECK(hipMemcpyAsync(owner, o.get(), nf * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipMemcpyAsync(neighbor, nb.get(), nf * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx.stream));
  CFDHIP_CHECK(hipStreamSynchronize(ctx.stream));
}

}  // namespace cfdhip
