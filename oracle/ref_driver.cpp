// Golden-vector driver for the reference solver (checker only).
//
// Linked against the reference's own CPU sources (oracle/Makefile.ref builds
// them from where they lie); never shipped, never run on the GPU path. Each
// command runs one piece of the SIMPLE pressure-velocity path through the
// reference's public API and dumps every array as raw little-endian files
// (<name>.f64 / <name>.i64) into out=DIR, for tests/golden/make_golden.py to
// pack into fixtures.
//
//   mesh    dim nx ny [nz]                     -> owner/neighbour/geometry (a1)
//   pcorr   dim nx ny [nz] seed                -> p' system with d == 1 (a9) + CG (a12)
//   simple  dim nx ny [nz] re iters flux psolver -> SIMPLE::solve fields+histories (a15)
//   stages  dim nx ny [nz] re warm flux psolver  -> one hand-driven SIMPLE iteration,
//                                                 every stage dumped (a3-a14)
//   piso    nx ny re dt warm steps              -> `warm` PISO::solveTimeStep calls from rest,
//                                                 the next step's transient momentum systems
//                                                 (TransientMomentum.cpp) from that state, then
//                                                 `steps` more PISO steps: per-step status, CFL,
//                                                 continuity, and the fields after each step
//   cpcorr  nx ny seed outlet                   -> assembleCompressiblePressureCorrection on
//                                                 seeded flux, face density, d, absolute
//                                                 pressure and temperature (cavity, or the
//                                                 right patch FixedValue with outlet=1), and
//                                                 its CG + Jacobi solution
//   fpcorr  nx ny seed outlet amp [mesh=multiblock] -> p' system on a distorted quad mesh (or
//                                                 two such blocks joined) with the
//                                                 non-orthogonal correction and explicit term,
//                                                 its CG + Jacobi solution, the corrected flux
//   cgtime  nx ny seed pmax                     -> seconds of pmax reference CG+Jacobi
//   simpletime dim nx ny re iters flux psolver  -> seconds of SIMPLE::solve and its stage times
//                                                 iterations on the seeded p' system
//                                                 (CPU baseline; prints one JSON line)
//   bicgstab in=DIR abs rel max [jacobi=1]      -> the reference's BiCGSTAB::solve on the CSR
//                                                 system of DIR (rowptr.i64 col.i64 val.f64
//                                                 b.f64 x0.f64): solution, history, status,
//                                                 iterations, restarts, residuals
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "cfd/algebra/BiCGSTAB.hpp"
#include "cfd/algebra/CG.hpp"
#include "cfd/algebra/Preconditioner.hpp"
#include "cfd/boundary/FixedGradient.hpp"
#include "cfd/boundary/FixedValue.hpp"
#include "cfd/compressible/CompressiblePressureCorrection.hpp"
#include "cfd/compressible/ThermodynamicProperties.hpp"
#include "cfd/boundary/MovingWall.hpp"
#include "cfd/boundary/Wall.hpp"
#include "cfd/discretization/Gradient.hpp"
#include "cfd/mesh/MeshGeometry.hpp"
#include "cfd/mesh/MultiBlockSpec.hpp"
#include "cfd/physics/ContinuityEquation.hpp"
#include "cfd/physics/FluidProperties.hpp"
#include "cfd/physics/MassFlux.hpp"
#include "cfd/pressure_velocity/PressureCorrectionEquation.hpp"
#include "cfd/pressure_velocity/RelaxedMomentum.hpp"
#include "cfd/pressure_velocity/RhieChow.hpp"
#include "cfd/pressure_velocity/PISO.hpp"
#include "cfd/pressure_velocity/SIMPLE.hpp"
#include "cfd/pressure_velocity/TransientMomentum.hpp"
#include "cfd/solver/TransientSolver.hpp"

using namespace cfd;
using cfd::algebra::LinearSolverSettings;
using cfd::algebra::SolverResult;
using cfd::algebra::SparseMatrix;
using cfd::algebra::Vector;
using cfd::fields::ScalarField;
using cfd::fields::SurfaceField;
using cfd::fields::VectorField;
using cfd::mesh::Mesh;
using cfd::mesh::MeshGeometry;

namespace {

std::map<std::string, std::string> args;
std::string outDir;

std::string arg(const std::string& k, const std::string& def = "") {
  auto it = args.find(k);
  if (it == args.end()) {
    if (def.empty()) {
      std::fprintf(stderr, "missing argument %s\n", k.c_str());
      std::exit(2);
    }
    return def;
  }
  return it->second;
}
long iarg(const std::string& k, const std::string& d = "") { return std::stol(arg(k, d)); }
double farg(const std::string& k, const std::string& d = "") { return std::stod(arg(k, d)); }

void dumpRaw(const std::string& name, const void* data, size_t bytes) {
  const std::string path = outDir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::perror(path.c_str());
    std::exit(1);
  }
  if (bytes) std::fwrite(data, 1, bytes, f);
  std::fclose(f);
}
void dumpF(const std::string& name, const std::vector<double>& v) {
  dumpRaw(name + ".f64", v.data(), v.size() * sizeof(double));
}
void dumpI(const std::string& name, const std::vector<int64_t>& v) {
  dumpRaw(name + ".i64", v.data(), v.size() * sizeof(int64_t));
}
template <class F>
std::vector<double> toVec(const F& field) {
  std::vector<double> v(field.size());
  for (size_t i = 0; i < field.size(); ++i) v[i] = field[i];
  return v;
}
void dumpVecField(const std::string& name, const VectorField& f, bool threeD) {
  std::vector<double> x(f.size()), y(f.size()), z(f.size());
  for (size_t i = 0; i < f.size(); ++i) {
    x[i] = f[i].x;
    y[i] = f[i].y;
    z[i] = f[i].z;
  }
  dumpF(name + "_x", x);
  dumpF(name + "_y", y);
  if (threeD) dumpF(name + "_z", z);
}
void dumpMatrix(const std::string& name, const SparseMatrix& m) {
  std::vector<int64_t> rp(m.rows() + 1), col(m.nonZeros());
  std::vector<double> val(m.nonZeros());
  for (size_t i = 0; i <= m.rows(); ++i) rp[i] = static_cast<int64_t>(m.rowOffsetsData()[i]);
  for (size_t k = 0; k < m.nonZeros(); ++k) {
    col[k] = static_cast<int64_t>(m.columnIndicesData()[k]);
    val[k] = m.valuesData()[k];
  }
  dumpI(name + "_rowptr", rp);
  dumpI(name + "_col", col);
  dumpF(name + "_val", val);
}
void dumpResult(const std::string& name, const SolverResult& r) {
  dumpF(name + "_x", toVec(r.solution));
  dumpF(name + "_hist", r.residualHistory);
  dumpI(name + "_meta", {static_cast<int64_t>(r.status), static_cast<int64_t>(r.iterations),
                         static_cast<int64_t>(r.restarts)});
  dumpF(name + "_res", {r.initialResidual, r.finalResidual});
}

// splitmix64 -> uniform [0,1) with 53 bits: (x >> 11) * 2^-53.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

struct Case {
  int dim;
  Mesh mesh;
  boundary::BoundaryConditionSet vbc;
  boundary::BoundaryConditionSet pbc;
};

Case makeCavity() {
  const int dim = static_cast<int>(iarg("dim", "2"));
  const size_t nx = iarg("nx"), ny = iarg("ny");
  Mesh mesh = dim == 3 ? MeshGeometry::createCartesian3D(nx, ny, iarg("nz"), 1.0, 1.0, 1.0)
                       : MeshGeometry::createCartesian2D(nx, ny, 1.0, 1.0);
  boundary::BoundaryConditionSet vbc, pbc;
  const double lid = farg("lid", "1.0");
  for (const auto& patch : mesh.boundaryPatches()) {
    const std::string& name = patch.name();
    const bool isLid = (dim == 2) ? name == "top" : name == "ymax";
    if (isLid)
      vbc.set(mesh, name, std::make_unique<boundary::MovingWall>(Vector3{lid, 0.0, 0.0}));
    else
      vbc.set(mesh, name, std::make_unique<boundary::Wall>());
    pbc.set(mesh, name, std::make_unique<boundary::FixedGradient>(0.0));
  }
  return Case{dim, std::move(mesh), std::move(vbc), std::move(pbc)};
}

LinearSolverSettings momentumSettings() {
  LinearSolverSettings s;
  s.type = algebra::LinearSolverType::BiCGSTAB;
  s.absoluteTolerance = farg("mabs", "1e-10");
  s.relativeTolerance = farg("mrel", "1e-8");
  s.maxIterations = iarg("mmax", "500");
  return s;
}

LinearSolverSettings pressureSettings() {
  LinearSolverSettings s;
  const std::string kind = arg("psolver", "cg_jacobi");
  s.type = kind == "bicgstab" ? algebra::LinearSolverType::BiCGSTAB : algebra::LinearSolverType::CG;
  s.preconditioner = kind == "cg_jacobi" ? algebra::PreconditionerType::Jacobi
                                         : algebra::PreconditionerType::None;
  s.absoluteTolerance = farg("pabs", kind == "bicgstab" ? "1e-10" : "0");
  s.relativeTolerance = farg("prel", "1e-8");
  s.maxIterations = iarg("pmax", "20000");
  return s;
}

pressure_velocity::SIMPLESettings simpleSettings() {
  pressure_velocity::SIMPLESettings s;
  s.maxIterations = iarg("iters", "1");
  s.velocityRelaxation = farg("au", "0.7");
  s.pressureRelaxation = farg("ap", "0.3");
  s.velocityTolerance = farg("vtol", "1e-6");
  s.pressureTolerance = farg("ptol", "1e-6");
  s.continuityTolerance = farg("ctol", "1e-6");
  s.momentumSolver = momentumSettings();
  s.pressureSolver = pressureSettings();
  const std::string flux = arg("flux", "rhie_chow");
  s.faceFlux = pressure_velocity::parseFaceFluxScheme(flux);
  return s;
}

void dumpMesh(const Mesh& m) {
  std::vector<int64_t> owner, neighbor, cellFaces, cellFaceOffsets{0};
  std::vector<double> fcx, fcy, fcz, sx, sy, sz, ccx, ccy, ccz, vol;
  for (const auto& f : m.faces()) {
    owner.push_back(static_cast<int64_t>(f.owner()));
    neighbor.push_back(f.neighbor().has_value() ? static_cast<int64_t>(*f.neighbor()) : -1);
    fcx.push_back(f.centroid().x);
    fcy.push_back(f.centroid().y);
    fcz.push_back(f.centroid().z);
    sx.push_back(f.areaVector().x);
    sy.push_back(f.areaVector().y);
    sz.push_back(f.areaVector().z);
  }
  for (const auto& cell : m.cells()) {
    ccx.push_back(cell.centroid().x);
    ccy.push_back(cell.centroid().y);
    ccz.push_back(cell.centroid().z);
    vol.push_back(cell.volume());
    for (Index fid : cell.faceIds()) cellFaces.push_back(static_cast<int64_t>(fid));
    cellFaceOffsets.push_back(static_cast<int64_t>(cellFaces.size()));
  }
  std::vector<int64_t> patchFaces, patchOffsets{0};
  for (const auto& p : m.boundaryPatches()) {
    for (Index fid : p.faceIds()) patchFaces.push_back(static_cast<int64_t>(fid));
    patchOffsets.push_back(static_cast<int64_t>(patchFaces.size()));
  }
  dumpI("owner", owner);
  dumpI("neighbor", neighbor);
  dumpI("cell_faces", cellFaces);
  dumpI("cell_face_offsets", cellFaceOffsets);
  dumpI("patch_faces", patchFaces);
  dumpI("patch_offsets", patchOffsets);
  dumpF("face_cx", fcx);
  dumpF("face_cy", fcy);
  dumpF("face_cz", fcz);
  dumpF("face_sx", sx);
  dumpF("face_sy", sy);
  dumpF("face_sz", sz);
  dumpF("cell_cx", ccx);
  dumpF("cell_cy", ccy);
  dumpF("cell_cz", ccz);
  dumpF("cell_vol", vol);
}

void cmdMesh() {
  const Case c = makeCavity();
  dumpMesh(c.mesh);
}

// p' system on a distorted structured-quad mesh (createStructuredQuad2D, interior
// vertices moved by a seeded fraction amp of the cell size) with the
// non-orthogonal correction and the explicit term of a seeded previous p'
// (its Green-Gauss gradient under the p' boundary conditions dumped as the
// input the device takes); its CG + Jacobi solution and the corrected flux.
void cmdFpcorr() {
  const size_t nx = iarg("nx"), ny = iarg("ny");
  const double amp = farg("amp", "0.2"), lx = farg("lx", "1.0"), ly = farg("ly", "1.0");
  SplitMix64 rng{static_cast<uint64_t>(iarg("seed", "11"))};
  // (nx + 1) x (ny + 1) vertices over [x0, x0 + lx] x [0, ly], interior ones moved
  auto lattice = [&](double x0) {
    std::vector<Vector2> verts;
    for (size_t j = 0; j <= ny; ++j)
      for (size_t i = 0; i <= nx; ++i) {
        double x = x0 + lx * static_cast<double>(i) / static_cast<double>(nx);
        double y = ly * static_cast<double>(j) / static_cast<double>(ny);
        if (i > 0 && i < nx && j > 0 && j < ny) {
          x += amp * (lx / static_cast<double>(nx)) * (2.0 * rng.uniform() - 1.0);
          y += amp * (ly / static_cast<double>(ny)) * (2.0 * rng.uniform() - 1.0);
        }
        verts.push_back(Vector2{x, y});
      }
    return verts;
  };
  // mesh=multiblock: two such blocks side by side (the second over
  // [lx, 2 lx]) joined conformally along x = lx (createMultiBlock2D)
  const Mesh m = [&] {
    if (arg("mesh", "quad") != "multiblock") return MeshGeometry::createStructuredQuad2D(nx, ny, lattice(0.0));
    mesh::MultiBlockSpec spec;
    spec.blocks.push_back(mesh::BlockSpec{"west", nx, ny, lattice(0.0)});
    spec.blocks.push_back(mesh::BlockSpec{"east", nx, ny, lattice(lx)});
    using S = mesh::BlockSide;
    spec.interfaces.push_back(mesh::BlockInterfaceSpec{{0, S::Right}, {1, S::Left}, false});
    spec.patches.push_back(mesh::BoundaryPatchSpec{"left", {{0, S::Left}}});
    spec.patches.push_back(mesh::BoundaryPatchSpec{"right", {{1, S::Right}}});
    spec.patches.push_back(mesh::BoundaryPatchSpec{"bottom", {{0, S::Bottom}, {1, S::Bottom}}});
    spec.patches.push_back(mesh::BoundaryPatchSpec{"top", {{0, S::Top}, {1, S::Top}}});
    return MeshGeometry::createMultiBlock2D(spec);
  }();
  boundary::BoundaryConditionSet pbc, gbc;
  const bool outlet = iarg("outlet", "0") != 0;
  for (const auto& patch : m.boundaryPatches()) {
    const bool fv = outlet && patch.name() == "right";
    if (fv) {
      pbc.set(m, patch.name(), std::make_unique<boundary::FixedValue>(0.0));
      gbc.set(m, patch.name(), std::make_unique<boundary::FixedValue>(0.0));
    } else {
      pbc.set(m, patch.name(), std::make_unique<boundary::FixedGradient>(0.0));
      gbc.set(m, patch.name(), std::make_unique<boundary::FixedGradient>(0.0));
    }
  }
  const size_t n = m.numberOfCells(), nf = m.numberOfFaces();
  SurfaceField flux(nf, 0.0);
  for (size_t f = 0; f < nf; ++f) flux[f] = 2.0 * rng.uniform() - 1.0;
  ScalarField du(n), dv(n), prev(n);
  for (size_t k = 0; k < n; ++k) du[k] = 0.5 + rng.uniform();
  for (size_t k = 0; k < n; ++k) dv[k] = 0.5 + rng.uniform();
  for (size_t k = 0; k < n; ++k) prev[k] = 2.0 * rng.uniform() - 1.0;
  const VectorField grad = discretization::gradient(m, prev, gbc);
  pressure_velocity::PressureCorrectionOptions opts;
  opts.nonOrthogonal = iarg("nonorth", "1") != 0;
  if (iarg("explicit", "1") != 0) opts.previousPressureCorrection = &prev;
  const double rho = farg("rho", "1.0");
  const auto asmb = pressure_velocity::assemblePressureCorrection(
      m, flux, du, dv, rho, iarg("ref", "0"), pbc, opts);
  dumpMesh(m);
  std::vector<double> gx(n), gy(n);
  for (size_t k = 0; k < n; ++k) {
    gx[k] = grad[k].x;
    gy[k] = grad[k].y;
  }
  dumpF("flux", toVec(flux));
  dumpF("du", toVec(du));
  dumpF("dv", toVec(dv));
  dumpF("grad_x", gx);
  dumpF("grad_y", gy);
  dumpMatrix("A", asmb.system.matrix());
  dumpF("rhs", toVec(asmb.system.rhs()));
  dumpF("face_coeff", toVec(asmb.faceCoefficient));
  dumpF("explicit_flux", toVec(asmb.explicitFaceFlux));
  LinearSolverSettings s;
  s.type = algebra::LinearSolverType::CG;
  s.absoluteTolerance = 0.0;
  s.relativeTolerance = farg("prel", "1e-12");
  s.maxIterations = iarg("pmax", "20000");
  algebra::CG cg(s, std::make_shared<algebra::JacobiPreconditioner>());
  const auto r = cg.solve(asmb.system);
  dumpResult("cg", r);
  const ScalarField pp = [&] {
    ScalarField f(n);
    for (size_t k = 0; k < n; ++k) f[k] = r.solution[k];
    return f;
  }();
  const SurfaceField corrected = pressure_velocity::correctFaceMassFlux(
      m, flux, asmb.faceCoefficient, pp, opts.previousPressureCorrection ? &asmb.explicitFaceFlux
                                                                         : nullptr);
  dumpF("flux_corrected", toVec(corrected));
  // the gradient's input and the other gradient forms of it: least squares
  // under the same (p') conditions, and both schemes under conditions with
  // non-zero values (FixedValue 0.3 on "right", FixedGradient -0.7 elsewhere)
  dumpF("prev", toVec(prev));
  auto dumpGrad = [&](const std::string& name, const VectorField& g) {
    std::vector<double> x(n), y(n);
    for (size_t k = 0; k < n; ++k) {
      x[k] = g[k].x;
      y[k] = g[k].y;
    }
    dumpF(name + "_x", x);
    dumpF(name + "_y", y);
  };
  using discretization::GradientScheme;
  dumpGrad("grad_ls", discretization::gradient(m, prev, gbc, GradientScheme::LeastSquares));
  boundary::BoundaryConditionSet vbc;
  for (const auto& patch : m.boundaryPatches()) {
    if (patch.name() == "right")
      vbc.set(m, patch.name(), std::make_unique<boundary::FixedValue>(0.3));
    else
      vbc.set(m, patch.name(), std::make_unique<boundary::FixedGradient>(-0.7));
  }
  dumpGrad("grad_v", discretization::gradient(m, prev, vbc, GradientScheme::GreenGauss));
  dumpGrad("grad_vls", discretization::gradient(m, prev, vbc, GradientScheme::LeastSquares));
  // the corrector passes of SIMPLE.cpp:876-924 on this system: pass 1 without
  // the explicit term, passes 2.. with grad(p' of the pass before) (scheme
  // `gscheme`), each solved by CG + Jacobi from zero; per pass its RHS,
  // explicit flux and solution, then the flux corrected with the last one
  const int passes = iarg("passes", "3");
  const GradientScheme gs =
      iarg("gscheme", "0") != 0 ? GradientScheme::LeastSquares : GradientScheme::GreenGauss;
  pressure_velocity::PressureCorrectionOptions popts;
  popts.nonOrthogonal = true;
  popts.gradientScheme = gs;
  ScalarField pPrev(n, 0.0);
  std::optional<pressure_velocity::PressureCorrectionAssembly> last;
  for (int k = 1; k <= passes; ++k) {
    popts.previousPressureCorrection = k > 1 ? &pPrev : nullptr;
    last = pressure_velocity::assemblePressureCorrection(m, flux, du, dv, rho, iarg("ref", "0"),
                                                         pbc, popts);
    const auto pr = cg.solve(last->system);
    dumpF("pass" + std::to_string(k) + "_rhs", toVec(last->system.rhs()));
    dumpF("pass" + std::to_string(k) + "_explicit", toVec(last->explicitFaceFlux));
    dumpResult("pass" + std::to_string(k), pr);
    for (size_t q = 0; q < n; ++q) pPrev[q] = pr.solution[q];
  }
  dumpF("passes_flux", toVec(pressure_velocity::correctFaceMassFlux(
                           m, flux, last->faceCoefficient, pPrev,
                           passes > 1 ? &last->explicitFaceFlux : nullptr)));
}

// p' operator with d == 1 (rho = 1, cell 0 pinned) and a seeded, mean-free RHS,
// solved by CG + Jacobi; the face flux passed in is a seeded field too so the
// assembled RHS (-continuity imbalance) is non-trivial.
void cmdPcorr() {
  const Case c = makeCavity();
  const Mesh& m = c.mesh;
  const size_t n = m.numberOfCells(), nf = m.numberOfFaces();
  SplitMix64 rng{static_cast<uint64_t>(iarg("seed", "20480001"))};
  SurfaceField flux(nf, 0.0);
  for (size_t f = 0; f < nf; ++f) flux[f] = 2.0 * rng.uniform() - 1.0;
  ScalarField ones(n, 1.0);
  pressure_velocity::PressureCorrectionOptions opts;
  const auto asmb = pressure_velocity::assemblePressureCorrection(
      m, flux, ones, ones, 1.0, 0, c.pbc, opts, c.dim == 3 ? &ones : nullptr);
  dumpF("flux", toVec(flux));
  dumpMatrix("A", asmb.system.matrix());
  dumpF("rhs", toVec(asmb.system.rhs()));
  dumpF("face_coeff", toVec(asmb.faceCoefficient));

  // Standalone solve: b_i = 2u_i - 1 from the seed, serial mean removed, b_0 = 0.
  Vector b(n);
  for (size_t i = 0; i < n; ++i) b[i] = 2.0 * rng.uniform() - 1.0;
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) sum += b[i];
  const double mean = sum / static_cast<double>(n);
  for (size_t i = 0; i < n; ++i) b[i] -= mean;
  b[0] = 0.0;
  dumpF("b", toVec(b));
  LinearSolverSettings s;
  s.type = algebra::LinearSolverType::CG;
  s.absoluteTolerance = 0.0;
  s.relativeTolerance = farg("prel", "1e-8");
  s.maxIterations = iarg("pmax", "20000");
  algebra::CG cg(s, std::make_shared<algebra::JacobiPreconditioner>());
  const algebra::LinearSystem sys(asmb.system.matrix(), b);
  const auto r = cg.solve(sys);
  dumpResult("cg", r);
  const Vector Ax = asmb.system.matrix().multiply(r.solution);
  dumpF("Ax", toVec(Ax));
}

void cmdCpcorr() {
  Case c = makeCavity();
  const Mesh& m = c.mesh;
  if (iarg("outlet", "0") != 0)
    c.pbc.replace(m, "right", std::make_unique<boundary::FixedValue>(0.0));
  const size_t n = m.numberOfCells(), nf = m.numberOfFaces();
  SplitMix64 rng{static_cast<uint64_t>(iarg("seed", "7"))};
  SurfaceField flux(nf, 0.0), rhoF(nf, 0.0);
  for (size_t f = 0; f < nf; ++f) flux[f] = 2.0 * rng.uniform() - 1.0;
  for (size_t f = 0; f < nf; ++f) rhoF[f] = 0.5 + rng.uniform();
  ScalarField d(n), pAbs(n), T(n);
  for (size_t k = 0; k < n; ++k) d[k] = 1e-3 * (0.5 + rng.uniform());
  for (size_t k = 0; k < n; ++k) pAbs[k] = 1e5 * (1.0 + 0.1 * (2.0 * rng.uniform() - 1.0));
  for (size_t k = 0; k < n; ++k) T[k] = 300.0 + 50.0 * (2.0 * rng.uniform() - 1.0);
  const double R = farg("gas", "287.05"), dtau = farg("dtau", "1e-3");
  const compressible::ThermodynamicProperties thermo(R, farg("cp", "1004.5"));
  pressure_velocity::PressureCorrectionOptions opts;
  const auto asmb = compressible::assembleCompressiblePressureCorrection(
      m, flux, rhoF, d, d, pAbs, T, thermo, dtau, iarg("ref", "0"), c.pbc, opts);
  dumpF("flux", toVec(flux));
  dumpF("face_density", toVec(rhoF));
  dumpF("d", toVec(d));
  dumpF("p_abs", toVec(pAbs));
  dumpF("T", toVec(T));
  dumpMatrix("A", asmb.system.matrix());
  dumpF("rhs", toVec(asmb.system.rhs()));
  dumpF("face_coeff", toVec(asmb.faceCoefficient));
  LinearSolverSettings s;
  s.type = algebra::LinearSolverType::CG;
  s.absoluteTolerance = 0.0;
  s.relativeTolerance = farg("prel", "1e-12");
  s.maxIterations = iarg("pmax", "20000");
  algebra::CG cg(s, std::make_shared<algebra::JacobiPreconditioner>());
  const auto r = cg.solve(asmb.system);
  dumpResult("cg", r);
}

void cmdSimple() {
  const Case c = makeCavity();
  const double re = farg("re", "100");
  const physics::FluidProperties fluid(1.0, farg("lid", "1.0") / re);
  pressure_velocity::SIMPLE simple(simpleSettings(), 0);
  const size_t n = c.mesh.numberOfCells();
  const auto res = simple.solve(c.mesh, fluid, c.vbc, c.pbc, VectorField(n, Vector3{}),
                                ScalarField(n, 0.0));
  dumpVecField("u", res.velocity, c.dim == 3);
  dumpF("p", toVec(res.pressure));
  dumpF("F", toVec(res.massFlux));
  dumpF("hist_u", res.uResidualHistory);
  dumpF("hist_v", res.vResidualHistory);
  dumpF("hist_w", res.wResidualHistory);
  dumpF("hist_p", res.pressureResidualHistory);
  dumpF("hist_c", res.continuityHistory);
  dumpI("meta", {static_cast<int64_t>(res.status), static_cast<int64_t>(res.iterations),
                 static_cast<int64_t>(res.momentumLinearIterations),
                 static_cast<int64_t>(res.pressureLinearIterations)});
  dumpF("scalars", {res.finalContinuityResidual, res.globalMassImbalance});
}

// PISO time steps on the 2D cavity (src/pressure_velocity/PISO.cpp:126-350), both
// linear solvers BiCGSTAB as PISO constructs them (pressure: psolver=bicgstab).
void cmdPiso() {
  const Case c = makeCavity();
  const Mesh& m = c.mesh;
  const size_t n = m.numberOfCells();
  const double re = farg("re", "100");
  const physics::FluidProperties fluid(1.0, farg("lid", "1.0") / re);
  const double dt = farg("dt");
  const pressure_velocity::PISOSettings settings{momentumSettings(), pressureSettings()};
  const pressure_velocity::PISO piso(m, fluid, c.vbc, c.pbc, settings, 0);
  solver::TransientState state{VectorField(n, Vector3{}), ScalarField(n, 0.0), SurfaceField()};
  state.massFlux = physics::calculateMassFlux(m, state.velocity, fluid, c.vbc);
  const long warm = iarg("warm", "0"), steps = iarg("steps", "1");
  for (long s = 0; s < warm; ++s) {
    const auto r = piso.solveTimeStep(state, dt);
    if (r.status != solver::TransientStepStatus::Converged) {
      std::fprintf(stderr, "warm step %ld failed\n", s);
      std::exit(1);
    }
    state = r.state;
  }
  dumpVecField("in_u", state.velocity, false);
  dumpF("in_p", toVec(state.pressure));
  dumpF("in_F", toVec(state.massFlux));
  const ScalarField mu(n, fluid.dynamicViscosity());
  const char* names[2] = {"mu", "mv"};
  for (int k = 0; k < 2; ++k) {
    ScalarField prev(n);
    for (size_t i = 0; i < n; ++i) prev[i] = k == 0 ? state.velocity[i].x : state.velocity[i].y;
    const auto a = pressure_velocity::assembleTransientMomentumComponent(
        m, state.velocity, state.pressure, state.massFlux, fluid, mu, c.vbc, c.pbc,
        k == 0 ? physics::VelocityComponent::U : physics::VelocityComponent::V, prev, dt);
    dumpMatrix(names[k], a.system.matrix());
    dumpF(std::string(names[k]) + "_rhs", toVec(a.system.rhs()));
    dumpF(std::string(names[k]) + "_diag", toVec(a.diagonal));
  }
  std::vector<double> scal;
  std::vector<int64_t> status;
  for (long s = 0; s < steps; ++s) {
    const auto r = piso.solveTimeStep(state, dt);
    status.push_back(static_cast<int64_t>(r.status));
    scal.push_back(r.maxCFL);
    scal.push_back(r.continuityResidual);
    scal.push_back(r.massImbalance);
    if (r.status != solver::TransientStepStatus::Converged) break;
    state = r.state;
    const std::string tag = "s" + std::to_string(s);
    dumpVecField(tag + "_u", state.velocity, false);
    dumpF(tag + "_p", toVec(state.pressure));
    dumpF(tag + "_F", toVec(state.massFlux));
  }
  dumpI("status", status);
  dumpF("scalars", scal);
}

// One outer iteration driven by hand through the same calls SIMPLE::solve makes
// on its CPU path, starting from the state after `warm` SIMPLE iterations.
void cmdStages() {
  const Case c = makeCavity();
  const Mesh& m = c.mesh;
  const bool threeD = c.dim == 3;
  const size_t n = m.numberOfCells();
  const double re = farg("re", "100");
  const physics::FluidProperties fluid(1.0, farg("lid", "1.0") / re);
  const auto settings = simpleSettings();
  const double au = settings.velocityRelaxation, ap = settings.pressureRelaxation;

  VectorField velocity(n, Vector3{});
  ScalarField pressure(n, 0.0);
  SurfaceField massFlux;
  const long warm = iarg("warm", "0");
  if (warm > 0) {
    auto s = settings;
    s.maxIterations = static_cast<Index>(warm);
    pressure_velocity::SIMPLE simple(s, 0);
    const auto res = simple.solve(m, fluid, c.vbc, c.pbc, velocity, pressure);
    velocity = res.velocity;
    pressure = res.pressure;
    massFlux = res.massFlux;
  } else {
    massFlux = physics::calculateMassFlux(m, velocity, fluid, c.vbc);
  }
  dumpVecField("in_u", velocity, threeD);
  dumpF("in_p", toVec(pressure));
  dumpF("in_F", toVec(massFlux));

  const ScalarField mu(n, fluid.dynamicViscosity());
  const auto comp = [&](const VectorField& v, int k) {
    ScalarField s(n);
    for (size_t i = 0; i < n; ++i) s[i] = k == 0 ? v[i].x : (k == 1 ? v[i].y : v[i].z);
    return s;
  };
  const int ncomp = threeD ? 3 : 2;
  const char* names[3] = {"mu", "mv", "mw"};
  std::vector<Vector> solutions;
  std::vector<ScalarField> response;
  algebra::BiCGSTAB momentumSolver(settings.momentumSolver);
  for (int k = 0; k < ncomp; ++k) {
    const ScalarField prev = comp(velocity, k);
    const auto asmb = pressure_velocity::assembleRelaxedMomentumComponent(
        m, velocity, pressure, massFlux, mu, c.vbc, c.pbc,
        static_cast<physics::VelocityComponent>(k), prev, au);
    dumpMatrix(names[k], asmb.system.matrix());
    dumpF(std::string(names[k]) + "_rhs", toVec(asmb.system.rhs()));
    dumpF(std::string(names[k]) + "_diag", toVec(asmb.diagonal));
    Vector x0(n);
    for (size_t i = 0; i < n; ++i) x0[i] = prev[i];
    const auto r = momentumSolver.solve(asmb.system, x0);
    dumpResult(std::string(names[k]) + "_solve", r);
    solutions.push_back(r.solution);
    response.push_back(pressure_velocity::computeMomentumResponseCoefficient(m, asmb.diagonal));
    dumpF(std::string("d") + "uvw"[k], toVec(response.back()));
  }
  VectorField vstar(n);
  for (size_t i = 0; i < n; ++i)
    vstar[i] = Vector3{solutions[0][i], solutions[1][i], threeD ? solutions[2][i] : 0.0};

  const auto faceFlux = pressure_velocity::resolveFaceFluxScheme(settings.faceFlux, c.dim);
  const ScalarField* dW = threeD ? &response[2] : nullptr;
  SurfaceField predictor;
  const VectorField gradP =
      discretization::gradient(m, pressure, c.pbc, discretization::GradientScheme::GreenGauss);
  dumpVecField("gradp", gradP, threeD);
  if (faceFlux == pressure_velocity::FaceFluxScheme::RhieChow) {
    predictor = pressure_velocity::rhieChowMassFlux(m, vstar, pressure, gradP, response[0],
                                                    response[1], dW, fluid, c.vbc, au);
  } else {
    predictor = physics::calculateMassFlux(m, vstar, fluid, c.vbc);
  }
  dumpF("Fstar", toVec(predictor));
  dumpF("Flinear", toVec(physics::calculateMassFlux(m, vstar, fluid, c.vbc)));

  pressure_velocity::PressureCorrectionOptions opts;
  const auto pasm = pressure_velocity::assemblePressureCorrection(
      m, predictor, response[0], response[1], fluid.density(), 0, c.pbc, opts, dW);
  dumpMatrix("pc", pasm.system.matrix());
  dumpF("pc_rhs", toVec(pasm.system.rhs()));
  dumpF("pc_face_coeff", toVec(pasm.faceCoefficient));

  const auto pset = settings.pressureSolver;
  std::unique_ptr<algebra::LinearSolver> psolver;
  if (pset.type == algebra::LinearSolverType::CG)
    psolver = std::make_unique<algebra::CG>(
        pset, pset.preconditioner == algebra::PreconditionerType::Jacobi
                  ? std::make_shared<algebra::JacobiPreconditioner>()
                  : nullptr);
  else
    psolver = std::make_unique<algebra::BiCGSTAB>(pset);
  const auto pr = psolver->solve(pasm.system);
  dumpResult("pc_solve", pr);
  ScalarField pprime(n);
  for (size_t i = 0; i < n; ++i) pprime[i] = pr.solution[i];

  ScalarField pnew(n);
  for (size_t i = 0; i < n; ++i) pnew[i] = pressure[i] + (ap * pprime[i]);
  dumpF("p_new", toVec(pnew));
  const VectorField gradPP = discretization::gradient(
      m, pprime,
      [&] {
        boundary::BoundaryConditionSet b;
        for (const auto& patch : m.boundaryPatches())
          b.set(m, patch.name(), std::make_unique<boundary::FixedGradient>(0.0));
        return b;
      }(),
      discretization::GradientScheme::GreenGauss);
  dumpVecField("gradpp", gradPP, threeD);
  const VectorField unew = pressure_velocity::correctVelocity(
      m, vstar, response[0], response[1], pprime, c.pbc,
      discretization::GradientScheme::GreenGauss, dW);
  dumpVecField("u_new", unew, threeD);
  const SurfaceField fnew =
      pressure_velocity::correctFaceMassFlux(m, predictor, pasm.faceCoefficient, pprime, nullptr);
  dumpF("F_new", toVec(fnew));
  const auto cont = physics::evaluateContinuity(m, fnew);
  dumpF("imbalance", toVec(cont.cellImbalance));
  double ss = 0.0;
  for (size_t i = 0; i < n; ++i) ss += cont.cellImbalance[i] * cont.cellImbalance[i];
  dumpF("cont_scalars", {std::sqrt(ss / static_cast<double>(n)), cont.globalNetFlux});
}

// CPU baseline: the reference's CG + Jacobi on the stand-alone p' system
// (d == 1, cell 0 pinned, seeded mean-free RHS), bounded to pmax iterations.
std::vector<char> readRaw(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    std::perror(path.c_str());
    std::exit(1);
  }
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<char> buf(static_cast<size_t>(size));
  if (size > 0 && std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
    std::perror(path.c_str());
    std::exit(1);
  }
  std::fclose(f);
  return buf;
}
template <class T>
std::vector<T> readArray(const std::string& dir, const std::string& name) {
  const std::vector<char> raw = readRaw(dir + "/" + name);
  std::vector<T> v(raw.size() / sizeof(T));
  if (!v.empty()) std::memcpy(v.data(), raw.data(), v.size() * sizeof(T));
  return v;
}

// BiCGSTAB::solve (src/algebra/BiCGSTAB.cpp:48-275) on a stored CSR system.
void cmdBicgstab() {
  const std::string in = arg("in");
  const auto rp = readArray<int64_t>(in, "rowptr.i64");
  const auto col = readArray<int64_t>(in, "col.i64");
  auto val = readArray<double>(in, "val.f64");
  const auto b = readArray<double>(in, "b.f64");
  const auto x0 = readArray<double>(in, "x0.f64");
  const size_t n = rp.size() - 1;
  std::vector<Index> offsets(rp.begin(), rp.end()), columns(col.begin(), col.end());
  Vector rhs(n), guess(n);
  for (size_t i = 0; i < n; ++i) {
    rhs[i] = b[i];
    guess[i] = x0[i];
  }
  const algebra::LinearSystem sys(SparseMatrix(n, n, std::move(val), std::move(columns),
                                               std::move(offsets)),
                                  rhs);
  LinearSolverSettings s;
  s.type = algebra::LinearSolverType::BiCGSTAB;
  s.absoluteTolerance = farg("abs");
  s.relativeTolerance = farg("rel");
  s.maxIterations = iarg("max");
  std::shared_ptr<algebra::Preconditioner> pre;
  if (iarg("jacobi", "0")) pre = std::make_shared<algebra::JacobiPreconditioner>();
  algebra::BiCGSTAB solver(s, pre);
  const SolverResult r = solver.solve(sys, guess);
  dumpResult("bcg", r);
}

void cmdCgTime() {
  const Case c = makeCavity();
  const Mesh& m = c.mesh;
  const size_t n = m.numberOfCells();
  ScalarField ones(n, 1.0);
  const SurfaceField zeroFlux(m.numberOfFaces(), 0.0);
  pressure_velocity::PressureCorrectionOptions opts;
  const auto asmb = pressure_velocity::assemblePressureCorrection(m, zeroFlux, ones, ones, 1.0, 0,
                                                                  c.pbc, opts, nullptr);
  SplitMix64 rng{static_cast<uint64_t>(iarg("seed", "20480001"))};
  Vector b(n);
  for (size_t i = 0; i < n; ++i) b[i] = 2.0 * rng.uniform() - 1.0;
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) sum += b[i];
  const double mean = sum / static_cast<double>(n);
  for (size_t i = 0; i < n; ++i) b[i] -= mean;
  b[0] = 0.0;
  LinearSolverSettings s;
  s.type = algebra::LinearSolverType::CG;
  s.absoluteTolerance = 0.0;
  s.relativeTolerance = 1e-8;
  s.maxIterations = iarg("pmax", "200");
  algebra::CG cg(s, std::make_shared<algebra::JacobiPreconditioner>());
  const algebra::LinearSystem sys(asmb.system.matrix(), b);
  const auto t0 = std::chrono::steady_clock::now();
  const auto r = cg.solve(sys);
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("{\"seconds\": %.9g, \"iterations\": %lld, \"status\": %d, \"final\": %.17g}\n",
              sec, static_cast<long long>(r.iterations), static_cast<int>(r.status),
              r.finalResidual);
}

// Wall time of SIMPLE::solve on the cavity (the CPU path, OpenMP SpMV), with
// the reference's own per-stage seconds (SIMPLEResult::stageSeconds).
void cmdSimpleTime() {
  const Case c = makeCavity();
  const double re = farg("re", "100");
  const physics::FluidProperties fluid(1.0, farg("lid", "1.0") / re);
  pressure_velocity::SIMPLE simple(simpleSettings(), 0);
  const size_t n = c.mesh.numberOfCells();
  const auto t0 = std::chrono::steady_clock::now();
  const auto res = simple.solve(c.mesh, fluid, c.vbc, c.pbc, VectorField(n, Vector3{}),
                                ScalarField(n, 0.0));
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const auto& st = res.stageSeconds;
  std::printf("{\"seconds\": %.9g, \"status\": %d, \"iterations\": %lld, "
              "\"momentum_iterations\": %lld, \"pressure_iterations\": %lld, "
              "\"setup\": %.9g, \"momentum_solve\": %.9g, \"pressure_solve\": %.9g, "
              "\"discretization\": %.9g, \"bookkeeping\": %.9g}\n",
              sec, static_cast<int>(res.status), static_cast<long long>(res.iterations),
              static_cast<long long>(res.momentumLinearIterations),
              static_cast<long long>(res.pressureLinearIterations), st.setup, st.momentumSolve,
              st.pressureSolve, st.discretization(), st.bookkeeping);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: ref_driver cmd key=value ... out=DIR\n");
    return 2;
  }
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    const auto eq = a.find('=');
    if (eq == std::string::npos) continue;
    args[a.substr(0, eq)] = a.substr(eq + 1);
  }
  outDir = arg("out", ".");
  const std::string cmd = argv[1];
  if (cmd == "mesh") cmdMesh();
  else if (cmd == "pcorr") cmdPcorr();
  else if (cmd == "simple") cmdSimple();
  else if (cmd == "stages") cmdStages();
  else if (cmd == "cgtime") cmdCgTime();
  else if (cmd == "simpletime") cmdSimpleTime();
  else if (cmd == "piso") cmdPiso();
  else if (cmd == "cpcorr") cmdCpcorr();
  else if (cmd == "fpcorr") cmdFpcorr();
  else if (cmd == "bicgstab") cmdBicgstab();
  else {
    std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
    return 2;
  }
  return 0;
}
