// cfd::gpu::GpuSimpleDiscretization (include/cfd/gpu/GpuSimpleDiscretization.hpp:56-262)
// bound to libcfdhip, compiled against the reference's own headers and linked
// in place of its CPU stub src/gpu/GpuSimpleDiscretization.cpp. A pimpl over
// the C++ adapter cfdhip::GpuSimpleDiscretization (include/cfdhip.hpp): this
// file only converts the reference's mesh, boundary-condition sets and fields
// into the adapter's plain shapes and back.
//
// prepare() accepts exactly the meshes the structured kernels reproduce: one
// structured block built as MeshGeometry::createCartesian2D/3D builds it. The
// device regenerates the mesh from (nx, ny, nz, lx, ly, lz) and prepare()
// compares every cell centroid and volume and every face's owner and
// neighbour with the caller's mesh, bit for bit; any difference, or a
// boundary condition outside {Wall, MovingWall, Inlet, Outlet} x {FixedValue,
// FixedGradient}, declines with the reason and SIMPLE runs its CPU path.
#include <cstring>
#include <memory>
#include <span>
#include <string>
#include <vector>

#include "cfd/boundary/BoundaryCondition.hpp"
#include "cfd/boundary/FixedGradient.hpp"
#include "cfd/boundary/FixedValue.hpp"
#include "cfd/boundary/Inlet.hpp"
#include "cfd/boundary/MovingWall.hpp"
#include "cfd/core/Exception.hpp"
#include "cfd/gpu/GPUBackend.hpp"
#include "cfd/gpu/GpuSimpleDiscretization.hpp"
#include "cfd/mesh/Mesh.hpp"
#include "cfdhip.hpp"

namespace cfd::gpu {

namespace {

using cfd::algebra::LinearSystem;
using cfd::algebra::SolverResult;
using cfd::algebra::SparseMatrix;
using cfd::algebra::Vector;
using cfd::boundary::BoundaryConditionType;
using cfd::fields::ScalarField;
using cfd::fields::SurfaceField;
using cfd::fields::VectorField;

template <class F>
auto translated(F&& fn) -> decltype(fn()) {
  try {
    return fn();
  } catch (const cfdhip::InvalidArgumentError& e) {
    throw cfd::InvalidArgumentError(e.what());
  } catch (const cfdhip::NumericalError& e) {
    throw cfd::NumericalError(e.what());
  }
}

bool sameBits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// The patch names of createCartesian2D / createCartesian3D, in the device's
// patch order (src/mesh/MeshGeometry.cpp:503-506, 1142).
const char* const kPatches2[4] = {"left", "right", "bottom", "top"};
const char* const kPatches3[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};

bool mapCondition(const cfd::boundary::BoundaryConditionSet& vbc,
                  const cfd::boundary::BoundaryConditionSet& pbc, const std::string& name,
                  cfdhip::PatchCondition& out, std::string& reason) {
  if (!vbc.has(name) || !pbc.has(name)) {
    reason = "patch '" + name + "' has no velocity or pressure condition";
    return false;
  }
  const auto& v = vbc.get(name);
  switch (v.type()) {
    case BoundaryConditionType::Wall: out.velocityKind = CFDHIP_WALL; break;
    case BoundaryConditionType::Outlet: out.velocityKind = CFDHIP_OUTLET; break;
    case BoundaryConditionType::MovingWall: {
      const auto& w = static_cast<const cfd::boundary::MovingWall&>(v).velocity();
      out.velocityKind = CFDHIP_MOVING_WALL;
      out.velocity[0] = w.x;
      out.velocity[1] = w.y;
      out.velocity[2] = w.z;
      break;
    }
    case BoundaryConditionType::Inlet: {
      const auto& w = static_cast<const cfd::boundary::Inlet&>(v).velocity();
      out.velocityKind = CFDHIP_INLET;
      out.velocity[0] = w.x;
      out.velocity[1] = w.y;
      out.velocity[2] = w.z;
      break;
    }
    default:
      reason = "velocity condition of patch '" + name + "' is not on the structured path";
      return false;
  }
  const auto& p = pbc.get(name);
  if (p.type() == BoundaryConditionType::FixedValue) {
    out.pressureKind = CFDHIP_FIXED_VALUE;
    out.pressureValue = static_cast<const cfd::boundary::FixedValue&>(p).value();
  } else if (p.type() == BoundaryConditionType::FixedGradient) {
    out.pressureKind = CFDHIP_FIXED_GRADIENT;
    out.pressureValue = static_cast<const cfd::boundary::FixedGradient&>(p).gradient();
  } else {
    reason = "pressure condition of patch '" + name + "' is not on the structured path";
    return false;
  }
  return true;
}

// The device's regenerated mesh against the caller's, bit for bit.
bool sameMesh(cfdhip::GpuSimpleDiscretization& d, const cfd::mesh::Mesh& mesh,
              std::string& reason) {
  const std::size_t n = mesh.numberOfCells(), nf = mesh.numberOfFaces();
  if (d.cells() != n || d.faces() != nf) {
    reason = "cell or face count differs from the uniform Cartesian grid";
    return false;
  }
  std::vector<double> centroid(3 * n), volume(n);
  std::vector<int64_t> owner(nf), neighbor(nf);
  d.meshCells(centroid, volume);
  d.meshFaces(owner, neighbor);
  for (std::size_t c = 0; c < n; ++c) {
    const auto& cell = mesh.cell(c);
    const auto& x = cell.centroid();
    if (!sameBits(x.x, centroid[3 * c]) || !sameBits(x.y, centroid[3 * c + 1]) ||
        !sameBits(x.z, centroid[3 * c + 2]) || !sameBits(cell.volume(), volume[c])) {
      reason = "cell geometry differs from the uniform Cartesian grid (cell " +
               std::to_string(c) + ")";
      return false;
    }
  }
  for (std::size_t f = 0; f < nf; ++f) {
    const auto& face = mesh.face(f);
    const int64_t nb = face.neighbor().has_value() ? static_cast<int64_t>(*face.neighbor()) : -1;
    if (static_cast<int64_t>(face.owner()) != owner[f] || nb != neighbor[f]) {
      reason = "face order differs from the uniform Cartesian grid (face " + std::to_string(f) +
               ")";
      return false;
    }
  }
  return true;
}

std::vector<double> component(const VectorField& u, int k) {
  std::vector<double> out(u.size());
  for (std::size_t i = 0; i < u.size(); ++i) out[i] = k == 0 ? u[i].x : (k == 1 ? u[i].y : u[i].z);
  return out;
}

LinearSystem toSystem(cfdhip::HostSystem&& s) {
  Vector rhs(s.rows);
  for (std::size_t i = 0; i < s.rows; ++i) rhs[i] = s.rhs[i];
  SparseMatrix A(s.rows, s.rows, std::move(s.values), std::move(s.columnIndices),
                 std::move(s.rowOffsets));
  return LinearSystem(std::move(A), std::move(rhs));
}

cfdhip::LinearSolverSettings toDevice(const cfd::algebra::LinearSolverSettings& s) {
  cfdhip::LinearSolverSettings d;
  d.absoluteTolerance = s.absoluteTolerance;
  d.relativeTolerance = s.relativeTolerance;
  d.maxIterations = s.maxIterations;
  d.type = s.type == cfd::algebra::LinearSolverType::CG ? cfdhip::LinearSolverType::CG
                                                         : cfdhip::LinearSolverType::BiCGSTAB;
  d.preconditioner = s.preconditioner == cfd::algebra::PreconditionerType::Jacobi
                         ? cfdhip::PreconditionerType::Jacobi
                         : cfdhip::PreconditionerType::None;
  return d;
}

SolverResult fromDevice(cfdhip::SolverResult&& r) {
  SolverResult out;  // solution intentionally empty: it stays on the device
  out.status = static_cast<cfd::algebra::SolverStatus>(static_cast<int>(r.status));
  out.iterations = r.iterations;
  out.initialResidual = r.initialResidual;
  out.finalResidual = r.finalResidual;
  out.residualHistory = std::move(r.residualHistory);
  out.restarts = r.restarts;
  out.backendUsed = cfd::algebra::LinearSolverBackend::GPU;
  return out;
}

}  // namespace

struct GpuSimpleDiscretization::Impl {
  bool available = false;
  bool ready = false;
  cfdhip::GpuSimpleDiscretization d;
};

GpuSimpleDiscretization::GpuSimpleDiscretization() : impl_(std::make_unique<Impl>()) {
  impl_->available = cfd::gpu::cudaAvailable();
}
GpuSimpleDiscretization::~GpuSimpleDiscretization() = default;
GpuSimpleDiscretization::GpuSimpleDiscretization(GpuSimpleDiscretization&&) noexcept = default;
GpuSimpleDiscretization& GpuSimpleDiscretization::operator=(GpuSimpleDiscretization&&) noexcept =
    default;

bool GpuSimpleDiscretization::active() const noexcept { return impl_ && impl_->available; }

bool GpuSimpleDiscretization::prepare(const cfd::mesh::Mesh& mesh,
                                      const cfd::boundary::BoundaryConditionSet& velocityBoundaries,
                                      const cfd::boundary::BoundaryConditionSet& pressureBoundaries,
                                      std::string& reason) {
  impl_->ready = false;
  if (!impl_->available) {
    reason = "no usable gfx950 device";
    return false;
  }
  const auto* grid = mesh.structuredGrid();
  if (grid == nullptr || mesh.structuredBlocks().size() > 1) {
    reason = "the structured kernels take one Cartesian block";
    return false;
  }
  const int dim = mesh.dimension();
  cfdhip::CartesianMesh cm;
  cm.dim = dim;
  cm.nx = grid->nx;
  cm.ny = grid->ny;
  cm.nz = dim == 3 ? grid->nz : 1;
  if (dim == 3) {
    const auto& hi = grid->vertex(grid->nx, grid->ny, grid->nz);
    cm.lx = hi.x;
    cm.ly = hi.y;
    cm.lz = hi.z;
  } else {
    const auto& hi = grid->vertex(grid->nx, grid->ny);
    cm.lx = hi.x;
    cm.ly = hi.y;
  }
  const int patchCount = dim == 3 ? 6 : 4;
  if (mesh.boundaryPatches().size() != static_cast<std::size_t>(patchCount)) {
    reason = "the structured kernels take the Cartesian grid's own boundary patches";
    return false;
  }
  std::vector<cfdhip::PatchCondition> patches(patchCount);
  for (int p = 0; p < patchCount; ++p) {
    const std::string name = dim == 3 ? kPatches3[p] : kPatches2[p];
    if (mesh.boundaryPatches()[p].name() != name) {
      reason = "boundary patch " + std::to_string(p) + " is not '" + name + "'";
      return false;
    }
    if (!mapCondition(velocityBoundaries, pressureBoundaries, name, patches[p], reason))
      return false;
  }
  if (!impl_->d.prepare(cm, patches, reason)) return false;
  if (!translated([&] { return sameMesh(impl_->d, mesh, reason); })) return false;
  impl_->ready = true;
  reason.clear();
  return true;
}

void GpuSimpleDiscretization::beginIteration(const VectorField& velocity,
                                             const ScalarField& pressure,
                                             const SurfaceField& massFlux,
                                             const ScalarField& effectiveViscosity) {
  if (!impl_->ready) return;
  const std::size_t n = velocity.size();
  const auto u = component(velocity, 0), v = component(velocity, 1), w = component(velocity, 2);
  translated([&] {
    impl_->d.uploadInitialState(u, v, w, std::span<const double>(pressure.data(), n),
                                std::span<const double>(massFlux.data(), massFlux.size()),
                                std::span<const double>(effectiveViscosity.data(), n));
  });
}

void GpuSimpleDiscretization::uploadInitialState(const VectorField& velocity,
                                                 const ScalarField& pressure,
                                                 const SurfaceField& massFlux,
                                                 const ScalarField& effectiveViscosity) {
  beginIteration(velocity, pressure, massFlux, effectiveViscosity);
}

GpuSimpleDiscretization::FieldAuthority GpuSimpleDiscretization::authority(
    PersistentField field) const noexcept {
  using A = cfdhip::GpuSimpleDiscretization::FieldAuthority;
  using P = cfdhip::GpuSimpleDiscretization::PersistentField;
  const P f = field == PersistentField::Velocity   ? P::Velocity
              : field == PersistentField::Pressure ? P::Pressure
              : field == PersistentField::MassFlux ? P::MassFlux
                                                   : P::Viscosity;
  const A a = impl_->d.authority(f);
  return a == A::DeviceOwned    ? FieldAuthority::DeviceOwned
         : a == A::Synchronized ? FieldAuthority::Synchronized
                                : FieldAuthority::HostOnly;
}

void GpuSimpleDiscretization::beginIterationResident() {
  if (!impl_->ready) return;
  translated([&] { impl_->d.beginIterationResident(); });
}

void GpuSimpleDiscretization::setViscosity(const ScalarField& effectiveViscosity) {
  if (!impl_->ready) return;
  translated([&] {
    impl_->d.setViscosity(
        std::span<const double>(effectiveViscosity.data(), effectiveViscosity.size()));
  });
}

void GpuSimpleDiscretization::updatePressure(cfd::Real relaxationAlpha) {
  if (!impl_->ready) return;
  translated([&] { impl_->d.updatePressure(relaxationAlpha); });
}

void GpuSimpleDiscretization::downloadVelocity(VectorField& out) {
  if (!impl_->ready) return;
  const std::size_t n = impl_->d.cells();
  std::vector<double> u(n), v(n), w(n);
  translated([&] { impl_->d.downloadVelocity(u, v, w); });
  out = VectorField(n);
  for (std::size_t c = 0; c < n; ++c) out[c] = cfd::Vector3{u[c], v[c], w[c]};
}

void GpuSimpleDiscretization::downloadPressure(ScalarField& out) {
  if (!impl_->ready) return;
  out = ScalarField(impl_->d.cells());
  translated([&] { impl_->d.downloadPressure(std::span<double>(out.data(), out.size())); });
}

void GpuSimpleDiscretization::downloadMassFlux(SurfaceField& out) {
  if (!impl_->ready) return;
  out = SurfaceField(impl_->d.faces());
  translated([&] { impl_->d.downloadMassFlux(std::span<double>(out.data(), out.size())); });
}

void GpuSimpleDiscretization::correctVelocityResident(cfd::Index gradientScheme,
                                                      bool threeDimensional) {
  if (!impl_->ready) return;
  translated([&] {
    impl_->d.correctVelocityResident(static_cast<int>(gradientScheme), threeDimensional);
  });
}

void GpuSimpleDiscretization::correctFaceMassFluxResident(bool withExplicitTerm) {
  if (!impl_->ready) return;
  translated([&] { impl_->d.correctFaceMassFluxResident(withExplicitTerm); });
}

bool GpuSimpleDiscretization::residentPressureAllFinite() {
  if (!impl_->ready) return true;
  return translated([&] { return impl_->d.residentPressureAllFinite(); });
}

LinearSystem GpuSimpleDiscretization::assembleMomentum(cfd::Index component,
                                                       const ScalarField& previousComponent,
                                                       cfd::Real relaxationAlpha,
                                                       cfd::Index convectionScheme,
                                                       bool applyNonOrthogonalCorrection) {
  return translated([&] {
    return toSystem(impl_->d.assembleMomentum(
        static_cast<int>(component),
        std::span<const double>(previousComponent.data(), previousComponent.size()),
        relaxationAlpha, static_cast<int>(convectionScheme), applyNonOrthogonalCorrection));
  });
}

void GpuSimpleDiscretization::setMomentumSolution(cfd::Index component, const Vector& solution) {
  if (!impl_->ready) return;
  translated([&] {
    impl_->d.setMomentumSolution(static_cast<int>(component),
                                 std::span<const double>(solution.data(), solution.size()));
  });
}

void GpuSimpleDiscretization::assembleMomentumResident(cfd::Index component,
                                                       cfd::Real relaxationAlpha,
                                                       cfd::Index convectionScheme,
                                                       bool applyNonOrthogonalCorrection) {
  if (!impl_->ready) return;
  translated([&] {
    impl_->d.assembleMomentumResident(static_cast<int>(component), relaxationAlpha,
                                      static_cast<int>(convectionScheme),
                                      applyNonOrthogonalCorrection);
  });
}

SolverResult GpuSimpleDiscretization::solveMomentumResident(
    cfd::Index component, const cfd::algebra::LinearSolverSettings& settings) {
  return translated([&] {
    return fromDevice(
        impl_->d.solveMomentumResident(static_cast<int>(component), toDevice(settings)));
  });
}

bool GpuSimpleDiscretization::residentMomentumPredictorAllFinite(bool threeDimensional) {
  return translated([&] { return impl_->d.residentMomentumPredictorAllFinite(threeDimensional); });
}

bool GpuSimpleDiscretization::residentVelocityAllFinite() {
  return translated([&] { return impl_->d.residentVelocityAllFinite(); });
}

void GpuSimpleDiscretization::computeResponseCoefficients(bool threeDimensional) {
  if (!impl_->ready) return;
  translated([&] { impl_->d.computeResponseCoefficients(threeDimensional); });
}

void GpuSimpleDiscretization::computePredictedFaceFlux(bool rhieChow, cfd::Real density,
                                                       cfd::Real relaxationAlpha,
                                                       cfd::Index gradientScheme,
                                                       bool threeDimensional) {
  if (!impl_->ready) return;
  translated([&] {
    impl_->d.computePredictedFaceFlux(rhieChow, density, relaxationAlpha,
                                      static_cast<int>(gradientScheme), threeDimensional);
  });
}

LinearSystem GpuSimpleDiscretization::assemblePressureCorrection(
    bool nonOrthogonal, cfd::Index referenceCell, cfd::Real density,
    const ScalarField* previousPressureCorrection, bool threeDimensional) {
  return translated([&] {
    const std::span<const double> previous =
        previousPressureCorrection != nullptr
            ? std::span<const double>(previousPressureCorrection->data(),
                                      previousPressureCorrection->size())
            : std::span<const double>();
    return toSystem(impl_->d.assemblePressureCorrection(nonOrthogonal, referenceCell, density,
                                                        previous, threeDimensional));
  });
}

void GpuSimpleDiscretization::setPressureCorrection(const Vector& pressureCorrection) {
  if (!impl_->ready) return;
  translated([&] {
    impl_->d.setPressureCorrection(
        std::span<const double>(pressureCorrection.data(), pressureCorrection.size()));
  });
}

void GpuSimpleDiscretization::assemblePressureCorrectionResident(bool nonOrthogonal,
                                                                 cfd::Index referenceCell,
                                                                 cfd::Real density,
                                                                 bool threeDimensional) {
  translated([&] {
    impl_->d.assemblePressureCorrectionResident(nonOrthogonal, referenceCell, density,
                                                threeDimensional);
  });
}

SolverResult GpuSimpleDiscretization::solvePressureCorrectionResident(
    const cfd::algebra::LinearSolverSettings& settings) {
  return translated(
      [&] { return fromDevice(impl_->d.solvePressureCorrectionResident(toDevice(settings))); });
}

bool GpuSimpleDiscretization::residentPressureCorrectionAllFinite() {
  return translated([&] { return impl_->d.residentPressureCorrectionAllFinite(); });
}

std::size_t GpuSimpleDiscretization::residentSolveBytes() const noexcept {
  return impl_ ? impl_->d.residentSolveBytes() : 0;
}

void GpuSimpleDiscretization::correctVelocity(cfd::Index gradientScheme, bool threeDimensional,
                                              VectorField& corrected) {
  if (!impl_->ready) return;
  const std::size_t n = impl_->d.cells();
  std::vector<double> out(3 * n);
  translated([&] {
    impl_->d.correctVelocity(static_cast<int>(gradientScheme), threeDimensional, out);
  });
  corrected = VectorField(n);
  for (std::size_t c = 0; c < n; ++c)
    corrected[c] = cfd::Vector3{out[c], out[n + c], out[2 * n + c]};
}

void GpuSimpleDiscretization::correctFaceMassFlux(bool withExplicitTerm, SurfaceField& corrected) {
  if (!impl_->ready) return;
  corrected = SurfaceField(impl_->d.faces());
  translated([&] {
    impl_->d.correctFaceMassFlux(withExplicitTerm,
                                 std::span<double>(corrected.data(), corrected.size()));
  });
}

}  // namespace cfd::gpu
