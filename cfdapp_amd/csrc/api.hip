// extern "C" entry points of libcfdhip (declared in include/cfdhip.h).
// Each catches every library exception, records the message on the context
// and returns the error code, so no exception crosses the C boundary.
#include <cmath>
#include <cstddef>
#include <cstdio>

#include "context.hpp"
#include "unet_net.hpp"

using namespace cfdhip;

struct cfdhip_ctx {
  Context c;
};

namespace cfdhip {

Context::~Context() {
  if (kw.hostState) (void)hipHostFree(kw.hostState);
  if (kw.hostState2) (void)hipHostFree(kw.hostState2);
  if (stream) (void)hipStreamDestroy(stream);
}

void Context::ensureFields() {
  const int64_t n = view.n, nf = view.nf;
  for (int id = 0; id < CFDHIP_FIELD_COUNT; ++id) {
    const bool face = id == CFDHIP_FIELD_FLUX || id == CFDHIP_FIELD_FLUX_STAR ||
                      id == CFDHIP_FIELD_PC_FACE_COEFF;
    field[id].ensure(face ? nf : n);
    CFDHIP_CHECK(hipMemsetAsync(field[id].get(), 0, sizeof(double) * (face ? nf : n), stream));
  }
  CFDHIP_CHECK(hipStreamSynchronize(stream));
}

int64_t KrylovWork::bytes() const {
  int64_t b = 0;
  for (const DeviceArray<double>* a : {&r, &p, &Ap, &z, &x, &this->b, &dinv, &rhat, &v, &s, &t,
                                       &phat, &shat, &partial, &hist, &scal})
    b += a->count * static_cast<int64_t>(sizeof(double));
  for (const DeviceArray<float>* a : {&r32, &x32, &dinv32, &w32, &dis32, &scaleTmp[0],
                                      &scaleTmp[1]})
    b += a->count * static_cast<int64_t>(sizeof(float));
  for (int q = 0; q < 8; ++q)
    b += d8[q].count * static_cast<int64_t>(sizeof(double)) +
         f8[q].count * static_cast<int64_t>(sizeof(float));
  b += state.count * static_cast<int64_t>(sizeof(KState));
  return b;
}

}  // namespace cfdhip

namespace {

const int kGfx950 = 950;

bool isGfx950(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

template <class F>
int guarded(cfdhip_ctx* ctx, F&& fn) {
  if (ctx == nullptr) return CFDHIP_ERR_INVALID;
  try {
    CFDHIP_CHECK(hipSetDevice(ctx->c.device));
    fn(ctx->c);
    ctx->c.lastError.clear();
    return CFDHIP_OK;
  } catch (const Error& e) {
    ctx->c.lastError = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.lastError = e.what();
    return CFDHIP_ERR_RUNTIME;
  }
}

int fieldCount(const Context& c, int id) {
  const bool face = id == CFDHIP_FIELD_FLUX || id == CFDHIP_FIELD_FLUX_STAR ||
                    id == CFDHIP_FIELD_PC_FACE_COEFF;
  return static_cast<int>(face ? c.view.nf : c.view.n);
}

cfdhip_solver_settings checkedSettings(const cfdhip_solver_settings* s) {
  require(s != nullptr, "solver settings are required");
  require(s->inner_tolerance >= 0.0 && s->inner_tolerance < 1.0,
          "LinearSolverSettings: inner_tolerance must be in [0, 1)");
  return *s;
}

void uploadCsr(Context& c, CsrDevice& M, int64_t n, const int64_t* rp, const int64_t* col,
               const double* val) {
  require(n > 0 && rp && col && val, "CSR arrays are required");
  require(rp[0] == 0, "SparseMatrix: rowOffsets[0] must be 0");
  for (int64_t r = 0; r < n; ++r) {
    require(rp[r] <= rp[r + 1], "SparseMatrix: rowOffsets must be non-decreasing");
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      require(col[k] >= 0 && col[k] < n, "SparseMatrix: column index out of range");
      require(k == rp[r] || col[k] > col[k - 1],
              "SparseMatrix: column indices within a row must be strictly ascending");
    }
  }
  const int64_t nnz = rp[n];
  M.n = n;
  M.rp.ensure(n + 1);
  M.col.ensure(nnz > 0 ? nnz : 1);
  M.val.ensure(nnz > 0 ? nnz : 1);
  M.val.count = nnz;
  CFDHIP_CHECK(hipMemcpyAsync(M.rp.get(), rp, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                              c.stream));
  if (nnz > 0) {
    CFDHIP_CHECK(hipMemcpyAsync(M.col.get(), col, nnz * sizeof(int64_t), hipMemcpyHostToDevice,
                                c.stream));
    CFDHIP_CHECK(hipMemcpyAsync(M.val.get(), val, nnz * sizeof(double), hipMemcpyHostToDevice,
                                c.stream));
  }
}

}  // namespace

extern "C" {

int cfdhip_device_available(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
  for (int d = 0; d < count; ++d)
    if (isGfx950(d)) return 1;
  (void)kGfx950;
  return 0;
}

cfdhip_ctx* cfdhip_create(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
  if (!isGfx950(device)) return nullptr;
  auto* ctx = new (std::nothrow) cfdhip_ctx;
  if (!ctx) return nullptr;
  try {
    ctx->c.device = device;
    CFDHIP_CHECK(hipSetDevice(device));
    CFDHIP_CHECK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    ctx->c.flags.ensure(4);
    CFDHIP_CHECK(hipMemsetAsync(ctx->c.flags.get(), 0, 4 * sizeof(unsigned), ctx->c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  } catch (...) {
    delete ctx;
    return nullptr;
  }
  return ctx;
}

void cfdhip_destroy(cfdhip_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  delete ctx;
}

const char* cfdhip_last_error(const cfdhip_ctx* ctx) {
  return ctx ? ctx->c.lastError.c_str() : "null context";
}

int cfdhip_synchronize(cfdhip_ctx* ctx) {
  return guarded(ctx, [](Context& c) { CFDHIP_CHECK(hipStreamSynchronize(c.stream)); });
}

int cfdhip_mesh_cartesian(cfdhip_ctx* ctx, int dim, int64_t nx, int64_t ny, int64_t nz, double lx,
                          double ly, double lz) {
  return guarded(ctx, [&](Context& c) { buildCartesianPlan(c, dim, nx, ny, nz, lx, ly, lz); });
}

int cfdhip_mesh_sizes(cfdhip_ctx* ctx, int64_t* out) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(out != nullptr, "out is required");
    out[0] = c.view.n;
    out[1] = c.view.nf;
    out[2] = c.view.npatch;
    out[3] = c.view.dim;
  });
}

int cfdhip_mesh_faces(cfdhip_ctx* ctx, int64_t* owner, int64_t* neighbor) {
  return guarded(ctx, [&](Context& c) {
    require(owner && neighbor, "owner/neighbor buffers are required");
    exportFaces(c, owner, neighbor);
  });
}

int cfdhip_mesh_cells(cfdhip_ctx* ctx, double* centroid, double* volume) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(centroid != nullptr, "centroid buffer is required");
    const View& v = c.view;
    // MeshGeometry::createCartesian2D/3D: (i + 0.5) * delta per axis, the 3D
    // coordinates offset by the zero origin (0.0 + x == x for these x >= 0)
    for (int64_t k = 0; k < v.nz; ++k)
      for (int64_t j = 0; j < v.ny; ++j)
        for (int64_t i = 0; i < v.nx; ++i) {
          const int64_t id = i + v.nx * (j + v.ny * k);
          centroid[3 * id] = (static_cast<double>(i) + 0.5) * c.delta[0];
          centroid[3 * id + 1] = (static_cast<double>(j) + 0.5) * c.delta[1];
          centroid[3 * id + 2] = v.dim == 3 ? (static_cast<double>(k) + 0.5) * c.delta[2] : 0.0;
          if (volume != nullptr) volume[id] = v.vol;
        }
  });
}

int cfdhip_set_boundary(cfdhip_ctx* ctx, int patch, int velocity_kind, const double* velocity3,
                        int pressure_kind, double pressure_value) {
  return guarded(ctx, [&](Context& c) {
    setBoundary(c, patch, velocity_kind, velocity3, pressure_kind, pressure_value);
  });
}

int cfdhip_set_fluid(cfdhip_ctx* ctx, double density, double viscosity) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(std::isfinite(density) && density > 0.0, "density must be finite and > 0");
    require(std::isfinite(viscosity) && viscosity > 0.0,
            "effectiveViscosity must be finite and > 0");
    c.rho = density;
    c.mu = viscosity;
    c.view.rho = density;
    std::vector<double> mu(c.view.n, viscosity + 0.0);
    CFDHIP_CHECK(hipMemcpyAsync(c.f(CFDHIP_FIELD_VISCOSITY), mu.data(), mu.size() * sizeof(double),
                                hipMemcpyHostToDevice, c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
    c.fluidSet = true;
  });
}

int cfdhip_field_upload(cfdhip_ctx* ctx, int field, const double* host, int64_t count) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(field >= 0 && field < CFDHIP_FIELD_COUNT, "unknown field");
    require(host != nullptr && count == fieldCount(c, field), "field size mismatch");
    CFDHIP_CHECK(hipMemcpyAsync(c.f(field), host, count * sizeof(double), hipMemcpyHostToDevice,
                                c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
    if (field == CFDHIP_FIELD_VISCOSITY) c.fluidSet = true;
  });
}

int cfdhip_field_device(cfdhip_ctx* ctx, int field, void** out) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(field >= 0 && field < CFDHIP_FIELD_COUNT, "unknown field");
    require(out != nullptr, "out is required");
    *out = c.f(field);
  });
}

int cfdhip_field_download(cfdhip_ctx* ctx, int field, double* host, int64_t count) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(field >= 0 && field < CFDHIP_FIELD_COUNT, "unknown field");
    require(host != nullptr && count == fieldCount(c, field), "field size mismatch");
    CFDHIP_CHECK(hipMemcpyAsync(host, c.f(field), count * sizeof(double), hipMemcpyDeviceToHost,
                                c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_gradient(cfdhip_ctx* ctx, int field, int bc) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(field >= 0 && field < CFDHIP_FIELD_COUNT && fieldCount(c, field) == c.view.n,
            "gradient needs a cell field");
    require(bc == 0 || bc == 1, "bc must be 0 (pressure) or 1 (p')");
    launchGradient(c, c.f(field), bc, c.f(CFDHIP_FIELD_GRAD_X), c.f(CFDHIP_FIELD_GRAD_Y),
                   c.f(CFDHIP_FIELD_GRAD_Z));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_mass_flux(cfdhip_ctx* ctx) {
  return guarded(ctx, [&](Context& c) {
    launchMassFlux(c, c.f(CFDHIP_FIELD_U), c.f(CFDHIP_FIELD_V), c.f(CFDHIP_FIELD_W),
                   c.f(CFDHIP_FIELD_FLUX));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_assemble_momentum(cfdhip_ctx* ctx, double alpha) {
  return guarded(ctx, [&](Context& c) { launchMomentum(c, alpha); });
}

int cfdhip_assemble_momentum_component(cfdhip_ctx* ctx, int component, double alpha,
                                       int convection_scheme, int non_orthogonal,
                                       const double* previous_component) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(component >= 0 && component < c.view.dim, "component out of range");
    // cfd::discretization::ConvectionScheme: 0 = Upwind, the scheme these
    // kernels implement (MomentumEquation.cpp:264-327)
    require(convection_scheme == 0,
            "the structured momentum kernels implement the Upwind convection scheme only",
            CFDHIP_ERR_UNSUPPORTED);
    require(non_orthogonal == 0,
            "the explicit non-orthogonal momentum correction is not on the structured path",
            CFDHIP_ERR_UNSUPPORTED);
    const double* old = nullptr;
    if (previous_component != nullptr) {
      c.stage.ensure(3 * c.view.n);
      CFDHIP_CHECK(hipMemcpyAsync(c.stage.get(), previous_component, c.view.n * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
      old = c.stage.get();
    }
    launchMomentum(c, alpha, 0.0, component, old);
  });
}

int cfdhip_assemble_transient_momentum(cfdhip_ctx* ctx, double dt) {
  return guarded(ctx, [&](Context& c) {
    require(std::isfinite(dt) && dt > 0.0,
            "assembleTransientMomentumComponent: dt must be finite and > 0");
    require(c.view.dim == 2, "assembleTransientMomentumComponent: the mesh must be 2D",
            CFDHIP_ERR_UNSUPPORTED);
    launchMomentum(c, 1.0, dt);
  });
}

int cfdhip_solve_momentum(cfdhip_ctx* ctx, int component, const cfdhip_solver_settings* s,
                          cfdhip_solve_result* r) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(component >= 0 && component < c.view.dim, "component out of range");
    const cfdhip_solver_settings set = checkedSettings(s);
    double* x = c.f(CFDHIP_FIELD_USTAR + component);
    CFDHIP_CHECK(hipMemcpyAsync(x, c.f(CFDHIP_FIELD_U + component), c.view.n * sizeof(double),
                                hipMemcpyDeviceToDevice, c.stream));
    const cfdhip_solve_result res =
        solveStencil(c, set, c.f(CFDHIP_FIELD_MOM_RHS_U + component), x);
    if (r) *r = res;
  });
}

int cfdhip_response(cfdhip_ctx* ctx) {
  return guarded(ctx, [&](Context& c) {
    launchResponse(c);
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_predicted_flux(cfdhip_ctx* ctx, int rhie_chow, double alpha) {
  return guarded(ctx, [&](Context& c) {
    launchPredictedFlux(c, rhie_chow != 0, alpha);
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_assemble_pressure(cfdhip_ctx* ctx, int64_t reference_cell) {
  return guarded(ctx, [&](Context& c) { launchPressureAssembly(c, reference_cell); });
}

int cfdhip_assemble_compressible_pressure(cfdhip_ctx* ctx, int64_t reference_cell,
                                          const double* face_density,
                                          const double* pressure_absolute,
                                          const double* temperature, double gas_constant,
                                          double pseudo_time_step) {
  return guarded(ctx, [&](Context& c) {
    require(face_density != nullptr && pressure_absolute != nullptr && temperature != nullptr,
            "assembleCompressiblePressureCorrection: face density, absolute pressure and "
            "temperature are required");
    launchCompressiblePressureAssembly(c, reference_cell, face_density, pressure_absolute,
                                       temperature, gas_constant, pseudo_time_step);
  });
}

int cfdhip_fmesh_set(cfdhip_ctx* ctx, int dim, int64_t n_cells, int64_t n_faces,
                     const int64_t* owner, const int64_t* neighbor, const double* face_centroid,
                     const double* area_vector, const double* cell_centroid,
                     const int64_t* cell_face_offsets, const int64_t* cell_faces,
                     const int32_t* fixed_value) {
  return guarded(ctx, [&](Context& c) {
    fmeshSet(c, dim, n_cells, n_faces, owner, neighbor, face_centroid, area_vector, cell_centroid,
             cell_face_offsets, cell_faces, fixed_value);
  });
}

int cfdhip_fmesh_assemble_pressure(cfdhip_ctx* ctx, const double* flux_star,
                                   const double* face_density, double density, const double* du,
                                   const double* dv, const double* dw, int non_orthogonal,
                                   const double* grad_previous, const double* extra_diagonal,
                                   int64_t reference_cell) {
  return guarded(ctx, [&](Context& c) {
    fmeshAssemblePressure(c, flux_star, face_density, density, du, dv, dw, non_orthogonal,
                          grad_previous, extra_diagonal, reference_cell);
  });
}

int cfdhip_fmesh_pressure_system(cfdhip_ctx* ctx, int64_t* rowptr, int64_t* col, double* val,
                                 int64_t capacity, int64_t* nnz, double* rhs, double* face_coeff,
                                 double* explicit_flux) {
  return guarded(ctx, [&](Context& c) {
    const int64_t k = fmeshSystem(c, rowptr, col, val, capacity, rhs, face_coeff, explicit_flux);
    if (nnz) *nnz = k;
  });
}

int cfdhip_fmesh_solve_pressure(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, double* x,
                                cfdhip_solve_result* r) {
  return guarded(ctx, [&](Context& c) {
    const cfdhip_solve_result res = fmeshSolve(c, checkedSettings(s), x);
    if (r) *r = res;
  });
}

int cfdhip_fmesh_correct_flux(cfdhip_ctx* ctx, const double* pprime, double* flux) {
  return guarded(ctx, [&](Context& c) { fmeshCorrectFlux(c, pprime, flux); });
}

int cfdhip_fmesh_set_cell_volumes(cfdhip_ctx* ctx, const double* cell_volume) {
  return guarded(ctx, [&](Context& c) { fmeshSetVolumes(c, cell_volume); });
}

int cfdhip_fmesh_gradient(cfdhip_ctx* ctx, const double* field, const int32_t* bc_kind,
                          const double* bc_value, int scheme, double* grad) {
  return guarded(ctx, [&](Context& c) {
    require(c.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
    FaceMesh& m = *c.fmesh;
    require(field != nullptr && grad != nullptr, "gradient: field and output are required");
    require((bc_kind == nullptr) == (bc_value == nullptr),
            "gradient: condition kinds and values come together");
    DeviceArray<double> phi, out;
    phi.ensure(m.n);
    out.ensure(3 * m.n);
    CFDHIP_CHECK(hipMemcpyAsync(phi.get(), field, m.n * sizeof(double), hipMemcpyHostToDevice,
                                c.stream));
    if (bc_kind == nullptr) {
      fmeshPprimeConditions(c, m);
    } else {
      for (int64_t f = 0; f < m.nf; ++f)
        require(m.hNeigh[f] >= 0 || bc_kind[f] == 0 || bc_kind[f] == 1,
                "gradient: a boundary condition kind is 0 (FixedGradient) or 1 (FixedValue)");
      m.bcKind.ensure(m.nf);
      m.bcVal.ensure(m.nf);
      CFDHIP_CHECK(hipMemcpyAsync(m.bcKind.get(), bc_kind, m.nf * sizeof(int32_t),
                                  hipMemcpyHostToDevice, c.stream));
      CFDHIP_CHECK(hipMemcpyAsync(m.bcVal.get(), bc_value, m.nf * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
    }
    fmeshGradient(c, phi.get(), m.bcKind.get(), m.bcVal.get(), scheme, out.get());
    CFDHIP_CHECK(hipMemcpyAsync(grad, out.get(), 3 * m.n * sizeof(double), hipMemcpyDeviceToHost,
                                c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_fmesh_assemble_pressure_pass(cfdhip_ctx* ctx, const double* previous_pprime,
                                        int gradient_scheme) {
  return guarded(ctx, [&](Context& c) {
    require(c.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
    FaceMesh& m = *c.fmesh;
    DeviceArray<double> prev;
    const double* p = m.x.get();
    if (previous_pprime != nullptr) {
      prev.ensure(m.n);
      CFDHIP_CHECK(hipMemcpyAsync(prev.get(), previous_pprime, m.n * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
      p = prev.get();
    } else {
      require(m.solved, "face-list p': no previous p' given and none solved");
    }
    fmeshAssemblePass(c, p, gradient_scheme);
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_fmesh_pressure_passes(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, int passes,
                                 int gradient_scheme, double* x, cfdhip_solve_result* results) {
  return guarded(ctx, [&](Context& c) {
    require(c.fmesh != nullptr, "no face-list mesh: call cfdhip_fmesh_set first");
    FaceMesh& m = *c.fmesh;
    require(passes >= 1, "nonOrthogonalCorrections: at least one pass");
    require(m.assembled && !m.hasExplicit,
            "face-list p': the passes start from a first-pass assembly (no explicit term)");
    const cfdhip_solver_settings set = checkedSettings(s);
    for (int k = 0; k < passes; ++k) {
      if (k > 0) fmeshAssemblePass(c, m.x.get(), gradient_scheme);
      const cfdhip_solve_result r = fmeshSolve(c, set, nullptr);
      if (results) results[k] = r;
      // SIMPLE.cpp:904-911: a pass that does not converge ends the loop
      if (r.status != CFDHIP_CONVERGED) {
        for (int q = k + 1; results && q < passes; ++q) {
          results[q] = cfdhip_solve_result{};
          results[q].status = -1;
        }
        break;
      }
    }
    if (x) {
      CFDHIP_CHECK(hipMemcpyAsync(x, m.x.get(), m.n * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
      CFDHIP_CHECK(hipStreamSynchronize(c.stream));
    }
  });
}

int cfdhip_solve_pressure(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, cfdhip_solve_result* r) {
  return guarded(ctx, [&](Context& c) {
    const cfdhip_solver_settings set = checkedSettings(s);
    const cfdhip_solve_result res =
        solvePressure(c, set, c.f(CFDHIP_FIELD_PC_RHS), c.f(CFDHIP_FIELD_PPRIME), true);
    if (r) *r = res;
  });
}

int cfdhip_update_pressure(cfdhip_ctx* ctx, double alpha) {
  return guarded(ctx, [&](Context& c) {
    launchUpdatePressure(c, alpha);
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_correct(cfdhip_ctx* ctx) {
  return guarded(ctx, [&](Context& c) {
    launchCorrect(c);
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_correct_velocity(cfdhip_ctx* ctx, double* out) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    const int64_t n = c.view.n;
    if (out == nullptr) {
      launchCorrectVelocity(c);
    } else {
      c.stage.ensure(3 * n);
      double* dest[3] = {c.stage.get(), c.stage.get() + n, c.stage.get() + 2 * n};
      // a 2D correction leaves W = 0, as the reference's Vector2 velocity
      if (c.view.dim == 2) CFDHIP_CHECK(hipMemsetAsync(dest[2], 0, n * sizeof(double), c.stream));
      launchCorrectVelocity(c, dest);
      CFDHIP_CHECK(hipMemcpyAsync(out, c.stage.get(), 3 * n * sizeof(double),
                                  hipMemcpyDeviceToHost, c.stream));
    }
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_correct_flux(cfdhip_ctx* ctx, double* out) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    const int64_t nf = c.view.nf;
    if (out == nullptr) {
      launchCorrectFlux(c);
    } else {
      c.stage.ensure(nf > 3 * c.view.n ? nf : 3 * c.view.n);
      launchCorrectFlux(c, c.stage.get());
      CFDHIP_CHECK(hipMemcpyAsync(out, c.stage.get(), nf * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
    }
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_fields_all_finite(cfdhip_ctx* ctx, const int* fields, int count, int* all_finite) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(all_finite != nullptr && count >= 0 && (count == 0 || fields != nullptr),
            "a field list and an output are required");
    require(count <= CFDHIP_FIELD_COUNT, "too many fields");
    const double* arrays[CFDHIP_FIELD_COUNT];
    int64_t counts[CFDHIP_FIELD_COUNT];
    for (int k = 0; k < count; ++k) {
      require(fields[k] >= 0 && fields[k] < CFDHIP_FIELD_COUNT, "unknown field");
      arrays[k] = c.f(fields[k]);
      counts[k] = fieldCount(c, fields[k]);
    }
    *all_finite = countNonFinite(c, arrays, counts, count) == 0 ? 1 : 0;
  });
}

int cfdhip_set_density(cfdhip_ctx* ctx, double density) {
  return guarded(ctx, [&](Context& c) {
    c.requireMesh();
    require(std::isfinite(density) && density > 0.0, "density must be finite and > 0");
    c.rho = density;
    c.view.rho = density;
  });
}

int cfdhip_resident_bytes(cfdhip_ctx* ctx, int64_t* bytes) {
  return guarded(ctx, [&](Context& c) {
    require(bytes != nullptr, "bytes is required");
    *bytes = c.kw.bytes();
  });
}

int cfdhip_continuity(cfdhip_ctx* ctx, double* out) {
  return guarded(ctx, [&](Context& c) {
    double tmp[2];
    launchContinuity(c, tmp);
    if (out) {
      out[0] = tmp[0];
      out[1] = tmp[1];
    }
  });
}

int cfdhip_simple_run(cfdhip_ctx* ctx, const double* settings,
                      const cfdhip_solver_settings* momentum,
                      const cfdhip_solver_settings* pressure, double* meta, double* history,
                      int64_t cap) {
  return guarded(ctx, [&](Context& c) {
    require(settings && momentum && pressure && meta, "settings, solvers and meta are required");
    simpleRun(c, settings, checkedSettings(momentum), checkedSettings(pressure), meta, history,
              cap);
  });
}

int cfdhip_piso_run(cfdhip_ctx* ctx, double dt, int64_t steps, int64_t reference_cell,
                    const cfdhip_solver_settings* momentum, const cfdhip_solver_settings* pressure,
                    double* meta, int64_t* taken) {
  return guarded(ctx, [&](Context& c) {
    require(momentum && pressure && meta && taken, "solvers, meta and taken are required");
    // LinearSolver::validateSettings (src/algebra/LinearSolver.cpp:9-28): PISO
    // reports bad settings as InvalidConfiguration (PISO.cpp:142-163)
    auto valid = [](const cfdhip_solver_settings& s) {
      return s.absolute_tolerance >= 0.0 && s.relative_tolerance >= 0.0 &&
             (s.absolute_tolerance > 0.0 || s.relative_tolerance > 0.0) && s.max_iterations >= 1;
    };
    if (!valid(*momentum) || !valid(*pressure)) {
      for (int k = 0; k < 6; ++k) meta[k] = 0.0;
      meta[0] = 4;  // TransientStepStatus::InvalidConfiguration
      *taken = 1;
      return;
    }
    *taken = pisoRun(c, dt, steps, reference_cell, *momentum, *pressure, meta);
  });
}

int cfdhip_pressure_csr(cfdhip_ctx* ctx, int64_t* rowptr, int64_t* col, double* val,
                        int64_t capacity, int64_t* nnz) {
  return guarded(ctx, [&](Context& c) {
    require(rowptr && col && val && nnz, "CSR buffers are required");
    *nnz = exportPressureCsr(c, rowptr, col, val, capacity);
  });
}

int cfdhip_momentum_csr(cfdhip_ctx* ctx, int64_t* rowptr, int64_t* col, double* val,
                        int64_t capacity, int64_t* nnz) {
  return guarded(ctx, [&](Context& c) {
    require(rowptr && col && val && nnz, "CSR buffers are required");
    *nnz = exportMomentumCsr(c, rowptr, col, val, capacity);
  });
}

int cfdhip_csr_spmv(cfdhip_ctx* ctx, int64_t n, const int64_t* rowptr, const int64_t* col,
                    const double* val, const double* x, double* y) {
  return guarded(ctx, [&](Context& c) {
    require(x && y, "x and y are required");
    CsrDevice M;
    uploadCsr(c, M, n, rowptr, col, val);
    DeviceArray<double> dx, dy;
    dx.ensure(n);
    dy.ensure(n);
    CFDHIP_CHECK(hipMemcpyAsync(dx.get(), x, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    csrSpmv(c, M, dx.get(), dy.get());
    CFDHIP_CHECK(hipMemcpyAsync(y, dy.get(), n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_csr_solve(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, int64_t n,
                     const int64_t* rowptr, const int64_t* col, const double* val,
                     const double* b, const double* x0, double* x, cfdhip_solve_result* r) {
  return guarded(ctx, [&](Context& c) {
    const cfdhip_solver_settings set = checkedSettings(s);
    require(b && x, "b and x are required");
    CsrDevice M;
    uploadCsr(c, M, n, rowptr, col, val);
    DeviceArray<double> db, dx;
    db.ensure(n);
    dx.ensure(n);
    CFDHIP_CHECK(hipMemcpyAsync(db.get(), b, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (x0) {
      CFDHIP_CHECK(
          hipMemcpyAsync(dx.get(), x0, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    } else {
      CFDHIP_CHECK(hipMemsetAsync(dx.get(), 0, n * sizeof(double), c.stream));
    }
    const cfdhip_solve_result res = solveCsr(c, set, M, db.get(), dx.get());
    CFDHIP_CHECK(hipMemcpyAsync(x, dx.get(), n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
    if (r) *r = res;
  });
}

int cfdhip_pressure_spmv(cfdhip_ctx* ctx, const double* x, double* y) {
  return guarded(ctx, [&](Context& c) {
    require(x && y, "x and y are required");
    const int64_t n = c.view.n;
    DeviceArray<double> dx, dy;
    dx.ensure(n);
    dy.ensure(n);
    CFDHIP_CHECK(hipMemcpyAsync(dx.get(), x, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    pressureSpmv(c, dx.get(), dy.get());
    CFDHIP_CHECK(hipMemcpyAsync(y, dy.get(), n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFDHIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int cfdhip_pressure_spmv_device(cfdhip_ctx* ctx, const double* x, double* y, int repeat) {
  return guarded(ctx, [&](Context& c) {
    require(x && y, "x and y are required");
    require(repeat >= 1, "repeat must be >= 1");
    c.requireMesh();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c.timing) {
      CFDHIP_CHECK(hipEventCreate(&e0));
      CFDHIP_CHECK(hipEventCreate(&e1));
      CFDHIP_CHECK(hipEventRecord(e0, c.stream));
    }
    for (int k = 0; k < repeat; ++k) pressureSpmv(c, x, y);
    if (c.timing) {
      CFDHIP_CHECK(hipEventRecord(e1, c.stream));
      CFDHIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      CFDHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      c.spmvCount = repeat;
      c.spmvMeanUs = 1000.0 * ms / repeat;
      c.timedIters = 0;
    }
  });
}

int cfdhip_pressure_solve_rhs(cfdhip_ctx* ctx, const cfdhip_solver_settings* s, const double* b,
                              double* x, cfdhip_solve_result* r) {
  return guarded(ctx, [&](Context& c) {
    const cfdhip_solver_settings set = checkedSettings(s);
    c.requireMesh();
    const int64_t n = c.view.n;
    if (b) {
      CFDHIP_CHECK(hipMemcpyAsync(c.f(CFDHIP_FIELD_PC_RHS), b, n * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
    }
    const cfdhip_solve_result res =
        solvePressure(c, set, c.f(CFDHIP_FIELD_PC_RHS), c.f(CFDHIP_FIELD_PPRIME), true);
    if (x) {
      CFDHIP_CHECK(hipMemcpyAsync(x, c.f(CFDHIP_FIELD_PPRIME), n * sizeof(double),
                                  hipMemcpyDeviceToHost, c.stream));
      CFDHIP_CHECK(hipStreamSynchronize(c.stream));
    }
    if (r) *r = res;
  });
}

int cfdhip_solve_history(cfdhip_ctx* ctx, double* history, int64_t cap, int64_t* length) {
  return guarded(ctx, [&](Context& c) {
    const int64_t len = static_cast<int64_t>(c.lastHistory.size());
    if (length) *length = len;
    if (history)
      for (int64_t i = 0; i < len && i < cap; ++i) history[i] = c.lastHistory[i];
  });
}

int cfdhip_set_host_preconditioner(cfdhip_ctx* ctx, cfdhip_host_preconditioner fn, void* user) {
  return guarded(ctx, [&](Context& c) {
    c.hostPc = fn;
    c.hostPcUser = fn != nullptr ? user : nullptr;
  });
}

int cfdhip_set_timing(cfdhip_ctx* ctx, int enabled) {
  return guarded(ctx, [&](Context& c) {
    c.timing = enabled != 0;
    if (c.net) c.net->setTiming(c.timing);
  });
}

int cfdhip_preconditioner_timing(cfdhip_ctx* ctx, double* ms, int64_t* calls) {
  return guarded(ctx, [&](Context& c) {
    require(ms != nullptr && calls != nullptr, "output arrays are required");
    for (int l = 0; l <= kNumLayers; ++l) {
      ms[l] = c.net ? c.net->layerMs[l] : 0.0;
      calls[l] = c.net ? c.net->layerCalls[l] : 0;
    }
  });
}

int cfdhip_kernel_timing(cfdhip_ctx* ctx, double* out) {
  return guarded(ctx, [&](Context& c) {
    require(out != nullptr, "out is required");
    out[0] = c.spmvCount;
    out[1] = c.spmvMeanUs;
    out[2] = c.timedIters;
  });
}

}  // extern "C"

// The C layout of the settings the host mirrors (ctypes, the C++ adapter) rely on.
static_assert(sizeof(cfdhip_solver_settings) == 56, "cfdhip_solver_settings layout");
static_assert(offsetof(cfdhip_solver_settings, inner_tolerance) == 48,
              "cfdhip_solver_settings layout");
