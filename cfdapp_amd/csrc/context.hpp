// The context: device, stream and the resident state of one case.
#pragma once

#include <memory>

#include "common.hpp"

namespace cfdhip {

class Unet;

// Device-resident state of one Krylov solve. Written by the finalising block
// of each reduction kernel; the host only reads it every check_interval
// iterations (and at the end) through a pinned mirror.
struct KState {
  double b0, absTol, relTol;
  double rz, pAp, alpha, beta, rnorm;
  // BiCGSTAB recurrence (src/algebra/BiCGSTAB.cpp:48-275)
  double rho, rhoOld, omega, rHatNorm, restartResidual, vnorm, snorm;
  double tDotT;
  int64_t iter, maxIter, completed, restarts, histLen;
  int32_t status;  // CFDHIP_CONVERGED ... ; -1 while running
  int32_t done;    // 1 once status is final
  int32_t restart; // 1: the host must run the true-residual restart
  int32_t sConverged;
  // flexible CG with the learned preconditioner: stagnation watch
  double best;         // smallest residual norm so far
  int64_t stagnant;    // iterations since it last decreased
  int64_t window;      // stagnant iterations that end the solve (0: never)
  int32_t stagnated;   // 1 when the window ran out
  // mixed-precision PCG with residual replacement: replaceRatio > 0 stops the
  // FP32 recurrence (done, replace = 1) once its residual has fallen below
  // replaceRatio times its largest value since the last FP64 replacement
  // (rmax), or below the target; the host then recomputes the residual in FP64
  int32_t replace;
  int32_t restartOnReplace;  // 1: a replacement restarts the direction (p = u)
  double replaceRatio, rmax;
  int64_t dirStart;          // iteration count at which the current direction sequence began
};

// Work vectors (allocated once per size, reused across solves).
struct KrylovWork {
  DeviceArray<double> r, p, Ap, z, x, b, dinv, rhat, v, s, t, phat, shat;
  DeviceArray<float> r32, x32, dinv32, w32, dis32;
  DeviceArray<float> scaleTmp[2];  // smoothed preconditioner scaling (learned FCG)
  DeviceArray<double> d8[8];  // single-kernel PCG buffers (FP64)
  DeviceArray<float> f8[8];   // single-kernel PCG buffers (FP32)
  DeviceArray<double> partial;
  DeviceArray<KState> state;
  DeviceArray<double> hist;
  DeviceArray<double> scal;  // scalar results of one-off reductions
  KState* hostState = nullptr;  // pinned mirror
  KState* hostState2 = nullptr; // a second one: two chunks of a solve in flight
  // device bytes held by the work vectors and reduction buffers above
  int64_t bytes() const;
};

struct CsrDevice {
  DeviceArray<int64_t> rp, col;
  DeviceArray<double> val;
  int64_t n = 0;
};

// Cell gradients on a face-list mesh (fgrad.hip): what depends only on the
// geometry is computed once on the host, in the reference's expressions; the
// field arithmetic runs on the device.
struct FaceGradPlan {
  bool built = false;
  bool transfer = false;  // some boundary fit line meets the boundary off the face centroid
  int64_t nSkew = 0, nClaims = 0;
  DeviceArray<double> wts;          // 2 per face: owner / neighbour centroid to face distance
  DeviceArray<int64_t> skewFace;    // internal faces with a non-zero skew vector
  DeviceArray<double> skewGeo;      // 4 per skewed face: t, skew vector
  DeviceArray<int32_t> slotClaim;   // per cell-face slot: the claim on that face, or -1
  DeviceArray<int64_t> claimFaces;  // 3 per claim: boundary face, opposite face, far cell
  DeviceArray<double> claimGeo;     // 6 per claim: far distance, w, back distance, offset
  // oblique Neumann-type boundary faces, for the conditions last used
  std::vector<int32_t> oblKinds;
  int64_t nObl = 0;
  DeviceArray<int64_t> oblFaces;
  DeviceArray<double> oblGeo;       // 8 per face: oblique (0/1), dn, unit normal, tangential offset
};

// A general face-list mesh (fmesh.hip) and its p' system.
struct FaceMesh {
  int dim = 2;
  int64_t n = 0, nf = 0;
  bool open = false;  // a FixedValue pressure face exists: no pinned cell
  std::vector<int64_t> hOwner, hNeigh, hRp, hCol, hCfOff, hCf;
  std::vector<int32_t> hFixed;
  std::vector<double> hFc, hSf, hCc;
  DeviceArray<int64_t> owner, neigh, cfOff, cf, cfSorted, slotOff, slotCode;
  DeviceArray<double> fc, sf, cc, vol;
  DeviceArray<int> fixed;
  DeviceArray<double> flux, du, dv, dw, rhoF, grad, extra, coeff, expl, rhs, x;
  CsrDevice A;
  int64_t planPin = -2;
  bool planExtra = false, assembled = false, hasExplicit = false, solved = false;
  bool hasVolumes = false;
  // the last assembly's inputs beyond the uploaded arrays (corrector passes)
  double rho = 1.0;
  int nonOrth = 1;
  bool hasRhoF = false, hasDw = false, hasExtra = false;
  int64_t ref = 0;
  // gradients
  FaceGradPlan gp;
  DeviceArray<double> fv, gradA, gradB, field, bcVal, gg;
  DeviceArray<int32_t> bcKind;
};

// inputs of the compressible p' assembly, kept on the device
struct CompressibleWork {
  DeviceArray<double> rhoF, pAbs, T, extra;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string lastError;

  bool hasMesh = false;
  View view{};
  DeviceArray<double> tables[3];
  double delta[3] = {0.0, 0.0, 0.0};  // uniform cell size per axis
  double rho = 1.0, mu = 0.0;
  bool fluidSet = false;
  bool bcSet[6] = {false, false, false, false, false, false};
  int64_t pin = -1;         // p' reference cell of the last assembly (-1: none)
  bool pressureAssembled = false, momentumAssembled = false;

  DeviceArray<double> field[CFDHIP_FIELD_COUNT];
  DeviceArray<double> momCoef;  // (2*dim+1) * n, CSR column order
  DeviceArray<unsigned> flags;  // [0] non-finite counter
  DeviceArray<double> stage;    // a caller's phi_old / correction outputs on their way to the host

  KrylovWork kw;
  CompressibleWork comp;
  std::unique_ptr<FaceMesh> fmesh;
  std::vector<double> lastHistory;
  std::unique_ptr<Unet> net;  // learned p' preconditioner (weights loaded by the caller)
  // CFDHIP_PREC_HOST: the caller's host preconditioner and its staging vectors
  cfdhip_host_preconditioner hostPc = nullptr;
  void* hostPcUser = nullptr;
  std::vector<double> hostPcIn, hostPcOut;
  int64_t lastFallback = 0;

  // inside the resident SIMPLE loop: its assembly checks already cover the
  // inputs of its solves (no input scan), and it reads neither the true
  // residual nor the history of a solve (neither is computed or copied)
  bool inLoop = false;

  // timing of the pressure SpMV kernel
  bool timing = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double spmvCount = 0, spmvMeanUs = 0, timedIters = 0;

  ~Context();
  int64_t cellCount() const { return view.n; }
  int64_t faceCount() const { return view.nf; }
  double* f(int id) const { return field[id].get(); }
  void requireMesh() const { require(hasMesh, "no mesh: call cfdhip_mesh_cartesian first"); }
  void ensureFields();
};

// plan.hip
void buildCartesianPlan(Context& ctx, int dim, int64_t nx, int64_t ny, int64_t nz, double lx,
                        double ly, double lz);
void setBoundary(Context& ctx, int patch, int vkind, const double* v3, int pkind, double pvalue);
void exportFaces(Context& ctx, int64_t* owner, int64_t* neighbor);

// assembly.hip
void launchGradient(Context& ctx, const double* phi, int bcKind, double* gx, double* gy, double* gz);
void launchMassFlux(Context& ctx, const double* u, const double* v, const double* w, double* F);
// dt > 0: the transient momentum system (TransientMomentum.cpp), alpha 1
// oldComponent >= 0: oldValues (device, n) replaces that component's resident
// velocity as the relaxation's phi_old
void launchMomentum(Context& ctx, double alpha, double dt = 0.0, int oldComponent = -1,
                    const double* oldValues = nullptr);
void launchResponse(Context& ctx);
void launchPredictedFlux(Context& ctx, bool rhieChow, double alpha);
// rhoF: per-face density (nullptr: the fluid's); extra: an additional diagonal
// added after the face terms (nullptr: none)
void launchPressureAssembly(Context& ctx, int64_t ref, const double* rhoF = nullptr,
                            const double* extra = nullptr);
// assembleCompressiblePressureCorrection: per-face density (faces), absolute
// pressure and temperature (cells), host or device pointers
void launchCompressiblePressureAssembly(Context& ctx, int64_t ref, const double* rhoF,
                                        const double* pAbs, const double* T, double R,
                                        double dtau);
void launchUpdatePressure(Context& ctx, double alpha);
void launchCorrect(Context& ctx);
// the two halves of launchCorrect; a non-null destination receives the result
// (velocity: 3 component arrays of n) instead of U/V/W or FLUX
void launchCorrectVelocity(Context& ctx, double* const dest[3] = nullptr);
void launchCorrectFlux(Context& ctx, double* dest = nullptr);
void launchContinuity(Context& ctx, double* out);
// max over cells of the Courant number of FLUX (cfl) or of |IMBALANCE|
double launchMaxCell(Context& ctx, bool cfl, double dt);
int64_t exportPressureCsr(Context& ctx, int64_t* rp, int64_t* col, double* val, int64_t cap);
int64_t exportMomentumCsr(Context& ctx, int64_t* rp, int64_t* col, double* val, int64_t cap);
unsigned countNonFinite(Context& ctx, const double* const* arrays, const int64_t* counts, int k);
// fmesh.hip: the p' system on a general face-list mesh
void fmeshSet(Context& ctx, int dim, int64_t n, int64_t nf, const int64_t* owner,
              const int64_t* neighbor, const double* faceCentroid, const double* areaVector,
              const double* cellCentroid, const int64_t* cellFaceOffsets,
              const int64_t* cellFaces, const int32_t* fixedValue);
void fmeshAssemblePressure(Context& ctx, const double* fluxStar, const double* faceDensity,
                           double rho, const double* du, const double* dv, const double* dw,
                           int nonOrthogonal, const double* gradPrev, const double* extraDiag,
                           int64_t ref);
int64_t fmeshSystem(Context& ctx, int64_t* rowptr, int64_t* col, double* val, int64_t cap,
                    double* rhs, double* faceCoeff, double* explicitFlux);
cfdhip_solve_result fmeshSolve(Context& ctx, const cfdhip_solver_settings& s, double* x);
void fmeshCorrectFlux(Context& ctx, const double* pprime, double* flux);
// the assembly again, on the last call's inputs, with the explicit term of the
// device gradient `gradDev` (3 per cell) or none
void fmeshReassemble(Context& ctx, const double* gradDev);
// fgrad.hip: cell gradients on the face-list mesh (device arrays): field[n],
// per-face condition kind (1 FixedValue, 0 FixedGradient) and value, scheme
// 0 Green-Gauss / 1 least squares; out[3 n]
void fmeshSetVolumes(Context& ctx, const double* cellVolume);
void fmeshGradient(Context& ctx, const double* field, const int32_t* kind, const double* value,
                   int scheme, double* out);
// the p' conditions of the explicit term into FaceMesh::bcKind / bcVal
void fmeshPprimeConditions(Context& ctx, FaceMesh& m);
// a corrector pass's assembly: the explicit term from the gradient (scheme)
// of the previous p' (device, n values)
void fmeshAssemblePass(Context& ctx, const double* previousDev, int scheme);

// krylov.hip
enum class OpKind { Pressure, Stencil, Csr };
cfdhip_solve_result solvePressure(Context& ctx, const cfdhip_solver_settings& s, const double* b,
                                  double* x, bool zeroGuess);
cfdhip_solve_result solveStencil(Context& ctx, const cfdhip_solver_settings& s, const double* b,
                                 double* x);
cfdhip_solve_result solveCsr(Context& ctx, const cfdhip_solver_settings& s, const CsrDevice& A,
                             const double* b, double* x);
void pressureSpmv(Context& ctx, const double* x, double* y);
void csrSpmv(Context& ctx, const CsrDevice& A, const double* x, double* y);

// cg1.hip: single-kernel 2D Jacobi-PCG iteration (T = double / float);
// work = r0 r1 w0 w1 p0 p1 s0 s1
// Optional copies of the iterate x every `stride` iterations (training data).
struct Snapshots {
  double* buf = nullptr;  // cap x n
  int64_t stride = 0, cap = 0, taken = 0;
};
template <class T>
cfdhip_solve_result singleKernelPcg2d(Context& ctx, const cfdhip_solver_settings& set,
                                      const T* diag, const T* coeff, const T* b, T* x,
                                      T* const work[8], Snapshots* snap = nullptr);
// The same FP64 solve on caller-provided buffers (partials, device state, its
// host mirror, history); events == false: no HIP events, graphs or history
// copy (the no-allocation entry point).
struct PcgBuffers {
  double* partial;
  int64_t partialCap;
  KState* state;
  KState* host;
  double* hist;
  int64_t histCap;
  bool events;
  // a second device state: lets small grids run the fused iteration kernel
  // (no finalize launch between iterations); null keeps one finalize per
  // iteration
  KState* state2 = nullptr;
  // a second pinned host mirror: lets the fused chunks run pipelined (the
  // next chunk queued before the host reads the last one's state)
  KState* host2 = nullptr;
};
cfdhip_solve_result pcg2dOnBuffers(Context& ctx, const cfdhip_solver_settings& set,
                                   const double* diag, const double* coeff, const double* b,
                                   double* x, double* const work[8], const PcgBuffers& buf,
                                   Snapshots* snap = nullptr);
// cg1.hip: the mixed-precision solve (FP32 iteration, FP64 residual replacement)
cfdhip_solve_result mixedPcg2d(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                               double* x);
// pcg2d.hip
void castToFloat(Context& ctx, int64_t n, const double* src, float* dst);
// cg3.hip: 3D Jacobi-PCG (work = r s w p u) and its mixed-precision refinement
template <class T>
cfdhip_solve_result pcg3d(Context& ctx, const cfdhip_solver_settings& set, const T* diag,
                          const T* coeff, const T* b, T* x, T* const work[5]);
cfdhip_solve_result mixedPcg3d(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                               double* x);
// fcg.hip: flexible CG with the learned preconditioner (FP32 inner, FP64
// refinement, Jacobi-PCG fallback on stagnation)
cfdhip_solve_result learnedFcg2d(Context& ctx, const cfdhip_solver_settings& set, const double* b,
                                 double* x);
void loadLearnedWeights(Context& ctx, const float* weights);
// the learned preconditioner's scaling S of a Neumann-padded network on a 2D
// p' system (nx x ny, face couplings at x-faces then y-faces from offY):
// S^-2 the interior-equivalent diagonal geometric-mean filtered, D^-1/2 at
// the pinned cell; tmp0, tmp1: two work images
void learnedScaling(hipStream_t stream, int64_t nx, int64_t ny, int64_t offY, int64_t pin,
                    const double* diag, const double* coeff, float* tmp0, float* tmp1,
                    float* out);

// simple.hip
// piso.hip: PISO time steps on the resident state; meta 6 per step
int64_t pisoRun(Context& ctx, double dt, int64_t steps, int64_t ref,
                const cfdhip_solver_settings& mom, const cfdhip_solver_settings& pres,
                double* meta);
void simpleRun(Context& ctx, const double* settings, const cfdhip_solver_settings& mom,
               const cfdhip_solver_settings& pres, double* meta, double* history, int64_t cap);

}  // namespace cfdhip
