// The batch of independent 2D p' systems (cfdhip_batch_*): the systems
// added one by one (FP64, device-resident) and the buffers of a solve;
// batch.hip solves it by Jacobi-PCG, batch_learned.hip by flexible CG with
// the learned preconditioner.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "context.hpp"
#include "unet_net.hpp"

namespace cfdhip {

// Per-case state of the batched Jacobi-PCG (batch.hip).
struct CaseState {
  double b0, rTrue, absTol, relTol, innerFloor, innerTol;
  double ib0, rz, alpha, beta, rnorm;
  int64_t iter, total, cycles, maxIter, steps, doneStep;
  int32_t phase, cur, status, done;
};

}  // namespace cfdhip

struct cfdhip_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string lastError;
  struct Sys {
    int64_t nx, ny, pin;
    // sides with a FixedValue pressure patch (bits 1 left, 2 right, 4 bottom,
    // 8 top): the learned preconditioner's odd reflection there
    int sides = 0;
    cfdhip::DeviceArray<double> diag, coeff, b;
  };
  std::vector<std::unique_ptr<Sys>> sys;
  // solve-time buffers
  cfdhip::DeviceArray<double> arena64;
  cfdhip::DeviceArray<float> arena32;
  cfdhip::DeviceArray<unsigned char> caseBuf;
  cfdhip::DeviceArray<cfdhip::CaseState> states;
  cfdhip::DeviceArray<double> partial;
  cfdhip::DeviceArray<int32_t> blockCase, blockTile, active;
  cfdhip::DeviceArray<int64_t> tile0, tiles;
  cfdhip::CaseState* hostStates = nullptr;
  int32_t* hostTable = nullptr;  // pinned staging of the block table and active list
  int64_t hostTableCap = 0;
  int lastRows = 0;
  std::vector<int64_t> xOff;  // solution of each case in arena64 (after a solve)
  // the learned preconditioner (batch_learned.hip)
  std::unique_ptr<cfdhip::Unet> net;
  cfdhip::DeviceArray<float> act, scale32;
  cfdhip::DeviceArray<unsigned char> table;
  cfdhip::DeviceArray<cfdhip::KState> kstates;
  cfdhip::DeviceArray<double> norms;
  void* hostStaging = nullptr;  // pinned staging of the learned solve's tables
  int64_t hostStagingBytes = 0;

  ~cfdhip_batch() {
    net.reset();
    if (hostStates) (void)hipHostFree(hostStates);
    if (hostTable) (void)hipHostFree(hostTable);
    if (hostStaging) (void)hipHostFree(hostStaging);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace cfdhip {

// Batched flexible CG with the learned preconditioner (batch_learned.hip).
void solveBatchLearned(cfdhip_batch& B, const cfdhip_solver_settings& set, int useGraph,
                       cfdhip_solve_result* results, double* latencyMs, double* stats);
// Batched Jacobi-PCG of every case (batch.hip); precision CFDHIP_MIXED: FP32
// inner iterations.
void solveBatchJacobi(cfdhip_batch& B, const cfdhip_solver_settings& set, int rows,
                      int useGraph, cfdhip_solve_result* results, double* latencyMs,
                      double* stats);

}  // namespace cfdhip
