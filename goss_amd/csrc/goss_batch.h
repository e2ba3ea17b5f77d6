// Mixed-population batch: node populations of different cell models and
// solvers advanced together by one persistent kernel launch per block of
// time steps (goss_batch.hip), the launches of one forward call captured in
// a hipGraph (goss_hip_api.hip).
//
// The reference advances one population per ODESystemSolver::forward call
// (ODESystemSolver.cpp:67-79); a batch runs several such systems at once.
// Each node executes exactly the per-node code of its population's own
// kernel (explicit_node / rkf32_node), so a node's result does not depend on
// whether it was advanced alone or in a batch.
#pragma once

#include <hip/hip_runtime.h>

#include "goss_kernels.h"
#include "goss_implicit.h"
#include "goss_adaptive.h"

namespace goss_hip {

constexpr int kMaxBatchPops = 16;

// One population of a batch, in device memory.  Chunks of blockDim.x nodes
// [chunk0, chunk0 + ceil(N / blockDim.x)) of the launch belong to it.
struct BatchPop {
  int kind;               // row of the fused kernel's kind table
  int nouter;             // outer steps per launch (time block)
  unsigned N, ld, chunk0;
  void* y;                // states [S][ld], FP32
  const void* fp;         // field parameters [NFP][ld], FP32
  const double2* steps;   // all (t, interval) pairs of the call; block b uses steps + b * nouter
  unsigned long sp_off;   // byte offset of the population's SharedParams in the parameter blob
  double ldt_0, delta;    // explicit solvers
  AdaptiveParams ap;      // RKF32
  ImplicitBuffers B;      // RKF32: per-node controller state
};

// A (model variant, solver) pair the fused kernel is compiled for (FP32).
struct BatchKind {
  const char* variant;
  int solver;
  unsigned long sp_bytes;                          // sizeof its SharedParams
  void (*make_params)(const double* shared, void* out);  // host: fill the SharedParams
};

const BatchKind* batch_kinds(int* count);

// One launch: time block `tblock` of every population.
hipError_t launch_batch(const BatchPop* pops, int npop, unsigned nchunks, int tblock, const char* params,
                        unsigned grid, unsigned block, hipStream_t stream);

}  // namespace goss_hip
