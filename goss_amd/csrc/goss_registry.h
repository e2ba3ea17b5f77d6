// Registry of the model variants compiled into libgoss_hip.so.
//
// Each variant is compiled in two translation units (variant_kernels.hip with
// GOSS_VARIANT=<name> and GOSS_PREC_F64 or GOSS_PREC_F32); each fills its
// part of the variant's launcher table.
#pragma once

#include <hip/hip_runtime.h>

#include "goss_kernels.h"
#include "goss_implicit.h"
#include "goss_adaptive.h"
#include "goss_adjoint.h"

namespace goss_hip {

struct ExplicitArgs {
  void* y;                 // states, SoA [S][ld] of the precision's type
  void* ylo;               // FP32C: low parts [NCOMP][ld]
  const void* fp;          // field parameters [NFP][ld]
  const double2* steps;    // device (t, interval) pairs
  int nouter;
  const double* shared;    // host: the P model parameter values
  double ldt_0;
  double delta;
  unsigned N, ld;
  FieldOut fo;
  hipStream_t stream;
};

typedef hipError_t (*ExplicitLaunch)(const ExplicitArgs&);

struct ImplicitArgs {
  double* y;               // FP64 states [S][ld]
  const double* fp;        // FP64 field parameters [NFP][ld]
  const double2* steps;
  int nouter;
  const double* shared;
  ImplicitParams ip;
  ImplicitBuffers B;
  unsigned N, ld;
  FieldOut fo;
  hipStream_t stream;
};

typedef hipError_t (*ImplicitLaunch)(const ImplicitArgs&);

struct AdaptiveArgs {
  void* y;                 // states [S][ld] of the precision's type
  const void* fp;          // field parameters [NFP][ld]
  const double2* steps;
  int nouter;
  const double* shared;
  AdaptiveParams ap;
  ImplicitBuffers B;       // per-node controller state (ds, is, error)
  unsigned N, ld;
  FieldOut fo;
  hipStream_t stream;
};

typedef hipError_t (*AdaptiveLaunch)(const AdaptiveArgs&);

// Adjoint of explicit stepping (ExplicitEuler, RK2, RK4, RL1, RL2, GRL1, GRL2): record (forward
// with checkpoints and samples) and backward (the reverse sweep).  Precision
// FP64 or FP32.
struct AdjointArgs {
  void* y;                 // states [S][ld] (record: advanced in place)
  const void* fp;          // field parameters [NFP][ld]
  const double2* steps;    // record: this launch's steps; backward: all steps
  int step0, nouter;       // record: global index of steps[0], steps in this launch
  const double* shared;
  double delta;
  unsigned N, ld;
  AdjointShape shape;
  void* ckpt;              // [n_ckpt][S][ld]
  double* samples;         // record: [n_samples][N]
  const double* grad_samples;  // backward: [n_samples][N]
  void* seg;               // backward: [ckpt_every][S][ld]
  void* seg2;              // backward, ckpt_every > 1: a second [ckpt_every][S][ld] buffer (the pipeline's other half)
  void* tape;              // backward: FP64 tape [vjp_tape_slots][ld], then work rows [vjp_work_rows][ld]
  void* lam;               // backward: adjoint states between segments [S][ld]
  double* pbar;            // backward: parameter gradients between segments [NFP][ld]
  double* grad_y0;         // backward: [N][S]
  double* grad_fp;         // backward: [N][NFP]
  hipStream_t stream;
  // backward, ckpt_every > 1: the segment recomputes run on `aux`, one segment ahead of the
  // reverse sweeps on `stream`.  ev_start: recorded on `stream` before the call; ev_rc[b] /
  // ev_rev[b]: recompute into / reverse sweep out of segment buffer b done.
  hipStream_t aux;
  hipEvent_t ev_start;
  hipEvent_t ev_rc[2];
  hipEvent_t ev_rev[2];
};

typedef hipError_t (*AdjointLaunch)(const AdjointArgs&);

// Model evaluation at n points, one per lane (ODE.h:55-77): kind 0 eval
// (out0 = values [n][S]), 1 linearized_eval (out0 = linear, out1 = rhs, both
// [n][S]), 2 the forward-difference Jacobian of ODE::compute_jacobian
// (ODE.cpp:47-67; out0 = [n][S*S] row-major), 3 eval_monitored (out0 =
// [n][num_monitored], the intermediates listed in `monitored`).  Array
// pointers are device pointers; params holds the P model parameters (host),
// field parameters included.
constexpr int kMaxMonitored = 64;
struct MonitoredList {
  int n;
  int index[kMaxMonitored];  // into the variant's intermediates
};
struct ModelEvalArgs {
  int kind;
  int only_linear;
  MonitoredList monitored;
  const double* states;  // [n][S]
  const double* times;   // [n]
  unsigned long n;
  const double* params;  // host, P values
  double* out0;
  double* out1;
  hipStream_t stream;
};

typedef hipError_t (*ModelEvalLaunch)(const ModelEvalArgs&);

// Solvers with a device adjoint (every explicit fixed-step solver), in the order of the adjoint tables.
constexpr int kAdjointSolvers[] = {kGRL1, kRL1, kExplicitEuler, kRK4, kRK2, kRL2, kGRL2};
constexpr int kNumAdjointSolvers = 7;
inline int adjoint_index(const int solver)
{
  for (int k = 0; k < kNumAdjointSolvers; ++k)
    if (kAdjointSolvers[k] == solver) return k;
  return -1;
}

struct VariantInfo {
  const char* name;
  int S, P, NFS, NFP, NCOMP;
  int jac_nnz, lu_nnz;  // structure of the Jacobian and of its LU factors (implicit solvers)
  int vjp_tape_slots[kNumAdjointSolvers];  // slots of the FP64 adjoint's per-node tape
  int vjp_work_rows[kNumAdjointSolvers];   // per-node work rows of the step adjoint (precision R)
  const int* fsi;
  const int* fpi;
  const double* defaults;
  const double* ic;
  const unsigned char* linear_terms;
  const unsigned char* differential_states;  // mass-matrix diagonal (0: algebraic state of a DAE)
  int is_dae;
  const char* const* state_names;
  const char* const* parameter_names;
  const char* const* field_state_names;
  const char* const* field_parameter_names;
  ExplicitLaunch explicit_launch[kNumSolvers][kNumPrecisions];
  ImplicitLaunch implicit_launch[kNumImplicitSlots][kNumJacobianModes];  // [implicit_slot(solver)][JacobianMode]
  AdaptiveLaunch rkf32_launch[kNumPrecisions];  // FP64 and FP32
  AdjointLaunch adjoint_record[kNumAdjointSolvers][kNumPrecisions];  // FP64 and FP32
  AdjointLaunch adjoint_backward[kNumAdjointSolvers][kNumPrecisions];
  ModelEvalLaunch model_eval;  // FP64
  int num_intermediates;       // named intermediates (the values eval_monitored can select)
  const char* const* intermediate_names;
};

// The variants, in registration order (compiled in, then loaded at run time).
int variant_count();
VariantInfo* variant_at(int index);

}  // namespace goss_hip
