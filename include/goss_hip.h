/* goss_hip.h -- C ABI of the MI355X implementation of goss's
 * ODESystemSolver::forward hot path.
 *
 * One goss_hip_system advances one generated cell-model variant over many
 * nodes on one GPU.  Every entry point replaces a member of the reference's
 * goss::ODESystemSolver (cpp/goss/ODESystemSolver.h:38-186) or of the solver /
 * model objects it owns; the citation is given beside each declaration.
 *
 * Conventions
 *  - All functions return 0 on success and a negative GOSS_HIP_E* code on
 *    failure; goss_hip_last_error() returns the message of the last failure
 *    on the calling thread.  Where the reference throws std::runtime_error
 *    the message text is the reference's.
 *  - Host buffers are caller-owned, host-resident `double*` of the sizes the
 *    reference uses (num_nodes * num_field_states, ...); as in the reference,
 *    no sizes are passed.  The *_device variants take device pointers.
 *  - "tangled" storage: element (node, i) at node * F + i; untangled at
 *    num_nodes * i + node (ODESystemSolver.h:244).
 *  - A system owns one HIP stream; calls on one system are not re-entrant.
 */
#ifndef GOSS_HIP_H
#define GOSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GOSS_HIP_OK 0
#define GOSS_HIP_EINVAL -1   /* bad argument (unknown name, index, size) */
#define GOSS_HIP_EDEVICE -2  /* HIP runtime error */
#define GOSS_HIP_ESOLVER -3  /* numerical failure reported by a solver */
#define GOSS_HIP_EUNSUPPORTED -4

/* Solvers (python/goss/solvers.py:423-436 names). */
typedef enum goss_hip_solver {
  GOSS_HIP_EXPLICIT_EULER = 0, /* ExplicitEuler.cpp:64-87 */
  GOSS_HIP_RK2 = 1,            /* RK2.cpp:62-92 */
  GOSS_HIP_RK4 = 2,            /* RK4.cpp:68-102 */
  GOSS_HIP_RL1 = 3,            /* RL1.cpp:59-79 */
  GOSS_HIP_RL2 = 4,            /* RL2.cpp:64-87 */
  GOSS_HIP_GRL1 = 5,           /* GRL1.cpp:61-82 */
  GOSS_HIP_GRL2 = 6,           /* GRL2.cpp:63-86 */
  GOSS_HIP_IMPLICIT_EULER = 7, /* ImplicitEuler.cpp:57-152 */
  GOSS_HIP_ESDIRK23A = 8,      /* ESDIRK23a.cpp:280-316 */
  GOSS_HIP_RKF32 = 9,          /* RKF32.cpp:105-222 */
  GOSS_HIP_THETA = 10,         /* ThetaSolver.cpp:48-155 */
  GOSS_HIP_BASIC_IMPLICIT_EULER = 11, /* BasicImplicitEuler.cpp:62-181 (FP64) */
  GOSS_HIP_ESDIRK4O32 = 12     /* ESDIRK4O32.cpp:92-298 (FP64) */
} goss_hip_solver;

/* Arithmetic of the per-node work.  FP64 is the reference's arithmetic.
 * FP32C computes in FP32 and carries the states whose derivative is not
 * affine in themselves (linear_terms == 0) as FP32 hi + lo pairs.
 * The implicit solvers run in FP64; RKF32 in FP64 or FP32 (its step
 * sizes and local time stay FP64). */
typedef enum goss_hip_precision {
  GOSS_HIP_FP64 = 0,
  GOSS_HIP_FP32 = 1,
  GOSS_HIP_FP32C = 2
} goss_hip_precision;

/* Public solver tunables: ODESolver::_ldt (ODESolver.h:119-128),
 * GRL1::delta (GRL1.h:39), ImplicitODESolver (ImplicitODESolver.h:36-41),
 * ImplicitEuler (ImplicitEuler.h:36-37), ESDIRK23a (ESDIRK23a.h:37-38),
 * AdaptiveExplicitSolver tolerances (AdaptiveExplicitSolver.cpp:52-53). */
typedef struct goss_hip_solver_params {
  double internal_time_step; /* <= 0: one step per forward interval */
  double delta;
  double eta_0;
  double kappa;
  double relative_tolerance;
  double max_relative_previous_residual;
  double min_dt;
  double atol;
  double rtol;
  int max_iterations;
  int always_recompute_jacobian;
  int num_refinements_without_always_recomputing_jacobian;
  /* Implicit solvers: how the Jacobian is formed (GOSS_HIP_JACOBIAN_*).  FD is
   * the reference's default (ODE.cpp:47-67); ANALYTIC the generated analytic
   * Jacobian; FD_FP32 (ImplicitEuler, ESDIRK23a, ThetaSolver) has no
   * counterpart in the reference: forward differences of the model's FP32
   * right-hand side.  The Newton iteration, its residual, the LU and the
   * substitutions are FP64 in all. */
  int jacobian;
  double theta; /* ThetaSolver::theta (ThetaSolver.h:37) */
} goss_hip_solver_params;

#define GOSS_HIP_JACOBIAN_FD 0      /* forward differences of the FP64 right-hand side */
#define GOSS_HIP_JACOBIAN_FD_FP32 1 /* forward differences of the FP32 right-hand side */
/* The model's analytic Jacobian: ODE::compute_jacobian (ODE.h:71) as overridden
 * by a plugin generated with generate_jacobian (codegeneration.py:186, 622-641).
 * The generated entries are evaluated and stored over the Jacobian's structure.
 * Every implicit solver. */
#define GOSS_HIP_JACOBIAN_ANALYTIC 2

/* Sizes and names of a generated model variant
 * (ParameterizedODE.h:56-156 getters). */
typedef struct goss_hip_model_info {
  int num_states;
  int num_parameters;
  int num_field_states;
  int num_field_parameters;
  const char* const* state_names;
  const char* const* parameter_names;
  const char* const* field_state_names;
  const char* const* field_parameter_names;
  const int* field_state_indices;   /* get_field_state_indices() */
  const int* field_parameter_indices;
  const double* default_parameters; /* generated defaults */
  const double* initial_conditions; /* ODE::get_ic */
  const unsigned char* linear_terms;/* ODE::linear_term */
  const unsigned char* differential_states; /* ODE::differential_states (ODE.h:85-89): 0 for an
                                               algebraic state of a DAE */
  int is_dae;                       /* ODE::is_dae (ODE.h:80-83) */
} goss_hip_model_info;

typedef struct goss_hip_system goss_hip_system;

const char* goss_hip_last_error(void);

/* Number of GPUs visible (0 on a host without one). */
int goss_hip_device_count(int* count);

/* Model variants compiled into this library.  Replaces the model plugin
 * factory create_<Name>() (python/goss/codegeneration.py:109-112) and the
 * metadata getters of ParameterizedODE (ParameterizedODE.h:56-156). */
int goss_hip_model_count(void);
const char* goss_hip_model_name(int index);
int goss_hip_model_info_get(const char* variant, goss_hip_model_info* out);
/* Adds the model variant compiled into library_path (a model library built
 * at run time from a .ode file by goss_amd.jit; the counterpart of the
 * reference's JIT module, python/goss/compilemodule.py:19-58, whose plugin
 * pointer Python hands to _gosscpp.make_ode, wrapper.cpp:259-266).  The
 * library stays loaded until the process ends. */
int goss_hip_model_load(const char* library_path, const char* variant);

/* Model methods at n points, evaluated on the device (host arrays in and
 * out; params: the P model parameter values, NULL for the defaults):
 * ODE::eval (ODE.h:55), ODE::linearized_eval (ODE.h:61; with only_linear the
 * entries of the non-linear states keep the values passed in `linear`) and
 * ODE::compute_jacobian, the forward-difference Jacobian (ODE.cpp:47-67,
 * jac[p][i*S + j] = d values[i] / d states[j]).  states [n][S], times [n]. */
int goss_hip_model_eval(const char* variant, const double* params, const double* states, const double* times,
                        uint64_t n, double* values);
int goss_hip_model_linearized_eval(const char* variant, const double* params, const double* states,
                                   const double* times, uint64_t n, int only_linear, double* linear, double* rhs);
int goss_hip_model_compute_jacobian(const char* variant, const double* params, const double* states,
                                    const double* times, uint64_t n, double* jac);
/* ODE::compute_jacobian (ODE.h:71) of a plugin generated with generate_jacobian
 * (codegeneration.py:622-641): the generated analytic Jacobian, [n][S][S]. */
int goss_hip_model_compute_analytic_jacobian(const char* variant, const double* params, const double* states,
                                             const double* times, uint64_t n, double* jac);
/* The model's named intermediates (the values eval_monitored can select),
 * and ParameterizedODE::eval_monitored (ParameterizedODE.h:90-105) of the
 * intermediates listed in `monitored` (indices into those names; at most 64)
 * at n points: values [n][num_monitored]. */
int goss_hip_model_intermediates(const char* variant, int* count, const char* const** names);
int goss_hip_model_eval_monitored(const char* variant, const double* params, int num_monitored,
                                  const int* monitored, const double* states, const double* times, uint64_t n,
                                  double* values);
/* ODE::lu_factorize (ODE.cpp:69-100) and ODE::forward_backward_subst
 * (ODE.cpp:102-132) of n row-major num_states x num_states matrices, in place
 * / into x [n][num_states], on the device. */
int goss_hip_lu_factorize(int num_states, uint64_t n, double* mats);
int goss_hip_forward_backward_subst(int num_states, uint64_t n, const double* mats, const double* b, double* x);

/* Default tunables of a solver (GRL1.h:39, ImplicitODESolver.h:36-41, ...). */
int goss_hip_solver_params_default(int solver, goss_hip_solver_params* out);
/* Whether a solver is adaptive (ODESolver::is_adaptive, ODESolver.h:131-134). */
int goss_hip_solver_is_adaptive(int solver);

/* ODESystemSolver(num_nodes, solver, ode) (ODESystemSolver.cpp:45-59):
 * attaches the model and resets states and field parameters to defaults. */
int goss_hip_system_create(const char* variant, int solver, const goss_hip_solver_params* params, int precision,
                           uint32_t num_nodes, goss_hip_system** out);
int goss_hip_system_destroy(goss_hip_system* sys);

/* ODESystemSolver::forward(t, interval) (ODESystemSolver.cpp:67-79). */
int goss_hip_system_forward(goss_hip_system* sys, double t, double interval);
/* forward(t[i], t[i+1]-t[i]) for i = 0..n-2 fused into few launches: the
 * loop of ODESystemSolver::solve (ODESystemSolver.cpp:81-99) without the
 * per-step copy-out. */
int goss_hip_system_forward_times(goss_hip_system* sys, const double* t, uint64_t n);
/* No counterpart in the reference.  goss_hip_system_forward_times without any
 * device or pinned-host allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the time points and the (t, interval) pairs the kernels read are
 * staged in a fixed device array of the library (768 KiB per GPU, held until
 * the process ends) instead of the system's growable step buffers.  At most
 * 32,768 time points (32,767 steps) per call; more gives GOSS_HIP_EINVAL.
 * Results equal those of goss_hip_system_forward_times bit for bit.  The
 * array is shared by every system on one GPU: calls that overlap on one GPU
 * race (calls on one system are ordered by its stream; before calling this on
 * another system of the same GPU, synchronize the previous one). */
int goss_hip_system_forward_times_noalloc(goss_hip_system* sys, const double* t, uint64_t n);
/* ODESystemSolver::solve(field_states, t, n, tangled) (ODESystemSolver.cpp:81-99):
 * rows 1..n-1 of field_states[n][...] receive the field states after each
 * step; row 0 is left to the caller (python/goss/systemsolver.py:130-132).
 * Rows travel in chunks of 64 MiB: into pageable host memory through two
 * pinned bounce buffers of the system (held until it is destroyed), copied
 * out by up to 8 host threads while the next chunk steps. */
int goss_hip_system_solve(goss_hip_system* sys, double* field_states, const double* t, uint64_t n, int tangled);
/* ODESystemSolver::solve without waiting.  The reference's solve is
 * synchronous and writes to host memory; here field_states is pinned host
 * memory (goss_hip_host_alloc, or any allocation HIP has pinned) or device
 * memory, and the call returns once the steps and the copies of the rows are
 * enqueued.  Pinned destinations get the rows in chunks: the copy of one chunk
 * (on a second stream of the system) runs while the steps of the next chunk
 * run, so the host link and the GPU work together.  Device destinations are
 * written by the kernels directly.  The times t are read during the call.
 * Work enqueued on the system afterwards, and goss_hip_system_synchronize,
 * follow the copies.  goss_hip_system_solve takes the same path when
 * field_states is pinned or device memory, and waits. */
int goss_hip_system_solve_async(goss_hip_system* sys, double* field_states, const double* t, uint64_t n,
                                int tangled);
/* Pinned (page-locked) host memory for goss_hip_system_solve_async. */
int goss_hip_host_alloc(void** ptr, size_t bytes);
int goss_hip_host_free(void* ptr);

/* The solver's public tunables after creation (GRL1::delta, ImplicitODESolver's
 * eta_0 ..., ODESolver::set_internal_time_step; ODESolver.h:119-128): the
 * reference's system holds the solver object, so changes made to it apply to
 * the next forward; here they apply from the next launch.  Per-node solver
 * state (Newton eta, step-size history) is kept. */
int goss_hip_system_set_solver_params(goss_hip_system* sys, const goss_hip_solver_params* params);
int goss_hip_system_get_solver_params(goss_hip_system* sys, goss_hip_solver_params* params);

/* get/set_field_states (ODESystemSolver.cpp:134-160). */
int goss_hip_system_get_field_states(goss_hip_system* sys, double* out, int tangled);
int goss_hip_system_set_field_states(goss_hip_system* sys, const double* in, int tangled);
/* get/set_field_state_components (ODESystemSolver.cpp:101-132). */
int goss_hip_system_get_field_state_components(goss_hip_system* sys, double* out, uint32_t num_components,
                                               const uint32_t* components, int tangled);
int goss_hip_system_set_field_state_components(goss_hip_system* sys, const double* in, uint32_t num_components,
                                               const uint32_t* components, int tangled);
/* get/set_field_parameters (ODESystemSolver.cpp:162-190). */
int goss_hip_system_get_field_parameters(goss_hip_system* sys, double* out, int tangled);
int goss_hip_system_set_field_parameters(goss_hip_system* sys, const double* in, int tangled);
/* Device-pointer variants of the field-state accessors (no PCIe copy). */
int goss_hip_system_get_field_states_device(goss_hip_system* sys, double* dev_out, int tangled);
int goss_hip_system_set_field_states_device(goss_hip_system* sys, const double* dev_in, int tangled);
/* All states, node-major [num_nodes][num_states] FP64, to and from device
 * memory (no PCIe copy): the saved copies of DOLFINODESystemSolver.save_states
 * and restore_states (python/goss/dolfinutils.py:782-810).  For the FP32C
 * precision the FP64 value is the sum of the two parts, and setting it back
 * restores both exactly. */
int goss_hip_system_get_states_device(goss_hip_system* sys, double* dev_out);
int goss_hip_system_set_states_device(goss_hip_system* sys, const double* dev_in);
/* The operator-splitting exchange of DOLFINODESystemSolver (to_field_states,
 * from_field_states: python/goss/dolfinutils.py:670-748) on the device.
 * dofs [num_field_states][num_nodes] (host memory): the index, in a field
 * vector of num_values doubles, of field state c of node n -- the reference's
 * dof_maps[label][field_name] (dolfinutils.py:290-296) stacked in field-state
 * order.  Checked once on the host (every index in [0, num_values), none
 * repeated) and kept on the device with the system.  gather: field state c of
 * node n <- values[dofs[c][n]]; scatter: values[dofs[c][n]] <- field state c
 * of node n.  values is device memory of num_values doubles, which must equal
 * the dof map's; both calls return after the device work is done. */
int goss_hip_system_set_field_dofs(goss_hip_system* sys, const int32_t* dofs, uint64_t num_values);
int goss_hip_system_gather_field_states_device(goss_hip_system* sys, const double* dev_values, uint64_t num_values);
int goss_hip_system_scatter_field_states_device(goss_hip_system* sys, double* dev_values, uint64_t num_values);

/* states(): all states, node-major [num_nodes][num_states] as in
 * ODESystemSolver::_states (ODESystemSolver.h:74-96), copied out / in. */
int goss_hip_system_get_states(goss_hip_system* sys, double* out);
int goss_hip_system_set_states(goss_hip_system* sys, const double* in);

/* reset_default (ODESystemSolver.cpp:192-216). */
int goss_hip_system_reset_default(goss_hip_system* sys);

/* ParameterizedODE::set_parameter / get_parameter (ParameterizedODE.h:110-120):
 * a shared model parameter.  Unknown names fail with the reference's message. */
int goss_hip_system_set_parameter(goss_hip_system* sys, const char* name, double value);
int goss_hip_system_get_parameter(goss_hip_system* sys, const char* name, double* value);

/* set_num_threads / get_num_threads (ODESystemSolver.cpp:218-250): recorded
 * and returned; the device path does not use host threads. */
int goss_hip_system_set_num_threads(goss_hip_system* sys, uint32_t num_threads);
int goss_hip_system_get_num_threads(goss_hip_system* sys, uint32_t* num_threads);

int goss_hip_system_num_nodes(goss_hip_system* sys, uint32_t* num_nodes);
/* Outer steps fused into one kernel launch by forward_times / solve. */
int goss_hip_system_set_steps_per_launch(goss_hip_system* sys, uint32_t steps);
/* Sums over nodes of the per-node solver counters.
 * ImplicitEuler / ESDIRK23a / ThetaSolver / BasicImplicitEuler / ESDIRK4O32:
 * out[0] Newton iterations of the last stage/solve
 * (ImplicitODESolver::_newton_iterations), out[1] Jacobian evaluations
 * (num_jac_comp, ImplicitODESolver.h:65-68), out[2] rejected Newton solves /
 * steps (BasicImplicitEuler: sub-steps whose iteration did not converge;
 * ESDIRK4O32: failed stage solves plus steps rejected by the error test,
 * AdaptiveImplicitSolver::num_rejected), out[3] accepted ESDIRK23a /
 * ESDIRK4O32 steps (num_accepted).
 * RKF32 (RKF32.h:70-72): out[0] 0, out[1] nfevals, out[2] ndtsr, out[3] ndtsa.
 * Zeros for the fixed-step solvers. */
int goss_hip_system_get_solver_counters(goss_hip_system* sys, int64_t* out);
/* Discrete adjoint of explicit time stepping.  No counterpart in the
 * reference: it serves gradient-based calibration of per-node parameters from
 * the Python side (a torch.autograd.Function over these calls).
 * Requirements: an explicit fixed-step solver (ExplicitEuler, RK2, RK4, RL1,
 * RL2, GRL1, GRL2), precision FP64 or
 * FP32, internal_time_step <= 0 (one step per interval).  The times are those of goss_hip_system_forward_times: n points,
 * n-1 steps.  The first field state is sampled after every `sample_every`
 * steps.
 *
 * goss_hip_adjoint_buffer_sizes: bytes of the caller-allocated device buffers
 * (checkpoints: one state copy every `checkpoint_every` steps; segment: the
 * backward sweep's recomputed states -- two segments' worth when
 * checkpoint_every > 1, since the recompute of one segment overlaps the
 * reverse sweep of the next on a second stream of the system -- and, in FP64,
 * its per-node tape of intermediates) and the number of samples.
 * goss_hip_adjoint_record: advances the system like forward_times, writing the
 * checkpoints and samples[n_samples][N] (device, FP64).
 * goss_hip_adjoint_backward: given d loss / d samples (device, [n_samples][N]),
 * writes d loss / d initial states (device, [N][num_states]) and d loss / d
 * field parameters (device, [N][num_field_parameters]) for the run recorded
 * into `checkpoints` with the same times and field parameters.
 * checkpoint_every = 1 stores every state ((n-1) x num_states x N values of
 * the precision's size) and the backward then runs the reverse sweep alone,
 * reading them in place: no recomputation, at the cost of the memory. */
int goss_hip_adjoint_buffer_sizes(goss_hip_system* sys, uint64_t n, uint32_t sample_every, uint32_t checkpoint_every,
                                  uint64_t* checkpoint_bytes, uint64_t* segment_bytes, uint64_t* num_samples);
int goss_hip_adjoint_record(goss_hip_system* sys, const double* t, uint64_t n, uint32_t sample_every,
                            uint32_t checkpoint_every, void* checkpoints_dev, double* samples_dev);
int goss_hip_adjoint_backward(goss_hip_system* sys, const double* t, uint64_t n, uint32_t sample_every,
                              uint32_t checkpoint_every, const void* checkpoints_dev, const double* grad_samples_dev,
                              void* segment_dev, double* grad_states_dev, double* grad_field_parameters_dev);
/* Mixed-population batch.  The reference advances one population per
 * ODESystemSolver::forward call (ODESystemSolver.cpp:67-79); a batch advances
 * several systems -- different cell models and solvers, uneven sizes --
 * together, with one fused persistent kernel launch per block of time steps
 * and the launches of one call captured in a hipGraph.  Each node runs its own
 * system's per-node code, so results equal those of each system's own
 * goss_hip_system_forward_times over the same times.  Populations must be
 * FP32 and of a (model, solver) pair compiled into the fused kernel (TP06 M
 * and epi with GRL1, FitzHugh-Nagumo with RK4, Lorenz with RKF32); others
 * give GOSS_HIP_EUNSUPPORTED.  The batch references the systems, which must
 * outlive it.
 *
 * goss_hip_batch_forward advances population p from t0 to t1 in outer steps
 * of dt[p] (t_k = t0 + k dt[p], intervals t_k+1 - t_k; (t1 - t0) / dt[p] must
 * be a whole number).  The config's span is the simulated time per launch
 * (lowered to a common divisor of the step counts); block_size is the
 * threads per block (64..256, a multiple of 64) and nodes_per_lane the
 * number of block_size-node chunks each block takes per launch.
 * goss_hip_batch_autotune times candidate configs (dispatch orders, then
 * launch shapes) on [t0, t1] and keeps the fastest; the systems' states are
 * restored afterwards. */
typedef struct goss_hip_batch goss_hip_batch;
#define GOSS_HIP_BATCH_MAX_POPULATIONS 16
typedef struct goss_hip_batch_config {
  uint32_t block_size;
  uint32_t nodes_per_lane;
  double span;
  /* Dispatch order: population order[i] takes the i-th range of chunks (the first
   * range starts first; a population of long per-node latency belongs there). */
  uint32_t order[GOSS_HIP_BATCH_MAX_POPULATIONS];
} goss_hip_batch_config;
int goss_hip_batch_create(goss_hip_system* const* systems, uint32_t n, goss_hip_batch** out);
int goss_hip_batch_destroy(goss_hip_batch* batch);
int goss_hip_batch_set_config(goss_hip_batch* batch, const goss_hip_batch_config* config);
int goss_hip_batch_get_config(goss_hip_batch* batch, goss_hip_batch_config* config);
int goss_hip_batch_forward(goss_hip_batch* batch, double t0, double t1, const double* dt);
int goss_hip_batch_autotune(goss_hip_batch* batch, double t0, double t1, const double* dt,
                            goss_hip_batch_config* best);
/* Wait for all work of the system. */
int goss_hip_system_synchronize(goss_hip_system* sys);
/* Kernel timing, no counterpart in the reference (benchmarks and profiling):
 * when enabled, the solver kernels of each stepping call (forward,
 * forward_times, forward_times_noalloc, solve) are bracketed by HIP events on
 * the system's stream, recorded right before the first and right after the
 * last launch (solve: of its last chunk of rows).  get_kernel_timing returns
 * the elapsed milliseconds between them and the number of launches of the
 * last stepping call (it waits for that call's kernels). */
int goss_hip_system_set_kernel_timing(goss_hip_system* sys, int enable);
int goss_hip_system_get_kernel_timing(goss_hip_system* sys, double* kernel_ms, uint64_t* launches);
/* The system's HIP stream (hipStream_t) for callers that order their own work. */
int goss_hip_system_stream(goss_hip_system* sys, void** stream);

#ifdef __cplusplus
}
#endif
#endif /* GOSS_HIP_H */
