/* CPU oracle: a restatement of goss's ODESystemSolver and its solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP product never links, loads or calls
 * this library; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, as the checker.
 *
 * Each entry point follows the reference function cited beside it; the
 * arithmetic is written in the same order so that results agree bit for bit
 * with the reference compiled from its own sources (oracle/_ref).
 *
 * Layouts: states are node-major [N][S] like ODESystemSolver::_states
 * (ODESystemSolver.h:163); field parameters [N][P_field]
 * (ODESystemSolver.h:166).
 */
#ifndef GOSS_ORACLE_H
#define GOSS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_system oracle_system;

const char* oracle_last_error(void);

/* Model variants: the generated models compiled into the oracle. */
int oracle_variant_count(void);
const char* oracle_variant_name(int index);
/* Returns the variant index or -1. */
int oracle_variant_find(const char* name);
/* sizes[0..3] = num_states, num_parameters, num_field_states, num_field_parameters */
void oracle_variant_sizes(int v, int* sizes);
void oracle_variant_defaults(int v, double* parameters, double* initial_conditions);
void oracle_variant_field_indices(int v, int* field_state_indices, int* field_parameter_indices);
void oracle_variant_linear_terms(int v, unsigned char* flags);
/* ODE::differential_states (ODE.h:85-89): 0 for the algebraic states of a DAE. */
void oracle_variant_differential_states(int v, unsigned char* flags);

/* A model supplied at run time by another library (the reference-side
 * checker registers the reference's own gotran-generated classes this way):
 * sizes, metadata and its right-hand side, called with the full parameter
 * vector (P values, field parameters substituted per node).  The struct and
 * everything it points to must outlive the oracle's use of it.  Step adjoints
 * are not available for such a model.  differential_states may be NULL (every
 * state differential); compute_jacobian may be NULL (the solvers then form the
 * forward-difference Jacobian of ODE::compute_jacobian, ODE.cpp:47-67), else it
 * stands for a model class that overrides ODE::compute_jacobian, as the
 * reference's hand-written test models do. */
typedef struct oracle_external_model {
  const char* name;
  int num_states, num_parameters, num_field_states, num_field_parameters;
  const int* field_state_indices;
  const int* field_parameter_indices;
  const double* default_parameters;
  const double* initial_conditions;
  const unsigned char* linear_terms;
  const char* const* parameter_names;
  void* ctx;
  void (*eval)(void* ctx, const double* states, double t, const double* params, double* values);
  void (*linearized_eval)(void* ctx, const double* states, double t, const double* params, double* linear,
                          double* rhs, int only_linear);
  const unsigned char* differential_states;
  void (*compute_jacobian)(void* ctx, double* states, double t, const double* params, double* jac);
} oracle_external_model;
/* Returns the new variant index (or the existing one for the same struct), -1 on error. */
int oracle_register_external_model(const oracle_external_model* model);

/* Single-model evaluation with a full parameter vector (P values). */
void oracle_model_eval(int v, const double* states, double t, const double* params, double* values);
void oracle_model_linearized_eval(int v, const double* states, double t, const double* params,
                                  double* linear, double* rhs, int only_linear);
void oracle_model_analytic_jacobian(int v, const double* states, double t, const double* params, double* jac);
/* ODE::compute_jacobian (ODE.cpp:47-67): forward differences. */
void oracle_model_fd_jacobian(int v, double* states, double t, const double* params, double* jac);
void oracle_model_step_vjp(int v, int kind, int tape, const double* states, double t, const double* params,
                           double h, double delta, const double* lam, double* out, double* fp_bar);
/* ODE::lu_factorize (ODE.cpp:69-100) and forward_backward_subst (ODE.cpp:102-132). */
void oracle_lu_factorize(int n, double* mat);
void oracle_forward_backward_subst(int n, const double* mat, const double* b, double* x);

/* ODESystemSolver (ODESystemSolver.h:38-186).  solver: "ExplicitEuler",
 * "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2", "ImplicitEuler", "ESDIRK23a",
 * "RKF32".  Solver tunables by name (internal_time_step, delta, eta_0,
 * kappa, relative_tolerance, max_iterations, max_relative_previous_residual,
 * always_recompute_jacobian, min_dt,
 * num_refinements_without_always_recomputing_jacobian, atol, rtol).
 * per_node_state != 0 keeps implicit/adaptive solver state per node instead
 * of carrying it from node to node as the reference does. */
int oracle_system_create(const char* variant, const char* solver, unsigned num_nodes,
                         int n_params, const char* const* names, const double* values,
                         int per_node_state, oracle_system** out);
void oracle_system_destroy(oracle_system* s);
int oracle_system_set_parameter(oracle_system* s, const char* name, double value);
void oracle_system_set_num_threads(oracle_system* s, unsigned n);
int oracle_system_forward(oracle_system* s, double t, double interval);
int oracle_system_forward_times(oracle_system* s, const double* t, unsigned long n);
int oracle_system_solve(oracle_system* s, double* field_states, const double* t, unsigned long n, int tangled);
void oracle_system_get_states(const oracle_system* s, double* out);
void oracle_system_set_states(oracle_system* s, const double* in);
void oracle_system_get_field_states(const oracle_system* s, double* out, int tangled);
void oracle_system_set_field_states(oracle_system* s, const double* in, int tangled);
void oracle_system_get_field_state_components(const oracle_system* s, double* out, unsigned nc,
                                              const unsigned* comp, int tangled);
void oracle_system_set_field_state_components(oracle_system* s, const double* in, unsigned nc,
                                              const unsigned* comp, int tangled);
void oracle_system_get_field_parameters(const oracle_system* s, double* out, int tangled);
void oracle_system_set_field_parameters(oracle_system* s, const double* in, int tangled);
void oracle_system_reset_default(oracle_system* s);
/* Per-node counters of the implicit/adaptive solvers (Newton iterations,
 * Jacobian evaluations, rejected steps, accepted steps); zeros otherwise. */
void oracle_system_counters(const oracle_system* s, long* out4);
/* ESDIRK4O32: accepted steps, steps rejected by the error test, right-hand-side evaluations. */
void oracle_system_esdirk4o32_counters(const oracle_system* s, long* out3);

#ifdef __cplusplus
}
#endif
#endif
