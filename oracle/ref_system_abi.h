// Reference-side C entry points shared by the checkers built from the
// reference's own sources (TEST INFRASTRUCTURE ONLY; see oracle/Makefile):
// ODESystemSolver (ODESystemSolver.h:38-186) driving the reference solvers,
// and single-model calls (ODE.h:55-77, ParameterizedODE.h:90-156).  The
// including file defines, before the include,
//   std::shared_ptr<goss::ParameterizedODE> make_model(const std::string&);
// in the anonymous namespace, choosing which model classes it links.
#pragma once

#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "goss/ODESystemSolver.h"
#include "goss/ExplicitEuler.h"
#include "goss/RK2.h"
#include "goss/RK4.h"
#include "goss/RL1.h"
#include "goss/RL2.h"
#include "goss/GRL1.h"
#include "goss/GRL2.h"
#include "goss/ImplicitEuler.h"
#include "goss/ESDIRK23a.h"
#include "goss/RKF32.h"
#include "goss/ThetaSolver.h"
#include "goss/BasicImplicitEuler.h"
#include "goss/ESDIRK4O32.h"


namespace {

using namespace goss;

std::string g_error;

std::shared_ptr<ODESolver> make_solver(const std::string& s)
{
  if (s == "ExplicitEuler") return std::make_shared<ExplicitEuler>();
  if (s == "RK2") return std::make_shared<RK2>();
  if (s == "RK4") return std::make_shared<RK4>();
  if (s == "RL1") return std::make_shared<RL1>();
  if (s == "RL2") return std::make_shared<RL2>();
  if (s == "GRL1") return std::make_shared<GRL1>();
  if (s == "GRL2") return std::make_shared<GRL2>();
  if (s == "ImplicitEuler") return std::make_shared<ImplicitEuler>();
  if (s == "ESDIRK23a") return std::make_shared<ESDIRK23a>();
  if (s == "RKF32") return std::make_shared<RKF32>();
  if (s == "ThetaSolver") return std::make_shared<ThetaSolver>();
  if (s == "BasicImplicitEuler") return std::make_shared<BasicImplicitEuler>();
  if (s == "ESDIRK4O32") return std::make_shared<ESDIRK4O32>();
  throw std::runtime_error("unknown solver " + s);
}

// Public tunables of the reference solvers (GRL1.h:39, ImplicitODESolver.h:36-41,
// ImplicitEuler.h:36-37, ESDIRK23a.h:37-38, AdaptiveExplicitSolver.h).
void set_solver_param(ODESolver& s, const std::string& name, double value)
{
  if (name == "internal_time_step") { s.set_internal_time_step(value); return; }
  if (auto* g = dynamic_cast<GRL2*>(&s)) { if (name == "delta") { g->delta = value; return; } }
  if (auto* g = dynamic_cast<GRL1*>(&s)) { if (name == "delta") { g->delta = value; return; } }
  if (auto* ie = dynamic_cast<ImplicitEuler*>(&s)) {
    if (name == "min_dt") { ie->min_dt = value; return; }
    if (name == "num_refinements_without_always_recomputing_jacobian") {
      ie->num_refinements_without_always_recomputing_jacobian = (int)value; return; }
  }
  if (auto* th = dynamic_cast<ThetaSolver*>(&s)) {
    if (name == "theta") { th->theta = value; return; }
    if (name == "min_dt") { th->min_dt = value; return; }
    if (name == "num_refinements_without_always_recomputing_jacobian") {
      th->num_refinements_without_always_recomputing_jacobian = (int)value; return; }
  }
  if (auto* es = dynamic_cast<ESDIRK23a*>(&s)) {
    if (name == "min_dt") { es->min_dt = value; return; }
    if (name == "num_refinements_without_always_recomputing_jacobian") {
      es->num_refinements_without_always_recomputing_jacobian = (int)value; return; }
  }
  if (auto* im = dynamic_cast<ImplicitODESolver*>(&s)) {
    if (name == "eta_0") { im->eta_0 = value; return; }
    if (name == "kappa") { im->kappa = value; return; }
    if (name == "relative_tolerance") { im->relative_tolerance = value; return; }
    if (name == "max_iterations") { im->max_iterations = (int)value; return; }
    if (name == "max_relative_previous_residual") { im->max_relative_previous_residual = value; return; }
    if (name == "always_recompute_jacobian") { im->always_recompute_jacobian = value != 0; return; }
  }
  if (auto* ad = dynamic_cast<AdaptiveExplicitSolver*>(&s)) {
    if (name == "atol") { ad->set_tol(value, ad->get_rtol()); return; }
    if (name == "rtol") { ad->set_tol(ad->get_atol(), value); return; }
  }
  if (auto* ai = dynamic_cast<AdaptiveImplicitSolver*>(&s)) {
    if (name == "atol") { ai->set_tol(value, ai->get_rtol()); return; }
    if (name == "rtol") { ai->set_tol(ai->get_atol(), value); return; }
  }
  throw std::runtime_error("unknown solver parameter " + name);
}

struct RefSystem {
  std::shared_ptr<ParameterizedODE> ode;
  std::shared_ptr<ODESolver> solver;
  std::unique_ptr<ODESystemSolver> sys;
};

}  // namespace

#define REF_TRY(expr) try { expr; return 0; } catch (std::exception& e) { g_error = e.what(); return -1; }

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

// Create solver + model, apply solver and model parameters, then build the
// ODESystemSolver (its constructor attaches the model and resets defaults).
int ref_system_create(const char* variant, const char* solver, unsigned num_nodes,
                      int n_solver_params, const char* const* solver_param_names, const double* solver_param_values,
                      int n_model_params, const char* const* model_param_names, const double* model_param_values,
                      void** out)
{
  try {
    auto* r = new RefSystem;
    r->ode = make_model(variant);
    r->solver = make_solver(solver);
    for (int i = 0; i < n_model_params; ++i) r->ode->set_parameter(model_param_names[i], model_param_values[i]);
    // Solver tunables first: the system constructor attaches the model
    // (attach -> reset reads eta_0) and copies the internal time step into
    // its per-node vector (ODESystemSolver.cpp:52-58).
    for (int i = 0; i < n_solver_params; ++i) set_solver_param(*r->solver, solver_param_names[i], solver_param_values[i]);
    r->sys.reset(new ODESystemSolver(num_nodes, r->solver, r->ode));
    // AdaptiveExplicitSolver::reset (run by the attach above) restores the
    // default tolerances (AdaptiveExplicitSolver.cpp:52-53): apply them again.
    for (int i = 0; i < n_solver_params; ++i) {
      const std::string name(solver_param_names[i]);
      if (name == "atol" || name == "rtol") set_solver_param(*r->solver, name, solver_param_values[i]);
    }
    *out = r;
    return 0;
  } catch (std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

void ref_system_destroy(void* h) { delete static_cast<RefSystem*>(h); }

// ESDIRK4O32 counters of the system's own solver (serial mode):
// num_accepted, num_rejected (AdaptiveImplicitSolver.h) and nfevals
// (ESDIRK4O32.h).  Returns -1 for other solvers.
int ref_system_esdirk4o32_counters(void* h, long* out3)
{
  auto* r = static_cast<RefSystem*>(h);
  auto* es = dynamic_cast<ESDIRK4O32*>(r->solver.get());
  if (!es) return -1;
  out3[0] = es->get_num_accepted();
  out3[1] = es->get_num_rejected();
  out3[2] = es->nfevals;
  return 0;
}

// RKF32 counters of the system's own solver (serial mode): nfevals, ndtsa,
// ndtsr (RKF32.h:70-72).  Returns -1 for other solvers.
int ref_system_rkf32_counters(void* h, long* out3)
{
  auto* r = static_cast<RefSystem*>(h);
  auto* rk = dynamic_cast<RKF32*>(r->solver.get());
  if (!rk) return -1;
  out3[0] = rk->nfevals;
  out3[1] = rk->ndtsa;
  out3[2] = rk->ndtsr;
  return 0;
}

int ref_system_set_num_threads(void* h, unsigned n) { REF_TRY(static_cast<RefSystem*>(h)->sys->set_num_threads(n)) }

int ref_system_forward(void* h, double t, double interval)
{
  REF_TRY(static_cast<RefSystem*>(h)->sys->forward(t, interval))
}

// forward(t[i], t[i+1] - t[i]) for i < n-1: the loop of ODESystemSolver::solve
// (ODESystemSolver.cpp:81-99) without the per-step field-state copy.
int ref_system_forward_times(void* h, const double* t, unsigned long n)
{
  try {
    auto* r = static_cast<RefSystem*>(h);
    double tc = t[0];
    for (unsigned long it = 1; it < n; ++it) {
      const double tn = t[it];
      r->sys->forward(tc, tn - tc);
      tc = tn;
    }
    return 0;
  } catch (std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

int ref_system_solve(void* h, double* field_states, double* t, unsigned long n, int tangled)
{
  REF_TRY(static_cast<RefSystem*>(h)->sys->solve(field_states, t, n, tangled != 0))
}

void ref_system_get_states(void* h, double* out)
{
  auto* r = static_cast<RefSystem*>(h);
  const size_t n = size_t(r->sys->num_nodes()) * r->ode->num_states();
  std::memcpy(out, r->sys->states(), n * sizeof(double));
}

void ref_system_set_states(void* h, const double* in)
{
  auto* r = static_cast<RefSystem*>(h);
  const size_t n = size_t(r->sys->num_nodes()) * r->ode->num_states();
  std::memcpy(r->sys->states(), in, n * sizeof(double));
}

void ref_system_get_field_states(void* h, double* out, int tangled)
{ static_cast<RefSystem*>(h)->sys->get_field_states(out, tangled != 0); }
void ref_system_set_field_states(void* h, const double* in, int tangled)
{ static_cast<RefSystem*>(h)->sys->set_field_states(in, tangled != 0); }
void ref_system_get_field_parameters(void* h, double* out, int tangled)
{ static_cast<RefSystem*>(h)->sys->get_field_parameters(out, tangled != 0); }
void ref_system_set_field_parameters(void* h, const double* in, int tangled)
{ static_cast<RefSystem*>(h)->sys->set_field_parameters(in, tangled != 0); }
void ref_system_get_field_state_components(void* h, double* out, unsigned nc, const unsigned* comp, int tangled)
{ static_cast<RefSystem*>(h)->sys->get_field_state_components(out, nc, comp, tangled != 0); }
void ref_system_set_field_state_components(void* h, const double* in, unsigned nc, const unsigned* comp, int tangled)
{ static_cast<RefSystem*>(h)->sys->set_field_state_components(in, nc, comp, tangled != 0); }
int ref_system_reset_default(void* h) { REF_TRY(static_cast<RefSystem*>(h)->sys->reset_default()) }

// Single-model entry points (ODE.h:55-77 and ParameterizedODE.h:90-156).
int ref_model_create(const char* variant, void** out)
{
  try {
    *out = new std::shared_ptr<ParameterizedODE>(make_model(variant));
    return 0;
  } catch (std::exception& e) {
    g_error = e.what();
    return -1;
  }
}
void ref_model_destroy(void* h) { delete static_cast<std::shared_ptr<ParameterizedODE>*>(h); }
static ParameterizedODE& M(void* h) { return **static_cast<std::shared_ptr<ParameterizedODE>*>(h); }
int ref_model_set_parameter(void* h, const char* name, double v) { REF_TRY(M(h).set_parameter(name, v)) }
int ref_model_get_parameter(void* h, const char* name, double* v) { REF_TRY(*v = M(h).get_parameter(name)) }
void ref_model_set_field_parameters(void* h, const double* fp) { M(h).set_field_parameters(fp); }
void ref_model_eval(void* h, const double* y, double t, double* out) { M(h).eval(y, t, out); }
void ref_model_linearized_eval(void* h, const double* y, double t, double* lin, double* rhs, int only_linear)
{ M(h).linearized_eval(y, t, lin, rhs, only_linear != 0); }
void ref_model_compute_jacobian(void* h, double* y, double t, double* jac) { M(h).compute_jacobian(y, t, jac); }
void ref_model_lu_factorize(void* h, double* mat) { M(h).lu_factorize(mat); }
void ref_model_forward_backward_subst(void* h, const double* mat, const double* b, double* x)
{ M(h).forward_backward_subst(mat, b, x); }
int ref_model_linear_term(void* h, unsigned i) { return M(h).linear_term(i); }

}  // extern "C"
