// goss_hip.hpp -- C++ host API over the C ABI of goss_hip.h.
//
// Mirrors the reference's three host classes, with their member names,
// argument meaning, defaults and error behaviour:
//   goss_hip::ParameterizedODE  <- goss::ParameterizedODE (ParameterizedODE.h:33-203)
//   goss_hip::ODESolver         <- goss::ODESolver and its subclasses' tunables
//                                  (ODESolver.h:34-159, GRL1.h:39, ImplicitODESolver.h:36-41, ...)
//   goss_hip::ODESystemSolver   <- goss::ODESystemSolver (ODESystemSolver.h:38-186)
// so that code written against goss's ODESystemSolver compiles against this
// header with the namespace changed.  Differences a caller must know:
//   * a ParameterizedODE is one of the model variants compiled into
//     libgoss_hip.so (goss_hip_model_name), not a gotran-generated class;
//   * states() returns a host copy of the device states (the reference
//     returns its own host array); write changes back with set_states();
//   * parameter values set on a ParameterizedODE before the system is built
//     become the system's shared parameters; later changes go through the
//     system's set_parameter (the model object is shared as in the reference).
// Failures throw goss_hip::Error (a std::runtime_error) with the library's
// message, which is the reference's where the reference throws.
#ifndef GOSS_HIP_HPP
#define GOSS_HIP_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "goss_hip.h"

namespace goss_hip {

using uint = unsigned int;
using ulong = unsigned long;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc)
{
  if (rc != GOSS_HIP_OK) throw Error(rc, goss_hip_last_error());
}

// Names of the compiled model variants.
inline std::vector<std::string> model_names()
{
  std::vector<std::string> out;
  const int n = goss_hip_model_count();
  for (int i = 0; i < n; ++i) out.emplace_back(goss_hip_model_name(i));
  return out;
}

// goss::ParameterizedODE: sizes, names, initial conditions and parameter
// values of one compiled model variant.
class ParameterizedODE {
 public:
  explicit ParameterizedODE(const std::string& variant) : variant_(variant)
  {
    check(goss_hip_model_info_get(variant.c_str(), &info_));
    auto names = [](const char* const* p, int n) {
      std::vector<std::string> v;
      for (int i = 0; i < n; ++i) v.emplace_back(p[i]);
      return v;
    };
    state_names_ = names(info_.state_names, info_.num_states);
    parameter_names_ = names(info_.parameter_names, info_.num_parameters);
    field_state_names_ = names(info_.field_state_names, info_.num_field_states);
    field_parameter_names_ = names(info_.field_parameter_names, info_.num_field_parameters);
    values_.assign(info_.default_parameters, info_.default_parameters + info_.num_parameters);
  }

  const std::string& variant() const { return variant_; }
  uint num_states() const { return uint(info_.num_states); }
  uint num_parameters() const { return uint(info_.num_parameters); }
  uint num_field_states() const { return uint(info_.num_field_states); }
  uint num_field_parameters() const { return uint(info_.num_field_parameters); }
  const std::vector<std::string>& get_state_names() const { return state_names_; }
  const std::vector<std::string>& get_parameter_names() const { return parameter_names_; }
  const std::vector<std::string>& get_field_state_names() const { return field_state_names_; }
  const std::vector<std::string>& get_field_parameter_names() const { return field_parameter_names_; }
  std::vector<uint> get_field_state_indices() const
  {
    return std::vector<uint>(info_.field_state_indices, info_.field_state_indices + info_.num_field_states);
  }
  // ODE::get_ic
  void get_ic(double* values) const
  {
    for (int i = 0; i < info_.num_states; ++i) values[i] = info_.initial_conditions[i];
  }
  // ODE::linear_terms: 1 where d rhs_i / d state_i does not depend on state_i.
  void linear_terms(uint* indices) const
  {
    for (int i = 0; i < info_.num_states; ++i) indices[i] = info_.linear_terms[i];
  }
  // ParameterizedODE::get_parameter / set_parameter (ParameterizedODE.h:110-120,
  // unknown names: ParameterizedODE.h:185-195).
  double get_parameter(const std::string& name) const { return values_[index_of(name)]; }
  void set_parameter(const std::string& name, double value) { values_[index_of(name)] = value; }

 private:
  std::size_t index_of(const std::string& name) const
  {
    for (std::size_t k = 0; k < parameter_names_.size(); ++k)
      if (parameter_names_[k] == name) return k;
    throw Error(GOSS_HIP_EINVAL, "\"" + name + "\" is not a parameter in  this ODE.");
  }

  std::string variant_;
  goss_hip_model_info info_{};
  std::vector<std::string> state_names_, parameter_names_, field_state_names_, field_parameter_names_;
  std::vector<double> values_;
};

// goss::ODESolver: which solver, and its public tunables.
class ODESolver {
 public:
  // Solver names as in python/goss/solvers.py:423-436.
  explicit ODESolver(const std::string& name) : name_(name), kind_(kind_of(name))
  {
    check(goss_hip_solver_params_default(kind_, &params_));
  }
  ODESolver(const std::string& name, double ldt) : ODESolver(name) { set_internal_time_step(ldt); }

  const std::string& name() const { return name_; }
  int kind() const { return kind_; }
  // ODESolver::get/set_internal_time_step (ODESolver.h:119-128).
  double get_internal_time_step() const { return params_.internal_time_step; }
  void set_internal_time_step(double ldt) { params_.internal_time_step = ldt; }
  // ODESolver::is_adaptive (ODESolver.h:131-134).
  bool is_adaptive() const { return goss_hip_solver_is_adaptive(kind_) != 0; }
  // The remaining tunables (GRL1::delta, ImplicitODESolver's eta_0, kappa,
  // relative_tolerance, ..., AdaptiveExplicitSolver's tolerances) by field.
  goss_hip_solver_params& parameters() { return params_; }
  const goss_hip_solver_params& parameters() const { return params_; }
  // AdaptiveExplicitSolver::set_tol (AdaptiveExplicitSolver.h:71-75).
  void set_tol(double atol, double rtol = 1.0e-8)
  {
    params_.atol = atol;
    params_.rtol = rtol;
  }

  // ODESolver::attach (ODESolver.h:62-67): the model that forward and solve advance.  They run on a
  // one-node FP64 device system, created at the first call with the model's current parameter values
  // (shared and field) and the solver's tunables as they are then.
  void attach(std::shared_ptr<ParameterizedODE> ode)
  {
    release();
    ode_ = std::move(ode);
  }
  // ODESolver::forward (ODESolver.h:75): the states y of one node advanced over `interval` from t.
  void forward(double* y, double t, double interval)
  {
    goss_hip_system* sys = one_node();
    check(goss_hip_system_set_states(sys, y));
    check(goss_hip_system_forward(sys, t, interval));
    check(goss_hip_system_get_states(sys, y));
  }
  // ODESolver::solve (ODESolver.h:78-98): y0 advanced interval by interval over t[0 .. n-1]; row it of
  // y[n][num_states] receives the states at t[it] (row 0 is left as it is, as in the reference).
  void solve(double* y, double* y0, const double* t, unsigned long n)
  {
    const std::size_t S = one_node_states();
    for (unsigned long it = 1; it < n; ++it) {
      forward(y0, t[it - 1], t[it] - t[it - 1]);
      for (std::size_t j = 0; j < S; ++j) y[it * S + j] = y0[j];
    }
  }

  ODESolver(const ODESolver& o) : name_(o.name_), kind_(o.kind_), params_(o.params_), ode_(o.ode_) {}
  ODESolver& operator=(const ODESolver& o)
  {
    if (this != &o) {
      release();
      name_ = o.name_;
      kind_ = o.kind_;
      params_ = o.params_;
      ode_ = o.ode_;
    }
    return *this;
  }
  ~ODESolver() { release(); }

 private:
  goss_hip_system* one_node()
  {
    if (sys_) return sys_;
    if (!ode_) throw Error(GOSS_HIP_EINVAL, "ODESolver: no ODE attached");
    check(goss_hip_system_create(ode_->variant().c_str(), kind_, &params_, GOSS_HIP_FP64, 1, &sys_));
    try {
      for (const std::string& p : ode_->get_parameter_names())
        check(goss_hip_system_set_parameter(sys_, p.c_str(), ode_->get_parameter(p)));
      std::vector<double> fp;
      for (const std::string& p : ode_->get_field_parameter_names()) fp.push_back(ode_->get_parameter(p));
      if (!fp.empty()) check(goss_hip_system_set_field_parameters(sys_, fp.data(), 1));
    } catch (...) {
      release();
      throw;
    }
    return sys_;
  }
  std::size_t one_node_states()
  {
    one_node();
    return ode_->num_states();
  }
  void release()
  {
    if (sys_) goss_hip_system_destroy(sys_);
    sys_ = nullptr;
  }

  static int kind_of(const std::string& name)
  {
    static const char* const names[] = {"ExplicitEuler", "RK2",       "RK4",           "RL1",       "RL2",
                                        "GRL1",          "GRL2",      "ImplicitEuler", "ESDIRK23a", "RKF32"};
    for (int k = 0; k < 10; ++k)
      if (name == names[k]) return k;
    throw Error(GOSS_HIP_EINVAL, "unknown solver " + name);
  }

  std::string name_;
  int kind_;
  goss_hip_solver_params params_{};
  std::shared_ptr<ParameterizedODE> ode_;
  goss_hip_system* sys_ = nullptr;
};

// goss::ODESystemSolver on the device.
class ODESystemSolver {
 public:
  // ODESystemSolver(nodes, solver, ode) (ODESystemSolver.cpp:45-59): states
  // and field parameters start at the model's defaults; the model's current
  // parameter values become the system's shared parameters.
  ODESystemSolver(uint nodes, std::shared_ptr<ODESolver> solver, std::shared_ptr<ParameterizedODE> ode,
                  int precision = GOSS_HIP_FP64)
      : solver_(std::move(solver)), ode_(std::move(ode)), num_nodes_(nodes)
  {
    if (!solver_ || !ode_) throw Error(GOSS_HIP_EINVAL, "ODESystemSolver needs a solver and a model");
    check(goss_hip_system_create(ode_->variant().c_str(), solver_->kind(), &solver_->parameters(), precision, nodes,
                                 &sys_));
    try {
      for (const std::string& p : ode_->get_parameter_names()) {
        double v = 0.0;
        check(goss_hip_system_get_parameter(sys_, p.c_str(), &v));
        const double want = ode_->get_parameter(p);
        if (v != want) check(goss_hip_system_set_parameter(sys_, p.c_str(), want));
      }
      check(goss_hip_system_reset_default(sys_));
    } catch (...) {
      goss_hip_system_destroy(sys_);
      throw;
    }
  }
  ~ODESystemSolver()
  {
    if (sys_) goss_hip_system_destroy(sys_);
  }
  ODESystemSolver(const ODESystemSolver&) = delete;
  ODESystemSolver& operator=(const ODESystemSolver&) = delete;

  // Step all nodes an interval of time forward (ODESystemSolver.cpp:67-79).
  void forward(double t, double interval) { check(goss_hip_system_forward(sys_, t, interval)); }
  // forward(t[i], t[i+1] - t[i]) for every i, fused into few launches.
  void forward_times(const double* t, ulong num_timesteps)
  {
    check(goss_hip_system_forward_times(sys_, t, num_timesteps));
  }
  // Step all nodes through the times t, writing the field states after each
  // step into rows 1.. of field_states (ODESystemSolver.cpp:81-99).
  void solve(double* field_states, double* t, const ulong num_timesteps, bool tangled_storage = true)
  {
    check(goss_hip_system_solve(sys_, field_states, t, num_timesteps, tangled_storage));
  }
  void get_field_state_components(double* component_field_states, uint num_components, const uint* components,
                                  bool tangled_storage = true) const
  {
    check(goss_hip_system_get_field_state_components(sys_, component_field_states, num_components,
                                                     reinterpret_cast<const uint32_t*>(components), tangled_storage));
  }
  void set_field_state_components(const double* component_field_states, uint num_components, const uint* components,
                                  bool tangled_storage = true)
  {
    check(goss_hip_system_set_field_state_components(sys_, component_field_states, num_components,
                                                     reinterpret_cast<const uint32_t*>(components), tangled_storage));
  }
  void get_field_states(double* system_field_states, bool tangled_storage = true) const
  {
    check(goss_hip_system_get_field_states(sys_, system_field_states, tangled_storage));
  }
  void set_field_states(const double* system_field_states, bool tangled_storage = true)
  {
    check(goss_hip_system_set_field_states(sys_, system_field_states, tangled_storage));
  }
  void get_field_parameters(double* system_field_params, bool tangled_storage = true) const
  {
    check(goss_hip_system_get_field_parameters(sys_, system_field_params, tangled_storage));
  }
  void set_field_parameters(const double* system_field_params, bool tangled_storage = true)
  {
    check(goss_hip_system_set_field_parameters(sys_, system_field_params, tangled_storage));
  }
  // Initial conditions and the current values of the field parameters
  // (ODESystemSolver.cpp:192-216).
  void reset_default() { check(goss_hip_system_reset_default(sys_)); }
  void set_num_threads(uint num_threads) { check(goss_hip_system_set_num_threads(sys_, num_threads)); }
  uint get_num_threads() const
  {
    uint32_t n = 0;
    check(goss_hip_system_get_num_threads(sys_, &n));
    return n;
  }
  // A host copy of all states, node-major [num_nodes][num_states].
  const double* states() const
  {
    host_states_.resize(std::size_t(num_nodes_) * ode_->num_states());
    check(goss_hip_system_get_states(sys_, host_states_.data()));
    return host_states_.data();
  }
  const double* states(uint node) const { return states() + std::size_t(node) * ode_->num_states(); }
  void set_states(const double* states) { check(goss_hip_system_set_states(sys_, states)); }
  // A shared model parameter of this system (ParameterizedODE::set_parameter
  // through ode(), ParameterizedODE.h:110-120); the model object follows.
  void set_parameter(const std::string& name, double value)
  {
    check(goss_hip_system_set_parameter(sys_, name.c_str(), value));
    ode_->set_parameter(name, value);
  }
  double get_parameter(const std::string& name) const
  {
    double v = 0.0;
    check(goss_hip_system_get_parameter(sys_, name.c_str(), &v));
    return v;
  }
  std::shared_ptr<ParameterizedODE> ode() { return ode_; }
  std::shared_ptr<ODESolver> solver() { return solver_; }
  uint num_nodes() const { return num_nodes_; }
  // Outer steps fused into one kernel launch by forward_times and solve.
  void set_steps_per_launch(uint steps) { check(goss_hip_system_set_steps_per_launch(sys_, steps)); }
  void synchronize() { check(goss_hip_system_synchronize(sys_)); }
  goss_hip_system* handle() { return sys_; }

 private:
  std::shared_ptr<ODESolver> solver_;
  std::shared_ptr<ParameterizedODE> ode_;
  uint num_nodes_;
  goss_hip_system* sys_ = nullptr;
  mutable std::vector<double> host_states_;
};

}  // namespace goss_hip

#endif  // GOSS_HIP_HPP
