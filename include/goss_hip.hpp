// goss_hip.hpp -- C++ host API over the C ABI of goss_hip.h.
//
// Mirrors the reference's host classes, with their member names, public
// tunables, argument meaning, defaults and error behaviour:
//   goss_hip::ParameterizedODE   <- goss::ParameterizedODE (ParameterizedODE.h:33-203)
//   goss_hip::ODESolver          <- goss::ODESolver (ODESolver.h:34-159), and the solver classes
//     ExplicitEuler, RK2, RK4, RL1, RL2, GRL1 (delta, GRL1.h:39), GRL2,
//     ImplicitODESolver (eta_0, kappa, relative_tolerance, max_iterations,
//       max_relative_previous_residual, always_recompute_jacobian; ImplicitODESolver.h:36-41),
//     ImplicitEuler (num_refinements_without_always_recomputing_jacobian, min_dt; ImplicitEuler.h:36-37),
//     ThetaSolver (+ theta; ThetaSolver.h:35-37), ESDIRK23a (ESDIRK23a.h:37-38), BasicImplicitEuler,
//     ESDIRK4O32 (set_tol; ESDIRK4O32.h:31-90),
//     AdaptiveExplicitSolver / RKF32 (set_tol, get_atol, get_rtol; nfevals, ndtsa, ndtsr)
//   goss_hip::ODESystemSolver    <- goss::ODESystemSolver (ODESystemSolver.h:38-186)
// so that code written against goss's ODESystemSolver, e.g.
//   auto ode = std::make_shared<goss_hip::ParameterizedODE>("tp06_m_fp3");
//   auto solver = std::make_shared<goss_hip::GRL1>();
//   solver->delta = 1e-8;
//   goss_hip::ODESystemSolver system(N, solver, ode);
// compiles with the namespace changed.  Semantics a caller must know:
//   * a ParameterizedODE is a model variant of the device library: compiled in
//     (goss_hip_model_name) or loaded from a model library built at run time
//     (ParameterizedODE::load; the Python package's goss_amd.jit builds them);
//   * the system holds the solver and the model as the reference does: their
//     tunables and parameter values are read again at every forward and solve;
//   * states() returns a host mirror of the device states.  While a mirror is
//     in use, every stepping call first copies it to the device (so writes
//     through the pointer take effect) and copies the states back after, as
//     with the reference's host array; this costs two PCIe copies of all states
//     per call, so large systems use get_states / set_states instead.
// Failures throw goss_hip::Error (a std::runtime_error) with the library's
// message, which is the reference's where the reference throws.
#ifndef GOSS_HIP_HPP
#define GOSS_HIP_HPP

#include <cstdint>
#include <cstring>
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

// Names of the model variants known to the library.
inline std::vector<std::string> model_names()
{
  std::vector<std::string> out;
  const int n = goss_hip_model_count();
  for (int i = 0; i < n; ++i) out.emplace_back(goss_hip_model_name(i));
  return out;
}

// goss::ParameterizedODE: sizes, names, initial conditions and parameter
// values of one model variant, and its methods evaluated on the device.
class ParameterizedODE {
 public:
  explicit ParameterizedODE(const std::string& variant) : variant_(variant) { init(); }
  // A variant from a model library built at run time (goss_hip_model_load).
  static std::shared_ptr<ParameterizedODE> load(const std::string& library_path, const std::string& variant)
  {
    check(goss_hip_model_load(library_path.c_str(), variant.c_str()));
    return std::make_shared<ParameterizedODE>(variant);
  }

  const std::string& variant() const { return variant_; }
  uint num_states() const { return uint(info_.num_states); }
  uint num_parameters() const { return uint(info_.num_parameters); }
  uint num_field_states() const { return uint(info_.num_field_states); }
  uint num_field_parameters() const { return uint(info_.num_field_parameters); }
  uint num_monitored() const { return uint(monitored_.size()); }
  const std::vector<std::string>& get_state_names() const { return state_names_; }
  const std::vector<std::string>& get_parameter_names() const { return parameter_names_; }
  const std::vector<std::string>& get_field_state_names() const { return field_state_names_; }
  const std::vector<std::string>& get_field_parameter_names() const { return field_parameter_names_; }
  std::vector<std::string> get_monitored_names() const
  {
    std::vector<std::string> out;
    for (int k : monitored_) out.push_back(intermediates_[k]);
    return out;
  }
  std::vector<uint> get_field_state_indices() const
  {
    return std::vector<uint>(info_.field_state_indices, info_.field_state_indices + info_.num_field_states);
  }
  // ODE::get_ic
  void get_ic(double* values) const
  {
    for (int i = 0; i < info_.num_states; ++i) values[i] = info_.initial_conditions[i];
  }
  // ODE::linear_term
  bool linear_term(uint i) const { return info_.linear_terms[i] != 0; }
  // ODE::is_dae and ODE::differential_states (ODE.h:80-89)
  bool is_dae() const { return info_.is_dae != 0; }
  std::vector<unsigned char> differential_states() const
  {
    return std::vector<unsigned char>(info_.differential_states, info_.differential_states + info_.num_states);
  }
  // ParameterizedODE::get_parameter / set_parameter (ParameterizedODE.h:110-120,
  // unknown names: ParameterizedODE.h:185-195).
  double get_parameter(const std::string& name) const { return values_[index_of(name)]; }
  void set_parameter(const std::string& name, double value) { values_[index_of(name)] = value; }
  const std::vector<double>& parameter_values() const { return values_; }
  // The intermediates eval_monitored returns, by name (the reference chooses them when the class is generated).
  void set_monitored(const std::vector<std::string>& names)
  {
    std::vector<int> idx;
    for (const auto& n : names) {
      int k = -1;
      for (std::size_t j = 0; j < intermediates_.size(); ++j)
        if (intermediates_[j] == n) k = int(j);
      if (k < 0) throw Error(GOSS_HIP_EINVAL, "\"" + n + "\" is not an intermediate of this ODE.");
      idx.push_back(k);
    }
    monitored_ = idx;
  }
  // ODE::eval / linearized_eval / compute_jacobian (ODE.h:55-77), on the device.
  void eval(const double* states, double time, double* values) const
  {
    check(goss_hip_model_eval(variant_.c_str(), values_.data(), states, &time, 1, values));
  }
  double eval(uint idx, const double* states, double time) const
  {
    if (idx >= num_states()) throw Error(GOSS_HIP_EINVAL, "throw - Index out of range");
    std::vector<double> v(num_states());
    eval(states, time, v.data());
    return v[idx];
  }
  void linearized_eval(const double* states, double time, double* linearized, double* rhs, bool only_linear) const
  {
    check(goss_hip_model_linearized_eval(variant_.c_str(), values_.data(), states, &time, 1, only_linear, linearized,
                                         rhs));
  }
  void compute_jacobian(double* states, double time, double* jac) const
  {
    check(goss_hip_model_compute_jacobian(variant_.c_str(), values_.data(), states, &time, 1, jac));
  }
  void lu_factorize(double* mat) const { check(goss_hip_lu_factorize(int(num_states()), 1, mat)); }
  void forward_backward_subst(const double* mat, const double* b, double* x) const
  {
    check(goss_hip_forward_backward_subst(int(num_states()), 1, mat, b, x));
  }
  // ParameterizedODE::eval_monitored / monitored_values (ParameterizedODE.h:90-105).
  void eval_monitored(const double* states, double t, double* monitored) const
  {
    monitored_values(states, &t, monitored, 1);
  }
  void monitored_values(const double* states, const double* t, double* monitored, ulong num_timesteps) const
  {
    if (monitored_.empty() || num_timesteps == 0) return;
    check(goss_hip_model_eval_monitored(variant_.c_str(), values_.data(), int(monitored_.size()), monitored_.data(),
                                        states, t, num_timesteps, monitored));
  }

 private:
  void init()
  {
    check(goss_hip_model_info_get(variant_.c_str(), &info_));
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
    int n = 0;
    const char* const* inter = nullptr;
    check(goss_hip_model_intermediates(variant_.c_str(), &n, &inter));
    intermediates_ = names(inter, n);
  }
  std::size_t index_of(const std::string& name) const
  {
    for (std::size_t k = 0; k < parameter_names_.size(); ++k)
      if (parameter_names_[k] == name) return k;
    throw Error(GOSS_HIP_EINVAL, "\"" + name + "\" is not a parameter in  this ODE.");
  }

  std::string variant_;
  goss_hip_model_info info_{};
  std::vector<std::string> state_names_, parameter_names_, field_state_names_, field_parameter_names_;
  std::vector<std::string> intermediates_;
  std::vector<int> monitored_;
  std::vector<double> values_;
};

namespace detail {

// A device system kept in step with a solver's tunables and a model's
// parameter values: changes made to either between calls are pushed before
// the next launch (the reference's system holds the very objects).
struct Synced {
  goss_hip_solver_params params{};
  std::vector<double> values;
};

inline bool same(const goss_hip_solver_params& a, const goss_hip_solver_params& b)
{
  return std::memcmp(&a, &b, sizeof(a)) == 0;
}

inline void push(goss_hip_system* sys, Synced& s, const goss_hip_solver_params& params, const ParameterizedODE& ode,
                 bool first)
{
  if (first || !same(params, s.params)) check(goss_hip_system_set_solver_params(sys, &params));
  const std::vector<double>& now = ode.parameter_values();
  for (std::size_t k = 0; k < now.size(); ++k)
    if (first || now[k] != s.values[k])
      check(goss_hip_system_set_parameter(sys, ode.get_parameter_names()[k].c_str(), now[k]));
  s.params = params;
  s.values = now;
}

}  // namespace detail

// goss::ODESolver (ODESolver.h:34-159): the base of the solver classes.
class ODESolver {
 public:
  virtual ~ODESolver() { release(); }
  ODESolver(const ODESolver& o) : kind_(o.kind_), ldt_(o.ldt_), ode_(o.ode_) {}
  ODESolver& operator=(const ODESolver& o)
  {
    if (this != &o) {
      release();
      ldt_ = o.ldt_;
      ode_ = o.ode_;
    }
    return *this;
  }

  int kind() const { return kind_; }
  // ODESolver::get/set_internal_time_step (ODESolver.h:119-128).
  double get_internal_time_step() const { return ldt_; }
  void set_internal_time_step(double ldt) { ldt_ = ldt; }
  // ODESolver::is_adaptive (ODESolver.h:131-134).
  bool is_adaptive() const { return goss_hip_solver_is_adaptive(kind_) != 0; }
  // The tunables as the C ABI's parameter block.
  goss_hip_solver_params parameters() const
  {
    goss_hip_solver_params p{};
    check(goss_hip_solver_params_default(kind_, &p));
    p.internal_time_step = ldt_;
    fill(p);
    return p;
  }
  // ODESolver::attach (ODESolver.h:62-67) and reset (a fresh solver state).
  virtual void attach(std::shared_ptr<ParameterizedODE> ode)
  {
    release();
    ode_ = std::move(ode);
  }
  virtual void reset() { release(); }
  std::shared_ptr<ParameterizedODE> get_ode() const { return ode_; }
  uint num_states() const { return ode_ ? ode_->num_states() : 0; }

  // ODESolver::forward (ODESolver.h:75): y of one node advanced over `interval` from t, on a
  // one-node FP64 device system whose solver state carries from call to call.
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
    const std::size_t S = num_states();
    for (unsigned long it = 1; it < n; ++it) {
      forward(y0, t[it - 1], t[it] - t[it - 1]);
      for (std::size_t j = 0; j < S; ++j) y[it * S + j] = y0[j];
    }
  }

 protected:
  explicit ODESolver(int kind) : kind_(kind) {}
  // Copies the subclass's public tunables into the parameter block.
  virtual void fill(goss_hip_solver_params&) const {}
  // Summed solver counters of the one-node system (goss_hip_system_get_solver_counters).
  int64_t counter(int k) const
  {
    if (!sys_) return 0;
    int64_t c[4] = {0, 0, 0, 0};
    check(goss_hip_system_get_solver_counters(sys_, c));
    return c[k];
  }

 private:
  goss_hip_system* one_node()
  {
    if (!ode_) throw Error(GOSS_HIP_EINVAL, "ODESolver: no ODE attached");
    const goss_hip_solver_params p = parameters();
    const bool first = sys_ == nullptr;
    if (first) check(goss_hip_system_create(ode_->variant().c_str(), kind_, &p, GOSS_HIP_FP64, 1, &sys_));
    try {
      detail::push(sys_, synced_, p, *ode_, first);
      if (ode_->num_field_parameters()) {
        std::vector<double> fp;
        for (const std::string& name : ode_->get_field_parameter_names()) fp.push_back(ode_->get_parameter(name));
        check(goss_hip_system_set_field_parameters(sys_, fp.data(), 1));
      }
    } catch (...) {
      release();
      throw;
    }
    return sys_;
  }
  void release()
  {
    if (sys_) goss_hip_system_destroy(sys_);
    sys_ = nullptr;
  }

  int kind_;
  double ldt_ = -1.0;  // ODESolver.h:137
  std::shared_ptr<ParameterizedODE> ode_;
  goss_hip_system* sys_ = nullptr;
  detail::Synced synced_;
};

class ExplicitEuler : public ODESolver {
 public:
  ExplicitEuler() : ODESolver(GOSS_HIP_EXPLICIT_EULER) {}
};
class RK2 : public ODESolver {
 public:
  RK2() : ODESolver(GOSS_HIP_RK2) {}
};
class RK4 : public ODESolver {
 public:
  RK4() : ODESolver(GOSS_HIP_RK4) {}
};
class RL1 : public ODESolver {
 public:
  RL1() : ODESolver(GOSS_HIP_RL1) {}
};
class RL2 : public ODESolver {
 public:
  RL2() : ODESolver(GOSS_HIP_RL2) {}
};
class GRL1 : public ODESolver {
 public:
  GRL1() : ODESolver(GOSS_HIP_GRL1) {}
  double delta = 1e-8;  // GRL1.h:39

 protected:
  explicit GRL1(int kind) : ODESolver(kind) {}
  void fill(goss_hip_solver_params& p) const override { p.delta = delta; }
};
class GRL2 : public GRL1 {
 public:
  GRL2() : GRL1(GOSS_HIP_GRL2) {}
};

class ImplicitODESolver : public ODESolver {
 public:
  // ImplicitODESolver.h:36-41
  double eta_0 = 1.0;
  double kappa = 0.1;
  double relative_tolerance = 1.e-12;
  int max_iterations = 30;
  double max_relative_previous_residual = 0.01;
  bool always_recompute_jacobian = false;
  // How the Jacobian is formed: GOSS_HIP_JACOBIAN_FD, the reference's forward differences
  // (ODE.cpp:47-67), GOSS_HIP_JACOBIAN_ANALYTIC, the model's generated analytic Jacobian (a plugin
  // generated with generate_jacobian), or GOSS_HIP_JACOBIAN_FD_FP32 (no counterpart in the reference).
  int jacobian = GOSS_HIP_JACOBIAN_FD;
  // Number of Jacobian evaluations so far (ImplicitODESolver.h:65-68).
  int num_jac_comp() const { return int(counter(1)); }

 protected:
  explicit ImplicitODESolver(int kind) : ODESolver(kind) {}
  void fill(goss_hip_solver_params& p) const override
  {
    p.eta_0 = eta_0;
    p.kappa = kappa;
    p.relative_tolerance = relative_tolerance;
    p.max_iterations = max_iterations;
    p.max_relative_previous_residual = max_relative_previous_residual;
    p.always_recompute_jacobian = always_recompute_jacobian;
    p.jacobian = jacobian;
  }
};

class ImplicitEuler : public ImplicitODESolver {
 public:
  ImplicitEuler() : ImplicitODESolver(GOSS_HIP_IMPLICIT_EULER) {}
  int num_refinements_without_always_recomputing_jacobian = 2;  // ImplicitEuler.h:36
  double min_dt = 0.0001;                                        // ImplicitEuler.h:37

 protected:
  explicit ImplicitEuler(int kind) : ImplicitODESolver(kind) {}
  void fill(goss_hip_solver_params& p) const override
  {
    ImplicitODESolver::fill(p);
    p.num_refinements_without_always_recomputing_jacobian = num_refinements_without_always_recomputing_jacobian;
    p.min_dt = min_dt;
  }
};

// BasicImplicitEuler (BasicImplicitEuler.cpp:62-181): full Newton, a new Jacobian every iteration.
class BasicImplicitEuler : public ImplicitODESolver {
 public:
  BasicImplicitEuler() : ImplicitODESolver(GOSS_HIP_BASIC_IMPLICIT_EULER) {}
};

class ThetaSolver : public ImplicitEuler {
 public:
  ThetaSolver() : ImplicitEuler(GOSS_HIP_THETA) {}
  double theta = 0.5;  // ThetaSolver.h:37

 protected:
  void fill(goss_hip_solver_params& p) const override
  {
    ImplicitEuler::fill(p);
    p.theta = theta;
  }
};

class ESDIRK23a : public ImplicitODESolver {
 public:
  ESDIRK23a() : ImplicitODESolver(GOSS_HIP_ESDIRK23A) {}
  int num_refinements_without_always_recomputing_jacobian = 2;  // ESDIRK23a.h:37
  double min_dt = 0.001;                                         // ESDIRK23a.h:38
  // AdaptiveImplicitSolver::set_tol (AdaptiveImplicitSolver.h): recorded; ESDIRK23a's own
  // error test uses its fixed tol = 1e-5 (ESDIRK23a.cpp:34).
  void set_tol(double atol, double rtol = 1.0e-8)
  {
    atol_ = atol;
    rtol_ = rtol;
  }
  double get_atol() const { return atol_; }
  double get_rtol() const { return rtol_; }
  long ndtsa() const { return long(counter(3)); }
  long ndtsr() const { return long(counter(2)); }

 protected:
  void fill(goss_hip_solver_params& p) const override
  {
    ImplicitODESolver::fill(p);
    p.num_refinements_without_always_recomputing_jacobian = num_refinements_without_always_recomputing_jacobian;
    p.min_dt = min_dt;
  }

 private:
  double atol_ = 1.0e-5, rtol_ = 1.0e-8;
};

// ESDIRK4O32 (ESDIRK4O32.h:31-90) over AdaptiveImplicitSolver (AdaptiveImplicitSolver.h:31-150): the error
// test uses set_tol's atol / rtol.
class ESDIRK4O32 : public ImplicitODESolver {
 public:
  ESDIRK4O32() : ImplicitODESolver(GOSS_HIP_ESDIRK4O32) {}
  void set_tol(double atol, double rtol = 1.0e-8)
  {
    atol_ = atol;
    rtol_ = rtol;
  }
  double get_atol() const { return atol_; }
  double get_rtol() const { return rtol_; }
  int get_iord() const { return 3; }  // ESDIRK4O32.cpp:38
  long get_num_accepted() const { return long(counter(3)); }
  long get_num_rejected() const { return long(counter(2)); }

 protected:
  void fill(goss_hip_solver_params& p) const override
  {
    ImplicitODESolver::fill(p);
    p.atol = atol_;
    p.rtol = rtol_;
  }

 private:
  double atol_ = 1.0e-5, rtol_ = 1.0e-8;  // AdaptiveImplicitSolver::reset (AdaptiveImplicitSolver.cpp:61-62)
};

// AdaptiveExplicitSolver (AdaptiveExplicitSolver.h:35-130) and RKF32 (RKF32.h:34-100).
class AdaptiveExplicitSolver : public ODESolver {
 public:
  void set_tol(double atol, double rtol = 1.0e-8)
  {
    atol_ = atol;
    rtol_ = rtol;
  }
  double get_atol() const { return atol_; }
  double get_rtol() const { return rtol_; }
  void set_iord(int iord) { iord_ = iord; }
  int get_iord() const { return iord_; }

 protected:
  explicit AdaptiveExplicitSolver(int kind, int iord) : ODESolver(kind), iord_(iord) {}
  void fill(goss_hip_solver_params& p) const override
  {
    p.atol = atol_;
    p.rtol = rtol_;
  }

 private:
  double atol_ = 1.0e-5, rtol_ = 1.0e-8;  // AdaptiveExplicitSolver.cpp:52-53
  int iord_;
};

class RKF32 : public AdaptiveExplicitSolver {
 public:
  RKF32() : AdaptiveExplicitSolver(GOSS_HIP_RKF32, 3) {}
  long nfevals() const { return long(counter(1)); }
  long ndtsr() const { return long(counter(2)); }
  long ndtsa() const { return long(counter(3)); }
};

// A solver by its name (python/goss/solvers.py:476-490 solver_mapper).
inline std::shared_ptr<ODESolver> make_solver(const std::string& name)
{
  if (name == "ExplicitEuler") return std::make_shared<ExplicitEuler>();
  if (name == "RK2") return std::make_shared<RK2>();
  if (name == "RK4") return std::make_shared<RK4>();
  if (name == "RL1") return std::make_shared<RL1>();
  if (name == "RL2") return std::make_shared<RL2>();
  if (name == "GRL1") return std::make_shared<GRL1>();
  if (name == "GRL2") return std::make_shared<GRL2>();
  if (name == "ImplicitEuler") return std::make_shared<ImplicitEuler>();
  if (name == "ThetaSolver") return std::make_shared<ThetaSolver>();
  if (name == "ESDIRK23a") return std::make_shared<ESDIRK23a>();
  if (name == "RKF32") return std::make_shared<RKF32>();
  if (name == "BasicImplicitEuler") return std::make_shared<BasicImplicitEuler>();
  if (name == "ESDIRK4O32") return std::make_shared<ESDIRK4O32>();
  throw Error(GOSS_HIP_EINVAL, "unknown solver " + name);
}

// goss::ODESystemSolver on the device.
class ODESystemSolver {
 public:
  // ODESystemSolver(nodes, solver, ode) (ODESystemSolver.cpp:45-59): attaches the model to the
  // solver; states and field parameters start at the model's defaults (the model's current
  // parameter values, field parameters included).
  ODESystemSolver(uint nodes, std::shared_ptr<ODESolver> solver, std::shared_ptr<ParameterizedODE> ode,
                  int precision = GOSS_HIP_FP64)
      : solver_(std::move(solver)), ode_(std::move(ode)), num_nodes_(nodes)
  {
    if (!solver_ || !ode_) throw Error(GOSS_HIP_EINVAL, "ODESystemSolver needs a solver and a model");
    solver_->attach(ode_);
    const goss_hip_solver_params p = solver_->parameters();
    check(goss_hip_system_create(ode_->variant().c_str(), solver_->kind(), &p, precision, nodes, &sys_));
    try {
      detail::push(sys_, synced_, p, *ode_, true);
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
  void forward(double t, double interval)
  {
    before();
    check(goss_hip_system_forward(sys_, t, interval));
    after();
  }
  // forward(t[i], t[i+1] - t[i]) for every i, fused into few launches.
  void forward_times(const double* t, ulong num_timesteps)
  {
    before();
    check(goss_hip_system_forward_times(sys_, t, num_timesteps));
    after();
  }
  // Step all nodes through the times t, writing the field states after each
  // step into rows 1.. of field_states (ODESystemSolver.cpp:81-99).
  void solve(double* field_states, double* t, const ulong num_timesteps, bool tangled_storage = true)
  {
    before();
    check(goss_hip_system_solve(sys_, field_states, t, num_timesteps, tangled_storage));
    after();
  }
  // solve into pinned host memory (goss_hip_host_alloc) or device memory, returning once the steps and the
  // row copies are enqueued; synchronize() waits for them.
  void solve_async(double* field_states, const double* t, const ulong num_timesteps, bool tangled_storage = true)
  {
    before();
    check(goss_hip_system_solve_async(sys_, field_states, t, num_timesteps, tangled_storage));
    after();
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
    before();
    check(goss_hip_system_set_field_state_components(sys_, component_field_states, num_components,
                                                     reinterpret_cast<const uint32_t*>(components), tangled_storage));
    after();
  }
  void get_field_states(double* system_field_states, bool tangled_storage = true) const
  {
    check(goss_hip_system_get_field_states(sys_, system_field_states, tangled_storage));
  }
  void set_field_states(const double* system_field_states, bool tangled_storage = true)
  {
    before();
    check(goss_hip_system_set_field_states(sys_, system_field_states, tangled_storage));
    after();
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
  void reset_default()
  {
    detail::push(sys_, synced_, solver_->parameters(), *ode_, false);
    check(goss_hip_system_reset_default(sys_));
    after();
  }
  void set_num_threads(uint num_threads) { check(goss_hip_system_set_num_threads(sys_, num_threads)); }
  uint get_num_threads() const
  {
    uint32_t n = 0;
    check(goss_hip_system_get_num_threads(sys_, &n));
    return n;
  }
  // All states, node-major [num_nodes][num_states]: a host mirror kept in step with the device
  // (see the header comment).
  double* states()
  {
    if (!mirror_on_) {
      host_states_.resize(std::size_t(num_nodes_) * ode_->num_states());
      check(goss_hip_system_get_states(sys_, host_states_.data()));
      mirror_on_ = true;
    }
    return host_states_.data();
  }
  double* states(uint node) { return states() + std::size_t(node) * ode_->num_states(); }
  // Copies of all states without the mirror.
  void get_states(double* out) const { check(goss_hip_system_get_states(sys_, out)); }
  void set_states(const double* in)
  {
    check(goss_hip_system_set_states(sys_, in));
    after();
  }
  std::shared_ptr<ParameterizedODE> ode() { return ode_; }
  std::shared_ptr<ODESolver> solver() { return solver_; }
  uint num_nodes() const { return num_nodes_; }
  // A shared parameter as the device system holds it.
  double get_parameter(const std::string& name) const
  {
    double v = 0.0;
    check(goss_hip_system_get_parameter(sys_, name.c_str(), &v));
    return v;
  }
  // Outer steps fused into one kernel launch by forward_times and solve.
  void set_steps_per_launch(uint steps) { check(goss_hip_system_set_steps_per_launch(sys_, steps)); }
  void synchronize() { check(goss_hip_system_synchronize(sys_)); }
  goss_hip_system* handle() { return sys_; }

 private:
  // The solver's tunables and the model's parameter values as they are now, and the mirror's writes.
  void before()
  {
    detail::push(sys_, synced_, solver_->parameters(), *ode_, false);
    if (mirror_on_) check(goss_hip_system_set_states(sys_, host_states_.data()));
  }
  void after()
  {
    if (mirror_on_) check(goss_hip_system_get_states(sys_, host_states_.data()));
  }

  std::shared_ptr<ODESolver> solver_;
  std::shared_ptr<ParameterizedODE> ode_;
  uint num_nodes_;
  goss_hip_system* sys_ = nullptr;
  detail::Synced synced_;
  std::vector<double> host_states_;
  bool mirror_on_ = false;
};

}  // namespace goss_hip

#endif  // GOSS_HIP_HPP
