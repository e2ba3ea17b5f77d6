// Reference checker built around the reference's own gotran-generated model
// classes (TEST INFRASTRUCTURE ONLY; built by oracle/Makefile into
// _ref/libgoss_refpin.so from the sources where they lie).
//
// The models are the generated headers the reference ships with its C++
// tests -- cpp/tests/fhn.h, beeler_reuter_1977.h, Winslow.h (field states and
// field parameters) and Panfilov.h (a field state) -- not this project's
// code generator -- and the reference's hand-written DAE test model
// cpp/tests/van_der_Pol.h (a plain ODE with an algebraic state and its own
// analytic compute_jacobian).  Three things are exported:
//   * the ODESystemSolver C ABI of ref_system_abi.h over the reference
//     solvers (cpp/goss/*.cpp), with these models;
//   * each model's metadata and right-hand side as an oracle_external_model,
//     so the oracle's solver restatement can run on the very same classes
//     (oracle_register_external_model);
//   * refpin_ode_run: a reference solver advancing one ODE (the van der Pol
//     DAE, which ODESystemSolver cannot take: it is not a ParameterizedODE)
//     through ODESolver::forward, as the reference's DAE test does
//     (cpp/tests/ode.cpp:56-78).
// Comparing the two pins the oracle's solvers against the reference with no
// code of this project inside the model.
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "goss/ParameterizedODE.h"

#include "tests/Panfilov.h"
#include "tests/Winslow.h"
#include "tests/beeler_reuter_1977.h"
#include "tests/fhn.h"
#include "tests/van_der_Pol.h"

#include "goss_oracle.h"

namespace {

using namespace goss;

std::shared_ptr<ParameterizedODE> make_model(const std::string& v)
{
  if (v == "ref_fhn") return std::make_shared<Fhn>();
  if (v == "ref_beeler_reuter_1977") return std::make_shared<Beeler_reuter_1977>();
  if (v == "ref_winslow") return std::make_shared<Winslow>();
  if (v == "ref_panfilov") return std::make_shared<Panfilov>();
  throw std::runtime_error("unknown model variant " + v);
}

const char* const kModelNames[] = {"ref_fhn", "ref_beeler_reuter_1977", "ref_winslow", "ref_panfilov"};
constexpr int kNumModels = 4;

// One model's metadata, read from a default-constructed instance, and the
// oracle_external_model that points into it.
struct ExternalModel {
  std::string name;
  std::vector<std::string> pnames_s;
  std::vector<const char*> pnames;
  std::vector<int> fsi, fpi;
  std::vector<double> defaults, ic;
  std::vector<unsigned char> linear;
  oracle_external_model ops{};
};

// A model instance per thread and model, with the parameter vector it holds;
// eval pushes only the parameters that changed (set_parameter by name,
// ParameterizedODE.h:111-114).
struct ThreadModel {
  std::shared_ptr<ParameterizedODE> m;
  std::vector<double> held;
};

ParameterizedODE& model_with(const ExternalModel& em, const double* p)
{
  thread_local std::vector<ThreadModel> tm(kNumModels);
  int k = 0;
  while (k < kNumModels && em.name != kModelNames[k]) ++k;
  ThreadModel& t = tm[k];
  if (!t.m) {
    t.m = make_model(em.name);
    t.held = em.defaults;
  }
  for (size_t i = 0; i < t.held.size(); ++i) {
    if (p[i] != t.held[i] || std::signbit(p[i]) != std::signbit(t.held[i])) {
      t.m->set_parameter(em.pnames_s[i], p[i]);
      t.held[i] = p[i];
    }
  }
  return *t.m;
}

void ext_eval(void* ctx, const double* y, double t, const double* p, double* out)
{
  model_with(*static_cast<ExternalModel*>(ctx), p).eval(y, t, out);
}

void ext_linearized(void* ctx, const double* y, double t, const double* p, double* lin, double* rhs, int only_linear)
{
  model_with(*static_cast<ExternalModel*>(ctx), p).linearized_eval(y, t, lin, rhs, only_linear != 0);
}

ExternalModel* external(int k)
{
  static std::vector<std::unique_ptr<ExternalModel>> all(kNumModels);
  if (!all[k]) {
    auto em = std::make_unique<ExternalModel>();
    em->name = kModelNames[k];
    auto m = make_model(em->name);
    const uint S = m->num_states(), P = m->num_parameters();
    em->pnames_s = m->get_parameter_names();
    for (auto& s : em->pnames_s) em->pnames.push_back(s.c_str());
    for (uint i = 0; i < P; ++i) em->defaults.push_back(m->get_parameter(em->pnames_s[i]));
    DoubleVector ic;
    m->get_ic(&ic);
    em->ic.assign(ic.data.get(), ic.data.get() + S);
    for (uint i = 0; i < S; ++i) em->linear.push_back((unsigned char)m->linear_term(i));
    for (uint i : m->get_field_state_indices()) em->fsi.push_back(int(i));
    for (auto& fp : m->get_field_parameter_names()) {
      int idx = -1;
      for (uint i = 0; i < P; ++i)
        if (em->pnames_s[i] == fp) idx = int(i);
      if (idx < 0) throw std::runtime_error("field parameter " + fp + " is not a parameter");
      em->fpi.push_back(idx);
    }
    oracle_external_model& o = em->ops;
    o.name = em->name.c_str();
    o.num_states = int(S);
    o.num_parameters = int(P);
    o.num_field_states = int(em->fsi.size());
    o.num_field_parameters = int(em->fpi.size());
    o.field_state_indices = em->fsi.data();
    o.field_parameter_indices = em->fpi.data();
    o.default_parameters = em->defaults.data();
    o.initial_conditions = em->ic.data();
    o.linear_terms = em->linear.data();
    o.parameter_names = em->pnames.data();
    o.ctx = em.get();
    o.eval = ext_eval;
    o.linearized_eval = ext_linearized;
    all[k] = std::move(em);
  }
  return all[k].get();
}

// The van der Pol DAE (an ODE, no parameters): one instance per thread.
goss::ODE& van_der_pol()
{
  thread_local std::shared_ptr<goss::ODE> m = std::make_shared<VanDerPol>();
  return *m;
}

void vdp_eval(void*, const double* y, double t, const double*, double* out) { van_der_pol().eval(y, t, out); }
void vdp_linearized(void*, const double* y, double t, const double*, double* lin, double* rhs, int only_linear)
{
  van_der_pol().linearized_eval(y, t, lin, rhs, only_linear != 0);
}
void vdp_jacobian(void*, double* y, double t, const double*, double* jac)
{
  van_der_pol().compute_jacobian(y, t, jac);
}

struct OdeModel {
  std::vector<double> ic;
  std::vector<unsigned char> linear, diff;
  oracle_external_model ops{};
};

const oracle_external_model* van_der_pol_ops()
{
  static std::unique_ptr<OdeModel> em;
  if (!em) {
    em = std::make_unique<OdeModel>();
    auto& m = van_der_pol();
    const uint S = m.num_states();
    DoubleVector ic;
    m.get_ic(&ic);
    em->ic.assign(ic.data.get(), ic.data.get() + S);
    for (uint i = 0; i < S; ++i) {
      em->linear.push_back((unsigned char)m.linear_term(i));
      em->diff.push_back(m.differential_states()[i]);
    }
    oracle_external_model& o = em->ops;
    o.name = "ref_van_der_pol";
    o.num_states = int(S);
    o.initial_conditions = em->ic.data();
    o.linear_terms = em->linear.data();
    o.differential_states = em->diff.data();
    o.eval = vdp_eval;
    o.linearized_eval = vdp_linearized;
    o.compute_jacobian = vdp_jacobian;
  }
  return &em->ops;
}

}  // namespace

#include "ref_system_abi.h"

extern "C" {

int refpin_model_count(void) { return kNumModels; }
const char* refpin_model_name(int k) { return (k >= 0 && k < kNumModels) ? kModelNames[k] : nullptr; }

// The model's oracle_external_model (valid while this library stays loaded).
const oracle_external_model* refpin_model_ops(const char* name)
{
  try {
    if (name && std::string(name) == "ref_van_der_pol") return van_der_pol_ops();
    for (int k = 0; k < kNumModels; ++k)
      if (name && std::string(name) == kModelNames[k]) return &external(k)->ops;
    g_error = std::string("unknown model ") + (name ? name : "(null)");
  } catch (std::exception& e) {
    g_error = e.what();
  }
  return nullptr;
}

// One reference solver advancing the van der Pol DAE from y (in place) over
// nsteps intervals of dt from t0, with solver tunables by name:
// solver->forward(y, t, dt); t += dt (cpp/tests/ode.cpp:61-78).
int refpin_ode_run(const char* solver, int n_params, const char* const* names, const double* values, double* y,
                   double t0, double dt, int nsteps)
{
  try {
    auto s = make_solver(solver);
    for (int i = 0; i < n_params; ++i) set_solver_param(*s, names[i], values[i]);
    s->attach(std::make_shared<VanDerPol>());
    double t = t0;
    for (int i = 0; i < nsteps; ++i) {
      s->forward(y, t, dt);
      t += dt;
    }
    return 0;
  } catch (std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

}  // extern "C"
