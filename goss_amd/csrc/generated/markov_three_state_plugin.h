// Generated by goss_amd.codegen from markov_three_state.ode -- do not edit.
#ifndef MARKOV_THREE_STATE_PLUGIN_H_IS_INCLUDED
#define MARKOV_THREE_STATE_PLUGIN_H_IS_INCLUDED
#include <memory>
#include <stdexcept>
#include <cmath>
#include "goss/ParameterizedODE.h"
#include "markov_three_state.h"

namespace goss {

class Markov_three_state : public ParameterizedODE
{
  struct Acc {
    const double* v;
    const double* c;
    template <int k> double get() const { return v[k]; }
    template <int k> double getd() const { return v[k]; }
    template <int k> double cst() const { return c[k]; }
    template <int k> double cstd() const { return c[k]; }
  };
  static constexpr int K = goss_models::markov_three_state::num_constants > 0 ? goss_models::markov_three_state::num_constants : 1;
public:
  Markov_three_state() : ParameterizedODE(3, 6, 1, 1, 0)
  {
    for (int i = 0; i < 3; ++i) _state_names[i] = goss_models::markov_three_state::state_names[i];
    for (int i = 0; i < 6; ++i) {
      _parameter_names[i] = goss_models::markov_three_state::parameter_names[i];
      _p[i] = goss_models::markov_three_state::default_parameters[i];
    }
    _field_state_names[0] = "O";
    _field_state_indices[0] = 1;
    _field_parameter_names[0] = "k_oi";
    _bind();
    for (int i = 0; i < 3; ++i) _linear_terms[i] = goss_models::markov_three_state::linear_terms[i];
  }
  Markov_three_state(const Markov_three_state& ode) : ParameterizedODE(ode)
  {
    for (int i = 0; i < 6; ++i) _p[i] = ode._p[i];
    analytic_jacobian = ode.analytic_jacobian;
    _bind();
  }
  // Parameters can change through _param_to_value at any time, so the
  // parameter-only constants are recomputed on every evaluation.
  void eval(const double* states, double time, double* values)
  {
    double c[K];
    goss_models::markov_three_state::compute_constants<double>(Acc{_p, nullptr}, c);
    goss_models::markov_three_state::eval<double>(states, time, Acc{_p, c}, values);
  }
  void linearized_eval(const double* states, double time, double* linear, double* rhs,
                       bool only_linear) const
  {
    double c[K];
    goss_models::markov_three_state::compute_constants<double>(Acc{_p, nullptr}, c);
    goss_models::markov_three_state::linearized_eval<double>(states, time, Acc{_p, c}, linear, rhs, only_linear);
  }
  void get_ic(goss::DoubleVector* values) const
  {
    values->n = _num_states;
    values->data.reset(new double[_num_states]);
    for (int i = 0; i < 3; ++i) values->data[i] = goss_models::markov_three_state::initial_conditions[i];
  }
  // ODE::compute_jacobian (ODE.h:71): the analytic Jacobian when analytic_jacobian is set, as a plugin
  // generated with generate_jacobian would have it (codegeneration.py:186, 622-641); otherwise the
  // forward differences of the base class, as gotran's default plugin.
  void compute_jacobian(double* states, double time, double* jac)
  {
    if (!analytic_jacobian) {
      ODE::compute_jacobian(states, time, jac);
      return;
    }
    double c[K];
    goss_models::markov_three_state::compute_constants<double>(Acc{_p, nullptr}, c);
    goss_models::markov_three_state::compute_jacobian<double>(states, time, Acc{_p, c}, jac);
  }
  bool analytic_jacobian = false;
  std::shared_ptr<ODE> copy() const
  { return std::make_shared<Markov_three_state>(*this); }
  void eval_monitored(const double*, double, double*) const
  { throw std::runtime_error("No monitored in the 'Markov_three_state' model."); }
  void set_field_parameters(const double* field_params)
  {
    _p[4] = field_params[0];
  }
private:
  double _p[6];
  void _bind()
  {
    for (int i = 0; i < 6; ++i) _param_to_value[_parameter_names[i]] = &_p[i];
  }
};

}  // namespace goss

extern "C" __attribute__((visibility("default"))) goss::ParameterizedODE* create_Markov_three_state()
{
  return new goss::Markov_three_state;
}
#endif
