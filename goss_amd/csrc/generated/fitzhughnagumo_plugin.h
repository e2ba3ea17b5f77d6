// Generated by goss_amd.codegen from fitzhughnagumo.ode -- do not edit.
#ifndef FITZHUGHNAGUMO_PLUGIN_H_IS_INCLUDED
#define FITZHUGHNAGUMO_PLUGIN_H_IS_INCLUDED
#include <memory>
#include <stdexcept>
#include <cmath>
#include "goss/ParameterizedODE.h"
#include "fitzhughnagumo.h"

namespace goss {

class Fitzhughnagumo : public ParameterizedODE
{
  struct Acc {
    const double* v;
    const double* c;
    template <int k> double get() const { return v[k]; }
    template <int k> double getd() const { return v[k]; }
    template <int k> double cst() const { return c[k]; }
    template <int k> double cstd() const { return c[k]; }
  };
  static constexpr int K = goss_models::fitzhughnagumo::num_constants > 0 ? goss_models::fitzhughnagumo::num_constants : 1;
public:
  Fitzhughnagumo() : ParameterizedODE(2, 11, 1, 0, 0)
  {
    for (int i = 0; i < 2; ++i) _state_names[i] = goss_models::fitzhughnagumo::state_names[i];
    for (int i = 0; i < 11; ++i) {
      _parameter_names[i] = goss_models::fitzhughnagumo::parameter_names[i];
      _p[i] = goss_models::fitzhughnagumo::default_parameters[i];
    }
    _field_state_names[0] = "V";
    _field_state_indices[0] = 0;
    _bind();
    for (int i = 0; i < 2; ++i) _linear_terms[i] = goss_models::fitzhughnagumo::linear_terms[i];
  }
  Fitzhughnagumo(const Fitzhughnagumo& ode) : ParameterizedODE(ode)
  {
    for (int i = 0; i < 11; ++i) _p[i] = ode._p[i];
    analytic_jacobian = ode.analytic_jacobian;
    _bind();
  }
  // Parameters can change through _param_to_value at any time, so the
  // parameter-only constants are recomputed on every evaluation.
  void eval(const double* states, double time, double* values)
  {
    double c[K];
    goss_models::fitzhughnagumo::compute_constants<double>(Acc{_p, nullptr}, c);
    goss_models::fitzhughnagumo::eval<double>(states, time, Acc{_p, c}, values);
  }
  void linearized_eval(const double* states, double time, double* linear, double* rhs,
                       bool only_linear) const
  {
    double c[K];
    goss_models::fitzhughnagumo::compute_constants<double>(Acc{_p, nullptr}, c);
    goss_models::fitzhughnagumo::linearized_eval<double>(states, time, Acc{_p, c}, linear, rhs, only_linear);
  }
  void get_ic(goss::DoubleVector* values) const
  {
    values->n = _num_states;
    values->data.reset(new double[_num_states]);
    for (int i = 0; i < 2; ++i) values->data[i] = goss_models::fitzhughnagumo::initial_conditions[i];
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
    goss_models::fitzhughnagumo::compute_constants<double>(Acc{_p, nullptr}, c);
    goss_models::fitzhughnagumo::compute_jacobian<double>(states, time, Acc{_p, c}, jac);
  }
  bool analytic_jacobian = false;
  std::shared_ptr<ODE> copy() const
  { return std::make_shared<Fitzhughnagumo>(*this); }
  void eval_monitored(const double*, double, double*) const
  { throw std::runtime_error("No monitored in the 'Fitzhughnagumo' model."); }
  void set_field_parameters(const double* field_params)
  {
    (void)field_params;
  }
private:
  double _p[11];
  void _bind()
  {
    for (int i = 0; i < 11; ++i) _param_to_value[_parameter_names[i]] = &_p[i];
  }
};

}  // namespace goss

extern "C" __attribute__((visibility("default"))) goss::ParameterizedODE* create_Fitzhughnagumo()
{
  return new goss::Fitzhughnagumo;
}
#endif
