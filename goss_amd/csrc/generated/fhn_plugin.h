// Generated by goss_amd.codegen from fhn.ode -- do not edit.
#ifndef FHN_PLUGIN_H_IS_INCLUDED
#define FHN_PLUGIN_H_IS_INCLUDED
#include <memory>
#include <stdexcept>
#include <cmath>
#include "goss/ParameterizedODE.h"
#include "fhn.h"

namespace goss {

class Fhn : public ParameterizedODE
{
  struct Acc { const double* v; template <int k> double get() const { return v[k]; } };
public:
  Fhn() : ParameterizedODE(2, 11, 1, 0, 0)
  {
    for (int i = 0; i < 2; ++i) _state_names[i] = goss_models::fhn::state_names[i];
    for (int i = 0; i < 11; ++i) {
      _parameter_names[i] = goss_models::fhn::parameter_names[i];
      _p[i] = goss_models::fhn::default_parameters[i];
    }
    _field_state_names[0] = "v";
    _field_state_indices[0] = 0;
    _bind();
    for (int i = 0; i < 2; ++i) _linear_terms[i] = goss_models::fhn::linear_terms[i];
  }
  Fhn(const Fhn& ode) : ParameterizedODE(ode)
  {
    for (int i = 0; i < 11; ++i) _p[i] = ode._p[i];
    _bind();
  }
  void eval(const double* states, double time, double* values)
  { goss_models::fhn::eval<double>(states, time, Acc{_p}, values); }
  void linearized_eval(const double* states, double time, double* linear, double* rhs,
                       bool only_linear) const
  { goss_models::fhn::linearized_eval<double>(states, time, Acc{_p}, linear, rhs, only_linear); }
  void get_ic(goss::DoubleVector* values) const
  {
    values->n = _num_states;
    values->data.reset(new double[_num_states]);
    for (int i = 0; i < 2; ++i) values->data[i] = goss_models::fhn::initial_conditions[i];
  }
  std::shared_ptr<ODE> copy() const
  { return std::make_shared<Fhn>(*this); }
  void eval_monitored(const double*, double, double*) const
  { throw std::runtime_error("No monitored in the 'Fhn' model."); }
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

extern "C" __attribute__((visibility("default"))) goss::ParameterizedODE* create_Fhn()
{
  return new goss::Fhn;
}
#endif
