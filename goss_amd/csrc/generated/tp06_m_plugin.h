// Generated by goss_amd.codegen from tentusscher_panfilov_2006_M_cell.ode -- do not edit.
#ifndef TP06_M_PLUGIN_H_IS_INCLUDED
#define TP06_M_PLUGIN_H_IS_INCLUDED
#include <memory>
#include <stdexcept>
#include <cmath>
#include "goss/ParameterizedODE.h"
#include "tp06_m.h"

namespace goss {

class Tp06_m : public ParameterizedODE
{
  struct Acc { const double* v; template <int k> double get() const { return v[k]; } };
public:
  Tp06_m() : ParameterizedODE(19, 53, 1, 0, 0)
  {
    for (int i = 0; i < 19; ++i) _state_names[i] = goss_models::tp06_m::state_names[i];
    for (int i = 0; i < 53; ++i) {
      _parameter_names[i] = goss_models::tp06_m::parameter_names[i];
      _p[i] = goss_models::tp06_m::default_parameters[i];
    }
    _field_state_names[0] = "V";
    _field_state_indices[0] = 17;
    _bind();
    for (int i = 0; i < 19; ++i) _linear_terms[i] = goss_models::tp06_m::linear_terms[i];
  }
  Tp06_m(const Tp06_m& ode) : ParameterizedODE(ode)
  {
    for (int i = 0; i < 53; ++i) _p[i] = ode._p[i];
    _bind();
  }
  void eval(const double* states, double time, double* values)
  { goss_models::tp06_m::eval<double>(states, time, Acc{_p}, values); }
  void linearized_eval(const double* states, double time, double* linear, double* rhs,
                       bool only_linear) const
  { goss_models::tp06_m::linearized_eval<double>(states, time, Acc{_p}, linear, rhs, only_linear); }
  void get_ic(goss::DoubleVector* values) const
  {
    values->n = _num_states;
    values->data.reset(new double[_num_states]);
    for (int i = 0; i < 19; ++i) values->data[i] = goss_models::tp06_m::initial_conditions[i];
  }
  std::shared_ptr<ODE> copy() const
  { return std::make_shared<Tp06_m>(*this); }
  void eval_monitored(const double*, double, double*) const
  { throw std::runtime_error("No monitored in the 'Tp06_m' model."); }
  void set_field_parameters(const double* field_params)
  {
    (void)field_params;
  }
private:
  double _p[53];
  void _bind()
  {
    for (int i = 0; i < 53; ++i) _param_to_value[_parameter_names[i]] = &_p[i];
  }
};

}  // namespace goss

extern "C" __attribute__((visibility("default"))) goss::ParameterizedODE* create_Tp06_m()
{
  return new goss::Tp06_m;
}
#endif
