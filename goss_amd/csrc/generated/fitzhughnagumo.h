// Generated by goss_amd.codegen from fitzhughnagumo.ode -- do not edit.
// Variant fitzhughnagumo: field states ['V'], field parameters [].
#pragma once
#include "goss_model_support.h"

namespace goss_models { namespace fitzhughnagumo {

constexpr int num_states = 2;
constexpr int num_parameters = 11;
constexpr int num_field_states = 1;
constexpr int num_field_parameters = 0;
constexpr int field_state_indices[] = {0};
constexpr int field_parameter_indices[] = {-1};
static const char* const state_names[] = {"V", "s"};
static const char* const parameter_names[] = {"V_peak", "V_rest", "a", "b", "c_1", "c_2", "c_3", "stim_amplitude", "stim_duration", "stim_period", "stim_start"};
static const char* const field_state_names[] = {"V"};
static const char* const field_parameter_names[] = {nullptr};
constexpr double default_parameters[] = {40.0, -85.0, 0.13, 0.013, 0.26, 0.1, 1.0, 0.0, 1.0, 1000.0, 1.0};
constexpr double initial_conditions[] = {-85.0, 0.0};
constexpr unsigned char linear_terms[] = {0, 1};

// Right-hand side: values[i] = d states[i] / dt.  Intermediates that depend on time but on
// no state are computed in the type of `time` (treal).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval(const real* __restrict__ states, const treal time, const PA& pa,
                              real* __restrict__ values)
{
  (void)time; (void)pa;
  const real V = states[0];
  const real s = states[1];
  const real V_peak = pa.template get<0>();
  const real V_rest = pa.template get<1>();
  const real a = pa.template get<2>();
  const real b = pa.template get<3>();
  const real c_1 = pa.template get<4>();
  const real c_2 = pa.template get<5>();
  const real c_3 = pa.template get<6>();
  const real stim_amplitude = pa.template get<7>();
  const real stim_duration = pa.template get<8>();
  const real stim_start = pa.template get<10>();
  const real V_amp = V_peak - V_rest;
  const real V_th = V_rest + V_amp*a;
  const real I = -c_2*s*(V - V_rest)/V_amp + c_1*(V - V_rest)*(V - V_th)*(V_peak - V)/m_sq(V_amp);
  const treal i_Stim = treal(stim_amplitude)*(treal(1.0) - treal(1.0)/(treal(1.0) + m_exp(treal(5.0)*time - treal(5.0)*treal(stim_start))))/(treal(1.0) + m_exp(treal(5.0)*time - treal(5.0)*treal(stim_duration) - treal(5.0)*treal(stim_start)));
  values[0] = I + real(i_Stim);
  values[1] = b*(V - V_rest - c_3*s);
}

// Linearized right-hand side: rhs[i] as in eval and linear[i] = d rhs[i] / d states[i].
// Entries with linear_terms[i] == 1 are written before the only_linear return.
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void linearized_eval(const real* __restrict__ states, const treal time, const PA& pa,
                                         real* __restrict__ linear, real* __restrict__ rhs,
                                         const bool only_linear)
{
  (void)time; (void)pa;
  const real V = states[0];
  const real s = states[1];
  const real V_peak = pa.template get<0>();
  const real V_rest = pa.template get<1>();
  const real a = pa.template get<2>();
  const real b = pa.template get<3>();
  const real c_1 = pa.template get<4>();
  const real c_2 = pa.template get<5>();
  const real c_3 = pa.template get<6>();
  const real stim_amplitude = pa.template get<7>();
  const real stim_duration = pa.template get<8>();
  const real stim_start = pa.template get<10>();
  const real V_amp = V_peak - V_rest;
  const real V_th = V_rest + V_amp*a;
  const real I = -c_2*s*(V - V_rest)/V_amp + c_1*(V - V_rest)*(V - V_th)*(V_peak - V)/m_sq(V_amp);
  const treal i_Stim = treal(stim_amplitude)*(treal(1.0) - treal(1.0)/(treal(1.0) + m_exp(treal(5.0)*time - treal(5.0)*treal(stim_start))))/(treal(1.0) + m_exp(treal(5.0)*time - treal(5.0)*treal(stim_duration) - treal(5.0)*treal(stim_start)));
  rhs[0] = I + real(i_Stim);
  rhs[1] = b*(V - V_rest - c_3*s);
  const real dds_dt_ds = -b*c_3;
  linear[1] = dds_dt_ds;
  if (only_linear) return;
  const real dI_dV = -c_2*s/V_amp + c_1*(V - V_rest)*(V_peak - V)/m_sq(V_amp) + c_1*(V - V_th)*(V_peak - V)/m_sq(V_amp) - c_1*(V - V_rest)*(V - V_th)/m_sq(V_amp);
  linear[0] = dI_dV;
}

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, real* __restrict__ jac)
{
  (void)time; (void)pa;
  const real V = states[0];
  const real s = states[1];
  const real V_peak = pa.template get<0>();
  const real V_rest = pa.template get<1>();
  const real a = pa.template get<2>();
  const real b = pa.template get<3>();
  const real c_1 = pa.template get<4>();
  const real c_2 = pa.template get<5>();
  const real c_3 = pa.template get<6>();
  const real V_amp = V_peak - V_rest;
  const real V_th = V_rest + V_amp*a;
  const real dI_dV = -c_2*s/V_amp + c_1*(V - V_rest)*(V_peak - V)/m_sq(V_amp) + c_1*(V - V_th)*(V_peak - V)/m_sq(V_amp) - c_1*(V - V_rest)*(V - V_th)/m_sq(V_amp);
  const real dI_ds = -c_2*(V - V_rest)/V_amp;
  const real dds_dt_ds = -b*c_3;
  jac[0] = dI_dV;
  jac[1] = dI_ds;
  jac[2] = b;
  jac[3] = dds_dt_ds;
}

}}  // namespace goss_models::fitzhughnagumo
