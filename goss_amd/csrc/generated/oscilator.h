// Generated by goss_amd.codegen from oscilator.ode -- do not edit.
// Variant oscilator: field states ['x', 'y'], field parameters ['a', 'b'].
#pragma once
#include "goss_model_support.h"

namespace goss_models { namespace oscilator {

constexpr int num_states = 2;
constexpr int num_parameters = 2;
constexpr int num_constants = 0;
constexpr int num_field_states = 2;
constexpr int num_field_parameters = 2;
constexpr int field_state_indices[] = {0, 1};
constexpr int field_parameter_indices[] = {0, 1};
static const char* const state_names[] = {"x", "y"};
static const char* const parameter_names[] = {"a", "b"};
static const char* const field_state_names[] = {"x", "y"};
static const char* const field_parameter_names[] = {"a", "b"};
constexpr double default_parameters[] = {1.0, 1.0};
constexpr double initial_conditions[] = {1.0, 0.0};
constexpr unsigned char linear_terms[] = {1, 1};

// Subexpressions of shared (non-field) parameters only, evaluated once per launch.
// c[k] may use c[j] for j < k.  Called with real = double on the host.
template <typename real, typename PA>
GOSS_HD GOSS_INLINE void compute_constants(const PA& pa, real* __restrict__ c)
{
  (void)pa; (void)c;
}

// Right-hand side: values[i] = d states[i] / dt.  Intermediates that depend on time but on
// no state are computed in the type of `time` (treal).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval(const real* __restrict__ states, const treal time, const PA& pa,
                              real* __restrict__ values)
{
  (void)time; (void)pa;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  values[0] = -a*y;
  values[1] = b*x;
}

// Linearized right-hand side: rhs[i] as in eval and linear[i] = d rhs[i] / d states[i].
// Entries with linear_terms[i] == 1 are written before the only_linear return.
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void linearized_eval(const real* __restrict__ states, const treal time, const PA& pa,
                                         real* __restrict__ linear, real* __restrict__ rhs,
                                         const bool only_linear)
{
  (void)time; (void)pa;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  rhs[0] = -a*y;
  rhs[1] = b*x;
  linear[0] = real(0.0);
  linear[1] = real(0.0);
  if (only_linear) return;
}

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, real* __restrict__ jac)
{
  (void)time; (void)pa;
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  const real ddx_dt_dy = -a;
  jac[0] = real(0.0);
  jac[1] = ddx_dt_dy;
  jac[2] = b;
  jac[3] = real(0.0);
}

}}  // namespace goss_models::oscilator
