// Generated by goss_amd.codegen from lorentz.ode -- do not edit.
// Variant lorentz: field states ['x'], field parameters [].
#pragma once
#include "goss_model_support.h"

namespace goss_models { namespace lorentz {

constexpr int num_states = 3;
constexpr int num_parameters = 3;
constexpr int num_constants = 0;
constexpr int num_field_states = 1;
constexpr int num_field_parameters = 0;
constexpr int field_state_indices[] = {0};
constexpr int field_parameter_indices[] = {-1};
static const char* const state_names[] = {"x", "y", "z"};
static const char* const parameter_names[] = {"beta", "rho", "sigma"};
static const char* const field_state_names[] = {"x"};
static const char* const field_parameter_names[] = {nullptr};
constexpr double default_parameters[] = {2.6666666666666665, 28.0, 10.0};
constexpr double initial_conditions[] = {1.0, 0.0, 0.0};
constexpr unsigned char linear_terms[] = {1, 1, 1};

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
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  values[0] = sigma*(y - x);
  values[1] = -y + x*(rho - z);
  values[2] = x*y - beta*z;
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
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  rhs[0] = sigma*(y - x);
  rhs[1] = -y + x*(rho - z);
  rhs[2] = x*y - beta*z;
  linear[0] = -sigma;
  linear[1] = -real(1.0);
  linear[2] = -beta;
  if (only_linear) return;
}

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, real* __restrict__ jac)
{
  (void)time; (void)pa;
  const real x = states[0];
  const real y = states[1];
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  const real ddy_dt_dx = rho - z;
  const real ddy_dt_dz = -x;
  jac[0] = -sigma;
  jac[1] = sigma;
  jac[2] = real(0.0);
  jac[3] = ddy_dt_dx;
  jac[4] = -real(1.0);
  jac[5] = ddy_dt_dz;
  jac[6] = y;
  jac[7] = x;
  jac[8] = -beta;
}

}}  // namespace goss_models::lorentz
