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
constexpr unsigned char linear_terms[] = {0, 0};
// Diagonal of the mass matrix (ODE::_differential_states): 0 for the algebraic states of a DAE.
constexpr unsigned char differential_states[] = {1, 1};
constexpr bool is_dae = false;

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

constexpr int num_intermediates = 1;
static const char* const intermediate_names[] = {"energy"};
// Every named intermediate of the model at (states, time): the values a caller may monitor
// (ParameterizedODE::eval_monitored, ParameterizedODE.h:90-105, selects among them).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_intermediates(const real* __restrict__ states, const treal time, const PA& pa,
                                            real* __restrict__ out)
{
  (void)time; (void)pa; (void)states; (void)out;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  out[0] = real(a*y + b*x);
}

// Structure of the Jacobian d values[j] / d states[i]: values[j] reads states[i] only for the
// entries below (through any operand of the emitted right-hand side), so every other entry of a
// forward-difference Jacobian is exactly zero.  And the structure of its no-pivot LU factors
// (the entries, the diagonal and the fill).
constexpr int jac_nnz = 2;
constexpr int lu_nnz = 4;
// jac_cols[i]: bit j set when values[j] reads states[i].  Jacobian storage is column-major over the
// structure: entry (j, i) at jac_col_start[i] + (number of bits of jac_cols[i] below j).
constexpr unsigned jac_cols[] = {0x2u, 0x1u};
constexpr int jac_col_start[] = {0, 1};
// LU storage is row-major over its structure: lu_slot[j * num_states + i], -1 outside.
constexpr short lu_slot[] = {0, 1, 2, 3};

// a = M + scale * J on the LU structure (fill entries zero), M the diagonal mass matrix
// (differential_states): the dense scale and add_mass_matrix of
// ImplicitODESolver::compute_factorized_jacobian (ImplicitODESolver.cpp:92-105, 258-262).  The
// algebraic states of a DAE add 0.0, as the reference's weight * 0 does.
template <typename real>
GOSS_HD GOSS_INLINE void lu_from_jacobian(const real* __restrict__ J, const real scale, real* __restrict__ a)
{
  a[0] = real(1.0);
  a[1] = J[1] * scale;
  a[2] = J[0] * scale;
  a[3] = real(1.0);
}

// ODE::lu_factorize (ODE.cpp:69-100) in place on a, over the structure.
template <typename real>
GOSS_HD GOSS_INLINE void lu_factorize_sparse(real* __restrict__ a)
{
  real acc;
  a[2] = a[2] / a[0];
  acc = a[2] * a[1];
  a[3] = a[3] - acc;
  (void)acc;
}

// ODE::forward_backward_subst (ODE.cpp:102-132) in place on x, over the structure.
template <typename real>
GOSS_HD GOSS_INLINE void lu_solve_sparse(const real* __restrict__ a, real* __restrict__ x)
{
  real acc;
  acc = a[2] * x[0];
  x[1] = x[1] - acc;
  x[1] = x[1] / a[3];
  acc = a[1] * x[1];
  x[0] = (x[0] - acc) / a[0];
  (void)acc;
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
  if (only_linear) return;
  linear[0] = real(0.0);
  linear[1] = real(0.0);
}

// linearized_eval (all entries) plus, for the gate states (gate_states[i] = 1: rhs[i] = (A - y_i) r),
// gate_target[i] = A.  A GRL stage of a gate is then y0_i + (y_i - A) expm1(linear[i] h).
constexpr unsigned char gate_states[] = {0, 0};
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void linearized_eval_gates(const real* __restrict__ states, const treal time, const PA& pa,
                                               real* __restrict__ linear, real* __restrict__ rhs,
                                               real* __restrict__ gate_target)
{
  (void)time; (void)pa; (void)gate_target;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  rhs[0] = -a*y;
  rhs[1] = b*x;
  linear[0] = real(0.0);
  linear[1] = real(0.0);
}

// Discrete adjoints of one step of size h at `time` (one step per interval):
//   *_vjp:  lam_prev = (d y'/d y)^T lam,  fp_bar += (d y'/d p_field)^T lam;
//   grl1: y'_i = y_i + rhs_i*h*phi1(linear_i*h) when |linear_i| > delta, else y_i + rhs_i*h (GRL1.h:72-84);
//   rl1:  the same with the exponential update for the linear_terms states only (RL1.h:68-77);
//   for a gate (rhs_i = (A - y_i) r) the exponential update is A + (y_i - A) exp(linear_i*h).
//   eval_vjp: y_bar += (d rhs/d y)^T seed, fp_bar += (d rhs/d p_field)^T seed (the explicit schemes).
// Reverse accumulation over the generated statements; the *_tape forms keep every intermediate
// and adjoint in a caller-provided tape (for targets whose registers cannot hold the sweep).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void grl1_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const real* __restrict__ lam__,
                                  real* __restrict__ lam_prev__, real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  const real rhs__0 = -a*y;
  const real rhs__1 = b*x;
  const real lin__0 = real(0.0);
  const real lin__1 = real(0.0);
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(lin__0), m_rcp(real(lin__0)), real(rhs__0), lam__[0], h__, hinv__, delta__, bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  real a__b = real(0.0);
  real b__b = real(0.0);
  real x__b = real(0.0);
  real y__b = real(0.0);
  b__b += bbar__1*(x);
  x__b += bbar__1*(b);
  a__b += bbar__0*(-y);
  y__b += bbar__0*(-a);
  lam_prev__[0] = lam__[0] + x__b;
  lam_prev__[1] = lam__[1] + y__b;
  fp_bar__[0] += a__b;
  fp_bar__[1] += b__b;
}

template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>
GOSS_HD GOSS_INLINE void grl1_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const LV& lam__,
                                  const OV& lam_prev__, real* __restrict__ fp_bar__,
                                  const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  tape.put(0, -a*y);
  tape.put(1, b*x);
  tape.put(2, real(0.0));
  tape.put(3, real(0.0));
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(tape.get(2)), m_rcp(real(tape.get(2))), real(tape.get(0)), lam__[0], h__, hinv__, delta__, bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(3)), m_rcp(real(tape.get(3))), real(tape.get(1)), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  tape.put(4, real(0.0));
  tape.put(5, real(0.0));
  tape.put(6, real(0.0));
  tape.put(7, real(0.0));
  tape.add(5, bbar__1*(x));
  tape.add(6, bbar__1*(b));
  tape.add(4, bbar__0*(-y));
  tape.add(7, bbar__0*(-a));
  lam_prev__[0] = lam__[0] + tape.get(6);
  lam_prev__[1] = lam__[1] + tape.get(7);
  fp_bar__[0] += tape.get(4);
  fp_bar__[1] += tape.get(5);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void rl1_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const real* __restrict__ lam__,
                                  real* __restrict__ lam_prev__, real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  const real rhs__0 = -a*y;
  const real rhs__1 = b*x;
  const real bbar__0 = lam__[0]*h__;
  const real bbar__1 = lam__[1]*h__;
  real a__b = real(0.0);
  real b__b = real(0.0);
  real x__b = real(0.0);
  real y__b = real(0.0);
  b__b += bbar__1*(x);
  x__b += bbar__1*(b);
  a__b += bbar__0*(-y);
  y__b += bbar__0*(-a);
  lam_prev__[0] = lam__[0] + x__b;
  lam_prev__[1] = lam__[1] + y__b;
  fp_bar__[0] += a__b;
  fp_bar__[1] += b__b;
}

template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>
GOSS_HD GOSS_INLINE void rl1_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const LV& lam__,
                                  const OV& lam_prev__, real* __restrict__ fp_bar__,
                                  const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  tape.put(0, -a*y);
  tape.put(1, b*x);
  const real bbar__0 = lam__[0]*h__;
  const real bbar__1 = lam__[1]*h__;
  tape.put(2, real(0.0));
  tape.put(3, real(0.0));
  tape.put(4, real(0.0));
  tape.put(5, real(0.0));
  tape.add(3, bbar__1*(x));
  tape.add(4, bbar__1*(b));
  tape.add(2, bbar__0*(-y));
  tape.add(5, bbar__0*(-a));
  lam_prev__[0] = lam__[0] + tape.get(4);
  lam_prev__[1] = lam__[1] + tape.get(5);
  fp_bar__[0] += tape.get(2);
  fp_bar__[1] += tape.get(3);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real* __restrict__ seed__, real* __restrict__ y_bar__,
                                  real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  const real rhs__0 = -a*y;
  const real rhs__1 = b*x;
  real a__b = real(0.0);
  real b__b = real(0.0);
  real x__b = real(0.0);
  real y__b = real(0.0);
  b__b += seed__[1]*(x);
  x__b += seed__[1]*(b);
  a__b += seed__[0]*(-y);
  y__b += seed__[0]*(-a);
  y_bar__[0] += x__b;
  y_bar__[1] += y__b;
  fp_bar__[0] += a__b;
  fp_bar__[1] += b__b;
}

template <typename real, typename PA, typename treal, typename TP, typename SV, typename YV>
GOSS_HD GOSS_INLINE void eval_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                       const SV& seed__, const YV& y_bar__, real* __restrict__ fp_bar__,
                                       const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real x = states[0];
  const real y = states[1];
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  tape.put(0, -a*y);
  tape.put(1, b*x);
  tape.put(2, real(0.0));
  tape.put(3, real(0.0));
  tape.put(4, real(0.0));
  tape.put(5, real(0.0));
  tape.add(3, seed__[1]*(x));
  tape.add(4, seed__[1]*(b));
  tape.add(2, seed__[0]*(-y));
  tape.add(5, seed__[0]*(-a));
  y_bar__[0] += tape.get(4);
  y_bar__[1] += tape.get(5);
  fp_bar__[0] += tape.get(2);
  fp_bar__[1] += tape.get(3);
}

constexpr int grl1_vjp_tape_slots = 8;
constexpr int rl1_vjp_tape_slots = 6;
constexpr int eval_vjp_tape_slots = 6;

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
// `jac` is a pointer or any object whose jac[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JAC>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, JAC jac)
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

// The analytic Jacobian over its structure: J[slot] for the entries of jac_cols / jac_col_start
// (d values[j] / d states[i] at jac_col_start[i] + rank of j in jac_cols[i]).  `J` is a pointer
// or any object whose J[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JS>
GOSS_HD GOSS_INLINE void compute_jacobian_structured(const real* __restrict__ states, const treal time,
                                                     const PA& pa, const JS& J)
{
  (void)time; (void)pa;
  const real a = pa.template get<0>();
  const real b = pa.template get<1>();
  J[0] = b;
  const real ddx_dt_dy = -a;
  J[1] = ddx_dt_dy;
}

}}  // namespace goss_models::oscilator
