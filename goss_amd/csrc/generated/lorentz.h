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
// Diagonal of the mass matrix (ODE::_differential_states): 0 for the algebraic states of a DAE.
constexpr unsigned char differential_states[] = {1, 1, 1};
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
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  values[0] = sigma*(y - x);
  values[1] = -y + x*(rho - z);
  values[2] = x*y - beta*z;
}

constexpr int num_intermediates = 0;
static const char* const intermediate_names[] = {nullptr};
// Every named intermediate of the model at (states, time): the values a caller may monitor
// (ParameterizedODE::eval_monitored, ParameterizedODE.h:90-105, selects among them).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_intermediates(const real* __restrict__ states, const treal time, const PA& pa,
                                            real* __restrict__ out)
{
  (void)time; (void)pa; (void)states; (void)out;
}

// Structure of the Jacobian d values[j] / d states[i]: values[j] reads states[i] only for the
// entries below (through any operand of the emitted right-hand side), so every other entry of a
// forward-difference Jacobian is exactly zero.  And the structure of its no-pivot LU factors
// (the entries, the diagonal and the fill).
constexpr int jac_nnz = 8;
constexpr int lu_nnz = 8;
// jac_cols[i]: bit j set when values[j] reads states[i].  Jacobian storage is column-major over the
// structure: entry (j, i) at jac_col_start[i] + (number of bits of jac_cols[i] below j).
constexpr unsigned jac_cols[] = {0x7u, 0x7u, 0x6u};
constexpr int jac_col_start[] = {0, 3, 6};
// LU storage is row-major over its structure: lu_slot[j * num_states + i], -1 outside.
constexpr short lu_slot[] = {0, 1, -1, 2, 3, 4, 5, 6, 7};

// a = M + scale * J on the LU structure (fill entries zero), M the diagonal mass matrix
// (differential_states): the dense scale and add_mass_matrix of
// ImplicitODESolver::compute_factorized_jacobian (ImplicitODESolver.cpp:92-105, 258-262).  The
// algebraic states of a DAE add 0.0, as the reference's weight * 0 does.
template <typename real>
GOSS_HD GOSS_INLINE void lu_from_jacobian(const real* __restrict__ J, const real scale, real* __restrict__ a)
{
  a[0] = J[0] * scale + real(1.0);
  a[1] = J[3] * scale;
  a[2] = J[1] * scale;
  a[3] = J[4] * scale + real(1.0);
  a[4] = J[6] * scale;
  a[5] = J[2] * scale;
  a[6] = J[5] * scale;
  a[7] = J[7] * scale + real(1.0);
}

// ODE::lu_factorize (ODE.cpp:69-100) in place on a, over the structure.
template <typename real>
GOSS_HD GOSS_INLINE void lu_factorize_sparse(real* __restrict__ a)
{
  real acc;
  a[2] = a[2] / a[0];
  acc = a[2] * a[1];
  a[3] = a[3] - acc;
  a[5] = a[5] / a[0];
  acc = a[5] * a[1];
  a[6] = (a[6] - acc) / a[3];
  acc = a[6] * a[4];
  a[7] = a[7] - acc;
  (void)acc;
}

// ODE::forward_backward_subst (ODE.cpp:102-132) in place on x, over the structure.
template <typename real>
GOSS_HD GOSS_INLINE void lu_solve_sparse(const real* __restrict__ a, real* __restrict__ x)
{
  real acc;
  acc = a[2] * x[0];
  x[1] = x[1] - acc;
  acc = a[5] * x[0];
  acc = acc + a[6] * x[1];
  x[2] = x[2] - acc;
  x[2] = x[2] / a[7];
  acc = a[4] * x[2];
  x[1] = (x[1] - acc) / a[3];
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

// linearized_eval (all entries) plus, for the gate states (gate_states[i] = 1: rhs[i] = (A - y_i) r),
// gate_target[i] = A.  A GRL stage of a gate is then y0_i + (y_i - A) expm1(linear[i] h).
constexpr unsigned char gate_states[] = {1, 0, 0};
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void linearized_eval_gates(const real* __restrict__ states, const treal time, const PA& pa,
                                               real* __restrict__ linear, real* __restrict__ rhs,
                                               real* __restrict__ gate_target)
{
  (void)time; (void)pa; (void)gate_target;
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
  gate_target[0] = y;
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
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  const real rhs__1 = -y + x*(rho - z);
  const real rhs__2 = x*y - beta*z;
  const real lin__0 = -sigma;
  const real lin__1 = -real(1.0);
  const real lin__2 = -beta;
  const real hinv__ = m_rcp(h__);
  real ey__0, ga__0, gc__0;
  m_grl1_gate_seeds(real(lin__0), x, real(y), lam__[0], h__, delta__, ey__0, ga__0, gc__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(lin__2), m_rcp(real(lin__2)), real(rhs__2), lam__[2], h__, hinv__, delta__, bbar__2, abar__2);
  real x__b = real(0.0);
  real y__b = ga__0;
  real z__b = real(0.0);
  x__b += bbar__2*(y);
  y__b += bbar__2*(x);
  z__b += bbar__2*(-beta);
  x__b += bbar__1*(rho - z);
  y__b += bbar__1*(-real(1.0));
  z__b += bbar__1*(-x);
  lam_prev__[0] = ey__0 + x__b;
  lam_prev__[1] = lam__[1] + y__b;
  lam_prev__[2] = lam__[2] + z__b;
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
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  tape.put(0, -y + x*(rho - z));
  tape.put(1, x*y - beta*z);
  tape.put(2, -sigma);
  tape.put(3, -real(1.0));
  tape.put(4, -beta);
  const real hinv__ = m_rcp(h__);
  real ey__0, ga__0, gc__0;
  m_grl1_gate_seeds(real(tape.get(2)), x, real(y), lam__[0], h__, delta__, ey__0, ga__0, gc__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(3)), m_rcp(real(tape.get(3))), real(tape.get(0)), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(tape.get(4)), m_rcp(real(tape.get(4))), real(tape.get(1)), lam__[2], h__, hinv__, delta__, bbar__2, abar__2);
  tape.put(5, real(0.0));
  tape.put(6, ga__0);
  tape.put(7, real(0.0));
  tape.add(5, bbar__2*(y));
  tape.add(6, bbar__2*(x));
  tape.add(7, bbar__2*(-beta));
  tape.add(5, bbar__1*(rho - z));
  tape.add(6, bbar__1*(-real(1.0)));
  tape.add(7, bbar__1*(-x));
  lam_prev__[0] = ey__0 + tape.get(5);
  lam_prev__[1] = lam__[1] + tape.get(6);
  lam_prev__[2] = lam__[2] + tape.get(7);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void rl1_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const real* __restrict__ lam__,
                                  real* __restrict__ lam_prev__, real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real x = states[0];
  const real y = states[1];
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  const real rhs__1 = -y + x*(rho - z);
  const real rhs__2 = x*y - beta*z;
  const real lin__0 = -sigma;
  const real lin__1 = -real(1.0);
  const real lin__2 = -beta;
  const real hinv__ = m_rcp(h__);
  real ey__0, ga__0, gc__0;
  m_grl1_gate_seeds(real(lin__0), x, real(y), lam__[0], h__, real(-1.0), ey__0, ga__0, gc__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, real(-1.0), bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(lin__2), m_rcp(real(lin__2)), real(rhs__2), lam__[2], h__, hinv__, real(-1.0), bbar__2, abar__2);
  real x__b = real(0.0);
  real y__b = ga__0;
  real z__b = real(0.0);
  x__b += bbar__2*(y);
  y__b += bbar__2*(x);
  z__b += bbar__2*(-beta);
  x__b += bbar__1*(rho - z);
  y__b += bbar__1*(-real(1.0));
  z__b += bbar__1*(-x);
  lam_prev__[0] = ey__0 + x__b;
  lam_prev__[1] = lam__[1] + y__b;
  lam_prev__[2] = lam__[2] + z__b;
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
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  tape.put(0, -y + x*(rho - z));
  tape.put(1, x*y - beta*z);
  tape.put(2, -sigma);
  tape.put(3, -real(1.0));
  tape.put(4, -beta);
  const real hinv__ = m_rcp(h__);
  real ey__0, ga__0, gc__0;
  m_grl1_gate_seeds(real(tape.get(2)), x, real(y), lam__[0], h__, real(-1.0), ey__0, ga__0, gc__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(3)), m_rcp(real(tape.get(3))), real(tape.get(0)), lam__[1], h__, hinv__, real(-1.0), bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(tape.get(4)), m_rcp(real(tape.get(4))), real(tape.get(1)), lam__[2], h__, hinv__, real(-1.0), bbar__2, abar__2);
  tape.put(5, real(0.0));
  tape.put(6, ga__0);
  tape.put(7, real(0.0));
  tape.add(5, bbar__2*(y));
  tape.add(6, bbar__2*(x));
  tape.add(7, bbar__2*(-beta));
  tape.add(5, bbar__1*(rho - z));
  tape.add(6, bbar__1*(-real(1.0)));
  tape.add(7, bbar__1*(-x));
  lam_prev__[0] = ey__0 + tape.get(5);
  lam_prev__[1] = lam__[1] + tape.get(6);
  lam_prev__[2] = lam__[2] + tape.get(7);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real* __restrict__ seed__, real* __restrict__ y_bar__,
                                  real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real x = states[0];
  const real y = states[1];
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  const real rhs__0 = sigma*(y - x);
  const real rhs__1 = -y + x*(rho - z);
  const real rhs__2 = x*y - beta*z;
  real x__b = real(0.0);
  real y__b = real(0.0);
  real z__b = real(0.0);
  x__b += seed__[2]*(y);
  y__b += seed__[2]*(x);
  z__b += seed__[2]*(-beta);
  x__b += seed__[1]*(rho - z);
  y__b += seed__[1]*(-real(1.0));
  z__b += seed__[1]*(-x);
  x__b += seed__[0]*(-sigma);
  y__b += seed__[0]*(sigma);
  y_bar__[0] += x__b;
  y_bar__[1] += y__b;
  y_bar__[2] += z__b;
}

template <typename real, typename PA, typename treal, typename TP, typename SV, typename YV>
GOSS_HD GOSS_INLINE void eval_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                       const SV& seed__, const YV& y_bar__, real* __restrict__ fp_bar__,
                                       const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real x = states[0];
  const real y = states[1];
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  tape.put(0, sigma*(y - x));
  tape.put(1, -y + x*(rho - z));
  tape.put(2, x*y - beta*z);
  tape.put(3, real(0.0));
  tape.put(4, real(0.0));
  tape.put(5, real(0.0));
  tape.add(3, seed__[2]*(y));
  tape.add(4, seed__[2]*(x));
  tape.add(5, seed__[2]*(-beta));
  tape.add(3, seed__[1]*(rho - z));
  tape.add(4, seed__[1]*(-real(1.0)));
  tape.add(5, seed__[1]*(-x));
  tape.add(3, seed__[0]*(-sigma));
  tape.add(4, seed__[0]*(sigma));
  y_bar__[0] += tape.get(3);
  y_bar__[1] += tape.get(4);
  y_bar__[2] += tape.get(5);
}

constexpr int grl1_vjp_tape_slots = 8;
constexpr int rl1_vjp_tape_slots = 8;
constexpr int eval_vjp_tape_slots = 6;

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
// `jac` is a pointer or any object whose jac[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JAC>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, JAC jac)
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

// The analytic Jacobian over its structure: J[slot] for the entries of jac_cols / jac_col_start
// (d values[j] / d states[i] at jac_col_start[i] + rank of j in jac_cols[i]).  `J` is a pointer
// or any object whose J[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JS>
GOSS_HD GOSS_INLINE void compute_jacobian_structured(const real* __restrict__ states, const treal time,
                                                     const PA& pa, const JS& J)
{
  (void)time; (void)pa;
  const real x = states[0];
  const real y = states[1];
  const real z = states[2];
  const real beta = pa.template get<0>();
  const real rho = pa.template get<1>();
  const real sigma = pa.template get<2>();
  J[0] = -sigma;
  J[2] = y;
  J[3] = sigma;
  J[4] = -real(1.0);
  J[5] = x;
  J[7] = -beta;
  const real ddy_dt_dx = rho - z;
  J[1] = ddy_dt_dx;
  const real ddy_dt_dz = -x;
  J[6] = ddy_dt_dz;
}

}}  // namespace goss_models::lorentz
