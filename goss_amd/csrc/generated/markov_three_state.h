// Generated by goss_amd.codegen from markov_three_state.ode -- do not edit.
// Variant markov_three_state: field states ['O'], field parameters ['k_oi'].
#pragma once
#include "goss_model_support.h"

namespace goss_models { namespace markov_three_state {

constexpr int num_states = 3;
constexpr int num_parameters = 6;
constexpr int num_constants = 0;
constexpr int num_field_states = 1;
constexpr int num_field_parameters = 1;
constexpr int field_state_indices[] = {1};
constexpr int field_parameter_indices[] = {4};
static const char* const state_names[] = {"C", "O", "I"};
static const char* const parameter_names[] = {"a_co", "k_co", "k_io", "k_oc", "k_oi", "w_co"};
static const char* const field_state_names[] = {"O"};
static const char* const field_parameter_names[] = {"k_oi"};
constexpr double default_parameters[] = {0.5, 2.0, 0.1, 0.5, 0.3, 3.0};
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
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  values[0] = O*k_oc - C*real(rates_C_O);
  values[1] = C*real(rates_C_O) + I*k_io - O*(k_oc + rates_O_I);
  values[2] = O*rates_O_I - I*k_io;
}

constexpr int num_intermediates = 5;
static const char* const intermediate_names[] = {"k_co_t", "rates_C_O", "rates_O_C", "rates_O_I", "rates_I_O"};
// Every named intermediate of the model at (states, time): the values a caller may monitor
// (ParameterizedODE::eval_monitored, ParameterizedODE.h:90-105, selects among them).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_intermediates(const real* __restrict__ states, const treal time, const PA& pa,
                                            real* __restrict__ out)
{
  (void)time; (void)pa; (void)states; (void)out;
  const real a_co = pa.template get<0>();
  const real k_co = pa.template get<1>();
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const real w_co = pa.template get<5>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  out[0] = real(k_co*(real(1.0) + a_co*m_sin(real(time)*w_co)));
  out[1] = real(real(k_co_t));
  out[2] = real(k_oc);
  out[3] = real(k_oi);
  out[4] = real(k_io);
}

// Structure of the Jacobian d values[j] / d states[i]: values[j] reads states[i] only for the
// entries below (through any operand of the emitted right-hand side), so every other entry of a
// forward-difference Jacobian is exactly zero.  And the structure of its no-pivot LU factors
// (the entries, the diagonal and the fill).
constexpr int jac_nnz = 7;
constexpr int lu_nnz = 7;
// jac_cols[i]: bit j set when values[j] reads states[i].  Jacobian storage is column-major over the
// structure: entry (j, i) at jac_col_start[i] + (number of bits of jac_cols[i] below j).
constexpr unsigned jac_cols[] = {0x3u, 0x7u, 0x6u};
constexpr int jac_col_start[] = {0, 2, 5};
// LU storage is row-major over its structure: lu_slot[j * num_states + i], -1 outside.
constexpr short lu_slot[] = {0, 1, -1, 2, 3, 4, -1, 5, 6};

// a = M + scale * J on the LU structure (fill entries zero), M the diagonal mass matrix
// (differential_states): the dense scale and add_mass_matrix of
// ImplicitODESolver::compute_factorized_jacobian (ImplicitODESolver.cpp:92-105, 258-262).  The
// algebraic states of a DAE add 0.0, as the reference's weight * 0 does.
template <typename real>
GOSS_HD GOSS_INLINE void lu_from_jacobian(const real* __restrict__ J, const real scale, real* __restrict__ a)
{
  a[0] = J[0] * scale + real(1.0);
  a[1] = J[2] * scale;
  a[2] = J[1] * scale;
  a[3] = J[3] * scale + real(1.0);
  a[4] = J[5] * scale;
  a[5] = J[4] * scale;
  a[6] = J[6] * scale + real(1.0);
}

// ODE::lu_factorize (ODE.cpp:69-100) in place on a, over the structure.
template <typename real>
GOSS_HD GOSS_INLINE void lu_factorize_sparse(real* __restrict__ a)
{
  real acc;
  a[2] = a[2] / a[0];
  acc = a[2] * a[1];
  a[3] = a[3] - acc;
  a[5] = a[5] / a[3];
  acc = a[5] * a[4];
  a[6] = a[6] - acc;
  (void)acc;
}

// ODE::forward_backward_subst (ODE.cpp:102-132) in place on x, over the structure.
template <typename real>
GOSS_HD GOSS_INLINE void lu_solve_sparse(const real* __restrict__ a, real* __restrict__ x)
{
  real acc;
  acc = a[2] * x[0];
  x[1] = x[1] - acc;
  acc = a[5] * x[1];
  x[2] = x[2] - acc;
  x[2] = x[2] / a[6];
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
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  rhs[0] = O*k_oc - C*real(rates_C_O);
  rhs[1] = C*real(rates_C_O) + I*k_io - O*(k_oc + rates_O_I);
  rhs[2] = O*rates_O_I - I*k_io;
  const treal ddC_dt_dC = -rates_C_O;
  const real ddO_dt_dO = -k_oc - rates_O_I;
  linear[0] = real(ddC_dt_dC);
  linear[1] = ddO_dt_dO;
  linear[2] = -k_io;
  if (only_linear) return;
}

// linearized_eval (all entries) plus, for the gate states (gate_states[i] = 1: rhs[i] = (A - y_i) r),
// gate_target[i] = A.  A GRL stage of a gate is then y0_i + (y_i - A) expm1(linear[i] h).
constexpr unsigned char gate_states[] = {0, 0, 0};
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void linearized_eval_gates(const real* __restrict__ states, const treal time, const PA& pa,
                                               real* __restrict__ linear, real* __restrict__ rhs,
                                               real* __restrict__ gate_target)
{
  (void)time; (void)pa; (void)gate_target;
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  rhs[0] = O*k_oc - C*real(rates_C_O);
  rhs[1] = C*real(rates_C_O) + I*k_io - O*(k_oc + rates_O_I);
  rhs[2] = O*rates_O_I - I*k_io;
  const treal ddC_dt_dC = -rates_C_O;
  const real ddO_dt_dO = -k_oc - rates_O_I;
  linear[0] = real(ddC_dt_dC);
  linear[1] = ddO_dt_dO;
  linear[2] = -k_io;
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
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  const treal ddC_dt_dC = -rates_C_O;
  const real ddO_dt_dO = -k_oc - rates_O_I;
  const real rhs__0 = O*k_oc - C*real(rates_C_O);
  const real rhs__1 = C*real(rates_C_O) + I*k_io - O*(k_oc + rates_O_I);
  const real rhs__2 = O*rates_O_I - I*k_io;
  const treal lin__0 = ddC_dt_dC;
  const real lin__1 = ddO_dt_dO;
  const real lin__2 = -k_io;
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(lin__0), m_rcp(real(lin__0)), real(rhs__0), lam__[0], h__, hinv__, delta__, bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(lin__2), m_rcp(real(lin__2)), real(rhs__2), lam__[2], h__, hinv__, delta__, bbar__2, abar__2);
  real C__b = real(0.0);
  real I__b = real(0.0);
  real O__b = real(0.0);
  real ddO_dt_dO__b = real(0.0);
  real k_oi__b = real(0.0);
  real rates_O_I__b = real(0.0);
  ddO_dt_dO__b += abar__1*(real(1.0));
  I__b += bbar__2*(-k_io);
  O__b += bbar__2*(rates_O_I);
  rates_O_I__b += bbar__2*(O);
  C__b += bbar__1*(real(rates_C_O));
  I__b += bbar__1*(k_io);
  O__b += bbar__1*(-k_oc - rates_O_I);
  rates_O_I__b += bbar__1*(-O);
  C__b += bbar__0*(-real(rates_C_O));
  O__b += bbar__0*(k_oc);
  rates_O_I__b += ddO_dt_dO__b*(-real(1.0));
  k_oi__b += rates_O_I__b*(real(1.0));
  lam_prev__[0] = lam__[0] + C__b;
  lam_prev__[1] = lam__[1] + O__b;
  lam_prev__[2] = lam__[2] + I__b;
  fp_bar__[0] += k_oi__b;
}

template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>
GOSS_HD GOSS_INLINE void grl1_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const LV& lam__,
                                  const OV& lam_prev__, real* __restrict__ fp_bar__,
                                  const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  tape.put(0, real(k_co_t));
  const treal rates_C_O = k_co_t;
  tape.put(1, real(rates_C_O));
  tape.put(2, k_oi);
  const treal ddC_dt_dC = -rates_C_O;
  tape.put(3, real(ddC_dt_dC));
  tape.put(4, -k_oc - tape.get(2));
  tape.put(5, O*k_oc - C*tape.get(1));
  tape.put(6, C*tape.get(1) + I*k_io - O*(k_oc + tape.get(2)));
  tape.put(7, O*tape.get(2) - I*k_io);
  const treal lin__0 = ddC_dt_dC;
  tape.put(8, real(lin__0));
  tape.put(9, tape.get(4));
  tape.put(10, -k_io);
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(tape.get(8)), m_rcp(real(tape.get(8))), real(tape.get(5)), lam__[0], h__, hinv__, delta__, bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(9)), m_rcp(real(tape.get(9))), real(tape.get(6)), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(tape.get(10)), m_rcp(real(tape.get(10))), real(tape.get(7)), lam__[2], h__, hinv__, delta__, bbar__2, abar__2);
  tape.put(11, real(0.0));
  tape.put(12, real(0.0));
  tape.put(13, real(0.0));
  tape.put(14, real(0.0));
  tape.put(15, real(0.0));
  tape.put(16, real(0.0));
  tape.add(14, abar__1*(real(1.0)));
  tape.add(12, bbar__2*(-k_io));
  tape.add(13, bbar__2*(tape.get(2)));
  tape.add(16, bbar__2*(O));
  tape.add(11, bbar__1*(tape.get(1)));
  tape.add(12, bbar__1*(k_io));
  tape.add(13, bbar__1*(-k_oc - tape.get(2)));
  tape.add(16, bbar__1*(-O));
  tape.add(11, bbar__0*(-tape.get(1)));
  tape.add(13, bbar__0*(k_oc));
  tape.add(16, tape.get(14)*(-real(1.0)));
  tape.add(15, tape.get(16)*(real(1.0)));
  lam_prev__[0] = lam__[0] + tape.get(11);
  lam_prev__[1] = lam__[1] + tape.get(13);
  lam_prev__[2] = lam__[2] + tape.get(12);
  fp_bar__[0] += tape.get(15);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void rl1_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const real* __restrict__ lam__,
                                  real* __restrict__ lam_prev__, real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  const treal ddC_dt_dC = -rates_C_O;
  const real ddO_dt_dO = -k_oc - rates_O_I;
  const real rhs__0 = O*k_oc - C*real(rates_C_O);
  const real rhs__1 = C*real(rates_C_O) + I*k_io - O*(k_oc + rates_O_I);
  const real rhs__2 = O*rates_O_I - I*k_io;
  const treal lin__0 = ddC_dt_dC;
  const real lin__1 = ddO_dt_dO;
  const real lin__2 = -k_io;
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(lin__0), m_rcp(real(lin__0)), real(rhs__0), lam__[0], h__, hinv__, real(-1.0), bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, real(-1.0), bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(lin__2), m_rcp(real(lin__2)), real(rhs__2), lam__[2], h__, hinv__, real(-1.0), bbar__2, abar__2);
  real C__b = real(0.0);
  real I__b = real(0.0);
  real O__b = real(0.0);
  real ddO_dt_dO__b = real(0.0);
  real k_oi__b = real(0.0);
  real rates_O_I__b = real(0.0);
  ddO_dt_dO__b += abar__1*(real(1.0));
  I__b += bbar__2*(-k_io);
  O__b += bbar__2*(rates_O_I);
  rates_O_I__b += bbar__2*(O);
  C__b += bbar__1*(real(rates_C_O));
  I__b += bbar__1*(k_io);
  O__b += bbar__1*(-k_oc - rates_O_I);
  rates_O_I__b += bbar__1*(-O);
  C__b += bbar__0*(-real(rates_C_O));
  O__b += bbar__0*(k_oc);
  rates_O_I__b += ddO_dt_dO__b*(-real(1.0));
  k_oi__b += rates_O_I__b*(real(1.0));
  lam_prev__[0] = lam__[0] + C__b;
  lam_prev__[1] = lam__[1] + O__b;
  lam_prev__[2] = lam__[2] + I__b;
  fp_bar__[0] += k_oi__b;
}

template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>
GOSS_HD GOSS_INLINE void rl1_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const LV& lam__,
                                  const OV& lam_prev__, real* __restrict__ fp_bar__,
                                  const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  tape.put(0, real(k_co_t));
  const treal rates_C_O = k_co_t;
  tape.put(1, real(rates_C_O));
  tape.put(2, k_oi);
  const treal ddC_dt_dC = -rates_C_O;
  tape.put(3, real(ddC_dt_dC));
  tape.put(4, -k_oc - tape.get(2));
  tape.put(5, O*k_oc - C*tape.get(1));
  tape.put(6, C*tape.get(1) + I*k_io - O*(k_oc + tape.get(2)));
  tape.put(7, O*tape.get(2) - I*k_io);
  const treal lin__0 = ddC_dt_dC;
  tape.put(8, real(lin__0));
  tape.put(9, tape.get(4));
  tape.put(10, -k_io);
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(tape.get(8)), m_rcp(real(tape.get(8))), real(tape.get(5)), lam__[0], h__, hinv__, real(-1.0), bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(9)), m_rcp(real(tape.get(9))), real(tape.get(6)), lam__[1], h__, hinv__, real(-1.0), bbar__1, abar__1);
  real bbar__2, abar__2;
  m_grl1_seeds(real(tape.get(10)), m_rcp(real(tape.get(10))), real(tape.get(7)), lam__[2], h__, hinv__, real(-1.0), bbar__2, abar__2);
  tape.put(11, real(0.0));
  tape.put(12, real(0.0));
  tape.put(13, real(0.0));
  tape.put(14, real(0.0));
  tape.put(15, real(0.0));
  tape.put(16, real(0.0));
  tape.add(14, abar__1*(real(1.0)));
  tape.add(12, bbar__2*(-k_io));
  tape.add(13, bbar__2*(tape.get(2)));
  tape.add(16, bbar__2*(O));
  tape.add(11, bbar__1*(tape.get(1)));
  tape.add(12, bbar__1*(k_io));
  tape.add(13, bbar__1*(-k_oc - tape.get(2)));
  tape.add(16, bbar__1*(-O));
  tape.add(11, bbar__0*(-tape.get(1)));
  tape.add(13, bbar__0*(k_oc));
  tape.add(16, tape.get(14)*(-real(1.0)));
  tape.add(15, tape.get(16)*(real(1.0)));
  lam_prev__[0] = lam__[0] + tape.get(11);
  lam_prev__[1] = lam__[1] + tape.get(13);
  lam_prev__[2] = lam__[2] + tape.get(12);
  fp_bar__[0] += tape.get(15);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real* __restrict__ seed__, real* __restrict__ y_bar__,
                                  real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  const real rhs__0 = O*k_oc - C*real(rates_C_O);
  const real rhs__1 = C*real(rates_C_O) + I*k_io - O*(k_oc + rates_O_I);
  const real rhs__2 = O*rates_O_I - I*k_io;
  real C__b = real(0.0);
  real I__b = real(0.0);
  real O__b = real(0.0);
  real k_oi__b = real(0.0);
  real rates_O_I__b = real(0.0);
  I__b += seed__[2]*(-k_io);
  O__b += seed__[2]*(rates_O_I);
  rates_O_I__b += seed__[2]*(O);
  C__b += seed__[1]*(real(rates_C_O));
  I__b += seed__[1]*(k_io);
  O__b += seed__[1]*(-k_oc - rates_O_I);
  rates_O_I__b += seed__[1]*(-O);
  C__b += seed__[0]*(-real(rates_C_O));
  O__b += seed__[0]*(k_oc);
  k_oi__b += rates_O_I__b*(real(1.0));
  y_bar__[0] += C__b;
  y_bar__[1] += O__b;
  y_bar__[2] += I__b;
  fp_bar__[0] += k_oi__b;
}

template <typename real, typename PA, typename treal, typename TP, typename SV, typename YV>
GOSS_HD GOSS_INLINE void eval_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                       const SV& seed__, const YV& y_bar__, real* __restrict__ fp_bar__,
                                       const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real C = states[0];
  const real O = states[1];
  const real I = states[2];
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  tape.put(0, real(k_co_t));
  const treal rates_C_O = k_co_t;
  tape.put(1, real(rates_C_O));
  tape.put(2, k_oi);
  tape.put(3, O*k_oc - C*tape.get(1));
  tape.put(4, C*tape.get(1) + I*k_io - O*(k_oc + tape.get(2)));
  tape.put(5, O*tape.get(2) - I*k_io);
  tape.put(6, real(0.0));
  tape.put(7, real(0.0));
  tape.put(8, real(0.0));
  tape.put(9, real(0.0));
  tape.put(10, real(0.0));
  tape.add(7, seed__[2]*(-k_io));
  tape.add(8, seed__[2]*(tape.get(2)));
  tape.add(10, seed__[2]*(O));
  tape.add(6, seed__[1]*(tape.get(1)));
  tape.add(7, seed__[1]*(k_io));
  tape.add(8, seed__[1]*(-k_oc - tape.get(2)));
  tape.add(10, seed__[1]*(-O));
  tape.add(6, seed__[0]*(-tape.get(1)));
  tape.add(8, seed__[0]*(k_oc));
  tape.add(9, tape.get(10)*(real(1.0)));
  y_bar__[0] += tape.get(6);
  y_bar__[1] += tape.get(8);
  y_bar__[2] += tape.get(7);
  fp_bar__[0] += tape.get(9);
}

constexpr int grl1_vjp_tape_slots = 17;
constexpr int rl1_vjp_tape_slots = 17;
constexpr int eval_vjp_tape_slots = 11;

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
// `jac` is a pointer or any object whose jac[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JAC>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, JAC jac)
{
  (void)time; (void)pa;
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  const real rates_O_I = k_oi;
  const treal ddC_dt_dC = -rates_C_O;
  const real ddO_dt_dO = -k_oc - rates_O_I;
  jac[0] = real(ddC_dt_dC);
  jac[1] = k_oc;
  jac[2] = real(0.0);
  jac[3] = real(rates_C_O);
  jac[4] = ddO_dt_dO;
  jac[5] = k_io;
  jac[6] = real(0.0);
  jac[7] = rates_O_I;
  jac[8] = -k_io;
}

// The analytic Jacobian over its structure: J[slot] for the entries of jac_cols / jac_col_start
// (d values[j] / d states[i] at jac_col_start[i] + rank of j in jac_cols[i]).  `J` is a pointer
// or any object whose J[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JS>
GOSS_HD GOSS_INLINE void compute_jacobian_structured(const real* __restrict__ states, const treal time,
                                                     const PA& pa, const JS& J)
{
  (void)time; (void)pa;
  const real k_io = pa.template get<2>();
  const real k_oc = pa.template get<3>();
  const real k_oi = pa.template get<4>();
  J[2] = k_oc;
  J[5] = k_io;
  J[6] = -k_io;
  const treal k_co_t = pa.template getd<1>()*(treal(1.0) + pa.template getd<0>()*m_sin(time*pa.template getd<5>()));
  const treal rates_C_O = k_co_t;
  J[1] = real(rates_C_O);
  const real rates_O_I = k_oi;
  J[4] = rates_O_I;
  const treal ddC_dt_dC = -rates_C_O;
  J[0] = real(ddC_dt_dC);
  const real ddO_dt_dO = -k_oc - rates_O_I;
  J[3] = ddO_dt_dO;
}

}}  // namespace goss_models::markov_three_state
