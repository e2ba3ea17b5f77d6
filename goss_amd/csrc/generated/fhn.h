// Generated by goss_amd.codegen from fhn.ode -- do not edit.
// Variant fhn: field states ['v'], field parameters [].
#pragma once
#include "goss_model_support.h"

namespace goss_models { namespace fhn {

constexpr int num_states = 2;
constexpr int num_parameters = 11;
constexpr int num_constants = 7;
constexpr int num_field_states = 1;
constexpr int num_field_parameters = 0;
constexpr int field_state_indices[] = {0};
constexpr int field_parameter_indices[] = {-1};
static const char* const state_names[] = {"v", "s"};
static const char* const parameter_names[] = {"a", "b", "c_1", "c_2", "c_3", "stim_amplitude", "stim_duration", "stim_period", "stim_start", "v_peak", "v_rest"};
static const char* const field_state_names[] = {"v"};
static const char* const field_parameter_names[] = {nullptr};
constexpr double default_parameters[] = {0.13, 0.013, 0.26, 0.1, 1.0, 50.0, 1.0, 1000.0, 1.0, 40.0, -85.0};
constexpr double initial_conditions[] = {-85.0, 0.0};
constexpr unsigned char linear_terms[] = {0, 1};
// Diagonal of the mass matrix (ODE::_differential_states): 0 for the algebraic states of a DAE.
constexpr unsigned char differential_states[] = {1, 1};
constexpr bool is_dae = false;

// Subexpressions of shared (non-field) parameters only, evaluated once per launch.
// c[k] may use c[j] for j < k.  Called with real = double on the host.
template <typename real, typename PA>
GOSS_HD GOSS_INLINE void compute_constants(const PA& pa, real* __restrict__ c)
{
  (void)pa; (void)c;
  const real a = pa.template getd<0>();
  const real b = pa.template getd<1>();
  const real c_1 = pa.template getd<2>();
  const real c_2 = pa.template getd<3>();
  const real c_3 = pa.template getd<4>();
  const real stim_duration = pa.template getd<6>();
  const real stim_start = pa.template getd<8>();
  const real v_peak = pa.template getd<9>();
  const real v_rest = pa.template getd<10>();
  c[0] = v_peak - v_rest;
  c[1] = v_rest + a*c[0];
  c[2] = -m_div(c_2, c[0]);
  c[3] = m_div(c_1, m_sq(c[0]));
  c[4] = -real(5.0)*stim_duration - real(5.0)*stim_start;
  c[5] = -b*c_3;
  c[6] = -m_div(c_1, m_sq(c[0]));
}

// Right-hand side: values[i] = d states[i] / dt.  Intermediates that depend on time but on
// no state are computed in the type of `time` (treal).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval(const real* __restrict__ states, const treal time, const PA& pa,
                              real* __restrict__ values)
{
  (void)time; (void)pa;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real I = c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v);
  const treal i_Stim = m_div(pa.template getd<5>()*(treal(1.0) - m_rcp(treal(1.0) + m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>()))), (treal(1.0) + m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time)));
  values[0] = I + real(i_Stim);
  values[1] = b*(v - v_rest - c_3*s);
}

constexpr int num_intermediates = 4;
static const char* const intermediate_names[] = {"v_amp", "v_th", "I", "i_Stim"};
// Every named intermediate of the model at (states, time): the values a caller may monitor
// (ParameterizedODE::eval_monitored, ParameterizedODE.h:90-105, selects among them).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_intermediates(const real* __restrict__ states, const treal time, const PA& pa,
                                            real* __restrict__ out)
{
  (void)time; (void)pa; (void)states; (void)out;
  const real v = states[0];
  const real s = states[1];
  const real stim_amplitude = pa.template get<5>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__0 = pa.template cst<0>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  out[0] = real(c__0);
  out[1] = real(c__1);
  out[2] = real(c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v));
  out[3] = real(m_div(stim_amplitude*(real(1.0) - m_rcp(real(1.0) + m_exp_l2(real(7.213475204444817)*real(time) - real(7.213475204444817)*stim_start, real(5.0)*real(time) - real(5.0)*stim_start))), (real(1.0) + m_exp_l2(real(1.4426950408889634)*c__4 + real(7.213475204444817)*real(time), c__4 + real(5.0)*real(time)))));
}

// Structure of the Jacobian d values[j] / d states[i]: values[j] reads states[i] only for the
// entries below (through any operand of the emitted right-hand side), so every other entry of a
// forward-difference Jacobian is exactly zero.  And the structure of its no-pivot LU factors
// (the entries, the diagonal and the fill).
constexpr int jac_nnz = 4;
constexpr int lu_nnz = 4;
// jac_cols[i]: bit j set when values[j] reads states[i].  Jacobian storage is column-major over the
// structure: entry (j, i) at jac_col_start[i] + (number of bits of jac_cols[i] below j).
constexpr unsigned jac_cols[] = {0x3u, 0x3u};
constexpr int jac_col_start[] = {0, 2};
// LU storage is row-major over its structure: lu_slot[j * num_states + i], -1 outside.
constexpr short lu_slot[] = {0, 1, 2, 3};

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
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__5 = pa.template cst<5>();
  const real c__6 = pa.template cst<6>();
  const real I = c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v);
  const treal i_Stim = m_div(pa.template getd<5>()*(treal(1.0) - m_rcp(treal(1.0) + m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>()))), (treal(1.0) + m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time)));
  rhs[0] = I + real(i_Stim);
  rhs[1] = b*(v - v_rest - c_3*s);
  linear[1] = c__5;
  if (only_linear) return;
  const real dI_dv = c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) + c__6*(v - c__1)*(v - v_rest);
  linear[0] = dI_dv;
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
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__5 = pa.template cst<5>();
  const real c__6 = pa.template cst<6>();
  const real I = c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v);
  const treal i_Stim = m_div(pa.template getd<5>()*(treal(1.0) - m_rcp(treal(1.0) + m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>()))), (treal(1.0) + m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time)));
  rhs[0] = I + real(i_Stim);
  rhs[1] = b*(v - v_rest - c_3*s);
  const real dI_dv = c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) + c__6*(v - c__1)*(v - v_rest);
  linear[0] = dI_dv;
  linear[1] = c__5;
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
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  const real c__5 = pa.template cst<5>();
  const real c__6 = pa.template cst<6>();
  const real I = c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v);
  const treal tr__0 = m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>());
  const treal tr__1 = treal(1.0) + tr__0;
  const treal tr__2 = m_rcp(tr__1);
  const treal tr__3 = m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time);
  const treal tr__4 = treal(1.0) + tr__3;
  const treal tr__5 = m_rcp(tr__4);
  const treal i_Stim = pa.template getd<5>()*tr__5*(treal(1.0) - tr__2);
  const real dI_dv = c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) + c__6*(v - c__1)*(v - v_rest);
  const real rhs__0 = I + real(i_Stim);
  const real rhs__1 = b*(v - v_rest - c_3*s);
  const real lin__0 = dI_dv;
  const real lin__1 = c__5;
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(lin__0), m_rcp(real(lin__0)), real(rhs__0), lam__[0], h__, hinv__, delta__, bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  real I__b = real(0.0);
  real dI_dv__b = real(0.0);
  real s__b = real(0.0);
  real v__b = real(0.0);
  dI_dv__b += abar__0*(real(1.0));
  s__b += bbar__1*(-b*c_3);
  v__b += bbar__1*(b);
  I__b += bbar__0*(real(1.0));
  s__b += dI_dv__b*(c__2);
  const real cs__0 = v - c__1;
  const real cs__1 = v - v_rest;
  const real cs__2 = c__3*cs__0;
  const real cs__3 = v_peak - v;
  const real cs__4 = c__3*cs__1;
  v__b += dI_dv__b*(-cs__2 - cs__4 + c__6*cs__0 + c__6*cs__1 + real(2.0)*c__3*cs__3);
  s__b += I__b*(c__2*cs__1);
  v__b += I__b*(c__2*s + cs__2*cs__3 + cs__3*cs__4 - cs__1*cs__2);
  lam_prev__[0] = lam__[0] + v__b;
  lam_prev__[1] = lam__[1] + s__b;
}

template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>
GOSS_HD GOSS_INLINE void grl1_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const LV& lam__,
                                  const OV& lam_prev__, real* __restrict__ fp_bar__,
                                  const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  const real c__5 = pa.template cst<5>();
  const real c__6 = pa.template cst<6>();
  tape.put(0, c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v));
  const treal tr__0 = m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>());
  tape.put(1, real(tr__0));
  const treal tr__1 = treal(1.0) + tr__0;
  tape.put(2, real(tr__1));
  const treal tr__2 = m_rcp(tr__1);
  tape.put(3, real(tr__2));
  const treal tr__3 = m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time);
  tape.put(4, real(tr__3));
  const treal tr__4 = treal(1.0) + tr__3;
  tape.put(5, real(tr__4));
  const treal tr__5 = m_rcp(tr__4);
  tape.put(6, real(tr__5));
  const treal i_Stim = pa.template getd<5>()*tr__5*(treal(1.0) - tr__2);
  tape.put(7, real(i_Stim));
  tape.put(8, c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) + c__6*(v - c__1)*(v - v_rest));
  tape.put(9, tape.get(0) + tape.get(7));
  tape.put(10, b*(v - v_rest - c_3*s));
  tape.put(11, tape.get(8));
  tape.put(12, c__5);
  const real hinv__ = m_rcp(h__);
  real bbar__0, abar__0;
  m_grl1_seeds(real(tape.get(11)), m_rcp(real(tape.get(11))), real(tape.get(9)), lam__[0], h__, hinv__, delta__, bbar__0, abar__0);
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(12)), m_rcp(real(tape.get(12))), real(tape.get(10)), lam__[1], h__, hinv__, delta__, bbar__1, abar__1);
  tape.put(13, real(0.0));
  tape.put(14, real(0.0));
  tape.put(15, real(0.0));
  tape.put(16, real(0.0));
  tape.add(14, abar__0*(real(1.0)));
  tape.add(15, bbar__1*(-b*c_3));
  tape.add(16, bbar__1*(b));
  tape.add(13, bbar__0*(real(1.0)));
  tape.add(15, tape.get(14)*(c__2));
  tape.add(16, tape.get(14)*(c__6*(v - c__1) + c__6*(v - v_rest) - c__3*(v - c__1) - c__3*(v - v_rest) + real(2.0)*c__3*(v_peak - v)));
  tape.add(15, tape.get(13)*(c__2*(v - v_rest)));
  tape.add(16, tape.get(13)*(c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) - c__3*(v - c__1)*(v - v_rest)));
  lam_prev__[0] = lam__[0] + tape.get(16);
  lam_prev__[1] = lam__[1] + tape.get(15);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void rl1_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const real* __restrict__ lam__,
                                  real* __restrict__ lam_prev__, real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  const real c__5 = pa.template cst<5>();
  const real I = c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v);
  const treal tr__0 = m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>());
  const treal tr__1 = treal(1.0) + tr__0;
  const treal tr__2 = m_rcp(tr__1);
  const treal tr__3 = m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time);
  const treal tr__4 = treal(1.0) + tr__3;
  const treal tr__5 = m_rcp(tr__4);
  const treal i_Stim = pa.template getd<5>()*tr__5*(treal(1.0) - tr__2);
  const real rhs__0 = I + real(i_Stim);
  const real rhs__1 = b*(v - v_rest - c_3*s);
  const real lin__1 = c__5;
  const real hinv__ = m_rcp(h__);
  const real bbar__0 = lam__[0]*h__;
  real bbar__1, abar__1;
  m_grl1_seeds(real(lin__1), m_rcp(real(lin__1)), real(rhs__1), lam__[1], h__, hinv__, real(-1.0), bbar__1, abar__1);
  real I__b = real(0.0);
  real s__b = real(0.0);
  real v__b = real(0.0);
  s__b += bbar__1*(-b*c_3);
  v__b += bbar__1*(b);
  I__b += bbar__0*(real(1.0));
  const real cs__0 = v - v_rest;
  s__b += I__b*(c__2*cs__0);
  const real cs__1 = v - c__1;
  const real cs__2 = c__3*(v_peak - v);
  v__b += I__b*(c__2*s + cs__0*cs__2 + cs__1*cs__2 - c__3*cs__0*cs__1);
  lam_prev__[0] = lam__[0] + v__b;
  lam_prev__[1] = lam__[1] + s__b;
}

template <typename real, typename PA, typename treal, typename TP, typename LV, typename OV>
GOSS_HD GOSS_INLINE void rl1_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real h__, const real delta__, const LV& lam__,
                                  const OV& lam_prev__, real* __restrict__ fp_bar__,
                                  const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__; (void)delta__;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  const real c__5 = pa.template cst<5>();
  tape.put(0, c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v));
  const treal tr__0 = m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>());
  tape.put(1, real(tr__0));
  const treal tr__1 = treal(1.0) + tr__0;
  tape.put(2, real(tr__1));
  const treal tr__2 = m_rcp(tr__1);
  tape.put(3, real(tr__2));
  const treal tr__3 = m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time);
  tape.put(4, real(tr__3));
  const treal tr__4 = treal(1.0) + tr__3;
  tape.put(5, real(tr__4));
  const treal tr__5 = m_rcp(tr__4);
  tape.put(6, real(tr__5));
  const treal i_Stim = pa.template getd<5>()*tr__5*(treal(1.0) - tr__2);
  tape.put(7, real(i_Stim));
  tape.put(8, tape.get(0) + tape.get(7));
  tape.put(9, b*(v - v_rest - c_3*s));
  tape.put(10, c__5);
  const real hinv__ = m_rcp(h__);
  const real bbar__0 = lam__[0]*h__;
  real bbar__1, abar__1;
  m_grl1_seeds(real(tape.get(10)), m_rcp(real(tape.get(10))), real(tape.get(9)), lam__[1], h__, hinv__, real(-1.0), bbar__1, abar__1);
  tape.put(11, real(0.0));
  tape.put(12, real(0.0));
  tape.put(13, real(0.0));
  tape.add(12, bbar__1*(-b*c_3));
  tape.add(13, bbar__1*(b));
  tape.add(11, bbar__0*(real(1.0)));
  tape.add(12, tape.get(11)*(c__2*(v - v_rest)));
  tape.add(13, tape.get(11)*(c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) - c__3*(v - c__1)*(v - v_rest)));
  lam_prev__[0] = lam__[0] + tape.get(13);
  lam_prev__[1] = lam__[1] + tape.get(12);
}

template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval_vjp(const real* __restrict__ states, const treal time, const PA& pa,
                                  const real* __restrict__ seed__, real* __restrict__ y_bar__,
                                  real* __restrict__ fp_bar__)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  const real I = c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v);
  const treal tr__0 = m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>());
  const treal tr__1 = treal(1.0) + tr__0;
  const treal tr__2 = m_rcp(tr__1);
  const treal tr__3 = m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time);
  const treal tr__4 = treal(1.0) + tr__3;
  const treal tr__5 = m_rcp(tr__4);
  const treal i_Stim = pa.template getd<5>()*tr__5*(treal(1.0) - tr__2);
  const real rhs__0 = I + real(i_Stim);
  const real rhs__1 = b*(v - v_rest - c_3*s);
  real I__b = real(0.0);
  real s__b = real(0.0);
  real v__b = real(0.0);
  s__b += seed__[1]*(-b*c_3);
  v__b += seed__[1]*(b);
  I__b += seed__[0]*(real(1.0));
  const real cs__0 = v - v_rest;
  s__b += I__b*(c__2*cs__0);
  const real cs__1 = v - c__1;
  const real cs__2 = c__3*(v_peak - v);
  v__b += I__b*(c__2*s + cs__0*cs__2 + cs__1*cs__2 - c__3*cs__0*cs__1);
  y_bar__[0] += v__b;
  y_bar__[1] += s__b;
}

template <typename real, typename PA, typename treal, typename TP, typename SV, typename YV>
GOSS_HD GOSS_INLINE void eval_vjp_tape(const real* __restrict__ states, const treal time, const PA& pa,
                                       const SV& seed__, const YV& y_bar__, real* __restrict__ fp_bar__,
                                       const TP& tape)
{
  (void)time; (void)pa; (void)fp_bar__;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real c_3 = pa.template get<4>();
  const real stim_start = pa.template get<8>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__4 = pa.template cst<4>();
  tape.put(0, c__2*s*(v - v_rest) + c__3*(v - c__1)*(v - v_rest)*(v_peak - v));
  const treal tr__0 = m_exp_l2(treal(7.213475204444817)*time - treal(7.213475204444817)*pa.template getd<8>(), treal(5.0)*time - treal(5.0)*pa.template getd<8>());
  tape.put(1, real(tr__0));
  const treal tr__1 = treal(1.0) + tr__0;
  tape.put(2, real(tr__1));
  const treal tr__2 = m_rcp(tr__1);
  tape.put(3, real(tr__2));
  const treal tr__3 = m_exp_l2(treal(1.4426950408889634)*pa.template cstd<4>() + treal(7.213475204444817)*time, pa.template cstd<4>() + treal(5.0)*time);
  tape.put(4, real(tr__3));
  const treal tr__4 = treal(1.0) + tr__3;
  tape.put(5, real(tr__4));
  const treal tr__5 = m_rcp(tr__4);
  tape.put(6, real(tr__5));
  const treal i_Stim = pa.template getd<5>()*tr__5*(treal(1.0) - tr__2);
  tape.put(7, real(i_Stim));
  tape.put(8, tape.get(0) + tape.get(7));
  tape.put(9, b*(v - v_rest - c_3*s));
  tape.put(10, real(0.0));
  tape.put(11, real(0.0));
  tape.put(12, real(0.0));
  tape.add(11, seed__[1]*(-b*c_3));
  tape.add(12, seed__[1]*(b));
  tape.add(10, seed__[0]*(real(1.0)));
  tape.add(11, tape.get(10)*(c__2*(v - v_rest)));
  tape.add(12, tape.get(10)*(c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) - c__3*(v - c__1)*(v - v_rest)));
  y_bar__[0] += tape.get(12);
  y_bar__[1] += tape.get(11);
}

constexpr int grl1_vjp_tape_slots = 17;
constexpr int rl1_vjp_tape_slots = 14;
constexpr int eval_vjp_tape_slots = 13;

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
// `jac` is a pointer or any object whose jac[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JAC>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, JAC jac)
{
  (void)time; (void)pa;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__5 = pa.template cst<5>();
  const real c__6 = pa.template cst<6>();
  const real dI_dv = c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) + c__6*(v - c__1)*(v - v_rest);
  const real dI_ds = c__2*(v - v_rest);
  jac[0] = dI_dv;
  jac[1] = dI_ds;
  jac[2] = b;
  jac[3] = c__5;
}

// The analytic Jacobian over its structure: J[slot] for the entries of jac_cols / jac_col_start
// (d values[j] / d states[i] at jac_col_start[i] + rank of j in jac_cols[i]).  `J` is a pointer
// or any object whose J[k] = value stores entry k.
template <typename real, typename PA, typename treal, typename JS>
GOSS_HD GOSS_INLINE void compute_jacobian_structured(const real* __restrict__ states, const treal time,
                                                     const PA& pa, const JS& J)
{
  (void)time; (void)pa;
  const real v = states[0];
  const real s = states[1];
  const real b = pa.template get<1>();
  const real v_peak = pa.template get<9>();
  const real v_rest = pa.template get<10>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real c__3 = pa.template cst<3>();
  const real c__5 = pa.template cst<5>();
  const real c__6 = pa.template cst<6>();
  J[1] = b;
  J[3] = c__5;
  const real dI_dv = c__2*s + c__3*(v - c__1)*(v_peak - v) + c__3*(v - v_rest)*(v_peak - v) + c__6*(v - c__1)*(v - v_rest);
  J[0] = dI_dv;
  const real dI_ds = c__2*(v - v_rest);
  J[2] = dI_ds;
}

}}  // namespace goss_models::fhn
