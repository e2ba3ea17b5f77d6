// Generated by goss_amd.codegen from beeler_reuter_1977.ode -- do not edit.
// Variant beeler_reuter_1977: field states ['V'], field parameters [].
#pragma once
#include "goss_model_support.h"

namespace goss_models { namespace beeler_reuter_1977 {

constexpr int num_states = 8;
constexpr int num_parameters = 10;
constexpr int num_constants = 3;
constexpr int num_field_states = 1;
constexpr int num_field_parameters = 0;
constexpr int field_state_indices[] = {7};
constexpr int field_parameter_indices[] = {-1};
static const char* const state_names[] = {"m", "h", "j", "Cai", "d", "f", "x1", "V"};
static const char* const parameter_names[] = {"E_Na", "g_Na", "g_Nac", "g_s", "IstimAmplitude", "IstimEnd", "IstimPeriod", "IstimPulseDuration", "IstimStart", "C"};
static const char* const field_state_names[] = {"V"};
static const char* const field_parameter_names[] = {nullptr};
constexpr double default_parameters[] = {50.0, 0.04, 3e-05, 0.0005, 0.5, 50000.0, 1000.0, 1.0, 10.0, 0.01};
constexpr double initial_conditions[] = {0.011, 0.988, 0.975, 0.0001, 0.003, 0.994, 0.0001, -84.624};
constexpr unsigned char linear_terms[] = {1, 1, 1, 0, 1, 1, 1, 0};

// Subexpressions of shared (non-field) parameters only, evaluated once per launch.
// c[k] may use c[j] for j < k.  Called with real = double on the host.
template <typename real, typename PA>
GOSS_HD GOSS_INLINE void compute_constants(const PA& pa, real* __restrict__ c)
{
  (void)pa; (void)c;
  const real IstimPeriod = pa.template getd<6>();
  const real C = pa.template getd<9>();
  c[0] = m_rcp(IstimPeriod);
  c[1] = m_rcp(C);
  c[2] = -m_rcp(C);
}

// Right-hand side: values[i] = d states[i] / dt.  Intermediates that depend on time but on
// no state are computed in the type of `time` (treal).
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void eval(const real* __restrict__ states, const treal time, const PA& pa,
                              real* __restrict__ values)
{
  (void)time; (void)pa;
  const real m = states[0];
  const real h = states[1];
  const real j = states[2];
  const real Cai = states[3];
  const real d = states[4];
  const real f = states[5];
  const real x1 = states[6];
  const real V = states[7];
  const real E_Na = pa.template get<0>();
  const real g_Na = pa.template get<1>();
  const real g_Nac = pa.template get<2>();
  const real g_s = pa.template get<3>();
  const real c__1 = pa.template cst<1>();
  const real i_Na = (V - E_Na)*(g_Nac + g_Na*h*j*m_cube(m));
  const real V_eff_0 = (m_fabs(real(47.0) + V) < real(0.01) ? real(0.01) : real(47.0) + V);
  const real alpha_m = m_div((-real(47.0) - V), (-real(1.0) + m_exp(-real(0.1)*V_eff_0)));
  const real beta_m = real(0.7095526727489909)*m_exp(-real(0.056)*V);
  const real alpha_h = real(5.497962438709065e-10)*m_exp(-real(0.25)*V);
  const real beta_h = m_div(real(1.7), (real(1.0) + real(0.1580253208896478)*m_exp(-real(0.082)*V)));
  const real alpha_j = m_div(real(1.8690473007222892e-10)*m_exp(-real(0.25)*V), (real(1.0) + real(1.6788275299956603e-07)*m_exp(-real(0.2)*V)));
  const real beta_j = m_div(real(0.3), (real(1.0) + real(0.040762203978366204)*m_exp(-real(0.1)*V)));
  const real E_s = -real(82.3) - real(13.0287)*m_log(real(0.001)*Cai);
  const real i_s = d*f*g_s*(V - E_s);
  const real alpha_d = m_div(real(0.095)*m_exp(real(0.05) - real(0.01)*V), (real(1.0) + real(1.4332881385696572)*m_exp(-real(0.07199424046076314)*V)));
  const real beta_d = m_div(real(0.07)*m_exp(-real(0.7457627118644068) - real(0.01694915254237288)*V), (real(1.0) + m_exp(real(2.2) + real(0.05)*V)));
  const real alpha_f = m_div(real(0.012)*m_exp(-real(0.224) - real(0.008)*V), (real(1.0) + real(66.5465065250986)*m_exp(real(0.14992503748125938)*V)));
  const real beta_f = m_div(real(0.0065)*m_exp(-real(0.6) - real(0.02)*V), (real(1.0) + m_exp(-real(6.0) - real(0.2)*V)));
  const real i_x1 = real(0.0019727757115328517)*x1*(-real(1.0) + real(21.75840239619708)*m_exp(real(0.04)*V))*m_exp(-real(0.04)*V);
  const real alpha_x1 = m_div(real(0.031158410986342627)*m_exp(real(0.08264462809917356)*V), (real(1.0) + real(17.41170806332765)*m_exp(real(0.05714285714285714)*V)));
  const real beta_x1 = m_div(real(0.0003916464405623223)*m_exp(-real(0.05998800239952009)*V), (real(1.0) + m_exp(-real(0.8) - real(0.04)*V)));
  const real V_eff_1 = (m_fabs(real(23.0) + V) < real(0.01) ? real(0.01) : real(23.0) + V);
  const real i_K1 = m_div(real(0.0035)*(real(4.6000000000000005) + real(0.2)*V), (real(1.0) - m_exp(-real(0.04)*V_eff_1))) + m_div(real(0.0035)*(-real(4.0) + real(119.85640018958804)*m_exp(real(0.04)*V)), (real(8.331137487687693)*m_exp(real(0.04)*V) + real(69.4078518387552)*m_exp(real(0.08)*V)));
  const treal Istim = (pa.template getd<5>() >= time && pa.template getd<8>() <= time && pa.template getd<7>() >= time - pa.template getd<8>() - pa.template getd<6>()*m_floor(pa.template cstd<0>()*(time - pa.template getd<8>())) ? pa.template getd<4>() : treal(0.0));
  values[0] = alpha_m*(real(1.0) - m) - beta_m*m;
  values[1] = alpha_h*(real(1.0) - h) - beta_h*h;
  values[2] = alpha_j*(real(1.0) - j) - beta_j*j;
  values[3] = real(7.000000000000001e-06) - real(0.07)*Cai - real(0.01)*i_s;
  values[4] = alpha_d*(real(1.0) - d) - beta_d*d;
  values[5] = alpha_f*(real(1.0) - f) - beta_f*f;
  values[6] = alpha_x1*(real(1.0) - x1) - beta_x1*x1;
  values[7] = c__1*(real(Istim) - i_K1 - i_Na - i_s - i_x1);
}

// Linearized right-hand side: rhs[i] as in eval and linear[i] = d rhs[i] / d states[i].
// Entries with linear_terms[i] == 1 are written before the only_linear return.
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void linearized_eval(const real* __restrict__ states, const treal time, const PA& pa,
                                         real* __restrict__ linear, real* __restrict__ rhs,
                                         const bool only_linear)
{
  (void)time; (void)pa;
  const real m = states[0];
  const real h = states[1];
  const real j = states[2];
  const real Cai = states[3];
  const real d = states[4];
  const real f = states[5];
  const real x1 = states[6];
  const real V = states[7];
  const real E_Na = pa.template get<0>();
  const real g_Na = pa.template get<1>();
  const real g_Nac = pa.template get<2>();
  const real g_s = pa.template get<3>();
  const real c__1 = pa.template cst<1>();
  const real c__2 = pa.template cst<2>();
  const real i_Na = (V - E_Na)*(g_Nac + g_Na*h*j*m_cube(m));
  const real V_eff_0 = (m_fabs(real(47.0) + V) < real(0.01) ? real(0.01) : real(47.0) + V);
  const real alpha_m = m_div((-real(47.0) - V), (-real(1.0) + m_exp(-real(0.1)*V_eff_0)));
  const real beta_m = real(0.7095526727489909)*m_exp(-real(0.056)*V);
  const real alpha_h = real(5.497962438709065e-10)*m_exp(-real(0.25)*V);
  const real beta_h = m_div(real(1.7), (real(1.0) + real(0.1580253208896478)*m_exp(-real(0.082)*V)));
  const real alpha_j = m_div(real(1.8690473007222892e-10)*m_exp(-real(0.25)*V), (real(1.0) + real(1.6788275299956603e-07)*m_exp(-real(0.2)*V)));
  const real beta_j = m_div(real(0.3), (real(1.0) + real(0.040762203978366204)*m_exp(-real(0.1)*V)));
  const real E_s = -real(82.3) - real(13.0287)*m_log(real(0.001)*Cai);
  const real i_s = d*f*g_s*(V - E_s);
  const real alpha_d = m_div(real(0.095)*m_exp(real(0.05) - real(0.01)*V), (real(1.0) + real(1.4332881385696572)*m_exp(-real(0.07199424046076314)*V)));
  const real beta_d = m_div(real(0.07)*m_exp(-real(0.7457627118644068) - real(0.01694915254237288)*V), (real(1.0) + m_exp(real(2.2) + real(0.05)*V)));
  const real alpha_f = m_div(real(0.012)*m_exp(-real(0.224) - real(0.008)*V), (real(1.0) + real(66.5465065250986)*m_exp(real(0.14992503748125938)*V)));
  const real beta_f = m_div(real(0.0065)*m_exp(-real(0.6) - real(0.02)*V), (real(1.0) + m_exp(-real(6.0) - real(0.2)*V)));
  const real i_x1 = real(0.0019727757115328517)*x1*(-real(1.0) + real(21.75840239619708)*m_exp(real(0.04)*V))*m_exp(-real(0.04)*V);
  const real alpha_x1 = m_div(real(0.031158410986342627)*m_exp(real(0.08264462809917356)*V), (real(1.0) + real(17.41170806332765)*m_exp(real(0.05714285714285714)*V)));
  const real beta_x1 = m_div(real(0.0003916464405623223)*m_exp(-real(0.05998800239952009)*V), (real(1.0) + m_exp(-real(0.8) - real(0.04)*V)));
  const real V_eff_1 = (m_fabs(real(23.0) + V) < real(0.01) ? real(0.01) : real(23.0) + V);
  const real i_K1 = m_div(real(0.0035)*(real(4.6000000000000005) + real(0.2)*V), (real(1.0) - m_exp(-real(0.04)*V_eff_1))) + m_div(real(0.0035)*(-real(4.0) + real(119.85640018958804)*m_exp(real(0.04)*V)), (real(8.331137487687693)*m_exp(real(0.04)*V) + real(69.4078518387552)*m_exp(real(0.08)*V)));
  const treal Istim = (pa.template getd<5>() >= time && pa.template getd<8>() <= time && pa.template getd<7>() >= time - pa.template getd<8>() - pa.template getd<6>()*m_floor(pa.template cstd<0>()*(time - pa.template getd<8>())) ? pa.template getd<4>() : treal(0.0));
  rhs[0] = alpha_m*(real(1.0) - m) - beta_m*m;
  rhs[1] = alpha_h*(real(1.0) - h) - beta_h*h;
  rhs[2] = alpha_j*(real(1.0) - j) - beta_j*j;
  rhs[3] = real(7.000000000000001e-06) - real(0.07)*Cai - real(0.01)*i_s;
  rhs[4] = alpha_d*(real(1.0) - d) - beta_d*d;
  rhs[5] = alpha_f*(real(1.0) - f) - beta_f*f;
  rhs[6] = alpha_x1*(real(1.0) - x1) - beta_x1*x1;
  rhs[7] = c__1*(real(Istim) - i_K1 - i_Na - i_s - i_x1);
  const real ddm_dt_dm = -alpha_m - beta_m;
  const real ddh_dt_dh = -alpha_h - beta_h;
  const real ddj_dt_dj = -alpha_j - beta_j;
  const real ddd_dt_dd = -alpha_d - beta_d;
  const real ddf_dt_df = -alpha_f - beta_f;
  const real ddx1_dt_dx1 = -alpha_x1 - beta_x1;
  linear[0] = ddm_dt_dm;
  linear[1] = ddh_dt_dh;
  linear[2] = ddj_dt_dj;
  linear[4] = ddd_dt_dd;
  linear[5] = ddf_dt_df;
  linear[6] = ddx1_dt_dx1;
  if (only_linear) return;
  const real dE_s_dCai = -m_div(real(13.0287), Cai);
  const real di_s_dCai = -d*dE_s_dCai*f*g_s;
  const real ddCai_dt_dCai = -real(0.07) - real(0.01)*di_s_dCai;
  const real dV_eff_1_dV = (m_fabs(real(23.0) + V) < real(0.01) ? real(0.0) : real(1.0));
  const real di_K1_dV = m_div(real(0.0007000000000000001), (real(1.0) - m_exp(-real(0.04)*V_eff_1))) + m_div(real(0.016779896026542326)*m_exp(real(0.04)*V), (real(8.331137487687693)*m_exp(real(0.04)*V) + real(69.4078518387552)*m_exp(real(0.08)*V))) + m_div(real(0.0035)*(-real(4.0) + real(119.85640018958804)*m_exp(real(0.04)*V))*(-real(0.3332454995075077)*m_exp(real(0.04)*V) - real(5.552628147100417)*m_exp(real(0.08)*V)), m_sq(real(8.331137487687693)*m_exp(real(0.04)*V) + real(69.4078518387552)*m_exp(real(0.08)*V))) - m_div(real(0.00014000000000000001)*dV_eff_1_dV*(real(4.6000000000000005) + real(0.2)*V)*m_exp(-real(0.04)*V_eff_1), m_sq(real(1.0) - m_exp(-real(0.04)*V_eff_1)));
  const real di_Na_dV = g_Nac + g_Na*h*j*m_cube(m);
  const real di_s_dV = d*f*g_s;
  const real di_x1_dV = real(0.0017169779107590322)*x1 - real(7.891102846131407e-05)*x1*(-real(1.0) + real(21.75840239619708)*m_exp(real(0.04)*V))*m_exp(-real(0.04)*V);
  const real ddV_dt_dV = c__2*di_K1_dV + c__2*di_Na_dV + c__2*di_s_dV + c__2*di_x1_dV;
  linear[3] = ddCai_dt_dCai;
  linear[7] = ddV_dt_dV;
}

// Analytic Jacobian: jac[i*num_states + j] = d values[i] / d states[j].
template <typename real, typename PA, typename treal>
GOSS_HD GOSS_INLINE void compute_jacobian(const real* __restrict__ states, const treal time,
                                          const PA& pa, real* __restrict__ jac)
{
  (void)time; (void)pa;
  const real m = states[0];
  const real h = states[1];
  const real j = states[2];
  const real Cai = states[3];
  const real d = states[4];
  const real f = states[5];
  const real x1 = states[6];
  const real V = states[7];
  const real E_Na = pa.template get<0>();
  const real g_Na = pa.template get<1>();
  const real g_Nac = pa.template get<2>();
  const real g_s = pa.template get<3>();
  const real c__2 = pa.template cst<2>();
  const real V_eff_0 = (m_fabs(real(47.0) + V) < real(0.01) ? real(0.01) : real(47.0) + V);
  const real alpha_m = m_div((-real(47.0) - V), (-real(1.0) + m_exp(-real(0.1)*V_eff_0)));
  const real beta_m = real(0.7095526727489909)*m_exp(-real(0.056)*V);
  const real alpha_h = real(5.497962438709065e-10)*m_exp(-real(0.25)*V);
  const real beta_h = m_div(real(1.7), (real(1.0) + real(0.1580253208896478)*m_exp(-real(0.082)*V)));
  const real alpha_j = m_div(real(1.8690473007222892e-10)*m_exp(-real(0.25)*V), (real(1.0) + real(1.6788275299956603e-07)*m_exp(-real(0.2)*V)));
  const real beta_j = m_div(real(0.3), (real(1.0) + real(0.040762203978366204)*m_exp(-real(0.1)*V)));
  const real E_s = -real(82.3) - real(13.0287)*m_log(real(0.001)*Cai);
  const real alpha_d = m_div(real(0.095)*m_exp(real(0.05) - real(0.01)*V), (real(1.0) + real(1.4332881385696572)*m_exp(-real(0.07199424046076314)*V)));
  const real beta_d = m_div(real(0.07)*m_exp(-real(0.7457627118644068) - real(0.01694915254237288)*V), (real(1.0) + m_exp(real(2.2) + real(0.05)*V)));
  const real alpha_f = m_div(real(0.012)*m_exp(-real(0.224) - real(0.008)*V), (real(1.0) + real(66.5465065250986)*m_exp(real(0.14992503748125938)*V)));
  const real beta_f = m_div(real(0.0065)*m_exp(-real(0.6) - real(0.02)*V), (real(1.0) + m_exp(-real(6.0) - real(0.2)*V)));
  const real alpha_x1 = m_div(real(0.031158410986342627)*m_exp(real(0.08264462809917356)*V), (real(1.0) + real(17.41170806332765)*m_exp(real(0.05714285714285714)*V)));
  const real beta_x1 = m_div(real(0.0003916464405623223)*m_exp(-real(0.05998800239952009)*V), (real(1.0) + m_exp(-real(0.8) - real(0.04)*V)));
  const real V_eff_1 = (m_fabs(real(23.0) + V) < real(0.01) ? real(0.01) : real(23.0) + V);
  const real ddm_dt_dm = -alpha_m - beta_m;
  const real dV_eff_0_dV = (m_fabs(real(47.0) + V) < real(0.01) ? real(0.0) : real(1.0));
  const real dalpha_m_dV = -m_rcp(-real(1.0) + m_exp(-real(0.1)*V_eff_0)) + m_div(real(0.1)*dV_eff_0_dV*(-real(47.0) - V)*m_exp(-real(0.1)*V_eff_0), m_sq(-real(1.0) + m_exp(-real(0.1)*V_eff_0)));
  const real dbeta_m_dV = -real(0.03973494967394349)*m_exp(-real(0.056)*V);
  const real ddm_dt_dV = dalpha_m_dV*(real(1.0) - m) - dbeta_m_dV*m;
  const real ddh_dt_dh = -alpha_h - beta_h;
  const real dalpha_h_dV = -real(1.3744906096772661e-10)*m_exp(-real(0.25)*V);
  const real dbeta_h_dV = m_div(real(0.022028729732016906)*m_exp(-real(0.082)*V), m_sq(real(1.0) + real(0.1580253208896478)*m_exp(-real(0.082)*V)));
  const real ddh_dt_dV = dalpha_h_dV*(real(1.0) - h) - dbeta_h_dV*h;
  const real ddj_dt_dj = -alpha_j - beta_j;
  const real dalpha_j_dV = m_div(real(6.275616126633314e-18)*m_exp(-real(0.45)*V), m_sq(real(1.0) + real(1.6788275299956603e-07)*m_exp(-real(0.2)*V))) - m_div(real(4.672618251805723e-11)*m_exp(-real(0.25)*V), (real(1.0) + real(1.6788275299956603e-07)*m_exp(-real(0.2)*V)));
  const real dbeta_j_dV = m_div(real(0.001222866119350986)*m_exp(-real(0.1)*V), m_sq(real(1.0) + real(0.040762203978366204)*m_exp(-real(0.1)*V)));
  const real ddj_dt_dV = dalpha_j_dV*(real(1.0) - j) - dbeta_j_dV*j;
  const real dE_s_dCai = -m_div(real(13.0287), Cai);
  const real di_s_dCai = -d*dE_s_dCai*f*g_s;
  const real ddCai_dt_dCai = -real(0.07) - real(0.01)*di_s_dCai;
  const real di_s_dd = f*g_s*(V - E_s);
  const real ddCai_dt_dd = -real(0.01)*di_s_dd;
  const real di_s_df = d*g_s*(V - E_s);
  const real ddCai_dt_df = -real(0.01)*di_s_df;
  const real di_s_dV = d*f*g_s;
  const real ddCai_dt_dV = -real(0.01)*di_s_dV;
  const real ddd_dt_dd = -alpha_d - beta_d;
  const real dalpha_d_dV = -m_div(real(0.00095)*m_exp(real(0.05) - real(0.01)*V), (real(1.0) + real(1.4332881385696572)*m_exp(-real(0.07199424046076314)*V))) + m_div(real(0.009802906635285633)*m_exp(-real(0.07199424046076314)*V)*m_exp(real(0.05) - real(0.01)*V), m_sq(real(1.0) + real(1.4332881385696572)*m_exp(-real(0.07199424046076314)*V)));
  const real dbeta_d_dV = -m_div(real(0.0011864406779661019)*m_exp(-real(0.7457627118644068) - real(0.01694915254237288)*V), (real(1.0) + m_exp(real(2.2) + real(0.05)*V))) - m_div(real(0.0035000000000000005)*m_exp(-real(0.7457627118644068) - real(0.01694915254237288)*V)*m_exp(real(2.2) + real(0.05)*V), m_sq(real(1.0) + m_exp(real(2.2) + real(0.05)*V)));
  const real ddd_dt_dV = dalpha_d_dV*(real(1.0) - d) - d*dbeta_d_dV;
  const real ddf_dt_df = -alpha_f - beta_f;
  const real dalpha_f_dV = -m_div(real(9.6e-05)*m_exp(-real(0.224) - real(0.008)*V), (real(1.0) + real(66.5465065250986)*m_exp(real(0.14992503748125938)*V))) - m_div(real(0.11972384982026733)*m_exp(real(0.14992503748125938)*V)*m_exp(-real(0.224) - real(0.008)*V), m_sq(real(1.0) + real(66.5465065250986)*m_exp(real(0.14992503748125938)*V)));
  const real dbeta_f_dV = -m_div(real(0.00013)*m_exp(-real(0.6) - real(0.02)*V), (real(1.0) + m_exp(-real(6.0) - real(0.2)*V))) + m_div(real(0.0013)*m_exp(-real(6.0) - real(0.2)*V)*m_exp(-real(0.6) - real(0.02)*V), m_sq(real(1.0) + m_exp(-real(6.0) - real(0.2)*V)));
  const real ddf_dt_dV = dalpha_f_dV*(real(1.0) - f) - dbeta_f_dV*f;
  const real ddx1_dt_dx1 = -alpha_x1 - beta_x1;
  const real dalpha_x1_dV = m_div(real(0.0025750752881274898)*m_exp(real(0.08264462809917356)*V), (real(1.0) + real(17.41170806332765)*m_exp(real(0.05714285714285714)*V))) - m_div(real(0.031001208903507358)*m_exp(real(0.1397874852420307)*V), m_sq(real(1.0) + real(17.41170806332765)*m_exp(real(0.05714285714285714)*V)));
  const real dbeta_x1_dV = -m_div(real(2.3494087616216093e-05)*m_exp(-real(0.05998800239952009)*V), (real(1.0) + m_exp(-real(0.8) - real(0.04)*V))) + m_div(real(1.5665857622492893e-05)*m_exp(-real(0.05998800239952009)*V)*m_exp(-real(0.8) - real(0.04)*V), m_sq(real(1.0) + m_exp(-real(0.8) - real(0.04)*V)));
  const real ddx1_dt_dV = dalpha_x1_dV*(real(1.0) - x1) - dbeta_x1_dV*x1;
  const real di_Na_dm = real(3.0)*g_Na*h*j*m_sq(m)*(V - E_Na);
  const real ddV_dt_dm = c__2*di_Na_dm;
  const real di_Na_dh = g_Na*j*m_cube(m)*(V - E_Na);
  const real ddV_dt_dh = c__2*di_Na_dh;
  const real di_Na_dj = g_Na*h*m_cube(m)*(V - E_Na);
  const real ddV_dt_dj = c__2*di_Na_dj;
  const real ddV_dt_dCai = c__2*di_s_dCai;
  const real ddV_dt_dd = c__2*di_s_dd;
  const real ddV_dt_df = c__2*di_s_df;
  const real di_x1_dx1 = real(0.0019727757115328517)*(-real(1.0) + real(21.75840239619708)*m_exp(real(0.04)*V))*m_exp(-real(0.04)*V);
  const real ddV_dt_dx1 = c__2*di_x1_dx1;
  const real dV_eff_1_dV = (m_fabs(real(23.0) + V) < real(0.01) ? real(0.0) : real(1.0));
  const real di_K1_dV = m_div(real(0.0007000000000000001), (real(1.0) - m_exp(-real(0.04)*V_eff_1))) + m_div(real(0.016779896026542326)*m_exp(real(0.04)*V), (real(8.331137487687693)*m_exp(real(0.04)*V) + real(69.4078518387552)*m_exp(real(0.08)*V))) + m_div(real(0.0035)*(-real(4.0) + real(119.85640018958804)*m_exp(real(0.04)*V))*(-real(0.3332454995075077)*m_exp(real(0.04)*V) - real(5.552628147100417)*m_exp(real(0.08)*V)), m_sq(real(8.331137487687693)*m_exp(real(0.04)*V) + real(69.4078518387552)*m_exp(real(0.08)*V))) - m_div(real(0.00014000000000000001)*dV_eff_1_dV*(real(4.6000000000000005) + real(0.2)*V)*m_exp(-real(0.04)*V_eff_1), m_sq(real(1.0) - m_exp(-real(0.04)*V_eff_1)));
  const real di_Na_dV = g_Nac + g_Na*h*j*m_cube(m);
  const real di_x1_dV = real(0.0017169779107590322)*x1 - real(7.891102846131407e-05)*x1*(-real(1.0) + real(21.75840239619708)*m_exp(real(0.04)*V))*m_exp(-real(0.04)*V);
  const real ddV_dt_dV = c__2*di_K1_dV + c__2*di_Na_dV + c__2*di_s_dV + c__2*di_x1_dV;
  jac[0] = ddm_dt_dm;
  jac[1] = real(0.0);
  jac[2] = real(0.0);
  jac[3] = real(0.0);
  jac[4] = real(0.0);
  jac[5] = real(0.0);
  jac[6] = real(0.0);
  jac[7] = ddm_dt_dV;
  jac[8] = real(0.0);
  jac[9] = ddh_dt_dh;
  jac[10] = real(0.0);
  jac[11] = real(0.0);
  jac[12] = real(0.0);
  jac[13] = real(0.0);
  jac[14] = real(0.0);
  jac[15] = ddh_dt_dV;
  jac[16] = real(0.0);
  jac[17] = real(0.0);
  jac[18] = ddj_dt_dj;
  jac[19] = real(0.0);
  jac[20] = real(0.0);
  jac[21] = real(0.0);
  jac[22] = real(0.0);
  jac[23] = ddj_dt_dV;
  jac[24] = real(0.0);
  jac[25] = real(0.0);
  jac[26] = real(0.0);
  jac[27] = ddCai_dt_dCai;
  jac[28] = ddCai_dt_dd;
  jac[29] = ddCai_dt_df;
  jac[30] = real(0.0);
  jac[31] = ddCai_dt_dV;
  jac[32] = real(0.0);
  jac[33] = real(0.0);
  jac[34] = real(0.0);
  jac[35] = real(0.0);
  jac[36] = ddd_dt_dd;
  jac[37] = real(0.0);
  jac[38] = real(0.0);
  jac[39] = ddd_dt_dV;
  jac[40] = real(0.0);
  jac[41] = real(0.0);
  jac[42] = real(0.0);
  jac[43] = real(0.0);
  jac[44] = real(0.0);
  jac[45] = ddf_dt_df;
  jac[46] = real(0.0);
  jac[47] = ddf_dt_dV;
  jac[48] = real(0.0);
  jac[49] = real(0.0);
  jac[50] = real(0.0);
  jac[51] = real(0.0);
  jac[52] = real(0.0);
  jac[53] = real(0.0);
  jac[54] = ddx1_dt_dx1;
  jac[55] = ddx1_dt_dV;
  jac[56] = ddV_dt_dm;
  jac[57] = ddV_dt_dh;
  jac[58] = ddV_dt_dj;
  jac[59] = ddV_dt_dCai;
  jac[60] = ddV_dt_dd;
  jac[61] = ddV_dt_df;
  jac[62] = ddV_dt_dx1;
  jac[63] = ddV_dt_dV;
}

}}  // namespace goss_models::beeler_reuter_1977
