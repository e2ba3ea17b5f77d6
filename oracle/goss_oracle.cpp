// CPU oracle: restatement of goss's ODESystemSolver and solvers.
// TEST INFRASTRUCTURE ONLY -- see goss_oracle.h.
#include "goss_oracle.h"

#include <cmath>
#include <cstring>
#include <deque>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "tp06_m.h"
#include "tp06_m_fp3.h"
#include "tp06_m_fp6.h"
#include "tp06_epi.h"
#include "fitzhughnagumo.h"
#include "lorentz.h"
#include "fhn.h"
#include "beeler_reuter_1977.h"
#include "oscilator.h"
#include "linear_decay.h"
#include "van_der_pol_dae.h"
#include "markov_three_state.h"

namespace {

std::string g_err;

// Parameter accessor of the generated model code: v = the node's parameter
// vector, c = its parameter-only constants (compute_constants).
struct Acc {
  const double* v;
  const double* c;
  template <int k> double get() const { return v[k]; }
  template <int k> double getd() const { return v[k]; }
  template <int k> double cst() const { return c[k]; }
  template <int k> double cstd() const { return c[k]; }
};

constexpr int kMaxConstants = 512;

struct VariantOps {
  const char* name;
  int S, P, NFS, NFP;
  const int* fsi;
  const int* fpi;
  const double* defaults;
  const double* ic;
  const unsigned char* linear_terms;
  const unsigned char* diff;  // ODE::differential_states (ODE.h:85-89); nullptr: every state differential
  const char* const* pnames;
  int K;  // number of parameter-only constants
  void (*constants)(const double* p, double* c);
  void (*eval_fn)(const double*, double, const double*, const double*, double*);
  void (*linearized_fn)(const double*, double, const double*, const double*, double*, double*, bool);
  void (*jacobian)(const double*, double, const double*, const double*, double*);
  // Generated step adjoints (test infrastructure for the device adjoint): kind 0 grl1_vjp, 1 rl1_vjp,
  // 2 eval_vjp (lam is the seed, out accumulates y_bar); tape selects the *_tape form.
  void (*step_vjp)(int kind, bool tape, const double*, double, const double*, const double*, double, double,
                   const double*, double*, double*);
  // A model registered at run time (oracle_register_external_model): its
  // right-hand side is called through these entries with the full parameter
  // vector; the generated entries above are then unused.
  const oracle_external_model* ext = nullptr;
  // The generated analytic Jacobian stands for ODE::compute_jacobian, as in a plugin generated with
  // generate_jacobian (codegeneration.py:186, 622-641).
  bool analytic_jacobian = false;

  // The mass matrix's diagonal entry of state i, as the reference reads it
  // (_ode->differential_states()[i], an unsigned char).
  double mass(const int i) const { return diff ? double(diff[i]) : 1.0; }
  bool is_dae() const
  {
    for (int i = 0; diff && i < S; ++i)
      if (!diff[i]) return true;
    return false;
  }

  void eval(const double* y, double t, const double* p, const double* c, double* out) const
  {
    if (ext) ext->eval(ext->ctx, y, t, p, out);
    else eval_fn(y, t, p, c, out);
  }
  void linearized(const double* y, double t, const double* p, const double* c, double* lin, double* rhs,
                  bool only_linear) const
  {
    if (ext) ext->linearized_eval(ext->ctx, y, t, p, lin, rhs, only_linear ? 1 : 0);
    else linearized_fn(y, t, p, c, lin, rhs, only_linear);
  }
};

// Host tape for grl1_vjp_tape (slot k at v[k]).
struct HostTape {
  double* v;
  double get(int k) const { return v[k]; }
  void put(int k, double x) const { v[k] = x; }
  void add(int k, double x) const { v[k] += x; }
};

#define GOSS_ORACLE_VARIANT(V)                                                                         \
  VariantOps{#V, goss_models::V::num_states, goss_models::V::num_parameters,                           \
             goss_models::V::num_field_states, goss_models::V::num_field_parameters,                   \
             goss_models::V::field_state_indices, goss_models::V::field_parameter_indices,             \
             goss_models::V::default_parameters, goss_models::V::initial_conditions,                   \
             goss_models::V::linear_terms, goss_models::V::differential_states,                        \
             goss_models::V::parameter_names,                                                          \
             goss_models::V::num_constants,                                                            \
             [](const double* p, double* c) { goss_models::V::compute_constants<double>(Acc{p, nullptr}, c); }, \
             [](const double* y, double t, const double* p, const double* c, double* out) {            \
               goss_models::V::eval<double>(y, t, Acc{p, c}, out);                                     \
             },                                                                                        \
             [](const double* y, double t, const double* p, const double* c, double* lin, double* rhs,  \
                bool ol) { goss_models::V::linearized_eval<double>(y, t, Acc{p, c}, lin, rhs, ol); },   \
             [](const double* y, double t, const double* p, const double* c, double* jac) {            \
               goss_models::V::compute_jacobian<double>(y, t, Acc{p, c}, jac);                         \
             },                                                                                        \
             [](int kind, bool tape, const double* y, double t, const double* p, const double* c, double h,  \
                double delta, const double* lam, double* out, double* fp_bar) {                         \
               namespace NS = goss_models::V;                                                          \
               const Acc a{p, c};                                                                      \
               int n = kind == 0 ? NS::grl1_vjp_tape_slots                                             \
                                 : (kind == 1 ? NS::rl1_vjp_tape_slots : NS::eval_vjp_tape_slots);     \
               std::vector<double> tp(size_t(tape ? n : 0));                                           \
               const HostTape T{tp.data()};                                                            \
               if (kind == 0 && !tape) NS::grl1_vjp<double>(y, t, a, h, delta, lam, out, fp_bar);      \
               if (kind == 0 && tape) NS::grl1_vjp_tape<double>(y, t, a, h, delta, lam, out, fp_bar, T); \
               if (kind == 1 && !tape) NS::rl1_vjp<double>(y, t, a, h, delta, lam, out, fp_bar);       \
               if (kind == 1 && tape) NS::rl1_vjp_tape<double>(y, t, a, h, delta, lam, out, fp_bar, T); \
               if (kind == 2 && !tape) NS::eval_vjp<double>(y, t, a, lam, out, fp_bar);                \
               if (kind == 2 && tape) NS::eval_vjp_tape<double>(y, t, a, lam, out, fp_bar, T);         \
             }}

// The generated models, then the models registered at run time (a deque:
// registration never moves the entries systems point to).
std::deque<VariantOps>& variant_table()
{
  static std::deque<VariantOps> t{
      GOSS_ORACLE_VARIANT(tp06_m),         GOSS_ORACLE_VARIANT(tp06_m_fp3),
      GOSS_ORACLE_VARIANT(tp06_m_fp6),     GOSS_ORACLE_VARIANT(tp06_epi),
      GOSS_ORACLE_VARIANT(fitzhughnagumo), GOSS_ORACLE_VARIANT(lorentz),
      GOSS_ORACLE_VARIANT(fhn),            GOSS_ORACLE_VARIANT(beeler_reuter_1977),
      GOSS_ORACLE_VARIANT(oscilator),      GOSS_ORACLE_VARIANT(linear_decay),
      GOSS_ORACLE_VARIANT(van_der_pol_dae), GOSS_ORACLE_VARIANT(markov_three_state),
  };
  return t;
}
#define kVariants variant_table()
#define kNumVariants int(variant_table().size())

// ---------------------------------------------------------------------------
// Dense helpers of ODE.cpp
// ---------------------------------------------------------------------------

// ODE::lu_factorize (ODE.cpp:69-100): in-place Doolittle LU without pivoting,
// unit lower triangle stored below the diagonal.
void lu_factorize(int n, double* a)
{
  for (int k = 1; k < n; k++) {
    for (int i = 0; i <= k - 1; ++i) {
      double acc = 0.0;
      for (int r = 0; r <= i - 1; r++) acc += a[i * n + r] * a[r * n + k];
      a[i * n + k] -= acc;
      acc = 0.0;
      for (int r = 0; r <= i - 1; r++) acc += a[k * n + r] * a[r * n + i];
      a[k * n + i] = (a[k * n + i] - acc) / a[i * n + i];
    }
    double acc = 0.0;
    for (int r = 0; r <= k - 1; r++) acc += a[k * n + r] * a[r * n + k];
    a[k * n + k] -= acc;
  }
}

// ODE::forward_backward_subst (ODE.cpp:102-132); b and x may alias.
void fb_subst(int n, const double* a, const double* b, double* x)
{
  x[0] = b[0];
  for (int i = 1; i < n; ++i) {
    double acc = 0.0;
    for (int j = 0; j <= i - 1; ++j) acc = acc + a[i * n + j] * x[j];
    x[i] = b[i] - acc;
  }
  x[n - 1] = x[n - 1] / a[(n - 1) * n + (n - 1)];
  for (int i = n - 2; i >= 0; i--) {
    double acc = 0;
    for (int j = i + 1; j < n; ++j) acc = acc + a[i * n + j] * x[j];
    x[i] = (x[i] - acc) / a[i * n + i];
  }
}

// ODE::compute_jacobian (ODE.cpp:47-67): forward differences with
// delta_i = sqrt(1e-15 * max(1e-5, |y_i|)); f1/f2 are caller scratch.
void fd_jacobian(const VariantOps& m, double* y, double t, const double* p, const double* c, double* jac, double* f1,
                 double* f2)
{
  const int S = m.S;
  m.eval(y, t, p, c, f1);
  for (int i = 0; i < S; ++i) {
    const double ysafe = y[i];
    const double mx = 1e-5 > std::fabs(ysafe) ? 1e-5 : std::fabs(ysafe);
    const double delta = std::sqrt(1e-15 * mx);
    y[i] += delta;
    m.eval(y, t, p, c, f2);
    for (int j = 0; j < S; ++j) jac[j * S + i] = (f2[j] - f1[j]) / delta;
    y[i] = ysafe;
  }
}

// The model's ODE::compute_jacobian: a model class that overrides it (the
// reference's hand-written test models, or a generated model with its
// analytic Jacobian) is called; otherwise the forward differences above.
void ode_jacobian(const VariantOps& m, double* y, double t, const double* p, const double* c, double* jac,
                  double* f1, double* f2)
{
  if (m.ext && m.ext->compute_jacobian) m.ext->compute_jacobian(m.ext->ctx, y, t, p, jac);
  else if (m.analytic_jacobian) m.jacobian(y, t, p, c, jac);
  else fd_jacobian(m, y, t, p, c, jac, f1, f2);
}

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------

enum SolverKind {
  EE, RK2, RK4, RL1, RL2, GRL1, GRL2, IMPLICIT_EULER, ESDIRK23A, RKF32, THETA, BASIC_IMPLICIT_EULER, ESDIRK4O32,
  NUM_SOLVERS
};

int solver_kind(const std::string& s)
{
  static const char* names[] = {"ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2",
                                "ImplicitEuler", "ESDIRK23a", "RKF32", "ThetaSolver", "BasicImplicitEuler",
                                "ESDIRK4O32"};
  for (int i = 0; i < NUM_SOLVERS; ++i)
    if (s == names[i]) return i;
  return -1;
}

struct SolverParams {
  double ldt = -1.0;  // ODESolver::_ldt (ODESolver.h:137)
  double delta = 1e-8;  // GRL1.h:39
  // ImplicitODESolver.h:36-41
  double eta_0 = 1.0;
  double kappa = 0.1;
  double relative_tolerance = 1.e-12;
  int max_iterations = 30;
  double max_relative_previous_residual = 0.01;
  bool always_recompute_jacobian = false;
  // ImplicitEuler.h:36-37 / ESDIRK23a.h:37-38
  int num_refinements_without_always_recomputing_jacobian = 2;
  double min_dt = 0.0001;
  // AdaptiveExplicitSolver::reset (AdaptiveExplicitSolver.cpp:46-65), set_tol (AdaptiveExplicitSolver.h:71-75)
  double atol = 1.0e-5;
  double rtol = 1.0e-8;
  double theta = 0.5;  // ThetaSolver.h:37
};

constexpr double GOSS_EPS = 3.0e-16;  // constants.h:22

// Mutable state of one implicit solver instance (ImplicitODESolver.h:83-101,
// ImplicitEuler.h:66-68, ESDIRK23a.h:94-115).  The reference keeps one such
// instance per thread and carries it from node to node; the per-node mode
// keeps one per node.
struct ImplicitState {
  std::vector<double> jac, jfac;
  // ImplicitODESolver
  double eta = 1.0;
  bool recompute_jacobian = true;
  int newton_iterations = 0;
  long rejects = 0, jac_comp = 0;
  bool always_recompute_jacobian = false;  // sticky copy of the public parameter
  // ImplicitEuler
  bool justrefined = false;
  // ESDIRK23a (ctor values, ESDIRK23a.cpp:27-44)
  double es_eta = 1.0;
  double loc_error = 1e9;
  double prev_dt = 0.1;
  bool first_step = true;
  bool new_jacobian = true;
  long nfevals = 0, num_rejected = 0;
  long accepted = 0;
  // RKF32 / AdaptiveExplicitSolver (AdaptiveExplicitSolver.cpp:28-31, 46-65;
  // RKF32.cpp:91-103): the initial step guess _ldt and facmax persist from
  // one forward call to the next.
  double rk_ldt = 0.1;
  double facmax = 2.0;
  long rk_nfevals = 0, ndtsa = 0, ndtsr = 0;
  // ESDIRK4O32 over AdaptiveImplicitSolver: ESDIRK4O32::reset sets the step
  // guess _ldt to -1 (ESDIRK4O32.cpp:84-94); AdaptiveImplicitSolver::reset
  // (AdaptiveImplicitSolver.cpp:43-66) facmax 2, err_old -1.
  double ai_ldt = -1.0;
  double ai_facmax = 2.0;
  double err_old = -1.0;
  double dt_old = 0.0;
  long e4_nfevals = 0;
  std::vector<double> e4_yn;  // ESDIRK4O32's member yn (ESDIRK4O32.h)
  void init(int S, const SolverParams& sp)
  {
    jac.assign(size_t(S) * S, 0.0);
    jfac.assign(size_t(S) * S, 0.0);
    eta = sp.eta_0;  // ImplicitODESolver::reset (ImplicitODESolver.cpp:80-90)
    always_recompute_jacobian = sp.always_recompute_jacobian;
  }
};

// Sub-step count of the fixed-step solvers, e.g. GRL1.cpp:65-67.
inline unsigned long substeps(double ldt_0, double dt)
{
  return ldt_0 > 0 ? (unsigned long)std::ceil(dt / ldt_0 - 1.0E-12) : 1UL;
}

// Per-thread scratch, sized to the model.
struct Work {
  std::vector<double> f1, f2, k1, k2, k3, k4, tmp, y2;
  std::vector<double> z1, z2, z3, z4, yh, prev, b, dz, yz, fy, y_int, z_int, prev_dz, ff1, ff2, ki, yn, e;
  ImplicitState carried;  // the per-thread solver instance of the reference
  void size(int S)
  {
    for (auto* v : {&f1, &f2, &k1, &k2, &k3, &k4, &tmp, &y2, &z1, &z2, &z3, &z4, &yh, &prev, &b, &dz, &yz, &fy,
                    &y_int, &z_int, &prev_dz, &ff1, &ff2, &ki, &yn, &e})
      v->assign(S, 0.0);
  }
};

// One node: solver.forward(y, t, interval) with the node's parameter vector.
void forward_explicit(int kind, const VariantOps& m, const SolverParams& sp, double* y, double t, double dt,
                      const double* p, const double* c, Work& w)
{
  const int S = m.S;
  const unsigned long nsteps = substeps(sp.ldt, dt);
  const double ldt = dt / nsteps;
  double lt = t;
  double* f1 = w.f1.data();
  double* f2 = w.f2.data();
  switch (kind) {
  case EE:  // ExplicitEuler.cpp:64-87
    for (unsigned long step = 0; step < nsteps; ++step) {
      m.eval(y, lt, p, c, f1);
      for (int i = 0; i < S; ++i) y[i] += ldt * f1[i];
      lt += ldt;
    }
    break;
  case RK2:  // RK2.cpp:62-92
    for (unsigned long step = 0; step < nsteps; ++step) {
      double* k1 = w.k1.data();
      double* tmp = w.tmp.data();
      m.eval(y, lt, p, c, k1);
      for (int i = 0; i < S; ++i) tmp[i] = y[i] + 0.5 * ldt * k1[i];
      m.eval(tmp, lt + 0.5 * ldt, p, c, k1);
      for (int i = 0; i < S; ++i) y[i] += ldt * k1[i];
      lt += ldt;
    }
    break;
  case RK4:  // RK4.cpp:68-102, axpy RK4.h:70-74
    for (unsigned long step = 0; step < nsteps; ++step) {
      double *k1 = w.k1.data(), *k2 = w.k2.data(), *k3 = w.k3.data(), *k4 = w.k4.data(), *tmp = w.tmp.data();
      m.eval(y, lt, p, c, k1);
      const double h = 0.5 * ldt;
      for (int i = 0; i < S; ++i) tmp[i] = y[i] + h * k1[i];
      m.eval(tmp, lt + 0.5 * ldt, p, c, k2);
      for (int i = 0; i < S; ++i) tmp[i] = y[i] + h * k2[i];
      m.eval(tmp, lt + 0.5 * ldt, p, c, k3);
      for (int i = 0; i < S; ++i) tmp[i] = y[i] + ldt * k3[i];
      m.eval(tmp, lt + ldt, p, c, k4);
      for (int i = 0; i < S; ++i) y[i] += ldt * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
      lt += ldt;
    }
    break;
  case RL1:  // RL1.cpp:59-79, _one_step RL1.h:68-77
  case RL2: {  // RL2.cpp:64-87
    auto rl_step = [&](double* out, const double* at, const double* y0, double tt, double h) {
      m.linearized(at, tt, p, c, f1, f2, true);
      for (int i = 0; i < S; ++i)
        out[i] = m.linear_terms[i] ? y0[i] + f2[i] / f1[i] * (std::exp(f1[i] * h) - 1.0) : y0[i] + f2[i] * h;
    };
    for (unsigned long step = 0; step < nsteps; ++step) {
      if (kind == RL1) {
        rl_step(y, y, y, lt, ldt);
      } else {
        double* y2 = w.y2.data();
        rl_step(y2, y, y, lt, ldt * 0.5);
        rl_step(y, y2, y, lt, ldt);  // second stage evaluated at time lt (RL2.cpp:79-82)
      }
      lt += ldt;
    }
    break;
  }
  case GRL1:  // GRL1.cpp:61-82, _one_step GRL1.h:72-84
  case GRL2: {  // GRL2.cpp:63-86 (delta is the member value; GRL2.cpp:67 reads an uninitialised local)
    const double delta = sp.delta;
    auto grl_step = [&](double* out, const double* at, const double* y0, double tt, double h) {
      m.linearized(at, tt, p, c, f1, f2, false);
      for (int i = 0; i < S; ++i)
        out[i] = (std::fabs(f1[i]) > delta) ? y0[i] + f2[i] / f1[i] * (std::exp(f1[i] * h) - 1.0)
                                            : y0[i] + f2[i] * h;
    };
    for (unsigned long step = 0; step < nsteps; ++step) {
      if (kind == GRL1) {
        grl_step(y, y, y, lt, ldt);
      } else {
        double* y2 = w.y2.data();
        grl_step(y2, y, y, lt, ldt * 0.5);
        grl_step(y, y2, y, lt, ldt);
      }
      lt += ldt;
    }
    break;
  }
  default:
    throw std::runtime_error("solver not available in the oracle");
  }
}


// ---------------------------------------------------------------------------
// Implicit solvers
// ---------------------------------------------------------------------------

inline double l2norm(const double* v, int S)
{
  // ImplicitODESolver::norm (ImplicitODESolver.cpp:247-256).
  double acc = 0;
  for (int i = 0; i < S; ++i) acc += v[i] * v[i];
  return std::sqrt(acc);
}

// ImplicitODESolver::compute_factorized_jacobian (ImplicitODESolver.cpp:92-105):
// the model's Jacobian at y, scaled by -dt*alpha, the diagonal mass matrix
// added (add_mass_matrix, :258-262), LU factorized in place.
void ie_factorized_jacobian(const VariantOps& m, ImplicitState& st, double* y, double t, double dt, double alpha,
                            const double* p, const double* c, Work& w)
{
  const int S = m.S;
  ode_jacobian(m, y, t, p, c, st.jac.data(), w.ff1.data(), w.ff2.data());
  const double scale = -dt * alpha;
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j) st.jac[i * S + j] *= scale;
  for (int i = 0; i < S; ++i) st.jac[i * S + i] += 1.0 * m.mass(i);  // add_mass_matrix (:258-262)
  lu_factorize(S, st.jac.data());
  st.jac_comp += 1;
}

// ImplicitODESolver::newton_solve (ImplicitODESolver.cpp:114-245).
bool newton_solve(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* z, const double* prev,
                  const double* y0, double t, double dt, double alpha, bool always_recompute, const double* p,
                  const double* c, Work& w)
{
  const int S = m.S;
  bool step_ok = true;
  st.newton_iterations = 0;
  double relative_previous_residual = 1.0, relative_residual = 1.0, previous_residual = 1.0,
         initial_residual = 1.0, residual;
  double* yz = w.yz.data();
  double* f1 = w.fy.data();
  double* b = w.b.data();
  double* dz = w.dz.data();
  do {
    for (int i = 0; i < S; ++i) yz[i] = y0[i] + z[i];
    m.eval(yz, t, p, c, f1);
    for (int i = 0; i < S; ++i) b[i] = -z[i] * m.mass(i) + dt * (prev[i] + alpha * f1[i]);
    residual = l2norm(b, S);
    if (relative_residual < sp.relative_tolerance) break;
    if (st.recompute_jacobian || always_recompute) {
      ie_factorized_jacobian(m, st, yz, t, dt, alpha, p, c, w);
      st.recompute_jacobian = false;
    }
    fb_subst(S, st.jac.data(), b, dz);
    if (st.newton_iterations == 0) {
      initial_residual = residual;
      st.eta = st.eta > GOSS_EPS ? st.eta : GOSS_EPS;
      st.eta = std::pow(st.eta, 0.8);
    } else {
      relative_previous_residual = residual / previous_residual;
      st.recompute_jacobian = relative_previous_residual >= sp.max_relative_previous_residual;
      if (relative_previous_residual >= 1) {
        step_ok = false;
        st.rejects++;
        st.recompute_jacobian = true;
        break;
      }
      const double scaled = std::max(std::pow(relative_previous_residual, sp.max_iterations - st.newton_iterations),
                                     GOSS_EPS);
      if (residual > (sp.kappa * sp.relative_tolerance * (1 - relative_previous_residual) / scaled))
        st.recompute_jacobian = true;
      st.eta = relative_previous_residual / (1.0 - relative_previous_residual);
    }
    if (st.newton_iterations > sp.max_iterations) {
      st.recompute_jacobian = true;
      st.rejects++;
      step_ok = false;
      break;
    }
    for (int i = 0; i < S; ++i) z[i] += dz[i];
    relative_residual = residual / initial_residual;
    previous_residual = residual;
    st.newton_iterations++;
  } while (st.eta * relative_residual >= sp.kappa * sp.relative_tolerance);
  return step_ok;
}

// ImplicitEuler::forward (ImplicitEuler.cpp:57-152).
void implicit_euler_forward(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* y, double t,
                            double dt, const double* p, const double* c, Work& w)
{
  const int S = m.S;
  const double t_end = t + dt;
  const double ldt_0 = sp.ldt;
  double ldt = ldt_0 > 0 ? ldt_0 : dt;
  int num_refinements = 0;
  const double eps = GOSS_EPS * 1000;
  double* z1 = w.z1.data();
  double* prev = w.prev.data();
  for (int i = 0; i < S; ++i) prev[i] = 0.0;
  while (true) {
    for (int i = 0; i < S; ++i) z1[i] = 0.0;
    if (num_refinements > sp.num_refinements_without_always_recomputing_jacobian)
      st.always_recompute_jacobian = true;
    const bool step_ok = newton_solve(m, sp, st, z1, prev, y, t + ldt, ldt, 1.0, st.always_recompute_jacobian, p,
                                      c, w);
    if (step_ok) {
      for (int i = 0; i < S; ++i) y[i] += z1[i];
      t += ldt;
      if (std::fabs(t - t_end) < eps) break;
      if (!st.justrefined) {
        const double tmp = 2.0 * ldt;
        if (ldt_0 > 0. && tmp >= ldt_0)
          ldt = ldt_0;
        else
          ldt = tmp;
      } else {
        st.justrefined = false;
      }
      if ((t + ldt + GOSS_EPS) > t_end) ldt = t_end - t;
    } else {
      ldt /= 2.0;
      if (ldt < sp.min_dt) {
        std::ostringstream msg;  // ImplicitEuler.cpp:132-141
        msg << "Newtons solver failed to converge as dt become smaller than 'min_dt' " << sp.min_dt << std::endl;
        throw std::runtime_error(msg.str());
      }
      st.recompute_jacobian = true;
      st.justrefined = true;
      num_refinements += 1;
    }
  }
}

// ThetaSolver::forward (ThetaSolver.cpp:48-155): a theta step
// z = dt ((1 - theta) f(y, t) + theta f(y + z, t + theta dt)) solved by
// newton_solve with prev = (1 - theta) f(y, t) and alpha = theta.  The
// public always_recompute_jacobian is set while num_refinements <=
// num_refinements_without_always_recomputing_jacobian (ThetaSolver.cpp:82-83,
// the opposite test of ImplicitEuler's), so from the first step on the
// Jacobian is recomputed at every Newton iteration.
void theta_forward(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* y, double t, double dt,
                   const double* p, const double* c, Work& w)
{
  const int S = m.S;
  const double theta = sp.theta;
  const double t_end = t + dt;
  const double ldt_0 = sp.ldt;
  double ldt = ldt_0 > 0 ? ldt_0 : dt;
  int num_refinements = 0;
  const double eps = GOSS_EPS * 1000;
  double* z1 = w.z1.data();
  double* prev = w.prev.data();
  double* ft1 = w.ff1.data();
  for (int i = 0; i < S; ++i) prev[i] = 0.0;
  while (true) {
    for (int i = 0; i < S; ++i) z1[i] = 0.0;
    if (std::abs(theta - 1.0) > GOSS_EPS) {
      m.eval(y, t, p, c, ft1);
      for (int i = 0; i < S; ++i) prev[i] = (1 - theta) * ft1[i];
    }
    if (num_refinements <= sp.num_refinements_without_always_recomputing_jacobian)
      st.always_recompute_jacobian = true;
    const bool step_ok = newton_solve(m, sp, st, z1, prev, y, t + theta * ldt, ldt, theta,
                                      st.always_recompute_jacobian, p, c, w);
    if (step_ok) {
      for (int i = 0; i < S; ++i) y[i] += z1[i];
      t += ldt;
      if (std::fabs(t - t_end) < eps) break;
      if (!st.justrefined) {
        const double tmp = 2.0 * ldt;
        if (ldt_0 > 0. && tmp >= ldt_0)
          ldt = ldt_0;
        else
          ldt = tmp;
      } else {
        st.justrefined = false;
      }
      if ((t + ldt + GOSS_EPS) > t_end) ldt = t_end - t;
    } else {
      ldt /= 2.0;
      if (ldt < sp.min_dt) {
        std::ostringstream msg;  // ThetaSolver.cpp:131-141
        msg << "Newtons solver failed to converge as dt become smaller than 'min_dt' " << sp.min_dt << std::endl;
        throw std::runtime_error(msg.str());
      }
      st.justrefined = true;
      num_refinements += 1;
    }
  }
}

// ESDIRK23a coefficients (ESDIRK23a.cpp:27-44).
struct EsdirkCoef {
  double gamma, a21, a22, a31, a32, a33, a41, a42, a43, a44, b1, b2, b3, b4, bh1, bh2, bh3, c2, c3, c4;
  double s_fac = 0.9, tol = 1.0e-5, etamin = 1.0e-10, kappa = 1.0e-1;
  EsdirkCoef()
  {
    gamma = 0.43586652150845899942;
    a21 = gamma;
    a22 = gamma;
    a31 = (-4 * gamma * gamma + 6 * gamma - 1) / (4 * gamma);
    a32 = (-2 * gamma + 1) / (4 * gamma);
    a33 = gamma;
    a41 = (6 * gamma - 1) / (12 * gamma);
    a42 = -1 / ((24 * gamma - 12) * gamma);
    a43 = (-6 * gamma * gamma + 6 * gamma - 1) / (6 * gamma - 3);
    a44 = gamma;
    b1 = a41;
    b2 = a42;
    b3 = a43;
    b4 = a44;
    bh1 = a31;
    bh2 = a32;
    bh3 = a33;
    c2 = 2.0 * gamma;
    c3 = 1.0;
    c4 = 1.0;
  }
};

// ESDIRK23a::compute_ode_jacobian (ESDIRK23a.h:57-62).
void es_ode_jacobian(const VariantOps& m, ImplicitState& st, double* y, double t, const double* p, const double* c,
                     Work& w)
{
  ode_jacobian(m, y, t, p, c, st.jac.data(), w.ff1.data(), w.ff2.data());
  st.jac_comp++;
  st.new_jacobian = true;
}

// ESDIRK23a::compute_factorized_jacobian(dt) (ESDIRK23a.cpp:101-115).
void es_factorized(const VariantOps& m, const EsdirkCoef& k, ImplicitState& st, double dt)
{
  const int S = m.S;
  st.jfac = st.jac;
  const double scale = -dt * k.a22;
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j) st.jfac[i * S + j] *= scale;
  for (int i = 0; i < S; ++i) st.jfac[i * S + i] += 1.0 * m.mass(i);
  lu_factorize(S, st.jfac.data());
}

// ESDIRK23a::compute_stage_val (ESDIRK23a.cpp:201-258).
bool es_stage(const VariantOps& m, const SolverParams& sp, const EsdirkCoef& k, ImplicitState& st, double* z,
              const double* y, const double* prev, double t_s, double dt, const double* p, const double* c, Work& w)
{
  const int S = m.S;
  st.newton_iterations = 0;
  bool conv_ok = true;
  double* y_int = w.y_int.data();
  double* z_int = w.z_int.data();
  double* dz = w.dz.data();
  double* prev_dz = w.prev_dz.data();
  double* b = w.b.data();
  for (int i = 0; i < S; ++i) dz[i] = 0.0;
  st.es_eta = std::pow(std::max(st.es_eta, k.etamin), 0.8);
  do {
    for (int i = 0; i < S; ++i) y_int[i] = y[i] + dt * prev[i] + k.gamma * dt * z[i];
    m.eval(y_int, t_s, p, c, z_int);
    st.nfevals++;
    for (int i = 0; i < S; ++i) b[i] = -z[i] + z_int[i];
    for (int i = 0; i < S; ++i) prev_dz[i] = dz[i];
    fb_subst(S, st.jfac.data(), b, dz);
    if (st.newton_iterations >= 1) {
      const double theta = l2norm(dz, S) / l2norm(prev_dz, S);
      if (!(theta <= 1)) {
        st.num_rejected++;
        conv_ok = false;
        break;
      }
      st.es_eta = theta / (1.0 - theta);
    }
    if (st.newton_iterations > sp.max_iterations) {
      st.num_rejected++;
      conv_ok = false;
      break;
    }
    for (int i = 0; i < S; ++i) z[i] += dz[i];
    st.newton_iterations++;
  } while (st.es_eta * l2norm(dz, S) > k.kappa * k.tol);
  return conv_ok;
}

// ESDIRK23a::do_step (ESDIRK23a.cpp:134-199).
bool es_do_step(const VariantOps& m, const SolverParams& sp, const EsdirkCoef& k, ImplicitState& st, double* y,
                double t0, double dt, const double* p, const double* c, Work& w)
{
  const int S = m.S;
  double *z1 = w.z1.data(), *z2 = w.z2.data(), *z3 = w.z3.data(), *z4 = w.z4.data(), *yh = w.yh.data(),
         *prev = w.prev.data();
  m.eval(y, t0, p, c, z1);
  st.nfevals++;
  for (int i = 0; i < S; i++) {
    z2[i] = z1[i];
    prev[i] = z1[i] * k.a21;
  }
  bool ok = es_stage(m, sp, k, st, z2, y, prev, t0 + k.c2 * dt, dt, p, c, w);
  if (ok) {
    for (int i = 0; i < S; i++) {
      z3[i] = z2[i];
      prev[i] = z1[i] * k.bh1 + z2[i] * k.bh2;
    }
    ok = es_stage(m, sp, k, st, z3, y, prev, t0 + k.c3 * dt, dt, p, c, w);
  }
  if (ok) {
    for (int i = 0; i < S; i++) {
      z4[i] = z3[i];
      prev[i] = z1[i] * k.b1 + z2[i] * k.b2 + z3[i] * k.b3;
    }
    ok = es_stage(m, sp, k, st, z4, y, prev, t0 + k.c4 * dt, dt, p, c, w);
  }
  if (ok) {
    for (int i = 0; i < S; i++)
      yh[i] = dt * ((k.b1 - k.bh1) * z1[i] + (k.b2 - k.bh2) * z2[i] + (k.b3 - k.gamma) * z3[i] + k.gamma * z4[i]);
    fb_subst(S, st.jfac.data(), yh, yh);
    st.loc_error = l2norm(yh, S);
    if (st.loc_error > k.tol) {
      st.num_rejected++;
      ok = false;
    } else {
      for (int i = 0; i < S; i++) y[i] = y[i] + dt * (k.b1 * z1[i] + k.b2 * z2[i] + k.b3 * z3[i] + k.gamma * z4[i]);
      st.accepted++;
    }
  }
  return ok;
}

// ESDIRK23a::reduce_time_step (ESDIRK23a.cpp:260-278).
void es_reduce(const EsdirkCoef& k, const SolverParams& sp, const ImplicitState& st, double& dt)
{
  double reduction_factor = 1.0;
  double e = st.loc_error;
  int passes = 0;  // ends the loop for an infinite loc_error (a finite one needs at most 313 passes)
  do {
    reduction_factor *= 0.6;
    e *= 0.1;
  } while (e > k.tol && ++passes < 400);
  dt = std::max(reduction_factor * dt, sp.min_dt);
}

// ESDIRK23a::advance_one_step (ESDIRK23a.cpp:117-132).
void es_advance(const VariantOps& m, const SolverParams& sp, const EsdirkCoef& k, ImplicitState& st, double* y,
                double t0, double& dt, const double* p, const double* c, Work& w)
{
  bool accepted = es_do_step(m, sp, k, st, y, t0, dt, p, c, w);
  while (!accepted && dt > sp.min_dt) {
    if (!st.new_jacobian) es_ode_jacobian(m, st, y, t0, p, c, w);
    es_reduce(k, sp, st, dt);
    es_factorized(m, k, st, dt);
    accepted = es_do_step(m, sp, k, st, y, t0, dt, p, c, w);
  }
  st.new_jacobian = false;
}

// ESDIRK23a::forward (ESDIRK23a.cpp:280-316).  The internal time step of the
// reference's ODESolver base is shadowed and unused (AdaptiveImplicitSolver.h:139).
void esdirk_forward(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* y, double t0,
                    double interval, const double* p, const double* c, Work& w)
{
  static const EsdirkCoef k;
  double local_t = 0;
  double dt = std::max(k.s_fac * st.prev_dt * std::pow((k.tol / st.loc_error), 0.25), sp.min_dt);
  dt = std::min(dt, interval);
  if (st.first_step) {
    st.prev_dt = dt;
    es_ode_jacobian(m, st, y, t0, p, c, w);
  }
  while (local_t < interval) {
    if (dt * 1.1 >= interval - local_t) dt = interval - local_t;
    if (st.newton_iterations >= 2) es_ode_jacobian(m, st, y, t0 + local_t, p, c, w);
    if (std::fabs((dt - st.prev_dt) / st.prev_dt) > 0.2 || st.new_jacobian) es_factorized(m, k, st, dt);
    es_advance(m, sp, k, st, y, t0 + local_t, dt, p, c, w);
    st.prev_dt = dt;
    local_t += dt;
    dt = std::max(k.s_fac * st.prev_dt * std::pow((k.tol / st.loc_error), 0.25), sp.min_dt);
  }
  st.first_step = false;
}

// BasicImplicitEuler::forward (BasicImplicitEuler.cpp:62-181) with its
// compute_factorized_jacobian (:46-60): full Newton with a new Jacobian at
// every iteration (scaled by -ldt, identity added, LU), residual
// (y - y_prev) - dt f(y, t + ldt).  The residual and the clock advance use the
// whole interval dt, the Jacobian the sub-step ldt, as in the reference.  A
// sub-step whose iteration does not converge within max_iterations keeps its
// last iterate; the reference prints a message, the oracle counts a rejection.
void basic_implicit_euler_forward(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* y,
                                  double t, double dt, const double* p, const double* c, Work& w)
{
  const int S = m.S;
  const double ldt_0 = sp.ldt;
  const unsigned long nsteps = ldt_0 > 0 ? (unsigned long)std::ceil(dt / ldt_0 - 1.0E-12) : 1UL;
  const double ldt = dt / nsteps;
  const double rtol = sp.relative_tolerance;
  double lt = t;
  double* prev = w.prev.data();
  double* f1 = w.f1.data();
  double* b = w.b.data();
  double* dz = w.dz.data();
  for (unsigned long step = 0; step < nsteps; ++step) {
    double relative_residual = 1.0, initial_residual = 1.0, residual;
    bool newton_converged = false;
    st.newton_iterations = 0;
    for (int i = 0; i < S; ++i) prev[i] = y[i];
    while (!newton_converged && st.newton_iterations < sp.max_iterations) {
      m.eval(y, lt + ldt, p, c, f1);
      for (int i = 0; i < S; ++i) b[i] = (y[i] - prev[i]) * m.mass(i) - dt * f1[i];
      residual = l2norm(b, S);
      if (st.newton_iterations == 0) initial_residual = residual;
      relative_residual = residual / initial_residual;
      if (relative_residual < rtol) {
        newton_converged = true;
        break;
      }
      ie_factorized_jacobian(m, st, y, lt + ldt, ldt, 1.0, p, c, w);
      fb_subst(S, st.jac.data(), b, dz);
      for (int i = 0; i < S; ++i) y[i] -= dz[i];
      ++st.newton_iterations;
    }
    if (!newton_converged)
      st.rejects++;
    else
      lt += dt;
  }
}

// AdaptiveImplicitSolver::dtinit (AdaptiveImplicitSolver.cpp:68-128).
double ai_dtinit(const VariantOps& m, const SolverParams& sp, double t, const double* y0, double* y1,
                 const double* f0, double* f1, double iord, const double* p, const double* c)
{
  const int S = m.S;
  double dnf = 0.0, dny = 0.0, sk, dt, tmp;
  for (int i = 0; i < S; ++i) {
    sk = sp.atol + sp.rtol * std::fabs(y0[i]);
    tmp = f0[i] / sk;
    dnf += tmp * tmp;
    tmp = y0[i] / sk;
    dny += tmp * tmp;
  }
  if (dnf <= 1.0e-10 || dny <= 1.0e-10)
    dt = 1.0e-6;
  else
    dt = 0.01 * std::sqrt(dny / dnf);
  for (int i = 0; i < S; ++i) y1[i] = y0[i] + dt * f0[i];
  m.eval(y1, t + dt, p, c, f1);
  double der2 = 0.0;
  for (int i = 0; i < S; ++i) {
    sk = sp.atol + sp.rtol * std::fabs(y1[i]);
    tmp = ((f1[i] - f0[i]) / sk);
    der2 += tmp * tmp;
  }
  der2 = std::sqrt(der2) / dt;
  double der12;
  if (std::fabs(der2) >= std::sqrt(dnf))
    der12 = std::fabs(der2);
  else
    der12 = std::sqrt(dnf);
  double dt1;
  if (der12 <= 1.0e-15) {
    if (1.0e-6 > std::fabs(dt) * 1.0e-3)
      dt1 = 1.0e-6;
    else
      dt1 = std::fabs(dt) * 1.0e-3;
  } else {
    dt1 = std::pow((0.01 / der12), (1.0 / iord));
  }
  if (100 * std::fabs(dt) < dt1) return 100 * std::fabs(dt);
  return dt1;
}

// Passes of the ESDIRK4O32 step loop after which the oracle gives up (the
// reference loops until it reaches the interval end, which a non-finite error
// or a step size halved to zero never does).
constexpr long kEsdirk4MaxPasses = 1000000;

// ESDIRK4O32::forward (ESDIRK4O32.cpp:92-298) with
// AdaptiveImplicitSolver::new_time_step (AdaptiveImplicitSolver.cpp:130-200),
// as the reference computes it:
//  * the second stage's right-hand side is evaluated at the stage increment z2
//    and the third stage starts from z2 + y (ESDIRK4O32.cpp:163-183);
//  * the solution lives in two buffers, the caller's y (A) and the solver's
//    member yn (B).  `y` points to A at the start; each accepted step swaps
//    the pointers `y` and `yn0` (:272-276), while the new solution is always
//    written to B by name (:255-262).  Once `y` points to B, the next step
//    overwrites y in place, the error test compares it with the stale A, and
//    an accepted step swaps back to A.  At the end `y` is copied to the
//    caller's buffer if it points to B (:290-297).
// The member yn persists in the solver instance (st.e4_yn); it is written
// before it is read in every call.
void esdirk4o32_forward(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* y, double t,
                        double interval, const double* p, const double* c, Work& w)
{
  static const EsdirkCoef k;  // the same tableau (ESDIRK4O32.cpp:29-37)
  const double iord = 3.0;    // ESDIRK4O32.cpp:38
  const double facmin = 0.5, facmaxb = 2.0, stabfac = 0.9, stabdown = 1.0, stabup = 1.2;
  const int S = m.S;
  if (st.e4_yn.size() != size_t(S)) st.e4_yn.assign(S, 0.0);
  double *z1 = w.z1.data(), *z2 = w.z2.data(), *z3 = w.z3.data(), *z4 = w.z4.data(), *yh = w.yh.data(),
         *prev = w.prev.data(), *f1 = w.f1.data();
  double* const ret = y;
  double* yn = st.e4_yn.data();  // the member vector (buffer B)
  double* yp = y;                // the pointer `y`
  double* yn0 = yn;              // the pointer `yn0`
  double ti = t;
  const double t_end = ti + interval;
  double dt = st.ai_ldt;
  bool reached_tend = false;
  if (dt < 0) {
    m.eval(yp, ti, p, c, z1);
    dt = ai_dtinit(m, sp, ti, yp, yn0, z1, z2, iord, p, c);
    st.e4_nfevals += 2;
  }
  long passes = 0;
  while (!reached_tend) {
    if (++passes > kEsdirk4MaxPasses) throw std::runtime_error("ESDIRK4O32 did not reach the end of the interval");
    m.eval(yp, ti, p, c, z1);
    st.e4_nfevals += 1;
    for (int i = 0; i < S; ++i) {
      prev[i] = k.a21 * z1[i];
      z2[i] = 0.0;
    }
    if (!newton_solve(m, sp, st, z2, prev, yp, ti + k.c2 * dt, dt, k.a22, sp.always_recompute_jacobian, p, c, w)) {
      st.recompute_jacobian = true;
      dt /= 2.0;
      continue;
    }
    for (int i = 0; i < S; ++i) z3[i] = z2[i] + yp[i];
    m.eval(z2, ti + k.c2 * dt, p, c, f1);
    st.e4_nfevals += 1;
    for (int i = 0; i < S; ++i) {
      z2[i] = f1[i];
      prev[i] = k.a31 * z1[i] + k.a32 * z2[i];
    }
    if (!newton_solve(m, sp, st, z3, prev, yp, ti + k.c3 * dt, dt, k.a33, sp.always_recompute_jacobian, p, c, w)) {
      st.recompute_jacobian = true;
      dt /= 2.0;
      continue;
    }
    for (int i = 0; i < S; ++i) z3[i] += yp[i];
    m.eval(z3, ti + k.c3 * dt, p, c, f1);
    st.e4_nfevals += 1;
    for (int i = 0; i < S; ++i) {
      z3[i] = f1[i];
      yh[i] = yp[i] + dt * (k.bh1 * z1[i] + k.bh2 * z2[i] + k.bh3 * z3[i]);
    }
    for (int i = 0; i < S; ++i) {
      z4[i] = yh[i] - yp[i];
      prev[i] = k.a41 * z1[i] + k.a42 * z2[i] + k.a43 * z3[i];
    }
    if (!newton_solve(m, sp, st, z4, prev, yp, ti + k.c4 * dt, dt, k.a44, sp.always_recompute_jacobian, p, c, w)) {
      st.recompute_jacobian = true;
      dt /= 2.0;
      continue;
    }
    for (int i = 0; i < S; ++i) z4[i] += yp[i];
    m.eval(z4, ti + k.c4 * dt, p, c, f1);
    st.e4_nfevals += 1;
    for (int i = 0; i < S; ++i) {
      yn[i] = yp[i] + dt * (k.b1 * z1[i] + k.b2 * z2[i] + k.b3 * z3[i] + k.b4 * f1[i]);
      yh[i] -= yn[i];
    }
    // new_time_step(y, yn0, yh, t_end)
    const double dt_prev = dt;
    st.recompute_jacobian = true;
    double err = 0.0;
    for (int i = 0; i < S; ++i) {
      const double yi_abs = std::fabs(yp[i]);
      const double yni_abs = std::fabs(yn0[i]);
      const double mx = std::max(yi_abs, yni_abs);
      const double sk = sp.atol + sp.rtol * mx;
      const double tmp = yh[i] / sk;
      err += tmp * tmp;
    }
    err = std::sqrt(err / S);
    bool step_accepted;
    if (err <= 1) {
      ti += dt;
      st.accepted += 1;
      step_accepted = true;
      if (std::fabs(ti - t_end) < 1e-14) reached_tend = true;
    } else {
      st.num_rejected += 1;
      step_accepted = false;
    }
    const double lstabfac = stabfac * (2 * sp.max_iterations + 1) / (2.0 * sp.max_iterations + st.newton_iterations);
    double fac = lstabfac * std::pow((1.0 / err), (1.0 / (iord + 1)));
    if (facmin > fac)
      fac = facmin;
    else if (fac > st.ai_facmax)
      fac = st.ai_facmax;
    if (!step_accepted)
      st.ai_facmax = 1.0;
    else
      st.ai_facmax = facmaxb;
    if (st.err_old > 0) fac *= dt / st.dt_old * std::pow((st.err_old / err), (1.0 / (iord + 1)));
    if (fac < stabup && fac > stabdown) {
      fac = 1.0;
      st.recompute_jacobian = false;
    }
    dt *= fac;
    st.ai_ldt = dt;
    if (ti + dt >= t_end) dt = t_end - ti;
    st.dt_old = dt_prev;
    st.err_old = err;
    if (step_accepted) std::swap(yp, yn0);
  }
  if (ret != yp)
    for (int i = 0; i < S; ++i) ret[i] = yp[i];
}

// ---------------------------------------------------------------------------
// RKF32 (RKF32.cpp:105-222) with AdaptiveExplicitSolver::new_time_step
// (AdaptiveExplicitSolver.cpp:168-244)
// ---------------------------------------------------------------------------

// Sub-steps one forward call may take before it is declared stuck.  The
// reference loops until |_t - t_end| < 1e-14 (AdaptiveExplicitSolver.cpp:206)
// and never returns when that test cannot be met (a step size that has
// underflowed after repeated rejections of a non-finite error).
constexpr long kRkf32MaxSteps = 1000000;

// AdaptiveExplicitSolver::dtinit (AdaptiveExplicitSolver.cpp:67-166) with _itol = 0: an initial step from
// the scaled norms of y0 and f0 and a difference estimate of the second derivative after one Euler step.
// y1 and f1 are scratch.
double rkf32_dtinit(const VariantOps& m, const SolverParams& sp, double t, const double* y0, double* y1,
                    const double* f0, double* f1, double iord, const double* p, const double* c)
{
  const int S = m.S;
  double dnf = 0.0, dny = 0.0, sk, dt, tmp;
  for (int i = 0; i < S; ++i) {
    sk = sp.atol + sp.rtol * std::fabs(y0[i]);
    tmp = f0[i] / sk;
    dnf += tmp * tmp;
    tmp = y0[i] / sk;
    dny += tmp * tmp;
  }
  if (dnf <= 1.0e-10 || dny <= 1.0e-10)
    dt = 1.0e-6;
  else
    dt = 0.01 * std::sqrt(dny / dnf);
  for (int i = 0; i < S; ++i) y1[i] = y0[i] + dt * f0[i];
  m.eval(y1, t + dt, p, c, f1);
  double der2 = 0.0;
  for (int i = 0; i < S; ++i) {
    sk = sp.atol + sp.rtol * std::fabs(y1[i]);
    tmp = (f1[i] - f0[i]) / sk;
    der2 += tmp * tmp;
  }
  der2 = std::sqrt(der2) / dt;
  double der12;
  if (std::fabs(der2) >= std::sqrt(dnf))
    der12 = std::fabs(der2);
  else
    der12 = std::sqrt(dnf);
  double dt1;
  if (der12 <= 1.0e-15) {
    if (std::fabs(dt) * 1.0e-3 < 1.0e-6)
      dt1 = 1.0e-6;
    else
      dt1 = std::fabs(dt) * 1.0e-3;
  } else {
    dt1 = std::pow((0.01 / der12), (1.0 / iord));
  }
  if (100 * std::fabs(dt) < dt1)
    return 100 * std::fabs(dt);
  else
    return dt1;
}

void rkf32_forward(const VariantOps& m, const SolverParams& sp, ImplicitState& st, double* y_out, double t,
                   double interval, const double* p, const double* c, Work& w)
{
  const int S = m.S;
  // Coefficients (RKF32.cpp:34-37).
  const double a21 = 1.0 / 2.0, a32 = 3.0 / 4.0, b1 = 2.0 / 9.0, b2 = 1.0 / 3.0, b3 = 4.0 / 9.0;
  const double bh1 = 7.0 / 24.0, bh2 = 1.0 / 4.0, bh3 = 1.0 / 3.0, bh4 = 1.0 / 8.0;
  const double d1 = b1 - bh1, d2 = b2 - bh2, d3 = b3 - bh3, d4 = -bh4, c2 = 1.0 / 2.0, c3 = 3.0 / 4.0;
  // AdaptiveExplicitSolver::reset values (AdaptiveExplicitSolver.cpp:58-64); _iord = 3 (RKF32.cpp:39).
  const double facmin = 0.5, facmaxb = 2.0, stabfac = 0.9, iord = 3.0;
  // The reference swaps the y / yn and k1 / k4 buffers on acceptance; the
  // values are copied here.
  double* y = y_out;
  double *k1 = w.k1.data(), *k2 = w.k2.data(), *k3 = w.k3.data(), *k4 = w.k4.data();
  double *ki = w.ki.data(), *yn = w.yn.data(), *e = w.e.data();
  const double t_end = t + interval;
  bool reached_tend = false;
  double lt = t;  // _t
  // `first` is set by reset() and never cleared (RKF32.cpp:126-131).
  m.eval(y, t, p, c, k1);
  st.rk_nfevals += 1;
  // _ldt is AdaptiveExplicitSolver::_ldt (0.1 after construction).  It turns negative when a step
  // overshoots t_end (the first step of a call is _ldt, which may exceed the interval) and the next
  // step, t_end - _t < 0, is stored as the guess (AdaptiveExplicitSolver.cpp:234-241); the next call
  // then takes its first step from dtinit (RKF32.cpp:134-141).
  double dt;
  if (st.rk_ldt < 0) {
    dt = rkf32_dtinit(m, sp, t, y, yn, k1, k2, iord, p, c);
    st.rk_nfevals += 1;
  } else {
    dt = st.rk_ldt;
  }
  long guard = 0;
  while (!reached_tend) {
    if (++guard > kRkf32MaxSteps) throw std::runtime_error("RKF32 did not reach the end of the interval");
    for (int i = 0; i < S; ++i) ki[i] = y[i] + dt * a21 * k1[i];
    m.eval(ki, t + c2 * dt, p, c, k2);  // stage times from the call's t (RKF32.cpp:155, 160)
    for (int i = 0; i < S; ++i) ki[i] = y[i] + dt * a32 * k2[i];
    m.eval(ki, t + c3 * dt, p, c, k3);
    for (int i = 0; i < S; ++i) yn[i] = y[i] + dt * (b1 * k1[i] + b2 * k2[i] + b3 * k3[i]);
    m.eval(yn, t + dt, p, c, k4);
    st.rk_nfevals += 3;
    for (int i = 0; i < S; ++i) e[i] = dt * (d1 * k1[i] + d2 * k2[i] + d3 * k3[i] + d4 * k4[i]);
    // new_time_step
    double err = 0.0;
    for (int i = 0; i < S; ++i) {
      const double yi_abs = std::fabs(y[i]), yni_abs = std::fabs(yn[i]);
      const double mx = yi_abs > yni_abs ? yi_abs : yni_abs;
      const double sk = sp.atol + sp.rtol * mx;
      const double tmp = e[i] / sk;
      err += tmp * tmp;
    }
    err = std::sqrt(err / S);
    if (std::isnan(err) || std::isinf(err)) err = 2000.0;
    bool accepted;
    if (err <= 1.0) {
      lt += dt;
      st.accepted += 1;
      accepted = true;
      if (std::fabs(lt - t_end) < 1e-14) reached_tend = true;
    } else {
      st.num_rejected += 1;
      accepted = false;
    }
    double fac = stabfac * std::pow(1.0 / err, 1.0 / (iord + 1));
    if (facmin > fac) fac = facmin;
    if (fac > st.facmax) fac = st.facmax;
    st.facmax = accepted ? facmaxb : 1.0;
    dt *= fac;
    if (lt + dt >= t_end)
      dt = t_end - lt;
    else if (lt + 1.5 * dt >= t_end)
      dt = (t_end - lt) / 2.0;
    else
      st.rk_ldt = dt;
    if (accepted) {
      st.ndtsa += 1;
      for (int i = 0; i < S; ++i) {
        y[i] = yn[i];
        k1[i] = k4[i];
      }
    } else {
      st.ndtsr += 1;
    }
  }
}

}  // namespace

struct oracle_system {
  const VariantOps* m;
  const VariantOps* base;  // the registered model; m points here or at `analytic`
  VariantOps analytic;     // the model with its analytic Jacobian (solver parameter jacobian = 2)
  int kind;
  SolverParams sp;
  unsigned N;
  unsigned num_threads = 0;
  int per_node_state = 0;
  std::vector<double> shared;  // P parameter values
  std::vector<double> states;  // [N][S]
  std::vector<double> fparams;  // [N][NFP]
  std::vector<Work> work;  // one per thread
  std::vector<ImplicitState> node_state;  // per-node mode
};

namespace {

void reset_default(oracle_system* s)
{
  // ODESystemSolver::reset_default (ODESystemSolver.cpp:192-216): initial
  // conditions and the current values of the field parameters.
  const VariantOps& m = *s->m;
  for (unsigned n = 0; n < s->N; ++n) {
    for (int i = 0; i < m.S; ++i) s->states[size_t(n) * m.S + i] = m.ic[i];
    for (int i = 0; i < m.NFP; ++i) s->fparams[size_t(n) * m.NFP + i] = s->shared[m.fpi[i]];
  }
}

void forward_node(oracle_system* s, unsigned node, double t, double interval, Work& w, double* prm)
{
  const VariantOps& m = *s->m;
  // ParameterizedODE::set_field_parameters (ODESystemSolver.h:219-224).
  for (int i = 0; i < m.NFP; ++i) prm[m.fpi[i]] = s->fparams[size_t(node) * m.NFP + i];
  double c[kMaxConstants];
  m.constants(prm, c);
  double* y = &s->states[size_t(node) * m.S];
  if (s->kind == IMPLICIT_EULER || s->kind == ESDIRK23A || s->kind == RKF32 || s->kind == THETA ||
      s->kind == BASIC_IMPLICIT_EULER || s->kind == ESDIRK4O32) {
    ImplicitState& st = s->per_node_state ? s->node_state[node] : w.carried;
    if (s->kind == IMPLICIT_EULER)
      implicit_euler_forward(m, s->sp, st, y, t, interval, prm, c, w);
    else if (s->kind == ESDIRK23A)
      esdirk_forward(m, s->sp, st, y, t, interval, prm, c, w);
    else if (s->kind == THETA)
      theta_forward(m, s->sp, st, y, t, interval, prm, c, w);
    else if (s->kind == BASIC_IMPLICIT_EULER)
      basic_implicit_euler_forward(m, s->sp, st, y, t, interval, prm, c, w);
    else if (s->kind == ESDIRK4O32)
      esdirk4o32_forward(m, s->sp, st, y, t, interval, prm, c, w);
    else
      rkf32_forward(m, s->sp, st, y, t, interval, prm, c, w);
    return;
  }
  forward_explicit(s->kind, m, s->sp, y, t, interval, prm, c, w);
}

// ODESystemSolver::forward (ODESystemSolver.cpp:67-79).
void system_forward(oracle_system* s, double t, double interval)
{
  const int P = s->m->P;
  if (s->num_threads > 0) {
    std::string err;
#pragma omp parallel num_threads(s->num_threads)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      std::vector<double> prm(s->shared);
      Work& w = s->work[tid];
#pragma omp for schedule(static)
      for (long node = 0; node < (long)s->N; node++) {
        try {
          forward_node(s, (unsigned)node, t, interval, w, prm.data());
        } catch (std::exception& e) {
#pragma omp critical
          err = e.what();
        }
      }
    }
    if (!err.empty()) throw std::runtime_error(err);
  } else {
    std::vector<double> prm(s->shared);
    (void)P;
    for (unsigned node = 0; node < s->N; node++) forward_node(s, node, t, interval, s->work[0], prm.data());
  }
}

int set_param(oracle_system* s, const std::string& name, double v)
{
  if (name == "internal_time_step") { s->sp.ldt = v; return 0; }
  if (name == "delta") { s->sp.delta = v; return 0; }
  if (name == "eta_0") { s->sp.eta_0 = v; return 0; }
  if (name == "kappa") { s->sp.kappa = v; return 0; }
  if (name == "relative_tolerance") { s->sp.relative_tolerance = v; return 0; }
  if (name == "max_iterations") { s->sp.max_iterations = int(v); return 0; }
  if (name == "max_relative_previous_residual") { s->sp.max_relative_previous_residual = v; return 0; }
  if (name == "always_recompute_jacobian") { s->sp.always_recompute_jacobian = v != 0; return 0; }
  if (name == "min_dt") { s->sp.min_dt = v; return 0; }
  if (name == "atol") { s->sp.atol = v; return 0; }
  if (name == "rtol") { s->sp.rtol = v; return 0; }
  if (name == "theta") { s->sp.theta = v; return 0; }
  if (name == "jacobian") {
    // 0: forward differences (ODE.cpp:47-67); 2: the model's analytic Jacobian.
    if (v == 0) { s->m = s->base; return 0; }
    if (v == 2) {
      if (s->base->ext) { g_err = "a model registered at run time brings its own compute_jacobian"; return -1; }
      s->analytic = *s->base;
      s->analytic.analytic_jacobian = true;
      s->m = &s->analytic;
      return 0;
    }
    g_err = "jacobian must be 0 (forward differences) or 2 (analytic)";
    return -1;
  }
  if (name == "num_refinements_without_always_recomputing_jacobian") {
    s->sp.num_refinements_without_always_recomputing_jacobian = int(v);
    return 0;
  }
  g_err = "unknown solver parameter " + name;
  return -1;
}

inline unsigned field_index(bool tangled, unsigned N, unsigned F, unsigned node, unsigned i)
{
  // ODESystemSolver.h:244.
  return tangled ? node * F + i : N * i + node;
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_variant_count(void) { return kNumVariants; }
const char* oracle_variant_name(int i) { return (i >= 0 && i < kNumVariants) ? kVariants[i].name : nullptr; }
int oracle_variant_find(const char* name)
{
  for (int i = 0; i < kNumVariants; ++i)
    if (std::strcmp(kVariants[i].name, name) == 0) return i;
  return -1;
}
void oracle_variant_sizes(int v, int* sz)
{
  sz[0] = kVariants[v].S;
  sz[1] = kVariants[v].P;
  sz[2] = kVariants[v].NFS;
  sz[3] = kVariants[v].NFP;
}
void oracle_variant_defaults(int v, double* p, double* ic)
{
  std::memcpy(p, kVariants[v].defaults, sizeof(double) * kVariants[v].P);
  std::memcpy(ic, kVariants[v].ic, sizeof(double) * kVariants[v].S);
}
void oracle_variant_field_indices(int v, int* fsi, int* fpi)
{
  for (int i = 0; i < kVariants[v].NFS; ++i) fsi[i] = kVariants[v].fsi[i];
  for (int i = 0; i < kVariants[v].NFP; ++i) fpi[i] = kVariants[v].fpi[i];
}
void oracle_variant_linear_terms(int v, unsigned char* f)
{
  std::memcpy(f, kVariants[v].linear_terms, kVariants[v].S);
}
void oracle_variant_differential_states(int v, unsigned char* f)
{
  for (int i = 0; i < kVariants[v].S; ++i) f[i] = kVariants[v].diff ? kVariants[v].diff[i] : 1;
}
void oracle_model_eval(int v, const double* y, double t, const double* p, double* out)
{
  double c[kMaxConstants];
  kVariants[v].constants(p, c);
  kVariants[v].eval(y, t, p, c, out);
}
void oracle_model_linearized_eval(int v, const double* y, double t, const double* p, double* lin, double* rhs,
                                  int only_linear)
{
  double c[kMaxConstants];
  kVariants[v].constants(p, c);
  kVariants[v].linearized(y, t, p, c, lin, rhs, only_linear != 0);
}
void oracle_model_analytic_jacobian(int v, const double* y, double t, const double* p, double* jac)
{
  double c[kMaxConstants];
  kVariants[v].constants(p, c);
  if (!kVariants[v].jacobian) {  // a model registered at run time has none
    for (int i = 0; i < kVariants[v].S * kVariants[v].S; ++i) jac[i] = std::nan("");
    return;
  }
  kVariants[v].jacobian(y, t, p, c, jac);
}
void oracle_model_fd_jacobian(int v, double* y, double t, const double* p, double* jac)
{
  std::vector<double> f1(kVariants[v].S), f2(kVariants[v].S);
  double c[kMaxConstants];
  kVariants[v].constants(p, c);
  fd_jacobian(kVariants[v], y, t, p, c, jac, f1.data(), f2.data());
}
// Generated step adjoints for one step y -> y' of size h at time t: kind 0 GRL1 and 1 RL1 give
// out = (d y'/d y)^T lam and fp_bar += (d y'/d p_field)^T lam (grl1_vjp, rl1_vjp); kind 2 gives
// out += (d rhs/d y)^T lam and fp_bar += (d rhs/d p_field)^T lam (eval_vjp).  tape: the *_tape form.
void oracle_model_step_vjp(int v, int kind, int tape, const double* y, double t, const double* p, double h,
                           double delta, const double* lam, double* out, double* fp_bar)
{
  double c[kMaxConstants];
  kVariants[v].constants(p, c);
  if (!kVariants[v].step_vjp) {  // a model registered at run time has none
    for (int i = 0; i < kVariants[v].S; ++i) out[i] = std::nan("");
    return;
  }
  kVariants[v].step_vjp(kind, tape != 0, y, t, p, c, h, delta, lam, out, fp_bar);
}
void oracle_lu_factorize(int n, double* mat) { lu_factorize(n, mat); }
void oracle_forward_backward_subst(int n, const double* mat, const double* b, double* x) { fb_subst(n, mat, b, x); }

int oracle_register_external_model(const oracle_external_model* e)
{
  if (!e || !e->name || !e->eval || !e->linearized_eval || e->num_states <= 0 || e->num_parameters < 0) {
    g_err = "incomplete external model";
    return -1;
  }
  const int found = oracle_variant_find(e->name);
  if (found >= 0) {
    if (kVariants[found].ext == e) return found;
    g_err = std::string("a model named ") + e->name + " is already registered";
    return -1;
  }
  VariantOps m{};
  m.name = e->name;
  m.S = e->num_states;
  m.P = e->num_parameters;
  m.NFS = e->num_field_states;
  m.NFP = e->num_field_parameters;
  m.fsi = e->field_state_indices;
  m.fpi = e->field_parameter_indices;
  m.defaults = e->default_parameters;
  m.ic = e->initial_conditions;
  m.linear_terms = e->linear_terms;
  m.diff = e->differential_states;
  m.pnames = e->parameter_names;
  m.K = 0;
  m.constants = [](const double*, double*) {};
  m.ext = e;
  variant_table().push_back(m);
  return kNumVariants - 1;
}

int oracle_system_create(const char* variant, const char* solver, unsigned num_nodes, int n_params,
                         const char* const* names, const double* values, int per_node_state, oracle_system** out)
{
  const int v = oracle_variant_find(variant);
  if (v < 0) { g_err = std::string("unknown model variant ") + variant; return -1; }
  const int kind = solver_kind(solver);
  if (kind < 0) { g_err = std::string("unknown solver ") + solver; return -1; }
  if (kVariants[v].K > kMaxConstants) { g_err = "model has too many constants"; return -1; }
  // Solvers that refuse a DAE when the model is attached (RL1.cpp:54-56; GRL1.cpp:56-58, RKF32.cpp:75-77,
  // ESDIRK4O32.cpp:72-74 and the other explicit solvers).
  if (kVariants[v].is_dae() && kind == RL1) { g_err = "cannot integrate a DAE ode with Rush Larsen method."; return -1; }
  if (kVariants[v].is_dae() && (kind <= GRL2 || kind == RKF32 || kind == ESDIRK4O32)) {
    g_err = "cannot integrate a DAE ode with an explicit solver.";
    return -1;
  }
  auto* s = new oracle_system;
  s->m = &kVariants[v];
  s->base = s->m;
  s->kind = kind;
  if (kind == ESDIRK23A) s->sp.min_dt = 0.001;  // ESDIRK23a.h:38
  s->N = num_nodes;
  s->per_node_state = per_node_state;
  s->shared.assign(s->m->defaults, s->m->defaults + s->m->P);
  for (int i = 0; i < n_params; ++i) {
    if (set_param(s, names[i], values[i]) != 0) { delete s; return -1; }
  }
  s->states.assign(size_t(num_nodes) * s->m->S, 0.0);
  s->fparams.assign(size_t(num_nodes) * s->m->NFP, 0.0);
  s->work.resize(1);
  s->work[0].size(s->m->S);
  s->work[0].carried.init(s->m->S, s->sp);
  if (per_node_state) {
    s->node_state.resize(num_nodes);
    for (auto& st : s->node_state) st.init(s->m->S, s->sp);
  }
  reset_default(s);
  *out = s;
  return 0;
}

void oracle_system_destroy(oracle_system* s) { delete s; }

int oracle_system_set_parameter(oracle_system* s, const char* name, double value)
{
  // A model parameter (ParameterizedODE::set_parameter, ParameterizedODE.h:111-114)
  // or a solver tunable.
  for (int k = 0; k < s->m->P; ++k) {
    if (std::strcmp(name, s->m->pnames[k]) == 0) {
      s->shared[k] = value;
      return 0;
    }
  }
  return set_param(s, name, value);
}

void oracle_system_set_num_threads(oracle_system* s, unsigned n)
{
  // ODESystemSolver::set_num_threads (ODESystemSolver.cpp:218-250).
  // Each thread gets a fresh copy of the solver (ESDIRK23a's copy constructor
  // restores first_step, loc_error, prev_dt and eta; ESDIRK23a.cpp:46-61).
  s->num_threads = n;
  s->work.resize(n > 0 ? n : 1);
  for (auto& w : s->work) {
    w.size(s->m->S);
    w.carried.init(s->m->S, s->sp);
  }
}

int oracle_system_forward(oracle_system* s, double t, double interval)
{
  try {
    system_forward(s, t, interval);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int oracle_system_forward_times(oracle_system* s, const double* t, unsigned long n)
{
  try {
    double tc = t[0];
    for (unsigned long it = 1; it < n; ++it) {
      const double tn = t[it];
      system_forward(s, tc, tn - tc);
      tc = tn;
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int oracle_system_solve(oracle_system* s, double* field_states, const double* t, unsigned long n, int tangled)
{
  // ODESystemSolver::solve (ODESystemSolver.cpp:81-99): row 0 is left to the caller.
  try {
    const size_t row = size_t(s->N) * s->m->NFS;
    double tc = t[0];
    for (unsigned long it = 1; it < n; ++it) {
      const double tn = t[it];
      system_forward(s, tc, tn - tc);
      oracle_system_get_field_states(s, field_states + it * row, tangled);
      tc = tn;
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void oracle_system_get_states(const oracle_system* s, double* out)
{
  std::memcpy(out, s->states.data(), s->states.size() * sizeof(double));
}
void oracle_system_set_states(oracle_system* s, const double* in)
{
  std::memcpy(s->states.data(), in, s->states.size() * sizeof(double));
}

void oracle_system_get_field_states(const oracle_system* s, double* out, int tangled)
{
  const unsigned F = s->m->NFS, S = s->m->S;
  for (unsigned n = 0; n < s->N; ++n)
    for (unsigned i = 0; i < F; ++i) out[field_index(tangled, s->N, F, n, i)] = s->states[size_t(n) * S + s->m->fsi[i]];
}
void oracle_system_set_field_states(oracle_system* s, const double* in, int tangled)
{
  const unsigned F = s->m->NFS, S = s->m->S;
  for (unsigned n = 0; n < s->N; ++n)
    for (unsigned i = 0; i < F; ++i) s->states[size_t(n) * S + s->m->fsi[i]] = in[field_index(tangled, s->N, F, n, i)];
}
void oracle_system_get_field_state_components(const oracle_system* s, double* out, unsigned nc, const unsigned* comp,
                                              int tangled)
{
  const unsigned F = s->m->NFS, S = s->m->S;
  for (unsigned n = 0; n < s->N; ++n)
    for (unsigned c = 0; c < nc; ++c) {
      const unsigned i = comp[c];
      out[field_index(tangled, s->N, F, n, i)] = s->states[size_t(n) * S + s->m->fsi[i]];
    }
}
void oracle_system_set_field_state_components(oracle_system* s, const double* in, unsigned nc, const unsigned* comp,
                                              int tangled)
{
  const unsigned F = s->m->NFS, S = s->m->S;
  for (unsigned n = 0; n < s->N; ++n)
    for (unsigned c = 0; c < nc; ++c) {
      const unsigned i = comp[c];
      s->states[size_t(n) * S + s->m->fsi[i]] = in[field_index(tangled, s->N, F, n, i)];
    }
}
void oracle_system_get_field_parameters(const oracle_system* s, double* out, int tangled)
{
  const unsigned F = s->m->NFP;
  for (unsigned n = 0; n < s->N; ++n)
    for (unsigned i = 0; i < F; ++i) out[field_index(tangled, s->N, F, n, i)] = s->fparams[size_t(n) * F + i];
}
void oracle_system_set_field_parameters(oracle_system* s, const double* in, int tangled)
{
  const unsigned F = s->m->NFP;
  for (unsigned n = 0; n < s->N; ++n)
    for (unsigned i = 0; i < F; ++i) s->fparams[size_t(n) * F + i] = in[field_index(tangled, s->N, F, n, i)];
}
void oracle_system_reset_default(oracle_system* s) { reset_default(s); }
// ESDIRK4O32: accepted steps, steps rejected by the error test, right-hand-side evaluations.
void oracle_system_esdirk4o32_counters(const oracle_system* s, long* out3)
{
  for (int i = 0; i < 3; ++i) out3[i] = 0;
  auto add = [&](const ImplicitState& st) {
    out3[0] += st.accepted;
    out3[1] += st.num_rejected;
    out3[2] += st.e4_nfevals;
  };
  if (s->per_node_state)
    for (const auto& st : s->node_state) add(st);
  else
    for (const auto& w : s->work) add(w.carried);
}

void oracle_system_counters(const oracle_system* s, long* out4)
{
  for (int i = 0; i < 4; ++i) out4[i] = 0;
  // Implicit solvers: Newton iterations of the last solve, Jacobians,
  // rejections, accepted steps.  RKF32: 0, RHS evaluations (nfevals),
  // rejected steps (ndtsr), accepted steps (ndtsa).
  const bool rk = s->kind == RKF32;
  auto add = [&](const ImplicitState& st) {
    out4[0] += rk ? 0 : st.newton_iterations;
    out4[1] += rk ? st.rk_nfevals : st.jac_comp;
    out4[2] += rk ? st.ndtsr : st.rejects + st.num_rejected;
    out4[3] += rk ? st.ndtsa : st.accepted;
  };
  if (s->per_node_state)
    for (const auto& st : s->node_state) add(st);
  else
    for (const auto& w : s->work) add(w.carried);
}

}  // extern "C"
