// CPU oracle: restatement of goss's ODESystemSolver and solvers.
// TEST INFRASTRUCTURE ONLY -- see goss_oracle.h.
#include "goss_oracle.h"

#include <cmath>
#include <cstring>
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
  const char* const* pnames;
  int K;  // number of parameter-only constants
  void (*constants)(const double* p, double* c);
  void (*eval)(const double*, double, const double*, const double*, double*);
  void (*linearized)(const double*, double, const double*, const double*, double*, double*, bool);
  void (*jacobian)(const double*, double, const double*, const double*, double*);
};

#define GOSS_ORACLE_VARIANT(V)                                                                         \
  VariantOps{#V, goss_models::V::num_states, goss_models::V::num_parameters,                           \
             goss_models::V::num_field_states, goss_models::V::num_field_parameters,                   \
             goss_models::V::field_state_indices, goss_models::V::field_parameter_indices,             \
             goss_models::V::default_parameters, goss_models::V::initial_conditions,                   \
             goss_models::V::linear_terms, goss_models::V::parameter_names,                            \
             goss_models::V::num_constants,                                                            \
             [](const double* p, double* c) { goss_models::V::compute_constants<double>(Acc{p, nullptr}, c); }, \
             [](const double* y, double t, const double* p, const double* c, double* out) {            \
               goss_models::V::eval<double>(y, t, Acc{p, c}, out);                                     \
             },                                                                                        \
             [](const double* y, double t, const double* p, const double* c, double* lin, double* rhs,  \
                bool ol) { goss_models::V::linearized_eval<double>(y, t, Acc{p, c}, lin, rhs, ol); },   \
             [](const double* y, double t, const double* p, const double* c, double* jac) {            \
               goss_models::V::compute_jacobian<double>(y, t, Acc{p, c}, jac);                         \
             }}

const VariantOps kVariants[] = {
    GOSS_ORACLE_VARIANT(tp06_m),         GOSS_ORACLE_VARIANT(tp06_m_fp3),
    GOSS_ORACLE_VARIANT(tp06_m_fp6),     GOSS_ORACLE_VARIANT(tp06_epi),
    GOSS_ORACLE_VARIANT(fitzhughnagumo), GOSS_ORACLE_VARIANT(lorentz),
    GOSS_ORACLE_VARIANT(fhn),            GOSS_ORACLE_VARIANT(beeler_reuter_1977),
    GOSS_ORACLE_VARIANT(oscilator),      GOSS_ORACLE_VARIANT(linear_decay),
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

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

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------

enum SolverKind { EE, RK2, RK4, RL1, RL2, GRL1, GRL2, IMPLICIT_EULER, ESDIRK23A, RKF32, NUM_SOLVERS };

int solver_kind(const std::string& s)
{
  static const char* names[] = {"ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2",
                                "ImplicitEuler", "ESDIRK23a", "RKF32"};
  for (int i = 0; i < NUM_SOLVERS; ++i)
    if (s == names[i]) return i;
  return -1;
}

struct SolverParams {
  double ldt = -1.0;  // ODESolver::_ldt (ODESolver.h:137)
  double delta = 1e-8;  // GRL1.h:39
};

// Sub-step count of the fixed-step solvers, e.g. GRL1.cpp:65-67.
inline unsigned long substeps(double ldt_0, double dt)
{
  return ldt_0 > 0 ? (unsigned long)std::ceil(dt / ldt_0 - 1.0E-12) : 1UL;
}

// Per-thread scratch, sized to the model.
struct Work {
  std::vector<double> f1, f2, k1, k2, k3, k4, tmp, y2;
  void size(int S)
  {
    for (auto* v : {&f1, &f2, &k1, &k2, &k3, &k4, &tmp, &y2}) v->assign(S, 0.0);
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

}  // namespace

struct oracle_system {
  const VariantOps* m;
  int kind;
  SolverParams sp;
  unsigned N;
  unsigned num_threads = 0;
  int per_node_state = 0;
  std::vector<double> shared;  // P parameter values
  std::vector<double> states;  // [N][S]
  std::vector<double> fparams;  // [N][NFP]
  std::vector<Work> work;  // one per thread
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
  forward_explicit(s->kind, m, s->sp, &s->states[size_t(node) * m.S], t, interval, prm, c, w);
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
  kVariants[v].jacobian(y, t, p, c, jac);
}
void oracle_model_fd_jacobian(int v, double* y, double t, const double* p, double* jac)
{
  std::vector<double> f1(kVariants[v].S), f2(kVariants[v].S);
  double c[kMaxConstants];
  kVariants[v].constants(p, c);
  fd_jacobian(kVariants[v], y, t, p, c, jac, f1.data(), f2.data());
}
void oracle_lu_factorize(int n, double* mat) { lu_factorize(n, mat); }
void oracle_forward_backward_subst(int n, const double* mat, const double* b, double* x) { fb_subst(n, mat, b, x); }

int oracle_system_create(const char* variant, const char* solver, unsigned num_nodes, int n_params,
                         const char* const* names, const double* values, int per_node_state, oracle_system** out)
{
  const int v = oracle_variant_find(variant);
  if (v < 0) { g_err = std::string("unknown model variant ") + variant; return -1; }
  const int kind = solver_kind(solver);
  if (kind < 0) { g_err = std::string("unknown solver ") + solver; return -1; }
  if (kVariants[v].K > kMaxConstants) { g_err = "model has too many constants"; return -1; }
  if (kind >= IMPLICIT_EULER) { g_err = std::string("solver not available in the oracle: ") + solver; return -1; }
  auto* s = new oracle_system;
  s->m = &kVariants[v];
  s->kind = kind;
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
  s->num_threads = n;
  s->work.resize(n > 0 ? n : 1);
  for (auto& w : s->work) w.size(s->m->S);
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
void oracle_system_counters(const oracle_system*, long* out4)
{
  for (int i = 0; i < 4; ++i) out4[i] = 0;
}

}  // extern "C"
