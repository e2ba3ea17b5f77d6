// CPU build of the device solver logic (goss_amd/csrc/goss_implicit.h) for
// tests: the per-node ImplicitEuler / ESDIRK23a code of the kernels runs on
// the host with the host math library.  With the same math library as the
// oracle, any difference from the oracle's per-node mode is a logic
// difference of the kernel code; what remains on the GPU is the last-bit
// difference of the device exp/log/pow.  The step adjoints of
// goss_adjoint.h (StepVjp, the composition the reverse-sweep kernel runs) are
// built here too, in FP64 (tape form, as on the device) and FP32.
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "goss_adaptive.h"
#include "goss_adjoint.h"
#include "goss_implicit.h"
#include "goss_model_wrap.h"
#include "tp06_m_fp6.h"
#include "tp06_m.h"
#include "tp06_epi.h"
#include "beeler_reuter_1977.h"
#include "fhn.h"
#include "lorentz.h"
#include "fitzhughnagumo.h"
#include "markov_three_state.h"

using namespace goss_hip;

namespace {

// Runs body(first, last, slot) over node ranges on up to 16 host threads
// (nodes are independent); slot indexes per-thread accumulators.
template <class F>
void parallel_nodes(unsigned N, F&& body)
{
  const unsigned T = std::max(1u, std::min({16u, std::thread::hardware_concurrency(), N}));
  std::vector<std::thread> th;
  for (unsigned k = 0; k < T; ++k) {
    const unsigned a = unsigned(size_t(N) * k / T), b = unsigned(size_t(N) * (k + 1) / T);
    th.emplace_back([=, &body] { body(a, b, k); });
  }
  for (auto& t : th) t.join();
}

// The generated namespaces wrapped as kernel model types (as in variant_kernels.hip).
GOSS_MODEL_TYPE(M_tp06_m, goss_models::tp06_m);
GOSS_MODEL_TYPE(M_tp06_epi, goss_models::tp06_epi);
GOSS_MODEL_TYPE(M_beeler_reuter_1977, goss_models::beeler_reuter_1977);
GOSS_MODEL_TYPE(M_fhn, goss_models::fhn);
GOSS_MODEL_TYPE(M_lorentz, goss_models::lorentz);
GOSS_MODEL_TYPE(M_fitzhughnagumo, goss_models::fitzhughnagumo);
GOSS_MODEL_TYPE(M_markov_three_state, goss_models::markov_three_state);

// RKF32 per node with the device logic (goss_adaptive.h); counters are
// nfevals, ndtsa, ndtsr summed over nodes.
template <class M>
int run_rkf32(unsigned N, const double* y0, const double* t, unsigned long n, const double* prm, double atol,
              double rtol, double* out_states, long* counters, const double* fp = nullptr)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  const SharedParamsOf<M, double> sp = make_shared_params<M, double>(prm);
  const AdaptiveParams ap{atol, rtol};
  for (int k = 0; k < 4; ++k) counters[k] = 0;
  std::vector<long> acc(16 * 4, 0);
  std::vector<int> failed(16, 0);
  parallel_nodes(N, [&](unsigned first, unsigned last, unsigned slot) {
  for (unsigned node = first; node < last; ++node) {
    // field parameters of the node (node-major [N][NFP]); the model's defaults without fp
    double fpv[NFP > 0 ? NFP : 1] = {0.0};
    for (int j = 0; j < NFP; ++j) fpv[j] = fp ? fp[size_t(node) * NFP + j] : prm[M::field_parameter_indices[j]];
    const DevAcc<M, double> dacc{sp, fpv};
    double ys[S], w[3 * S];
    for (int i = 0; i < S; ++i) ys[i] = y0[size_t(node) * S + i];
    const RVec<double> yv{ys, 1}, k1{w, 1}, yn{w + S, 1}, ep{w + 2 * S, 1};
    double ldt = 0.1, facmax = 2.0;
    int nfevals = 0, ndtsa = 0, ndtsr = 0;
    for (unsigned long k = 0; k + 1 < n; ++k) {
      if (!Rkf32Node<M, double>::forward(yv, k1, yn, ep, t[k], t[k + 1] - t[k], dacc, ap, ldt, facmax, nfevals,
                                         ndtsa, ndtsr)) {
        failed[slot] = 1;
        break;
      }
    }
    for (int i = 0; i < S; ++i) out_states[size_t(node) * S + i] = ys[i];
    acc[slot * 4 + 1] += nfevals;
    acc[slot * 4 + 2] += ndtsr;
    acc[slot * 4 + 3] += ndtsa;
  }
  });
  for (unsigned k = 0; k < 16; ++k)
    for (int j = 0; j < 4; ++j) counters[j] += acc[k * 4 + j];
  return *std::max_element(failed.begin(), failed.end());
}

template <class M, int JAC = kJacobianFD>
int run(int solver, unsigned N, const double* y0, const double* t, unsigned long n, const double* prm,
        const ImplicitParams& ip, double* out_states, long* counters)
{
  constexpr int S = M::num_states;
  const SharedParamsOf<M, double> sp = make_shared_params<M, double>(prm);
  const SharedParamsOf<M, float> spf = make_shared_params<M, float>(prm);
  const double fpv[1] = {0.0};
  const float fpvf[1] = {0.0f};
  const DevAcc<M, double> acc64{sp, fpv};
  const DevAcc<M, float> accf{spf, fpvf};
  const JacAcc<M, JAC> jacc{acc64, accf};
  using Acc = std::conditional_t<JAC == kJacobianFD, DevAcc<M, double>, JacAcc<M, JAC>>;
  const Acc& acc = [&]() -> const Acc& {
    if constexpr (JAC == kJacobianFD) return acc64;
    else return jacc;
  }();
  (void)jacc;
  for (int k = 0; k < 4; ++k) counters[k] = 0;
  std::vector<long> cnt(16 * 4, 0);
  std::vector<int> failed(16, 0);
  parallel_nodes(N, [&](unsigned first, unsigned last, unsigned slot) {
  std::vector<double> jac(M::jac_nnz), jfac(M::lu_nnz), work(size_t(implicit_work_vectors(solver)) * S);
  ImplicitBuffers B{};
  B.jac = jac.data();
  B.jfac = jfac.data();
  B.work = work.data();
  for (unsigned node = first; node < last; ++node) {
    double ys[S];
    for (int i = 0; i < S; ++i) ys[i] = y0[size_t(node) * S + i];
    // Adaptive-implicit controller rows of ESDIRK4O32 (its reset values).
    double dsv[kNumDState] = {};
    dsv[kAiLdt] = -1.0;
    dsv[kAiFacmax] = 2.0;
    dsv[kErrOld] = -1.0;
    dsv[kDtOld] = 0.0;
    NodeImplicitState st{};
    st.eta = ip.eta_0;
    st.es_eta = 1.0;
    st.loc_error = 1e9;
    st.prev_dt = 0.1;
    st.recompute = 1;
    st.always = ip.always_recompute_jacobian;
    st.first_step = 1;
    st.new_jacobian = 1;
    const NodeViews v = node_views(ys, B, S, 1, 0);
    for (unsigned long k = 0; k + 1 < n; ++k) {
      if (st.failed) break;
      const double tk = t[k], dt = t[k + 1] - t[k];
      if (solver == kImplicitEuler) {
        if (!ImplicitEulerNode<M>::forward(v, tk, dt, acc, ip, st)) {
          st.failed = 1;
          failed[slot] = 1;
        }
      } else if (solver == kThetaSolver) {
        if (!ThetaNode<M>::forward(v, tk, dt, acc, ip, st)) {
          st.failed = 1;
          failed[slot] = 1;
        }
      } else if (solver == kBasicImplicitEuler) {
        BasicImplicitEulerNode<M>::forward(v, tk, dt, acc, ip, st);
      } else if (solver == kESDIRK4O32) {
        if (!Esdirk4o32Node<M>::forward(v, VecView{dsv, 1}, tk, dt, acc, ip, st)) {
          st.failed = 1;
          failed[slot] = 1;
        }
      } else {
        EsdirkNode<M>::forward(v, tk, dt, acc, ip, st);
      }
    }
    for (int i = 0; i < S; ++i) out_states[size_t(node) * S + i] = ys[i];
    cnt[slot * 4 + 0] += st.newton_its;
    cnt[slot * 4 + 1] += st.jac_comp;
    cnt[slot * 4 + 2] += st.rejects;
    cnt[slot * 4 + 3] += st.accepted;
  }
  });
  for (unsigned k = 0; k < 16; ++k)
    for (int j = 0; j < 4; ++j) counters[j] += cnt[k * 4 + j];
  return *std::max_element(failed.begin(), failed.end());
}

GOSS_MODEL_TYPE(MW_tp06_m_fp6, goss_models::tp06_m_fp6);

// One step adjoint of StepVjp<M, SOLVER, R> at (y, t, h) with seed lam; params are all P parameters.
template <class M, int SOLVER, class R>
void step_vjp(const double* y, double t, double h, const double* lam, const double* prm, double delta, double* lp,
              double* pb)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  const SharedParamsOf<M, R> sp = make_shared_params<M, R>(prm);
  R fpv[NFP > 0 ? NFP : 1];
  for (int j = 0; j < NFP; ++j) fpv[j] = R(prm[M::field_parameter_indices[j]]);
  const DevAcc<M, R> acc{sp, fpv};
  R yr[S], lr[S], lpr[S], pbr[NFP > 0 ? NFP : 1];
  for (int i = 0; i < S; ++i) {
    yr[i] = R(y[i]);
    lr[i] = R(lam[i]);
  }
  for (int j = 0; j < NFP; ++j) pbr[j] = R(0);
  const int slots = sizeof(R) == 8 ? step_vjp_tape_slots<M, SOLVER>() : 0;
  std::vector<R> mem(size_t(slots + step_vjp_work_rows<M, SOLVER>()) + 1);
  const VTape<R> tape{mem.data(), 1u, 0u};
  const VTape<R> work{mem.data() + slots, 1u, 0u};
  StepVjp<M, SOLVER, R>::run(yr, t, R(h), acc, R(delta), lr, lpr, pbr, &tape, work);
  for (int i = 0; i < S; ++i) lp[i] = double(lpr[i]);
  for (int j = 0; j < NFP; ++j) pb[j] = double(pbr[j]);
}

}  // namespace

// The reverse-sweep kernel's step adjoint on the host, for tp06_m_fp6:
// solver GRL1 / RL1 / ExplicitEuler / RK4 / RK2 / RL2 / GRL2, fp64 nonzero for double.
extern "C" int kernel_logic_step_vjp(int solver, int fp64, const double* y, double t, double h, const double* lam,
                                     const double* prm, double delta, double* lp, double* pb)
{
  using M = MW_tp06_m_fp6;
#define GOSS_HARNESS_VJP(SOLVER)                                                        \
  if (solver == SOLVER) {                                                               \
    if (fp64) step_vjp<M, SOLVER, double>(y, t, h, lam, prm, delta, lp, pb);            \
    else step_vjp<M, SOLVER, float>(y, t, h, lam, prm, delta, lp, pb);                  \
    return 0;                                                                           \
  }
  GOSS_HARNESS_VJP(kGRL1)
  GOSS_HARNESS_VJP(kRL1)
  GOSS_HARNESS_VJP(kExplicitEuler)
  GOSS_HARNESS_VJP(kRK4)
  GOSS_HARNESS_VJP(kRK2)
  GOSS_HARNESS_VJP(kRL2)
  GOSS_HARNESS_VJP(kGRL2)
#undef GOSS_HARNESS_VJP
  return -2;
}

// The implicit solvers with the Jacobian from forward differences of the FP32 right-hand side
// (host float arithmetic), for tp06_m and tp06_epi.
extern "C" int kernel_logic_run_fp32_jacobian(const char* variant, int solver, unsigned N, const double* y0,
                                              const double* t, unsigned long n, const double* params,
                                              double relative_tolerance, double* out_states, long* counters)
{
  const double min_dt = solver == kESDIRK23a ? 1e-3 : 1e-4;
  ImplicitParams ip{-1.0, 1.0, 0.1, relative_tolerance, 0.01, min_dt, 30, 0, 2, 0.5, 1.0e-5, 1.0e-8};
  const std::string v(variant);
  if (v == "tp06_m") return run<M_tp06_m, kJacobianFDFP32>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "tp06_epi") return run<M_tp06_epi, kJacobianFDFP32>(solver, N, y0, t, n, params, ip, out_states, counters);
  return -2;
}

// The implicit solvers with the model's analytic Jacobian (compute_jacobian_structured, as the kernels
// evaluate it), reference defaults of the tunables.
extern "C" int kernel_logic_run_analytic_jacobian(const char* variant, int solver, unsigned N, const double* y0,
                                                  const double* t, unsigned long n, const double* params,
                                                  double* out_states, long* counters)
{
  const double min_dt = (solver == kESDIRK23a) ? 1e-3 : 1e-4;
  ImplicitParams ip{-1.0, 1.0, 0.1, 1e-12, 0.01, min_dt, 30, 0, 2, 0.5, 1.0e-5, 1.0e-8};
  const std::string v(variant);
  constexpr int A = kJacobianAnalytic;
  if (v == "tp06_m") return run<M_tp06_m, A>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "tp06_epi") return run<M_tp06_epi, A>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "beeler_reuter_1977")
    return run<M_beeler_reuter_1977, A>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "fhn") return run<M_fhn, A>(solver, N, y0, t, n, params, ip, out_states, counters);
  return -2;
}

extern "C" int kernel_logic_run_theta(const char* variant, unsigned N, const double* y0, const double* t,
                                      unsigned long n, const double* params, double ldt_0, double theta,
                                      double* out_states, long* counters)
{
  ImplicitParams ip{ldt_0, 1.0, 0.1, 1e-12, 0.01, 1e-4, 30, 0, 2, theta, 1.0e-5, 1.0e-8};
  const std::string v(variant);
  if (v == "tp06_m") return run<M_tp06_m>(kThetaSolver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "tp06_epi") return run<M_tp06_epi>(kThetaSolver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "beeler_reuter_1977")
    return run<M_beeler_reuter_1977>(kThetaSolver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "fhn") return run<M_fhn>(kThetaSolver, N, y0, t, n, params, ip, out_states, counters);
  return -2;
}

extern "C" int kernel_logic_run(const char* variant, int solver, unsigned N, const double* y0, const double* t,
                                unsigned long n, const double* params, double ldt_0, double eta_0, double kappa,
                                double relative_tolerance, double max_rel_prev, double min_dt, int max_iterations,
                                int always, int num_refinements, double* out_states, long* counters)
{
  ImplicitParams ip{ldt_0, eta_0, kappa, relative_tolerance, max_rel_prev, min_dt, max_iterations, always,
                    num_refinements, 0.5, 1.0e-5, 1.0e-8};
  const std::string v(variant);
  if (solver == kRKF32) {
    // atol / rtol arrive in the eta_0 / kappa slots.
    if (v == "tp06_m") return run_rkf32<M_tp06_m>(N, y0, t, n, params, eta_0, kappa, out_states, counters);
    if (v == "lorentz") return run_rkf32<M_lorentz>(N, y0, t, n, params, eta_0, kappa, out_states, counters);
    if (v == "fitzhughnagumo")
      return run_rkf32<M_fitzhughnagumo>(N, y0, t, n, params, eta_0, kappa, out_states, counters);
    return -2;
  }
  if (v == "tp06_m") return run<M_tp06_m>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "tp06_epi") return run<M_tp06_epi>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "beeler_reuter_1977") return run<M_beeler_reuter_1977>(solver, N, y0, t, n, params, ip, out_states, counters);
  if (v == "fhn") return run<M_fhn>(solver, N, y0, t, n, params, ip, out_states, counters);
  return -2;
}

// RKF32 with per-node field parameters fp ([N][NFP]).
extern "C" int kernel_logic_run_rkf32_fp(const char* variant, unsigned N, const double* y0, const double* fp,
                                         const double* t, unsigned long n, const double* params, double atol,
                                         double rtol, double* out_states, long* counters)
{
  const std::string v(variant);
  if (v == "markov_three_state")
    return run_rkf32<M_markov_three_state>(N, y0, t, n, params, atol, rtol, out_states, counters, fp);
  return -2;
}
