// A generated model namespace (generated/<variant>.h) wrapped as a type for
// the kernel templates, and the host-side evaluation of its shared parameters
// and parameter-only constants.
#pragma once

#include "goss_kernels.h"

// GOSS_MODEL_TYPE(TYPE, NSN) defines struct TYPE over namespace NSN.
#define GOSS_MODEL_TYPE(TYPE, NSN)                                                                                   \
  struct TYPE {                                                                                                      \
    static constexpr int num_states = NSN::num_states;                                                               \
    static constexpr int num_parameters = NSN::num_parameters;                                                       \
    static constexpr int num_constants = NSN::num_constants;                                                         \
    static constexpr int num_field_states = NSN::num_field_states;                                                   \
    static constexpr int num_field_parameters = NSN::num_field_parameters;                                           \
    static constexpr const int* field_state_indices = NSN::field_state_indices;                                      \
    static constexpr const int* field_parameter_indices = NSN::field_parameter_indices;                              \
    static constexpr const unsigned char* linear_terms = NSN::linear_terms;                                          \
    static constexpr const unsigned char* differential_states = NSN::differential_states;                            \
    static constexpr bool is_dae = NSN::is_dae;                                                                      \
    static constexpr const unsigned char* gate_states = NSN::gate_states;                                            \
    template <typename A>                                                                                            \
    static void compute_constants(const A& a, double* c)                                                             \
    {                                                                                                                \
      NSN::compute_constants<double>(a, c);                                                                          \
    }                                                                                                                \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void eval(const R* y, const T t, const A& a, R* out)                                  \
    {                                                                                                                \
      NSN::eval<R>(y, t, a, out);                                                                                    \
    }                                                                                                                \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void linearized_eval(const R* y, const T t, const A& a, R* lin, R* rhs, bool ol)      \
    {                                                                                                                \
      NSN::linearized_eval<R>(y, t, a, lin, rhs, ol);                                                                \
    }                                                                                                                \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void linearized_eval_gates(const R* y, const T t, const A& a, R* lin, R* rhs,         \
                                                          R* gate_target)                                            \
    {                                                                                                                \
      NSN::linearized_eval_gates<R>(y, t, a, lin, rhs, gate_target);                                                 \
    }                                                                                                                \
    template <typename R, typename A, typename T, typename J>                                                        \
    GOSS_HD GOSS_INLINE static void compute_jacobian(const R* y, const T t, const A& a, J jac)                       \
    {                                                                                                                \
      NSN::compute_jacobian<R>(y, t, a, jac);                                                                        \
    }                                                                                                                \
    template <typename R, typename A, typename T, typename J>                                                        \
    GOSS_HD GOSS_INLINE static void compute_jacobian_structured(const R* y, const T t, const A& a, const J& jac)     \
    {                                                                                                                \
      NSN::compute_jacobian_structured<R>(y, t, a, jac);                                                             \
    }                                                                                                                \
    static constexpr int num_intermediates = NSN::num_intermediates;                                                 \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void eval_intermediates(const R* y, const T t, const A& a, R* out)                    \
    {                                                                                                                \
      NSN::eval_intermediates<R>(y, t, a, out);                                                                      \
    }                                                                                                                \
    static constexpr int jac_nnz = NSN::jac_nnz;                                                                     \
    static constexpr int lu_nnz = NSN::lu_nnz;                                                                       \
    static constexpr const unsigned* jac_cols = NSN::jac_cols;                                                       \
    static constexpr const int* jac_col_start = NSN::jac_col_start;                                                  \
    GOSS_HD GOSS_INLINE static void lu_from_jacobian(const double* J, const double scale, double* a)                 \
    {                                                                                                                \
      NSN::lu_from_jacobian<double>(J, scale, a);                                                                    \
    }                                                                                                                \
    GOSS_HD GOSS_INLINE static void lu_factorize_sparse(double* a) { NSN::lu_factorize_sparse<double>(a); }         \
    GOSS_HD GOSS_INLINE static void lu_solve_sparse(const double* a, double* x)                                     \
    {                                                                                                                \
      NSN::lu_solve_sparse<double>(a, x);                                                                            \
    }                                                                                                                \
    static constexpr int grl1_vjp_tape_slots = NSN::grl1_vjp_tape_slots;                                             \
    static constexpr int rl1_vjp_tape_slots = NSN::rl1_vjp_tape_slots;                                               \
    static constexpr int eval_vjp_tape_slots = NSN::eval_vjp_tape_slots;                                             \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void rl1_vjp(const R* y, const T t, const A& a, const R h, const R delta,             \
                                            const R* lam, R* lam_prev, R* fp_bar)                                    \
    {                                                                                                                \
      NSN::rl1_vjp<R>(y, t, a, h, delta, lam, lam_prev, fp_bar);                                                     \
    }                                                                                                                \
    template <typename R, typename A, typename T, typename TP, typename LV, typename OV>                             \
    GOSS_HD GOSS_INLINE static void rl1_vjp_tape(const R* y, const T t, const A& a, const R h, const R delta,        \
                                                 const LV& lam, const OV& lam_prev, R* fp_bar, const TP& tape)       \
    {                                                                                                                \
      NSN::rl1_vjp_tape<R>(y, t, a, h, delta, lam, lam_prev, fp_bar, tape);                                          \
    }                                                                                                                \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void eval_vjp(const R* y, const T t, const A& a, const R* seed, R* y_bar, R* fp_bar)  \
    {                                                                                                                \
      NSN::eval_vjp<R>(y, t, a, seed, y_bar, fp_bar);                                                                \
    }                                                                                                                \
    template <typename R, typename A, typename T, typename TP, typename SV, typename YV>                             \
    GOSS_HD GOSS_INLINE static void eval_vjp_tape(const R* y, const T t, const A& a, const SV& seed, const YV& y_bar, \
                                                  R* fp_bar, const TP& tape)                                         \
    {                                                                                                                \
      NSN::eval_vjp_tape<R>(y, t, a, seed, y_bar, fp_bar, tape);                                                     \
    }                                                                                                                \
    template <typename R, typename A, typename T>                                                                    \
    GOSS_HD GOSS_INLINE static void grl1_vjp(const R* y, const T t, const A& a, const R h, const R delta,            \
                                             const R* lam, R* lam_prev, R* fp_bar)                                   \
    {                                                                                                                \
      NSN::grl1_vjp<R>(y, t, a, h, delta, lam, lam_prev, fp_bar);                                                    \
    }                                                                                                                \
    template <typename R, typename A, typename T, typename TP, typename LV, typename OV>                             \
    GOSS_HD GOSS_INLINE static void grl1_vjp_tape(const R* y, const T t, const A& a, const R h, const R delta,       \
                                                  const LV& lam, const OV& lam_prev, R* fp_bar, const TP& tape)      \
    {                                                                                                                \
      NSN::grl1_vjp_tape<R>(y, t, a, h, delta, lam, lam_prev, fp_bar, tape);                                         \
    }                                                                                                                \
  }

namespace goss_hip {

// Host-side accessor for compute_constants (all parameters in FP64).
struct HostAcc {
  const double* v;
  template <int k> double getd() const { return v[k]; }
};

// Shared parameters in the kernel's precision, with the parameter-only
// constants evaluated once on the host in FP64 (FP32 kernels also get the
// FP64 values for their time-only expressions).
template <class M, class R>
SharedParamsOf<M, R> make_shared_params(const double* shared)
{
  SharedParamsOf<M, R> sp;
  double c[M::num_constants > 0 ? M::num_constants : 1];
  M::compute_constants(HostAcc{shared}, c);
  for (int k = 0; k < M::num_parameters; ++k) sp.v[k] = R(shared[k]);
  for (int k = 0; k < M::num_constants; ++k) sp.c[k] = R(c[k]);
  if constexpr (sizeof(R) == 4) {
    for (int k = 0; k < M::num_parameters; ++k) sp.vd[k] = shared[k];
    for (int k = 0; k < M::num_constants; ++k) sp.cd[k] = c[k];
  }
  return sp;
}

}  // namespace goss_hip
