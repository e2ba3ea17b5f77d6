// Device kernels of the explicit solver family.
//
// One node per lane.  A node's states are loaded once from the structure-of-
// arrays buffer (row i holds state i of every node, row stride `ld`, so each
// wavefront reads 64 consecutive values of one state), kept in registers for
// all fused outer steps of a launch, and stored once at the end.  Field
// parameters are per-lane registers; shared parameters live in the kernel-
// argument block and are read with scalar loads.
//
// Precision policies:
//   double                 every operation in FP64 (the parity path);
//   float, COMP = false    FP32 states and arithmetic;
//   float, COMP = true     FP32 arithmetic; the states with linear_terms == 0
//                          (TP06: Ca_i, Ca_SR, Ca_ss, Na_i, V, K_i) are carried
//                          as unevaluated sums hi + lo, so that increments
//                          below the FP32 spacing of the state accumulate.
//
// Time is always FP64: the outer (t, interval) pairs come from the host and
// the sub-step clock lt += ldt is accumulated as in GRL1.cpp:71-81.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "goss_model_support.h"

namespace goss_hip {

enum SolverKind : int {
  kExplicitEuler = 0,
  kRK2 = 1,
  kRK4 = 2,
  kRL1 = 3,
  kRL2 = 4,
  kGRL1 = 5,
  kGRL2 = 6,
  kImplicitEuler = 7,
  kESDIRK23a = 8,
  kRKF32 = 9,
  kThetaSolver = 10,
  kBasicImplicitEuler = 11,
  kESDIRK4O32 = 12,
  kNumSolvers = 13,
};

enum Precision : int { kFP64 = 0, kFP32 = 1, kFP32C = 2, kNumPrecisions = 3 };

constexpr int kBlock = 256;

// Shared (non-field) parameters and the parameter-only constants of the
// generated code (compute_constants, evaluated on the host in FP64), passed by
// value in the kernel-argument block.  The FP32 family also carries FP64
// copies for the time-only expressions, which are evaluated in FP64.
template <class R, int P, int K>
struct SharedParams {
  R v[P > 0 ? P : 1];
  R c[K > 0 ? K : 1];
};

template <int P, int K>
struct SharedParams<float, P, K> {
  float v[P > 0 ? P : 1];
  float c[K > 0 ? K : 1];
  double vd[P > 0 ? P : 1];
  double cd[K > 0 ? K : 1];
};

template <class M>
constexpr int field_slot(int k)
{
  for (int j = 0; j < M::num_field_parameters; ++j)
    if (M::field_parameter_indices[j] == k) return j;
  return -1;
}

template <class M, class R>
using SharedParamsOf = SharedParams<R, M::num_parameters, M::num_constants>;

// Parameter accessor handed to the generated model code: field parameters
// from per-lane registers, shared parameters and constants from the kernel
// arguments.  getd/cstd serve the FP64 time-only expressions.
template <class M, class R>
struct DevAcc {
  const SharedParamsOf<M, R>& sp;
  const R* fpv;
  template <int k>
  GOSS_HD GOSS_INLINE R get() const
  {
    constexpr int j = field_slot<M>(k);
    if constexpr (j >= 0)
      return fpv[j];
    else
      return sp.v[k];
  }
  template <int k>
  GOSS_HD GOSS_INLINE double getd() const
  {
    constexpr int j = field_slot<M>(k);
    if constexpr (j >= 0)
      return double(fpv[j]);
    else if constexpr (sizeof(R) == 8)
      return sp.v[k];
    else
      return sp.vd[k];
  }
  template <int k>
  GOSS_HD GOSS_INLINE R cst() const { return sp.c[k]; }
  template <int k>
  GOSS_HD GOSS_INLINE double cstd() const
  {
    if constexpr (sizeof(R) == 8)
      return sp.c[k];
    else
      return sp.cd[k];
  }
};

// Slot of state i among the compensated states, or -1.
template <class M>
constexpr int comp_slot(int i)
{
  int c = 0;
  for (int s = 0; s < i; ++s) c += M::linear_terms[s] == 0;
  return M::linear_terms[i] == 0 ? c : -1;
}

template <class M>
constexpr int num_comp()
{
  int c = 0;
  for (int s = 0; s < M::num_states; ++s) c += M::linear_terms[s] == 0;
  return c;
}

// exp(x) - 1 in FP32 without the cancellation of __expf(x) - 1 near 0:
// Taylor polynomial to x^4 for |x| < 1/64 (truncation below 2.5e-9
// relative), hardware exp2 otherwise (cancellation error below 7e-6
// relative at |x| = 1/64, shrinking as |x| grows).
GOSS_HD GOSS_INLINE float expm1_f32(const float x)
{
  const float p = x * (1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f))));
  const float q = m_exp(x) - 1.0f;  // __expf on the device
  return m_select(fabsf(x) < 0.015625f, p, q);
}

// Exact sum hi + lo += inc (Knuth's TwoSum; 6 additions).
__device__ GOSS_INLINE void comp_add(float& hi, float& lo, const float inc)
{
  const float b = inc + lo;
  const float s = hi + b;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (b - bb);
  hi = s;
  lo = err;
}

// State vector of one node: values (hi) plus low parts for compensated
// states.  For double / plain float the low parts are never touched and are
// removed by the compiler.
template <class M, class R, bool COMP>
struct NodeState {
  static constexpr int S = M::num_states;
  R y[S];
  R lo[S];

  // y[i] += inc
  __device__ GOSS_INLINE void add(const int i, const R inc)
  {
    if constexpr (COMP) {
      if (M::linear_terms[i] == 0) {
        comp_add(y[i], lo[i], inc);
        return;
      }
    }
    y[i] += inc;
  }
};

// Increment of one Rush-Larsen style update, FP64: written as in GRL1.h:80-83
// so that y0 + increment rounds exactly like the reference.
GOSS_HD GOSS_INLINE double rl_update(const double y0, const double a, const double b, const double h)
{
  return y0 + b / a * (m_exp(a * h) - 1.0);
}

// Increment of one Rush-Larsen style update, FP32: rhs * expm1(lin h) / lin, or rhs * h when the
// state is not updated exponentially; both are computed and one is selected.
GOSS_HD GOSS_INLINE float rl_increment_f32(const float lin, const float rhs, const float h, const bool use_exp)
{
  const float e = expm1_f32(lin * h);
  const float inc_exp = rhs * e * m_rcp(lin);  // v_rcp_f32 on the device
  const float inc_lin = rhs * h;
  return m_select(use_exp, inc_exp, inc_lin);
}

// The same increment for a gate state, rhs = (A - y) r with linear = -r, evaluated at y:
// (y - A) expm1(x) with x = linear h, or (y - A) x (= rhs h) when the state is not updated
// exponentially.  No reciprocal of the diagonal and no product with rhs.
GOSS_HD GOSS_INLINE float rl_gate_increment_f32(const float y_minus_target, const float lin, const float h,
                                                const bool use_exp)
{
  const float x = lin * h;
  return y_minus_target * m_select(use_exp, expm1_f32(x), x);
}

// Whether state i takes the exponential update: |a| > delta (generalized) or the model's
// linear_terms flag (classic).
template <class M, bool GEN, class R>
GOSS_HD GOSS_INLINE bool rl_use_exp(const int i, const R lin, const R delta)
{
  if constexpr (GEN) return m_fabs(lin) > delta;
  else return M::linear_terms[i] != 0;
}

// ---------------------------------------------------------------------------
// One sub-step of each explicit scheme.  `lt` is the FP64 sub-step clock.
// ---------------------------------------------------------------------------

template <class M, int SOLVER, class R, bool COMP>
struct Step;

// ExplicitEuler.cpp:76-85
template <class M, class R, bool COMP>
struct Step<M, kExplicitEuler, R, COMP> {
  template <class A>
  __device__ GOSS_INLINE static void run(NodeState<M, R, COMP>& n, const double lt, const double ldt, const A& acc,
                                         const R)
  {
    constexpr int S = M::num_states;
    R f[S];
    M::template eval<R>(n.y, lt, acc, f);
    const R h = R(ldt);
#pragma unroll
    for (int i = 0; i < S; ++i) n.add(i, h * f[i]);
  }
};

// RK2.cpp:73-91 (midpoint)
template <class M, class R, bool COMP>
struct Step<M, kRK2, R, COMP> {
  template <class A>
  __device__ GOSS_INLINE static void run(NodeState<M, R, COMP>& n, const double lt, const double ldt, const A& acc,
                                         const R)
  {
    constexpr int S = M::num_states;
    R k1[S], tmp[S];
    M::template eval<R>(n.y, lt, acc, k1);
    const R hh = R(0.5 * ldt);
#pragma unroll
    for (int i = 0; i < S; ++i) tmp[i] = n.y[i] + hh * k1[i];
    M::template eval<R>(tmp, lt + 0.5 * ldt, acc, k1);
    const R h = R(ldt);
#pragma unroll
    for (int i = 0; i < S; ++i) n.add(i, h * k1[i]);
  }
};

// RK4.cpp:79-101
template <class M, class R, bool COMP>
struct Step<M, kRK4, R, COMP> {
  template <class A>
  __device__ GOSS_INLINE static void run(NodeState<M, R, COMP>& n, const double lt, const double ldt, const A& acc,
                                         const R)
  {
    constexpr int S = M::num_states;
    R k1[S], k2[S], k3[S], k4[S], tmp[S];
    const R hh = R(0.5 * ldt);
    const R h = R(ldt);
    M::template eval<R>(n.y, lt, acc, k1);
#pragma unroll
    for (int i = 0; i < S; ++i) tmp[i] = n.y[i] + hh * k1[i];
    M::template eval<R>(tmp, lt + 0.5 * ldt, acc, k2);
#pragma unroll
    for (int i = 0; i < S; ++i) tmp[i] = n.y[i] + hh * k2[i];
    M::template eval<R>(tmp, lt + 0.5 * ldt, acc, k3);
#pragma unroll
    for (int i = 0; i < S; ++i) tmp[i] = n.y[i] + h * k3[i];
    M::template eval<R>(tmp, lt + ldt, acc, k4);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      if constexpr (sizeof(R) == 8)
        n.add(i, h * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0);  // RK4.cpp:99-100
      else
        n.add(i, h * (k1[i] + R(2.0) * k2[i] + R(2.0) * k3[i] + k4[i]) * R(1.0 / 6.0));
    }
  }
};

// Rush-Larsen family: one stage evaluated at `at` and applied to y0 (the
// y0 of RL1.h:68-77 / GRL1.h:72-84).  In FP32 the generalized stage updates a
// gate state (rhs = (A - y) r, the model's gate_states) in the closed form
// (y - A) expm1(linear h), which is rhs expm1(linear h) / linear.  APPLY: the result goes to the node
// state n (whose values are y0); otherwise to `out`.  GEN selects the
// generalized form (full diagonal, |a| > delta) or the classic one
// (linear_terms flags, only_linear evaluation).
template <class M, class R, bool COMP, bool GEN, bool APPLY, class A>
__device__ GOSS_INLINE void rl_stage(R* out, NodeState<M, R, COMP>& n, const R* at, const double tt, const R h,
                                     const A& acc, const R delta)
{
  constexpr int S = M::num_states;
  // FP32 generalized stages take the gate form for the model's gate states.
  constexpr bool GATES = GEN && sizeof(R) == 4;
  R lin[S], rhs[S], tgt[S];
  if constexpr (GATES) M::template linearized_eval_gates<R>(at, tt, acc, lin, rhs, tgt);
  else M::template linearized_eval<R>(at, tt, acc, lin, rhs, !GEN);
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const bool use_exp = rl_use_exp<M, GEN>(i, lin[i], delta);
    const R y0 = n.y[i];
    if constexpr (sizeof(R) == 8) {
      // Same expression tree as the reference update; both values are computed and
      // one is selected (no lane-divergent branch, as for the model's piecewise terms).
      const double v_exp = rl_update(y0, lin[i], rhs[i], h);
      const double v_lin = y0 + rhs[i] * h;
      const double v = m_select(use_exp, v_exp, v_lin);
      if constexpr (APPLY) n.y[i] = v; else out[i] = v;
    } else {
      float inc;
      if (GATES && M::gate_states[i]) inc = rl_gate_increment_f32(at[i] - tgt[i], lin[i], h, use_exp);
      else inc = rl_increment_f32(lin[i], rhs[i], h, use_exp);
      if constexpr (APPLY) n.add(i, inc); else out[i] = y0 + inc;
    }
  }
}

// rl_stage without a node state (APPLY = false): out = the stage evaluated at `at` applied to y0.
// The step adjoints recompute RL2 / GRL2 half steps with it.
template <class M, class R, bool GEN, class A>
GOSS_HD GOSS_INLINE void rl_point(R* out, const R* y0, const R* at, const double tt, const R h, const A& acc,
                                  const R delta)
{
  constexpr int S = M::num_states;
  constexpr bool GATES = GEN && sizeof(R) == 4;
  R lin[S], rhs[S], tgt[S];
  if constexpr (GATES) M::template linearized_eval_gates<R>(at, tt, acc, lin, rhs, tgt);
  else M::template linearized_eval<R>(at, tt, acc, lin, rhs, !GEN);
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const bool use_exp = rl_use_exp<M, GEN>(i, lin[i], delta);
    if constexpr (sizeof(R) == 8) {
      const double v_exp = rl_update(y0[i], lin[i], rhs[i], h);
      const double v_lin = y0[i] + rhs[i] * h;
      out[i] = m_select(use_exp, v_exp, v_lin);
    } else if (GATES && M::gate_states[i]) {
      out[i] = y0[i] + rl_gate_increment_f32(at[i] - tgt[i], lin[i], h, use_exp);
    } else {
      out[i] = y0[i] + rl_increment_f32(lin[i], rhs[i], h, use_exp);
    }
  }
}

template <class M, int SOLVER, class R, bool COMP>
struct RLStep {
  static constexpr bool GEN = (SOLVER == kGRL1 || SOLVER == kGRL2);
  static constexpr bool TWO = (SOLVER == kRL2 || SOLVER == kGRL2);
  template <class A>
  __device__ GOSS_INLINE static void run(NodeState<M, R, COMP>& n, const double lt, const double ldt, const A& acc,
                                         const R delta)
  {
    constexpr int S = M::num_states;
    if constexpr (!TWO) {
      // RL1.cpp:73-77 / GRL1.cpp:73-81
      rl_stage<M, R, COMP, GEN, true>(nullptr, n, n.y, lt, R(ldt), acc, delta);
    } else {
      // RL2.cpp:75-85 / GRL2.cpp:75-85: half step to y2, then a full step
      // from y evaluated at y2 and at time lt.
      R y2[S];
      rl_stage<M, R, COMP, GEN, false>(y2, n, n.y, lt, R(ldt * 0.5), acc, delta);
      rl_stage<M, R, COMP, GEN, true>(nullptr, n, y2, lt, R(ldt), acc, delta);
    }
  }
};

template <class M, class R, bool COMP>
struct Step<M, kRL1, R, COMP> : RLStep<M, kRL1, R, COMP> {};
template <class M, class R, bool COMP>
struct Step<M, kRL2, R, COMP> : RLStep<M, kRL2, R, COMP> {};
template <class M, class R, bool COMP>
struct Step<M, kGRL1, R, COMP> : RLStep<M, kGRL1, R, COMP> {};
template <class M, class R, bool COMP>
struct Step<M, kGRL2, R, COMP> : RLStep<M, kGRL2, R, COMP> {};

// Number of sub-steps, GRL1.cpp:65-67 (and every fixed-step solver).
__device__ GOSS_INLINE unsigned long substeps(const double ldt_0, const double dt)
{
  return ldt_0 > 0 ? (unsigned long)ceil(dt / ldt_0 - 1.0E-12) : 1UL;
}

// Optional per-outer-step output of field states (ODESystemSolver::solve,
// ODESystemSolver.cpp:81-99): row k of out holds the field states after
// outer step k, tangled ([N][F]) or not ([F][N]).
struct FieldOut {
  double* out;  // nullptr: no output
  unsigned long row_stride;  // N * F
  int tangled;
};

// ---------------------------------------------------------------------------
// The fused explicit step loop of one node: `nouter` outer steps
// (t_k, interval_k), states in registers between the load and the store.
// Shared by explicit_kernel (one population) and the mixed-population batch
// kernel.
// ---------------------------------------------------------------------------
template <class M, int SOLVER, class R, bool COMP>
__device__ GOSS_INLINE void explicit_node(const unsigned node, R* __restrict__ y, R* __restrict__ ylo,
                                          const R* __restrict__ fp, const double2* __restrict__ steps,
                                          const int nouter, const SharedParamsOf<M, R>& sp, const double ldt_0,
                                          const R delta, const unsigned N, const unsigned ld, const FieldOut& fo)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;

  NodeState<M, R, COMP> st;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    st.y[i] = y[size_t(i) * ld + node];
    if constexpr (COMP) {
      const int c = comp_slot<M>(i);
      st.lo[i] = c >= 0 ? ylo[size_t(c) * ld + node] : R(0);
    }
  }
  R fpv[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) fpv[j] = fp[size_t(j) * ld + node];
  const DevAcc<M, R> acc{sp, fpv};

#pragma unroll 1
  for (int k = 0; k < nouter; ++k) {
    const double t = steps[k].x;
    const double dt = steps[k].y;
    const unsigned long nsteps = substeps(ldt_0, dt);
    const double ldt = dt / nsteps;
    double lt = t;
#pragma unroll 1
    for (unsigned long s = 0; s < nsteps; ++s) {
      Step<M, SOLVER, R, COMP>::run(st, lt, ldt, acc, delta);
      lt += ldt;
    }
    if (fo.out) {
      constexpr int F = M::num_field_states;
#pragma unroll
      for (int i = 0; i < F; ++i) {
        const int si = M::field_state_indices[i];
        double v = double(st.y[si]);
        if constexpr (COMP) v += double(st.lo[si]);
        const size_t idx = fo.tangled ? size_t(node) * F + i : size_t(N) * i + node;
        fo.out[size_t(k) * fo.row_stride + idx] = v;
      }
    }
  }

#pragma unroll
  for (int i = 0; i < S; ++i) {
    y[size_t(i) * ld + node] = st.y[i];
    if constexpr (COMP) {
      const int c = comp_slot<M>(i);
      if (c >= 0) ylo[size_t(c) * ld + node] = st.lo[i];
    }
  }
}

// Waves per SIMD the FP64 explicit kernels are compiled for: the register budget per lane is
// 512 / waves (VGPRs and AGPRs).  With 2 the TP06 GRL1 kernel spills ~250 B per lane to scratch
// but gains 22 % over 1 (5.84e9 -> 7.13e9 node-steps/s at 1M nodes); 3 and 4 lose (4.1e9, 3.3e9).
// The FP32 kernels fit two waves without a bound.
#ifndef GOSS_EXPLICIT_WAVES_FP64
#define GOSS_EXPLICIT_WAVES_FP64 2
#endif
// FP32 / FP32C: no bound (the TP06 GRL1 kernel fits two waves); 3 or 4 waves measured within 1 %.
#ifndef GOSS_EXPLICIT_WAVES_FP32
#define GOSS_EXPLICIT_WAVES_FP32 1
#endif

template <class M, int SOLVER, class R, bool COMP>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 ? GOSS_EXPLICIT_WAVES_FP64 : GOSS_EXPLICIT_WAVES_FP32)))
explicit_kernel(R* __restrict__ y, R* __restrict__ ylo, const R* __restrict__ fp, const double2* __restrict__ steps,
                const int nouter, const SharedParamsOf<M, R> sp, const double ldt_0, const R delta,
                const unsigned N, const unsigned ld, const FieldOut fo)
{
  const unsigned node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= N) return;
  explicit_node<M, SOLVER, R, COMP>(node, y, ylo, fp, steps, nouter, sp, ldt_0, delta, N, ld, fo);
}

}  // namespace goss_hip
