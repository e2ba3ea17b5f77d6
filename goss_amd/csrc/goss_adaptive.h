// Device kernel of the adaptive explicit solver RKF32 (RKF32.cpp:105-222) with
// the step-size controller of AdaptiveExplicitSolver::new_time_step
// (AdaptiveExplicitSolver.cpp:168-244).
//
// One node per lane.  Each node owns the controller state the reference
// keeps in its solver instance: the initial step guess _ldt and facmax,
// which carry over from one forward call to the next, and the counters
// nfevals / ndtsa / ndtsr.  Step sizes and the local time are FP64 in every
// precision; the states, stage values and the error vector use the
// precision R.
//
// The states and the vectors that live across RHS evaluations (k1, the
// partial sums of yn and of the error vector e) are HBM rows of stride ld,
// as in the implicit kernels; registers hold only one evaluation's input and
// output, which keeps a 19-state FP64 model free of register spills.  The
// partial sums keep the reference's left-to-right association:
// yn = y + dt*((b1 k1 + b2 k2) + b3 k3), e = dt*(((d1 k1 + d2 k2) + d3 k3) + d4 k4).
#pragma once

#include "goss_implicit.h"

namespace goss_hip {

// Sub-steps one forward call may take before the node is flagged as stuck.
// The reference loops until |_t - t_end| < 1e-14 (AdaptiveExplicitSolver.cpp:206)
// and does not return when that cannot be met (a step size underflowed by
// repeated rejections of a non-finite error).
constexpr long kRkf32MaxSteps = 1000000;

struct AdaptiveParams {
  double atol, rtol;  // AdaptiveExplicitSolver::set_tol (AdaptiveExplicitSolver.h:71-75)
};

// View of one node's vector of precision R (element i at a[i * ld]).
template <class R>
struct RVec {
  static constexpr bool in_registers = false;
  R* a;
  unsigned ld;
  GOSS_HD GOSS_INLINE R& operator[](int i) const { return a[size_t(i) * ld]; }
};

// A node's vector held in registers, for models small enough that the RKF32
// vectors (y, k1, yn, ep: 4 S values) take a few registers; every index is a
// compile-time constant after unrolling.
template <class R, int S>
struct RegVec {
  static constexpr bool in_registers = true;
  mutable R v[S];
  GOSS_HD GOSS_INLINE R& operator[](int i) const { return v[i]; }
};

// Models with at most this many states keep the RKF32 vectors in registers.
constexpr int kRkf32RegisterStates = 4;

// Work vectors of RKF32 per node (each S rows of the work buffer).
enum Rkf32Work : int { kRkK1 = 0, kRkYn, kRkEp, kNumRkf32Work };

template <class M, class R>
struct Rkf32Node {
  static constexpr int S = M::num_states;

  // AdaptiveExplicitSolver::dtinit (AdaptiveExplicitSolver.cpp:67-166) with _itol = 0, iord = 3: an
  // initial step from the scaled norms of y0 and f0 = k1 and a difference estimate of the second
  // derivative after one explicit Euler step.  Sums in the precision R, step sizes in FP64.
  template <class V, class A>
  GOSS_HD static double dtinit(const V& y0, const V& f0, const double t, const A& acc, const AdaptiveParams& ap)
  {
    constexpr double iord = 3.0;
    const R atol = R(ap.atol), rtol = R(ap.rtol);
    R dnf = 0, dny = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const R sk = atol + rtol * fabs(y0[i]);
      R tmp = f0[i] / sk;
      dnf += tmp * tmp;
      tmp = y0[i] / sk;
      dny += tmp * tmp;
    }
    double dt;
    if (double(dnf) <= 1.0e-10 || double(dny) <= 1.0e-10)
      dt = 1.0e-6;
    else
      dt = 0.01 * m_sqrt(double(dny / dnf));
    R y1[S], f1[S];
#pragma unroll
    for (int i = 0; i < S; ++i) y1[i] = y0[i] + R(dt) * f0[i];
    M::template eval<R>(y1, t + dt, acc, f1);
    R der2 = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const R sk = atol + rtol * fabs(y1[i]);
      const R tmp = (f1[i] - f0[i]) / sk;
      der2 += tmp * tmp;
    }
    const double der2d = m_sqrt(double(der2)) / dt;
    const double sdnf = m_sqrt(double(dnf));
    const double der12 = fabs(der2d) >= sdnf ? fabs(der2d) : sdnf;
    double dt1;
    if (der12 <= 1.0e-15)
      dt1 = fabs(dt) * 1.0e-3 < 1.0e-6 ? 1.0e-6 : fabs(dt) * 1.0e-3;
    else
      dt1 = m_pow(0.01 / der12, 1.0 / iord);
    return 100 * fabs(dt) < dt1 ? 100 * fabs(dt) : dt1;
  }

  // RKF32::forward for one node.  y, k1, yn and ep are HBM views (RVec: only
  // the input and output of one RHS evaluation are held in registers) or, for
  // small models, register vectors (RegVec).  Returns false when the sub-step
  // cap is hit.
  template <class V, class A>
  GOSS_HD static bool forward(const V& y, const V& k1, const V& yn, const V& ep,
                              const double t, const double interval, const A& acc, const AdaptiveParams& ap,
                              double& ldt_state, double& facmax, int& nfevals, int& ndtsa, int& ndtsr)
  {
    // Coefficients (RKF32.cpp:34-37).
    constexpr double a21 = 1.0 / 2.0, a32 = 3.0 / 4.0, b1 = 2.0 / 9.0, b2 = 1.0 / 3.0, b3 = 4.0 / 9.0;
    constexpr double bh1 = 7.0 / 24.0, bh2 = 1.0 / 4.0, bh3 = 1.0 / 3.0, bh4 = 1.0 / 8.0;
    constexpr double d1 = b1 - bh1, d2 = b2 - bh2, d3 = b3 - bh3, d4 = -bh4, c2 = 1.0 / 2.0, c3 = 3.0 / 4.0;
    // AdaptiveExplicitSolver::reset (AdaptiveExplicitSolver.cpp:58-64), _iord = 3 (RKF32.cpp:39).
    constexpr double facmin = 0.5, facmaxb = 2.0, stabfac = 0.9, iord = 3.0;
    const R atol = R(ap.atol), rtol = R(ap.rtol);
    const double t_end = t + interval;
    double lt = t;  // _t
    bool reached_tend = false;
    // `first` is set by reset() and never cleared (RKF32.cpp:126-131).
    {
      R in[S], out[S];
#pragma unroll
      for (int i = 0; i < S; ++i) in[i] = y[i];
      M::template eval<R>(in, t, acc, out);
#pragma unroll
      for (int i = 0; i < S; ++i) k1[i] = out[i];
    }
    nfevals += 1;
    // AdaptiveExplicitSolver::_ldt turns negative when a step overshoots the end of an interval and the
    // backward step t_end - _t is stored as the guess (AdaptiveExplicitSolver.cpp:234-241); the next
    // call then starts from dtinit (RKF32.cpp:134-141).
    double dt;
    if (ldt_state < 0) {
      dt = dtinit(y, k1, t, acc, ap);
      nfevals += 1;
    } else {
      dt = ldt_state;
    }
    long guard = 0;
#pragma unroll 1
    while (!reached_tend) {
      if (++guard > kRkf32MaxSteps) return false;
      const R rdt = R(dt);
      // Stage 2; yn and ep collect b1 k1 + b2 k2 and d1 k1 + d2 k2 (the
      // reference's left-to-right sums, continued after stage 3).
      {
        R ki[S], k2[S];
#pragma unroll
        for (int i = 0; i < S; ++i) ki[i] = y[i] + rdt * R(a21) * k1[i];
        M::template eval<R>(ki, t + c2 * dt, acc, k2);  // stage times from the call's t (RKF32.cpp:155, 160)
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const R k1i = k1[i];
          ki[i] = y[i] + rdt * R(a32) * k2[i];
          yn[i] = R(b1) * k1i + R(b2) * k2[i];
          ep[i] = R(d1) * k1i + R(d2) * k2[i];
        }
        // Stage 3.
        M::template eval<R>(ki, t + c3 * dt, acc, k2);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          yn[i] = y[i] + rdt * (yn[i] + R(b3) * k2[i]);
          ep[i] = ep[i] + R(d3) * k2[i];
        }
      }
      // FSAL evaluation at yn, then new_time_step: RMS of
      // e / (atol + rtol * max(|y|, |yn|)).
      R err = 0;
      {
        R in[S], k4[S];
#pragma unroll
        for (int i = 0; i < S; ++i) in[i] = yn[i];
        M::template eval<R>(in, t + dt, acc, k4);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const R e = rdt * (ep[i] + R(d4) * k4[i]);
          const R yi_abs = fabs(y[i]), yni_abs = fabs(in[i]);
          const R mx = yi_abs > yni_abs ? yi_abs : yni_abs;
          const R sk = atol + rtol * mx;
          const R tmp = e / sk;
          err += tmp * tmp;
          ep[i] = k4[i];  // kept for the FSAL swap below
        }
      }
      nfevals += 3;
      double errd = m_sqrt(double(err) / S);
      if (__builtin_isnan(errd) || __builtin_isinf(errd)) errd = 2000.0;
      bool accepted;
      if (errd <= 1.0) {
        lt += dt;
        accepted = true;
        if (fabs(lt - t_end) < 1e-14) reached_tend = true;
      } else {
        accepted = false;
      }
      double fac = stabfac * m_pow(1.0 / errd, 1.0 / (iord + 1));
      if (facmin > fac) fac = facmin;
      if (fac > facmax) fac = facmax;
      facmax = accepted ? facmaxb : 1.0;
      dt *= fac;
      if (lt + dt >= t_end)
        dt = t_end - lt;
      else if (lt + 1.5 * dt >= t_end)
        dt = (t_end - lt) / 2.0;
      else
        ldt_state = dt;
      if (accepted) {
        ndtsa += 1;
        if constexpr (V::in_registers) {
#pragma unroll
          for (int i = 0; i < S; ++i) {
            y[i] = yn[i];
            k1[i] = ep[i];
          }
        } else {
#pragma unroll 1
          for (int i = 0; i < S; ++i) {
            y[i] = yn[i];
            k1[i] = ep[i];
          }
        }
      } else {
        ndtsr += 1;
      }
    }
    return true;
  }
};

// The fused RKF32 step loop of one node: `nouter` outer steps (t_k,
// interval_k).  For models of more than kRkf32RegisterStates states the states
// stay in their SoA buffer and k1, yn and ep are rows of the work buffer;
// smaller models hold all four vectors in registers.  The controller state is
// read from and written back to the per-node rows of B.  Shared by rkf32_kernel (one population) and the
// mixed-population batch kernel.
template <class M, class R>
__device__ GOSS_INLINE void rkf32_node(const unsigned node, R* __restrict__ y, const R* __restrict__ fp,
                                       const double2* __restrict__ steps, const int nouter,
                                       const SharedParamsOf<M, R>& sp, const AdaptiveParams& ap,
                                       const ImplicitBuffers& B, const unsigned N, const unsigned ld,
                                       const FieldOut& fo)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  R fpv[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) fpv[j] = fp[size_t(j) * ld + node];
  const DevAcc<M, R> acc{sp, fpv};
  double ldt_state = B.ds[size_t(kRkLdt) * ld + node];
  double facmax = B.ds[size_t(kFacmax) * ld + node];
  int nfevals = B.is[size_t(kNfevals) * ld + node];
  int ndtsa = B.is[size_t(kAccepted) * ld + node];
  int ndtsr = B.is[size_t(kRejects) * ld + node];
  int failed = B.is[size_t(kFailed) * ld + node];

  auto run = [&](const auto& yv, const auto& k1, const auto& yn, const auto& ep) {
#pragma unroll 1
    for (int k = 0; k < nouter; ++k) {
      if (!failed) {
        if (!Rkf32Node<M, R>::forward(yv, k1, yn, ep, steps[k].x, steps[k].y, acc, ap, ldt_state, facmax, nfevals,
                                      ndtsa, ndtsr)) {
          failed = 1;
          atomicOr(B.error, 1);
        }
      }
      if (fo.out) {
        constexpr int F = M::num_field_states;
#pragma unroll
        for (int i = 0; i < F; ++i) {
          const size_t idx = fo.tangled ? size_t(node) * F + i : size_t(N) * i + node;
          fo.out[size_t(k) * fo.row_stride + idx] = double(yv[M::field_state_indices[i]]);
        }
      }
    }
  };
  if constexpr (S <= kRkf32RegisterStates) {
    RegVec<R, S> yv, k1, yn, ep;
#pragma unroll
    for (int i = 0; i < S; ++i) yv[i] = y[size_t(i) * ld + node];
    run(yv, k1, yn, ep);
#pragma unroll
    for (int i = 0; i < S; ++i) y[size_t(i) * ld + node] = yv[i];
  } else {
    R* const work = reinterpret_cast<R*>(B.work);
    const RVec<R> yv{y + node, ld};
    const RVec<R> k1{work + size_t(kRkK1) * S * ld + node, ld};
    const RVec<R> yn{work + size_t(kRkYn) * S * ld + node, ld};
    const RVec<R> ep{work + size_t(kRkEp) * S * ld + node, ld};
    run(yv, k1, yn, ep);
  }
  B.ds[size_t(kRkLdt) * ld + node] = ldt_state;
  B.ds[size_t(kFacmax) * ld + node] = facmax;
  B.is[size_t(kNfevals) * ld + node] = nfevals;
  B.is[size_t(kAccepted) * ld + node] = ndtsa;
  B.is[size_t(kRejects) * ld + node] = ndtsr;
  B.is[size_t(kFailed) * ld + node] = failed;
}

// Waves per SIMD of the RKF32 kernel.  More waves spill and were slower (TP06 M, 262,144 nodes,
// 20 intervals of 0.5 ms: FP64 1004 / 1069 / 1515 ms, FP32 364 / 436 / 576 ms with 1 / 2 / 4 waves).
#ifndef GOSS_RKF32_WAVES
#define GOSS_RKF32_WAVES 1
#endif

template <class M, class R>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GOSS_RKF32_WAVES)))
rkf32_kernel(R* __restrict__ y, const R* __restrict__ fp, const double2* __restrict__ steps, const int nouter,
             const SharedParamsOf<M, R> sp, const AdaptiveParams ap, const ImplicitBuffers B, const unsigned N,
             const unsigned ld, const FieldOut fo)
{
  const unsigned node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= N) return;
  rkf32_node<M, R>(node, y, fp, steps, nouter, sp, ap, B, N, ld, fo);
}

}  // namespace goss_hip
