// Device kernels of the implicit solvers: ImplicitEuler (ImplicitEuler.cpp:57-152
// over ImplicitODESolver::newton_solve, ImplicitODESolver.cpp:114-245),
// ThetaSolver (ThetaSolver.cpp:48-155), BasicImplicitEuler
// (BasicImplicitEuler.cpp:62-181), ESDIRK23a (ESDIRK23a.cpp:101-316) and
// ESDIRK4O32 (ESDIRK4O32.cpp:92-298 over AdaptiveImplicitSolver.cpp:68-200).
//
// One node per lane.  Each node owns its solver state (Newton eta, Jacobian
// reuse flags, ESDIRK step-size history, held in registers during a launch),
// its finite-difference Jacobian (ODE::compute_jacobian, ODE.cpp:47-67) and
// the LU factors of I - h a J (ODE::lu_factorize, ODE.cpp:69-100), and a few
// work vectors (stage values, Newton increments).  The Jacobian and the
// factors are stored over their structure only (the generated model's
// jac_nnz / lu_nnz entries: 76 and 82 of 361 for TP06): the entries a
// state's right-hand side does not read are exact zeros of the difference
// quotient, and the LU fill is fixed by that structure, so the dense
// routines' results are reproduced with the zero terms left out.  The
// factorisation and the substitutions are generated straight-line code over
// a register copy of the factors: all loads of an operand are issued at
// once and no load waits on the previous row.
// States, factors and work vectors live in HBM as structure of arrays:
// element i of every node in one row of stride ld, so a wavefront's 64 lanes
// touch 512 contiguous bytes per element.  Only the input and output of one
// right-hand-side evaluation are held in registers across the solver logic
// (the FP64 RHS of a cell model needs most of the register file by itself).
// The RHS is instantiated at one site per algorithmic role (Newton residual,
// first ESDIRK stage, Jacobian columns).
//
// The arithmetic is FP64 throughout and follows the reference's operation
// order (compiled with -ffp-contract=off).  Optionally (JacAcc below) the
// forward-difference Jacobian differences the model's FP32 right-hand side
// instead; the Newton iteration, its residual, the LU and the substitutions
// stay FP64 in every mode.
#pragma once

#include <type_traits>

#include "goss_kernels.h"

namespace goss_hip {

// How a node's Jacobian is formed.
enum JacobianMode : int {
  kJacobianFD = 0,      // ODE::compute_jacobian (ODE.cpp:47-67): FP64 forward differences
  kJacobianFDFP32 = 1,  // forward differences of the FP32 right-hand side, stored as FP64
  kJacobianAnalytic = 2,  // the generated analytic Jacobian (a plugin generated with generate_jacobian)
  kNumJacobianModes = 3
};

// Accessor of the implicit kernels in the kJacobianFDFP32 and kJacobianAnalytic
// modes: forwards the parameter reads of the FP64 model code to `acc`, carries
// the FP32 accessor for the FP32 Jacobian's evaluations and names the mode.
template <class M, int JAC>
struct JacAcc {
  static constexpr int jacobian_mode = JAC;
  const DevAcc<M, double>& acc;
  const DevAcc<M, float>& accf;
  template <int k>
  GOSS_HD GOSS_INLINE double get() const { return acc.template get<k>(); }
  template <int k>
  GOSS_HD GOSS_INLINE double getd() const { return acc.template getd<k>(); }
  template <int k>
  GOSS_HD GOSS_INLINE double cst() const { return acc.template cst<k>(); }
  template <int k>
  GOSS_HD GOSS_INLINE double cstd() const { return acc.template cstd<k>(); }
};

template <class A, class = void>
struct JacobianModeOf {
  static constexpr int value = kJacobianFD;
};
template <class A>
struct JacobianModeOf<A, std::void_t<decltype(A::jacobian_mode)>> {
  static constexpr int value = A::jacobian_mode;
};

constexpr double kGossEps = 3.0e-16;  // constants.h:22

// std::max / std::min semantics (a NaN first argument is returned), which the
// reference uses; fmax/fmin would drop a NaN instead.
GOSS_HD GOSS_INLINE double smax(const double a, const double b) { return (a < b) ? b : a; }
GOSS_HD GOSS_INLINE double smin(const double a, const double b) { return (b < a) ? b : a; }

// Public tunables (ImplicitODESolver.h:36-41, ImplicitEuler.h:36-37, ESDIRK23a.h:37-38).
struct ImplicitParams {
  double ldt_0;
  double eta_0;
  double kappa;
  double relative_tolerance;
  double max_relative_previous_residual;
  double min_dt;
  int max_iterations;
  int always_recompute_jacobian;
  int num_refinements;
  double theta;  // ThetaSolver.h:37
  double atol, rtol;  // AdaptiveImplicitSolver::set_tol (AdaptiveImplicitSolver.h), ESDIRK4O32's error test
};

// Per-node solver state, structure of arrays [field][ld] (shared by the
// implicit solvers and RKF32).
enum DState : int {
  kEta = 0,   // ImplicitODESolver::_eta
  kEsEta,     // ESDIRK23a::eta
  kLocError,  // ESDIRK23a::loc_error
  kPrevDt,    // ESDIRK23a::prev_dt
  kRkLdt,     // AdaptiveExplicitSolver::_ldt (RKF32 initial step guess)
  kFacmax,    // AdaptiveExplicitSolver::facmax
  kAiLdt,     // AdaptiveImplicitSolver::_ldt (ESDIRK4O32's step guess; -1 after reset)
  kAiFacmax,  // AdaptiveImplicitSolver::facmax
  kErrOld,    // AdaptiveImplicitSolver::err_old
  kDtOld,     // AdaptiveImplicitSolver::dt_old
  kNumDState
};
enum IState : int {
  kRecompute = 0,   // ImplicitODESolver::_recompute_jacobian
  kNewtonIts,       // _newton_iterations
  kAlways,          // always_recompute_jacobian (sticky, ImplicitEuler.cpp:81-82)
  kJustRefined,     // ImplicitEuler::_justrefined
  kFirstStep,       // ESDIRK23a::first_step
  kNewJacobian,     // ESDIRK23a::new_jacobian
  kJacComp,         // counters
  kRejects,
  kAccepted,
  kFailed,          // ImplicitEuler gave up below min_dt; RKF32 did not reach the interval end
  kNfevals,         // RKF32::nfevals
  kNumIState
};

// Work vectors per node (each S rows of the work buffer).
enum EsdirkWork : int { kEsZ1 = 0, kEsZ2, kEsZ3, kEsZ4, kEsPrev, kEsF1, kEsDZ, kNumEsdirkWork };
enum EulerWork : int { kIeZ = 0, kIeYZ, kIeB, kIeF1, kIePrev, kNumEulerWork };  // kIePrev: ThetaSolver only
// ESDIRK4O32: the Newton rows of EulerWork (kIeZ holds the stage being solved, z4 itself), then its own.
enum Esdirk4Work : int { kE4Z1 = kNumEulerWork, kE4Z2, kE4Z3, kE4Yn, kE4Yh, kNumEsdirk4Work };

__host__ __device__ constexpr int implicit_work_vectors(int solver)
{
  return solver == kESDIRK23a ? int(kNumEsdirkWork) : (solver == kESDIRK4O32 ? int(kNumEsdirk4Work) : int(kNumEulerWork));
}

constexpr int kNumImplicitSlots = 5;

// Slot of an implicit solver in VariantInfo::implicit_launch.
__host__ __device__ constexpr int implicit_slot(int solver)
{
  return solver == kImplicitEuler ? 0
         : solver == kESDIRK23a   ? 1
         : solver == kThetaSolver ? 2
         : solver == kBasicImplicitEuler ? 3
                                         : 4;
}

struct ImplicitBuffers {
  double* jac;   // [jac_nnz][ld]: the Jacobian over its structure
  double* jfac;  // [lu_nnz][ld]: the LU factors of I - h a J over their structure
  double* work;  // [implicit_work_vectors(solver) * S][ld]
  double* ds;    // [kNumDState][ld]
  int* is;       // [kNumIState][ld]
  int* error;    // one flag per system (set when a node fails)
};

// View of one node's vector (element i at a[i * ld]).
struct VecView {
  double* a;
  unsigned ld;
  GOSS_HD GOSS_INLINE double& operator[](int i) const { return a[size_t(i) * ld]; }
};

// Everything one node works on: its states, work vectors, Jacobian J and LU
// factors F (each a vector over its structure).
struct NodeViews {
  VecView y;
  double* work;
  unsigned ld;
  int S;
  VecView J, F;
  GOSS_HD GOSS_INLINE VecView w(int k) const { return VecView{work + size_t(k) * S * ld, ld}; }
};

template <class M>
struct Implicit {
  static constexpr int S = M::num_states;

  // ODE::compute_jacobian (ODE.cpp:47-67): forward differences, column by
  // column, delta_i = sqrt(1e-15 * max(1e-5, |y_i|)).  Pass i = -1 evaluates
  // f(y) into f1, pass i >= 0 the perturbed RHS of column i.
  // Stores column i's entries (rows of the structure) of (f - f1) / d.
  GOSS_HD static GOSS_INLINE void store_column(const VecView& J, const int i, const double* f, const VecView& f1,
                                               const double d)
  {
    const unsigned rows = M::jac_cols[i];
    int slot = M::jac_col_start[i];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      if (rows & (1u << j)) {
        J[slot] = (f[j] - f1[j]) / d;
        ++slot;
      }
    }
  }

  template <class A>
  GOSS_HD static void fd_jacobian(const VecView& y, const double t, const A& acc, const VecView& f1,
                                  const VecView& J)
  {
#pragma unroll 1
    for (int i = -1; i < S; ++i) {
      double delta = 0.0;
      if (i >= 0) {
        const double ysafe = y[i];
        const double mx = 1e-5 > fabs(ysafe) ? 1e-5 : fabs(ysafe);
        delta = m_sqrt(1e-15 * mx);
      }
      double yp[S], f[S];
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const double ysafe = y[j];
        yp[j] = (j == i) ? ysafe + delta : ysafe;
      }
      M::template eval<double>(yp, t, acc, f);
      if (i < 0) {
#pragma unroll
        for (int j = 0; j < S; ++j) f1[j] = f[j];
      } else {
        store_column(J, i, f, f1, delta);
      }
    }
  }

  // Forward differences of the FP32 right-hand side.  Column i perturbs state i
  // by sqrt(eps_FP32) * max(1e-5, |y_i|) (the reference's floor, with the FP32
  // epsilon in place of its 1e-15) and divides by the perturbation as it is
  // represented in FP32.  The base evaluation, also FP32, is kept in f1.
  // Returns whether every FP32 evaluation was finite.
  template <class A>
  GOSS_HD static bool fd_jacobian_f32(const VecView& y, const double t, const A& acc, const VecView& f1,
                                      const VecView& J)
  {
    constexpr float kSqrtEps = 3.4526698e-4f;  // sqrt(2^-23)
    constexpr float kMaxFinite = 3.4028235e38f;
    bool finite = true;
#pragma unroll 1
    for (int i = -1; i < S; ++i) {
      float yp[S], f[S];
      float dy = 0.0f;
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const float yj = float(y[j]);
        if (j == i) {
          const float mx = 1e-5f > fabsf(yj) ? 1e-5f : fabsf(yj);
          const float yn = yj + kSqrtEps * mx;
          dy = yn - yj;
          yp[j] = yn;
        } else {
          yp[j] = yj;
        }
      }
      M::template eval<float>(yp, t, acc.accf, f);
#pragma unroll
      for (int j = 0; j < S; ++j) finite = finite && (fabsf(f[j]) <= kMaxFinite);  // false for inf and NaN
      if (i < 0) {
#pragma unroll
        for (int j = 0; j < S; ++j) f1[j] = double(f[j]);
      } else {
        double fd[S];
#pragma unroll
        for (int j = 0; j < S; ++j) fd[j] = double(f[j]);
        store_column(J, i, fd, f1, double(dy));
      }
    }
    return finite;
  }

  // The Jacobian at y into J, as the accessor's mode says (f1 is overwritten).  Inlined: every
  // solver's kernel evaluates an analytic Jacobian under its own register budget (a function shared
  // by kernels of different budgets is compiled for the smallest).
  // In the FP32 mode a node whose FP32 right-hand side is not finite at any of
  // the difference points gets the FP64 forward-difference Jacobian instead:
  // FP32 resolves removable singularities of a model only to its own width
  // (the L-type calcium current of the TP06 epicardial model,
  // (V - 15) / (exp(c (V - 15)) - 1), is 0/0 in FP32 for |V - 15 mV| below
  // about 1e-6 mV), where FP64 is finite.
  template <class A>
  GOSS_HD static GOSS_INLINE void jacobian(const VecView& y, const double t, const A& acc, const VecView& f1,
                                           const VecView& J)
  {
    if constexpr (JacobianModeOf<A>::value == kJacobianFDFP32) {
      if (!fd_jacobian_f32(y, t, acc, f1, J)) fd_jacobian(y, t, acc, f1, J);
    } else if constexpr (JacobianModeOf<A>::value == kJacobianAnalytic) {
      // ODE::compute_jacobian overridden by the model (ODE.h:71; codegeneration.py:622-641), each entry
      // stored over the structure as soon as it is computed.
      (void)f1;
      double ys[S];
#pragma unroll
      for (int j = 0; j < S; ++j) ys[j] = y[j];
      M::template compute_jacobian_structured<double>(ys, t, acc.acc, J);
    } else {
      fd_jacobian(y, t, acc, f1, J);
    }
  }

  // F <- LU factors of I + scale * J (ImplicitODESolver.cpp:92-105 with
  // ODE.cpp:69-100); scale is -dt*alpha or -dt*gamma (ESDIRK23a.cpp:101-115).
  // J is read into registers, factorised there and stored once.
  GOSS_HD static void factorize(const VecView& J, const VecView& F, const double scale)
  {
    double j[M::jac_nnz], a[M::lu_nnz];
#pragma unroll
    for (int k = 0; k < M::jac_nnz; ++k) j[k] = J[k];
    M::lu_from_jacobian(j, scale, a);
    M::lu_factorize_sparse(a);
#pragma unroll
    for (int k = 0; k < M::lu_nnz; ++k) F[k] = a[k];
  }

  // ODE::forward_backward_subst (ODE.cpp:102-132) in place on xv, with the
  // factors and the vector read into registers first.
  GOSS_HD static void fb_subst(const VecView& F, const VecView& xv)
  {
    double a[M::lu_nnz], x[S];
#pragma unroll
    for (int k = 0; k < M::lu_nnz; ++k) a[k] = F[k];
#pragma unroll
    for (int i = 0; i < S; ++i) x[i] = xv[i];
    M::lu_solve_sparse(a, x);
#pragma unroll
    for (int i = 0; i < S; ++i) xv[i] = x[i];
  }

  GOSS_HD static GOSS_INLINE double norm(const double* v)
  {
    double acc = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) acc += v[i] * v[i];
    return m_sqrt(acc);
  }

  GOSS_HD static double norm(const VecView& v)
  {
    double acc = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) acc += v[i] * v[i];
    return m_sqrt(acc);
  }
};

// Per-node solver state, loaded into registers for a launch.
struct NodeImplicitState {
  double eta, es_eta, loc_error, prev_dt;
  int recompute, newton_its, always, justrefined, first_step, new_jacobian, jac_comp, rejects, accepted, failed,
      nfevals;
};

GOSS_HD GOSS_INLINE NodeImplicitState load_state(const ImplicitBuffers& B, unsigned ld, unsigned n)
{
  NodeImplicitState s;
  s.eta = B.ds[size_t(kEta) * ld + n];
  s.es_eta = B.ds[size_t(kEsEta) * ld + n];
  s.loc_error = B.ds[size_t(kLocError) * ld + n];
  s.prev_dt = B.ds[size_t(kPrevDt) * ld + n];
  s.recompute = B.is[size_t(kRecompute) * ld + n];
  s.newton_its = B.is[size_t(kNewtonIts) * ld + n];
  s.always = B.is[size_t(kAlways) * ld + n];
  s.justrefined = B.is[size_t(kJustRefined) * ld + n];
  s.first_step = B.is[size_t(kFirstStep) * ld + n];
  s.new_jacobian = B.is[size_t(kNewJacobian) * ld + n];
  s.jac_comp = B.is[size_t(kJacComp) * ld + n];
  s.rejects = B.is[size_t(kRejects) * ld + n];
  s.accepted = B.is[size_t(kAccepted) * ld + n];
  s.failed = B.is[size_t(kFailed) * ld + n];
  s.nfevals = B.is[size_t(kNfevals) * ld + n];
  return s;
}

GOSS_HD GOSS_INLINE void store_state(const ImplicitBuffers& B, unsigned ld, unsigned n, const NodeImplicitState& s)
{
  B.ds[size_t(kEta) * ld + n] = s.eta;
  B.ds[size_t(kEsEta) * ld + n] = s.es_eta;
  B.ds[size_t(kLocError) * ld + n] = s.loc_error;
  B.ds[size_t(kPrevDt) * ld + n] = s.prev_dt;
  B.is[size_t(kRecompute) * ld + n] = s.recompute;
  B.is[size_t(kNewtonIts) * ld + n] = s.newton_its;
  B.is[size_t(kAlways) * ld + n] = s.always;
  B.is[size_t(kJustRefined) * ld + n] = s.justrefined;
  B.is[size_t(kFirstStep) * ld + n] = s.first_step;
  B.is[size_t(kNewJacobian) * ld + n] = s.new_jacobian;
  B.is[size_t(kJacComp) * ld + n] = s.jac_comp;
  B.is[size_t(kRejects) * ld + n] = s.rejects;
  B.is[size_t(kAccepted) * ld + n] = s.accepted;
  B.is[size_t(kFailed) * ld + n] = s.failed;
  B.is[size_t(kNfevals) * ld + n] = s.nfevals;
}

// ---------------------------------------------------------------------------
// ImplicitEuler
// ---------------------------------------------------------------------------
template <class M>
struct ImplicitEulerNode {
  static constexpr int S = M::num_states;
  using I = Implicit<M>;

  // ImplicitODESolver::newton_solve (ImplicitODESolver.cpp:114-245) with the
  // diagonal mass matrix differential_states (the identity unless the model is
  // a DAE); z is the work vector kIeZ.  ImplicitEuler calls it
  // with prev = 0 and alpha = 1 (ImplicitEuler.cpp:85), ThetaSolver with
  // prev = (1 - theta) f(y, t) in kIePrev and alpha = theta
  // (ThetaSolver.cpp:86).
  template <bool PREV, class A>
  GOSS_HD static bool newton_solve(const NodeViews& v, const double t, const double dt, const double alpha,
                                   const A& acc, const ImplicitParams& p, NodeImplicitState& st, const bool always)
  {
    const VecView z = v.w(kIeZ), yzv = v.w(kIeYZ), bv = v.w(kIeB), f1 = v.w(kIeF1);
    bool step_ok = true;
    st.newton_its = 0;
    double relative_previous_residual = 1.0, relative_residual = 1.0, previous_residual = 1.0,
           initial_residual = 1.0, residual;
#pragma unroll 1
    do {
      {
        double yz[S], f[S], b[S];
#pragma unroll
        for (int i = 0; i < S; ++i) yz[i] = v.y[i] + z[i];
        M::template eval<double>(yz, t, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          if constexpr (PREV) {
            const VecView prev = v.w(kIePrev);
            b[i] = -z[i] * double(M::differential_states[i]) + dt * (prev[i] + alpha * f[i]);
          } else {
            b[i] = -z[i] * double(M::differential_states[i]) + dt * (0.0 + 1.0 * f[i]);
          }
          bv[i] = b[i];
          yzv[i] = yz[i];
        }
        residual = I::norm(b);
      }
      if (relative_residual < p.relative_tolerance) break;
      if (st.recompute || always) {
        I::jacobian(yzv, t, acc, f1, v.J);
        I::factorize(v.J, v.F, -dt * alpha);
        st.jac_comp += 1;
        st.recompute = 0;
      }
      I::fb_subst(v.F, bv);  // bv <- dz
      if (st.newton_its == 0) {
        initial_residual = residual;
        st.eta = st.eta > kGossEps ? st.eta : kGossEps;
        st.eta = m_pow(st.eta, 0.8);
      } else {
        relative_previous_residual = residual / previous_residual;
        st.recompute = relative_previous_residual >= p.max_relative_previous_residual;
        if (relative_previous_residual >= 1) {
          step_ok = false;
          st.rejects++;
          st.recompute = 1;
          break;
        }
        const double scaled = smax(m_pow(relative_previous_residual, double(p.max_iterations - st.newton_its)),
                                   kGossEps);
        if (residual > (p.kappa * p.relative_tolerance * (1 - relative_previous_residual) / scaled))
          st.recompute = 1;
        st.eta = relative_previous_residual / (1.0 - relative_previous_residual);
      }
      if (st.newton_its > p.max_iterations) {
        st.recompute = 1;
        st.rejects++;
        step_ok = false;
        break;
      }
#pragma unroll
      for (int i = 0; i < S; ++i) z[i] += bv[i];
      relative_residual = residual / initial_residual;
      previous_residual = residual;
      st.newton_its++;
    } while (st.eta * relative_residual >= p.kappa * p.relative_tolerance);
    return step_ok;
  }

  // ImplicitEuler::forward (ImplicitEuler.cpp:57-152).  Returns false when
  // the step size falls below min_dt (the reference throws).
  template <class A>
  GOSS_HD static bool forward(const NodeViews& v, double t, const double dt, const A& acc,
                              const ImplicitParams& p, NodeImplicitState& st)
  {
    const VecView z = v.w(kIeZ);
    const double t_end = t + dt;
    const double ldt_0 = p.ldt_0;
    double ldt = ldt_0 > 0 ? ldt_0 : dt;
    int num_refinements = 0;
    const double eps = kGossEps * 1000;
#pragma unroll 1
    while (true) {
#pragma unroll
      for (int i = 0; i < S; ++i) z[i] = 0.0;
      if (num_refinements > p.num_refinements) st.always = 1;
      const bool step_ok = newton_solve<false>(v, t + ldt, ldt, 1.0, acc, p, st, st.always != 0);
      if (step_ok) {
#pragma unroll
        for (int i = 0; i < S; ++i) v.y[i] += z[i];
        t += ldt;
        if (fabs(t - t_end) < eps) break;
        if (!st.justrefined) {
          const double tmp = 2.0 * ldt;
          if (ldt_0 > 0. && tmp >= ldt_0)
            ldt = ldt_0;
          else
            ldt = tmp;
        } else {
          st.justrefined = 0;
        }
        if ((t + ldt + kGossEps) > t_end) ldt = t_end - t;
      } else {
        ldt /= 2.0;
        if (ldt < p.min_dt) return false;
        st.recompute = 1;
        st.justrefined = 1;
        num_refinements += 1;
      }
    }
    return true;
  }
};

// ---------------------------------------------------------------------------
// ThetaSolver
// ---------------------------------------------------------------------------
template <class M>
struct ThetaNode {
  static constexpr int S = M::num_states;
  using E = ImplicitEulerNode<M>;

  // ThetaSolver::forward (ThetaSolver.cpp:48-155).  The sticky
  // always_recompute_jacobian is set while num_refinements <=
  // num_refinements_without_always_recomputing_jacobian (ThetaSolver.cpp:82-83),
  // i.e. from the first step on.  Returns false when the step size falls
  // below min_dt (the reference throws).
  template <class A>
  GOSS_HD static bool forward(const NodeViews& v, double t, const double dt, const A& acc, const ImplicitParams& p,
                              NodeImplicitState& st)
  {
    const VecView z = v.w(kIeZ), prev = v.w(kIePrev);
    const double theta = p.theta;
    const double t_end = t + dt;
    const double ldt_0 = p.ldt_0;
    double ldt = ldt_0 > 0 ? ldt_0 : dt;
    int num_refinements = 0;
    const double eps = kGossEps * 1000;
#pragma unroll
    for (int i = 0; i < S; ++i) prev[i] = 0.0;
#pragma unroll 1
    while (true) {
#pragma unroll
      for (int i = 0; i < S; ++i) z[i] = 0.0;
      if (fabs(theta - 1.0) > kGossEps) {
        double y[S], f[S];
#pragma unroll
        for (int i = 0; i < S; ++i) y[i] = v.y[i];
        M::template eval<double>(y, t, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) prev[i] = (1 - theta) * f[i];
      }
      if (num_refinements <= p.num_refinements) st.always = 1;
      const bool step_ok = E::template newton_solve<true>(v, t + theta * ldt, ldt, theta, acc, p, st, st.always != 0);
      if (step_ok) {
#pragma unroll
        for (int i = 0; i < S; ++i) v.y[i] += z[i];
        t += ldt;
        if (fabs(t - t_end) < eps) break;
        if (!st.justrefined) {
          const double tmp = 2.0 * ldt;
          if (ldt_0 > 0. && tmp >= ldt_0)
            ldt = ldt_0;
          else
            ldt = tmp;
        } else {
          st.justrefined = 0;
        }
        if ((t + ldt + kGossEps) > t_end) ldt = t_end - t;
      } else {
        ldt /= 2.0;
        if (ldt < p.min_dt) return false;
        st.justrefined = 1;
        num_refinements += 1;
      }
    }
    return true;
  }
};

// ---------------------------------------------------------------------------
// BasicImplicitEuler
// ---------------------------------------------------------------------------
template <class M>
struct BasicImplicitEulerNode {
  static constexpr int S = M::num_states;
  using I = Implicit<M>;

  // BasicImplicitEuler::forward (BasicImplicitEuler.cpp:62-181) with its
  // compute_factorized_jacobian (:46-60): full Newton, a new Jacobian at every
  // iteration, on the residual (y - y_prev) - dt f(y, t + ldt).  As in the
  // reference the residual and the clock advance use the whole interval dt and
  // the Jacobian the sub-step ldt (equal with the default internal time step).
  // A sub-step whose iteration does not converge within max_iterations keeps
  // its last iterate and is counted as rejected (the reference prints a
  // message and goes on).
  template <class A>
  GOSS_HD static void forward(const NodeViews& v, const double t, const double dt, const A& acc,
                              const ImplicitParams& p, NodeImplicitState& st)
  {
    const VecView prev = v.w(kIePrev), bv = v.w(kIeB), f1 = v.w(kIeF1);
    const double ldt_0 = p.ldt_0;
    const unsigned long nsteps = ldt_0 > 0 ? (unsigned long)ceil(dt / ldt_0 - 1.0E-12) : 1UL;
    const double ldt = dt / nsteps;
    const double rtol = p.relative_tolerance;
    double lt = t;
#pragma unroll 1
    for (unsigned long step = 0; step < nsteps; ++step) {
      double relative_residual = 1.0, initial_residual = 1.0, residual;
      bool converged = false;
      st.newton_its = 0;
#pragma unroll
      for (int i = 0; i < S; ++i) prev[i] = v.y[i];
#pragma unroll 1
      while (!converged && st.newton_its < p.max_iterations) {
        {
          double y[S], f[S];
#pragma unroll
          for (int i = 0; i < S; ++i) y[i] = v.y[i];
          M::template eval<double>(y, lt + ldt, acc, f);
          double b[S];
#pragma unroll
          for (int i = 0; i < S; ++i) {
            b[i] = (y[i] - prev[i]) * double(M::differential_states[i]) - dt * f[i];
            bv[i] = b[i];
          }
          residual = I::norm(b);
        }
        if (st.newton_its == 0) initial_residual = residual;
        relative_residual = residual / initial_residual;
        if (relative_residual < rtol) {
          converged = true;
          break;
        }
        I::jacobian(v.y, lt + ldt, acc, f1, v.J);
        I::factorize(v.J, v.F, -ldt * 1.0);
        st.jac_comp += 1;
        I::fb_subst(v.F, bv);  // bv <- dz
#pragma unroll
        for (int i = 0; i < S; ++i) v.y[i] -= bv[i];
        ++st.newton_its;
      }
      if (converged)
        lt += dt;
      else
        st.rejects++;
    }
  }
};

// ---------------------------------------------------------------------------
// ESDIRK23a
// ---------------------------------------------------------------------------
struct EsdirkCoefficients {
  double gamma, a21, a22, bh1, bh2, b1, b2, b3, c2, c3, c4;
};

__host__ __device__ inline EsdirkCoefficients esdirk_coefficients()
{
  // ESDIRK23a.cpp:27-44; ESDIRK4O32 uses the same tableau (ESDIRK4O32.cpp:29-37).
  EsdirkCoefficients k;
  const double gamma = 0.43586652150845899942;
  k.gamma = gamma;
  k.a21 = gamma;
  k.a22 = gamma;
  k.bh1 = (-4 * gamma * gamma + 6 * gamma - 1) / (4 * gamma);  // a31
  k.bh2 = (-2 * gamma + 1) / (4 * gamma);                       // a32
  k.b1 = (6 * gamma - 1) / (12 * gamma);                        // a41
  k.b2 = -1 / ((24 * gamma - 12) * gamma);                      // a42
  k.b3 = (-6 * gamma * gamma + 6 * gamma - 1) / (6 * gamma - 3);  // a43
  k.c2 = 2.0 * gamma;
  k.c3 = 1.0;
  k.c4 = 1.0;
  return k;
}

// ---------------------------------------------------------------------------
// ESDIRK4O32
// ---------------------------------------------------------------------------
// Passes of the ESDIRK4O32 step loop after which a node gives up: the
// reference loops until it reaches the interval end, which a non-finite error
// or a step size halved to zero never does.
constexpr long kEsdirk4MaxPasses = 1000000;

template <class M>
struct Esdirk4o32Node {
  static constexpr int S = M::num_states;
  using I = Implicit<M>;
  using E = ImplicitEulerNode<M>;
  // AdaptiveImplicitSolver constants (AdaptiveImplicitSolver.cpp:43-66) and _iord (ESDIRK4O32.cpp:38).
  static constexpr double facmin = 0.5, facmaxb = 2.0, stabfac = 0.9, stabdown = 1.0, stabup = 1.2;
  static constexpr double iord = 3.0;

  // AdaptiveImplicitSolver::dtinit (AdaptiveImplicitSolver.cpp:68-128): f0 = f(y0, t) is given in z1;
  // y1, f1 are scratch rows.
  template <class A>
  GOSS_HD static double dtinit(const NodeViews& v, const double t, const A& acc, const ImplicitParams& p)
  {
    const VecView f0 = v.w(kE4Z1), y1 = v.w(kE4Yn);
    double dnf = 0.0, dny = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const double sk = p.atol + p.rtol * fabs(v.y[i]);
      double tmp = f0[i] / sk;
      dnf += tmp * tmp;
      tmp = v.y[i] / sk;
      dny += tmp * tmp;
    }
    double dt;
    if (dnf <= 1.0e-10 || dny <= 1.0e-10)
      dt = 1.0e-6;
    else
      dt = 0.01 * m_sqrt(dny / dnf);
    double der2 = 0.0;
    {
      double y[S], f[S];
#pragma unroll
      for (int i = 0; i < S; ++i) {
        y[i] = v.y[i] + dt * f0[i];
        y1[i] = y[i];
      }
      M::template eval<double>(y, t + dt, acc, f);
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const double sk = p.atol + p.rtol * fabs(y[i]);
        const double tmp = ((f[i] - f0[i]) / sk);
        der2 += tmp * tmp;
      }
    }
    der2 = m_sqrt(der2) / dt;
    const double der12 = fabs(der2) >= m_sqrt(dnf) ? fabs(der2) : m_sqrt(dnf);
    double dt1;
    if (der12 <= 1.0e-15)
      dt1 = 1.0e-6 > fabs(dt) * 1.0e-3 ? 1.0e-6 : fabs(dt) * 1.0e-3;
    else
      dt1 = m_pow((0.01 / der12), (1.0 / iord));
    if (100 * fabs(dt) < dt1) return 100 * fabs(dt);
    return dt1;
  }

  // newton_solve (ImplicitODESolver.cpp:114-245) on stage zs (copied in and
  // out of the Newton row kIeZ) with prev in kIePrev and y0 = y.
  template <class A>
  GOSS_HD static bool stage(const VecView& zs, const double t, const double dt, const double alpha, const NodeViews& v,
                            const A& acc, const ImplicitParams& p, NodeImplicitState& st)
  {
    const VecView z = v.w(kIeZ);
    if (zs.a != z.a) {
#pragma unroll
      for (int i = 0; i < S; ++i) z[i] = zs[i];
    }
    const bool ok = E::template newton_solve<true>(v, t, dt, alpha, acc, p, st, p.always_recompute_jacobian != 0);
    if (zs.a != z.a) {
#pragma unroll
      for (int i = 0; i < S; ++i) zs[i] = z[i];
    }
    return ok;
  }

  // ESDIRK4O32::forward (ESDIRK4O32.cpp:92-298) with new_time_step
  // (AdaptiveImplicitSolver.cpp:130-200), as the reference computes it: the
  // second stage's right-hand side is evaluated at the increment z2, and the
  // solution lives in two buffers, the node's states (the caller's y, A) and
  // the row kE4Yn (the solver's member yn, B).  `y` starts on A; an accepted
  // step swaps the pointers `y` and `yn0` while the new solution is always
  // written to B, so once `y` is on B the next step overwrites it in place
  // and is tested against the stale A; at the end B is copied to A if `y`
  // is on B (ESDIRK4O32.cpp:255-297).  The solver's _t, _dt and step flags
  // are local to the call; _ldt, facmax, err_old and dt_old persist in the
  // node's rows `ds`.  Returns false when the pass limit is reached.
  template <class A>
  GOSS_HD static bool forward(const NodeViews& v, const VecView& ds, const double t, const double interval,
                              const A& acc, const ImplicitParams& p, NodeImplicitState& st)
  {
    const EsdirkCoefficients k = esdirk_coefficients();
    const double a33 = k.gamma, a44 = k.gamma, bh3 = k.gamma, b4 = k.gamma;
    const VecView z1 = v.w(kE4Z1), z2 = v.w(kE4Z2), z3 = v.w(kE4Z3), z4 = v.w(kIeZ), ynb = v.w(kE4Yn),
                  yh = v.w(kE4Yh), prev = v.w(kIePrev);
    bool y_on_b = false;  // `y` points to B
    double ti = t;
    const double t_end = ti + interval;
    double dt = ds[kAiLdt];
    double facmax = ds[kAiFacmax], err_old = ds[kErrOld], dt_old = ds[kDtOld];
    bool reached_tend = false;
    if (dt < 0) {
      {
        double y[S], f[S];
#pragma unroll
        for (int i = 0; i < S; ++i) y[i] = v.y[i];
        M::template eval<double>(y, ti, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) z1[i] = f[i];
      }
      dt = dtinit(v, ti, acc, p);  // y1 = yn0 = B
      st.nfevals += 2;
    }
    bool ok = true;
    long passes = 0;
#pragma unroll 1
    while (!reached_tend) {
      if (++passes > kEsdirk4MaxPasses) {
        ok = false;
        break;
      }
      NodeViews vy = v;
      vy.y = y_on_b ? ynb : v.y;
      const VecView yv = vy.y, yn0 = y_on_b ? v.y : ynb;
      {
        double y[S], f[S];
#pragma unroll
        for (int i = 0; i < S; ++i) y[i] = yv[i];
        M::template eval<double>(y, ti, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          z1[i] = f[i];
          prev[i] = k.a21 * f[i];
          z2[i] = 0.0;
        }
      }
      st.nfevals += 1;
      if (!stage(z2, ti + k.c2 * dt, dt, k.a22, vy, acc, p, st)) {
        st.recompute = 1;
        dt /= 2.0;
        continue;
      }
      {
        // The stage evaluation reads z2 itself (ESDIRK4O32.cpp:169), as the reference does.
        double zz[S], f[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
          zz[i] = z2[i];
          z3[i] = zz[i] + yv[i];
        }
        M::template eval<double>(zz, ti + k.c2 * dt, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          z2[i] = f[i];
          prev[i] = k.bh1 * z1[i] + k.bh2 * f[i];
        }
      }
      st.nfevals += 1;
      if (!stage(z3, ti + k.c3 * dt, dt, a33, vy, acc, p, st)) {
        st.recompute = 1;
        dt /= 2.0;
        continue;
      }
      {
        double zz[S], f[S];
#pragma unroll
        for (int i = 0; i < S; ++i) zz[i] = z3[i] + yv[i];
        M::template eval<double>(zz, ti + k.c3 * dt, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          z3[i] = f[i];
          const double y0 = yv[i];
          const double yhi = y0 + dt * (k.bh1 * z1[i] + k.bh2 * z2[i] + bh3 * f[i]);
          yh[i] = yhi;
          z4[i] = yhi - y0;
          prev[i] = k.b1 * z1[i] + k.b2 * z2[i] + k.b3 * f[i];
        }
      }
      st.nfevals += 1;
      if (!stage(z4, ti + k.c4 * dt, dt, a44, vy, acc, p, st)) {
        st.recompute = 1;
        dt /= 2.0;
        continue;
      }
      double err = 0.0;
      {
        double zz[S], f[S];
#pragma unroll
        for (int i = 0; i < S; ++i) zz[i] = z4[i] + yv[i];
        M::template eval<double>(zz, ti + k.c4 * dt, acc, f);
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const double yni = yv[i] + dt * (k.b1 * z1[i] + k.b2 * z2[i] + k.b3 * z3[i] + b4 * f[i]);
          ynb[i] = yni;  // B, which is y itself when y is on B
          const double e = yh[i] - yni;
          // new_time_step(y, yn0, e): sk = atol + rtol max(|y|, |yn0|) (std::max semantics).
          const double yi_abs = fabs(yv[i]), yni_abs = fabs(yn0[i]);
          const double mx = smax(yi_abs, yni_abs);
          const double sk = p.atol + p.rtol * mx;
          const double tmp = e / sk;
          err += tmp * tmp;
        }
      }
      st.nfevals += 1;
      // new_time_step (AdaptiveImplicitSolver.cpp:130-200).
      const double dt_prev = dt;
      st.recompute = 1;
      err = m_sqrt(err / S);
      bool accepted;
      if (err <= 1) {
        ti += dt;
        st.accepted++;
        accepted = true;
        if (fabs(ti - t_end) < 1e-14) reached_tend = true;
      } else {
        st.rejects++;
        accepted = false;
      }
      const double lstabfac = stabfac * (2 * p.max_iterations + 1) / (2.0 * p.max_iterations + st.newton_its);
      double fac = lstabfac * m_pow((1.0 / err), (1.0 / (iord + 1)));
      if (facmin > fac)
        fac = facmin;
      else if (fac > facmax)
        fac = facmax;
      facmax = accepted ? facmaxb : 1.0;
      if (err_old > 0) fac *= dt / dt_old * m_pow((err_old / err), (1.0 / (iord + 1)));
      if (fac < stabup && fac > stabdown) {
        fac = 1.0;
        st.recompute = 0;
      }
      dt *= fac;
      ds[kAiLdt] = dt;
      if (ti + dt >= t_end) dt = t_end - ti;
      dt_old = dt_prev;
      err_old = err;
      if (accepted) y_on_b = !y_on_b;
    }
    if (y_on_b) {
#pragma unroll
      for (int i = 0; i < S; ++i) v.y[i] = ynb[i];
    }
    ds[kAiFacmax] = facmax;
    ds[kErrOld] = err_old;
    ds[kDtOld] = dt_old;
    return ok;
  }
};

template <class M>
struct EsdirkNode {
  static constexpr int S = M::num_states;
  using I = Implicit<M>;
  static constexpr double s_fac = 0.9, tol = 1.0e-5, etamin = 1.0e-10, kappa = 1.0e-1;

  template <class A>
  GOSS_HD static void ode_jacobian(const NodeViews& v, const double t, const A& acc, NodeImplicitState& st)
  {
    I::jacobian(v.y, t, acc, v.w(kEsF1), v.J);
    st.jac_comp++;
    st.new_jacobian = 1;
  }

  GOSS_HD static void factorized(const EsdirkCoefficients& k, const double dt, const NodeViews& v)
  {
    I::factorize(v.J, v.F, -dt * k.a22);
  }

  // compute_stage_val (ESDIRK23a.cpp:201-258).  The Newton increment of the
  // previous iteration enters only through its norm, kept as prev_ndz.
  template <class A>
  GOSS_HD static bool stage(const EsdirkCoefficients& k, const VecView& z, const NodeViews& v, const double t_s,
                            const double dt, const A& acc, const ImplicitParams& p, NodeImplicitState& st)
  {
    const VecView prev = v.w(kEsPrev), dz = v.w(kEsDZ);
    st.newton_its = 0;
    bool conv_ok = true;
    st.es_eta = m_pow(smax(st.es_eta, etamin), 0.8);
    double ndz, prev_ndz = 0.0;
#pragma unroll 1
    do {
      {
        double y_int[S], z_int[S];
#pragma unroll
        for (int i = 0; i < S; ++i) y_int[i] = v.y[i] + dt * prev[i] + k.gamma * dt * z[i];
        M::template eval<double>(y_int, t_s, acc, z_int);
#pragma unroll
        for (int i = 0; i < S; ++i) dz[i] = -z[i] + z_int[i];
      }
      I::fb_subst(v.F, dz);
      ndz = I::norm(dz);
      if (st.newton_its >= 1) {
        const double theta = ndz / prev_ndz;
        if (!(theta <= 1)) {
          st.rejects++;
          conv_ok = false;
          break;
        }
        st.es_eta = theta / (1.0 - theta);
      }
      if (st.newton_its > p.max_iterations) {
        st.rejects++;
        conv_ok = false;
        break;
      }
#pragma unroll
      for (int i = 0; i < S; ++i) z[i] += dz[i];
      prev_ndz = ndz;
      st.newton_its++;
    } while (st.es_eta * ndz > kappa * tol);
    return conv_ok;
  }

  // do_step (ESDIRK23a.cpp:134-199): the three implicit stages run through
  // one loop over their coefficients.
  template <class A>
  GOSS_HD static bool do_step(const EsdirkCoefficients& k, const double t0, const double dt, const NodeViews& v,
                              const A& acc, const ImplicitParams& p, NodeImplicitState& st)
  {
    const VecView z1 = v.w(kEsZ1), z2 = v.w(kEsZ2), z3 = v.w(kEsZ3), z4 = v.w(kEsZ4), prev = v.w(kEsPrev);
    {
      double y0[S], f[S];
#pragma unroll
      for (int i = 0; i < S; i++) y0[i] = v.y[i];
      M::template eval<double>(y0, t0, acc, f);
#pragma unroll
      for (int i = 0; i < S; i++) z1[i] = f[i];
    }
    bool ok = true;
#pragma unroll 1
    for (int s = 1; s <= 3 && ok; ++s) {
      const VecView zs = v.w(kEsZ1 + s), zp = v.w(kEsZ1 + s - 1);
      double c1, c2 = 0.0, c3 = 0.0, cs;
      if (s == 1) {
        c1 = k.a21;
        cs = k.c2;
      } else if (s == 2) {
        c1 = k.bh1;
        c2 = k.bh2;
        cs = k.c3;
      } else {
        c1 = k.b1;
        c2 = k.b2;
        c3 = k.b3;
        cs = k.c4;
      }
#pragma unroll
      for (int i = 0; i < S; i++) {
        zs[i] = zp[i];
        double pv = z1[i] * c1;
        if (s >= 2) pv = pv + z2[i] * c2;
        if (s >= 3) pv = pv + z3[i] * c3;
        prev[i] = pv;
      }
      ok = stage(k, zs, v, t0 + cs * dt, dt, acc, p, st);
    }
    if (ok) {
      const VecView yh = v.w(kEsDZ);
#pragma unroll
      for (int i = 0; i < S; i++)
        yh[i] = dt * ((k.b1 - k.bh1) * z1[i] + (k.b2 - k.bh2) * z2[i] + (k.b3 - k.gamma) * z3[i] +
                      k.gamma * z4[i]);
      I::fb_subst(v.F, yh);
      st.loc_error = I::norm(yh);
      if (st.loc_error > tol) {
        st.rejects++;
        ok = false;
      } else {
#pragma unroll
        for (int i = 0; i < S; i++)
          v.y[i] = v.y[i] + dt * (k.b1 * z1[i] + k.b2 * z2[i] + k.b3 * z3[i] + k.gamma * z4[i]);
        st.accepted++;
      }
    }
    return ok;
  }

  // forward (ESDIRK23a.cpp:280-316) with advance_one_step (117-132) and
  // reduce_time_step (260-278) inlined; the first attempt and the retries
  // after a reduction share one do_step call.
  template <class A>
  GOSS_HD static void forward(const NodeViews& v, const double t0, const double interval, const A& acc,
                              const ImplicitParams& p, NodeImplicitState& st)
  {
    const EsdirkCoefficients k = esdirk_coefficients();
    double local_t = 0;
    double dt = smax(s_fac * st.prev_dt * m_pow((tol / st.loc_error), 0.25), p.min_dt);
    dt = smin(dt, interval);
    if (st.first_step) {
      st.prev_dt = dt;
      ode_jacobian(v, t0, acc, st);
    }
#pragma unroll 1
    while (local_t < interval) {
      if (dt * 1.1 >= interval - local_t) dt = interval - local_t;
      if (st.newton_its >= 2) ode_jacobian(v, t0 + local_t, acc, st);
      if (fabs((dt - st.prev_dt) / st.prev_dt) > 0.2 || st.new_jacobian) factorized(k, dt, v);
      const double ts = t0 + local_t;
#pragma unroll 1
      while (true) {
        const bool accepted = do_step(k, ts, dt, v, acc, p, st);
        if (accepted || !(dt > p.min_dt)) break;
        if (!st.new_jacobian) ode_jacobian(v, ts, acc, st);
        // reduce_time_step (ESDIRK23a.cpp:260-278).  For a finite loc_error
        // at most 313 passes bring e below tol; the cap only ends the loop
        // for an infinite loc_error, where the reference would not return.
        double reduction_factor = 1.0;
        double e = st.loc_error;
        int passes = 0;
        do {
          reduction_factor *= 0.6;
          e *= 0.1;
        } while (e > tol && ++passes < 400);
        dt = smax(reduction_factor * dt, p.min_dt);
        factorized(k, dt, v);
      }
      st.new_jacobian = 0;
      st.prev_dt = dt;
      local_t += dt;
      dt = smax(s_fac * st.prev_dt * m_pow((tol / st.loc_error), 0.25), p.min_dt);
    }
    st.first_step = 0;
  }
};

GOSS_HD GOSS_INLINE NodeViews node_views(double* y, const ImplicitBuffers& B, int S, unsigned ld, unsigned node)
{
  NodeViews v;
  v.y = VecView{y + node, ld};
  v.work = B.work + node;
  v.ld = ld;
  v.S = S;
  v.J = VecView{B.jac + node, ld};
  v.F = VecView{B.jfac + node, ld};
  return v;
}

// The fused implicit kernel: `nouter` outer steps (t_k, interval_k).  JAC
// selects the Jacobian (JacobianMode); spf, the FP32 shared parameters, is
// read only by the FP32 Jacobian evaluations.
// Waves per SIMD of the implicit kernels.  They wait on dependent memory most of the time, so more
// waves beat the extra scratch spills: TP06 epi at 262,144 nodes, ImplicitEuler 3.65e7 node-steps/s
// with 1 wave (256 VGPRs + 256 AGPRs), 5.15e7 with 2, 6.64e7 with 4, 6.05e7 with 6, 5.69e7 with 8.
#ifndef GOSS_IMPLICIT_WAVES
#define GOSS_IMPLICIT_WAVES 4
#endif
// With the analytic Jacobian, the solvers that form a Jacobian at most Newton iterations
// (ImplicitEuler, ThetaSolver, BasicImplicitEuler) run 2 waves of 256 VGPRs: the evaluation's
// temporaries spill less.  TP06 epi, 262,144 nodes, node-steps/s with 4 and 2 waves: ImplicitEuler
// 1.24e8 / 1.33e8, ThetaSolver 1.07e8 / 1.11e8, BasicImplicitEuler 2.24e7 / 2.98e7.  ESDIRK23a reuses
// its Jacobian and keeps 4 (3.83e7 / 3.46e7).
#ifndef GOSS_IMPLICIT_WAVES_ANALYTIC_NEWTON
#define GOSS_IMPLICIT_WAVES_ANALYTIC_NEWTON 2
#endif
constexpr int implicit_waves(const int solver, const int jac)
{
  return (jac == kJacobianAnalytic && (solver == kImplicitEuler || solver == kThetaSolver ||
                                       solver == kBasicImplicitEuler))
             ? GOSS_IMPLICIT_WAVES_ANALYTIC_NEWTON
             : GOSS_IMPLICIT_WAVES;
}

template <class M, int SOLVER, int JAC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(implicit_waves(SOLVER, JAC))))
implicit_kernel(double* __restrict__ y, const double* __restrict__ fp, const double2* __restrict__ steps,
                const int nouter, const SharedParamsOf<M, double> sp, const SharedParamsOf<M, float> spf,
                const ImplicitParams ip, const ImplicitBuffers B, const unsigned N, const unsigned ld,
                const FieldOut fo)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  const unsigned node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= N) return;
  double fpv[NFP > 0 ? NFP : 1];
  float fpvf[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) {
    fpv[j] = fp[size_t(j) * ld + node];
    fpvf[j] = float(fpv[j]);
  }
  const DevAcc<M, double> acc64{sp, fpv};
  const DevAcc<M, float> accf{spf, fpvf};
  const JacAcc<M, JAC> jacc{acc64, accf};
  using Acc = std::conditional_t<JAC == kJacobianFD, DevAcc<M, double>, JacAcc<M, JAC>>;
  const Acc& acc = [&]() -> const Acc& {
    if constexpr (JAC == kJacobianFD) return acc64;
    else return jacc;
  }();
  (void)jacc;
  NodeImplicitState st = load_state(B, ld, node);
  const NodeViews v = node_views(y, B, S, ld, node);

#pragma unroll 1
  for (int k = 0; k < nouter; ++k) {
    if (!st.failed) {
      const double t = steps[k].x;
      const double dt = steps[k].y;
      if constexpr (SOLVER == kImplicitEuler) {
        if (!ImplicitEulerNode<M>::forward(v, t, dt, acc, ip, st)) {
          st.failed = 1;
          atomicOr(B.error, 1);
        }
      } else if constexpr (SOLVER == kThetaSolver) {
        if (!ThetaNode<M>::forward(v, t, dt, acc, ip, st)) {
          st.failed = 1;
          atomicOr(B.error, 1);
        }
      } else if constexpr (SOLVER == kBasicImplicitEuler) {
        BasicImplicitEulerNode<M>::forward(v, t, dt, acc, ip, st);
      } else if constexpr (SOLVER == kESDIRK4O32) {
        if (!Esdirk4o32Node<M>::forward(v, VecView{B.ds + node, ld}, t, dt, acc, ip, st)) {
          st.failed = 1;
          atomicOr(B.error, 1);
        }
      } else {
        EsdirkNode<M>::forward(v, t, dt, acc, ip, st);
      }
    }
    if (fo.out) {
      constexpr int Fs = M::num_field_states;
#pragma unroll
      for (int i = 0; i < Fs; ++i) {
        const size_t idx = fo.tangled ? size_t(node) * Fs + i : size_t(N) * i + node;
        fo.out[size_t(k) * fo.row_stride + idx] = v.y[M::field_state_indices[i]];
      }
    }
  }
  store_state(B, ld, node, st);
}

}  // namespace goss_hip
