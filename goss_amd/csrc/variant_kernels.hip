// Kernel instantiations of one model variant in one precision family.
//
// Compiled once per (variant, family) with
//   -DGOSS_VARIANT=<name> and -DGOSS_PREC_F64 (FP64, -ffp-contract=off)
//                          or -DGOSS_PREC_F32 (FP32 and FP32C).
// Exports goss_register_<name>_<family>(VariantInfo&), called by the registry.
#include "goss_model_wrap.h"
#include "goss_layout.h"
#include "goss_registry.h"

#define GOSS_STR2(x) #x
#define GOSS_STR(x) GOSS_STR2(x)
#define GOSS_CAT3_(a, b, c) a##b##c
#define GOSS_CAT3(a, b, c) GOSS_CAT3_(a, b, c)

#include GOSS_STR(generated/GOSS_VARIANT.h)

namespace goss_hip {
namespace {

namespace NS = goss_models::GOSS_VARIANT;

// The generated namespace wrapped as a type for the kernel templates.
GOSS_MODEL_TYPE(M, NS);

template <int SOLVER, class R, bool COMP>
hipError_t launch_explicit(const ExplicitArgs& a)
{
  const SharedParamsOf<M, R> sp = make_shared_params<M, R>(a.shared);
  const dim3 grid((a.N + kBlock - 1) / kBlock);
  hipLaunchKernelGGL((explicit_kernel<M, SOLVER, R, COMP>), grid, dim3(kBlock), 0, a.stream,
                     static_cast<R*>(a.y), static_cast<R*>(a.ylo), static_cast<const R*>(a.fp), a.steps, a.nouter,
                     sp, a.ldt_0, R(a.delta), a.N, a.ld, a.fo);
  return hipGetLastError();
}

#if defined(GOSS_PREC_F64)
template <int SOLVER, int JAC>
hipError_t launch_implicit(const ImplicitArgs& a)
{
  const SharedParamsOf<M, double> sp = make_shared_params<M, double>(a.shared);
  const SharedParamsOf<M, float> spf = make_shared_params<M, float>(a.shared);
  const dim3 grid((a.N + kBlock - 1) / kBlock);
  hipLaunchKernelGGL((implicit_kernel<M, SOLVER, JAC>), grid, dim3(kBlock), 0, a.stream, a.y, a.fp, a.steps,
                     a.nouter, sp, spf, a.ip, a.B, a.N, a.ld, a.fo);
  return hipGetLastError();
}

// The five implicit solvers with one Jacobian mode.
template <int JAC>
void register_implicit(VariantInfo& v)
{
  v.implicit_launch[0][JAC] = &launch_implicit<kImplicitEuler, JAC>;
  v.implicit_launch[1][JAC] = &launch_implicit<kESDIRK23a, JAC>;
  v.implicit_launch[2][JAC] = &launch_implicit<kThetaSolver, JAC>;
  v.implicit_launch[3][JAC] = &launch_implicit<kBasicImplicitEuler, JAC>;
  v.implicit_launch[4][JAC] = &launch_implicit<kESDIRK4O32, JAC>;
}
#endif

template <class R>
hipError_t launch_rkf32(const AdaptiveArgs& a)
{
  const SharedParamsOf<M, R> sp = make_shared_params<M, R>(a.shared);
  const dim3 grid((a.N + kBlock - 1) / kBlock);
  hipLaunchKernelGGL((rkf32_kernel<M, R>), grid, dim3(kBlock), 0, a.stream, static_cast<R*>(a.y),
                     static_cast<const R*>(a.fp), a.steps, a.nouter, sp, a.ap, a.B, a.N, a.ld, a.fo);
  return hipGetLastError();
}

template <class R, int SOLVER>
hipError_t launch_adjoint_record(const AdjointArgs& a)
{
  const SharedParamsOf<M, R> sp = make_shared_params<M, R>(a.shared);
  const dim3 grid((a.N + kBlock - 1) / kBlock);
  hipLaunchKernelGGL((adjoint_record_kernel<M, R, SOLVER>), grid, dim3(kBlock), 0, a.stream, static_cast<R*>(a.y),
                     static_cast<const R*>(a.fp), a.steps, a.step0, a.nouter, sp, R(a.delta), a.N, a.ld, a.shape,
                     static_cast<R*>(a.ckpt), a.samples);
  return hipGetLastError();
}

template <class R, int SOLVER>
hipError_t launch_adjoint_backward(const AdjointArgs& a)
{
  const SharedParamsOf<M, R> sp = make_shared_params<M, R>(a.shared);
  const dim3 grid((a.N + kBlock - 1) / kBlock);
  if (a.shape.ckpt_every == 1) {
    // The record stored the state before every step: nothing to recompute.  The reverse
    // sweep reads the checkpoints in place, kReverseChunk steps per launch (the chunk plays
    // the role of a segment: k0 = chunk * kReverseChunk, states at ckpt + k * S * ld).
    AdjointShape shp = a.shape;
    shp.ckpt_every = kReverseChunk;
    const int nch = (shp.n + kReverseChunk - 1) / kReverseChunk;
    const size_t per = size_t(M::num_states) * a.ld;
    for (int c = nch - 1; c >= 0; --c) {
      const R* states = static_cast<const R*>(a.ckpt) + size_t(c) * kReverseChunk * per;
      hipLaunchKernelGGL((adjoint_reverse_kernel<M, R, SOLVER>), grid, dim3(kBlock), 0, a.stream,
                         static_cast<const R*>(a.fp), a.steps, sp, R(a.delta), a.N, a.ld, shp, c, a.grad_samples,
                         states, static_cast<R*>(a.tape), static_cast<R*>(a.lam), a.pbar, int(c == nch - 1),
                         int(c == 0), a.grad_y0, a.grad_fp);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  // Checkpoint segments, last to first.  The recompute of segment s-1 (on `aux`, into the other
  // segment buffer) runs while the reverse sweep of segment s (on `stream`) reads its own: the
  // two kernels fit on one SIMD together (a reverse-sweep wave and a recompute wave), so the
  // recompute fills issue slots the single reverse-sweep wave leaves idle.  Every kernel sees
  // the same inputs as in a serial order, so the results are the same bit for bit.
  const int nseg = (a.shape.n + a.shape.ckpt_every - 1) / a.shape.ckpt_every;
  R* const seg[2] = {static_cast<R*>(a.seg), static_cast<R*>(a.seg2)};
  hipError_t e = hipSuccess;
  auto recompute = [&](const int s) {
    const int b = s & 1;
    if (s + 2 < nseg && (e = hipStreamWaitEvent(a.aux, a.ev_rev[b], 0)) != hipSuccess) return;
    hipLaunchKernelGGL((adjoint_recompute_kernel<M, R, SOLVER>), grid, dim3(kBlock), 0, a.aux,
                       static_cast<const R*>(a.ckpt), static_cast<const R*>(a.fp), a.steps, sp, R(a.delta), a.N, a.ld,
                       a.shape, s, seg[b]);
    if ((e = hipGetLastError()) != hipSuccess) return;
    e = hipEventRecord(a.ev_rc[b], a.aux);
  };
  if ((e = hipEventRecord(a.ev_start, a.stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(a.aux, a.ev_start, 0)) != hipSuccess) return e;
  recompute(nseg - 1);
  if (e != hipSuccess) return e;
  for (int s = nseg - 1; s >= 0; --s) {
    const int b = s & 1;
    if (s > 0) {
      recompute(s - 1);
      if (e != hipSuccess) return e;
    }
    if ((e = hipStreamWaitEvent(a.stream, a.ev_rc[b], 0)) != hipSuccess) return e;
    hipLaunchKernelGGL((adjoint_reverse_kernel<M, R, SOLVER>), grid, dim3(kBlock), 0, a.stream, static_cast<const R*>(a.fp),
                       a.steps, sp, R(a.delta), a.N, a.ld, a.shape, s, a.grad_samples,
                       static_cast<const R*>(seg[b]), static_cast<R*>(a.tape), static_cast<R*>(a.lam),
                       a.pbar, int(s == nseg - 1), int(s == 0), a.grad_y0, a.grad_fp);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(a.ev_rev[b], a.stream)) != hipSuccess) return e;
  }
  return hipSuccess;
}

#if defined(GOSS_PREC_F64)
// ParameterizedODE::eval / linearized_eval and ODE::compute_jacobian
// (ODE.h:55-77, ODE.cpp:47-67) at n points, one per lane, node-major in and out.
struct FieldValues {
  double v[kMaxFieldParams > 0 ? kMaxFieldParams : 1];
};

template <int KIND>
__global__ void __launch_bounds__(kBlock) model_eval_kernel(const double* __restrict__ states,
                                                            const double* __restrict__ times, const unsigned long n,
                                                            const SharedParamsOf<M, double> sp, const FieldValues fv,
                                                            const int only_linear, const MonitoredList mon,
                                                            double* __restrict__ out0, double* __restrict__ out1)
{
  constexpr int S = M::num_states;
  const unsigned long p = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  double fpv[M::num_field_parameters > 0 ? M::num_field_parameters : 1];
#pragma unroll
  for (int j = 0; j < M::num_field_parameters; ++j) fpv[j] = fv.v[j];
  const DevAcc<M, double> acc{sp, fpv};
  double y[S], f[S];
#pragma unroll
  for (int i = 0; i < S; ++i) y[i] = states[p * S + i];
  const double t = times[p];
  if constexpr (KIND == 0) {
    M::template eval<double>(y, t, acc, f);
#pragma unroll
    for (int i = 0; i < S; ++i) out0[p * S + i] = f[i];
  } else if constexpr (KIND == 1) {
    double lin[S];
#pragma unroll
    for (int i = 0; i < S; ++i) lin[i] = 0.0;
    M::template linearized_eval<double>(y, t, acc, lin, f, only_linear != 0);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      out0[p * S + i] = lin[i];
      out1[p * S + i] = f[i];
    }
  } else if constexpr (KIND == 4) {
    // ODE::compute_jacobian (ODE.h:71) overridden by the generated analytic Jacobian.
    M::template compute_jacobian<double>(y, t, acc, out0 + p * S * S);
  } else if constexpr (KIND == 3) {
    // ParameterizedODE::eval_monitored (ParameterizedODE.h:90): the selected intermediates.
    double all[M::num_intermediates > 0 ? M::num_intermediates : 1];
    M::template eval_intermediates<double>(y, t, acc, all);
    for (int k = 0; k < mon.n; ++k) out0[p * mon.n + k] = all[mon.index[k]];
  } else {
    // ODE::compute_jacobian (ODE.cpp:47-67): forward differences column by column.
    double f2[S];
    M::template eval<double>(y, t, acc, f);
#pragma unroll 1
    for (int i = 0; i < S; ++i) {
      double yp[S];
#pragma unroll
      for (int k = 0; k < S; ++k) yp[k] = y[k];
      const double ysafe = y[i];
      const double mx = 1e-5 > fabs(ysafe) ? 1e-5 : fabs(ysafe);
      const double delta = m_sqrt(1e-15 * mx);
#pragma unroll
      for (int k = 0; k < S; ++k) yp[k] = (k == i) ? ysafe + delta : yp[k];
      M::template eval<double>(yp, t, acc, f2);
#pragma unroll
      for (int j = 0; j < S; ++j) out0[p * S * S + j * S + i] = (f2[j] - f[j]) / delta;
    }
  }
}

hipError_t launch_model_eval(const ModelEvalArgs& a)
{
  const SharedParamsOf<M, double> sp = make_shared_params<M, double>(a.params);
  FieldValues fv{};
  for (int j = 0; j < M::num_field_parameters; ++j) fv.v[j] = a.params[M::field_parameter_indices[j]];
  const dim3 grid(unsigned((a.n + kBlock - 1) / kBlock));
#define GOSS_MODEL_EVAL(K)                                                                                    \
  hipLaunchKernelGGL((model_eval_kernel<K>), grid, dim3(kBlock), 0, a.stream, a.states, a.times, a.n, sp, fv, \
                     a.only_linear, a.monitored, a.out0, a.out1)
  if (a.kind == 0)
    GOSS_MODEL_EVAL(0);
  else if (a.kind == 1)
    GOSS_MODEL_EVAL(1);
  else if (a.kind == 2)
    GOSS_MODEL_EVAL(2);
  else if (a.kind == 3)
    GOSS_MODEL_EVAL(3);
  else if (a.kind == 4)
    GOSS_MODEL_EVAL(4);
  else
    return hipErrorInvalidValue;
#undef GOSS_MODEL_EVAL
  return hipGetLastError();
}
#endif

template <class R, bool COMP>
void register_family(VariantInfo& v, int prec)
{
  v.explicit_launch[kExplicitEuler][prec] = &launch_explicit<kExplicitEuler, R, COMP>;
  v.explicit_launch[kRK2][prec] = &launch_explicit<kRK2, R, COMP>;
  v.explicit_launch[kRK4][prec] = &launch_explicit<kRK4, R, COMP>;
  v.explicit_launch[kRL1][prec] = &launch_explicit<kRL1, R, COMP>;
  v.explicit_launch[kRL2][prec] = &launch_explicit<kRL2, R, COMP>;
  v.explicit_launch[kGRL1][prec] = &launch_explicit<kGRL1, R, COMP>;
  v.explicit_launch[kGRL2][prec] = &launch_explicit<kGRL2, R, COMP>;
}

template <class R>
void register_adjoint(VariantInfo& v, const int prec)
{
  v.adjoint_record[adjoint_index(kGRL1)][prec] = &launch_adjoint_record<R, kGRL1>;
  v.adjoint_backward[adjoint_index(kGRL1)][prec] = &launch_adjoint_backward<R, kGRL1>;
  v.adjoint_record[adjoint_index(kRL1)][prec] = &launch_adjoint_record<R, kRL1>;
  v.adjoint_backward[adjoint_index(kRL1)][prec] = &launch_adjoint_backward<R, kRL1>;
  v.adjoint_record[adjoint_index(kExplicitEuler)][prec] = &launch_adjoint_record<R, kExplicitEuler>;
  v.adjoint_backward[adjoint_index(kExplicitEuler)][prec] = &launch_adjoint_backward<R, kExplicitEuler>;
  v.adjoint_record[adjoint_index(kRK4)][prec] = &launch_adjoint_record<R, kRK4>;
  v.adjoint_backward[adjoint_index(kRK4)][prec] = &launch_adjoint_backward<R, kRK4>;
  v.adjoint_record[adjoint_index(kRK2)][prec] = &launch_adjoint_record<R, kRK2>;
  v.adjoint_backward[adjoint_index(kRK2)][prec] = &launch_adjoint_backward<R, kRK2>;
  v.adjoint_record[adjoint_index(kRL2)][prec] = &launch_adjoint_record<R, kRL2>;
  v.adjoint_backward[adjoint_index(kRL2)][prec] = &launch_adjoint_backward<R, kRL2>;
  v.adjoint_record[adjoint_index(kGRL2)][prec] = &launch_adjoint_record<R, kGRL2>;
  v.adjoint_backward[adjoint_index(kGRL2)][prec] = &launch_adjoint_backward<R, kGRL2>;
}

void fill_info(VariantInfo& v)
{
  v.name = GOSS_STR(GOSS_VARIANT);
  v.S = NS::num_states;
  v.P = NS::num_parameters;
  v.NFS = NS::num_field_states;
  v.NFP = NS::num_field_parameters;
  v.NCOMP = num_comp<M>();
  v.jac_nnz = M::jac_nnz;
  v.lu_nnz = M::lu_nnz;
  v.num_intermediates = NS::num_intermediates;
  v.intermediate_names = NS::intermediate_names;
#define GOSS_ADJOINT_ROWS(SOLVER)                                                   \
  v.vjp_tape_slots[adjoint_index(SOLVER)] = step_vjp_tape_slots<M, SOLVER>();     \
  v.vjp_work_rows[adjoint_index(SOLVER)] = step_vjp_work_rows<M, SOLVER>();
  GOSS_ADJOINT_ROWS(kGRL1)
  GOSS_ADJOINT_ROWS(kRL1)
  GOSS_ADJOINT_ROWS(kExplicitEuler)
  GOSS_ADJOINT_ROWS(kRK4)
  GOSS_ADJOINT_ROWS(kRK2)
  GOSS_ADJOINT_ROWS(kRL2)
  GOSS_ADJOINT_ROWS(kGRL2)
#undef GOSS_ADJOINT_ROWS
  v.fsi = NS::field_state_indices;
  v.fpi = NS::field_parameter_indices;
  v.defaults = NS::default_parameters;
  v.ic = NS::initial_conditions;
  v.linear_terms = NS::linear_terms;
  v.differential_states = NS::differential_states;
  v.is_dae = NS::is_dae;
  v.state_names = NS::state_names;
  v.parameter_names = NS::parameter_names;
  v.field_state_names = NS::field_state_names;
  v.field_parameter_names = NS::field_parameter_names;
}

}  // namespace

#if defined(GOSS_PREC_F64) && defined(GOSS_PARITY_MATH)
// The parity build (tests only): the implicit solvers with the reference's FD
// Jacobian and RKF32, with host-identical FP64 elementary functions.
void GOSS_CAT3(goss_register_, GOSS_VARIANT, _f64)(VariantInfo& v)
{
  fill_info(v);
  register_implicit<kJacobianFD>(v);
  register_implicit<kJacobianAnalytic>(v);
  v.rkf32_launch[kFP64] = &launch_rkf32<double>;
}
#elif defined(GOSS_PREC_F64)
void GOSS_CAT3(goss_register_, GOSS_VARIANT, _f64)(VariantInfo& v)
{
  fill_info(v);
  register_family<double, false>(v, kFP64);
  register_implicit<kJacobianFD>(v);
  register_implicit<kJacobianAnalytic>(v);
  // The FP32 Jacobian for ImplicitEuler, ESDIRK23a and ThetaSolver.
  v.implicit_launch[0][kJacobianFDFP32] = &launch_implicit<kImplicitEuler, kJacobianFDFP32>;
  v.implicit_launch[1][kJacobianFDFP32] = &launch_implicit<kESDIRK23a, kJacobianFDFP32>;
  v.implicit_launch[2][kJacobianFDFP32] = &launch_implicit<kThetaSolver, kJacobianFDFP32>;
  v.rkf32_launch[kFP64] = &launch_rkf32<double>;
  register_adjoint<double>(v, kFP64);
  v.model_eval = &launch_model_eval;
}
#if defined(GOSS_PLUGIN)
// A model compiled at run time into its own library: the main library loads it
// (goss_hip_model_load) and calls this to fill the variant's FP64 entries.
extern "C" int goss_hip_plugin_register_f64(void* info, size_t info_size)
{
  if (info_size != sizeof(VariantInfo)) return -1;
  GOSS_CAT3(goss_register_, GOSS_VARIANT, _f64)(*static_cast<VariantInfo*>(info));
  return 0;
}
#endif
#elif defined(GOSS_PREC_F32)
void GOSS_CAT3(goss_register_, GOSS_VARIANT, _f32)(VariantInfo& v)
{
  fill_info(v);
  register_family<float, false>(v, kFP32);
  register_family<float, true>(v, kFP32C);
  v.rkf32_launch[kFP32] = &launch_rkf32<float>;
  register_adjoint<float>(v, kFP32);
}
#if defined(GOSS_PLUGIN)
extern "C" int goss_hip_plugin_register_f32(void* info, size_t info_size)
{
  if (info_size != sizeof(VariantInfo)) return -1;
  GOSS_CAT3(goss_register_, GOSS_VARIANT, _f32)(*static_cast<VariantInfo*>(info));
  return 0;
}
#endif
#else
#error "define GOSS_PREC_F64 or GOSS_PREC_F32"
#endif

}  // namespace goss_hip
