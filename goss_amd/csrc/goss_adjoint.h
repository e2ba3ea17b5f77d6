// Discrete adjoint of GRL1 time stepping (GRL1.cpp:61-82, one step per outer
// interval) for gradients of a loss on sampled field states with respect to
// the initial states and the per-node field parameters.
//
// Forward (adjoint_record_kernel): the GRL1 kernel's step, plus a checkpoint
// of all states every `ckpt_every` steps ([n_ckpt][S][ld]) and a sample of
// one field state every `sample_every` steps ([n_samples][N], FP64).
//
// Backward: the checkpoint segments are walked from the last to the first,
// two launches per segment.  adjoint_recompute_kernel recomputes the
// segment's states from its checkpoint with the same step code (bit-identical
// to the forward) into a segment buffer ([ckpt_every][S][ld]);
// adjoint_reverse_kernel runs the steps in reverse: at step k (y_k -> y_k+1) the loss gradient of a
// sample taken after step k is added to lam_k+1, and the generated grl1_vjp
// gives lam_k = lam_k+1 + (dy_k+1/dy_k)^T lam_k+1 and the field-parameter
// contribution (dy_k+1/dp)^T lam_k+1.  lam stays in registers within a
// segment and in HBM between segments; parameter gradients accumulate in
// FP64.  Splitting recompute and reverse sweep into two kernels keeps lam
// out of the recompute's register budget.  In FP32 the reverse
// sweep of one step runs in registers; in FP64 its intermediates and
// adjoints live in a per-node tape ([grl1_vjp_tape_slots][ld]) because they
// do not fit in the register file.
//
// One node per lane, no inter-lane communication: nodes are independent.
#pragma once

#include "goss_kernels.h"

namespace goss_hip {

// Per-node tape of grl1_vjp_tape: slot k at a[k * ld + lane].  Volatile
// accesses keep every intermediate in memory instead of registers (each read
// is a load), which bounds the registers the FP64 reverse sweep needs.
template <class R>
struct VTape {
  R* a;          // tape base (uniform across the wavefront)
  unsigned ld;
  unsigned lane; // node index
  GOSS_HD GOSS_INLINE volatile R* at(const int k) const
  {
    R* row = a + size_t(k) * ld;
#ifdef __HIP_DEVICE_COMPILE__
    // The row address is uniform; making it opaque per access (an empty asm with side
    // effects, kept in SGPRs) stops the compiler from keeping hundreds of slot addresses
    // alive across the sweep.  It emits no instruction.
    asm volatile("" : "+s"(row));
#endif
    return row + lane;
  }
  GOSS_HD GOSS_INLINE R get(const int k) const { return *at(k); }
  GOSS_HD GOSS_INLINE void put(const int k, const R v) const { *at(k) = v; }
  GOSS_HD GOSS_INLINE void add(const int k, const R v) const { put(k, get(k) + v); }
};

struct AdjointShape {
  int n;             // outer steps
  int ckpt_every;    // steps per checkpoint segment
  int sample_every;  // steps between samples (sample j after step (j+1)*sample_every)
  int field_index;   // state index that is sampled
};

template <class M, class R>
__global__ void __launch_bounds__(kBlock)
adjoint_record_kernel(R* __restrict__ y, const R* __restrict__ fp, const double2* __restrict__ steps, const int step0,
                      const int nouter, const SharedParamsOf<M, R> sp, const R delta, const unsigned N,
                      const unsigned ld, const AdjointShape shp, R* __restrict__ ckpt, double* __restrict__ samples)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  const unsigned node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= N) return;
  NodeState<M, R, false> st;
#pragma unroll
  for (int i = 0; i < S; ++i) st.y[i] = y[size_t(i) * ld + node];
  R fpv[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) fpv[j] = fp[size_t(j) * ld + node];
  const DevAcc<M, R> acc{sp, fpv};

#pragma unroll 1
  for (int k = 0; k < nouter; ++k) {
    const int gk = step0 + k;
    if (gk % shp.ckpt_every == 0) {
      R* c = ckpt + size_t(gk / shp.ckpt_every) * S * ld;
#pragma unroll
      for (int i = 0; i < S; ++i) c[size_t(i) * ld + node] = st.y[i];
    }
    Step<M, kGRL1, R, false>::run(st, steps[k].x, steps[k].y, acc, delta);
    if ((gk + 1) % shp.sample_every == 0) {
      R v = R(0);
#pragma unroll
      for (int i = 0; i < S; ++i) v = (i == shp.field_index) ? st.y[i] : v;  // static indices keep y in registers
      samples[size_t((gk + 1) / shp.sample_every - 1) * N + node] = double(v);
    }
  }
#pragma unroll
  for (int i = 0; i < S; ++i) y[size_t(i) * ld + node] = st.y[i];
}

// Backward, part 1 of a segment: recompute its states y_k0 .. y_k1-1 from the
// checkpoint into the segment buffer.
template <class M, class R>
__global__ void __launch_bounds__(kBlock)
adjoint_recompute_kernel(const R* __restrict__ ckpt, const R* __restrict__ fp, const double2* __restrict__ steps,
                         const SharedParamsOf<M, R> sp, const R delta, const unsigned N, const unsigned ld,
                         const AdjointShape shp, const int seg_index, R* __restrict__ seg)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  const unsigned node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= N) return;
  R fpv[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) fpv[j] = fp[size_t(j) * ld + node];
  const DevAcc<M, R> acc{sp, fpv};
  const int k0 = seg_index * shp.ckpt_every;
  const int k1 = min(shp.n, k0 + shp.ckpt_every);
  NodeState<M, R, false> st;
  const R* c = ckpt + size_t(seg_index) * S * ld;
#pragma unroll
  for (int i = 0; i < S; ++i) st.y[i] = c[size_t(i) * ld + node];
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
    R* out = seg + size_t(k - k0) * S * ld;
#pragma unroll
    for (int i = 0; i < S; ++i) out[size_t(i) * ld + node] = st.y[i];
    if (k + 1 < k1) Step<M, kGRL1, R, false>::run(st, steps[k].x, steps[k].y, acc, delta);
  }
}

// Backward, part 2 of a segment: the reverse sweep over its steps.  lam
// ([S][ld], precision R) and the parameter gradients ([NFP][ld], FP64) carry
// over between segments in HBM; `first` starts them at zero, `last` writes
// the results node-major.
template <class M, class R>
__global__ void __launch_bounds__(kBlock)
adjoint_reverse_kernel(const R* __restrict__ fp, const double2* __restrict__ steps, const SharedParamsOf<M, R> sp,
                       const R delta, const unsigned N, const unsigned ld, const AdjointShape shp, const int seg_index,
                       const double* __restrict__ grad_samples, const R* __restrict__ seg, R* __restrict__ tape,
                       R* __restrict__ lam_buf, double* __restrict__ pbar_buf, const int first, const int last,
                       double* __restrict__ grad_y0, double* __restrict__ grad_fp)
{
  constexpr int S = M::num_states;
  constexpr int NFP = M::num_field_parameters;
  const unsigned node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= N) return;
  R fpv[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) fpv[j] = fp[size_t(j) * ld + node];
  const DevAcc<M, R> acc{sp, fpv};
  R lam[S];
  double pbar[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int i = 0; i < S; ++i) lam[i] = first ? R(0) : lam_buf[size_t(i) * ld + node];
#pragma unroll
  for (int j = 0; j < NFP; ++j) pbar[j] = first ? 0.0 : pbar_buf[size_t(j) * ld + node];

  const int k0 = seg_index * shp.ckpt_every;
  const int k1 = min(shp.n, k0 + shp.ckpt_every);
#pragma unroll 1
  for (int k = k1 - 1; k >= k0; --k) {
    if ((k + 1) % shp.sample_every == 0) {
      const R gs = R(grad_samples[size_t((k + 1) / shp.sample_every - 1) * N + node]);
#pragma unroll
      for (int i = 0; i < S; ++i) lam[i] += (i == shp.field_index) ? gs : R(0);
    }
    R yk[S], lp[S], pb[NFP > 0 ? NFP : 1];
    const R* in = seg + size_t(k - k0) * S * ld;
#pragma unroll
    for (int i = 0; i < S; ++i) yk[i] = in[size_t(i) * ld + node];
#pragma unroll
    for (int j = 0; j < NFP; ++j) pb[j] = R(0);
    if constexpr (sizeof(R) == 8) {
      // FP64: the reverse sweep's intermediates do not fit in registers; they go to the tape.
      // The tape base is made opaque per step so that the compiler does not hoist the
      // ~440 per-slot addresses (VGPR pairs) out of the step loop; the empty asm emits nothing.
      R* tp = tape;
#ifdef __HIP_DEVICE_COMPILE__
      asm volatile("" : "+s"(tp));
#endif
      M::template grl1_vjp_tape<R>(yk, steps[k].x, acc, R(steps[k].y), delta, lam, lp, pb, VTape<R>{tp, ld, node});
    } else {
      M::template grl1_vjp<R>(yk, steps[k].x, acc, R(steps[k].y), delta, lam, lp, pb);
    }
#pragma unroll
    for (int i = 0; i < S; ++i) lam[i] = lp[i];
#pragma unroll
    for (int j = 0; j < NFP; ++j) pbar[j] += double(pb[j]);
  }
  if (last) {
#pragma unroll
    for (int i = 0; i < S; ++i) grad_y0[size_t(node) * S + i] = double(lam[i]);
#pragma unroll
    for (int j = 0; j < NFP; ++j) grad_fp[size_t(node) * NFP + j] = pbar[j];
  } else {
#pragma unroll
    for (int i = 0; i < S; ++i) lam_buf[size_t(i) * ld + node] = lam[i];
#pragma unroll
    for (int j = 0; j < NFP; ++j) pbar_buf[size_t(j) * ld + node] = pbar[j];
  }
}

}  // namespace goss_hip
