// Discrete adjoint of GRL1 time stepping (GRL1.cpp:61-82, one step per outer
// interval) for gradients of a loss on sampled field states with respect to
// the initial states and the per-node field parameters.
//
// Forward (adjoint_record_kernel): the GRL1 kernel's step, plus a checkpoint
// of all states every `ckpt_every` steps ([n_ckpt][S][ld]) and a sample of
// one field state every `sample_every` steps ([n_samples][N], FP64).
//
// Backward (adjoint_backward_kernel): one launch walks the checkpoint
// segments from the last to the first.  Per segment each lane recomputes the
// segment's states from its checkpoint with the same step code (bit-identical
// to the forward), keeps them in a segment buffer ([ckpt_every][S][ld]), then
// runs the steps in reverse: at step k (y_k -> y_k+1) the loss gradient of a
// sample taken after step k is added to lam_k+1, and the generated grl1_vjp
// gives lam_k = lam_k+1 + (dy_k+1/dy_k)^T lam_k+1 and the field-parameter
// contribution (dy_k+1/dp)^T lam_k+1.  lam stays in registers across the
// whole sweep; parameter gradients accumulate in FP64.
//
// One node per lane, no inter-lane communication: nodes are independent.
#pragma once

#include "goss_kernels.h"

namespace goss_hip {

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
      R* c = ckpt + size_t(gk / shp.ckpt_every) * S * ld + node;
#pragma unroll
      for (int i = 0; i < S; ++i) c[size_t(i) * ld] = st.y[i];
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

template <class M, class R>
__global__ void __launch_bounds__(kBlock)
adjoint_backward_kernel(const R* __restrict__ ckpt, const R* __restrict__ fp, const double2* __restrict__ steps,
                        const SharedParamsOf<M, R> sp, const R delta, const unsigned N, const unsigned ld,
                        const AdjointShape shp, const double* __restrict__ grad_samples, R* __restrict__ seg,
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
#pragma unroll
  for (int i = 0; i < S; ++i) lam[i] = R(0);
  double pbar[NFP > 0 ? NFP : 1];
#pragma unroll
  for (int j = 0; j < NFP; ++j) pbar[j] = 0.0;

  const int nseg = (shp.n + shp.ckpt_every - 1) / shp.ckpt_every;
#pragma unroll 1
  for (int s = nseg - 1; s >= 0; --s) {
    const int k0 = s * shp.ckpt_every;
    const int k1 = min(shp.n, k0 + shp.ckpt_every);
    // Recompute the segment's states y_k0 .. y_k1-1 from the checkpoint.
    {
      NodeState<M, R, false> st;
      const R* c = ckpt + size_t(s) * S * ld + node;
#pragma unroll
      for (int i = 0; i < S; ++i) st.y[i] = c[size_t(i) * ld];
#pragma unroll 1
      for (int k = k0; k < k1; ++k) {
        R* out = seg + size_t(k - k0) * S * ld + node;
#pragma unroll
        for (int i = 0; i < S; ++i) out[size_t(i) * ld] = st.y[i];
        if (k + 1 < k1) Step<M, kGRL1, R, false>::run(st, steps[k].x, steps[k].y, acc, delta);
      }
    }
    // Reverse sweep over the segment.
#pragma unroll 1
    for (int k = k1 - 1; k >= k0; --k) {
      if ((k + 1) % shp.sample_every == 0) {
        const R gs = R(grad_samples[size_t((k + 1) / shp.sample_every - 1) * N + node]);
#pragma unroll
        for (int i = 0; i < S; ++i) lam[i] += (i == shp.field_index) ? gs : R(0);
      }
      R yk[S], lp[S], pb[NFP > 0 ? NFP : 1];
      const R* in = seg + size_t(k - k0) * S * ld + node;
#pragma unroll
      for (int i = 0; i < S; ++i) yk[i] = in[size_t(i) * ld];
#pragma unroll
      for (int j = 0; j < NFP; ++j) pb[j] = R(0);
      M::template grl1_vjp<R>(yk, steps[k].x, acc, R(steps[k].y), delta, lam, lp, pb);
#pragma unroll
      for (int i = 0; i < S; ++i) lam[i] = lp[i];
#pragma unroll
      for (int j = 0; j < NFP; ++j) pbar[j] += double(pb[j]);
    }
  }
#pragma unroll
  for (int i = 0; i < S; ++i) grad_y0[size_t(node) * S + i] = double(lam[i]);
#pragma unroll
  for (int j = 0; j < NFP; ++j) grad_fp[size_t(node) * NFP + j] = pbar[j];
}

}  // namespace goss_hip
