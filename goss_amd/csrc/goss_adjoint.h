// Discrete adjoint of explicit time stepping -- GRL1 (GRL1.cpp:61-82), RL1
// (RL1.cpp:59-79), ExplicitEuler (ExplicitEuler.cpp:64-87) and RK4
// (RK4.cpp:68-102), one step per outer interval -- for gradients of a loss on
// sampled field states with respect to the initial states and the per-node
// field parameters.
//
// Forward (adjoint_record_kernel): the solver kernel's step, plus a checkpoint
// of all states every `ckpt_every` steps ([n_ckpt][S][ld]) and a sample of
// one field state every `sample_every` steps ([n_samples][N], FP64).
//
// Backward: the checkpoint segments are walked from the last to the first,
// two launches per segment.  adjoint_recompute_kernel recomputes the
// segment's states from its checkpoint with the same step code (bit-identical
// to the forward) into a segment buffer ([ckpt_every][S][ld]);
// adjoint_reverse_kernel runs the steps in reverse: at step k (y_k -> y_k+1) the loss gradient of a
// sample taken after step k is added to lam_k+1, and the step's adjoint
// (StepVjp: the generated grl1_vjp / rl1_vjp, or eval_vjp composed over the
// stages) gives lam_k = (dy_k+1/dy_k)^T lam_k+1 and the field-parameter
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

// S consecutive tape slots seen as a vector: v[i] is slot off + i.
template <class R>
struct TapeVec {
  VTape<R> t;
  int off;
  GOSS_HD GOSS_INLINE volatile R& operator[](const int i) const { return *t.at(off + i); }
};

// Adjoint of one step y -> y' of size h at time t: lp = (dy'/dy)^T lam and
// pb += (dy'/dp_field)^T lam.  TAPE: the generated *_tape forms, with the
// intermediates and adjoints in a per-node tape (FP64).
template <class M, int SOLVER, class R>
struct StepVjp;

// Tape slots one step's adjoint needs.
template <class M, int SOLVER>
constexpr int step_vjp_tape_slots()
{
  if constexpr (SOLVER == kGRL1 || SOLVER == kGRL2) return M::grl1_vjp_tape_slots;
  else if constexpr (SOLVER == kRL1 || SOLVER == kRL2) return M::rl1_vjp_tape_slots;
  else return M::eval_vjp_tape_slots;  // ExplicitEuler, RK2, RK4
}

template <class M, class R>
struct StepVjp<M, kGRL1, R> {
  template <class A, class TP>
  GOSS_HD GOSS_INLINE static void run(const R* y, double t, R h, const A& acc, R delta, const R* lam, R* lp, R* pb,
                                         const TP* tape, const TP&)
  {
    if constexpr (sizeof(R) == 8) M::template grl1_vjp_tape<R>(y, t, acc, h, delta, lam, lp, pb, *tape);
    else M::template grl1_vjp<R>(y, t, acc, h, delta, lam, lp, pb);
  }
};

template <class M, class R>
struct StepVjp<M, kRL1, R> {
  template <class A, class TP>
  GOSS_HD GOSS_INLINE static void run(const R* y, double t, R h, const A& acc, R delta, const R* lam, R* lp, R* pb,
                                         const TP* tape, const TP&)
  {
    if constexpr (sizeof(R) == 8) M::template rl1_vjp_tape<R>(y, t, acc, h, delta, lam, lp, pb, *tape);
    else M::template rl1_vjp<R>(y, t, acc, h, delta, lam, lp, pb);
  }
};

// y' = y + h f(y, t) (ExplicitEuler.cpp:76-85): lp = lam + f_y^T (h lam).
// In FP64 the seed is a work row (rows 0..S).
template <class M, class R>
struct StepVjp<M, kExplicitEuler, R> {
  template <class A, class TP>
  GOSS_HD GOSS_INLINE static void run(const R* y, double t, R h, const A& acc, R, const R* lam, R* lp, R* pb,
                                         const TP* tape, const TP& work)
  {
    constexpr int S = M::num_states;
    if constexpr (sizeof(R) == 8) {
      const TapeVec<R> seed_v{work, 0};
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] = lam[i];
        seed_v[i] = h * lam[i];
      }
      M::template eval_vjp_tape<R>(y, t, acc, seed_v, lp, pb, *tape);
    } else {
      R seed[S];
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] = lam[i];
        seed[i] = h * lam[i];
      }
      M::template eval_vjp<R>(y, t, acc, seed, lp, pb);
    }
  }
};

// RK4 (RK4.cpp:79-101): k1 = f(y, t), k2 = f(y + h/2 k1, t + h/2),
// k3 = f(y + h/2 k2, t + h/2), k4 = f(y + h k3, t + h),
// y' = y + h (k1 + 2 k2 + 2 k3 + k4) / 6.  The stage inputs are recomputed
// into the node's work rows (HBM: they live across the four reverse stages,
// which need the register file), then the stages are swept in reverse:
// kb_s = d y'/d k_s lam plus what later stages pass back through their inputs.
// In FP64 lam, the running result, the seed and the stage's y_bar are work
// rows too (rows 3S.. 7S), so that only one stage input is held in registers
// across a stage's sweep.
template <class M, class R>
struct StepVjp<M, kRK4, R> {
  template <class A, class TP>
  GOSS_HD GOSS_INLINE static void run(const R* y, double t, R h, const A& acc, R, const R* lam, R* lp, R* pb,
                                         const TP* tape, const TP& work)
  {
    constexpr int S = M::num_states;
    const double hd = double(h);
    const R hh = R(0.5) * h;
    {
      R k[S], yi[S];
      M::template eval<R>(y, t, acc, k);
#pragma unroll
      for (int i = 0; i < S; ++i) {
        yi[i] = y[i] + hh * k[i];
        work.put(i, yi[i]);
      }
      M::template eval<R>(yi, t + 0.5 * hd, acc, k);
#pragma unroll
      for (int i = 0; i < S; ++i) {
        yi[i] = y[i] + hh * k[i];
        work.put(S + i, yi[i]);
      }
      M::template eval<R>(yi, t + 0.5 * hd, acc, k);
#pragma unroll
      for (int i = 0; i < S; ++i) work.put(2 * S + i, y[i] + h * k[i]);
    }
    const R w1 = h / R(6), w2 = h / R(3);
    // Stage inputs, newest first, and the weight of lam in each stage's seed.
    const int in_row[4] = {2 * S, S, 0, -1};
    const R wl[4] = {w1, w2, w2, w1};
    const R wy[4] = {R(0), h, hh, hh};  // d (stage s input) / d k_(s-1) of the stage swept before
    if constexpr (sizeof(R) == 8) {
      const TapeVec<R> lam_v{work, 3 * S}, lp_v{work, 4 * S}, seed_v{work, 5 * S}, yb_v{work, 6 * S};
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lam_v[i] = lam[i];
        lp_v[i] = lam[i];
        yb_v[i] = R(0);
      }
#pragma unroll 1
      for (int s = 0; s < 4; ++s) {
        R yi[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
          seed_v[i] = wl[s] * lam_v[i] + wy[s] * yb_v[i];
          yb_v[i] = R(0);
          yi[i] = s < 3 ? R(work.get(in_row[s] + i)) : y[i];
        }
        const double ts = s == 0 ? t + hd : (s < 3 ? t + 0.5 * hd : t);
        M::template eval_vjp_tape<R>(yi, ts, acc, seed_v, yb_v, pb, *tape);
#pragma unroll
        for (int i = 0; i < S; ++i) lp_v[i] = lp_v[i] + yb_v[i];
      }
#pragma unroll
      for (int i = 0; i < S; ++i) lp[i] = lp_v[i];
    } else {
      R seed[S], yb[S], yi[S];
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] = lam[i];
        seed[i] = w1 * lam[i];  // kb4
        yb[i] = R(0);
        yi[i] = work.get(2 * S + i);
      }
      M::template eval_vjp<R>(yi, t + hd, acc, seed, yb, pb);  // stage 4: input y + h k3
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] += yb[i];
        seed[i] = w2 * lam[i] + h * yb[i];  // kb3
        yb[i] = R(0);
        yi[i] = work.get(S + i);
      }
      M::template eval_vjp<R>(yi, t + 0.5 * hd, acc, seed, yb, pb);  // stage 3: input y + h/2 k2
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] += yb[i];
        seed[i] = w2 * lam[i] + hh * yb[i];  // kb2
        yb[i] = R(0);
        yi[i] = work.get(i);
      }
      M::template eval_vjp<R>(yi, t + 0.5 * hd, acc, seed, yb, pb);  // stage 2: input y + h/2 k1
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] += yb[i];
        seed[i] = w1 * lam[i] + hh * yb[i];  // kb1
      }
      M::template eval_vjp<R>(y, t, acc, seed, lp, pb);  // stage 1: input y
    }
  }
};

// RK2, the midpoint rule (RK2.cpp:73-91): k1 = f(y, t), ym = y + h/2 k1,
// y' = y + h f(ym, t + h/2).  Reverse: yb = f_y(ym)^T (h lam); lp = lam + yb +
// f_y(y)^T (h/2 yb).  In FP64 ym, the seed and yb are work rows (0..3S).
template <class M, class R>
struct StepVjp<M, kRK2, R> {
  template <class A, class TP>
  GOSS_HD GOSS_INLINE static void run(const R* y, double t, R h, const A& acc, R, const R* lam, R* lp, R* pb,
                                         const TP* tape, const TP& work)
  {
    constexpr int S = M::num_states;
    const double hd = double(h);
    const R hh = R(0.5) * h;
    if constexpr (sizeof(R) == 8) {
      {
        R k[S];
        M::template eval<R>(y, t, acc, k);
#pragma unroll
        for (int i = 0; i < S; ++i) work.put(i, y[i] + hh * k[i]);
      }
      const TapeVec<R> seed_v{work, S}, yb_v{work, 2 * S};
      {
        R ym[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
          lp[i] = lam[i];
          seed_v[i] = h * lam[i];
          yb_v[i] = R(0);
          ym[i] = work.get(i);
        }
        M::template eval_vjp_tape<R>(ym, t + 0.5 * hd, acc, seed_v, yb_v, pb, *tape);
      }
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] += yb_v[i];
        seed_v[i] = hh * yb_v[i];
      }
      M::template eval_vjp_tape<R>(y, t, acc, seed_v, lp, pb, *tape);
    } else {
      R ym[S], seed[S], yb[S];
      {
        R k[S];
        M::template eval<R>(y, t, acc, k);
#pragma unroll
        for (int i = 0; i < S; ++i) ym[i] = y[i] + hh * k[i];
      }
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] = lam[i];
        seed[i] = h * lam[i];
        yb[i] = R(0);
      }
      M::template eval_vjp<R>(ym, t + 0.5 * hd, acc, seed, yb, pb);
#pragma unroll
      for (int i = 0; i < S; ++i) {
        lp[i] += yb[i];
        seed[i] = hh * yb[i];
      }
      M::template eval_vjp<R>(y, t, acc, seed, lp, pb);
    }
  }
};

// RL2 / GRL2 (RL2.cpp:75-85, GRL2.cpp:63-85): y2 = y + inc(y; h/2), then
// y' = y + inc(y2; h), both stages evaluated at time t.  With V(x, k, mu) the
// one-stage VJP mu + (d inc(x; k)/dx)^T mu (rl1_vjp / grl1_vjp):
//   lam2 = V(y2, h, lam) - lam,   lp = lam + V(y, h/2, lam2),
// and the parameter gradients of both calls add up.  y2, lam2 and copies of lam
// and y are work rows (0..4S), so that few vectors are live in registers at once.
template <class M, class R, bool GEN>
struct TwoStageRLVjp {
  template <class A, class TP, class MV, class OV>
  GOSS_HD GOSS_INLINE static void vjp(const R* x, double t, R k, const A& acc, R delta, const MV mu, const OV out,
                                     R* pb, const TP* tape)
  {
    if constexpr (GEN) {
      if constexpr (sizeof(R) == 8) M::template grl1_vjp_tape<R>(x, t, acc, k, delta, mu, out, pb, *tape);
      else M::template grl1_vjp<R>(x, t, acc, k, delta, mu, out, pb);
    } else {
      if constexpr (sizeof(R) == 8) M::template rl1_vjp_tape<R>(x, t, acc, k, delta, mu, out, pb, *tape);
      else M::template rl1_vjp<R>(x, t, acc, k, delta, mu, out, pb);
    }
  }
  template <class A, class TP>
  GOSS_HD GOSS_INLINE static void run(const R* y, double t, R h, const A& acc, R delta, const R* lam, R* lp, R* pb,
                                      const TP* tape, const TP& work)
  {
    constexpr int S = M::num_states;
    const R hh = R(0.5) * h;
    if constexpr (sizeof(R) == 8) {
      // Work rows: y2 (0..S), lam2 (S..2S), lam (2S..3S), y (3S..4S).  lam and y are copied
      // out first, so that their registers are free while y2 is recomputed (a full
      // linearized_eval) and while the first VJP runs; y2 is recomputed into rows too.
      const TapeVec<R> lam2_v{work, S}, lam_v{work, 2 * S}, y_v{work, 3 * S};
#pragma unroll
      for (int i = 0; i < S; ++i) lam_v[i] = lam[i];
      {
        R y2[S];
        rl_point<M, R, GEN>(y2, y, y, t, hh, acc, delta);
#pragma unroll
        for (int i = 0; i < S; ++i) work.put(i, y2[i]);
      }
#pragma unroll
      for (int i = 0; i < S; ++i) y_v[i] = y[i];
      {
        R y2[S];
#pragma unroll
        for (int i = 0; i < S; ++i) y2[i] = work.get(i);
        vjp(y2, t, h, acc, delta, lam_v, lam2_v, pb, tape);
      }
#pragma unroll
      for (int i = 0; i < S; ++i) lam2_v[i] = lam2_v[i] - lam_v[i];
      {
        R yr[S];
#pragma unroll
        for (int i = 0; i < S; ++i) yr[i] = y_v[i];
        vjp(yr, t, hh, acc, delta, lam2_v, lp, pb, tape);
      }
#pragma unroll
      for (int i = 0; i < S; ++i) lp[i] += lam_v[i];
    } else {
      // The register VJPs take arrays: the same rows, copied into registers call by call.
      const TapeVec<R> lam2_v{work, S}, lam_v{work, 2 * S}, y_v{work, 3 * S};
#pragma unroll
      for (int i = 0; i < S; ++i) lam_v[i] = lam[i];
      {
        R y2[S];
        rl_point<M, R, GEN>(y2, y, y, t, hh, acc, delta);
#pragma unroll
        for (int i = 0; i < S; ++i) work.put(i, y2[i]);
      }
#pragma unroll
      for (int i = 0; i < S; ++i) y_v[i] = y[i];
      {
        R y2[S], lr[S], g[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
          y2[i] = work.get(i);
          lr[i] = lam_v[i];
        }
        vjp(y2, t, h, acc, delta, lr, g, pb, tape);
#pragma unroll
        for (int i = 0; i < S; ++i) lam2_v[i] = g[i] - lr[i];
      }
      {
        R yr[S], l2[S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
          yr[i] = y_v[i];
          l2[i] = lam2_v[i];
        }
        vjp(yr, t, hh, acc, delta, l2, lp, pb, tape);
      }
#pragma unroll
      for (int i = 0; i < S; ++i) lp[i] += lam_v[i];
    }
  }
};

template <class M, class R>
struct StepVjp<M, kRL2, R> : TwoStageRLVjp<M, R, false> {};
template <class M, class R>
struct StepVjp<M, kGRL2, R> : TwoStageRLVjp<M, R, true> {};

// Per-node work rows a step's adjoint keeps in HBM (besides the FP64 tape).
template <class M, int SOLVER>
constexpr int step_vjp_work_rows()
{
  constexpr int S = M::num_states;
  return SOLVER == kRK4 ? 7 * S
       : SOLVER == kRK2 ? 3 * S
       : (SOLVER == kRL2 || SOLVER == kGRL2) ? 4 * S
       : SOLVER == kExplicitEuler ? S : 0;
}

// Steps per reverse-sweep launch when every state is checkpointed (checkpoint interval 1).
constexpr int kReverseChunk = 250;

struct AdjointShape {
  int n;             // outer steps
  int ckpt_every;    // steps per checkpoint segment
  int sample_every;  // steps between samples (sample j after step (j+1)*sample_every)
  int field_index;   // state index that is sampled
};

template <class M, class R, int SOLVER>
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
    Step<M, SOLVER, R, false>::run(st, steps[k].x, steps[k].y, acc, delta);
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
template <class M, class R, int SOLVER>
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
    if (k + 1 < k1) Step<M, SOLVER, R, false>::run(st, steps[k].x, steps[k].y, acc, delta);
  }
}

// Backward, part 2 of a segment: the reverse sweep over its steps.  lam
// ([S][ld], precision R) and the parameter gradients ([NFP][ld], FP64) carry
// over between segments in HBM; `first` starts them at zero, `last` writes
// the results node-major.
// Waves per SIMD of the reverse sweep.  It is VALU-bound; a 2-wave register budget spills and
// was slower (FP32 TP06 GRL1 at 65,536 nodes x 20,000 steps: backward 137.5 ms against 102.5 ms).
#ifndef GOSS_REVERSE_WAVES
#define GOSS_REVERSE_WAVES 1
#endif

template <class M, class R, int SOLVER>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GOSS_REVERSE_WAVES)))
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
    // FP64: the reverse sweep's intermediates do not fit in registers; they go to the tape.
    // The tape base is made opaque per step so that the compiler does not hoist the
    // per-slot addresses (VGPR pairs) out of the step loop; the empty asm emits nothing.
    // The work rows (RK4's stage inputs) follow the FP64 tape.
    R* tp = tape;
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+s"(tp));
#endif
    const VTape<R> vt{tp, ld, node};
    const VTape<R> work{tp + (sizeof(R) == 8 ? size_t(step_vjp_tape_slots<M, SOLVER>()) : 0) * ld, ld, node};
    StepVjp<M, SOLVER, R>::run(yk, steps[k].x, R(steps[k].y), acc, delta, lam, lp, pb, &vt, work);
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
