// The fused mixed-population kernel and its kind table.
//
// Compiled once with the FP32 family's flags, so that every population runs
// the same machine arithmetic as its own FP32 kernel.  Adding a (variant,
// solver) pair is one line of GOSS_BATCH_KINDS.
#include "goss_batch.h"

#include <cstring>

#include "goss_model_wrap.h"

#include "generated/tp06_m.h"
#include "generated/tp06_epi.h"
#include "generated/fitzhughnagumo.h"
#include "generated/lorentz.h"

// Models compiled into the fused kernel, and the (model, solver) pairs it runs.
#define GOSS_BATCH_MODELS(X) X(tp06_m) X(tp06_epi) X(fitzhughnagumo) X(lorentz)
#define GOSS_BATCH_KINDS(X) \
  X(tp06_m, kGRL1)          \
  X(tp06_epi, kGRL1)        \
  X(fitzhughnagumo, kRK4)   \
  X(lorentz, kRKF32)

namespace goss_hip {
namespace {

#define GOSS_BATCH_MODEL(V) GOSS_MODEL_TYPE(BM_##V, goss_models::V);
GOSS_BATCH_MODELS(GOSS_BATCH_MODEL)
#undef GOSS_BATCH_MODEL

enum BatchKindId : int {
#define GOSS_BATCH_ENUM(V, SOLVER) kBatch_##V##_##SOLVER,
  GOSS_BATCH_KINDS(GOSS_BATCH_ENUM)
#undef GOSS_BATCH_ENUM
  kNumBatchKinds
};

template <class M>
void make_params(const double* shared, void* out)
{
  const SharedParamsOf<M, float> sp = make_shared_params<M, float>(shared);
  std::memcpy(out, &sp, sizeof(sp));
}

const BatchKind kKinds[] = {
#define GOSS_BATCH_ROW(V, SOLVER) {#V, SOLVER, sizeof(SharedParamsOf<BM_##V, float>), &make_params<BM_##V>},
    GOSS_BATCH_KINDS(GOSS_BATCH_ROW)
#undef GOSS_BATCH_ROW
};

// The node's whole time block with its population's per-node code.
template <class M, int SOLVER>
__device__ GOSS_INLINE void batch_node(const BatchPop& P, const char* __restrict__ params, const int tblock,
                                       const unsigned node)
{
  const SharedParamsOf<M, float>& sp = *reinterpret_cast<const SharedParamsOf<M, float>*>(params + P.sp_off);
  const double2* steps = P.steps + size_t(tblock) * P.nouter;
  const FieldOut fo{nullptr, 0, 0};
  float* y = static_cast<float*>(P.y);
  const float* fp = static_cast<const float*>(P.fp);
  if constexpr (SOLVER == kRKF32)
    rkf32_node<M, float>(node, y, fp, steps, P.nouter, sp, P.ap, P.B, P.N, P.ld, fo);
  else
    explicit_node<M, SOLVER, float, false>(node, y, nullptr, fp, steps, P.nouter, sp, P.ldt_0, float(P.delta), P.N,
                                           P.ld, fo);
}

// Persistent over chunks: block b takes chunks b, b + gridDim.x, ...  Chunk c
// is blockDim.x consecutive nodes of the population whose chunk range holds c.
// The population index and kind are uniform across the block, so the switch
// is a scalar branch and every wave runs one population's code.
__global__ void __launch_bounds__(kBlock)
batch_kernel(const BatchPop* __restrict__ pops, const int npop, const unsigned nchunks, const int tblock,
             const char* __restrict__ params)
{
#pragma unroll 1
  for (unsigned c = blockIdx.x; c < nchunks; c += gridDim.x) {
    int p = 0;
#pragma unroll 1
    while (p + 1 < npop && c >= pops[p + 1].chunk0) ++p;
    const BatchPop& P = pops[p];
    const unsigned node = (c - P.chunk0) * blockDim.x + threadIdx.x;
    if (node >= P.N) continue;
    switch (P.kind) {
#define GOSS_BATCH_CASE(V, SOLVER) \
  case kBatch_##V##_##SOLVER: batch_node<BM_##V, SOLVER>(P, params, tblock, node); break;
      GOSS_BATCH_KINDS(GOSS_BATCH_CASE)
#undef GOSS_BATCH_CASE
      default: break;
    }
  }
}

}  // namespace

const BatchKind* batch_kinds(int* count)
{
  *count = int(sizeof(kKinds) / sizeof(kKinds[0]));
  return kKinds;
}

hipError_t launch_batch(const BatchPop* pops, int npop, unsigned nchunks, int tblock, const char* params,
                        unsigned grid, unsigned block, hipStream_t stream)
{
  if (block == 0 || block > unsigned(kBlock) || block % 64 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(block), 0, stream, pops, npop, nchunks, tblock, params);
  return hipGetLastError();
}

}  // namespace goss_hip
