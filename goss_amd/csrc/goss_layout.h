// Host-callable launchers of the layout kernels (goss_layout.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace goss_hip {

constexpr int kMaxStates = 128;
constexpr int kMaxFieldParams = 64;

// Per-variant index maps, passed by value as kernel arguments.
struct IndexMap {
  int S, F, NFP;
  int fsi[kMaxStates];   // field state -> state index
  int comp[kMaxStates];  // state -> compensated slot or -1
};

struct ComponentList {
  int n;
  int c[kMaxStates];
};

struct DefaultValues {
  double ic[kMaxStates];
  double fp[kMaxFieldParams];
};

enum class LayoutOp { kGet, kSet };

// buf is a device buffer of doubles in the caller's layout.  ylo is null
// unless the precision carries compensated low parts.
template <class R>
hipError_t layout_states(LayoutOp op, void* y, void* ylo, double* buf, const IndexMap& m, unsigned N, unsigned ld,
                         hipStream_t s);
template <class R>
hipError_t layout_field_states(LayoutOp op, void* y, void* ylo, double* buf, const IndexMap& m,
                               const ComponentList& cl, unsigned N, unsigned ld, int tangled, hipStream_t s);
// kGet: field states -> values[dofs] (scatter); kSet: values[dofs] -> field
// states (gather).  dofs [F][N], device memory.
template <class R>
hipError_t layout_field_dofs(LayoutOp op, void* y, void* ylo, double* values, const int* dofs, const IndexMap& m,
                             unsigned N, unsigned ld, hipStream_t s);
template <class R>
hipError_t layout_field_params(LayoutOp op, void* fp, double* buf, unsigned NFP, unsigned N, unsigned ld,
                               int tangled, hipStream_t s);
template <class R>
hipError_t layout_fill_defaults(void* y, void* ylo, void* fp, const IndexMap& m, const DefaultValues& d, unsigned N,
                                unsigned ld, hipStream_t s);

hipError_t layout_init_implicit(double* ds, int* is, unsigned N, unsigned ld, double eta_0, int always,
                                hipStream_t s);

}  // namespace goss_hip
