// Layout kernels: conversions between the caller's FP64 arrays (node-major
// states, tangled/untangled field arrays of ODESystemSolver.h:238-312) and the
// device structure-of-arrays buffers of each precision.
#include "goss_layout.h"
#include "goss_implicit.h"

namespace goss_hip {

namespace {

// Value of state i of node n (hi + lo for compensated FP32 states).
template <class R>
__device__ inline double load_state(const R* y, const R* ylo, const IndexMap& m, unsigned ld, int i, unsigned n)
{
  double v = double(y[size_t(i) * ld + n]);
  if (ylo && m.comp[i] >= 0) v += double(ylo[size_t(m.comp[i]) * ld + n]);
  return v;
}

template <class R>
__device__ inline void store_state(R* y, R* ylo, const IndexMap& m, unsigned ld, int i, unsigned n, double v)
{
  const R hi = R(v);
  y[size_t(i) * ld + n] = hi;
  if (ylo && m.comp[i] >= 0) ylo[size_t(m.comp[i]) * ld + n] = R(v - double(hi));
}

// Node-major [N][S] <-> SoA.  One thread per (node, state).
template <class R>
__global__ void states_to_aos(const R* __restrict__ y, const R* __restrict__ ylo, double* __restrict__ out,
                              const IndexMap m, unsigned N, unsigned ld)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * m.S) return;
  const unsigned n = unsigned(k / m.S);
  const int i = int(k % m.S);
  out[k] = load_state(y, ylo, m, ld, i, n);
}

template <class R>
__global__ void states_from_aos(R* __restrict__ y, R* __restrict__ ylo, const double* __restrict__ in,
                                const IndexMap m, unsigned N, unsigned ld)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * m.S) return;
  const unsigned n = unsigned(k / m.S);
  const int i = int(k % m.S);
  store_state(y, ylo, m, ld, i, n, in[k]);
}

// Field states: components c of the selected list, index formula of
// ODESystemSolver.h:244.  One thread per (node, component).
template <class R>
__global__ void field_states_get(const R* __restrict__ y, const R* __restrict__ ylo, double* __restrict__ out,
                                 const IndexMap m, const ComponentList cl, unsigned N, unsigned ld, int tangled)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * cl.n) return;
  const unsigned n = unsigned(k / cl.n);
  const int c = cl.c[k % cl.n];
  const size_t idx = tangled ? size_t(n) * m.F + c : size_t(N) * c + n;
  out[idx] = load_state(y, ylo, m, ld, m.fsi[c], n);
}

template <class R>
__global__ void field_states_set(R* __restrict__ y, R* __restrict__ ylo, const double* __restrict__ in,
                                 const IndexMap m, const ComponentList cl, unsigned N, unsigned ld, int tangled)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * cl.n) return;
  const unsigned n = unsigned(k / cl.n);
  const int c = cl.c[k % cl.n];
  const size_t idx = tangled ? size_t(n) * m.F + c : size_t(N) * c + n;
  store_state(y, ylo, m, ld, m.fsi[c], n, in[idx]);
}

// Field states <-> a field vector through a dof map: dofs[c * N + n] is the
// index in `values` of field state c of node n (validated on the host: in
// range and without repeats).  One thread per (component, node), so the dof
// map and the state rows are read and written contiguously.
template <class R>
__global__ void field_states_gather(R* __restrict__ y, R* __restrict__ ylo, const double* __restrict__ values,
                                    const int* __restrict__ dofs, const IndexMap m, unsigned N, unsigned ld)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * m.F) return;
  const int c = int(k / N);
  const unsigned n = unsigned(k % N);
  store_state(y, ylo, m, ld, m.fsi[c], n, values[dofs[k]]);
}

template <class R>
__global__ void field_states_scatter(const R* __restrict__ y, const R* __restrict__ ylo, double* __restrict__ values,
                                     const int* __restrict__ dofs, const IndexMap m, unsigned N, unsigned ld)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * m.F) return;
  const int c = int(k / N);
  const unsigned n = unsigned(k % N);
  values[dofs[k]] = load_state(y, ylo, m, ld, m.fsi[c], n);
}

// Field parameters [NFP][ld] <-> caller layout.
template <class R>
__global__ void field_params_get(const R* __restrict__ fp, double* __restrict__ out, unsigned NFP, unsigned N,
                                 unsigned ld, int tangled)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * NFP) return;
  const unsigned n = unsigned(k / NFP);
  const unsigned j = unsigned(k % NFP);
  out[tangled ? size_t(n) * NFP + j : size_t(N) * j + n] = double(fp[size_t(j) * ld + n]);
}

template <class R>
__global__ void field_params_set(R* __restrict__ fp, const double* __restrict__ in, unsigned NFP, unsigned N,
                                 unsigned ld, int tangled)
{
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= size_t(N) * NFP) return;
  const unsigned n = unsigned(k / NFP);
  const unsigned j = unsigned(k % NFP);
  fp[size_t(j) * ld + n] = R(in[tangled ? size_t(n) * NFP + j : size_t(N) * j + n]);
}

// reset_default (ODESystemSolver.cpp:192-216): every node gets the initial
// conditions and the default values of the field parameters.
template <class R>
__global__ void fill_defaults(R* __restrict__ y, R* __restrict__ ylo, R* __restrict__ fp, const IndexMap m,
                              const DefaultValues d, unsigned N, unsigned ld)
{
  const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  for (int i = 0; i < m.S; ++i) store_state(y, ylo, m, ld, i, n, d.ic[i]);
  for (int j = 0; j < m.NFP; ++j) fp[size_t(j) * ld + n] = R(d.fp[j]);
}

// Per-node implicit solver state at construction: ImplicitODESolver::reset
// (ImplicitODESolver.cpp:80-90: recompute Jacobian, eta = eta_0), the
// ESDIRK23a constructor values (ESDIRK23a.cpp:27-44: eta 1, loc_error 1e9,
// prev_dt 0.1, first_step, new_jacobian) and the adaptive controllers' state.
__global__ void init_implicit_state(double* __restrict__ ds, int* __restrict__ is, unsigned N, unsigned ld,
                                    double eta_0, int always)
{
  const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  ds[size_t(kEta) * ld + n] = eta_0;
  ds[size_t(kEsEta) * ld + n] = 1.0;
  ds[size_t(kLocError) * ld + n] = 1e9;
  ds[size_t(kPrevDt) * ld + n] = 0.1;
  // AdaptiveExplicitSolver constructor and reset (AdaptiveExplicitSolver.cpp:28-31, 58-63).
  ds[size_t(kRkLdt) * ld + n] = 0.1;
  ds[size_t(kFacmax) * ld + n] = 2.0;
  // ESDIRK4O32::reset and AdaptiveImplicitSolver::reset (ESDIRK4O32.cpp:84-94,
  // AdaptiveImplicitSolver.cpp:43-66): no step guess yet, facmax 2, no previous error.
  ds[size_t(kAiLdt) * ld + n] = -1.0;
  ds[size_t(kAiFacmax) * ld + n] = 2.0;
  ds[size_t(kErrOld) * ld + n] = -1.0;
  ds[size_t(kDtOld) * ld + n] = 0.0;
  for (int k = 0; k < kNumIState; ++k) is[size_t(k) * ld + n] = 0;
  is[size_t(kRecompute) * ld + n] = 1;
  is[size_t(kAlways) * ld + n] = always;
  is[size_t(kFirstStep) * ld + n] = 1;
  is[size_t(kNewJacobian) * ld + n] = 1;
}

inline dim3 grid_for(size_t work) { return dim3(unsigned((work + 255) / 256)); }

}  // namespace

hipError_t layout_init_implicit(double* ds, int* is, unsigned N, unsigned ld, double eta_0, int always, hipStream_t s)
{
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(init_implicit_state, grid_for(N), dim3(256), 0, s, ds, is, N, ld, eta_0, always);
  return hipGetLastError();
}

template <class R>
hipError_t layout_states(LayoutOp op, void* y, void* ylo, double* buf, const IndexMap& m, unsigned N, unsigned ld,
                         hipStream_t s)
{
  const size_t work = size_t(N) * m.S;
  if (work == 0) return hipSuccess;
  if (op == LayoutOp::kGet)
    hipLaunchKernelGGL(states_to_aos<R>, grid_for(work), dim3(256), 0, s, static_cast<const R*>(y),
                       static_cast<const R*>(ylo), buf, m, N, ld);
  else
    hipLaunchKernelGGL(states_from_aos<R>, grid_for(work), dim3(256), 0, s, static_cast<R*>(y),
                       static_cast<R*>(ylo), buf, m, N, ld);
  return hipGetLastError();
}

template <class R>
hipError_t layout_field_states(LayoutOp op, void* y, void* ylo, double* buf, const IndexMap& m,
                               const ComponentList& cl, unsigned N, unsigned ld, int tangled, hipStream_t s)
{
  const size_t work = size_t(N) * cl.n;
  if (work == 0) return hipSuccess;
  if (op == LayoutOp::kGet)
    hipLaunchKernelGGL(field_states_get<R>, grid_for(work), dim3(256), 0, s, static_cast<const R*>(y),
                       static_cast<const R*>(ylo), buf, m, cl, N, ld, tangled);
  else
    hipLaunchKernelGGL(field_states_set<R>, grid_for(work), dim3(256), 0, s, static_cast<R*>(y),
                       static_cast<R*>(ylo), buf, m, cl, N, ld, tangled);
  return hipGetLastError();
}

template <class R>
hipError_t layout_field_dofs(LayoutOp op, void* y, void* ylo, double* values, const int* dofs, const IndexMap& m,
                             unsigned N, unsigned ld, hipStream_t s)
{
  const size_t work = size_t(N) * m.F;
  if (work == 0) return hipSuccess;
  if (op == LayoutOp::kGet)
    hipLaunchKernelGGL(field_states_scatter<R>, grid_for(work), dim3(256), 0, s, static_cast<const R*>(y),
                       static_cast<const R*>(ylo), values, dofs, m, N, ld);
  else
    hipLaunchKernelGGL(field_states_gather<R>, grid_for(work), dim3(256), 0, s, static_cast<R*>(y),
                       static_cast<R*>(ylo), values, dofs, m, N, ld);
  return hipGetLastError();
}

template <class R>
hipError_t layout_field_params(LayoutOp op, void* fp, double* buf, unsigned NFP, unsigned N, unsigned ld,
                               int tangled, hipStream_t s)
{
  const size_t work = size_t(N) * NFP;
  if (work == 0) return hipSuccess;
  if (op == LayoutOp::kGet)
    hipLaunchKernelGGL(field_params_get<R>, grid_for(work), dim3(256), 0, s, static_cast<const R*>(fp), buf, NFP,
                       N, ld, tangled);
  else
    hipLaunchKernelGGL(field_params_set<R>, grid_for(work), dim3(256), 0, s, static_cast<R*>(fp), buf, NFP, N,
                       ld, tangled);
  return hipGetLastError();
}

template <class R>
hipError_t layout_fill_defaults(void* y, void* ylo, void* fp, const IndexMap& m, const DefaultValues& d, unsigned N,
                                unsigned ld, hipStream_t s)
{
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_defaults<R>, grid_for(N), dim3(256), 0, s, static_cast<R*>(y), static_cast<R*>(ylo),
                     static_cast<R*>(fp), m, d, N, ld);
  return hipGetLastError();
}

#define GOSS_LAYOUT_INST(R)                                                                                     \
  template hipError_t layout_states<R>(LayoutOp, void*, void*, double*, const IndexMap&, unsigned, unsigned,    \
                                       hipStream_t);                                                            \
  template hipError_t layout_field_states<R>(LayoutOp, void*, void*, double*, const IndexMap&,                  \
                                             const ComponentList&, unsigned, unsigned, int, hipStream_t);       \
  template hipError_t layout_field_params<R>(LayoutOp, void*, double*, unsigned, unsigned, unsigned, int,       \
                                             hipStream_t);                                                      \
  template hipError_t layout_field_dofs<R>(LayoutOp, void*, void*, double*, const int*, const IndexMap&,        \
                                           unsigned, unsigned, hipStream_t);                                    \
  template hipError_t layout_fill_defaults<R>(void*, void*, void*, const IndexMap&, const DefaultValues&,       \
                                              unsigned, unsigned, hipStream_t);
GOSS_LAYOUT_INST(float)
GOSS_LAYOUT_INST(double)

}  // namespace goss_hip
