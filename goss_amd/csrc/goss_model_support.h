// Math helpers behind the generated model code (goss_amd/csrc/generated/*.h).
//
// The generated expressions call m_exp, m_log, m_sq, ... so that one text
// serves the CPU model class (g++, double, glibc) and the HIP kernels
// (hipcc, float or double, ROCm device libraries).
//
// Device float: division, reciprocal, exp, log and sqrt use the hardware
// approximations (v_rcp_f32, v_exp_f32, v_log_f32, v_sqrt_f32) unless
// GOSS_F32_ACCURATE_MATH is defined.  Double always uses IEEE division and
// the ocml/glibc routines, the reference's arithmetic.
#pragma once

#include <cmath>
#include <type_traits>

#if defined(GOSS_PARITY_MATH)
#include "goss_parity_math.h"
#endif

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GOSS_HD __host__ __device__
#define GOSS_INLINE __attribute__((always_inline)) inline
#else
#define GOSS_HD
#define GOSS_INLINE inline
#endif

// Piecewise expressions print as GOSS_SELECT(cond, a, b).  In device code both
// arms are evaluated and one is selected, so a model evaluation has no
// lane-divergent branch.  With an if/else over a lane-dependent condition
// (V < -40 in the TP06 gates), hipcc (ROCm 7.2, gfx950) has placed register
// copies in the join block ahead of the exec-mask restore, so lanes from the
// other arm kept stale values: a state in an FP64 kernel, a per-node address
// in an FP32 one.  Host code keeps the ?: form.
template <class A, class B>
GOSS_HD GOSS_INLINE typename std::common_type<A, B>::type m_select(const bool c, const A a, const B b)
{
  return c ? a : b;
}
#if defined(__HIP_DEVICE_COMPILE__)
#define GOSS_SELECT(c, a, b) m_select((c), (a), (b))
#else
#define GOSS_SELECT(c, a, b) ((c) ? (a) : (b))
#endif

template <typename T> GOSS_HD GOSS_INLINE T m_sq(const T x) { return x * x; }
template <typename T> GOSS_HD GOSS_INLINE T m_cube(const T x) { return x * x * x; }
template <typename T> GOSS_HD GOSS_INLINE T m_sign(const T x) { return T((x > T(0)) - (x < T(0))); }
template <typename T> GOSS_HD GOSS_INLINE T m_heaviside(const T x) { return x > T(0) ? T(1) : (x < T(0) ? T(0) : T(0.5)); }

GOSS_HD GOSS_INLINE double m_div(const double a, const double b) { return a / b; }
GOSS_HD GOSS_INLINE double m_rcp(const double b) { return 1.0 / b; }
#if defined(GOSS_PARITY_MATH)
// Parity build (tests only): exp, log, pow and sqrt that round identically on
// the host and the GPU (goss_parity_math.h).
GOSS_HD GOSS_INLINE double m_exp(const double x) { return goss_pm::exp(x); }
GOSS_HD GOSS_INLINE double m_exp_l2(const double x2, const double x) { (void)x2; return goss_pm::exp(x); }
GOSS_HD GOSS_INLINE double m_log(const double x) { return goss_pm::log(x); }
GOSS_HD GOSS_INLINE double m_sqrt(const double x) { return goss_pm::sqrt(x); }
GOSS_HD GOSS_INLINE double m_pow(const double x, const double y) { return goss_pm::pow(x, y); }
#else
GOSS_HD GOSS_INLINE double m_exp(const double x) { return std::exp(x); }
// m_exp_l2(x * log2(e), x): the FP32 device kernels use the first argument (see below); double uses x.
GOSS_HD GOSS_INLINE double m_exp_l2(const double x2, const double x) { (void)x2; return std::exp(x); }
GOSS_HD GOSS_INLINE double m_log(const double x) { return std::log(x); }
GOSS_HD GOSS_INLINE double m_sqrt(const double x) { return std::sqrt(x); }
GOSS_HD GOSS_INLINE double m_pow(const double x, const double y) { return std::pow(x, y); }
#endif
GOSS_HD GOSS_INLINE double m_fabs(const double x) { return std::fabs(x); }
GOSS_HD GOSS_INLINE double m_floor(const double x) { return std::floor(x); }
GOSS_HD GOSS_INLINE double m_ceil(const double x) { return std::ceil(x); }
GOSS_HD GOSS_INLINE double m_fmax(const double x, const double y) { return std::fmax(x, y); }
GOSS_HD GOSS_INLINE double m_fmin(const double x, const double y) { return std::fmin(x, y); }
GOSS_HD GOSS_INLINE double m_sin(const double x) { return std::sin(x); }
GOSS_HD GOSS_INLINE double m_cos(const double x) { return std::cos(x); }
GOSS_HD GOSS_INLINE double m_tan(const double x) { return std::tan(x); }
GOSS_HD GOSS_INLINE double m_asin(const double x) { return std::asin(x); }
GOSS_HD GOSS_INLINE double m_acos(const double x) { return std::acos(x); }
GOSS_HD GOSS_INLINE double m_atan(const double x) { return std::atan(x); }
GOSS_HD GOSS_INLINE double m_sinh(const double x) { return std::sinh(x); }
GOSS_HD GOSS_INLINE double m_cosh(const double x) { return std::cosh(x); }
GOSS_HD GOSS_INLINE double m_tanh(const double x) { return std::tanh(x); }

#if defined(__HIP_DEVICE_COMPILE__) && !defined(GOSS_F32_ACCURATE_MATH)
// FP32 device kernels: hardware reciprocal (v_rcp_f32, 1 ulp), exp2/log2
// (v_exp_f32 / v_log_f32 with a scale) and square root (v_sqrt_f32).
GOSS_HD GOSS_INLINE float m_div(const float a, const float b) { return a * __builtin_amdgcn_rcpf(b); }
GOSS_HD GOSS_INLINE float m_rcp(const float b) { return __builtin_amdgcn_rcpf(b); }
GOSS_HD GOSS_INLINE float m_exp(const float x) { return __expf(x); }
GOSS_HD GOSS_INLINE float m_log(const float x) { return __logf(x); }
GOSS_HD GOSS_INLINE float m_sqrt(const float x) { return __builtin_amdgcn_sqrtf(x); }
// exp(x) given x2 = x * log2(e) with the constant folded into x's coefficients by the code
// generator: one v_exp_f32, no scaling multiply.
GOSS_HD GOSS_INLINE float m_exp_l2(const float x2, const float x) { (void)x; return __builtin_amdgcn_exp2f(x2); }
#else
GOSS_HD GOSS_INLINE float m_exp_l2(const float x2, const float x) { (void)x2; return std::exp(x); }
GOSS_HD GOSS_INLINE float m_div(const float a, const float b) { return a / b; }
GOSS_HD GOSS_INLINE float m_rcp(const float b) { return 1.0f / b; }
GOSS_HD GOSS_INLINE float m_exp(const float x) { return std::exp(x); }
GOSS_HD GOSS_INLINE float m_log(const float x) { return std::log(x); }
GOSS_HD GOSS_INLINE float m_sqrt(const float x) { return std::sqrt(x); }
#endif
GOSS_HD GOSS_INLINE float m_fabs(const float x) { return std::fabs(x); }
GOSS_HD GOSS_INLINE float m_floor(const float x) { return std::floor(x); }
GOSS_HD GOSS_INLINE float m_ceil(const float x) { return std::ceil(x); }
GOSS_HD GOSS_INLINE float m_pow(const float x, const float y) { return std::pow(x, y); }
GOSS_HD GOSS_INLINE float m_fmax(const float x, const float y) { return std::fmax(x, y); }
GOSS_HD GOSS_INLINE float m_fmin(const float x, const float y) { return std::fmin(x, y); }
GOSS_HD GOSS_INLINE float m_sin(const float x) { return std::sin(x); }
GOSS_HD GOSS_INLINE float m_cos(const float x) { return std::cos(x); }
GOSS_HD GOSS_INLINE float m_tan(const float x) { return std::tan(x); }
GOSS_HD GOSS_INLINE float m_asin(const float x) { return std::asin(x); }
GOSS_HD GOSS_INLINE float m_acos(const float x) { return std::acos(x); }
GOSS_HD GOSS_INLINE float m_atan(const float x) { return std::atan(x); }
GOSS_HD GOSS_INLINE float m_sinh(const float x) { return std::sinh(x); }
GOSS_HD GOSS_INLINE float m_cosh(const float x) { return std::cosh(x); }
GOSS_HD GOSS_INLINE float m_tanh(const float x) { return std::tanh(x); }

// phi1(x) = (exp(x) - 1) / x and its derivative phi1'(x) = (x exp(x) - (exp(x) - 1)) / x^2,
// used by the GRL1 adjoint (a GRL1 increment is rhs * h * phi1(linear * h)).  Near 0 both are
// evaluated from their Taylor series (phi1 = sum x^k/(k+1)!, phi1' = sum k x^(k-1)/(k+1)!),
// where the closed forms cancel; the switch points keep either form's relative error below
// 1e-12 (double) and 2e-6 (float).  Both forms are evaluated and one selected (no lane-divergent
// branch; the seeds below do the same).
template <typename T>
GOSS_HD GOSS_INLINE T m_phi1(const T x)
{
  const T small = sizeof(T) == 8 ? T(1e-3) : T(1e-1);
  const T series = T(1) + x * (T(1.0 / 2) + x * (T(1.0 / 6) + x * (T(1.0 / 24) + x * (T(1.0 / 120) + x * T(1.0 / 720)))));
  const T closed = std::expm1(x) / x;
  return m_select(m_fabs(x) < small, series, closed);
}
template <typename T>
GOSS_HD GOSS_INLINE T m_dphi1(const T x)
{
  const T small = sizeof(T) == 8 ? T(1e-3) : T(1e-1);
  const T series = T(1.0 / 2) + x * (T(1.0 / 3) + x * (T(1.0 / 8) + x * (T(1.0 / 30) + x * (T(1.0 / 144) + x * T(1.0 / 840)))));
  const T closed = (x * std::exp(x) - std::expm1(x)) / (x * x);
  return m_select(m_fabs(x) < small, series, closed);
}
// Seeds of the GRL1 step adjoint for one state: with x = linear * h and lh = lam * h,
//   bbar = lh * phi1(x),  abar = lh * h * rhs * phi1'(x)   when |linear| > delta,
//   bbar = lh,            abar = 0                          otherwise (the step is y + rhs * h).
// inv_linear = 1/linear and hinv = 1/h are passed in (a gate's 1/linear is -tau, with no
// reciprocal to take); this generic form evaluates phi1 and phi1' directly.
template <typename T>
GOSS_HD GOSS_INLINE void m_grl1_seeds(const T linear, const T inv_linear, const T rhs, const T lam, const T h,
                                      const T hinv, const T delta, T& bbar, T& abar)
{
  (void)inv_linear; (void)hinv;
  const bool ex = m_fabs(linear) > delta;
  const T lh = lam * h;
  const T b_ex = lh * m_phi1(linear * h);
  const T a_ex = lh * rhs * h * m_dphi1(linear * h);
  bbar = m_select(ex, b_ex, lh);
  abar = m_select(ex, a_ex, T(0));
}
// Seeds of the GRL1 step adjoint for a gate, whose step is y' = A + (y - A) exp(x), x = linear * h
// (y + (linear (y - A)) h when |linear| <= delta):
//   ey = lam dy'/dy,  ga = lam dy'/dA,  gc = lam dy'/dlinear.
// exp(x) - 1 comes from expm1, so 1 - exp(x) keeps its relative accuracy for slow gates.
template <typename T>
GOSS_HD GOSS_INLINE void m_grl1_gate_seeds(const T linear, const T y, const T a, const T lam, const T h,
                                           const T delta, T& ey, T& ga, T& gc)
{
  const bool ex = m_fabs(linear) > delta;
  const T x = linear * h;
  const T e1 = T(std::expm1(x));
  const T em1 = m_select(ex, e1, x);
  ey = lam * (T(1) + em1);
  ga = -lam * em1;
  gc = lam * h * (y - a) * m_select(ex, T(1) + em1, T(1));
}
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GOSS_F32_ACCURATE_MATH)
// FP32 device kernels: beyond |x| = 0.1 the closed forms use the hardware exp and reciprocal
// (exp(x) - 1 then loses at most ~1e-6 relative to cancellation).
template <>
GOSS_HD GOSS_INLINE float m_phi1<float>(const float x)
{
  const float p = 1.0f + x * (0.5f + x * (1.0f / 6 + x * (1.0f / 24 + x * (1.0f / 120 + x * (1.0f / 720)))));
  const float q = (__expf(x) - 1.0f) * __builtin_amdgcn_rcpf(x);
  return fabsf(x) < 0.1f ? p : q;
}
template <>
GOSS_HD GOSS_INLINE float m_dphi1<float>(const float x)
{
  const float p = 0.5f + x * (1.0f / 3 + x * (0.125f + x * (1.0f / 30 + x * (1.0f / 144 + x * (1.0f / 840)))));
  const float e = __expf(x);
  const float q = (x * e - (e - 1.0f)) * __builtin_amdgcn_rcpf(x * x);
  return fabsf(x) < 0.1f ? p : q;
}
// exp(x) - 1: Taylor polynomial to x^4 below |x| = 1/64, hardware exp beyond.
template <>
GOSS_HD GOSS_INLINE void m_grl1_gate_seeds<float>(const float linear, const float y, const float a, const float lam,
                                                  const float h, const float delta, float& ey, float& ga, float& gc)
{
  const bool ex = fabsf(linear) > delta;
  const float x = linear * h;
  const float p = x * (1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f))));
  const float q = __expf(x) - 1.0f;
  const float em1 = m_select(ex, m_select(fabsf(x) < 0.015625f, p, q), x);
  ey = lam * (1.0f + em1);
  ga = -lam * em1;
  gc = lam * h * (y - a) * m_select(ex, 1.0f + em1, 1.0f);
}
// One exponential for both, 1/x = inv_linear * hinv: phi1 = (e - 1) / x, phi1' = (e - phi1) / x.
// Below |x| = 0.05 the series (truncation below 1e-9 relative for phi1, 1e-7 for phi1'); above it
// the closed forms lose at most ~2.4e-6 relative to the cancellation in e - 1.
template <>
GOSS_HD GOSS_INLINE void m_grl1_seeds<float>(const float linear, const float inv_linear, const float rhs,
                                             const float lam, const float h, const float hinv, const float delta,
                                             float& bbar, float& abar)
{
  const float x = linear * h;
  const float e = __expf(x);
  const float r = inv_linear * hinv;
  const float pc = (e - 1.0f) * r;
  const float dpc = (e - pc) * r;
  const float ps = 1.0f + x * (0.5f + x * (1.0f / 6 + x * (1.0f / 24 + x * (1.0f / 120))));
  const float dps = 0.5f + x * (1.0f / 3 + x * (0.125f + x * (1.0f / 30)));
  const bool small = fabsf(x) < 0.05f;
  const bool ex = fabsf(linear) > delta;
  const float lh = lam * h;
  const float b_ex = lh * m_select(small, ps, pc);
  const float a_ex = lh * rhs * h * m_select(small, dps, dpc);
  bbar = m_select(ex, b_ex, lh);
  abar = m_select(ex, a_ex, 0.0f);
}
#endif
