// Shared helpers for the DOGPUP MI355X library: error state, launch checks,
// complex arithmetic on (re, im) pairs, 16-byte vector packs.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/dogpup.h"

namespace dogpup {

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define DP_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::dogpup::fail(DOGPUP_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #call, \
                            hipGetErrorString(e_));                                   \
  } while (0)

#define DP_LAUNCH_CHECK() DP_HIP(hipGetLastError())

#define DP_ARG(cond, ...)                                                             \
  do {                                                                                \
    if (!(cond)) return ::dogpup::fail(DOGPUP_ERR_ARG, __VA_ARGS__);                  \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// -------- complex helpers (unconjugated algebra, as the reference uses) --------
template <typename T>
struct cpx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cpx<T> cdiv(cpx<T> a, cpx<T> b) {
  T d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// A 16-byte pack of complex values: two complex float or one complex double.
template <typename T>
struct Pack;
template <>
struct Pack<float> {
  static constexpr int N = 2;
  using V = float4;
  __device__ static inline void load(const float* p, float re[2], float im[2]) {
    float4 v = *reinterpret_cast<const float4*>(p);
    re[0] = v.x; im[0] = v.y; re[1] = v.z; im[1] = v.w;
  }
  __device__ static inline void store(float* p, const float re[2], const float im[2]) {
    *reinterpret_cast<float4*>(p) = make_float4(re[0], im[0], re[1], im[1]);
  }
};
template <>
struct Pack<double> {
  static constexpr int N = 1;
  using V = double2;
  __device__ static inline void load(const double* p, double re[1], double im[1]) {
    double2 v = *reinterpret_cast<const double2*>(p);
    re[0] = v.x; im[0] = v.y;
  }
  __device__ static inline void store(double* p, const double re[1], const double im[1]) {
    *reinterpret_cast<double2*>(p) = make_double2(re[0], im[0]);
  }
};

inline unsigned div_up(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace dogpup
