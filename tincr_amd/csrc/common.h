// Shared definitions for the gfx950 LC-TDA kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace tda {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Storage / arithmetic precisions of the bulk products.
enum Prec : int { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16 = 2, PREC_FP16 = 3 };
// Internal: FP16 hi / lo split operands (about FP32 accuracy from FP16 MFMA),
// the FP32-accuracy phase of the precision ladder for the AO operator.
constexpr int PREC_F16X = 4;

inline size_t prec_bytes(int p) { return p == PREC_FP64 ? 8 : p == PREC_FP32 ? 4 : 2; }

// FP16 hi / lo split of a value: hi = fp16(v), lo = fp16((v - hi) 2^11).
__device__ inline void split_f16(float v, _Float16& hi, _Float16& lo) {
  hi = (_Float16)v;
  lo = (_Float16)((v - (float)hi) * 2048.f);
}
__device__ inline float join_f16(_Float16 hi, _Float16 lo) { return (float)hi + (float)lo * 4.8828125e-4f; }

// Thrown by launchers / the driver; the C ABI turns it into a status code.
struct Error {
  int code;
  std::string msg;
};

#define TDA_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      throw ::tda::Error{4, std::string(#call) + ": " + hipGetErrorString(e_) +   \
                                " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
    }                                                                              \
  } while (0)

#define TDA_CHECK(cond, code, what)                                 \
  do {                                                              \
    if (!(cond)) throw ::tda::Error{(code), std::string(what)};     \
  } while (0)

#define TDA_LAUNCH_CHECK() TDA_HIP(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// A two-level index map: x -> (x / div) * s1 + (x % div) * s0.  With
// div >= extent it is the plain strided map x * s0.
struct IdxMap {
  int64_t div;
  int64_t s1;
  int64_t s0;
  __host__ __device__ inline int64_t operator()(int64_t x) const {
    if (div >= (INT64_C(1) << 62)) return x * s0;  // plain map: no division
    if ((uint64_t)(x | div) < (UINT64_C(1) << 32)) {
      // 32-bit quotient: a few instructions instead of the 64-bit division routine
      const uint32_t q = (uint32_t)x / (uint32_t)div;
      return (int64_t)q * s1 + (x - (int64_t)q * div) * s0;
    }
    return (x / div) * s1 + (x % div) * s0;
  }
};

inline IdxMap plain(int64_t stride) { return IdxMap{INT64_C(1) << 62, 0, stride}; }
inline IdxMap split(int64_t div, int64_t s1, int64_t s0) { return IdxMap{div, s1, s0}; }

// Wave-level sum over 64 lanes.
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ inline T to_type(double x);
template <>
__device__ inline double to_type<double>(double x) { return x; }
template <>
__device__ inline float to_type<float>(double x) { return (float)x; }
template <>
__device__ inline _Float16 to_type<_Float16>(double x) { return (_Float16)(float)x; }
template <>
__device__ inline __bf16 to_type<__bf16>(double x) { return (__bf16)(float)x; }

template <typename T>
__device__ inline double to_double(T x) { return (double)(float)x; }
template <>
__device__ inline double to_double<double>(double x) { return x; }

}  // namespace tda
