// Shared device/host helpers for the unitair_amd HIP library (gfx950 only).
//
// State vectors are unitair "vector layout" tensors (*batch, 2**n), stored
// contiguously, complex values interleaved (re, im).  Qubit q of an n-qubit
// state is index bit n-1-q (big-endian: qubit 0 is the most significant bit),
// following unitair/states/conversions.py:43-45 and
// unitair/simulation/operations.py:729-731.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/unitair_amd.h"

namespace ua {

// ---------------------------------------------------------------- errors --
void set_error(const char* fmt, ...);
const char* get_error();

#define UA_CHECK_ARG(cond, ...)                 \
  do {                                          \
    if (!(cond)) {                              \
      ::ua::set_error(__VA_ARGS__);             \
      return UA_ERR_INVALID_ARGUMENT;           \
    }                                           \
  } while (0)

// Checks the launch that was just issued; returns a status code.
int check_launch(const char* what);

// ------------------------------------------------------------ amplitudes --
// Arithmetic over the four element types a state can have.  Real states
// (float32/float64) are valid in unitair (tests/hypothesis_strategies/
// tensors.py:66-76 draw complex64 or float32).
template <typename T> struct Num;

template <> struct Num<float> {
  using real = float;
  static constexpr bool complex = false;
  __device__ __forceinline__ static float zero() { return 0.f; }
  __device__ __forceinline__ static float mac(float acc, float u, float a) { return fmaf(u, a, acc); }
};
template <> struct Num<double> {
  using real = double;
  static constexpr bool complex = false;
  __device__ __forceinline__ static double zero() { return 0.0; }
  __device__ __forceinline__ static double mac(double acc, double u, double a) { return fma(u, a, acc); }
};
template <> struct Num<float2> {
  using real = float;
  static constexpr bool complex = true;
  __device__ __forceinline__ static float2 zero() { return make_float2(0.f, 0.f); }
  __device__ __forceinline__ static float2 mac(float2 acc, float2 u, float2 a) {
    acc.x = fmaf(u.x, a.x, acc.x);
    acc.x = fmaf(-u.y, a.y, acc.x);
    acc.y = fmaf(u.x, a.y, acc.y);
    acc.y = fmaf(u.y, a.x, acc.y);
    return acc;
  }
};
template <> struct Num<double2> {
  using real = double;
  static constexpr bool complex = true;
  __device__ __forceinline__ static double2 zero() { return make_double2(0.0, 0.0); }
  __device__ __forceinline__ static double2 mac(double2 acc, double2 u, double2 a) {
    acc.x = fma(u.x, a.x, acc.x);
    acc.x = fma(-u.y, a.y, acc.x);
    acc.y = fma(u.x, a.y, acc.y);
    acc.y = fma(u.y, a.x, acc.y);
    return acc;
  }
};

// Two adjacent elements moved as one vector access (16 B for complex64).
template <typename T> struct alignas(2 * sizeof(T)) Pair { T v[2]; };

inline size_t elem_size(int dtype) {
  switch (dtype) {
    case UA_FLOAT32: return 4;
    case UA_FLOAT64: return 8;
    case UA_COMPLEX64: return 8;
    case UA_COMPLEX128: return 16;
    default: return 0;
  }
}

// Inserts a zero bit at each of the ascending bit positions pos[0..k-1].
__device__ __forceinline__ uint64_t insert_zero_bits(uint64_t x, const int* pos, int k) {
  for (int p = 0; p < k; ++p) {
    const uint64_t low = x & ((uint64_t(1) << pos[p]) - 1);
    x = ((x >> pos[p]) << (pos[p] + 1)) | low;
  }
  return x;
}

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t threads, int block = kBlock) {
  return static_cast<unsigned>((threads + block - 1) / block);
}

}  // namespace ua
