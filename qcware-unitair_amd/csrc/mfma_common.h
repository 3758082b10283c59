// Matrix-core helpers shared by the dense-gate kernels.
//
// A complex k-qubit gate (D = 2^k) acting on a group of D amplitudes is, in
// real arithmetic, the 2D x 2D block
//     [ Re U  -Im U ]   acting on the columns   [ Re x ]
//     [ Im U   Re U ]                            [ Im x ]
// so a batch of groups is a real GEMM that v_mfma_f32_16x16x4f32 (complex64)
// and v_mfma_f64_16x16x4f64 (complex128) run.  Lane layouts of the two
// instructions, measured on gfx950:
//   A: lane l holds A[l%16][l/16]; B: lane l holds B[l/16][l%16];
//   result register r: f32 -> row 4*(l/16)+r, f64 -> row l/16+4r; col l%16.
#pragma once
#include "ua_common.h"

namespace ua {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T> struct RealOf;
template <> struct RealOf<float2> { using type = float; };
template <> struct RealOf<double2> { using type = double; };

template <typename R> struct Mfma;
template <> struct Mfma<float> {
  using acc_t = floatx4;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct Mfma<double> {
  using acc_t = doublex4;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// Entry (row, col) of the real 2D x 2D block of U, where u = U[row % D][col % D].
template <typename T>
__device__ __forceinline__ typename RealOf<T>::type real_block(T u, bool im_row, bool im_col) {
  return (!im_row && !im_col) ? u.x : (!im_row ? -u.y : (!im_col ? u.y : u.x));
}

}  // namespace ua
