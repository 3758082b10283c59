// Products of 1-qubit gates per qubit: ua_fuse_1q_ops
// (unitair/gates/matrix_algebra.py:6-51, fuse_single_qubit_operators, which
// multiplies with torch.matmul one pair at a time).  One thread per (batch
// entry, qubit slot) walks the gate list once and keeps its 2x2 product in
// registers; the output is laid out (batch, slot, 2, 2), the operand layout of
// ua_apply_1q_layer, so apply_to_qubits needs no stacking copy in between.
#include "ua_common.h"

namespace ua {
namespace {

struct FuseParams {
  const void* ops[UA_FUSE_MAX_OPS];
  int slot[UA_FUSE_MAX_OPS];
  int num_ops, num_slots, accumulate;
  int64_t batch, op_bstride;
};

template <typename T> struct Alg;
template <> struct Alg<float> {
  __device__ static float one() { return 1.f; }
  __device__ static float mul(float a, float b) { return a * b; }
  __device__ static float mad(float acc, float a, float b) { return fmaf(a, b, acc); }
};
template <> struct Alg<double> {
  __device__ static double one() { return 1.0; }
  __device__ static double mul(double a, double b) { return a * b; }
  __device__ static double mad(double acc, double a, double b) { return fma(a, b, acc); }
};
template <> struct Alg<float2> {
  __device__ static float2 one() { return make_float2(1.f, 0.f); }
  __device__ static float2 mul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  __device__ static float2 mad(float2 acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
  }
};
template <> struct Alg<double2> {
  __device__ static double2 one() { return make_double2(1.0, 0.0); }
  __device__ static double2 mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  __device__ static double2 mad(double2 acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
  }
};

template <typename T>
__global__ __launch_bounds__(kBlock) void fuse_1q_kernel(T* out, FuseParams p) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= p.batch * p.num_slots) return;
  const int64_t b = idx / p.num_slots;
  const int s = int(idx % p.num_slots);
  T* m = out + 4 * idx;
  T m00, m01, m10, m11;
  if (p.accumulate) {
    m00 = m[0]; m01 = m[1]; m10 = m[2]; m11 = m[3];
  } else {
    m00 = Alg<T>::one(); m11 = Alg<T>::one(); m01 = T{}; m10 = T{};
  }
  for (int i = 0; i < p.num_ops; ++i) {
    if (p.slot[i] != s) continue;
    const T* u = static_cast<const T*>(p.ops[i]) + b * p.op_bstride;
    const T u00 = u[0], u01 = u[1], u10 = u[2], u11 = u[3];
    // M <- U @ M
    const T n00 = Alg<T>::mad(Alg<T>::mul(u00, m00), u01, m10);
    const T n01 = Alg<T>::mad(Alg<T>::mul(u00, m01), u01, m11);
    const T n10 = Alg<T>::mad(Alg<T>::mul(u10, m00), u11, m10);
    const T n11 = Alg<T>::mad(Alg<T>::mul(u10, m01), u11, m11);
    m00 = n00; m01 = n01; m10 = n10; m11 = n11;
  }
  m[0] = m00; m[1] = m01; m[2] = m10; m[3] = m11;
}

}  // namespace
}  // namespace ua

extern "C" int ua_fuse_1q_ops(void* out, const void* const* ops, int dtype, int num_ops,
                              const int* slot, int num_slots, int64_t batch,
                              int64_t op_batch_stride, int accumulate, ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(elem_size(dtype) != 0, "unsupported dtype code %d", dtype);
  UA_CHECK_ARG(num_ops >= 0 && num_ops <= UA_FUSE_MAX_OPS,
               "num_ops=%d: at most %d operators per call", num_ops, UA_FUSE_MAX_OPS);
  UA_CHECK_ARG(num_slots >= 0 && batch >= 0 && op_batch_stride >= 0, "negative size");
  if (batch == 0 || num_slots == 0) return UA_OK;
  UA_CHECK_ARG(out, "NULL output");
  UA_CHECK_ARG(num_ops == 0 || (ops && slot), "NULL operator list");
  FuseParams p{};
  for (int i = 0; i < num_ops; ++i) {
    UA_CHECK_ARG(ops[i], "operator %d: NULL pointer", i);
    UA_CHECK_ARG(slot[i] >= 0 && slot[i] < num_slots, "operator %d: slot %d out of range", i,
                 slot[i]);
    p.ops[i] = ops[i];
    p.slot[i] = slot[i];
  }
  p.num_ops = num_ops;
  p.num_slots = num_slots;
  p.accumulate = accumulate ? 1 : 0;
  p.batch = batch;
  p.op_bstride = op_batch_stride;
  const int64_t threads = batch * num_slots;
  UA_CHECK_ARG(threads / kBlock < 0x7fffffffLL, "too many matrices");
  const unsigned grid = grid_for(threads);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case UA_FLOAT32: fuse_1q_kernel<float><<<grid, kBlock, 0, s>>>(static_cast<float*>(out), p); break;
    case UA_FLOAT64: fuse_1q_kernel<double><<<grid, kBlock, 0, s>>>(static_cast<double*>(out), p); break;
    case UA_COMPLEX64: fuse_1q_kernel<float2><<<grid, kBlock, 0, s>>>(static_cast<float2*>(out), p); break;
    default: fuse_1q_kernel<double2><<<grid, kBlock, 0, s>>>(static_cast<double2*>(out), p); break;
  }
  return check_launch("fuse_1q_kernel");
}
