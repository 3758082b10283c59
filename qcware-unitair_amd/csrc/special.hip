// In-place CZ and CNOT for circuit execution.
//
// In unitair these gates are dense 4x4 matrices (unitair/gates/gates.py:
// 170-213) applied through apply_operator, or the CZ-only multi_cz
// (unitair/simulation/operations.py:657-748).  Both are sign flips or swaps:
// run in place they touch only the amplitudes they change, a quarter (CZ) or
// a half (CNOT) of the state, and their results are bit-identical to the
// dense application (a multiply by 0/1/-1 and exact sums).  When one of the
// two qubits is index bit 0, the kernels move whole element pairs so that
// every access is a full 16-byte vector.
#include "ua_common.h"

namespace ua {
namespace {

struct SpecialParams {
  int64_t total;      // threads that own work
  int64_t bstride;    // elements between states
  int log2_tps;       // log2(threads per state)
  int lo, hi;         // ascending bit positions of the two qubits
  uint64_t set_bits;  // bits forced to 1 in every visited index
  uint64_t flip;      // CNOT: target bit (0 for CZ)
};

__device__ __forceinline__ float2 negate(float2 v) { return make_float2(-v.x, -v.y); }
__device__ __forceinline__ double2 negate(double2 v) { return make_double2(-v.x, -v.y); }
__device__ __forceinline__ float negate(float v) { return -v; }
__device__ __forceinline__ double negate(double v) { return -v; }

// One thread per G consecutive indices with both (CZ) or the control (CNOT)
// bit set.  VEC (G = 2) when bit 0 is not one of the two qubits.
template <typename T, bool CNOT, int G>
__global__ __launch_bounds__(kBlock) void special_kernel(T* state, SpecialParams p) {
  const int64_t tid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (tid >= p.total) return;
  const int64_t b = tid >> p.log2_tps;
  const uint64_t t = uint64_t(tid) & ((uint64_t(1) << p.log2_tps) - 1);
  const int pos[2] = {p.lo, p.hi};
  const uint64_t i = insert_zero_bits(t * G, pos, 2) | p.set_bits;
  T* s = state + b * p.bstride;
  if constexpr (G == 2) {
    Pair<T>* a = reinterpret_cast<Pair<T>*>(s + i);
    if constexpr (CNOT) {
      Pair<T>* c = reinterpret_cast<Pair<T>*>(s + (i | p.flip));
      const Pair<T> x = *a, y = *c;
      *a = y;
      *c = x;
    } else {
      Pair<T> x = *a;
      x.v[0] = negate(x.v[0]);
      x.v[1] = negate(x.v[1]);
      *a = x;
    }
  } else {
    if constexpr (CNOT) {
      const T x = s[i], y = s[i | p.flip];
      s[i] = y;
      s[i | p.flip] = x;
    } else {
      s[i] = negate(s[i]);
    }
  }
}

// Gates that involve index bit 0 (qubit n-1) as whole 16-byte pairs
// (x[2i], x[2i+1]): every access is one contiguous vector, so only full
// lines move.  One thread per pair with bit `hi` as MODE needs:
//   kPairCZ       bit hi set: negate the odd element;
//   kPairCnotT0   CNOT with target bit 0, control hi set: swap the halves;
//   kPairCnotC0   CNOT with control bit 0, target hi: pairs i (hi clear) and
//                 i | 2^hi exchange their odd elements.
enum { kPairCZ = 0, kPairCnotT0 = 1, kPairCnotC0 = 2 };

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void special_pair_kernel(T* state, SpecialParams p) {
  const int64_t tid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (tid >= p.total) return;
  const int64_t b = tid >> p.log2_tps;
  const uint64_t t = uint64_t(tid) & ((uint64_t(1) << p.log2_tps) - 1);
  const int pos[2] = {0, p.hi};
  const uint64_t i = insert_zero_bits(t, pos, 2);
  T* s = state + b * p.bstride;
  const uint64_t hb = uint64_t(1) << p.hi;
  if constexpr (MODE == kPairCZ) {
    Pair<T>* a = reinterpret_cast<Pair<T>*>(s + (i | hb));
    Pair<T> x = *a;
    x.v[1] = negate(x.v[1]);
    *a = x;
  } else if constexpr (MODE == kPairCnotT0) {
    Pair<T>* a = reinterpret_cast<Pair<T>*>(s + (i | hb));
    const Pair<T> x = *a;
    Pair<T> y;
    y.v[0] = x.v[1];
    y.v[1] = x.v[0];
    *a = y;
  } else {
    Pair<T>* a = reinterpret_cast<Pair<T>*>(s + i);
    Pair<T>* c = reinterpret_cast<Pair<T>*>(s + (i | hb));
    Pair<T> x = *a, y = *c;
    const T t1 = x.v[1];
    x.v[1] = y.v[1];
    y.v[1] = t1;
    *a = x;
    *c = y;
  }
}

template <typename T>
void launch_special_pair(void* state, SpecialParams p, int mode, hipStream_t s) {
  T* st = static_cast<T*>(state);
  if (mode == kPairCZ) special_pair_kernel<T, kPairCZ><<<grid_for(p.total), kBlock, 0, s>>>(st, p);
  else if (mode == kPairCnotT0) special_pair_kernel<T, kPairCnotT0><<<grid_for(p.total), kBlock, 0, s>>>(st, p);
  else special_pair_kernel<T, kPairCnotC0><<<grid_for(p.total), kBlock, 0, s>>>(st, p);
}

template <typename T, bool CNOT>
void launch_special(void* state, SpecialParams p, bool vec, hipStream_t s) {
  if (vec)
    special_kernel<T, CNOT, 2><<<grid_for(p.total), kBlock, 0, s>>>(static_cast<T*>(state), p);
  else
    special_kernel<T, CNOT, 1><<<grid_for(p.total), kBlock, 0, s>>>(static_cast<T*>(state), p);
}

}  // namespace

int apply_special_impl(bool cnot, void* state, int dtype, int n, int qa, int qb,
                       int64_t batch, hipStream_t s) {
  const size_t es = elem_size(dtype);
  UA_CHECK_ARG(es != 0, "unsupported dtype code %d", dtype);
  UA_CHECK_ARG(n >= 2 && n <= 62, "num_qubits=%d out of range", n);
  UA_CHECK_ARG(qa >= 0 && qa < n && qb >= 0 && qb < n && qa != qb,
               "qubits (%d, %d) invalid for %d qubits", qa, qb, n);
  UA_CHECK_ARG(batch >= 0, "negative batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(state != nullptr, "NULL state");
  const int ba = n - 1 - qa, bb = n - 1 - qb;  // CNOT: qa control, qb target
  SpecialParams p{};
  p.lo = ba < bb ? ba : bb;
  p.hi = ba < bb ? bb : ba;
  p.bstride = int64_t(1) << n;
  if (cnot) {
    p.set_bits = uint64_t(1) << ba;
    p.flip = uint64_t(1) << bb;
  } else {
    p.set_bits = (uint64_t(1) << ba) | (uint64_t(1) << bb);
    p.flip = 0;
  }
  const bool aligned = (reinterpret_cast<uintptr_t>(state) % (2 * es)) == 0;
  if (p.lo == 0 && aligned) {  // bit 0 involved: whole pairs
    const int mode = !cnot ? kPairCZ : (bb == 0 ? kPairCnotT0 : kPairCnotC0);
    p.log2_tps = n - 2;
    p.total = batch << p.log2_tps;
    switch (dtype) {
      case UA_FLOAT32: launch_special_pair<float>(state, p, mode, s); break;
      case UA_FLOAT64: launch_special_pair<double>(state, p, mode, s); break;
      case UA_COMPLEX64: launch_special_pair<float2>(state, p, mode, s); break;
      case UA_COMPLEX128: launch_special_pair<double2>(state, p, mode, s); break;
    }
    return check_launch(cnot ? "cnot_pair_kernel" : "cz_pair_kernel");
  }
  const bool vec = p.lo >= 1 && aligned;
  p.log2_tps = n - 2 - (vec ? 1 : 0);
  p.total = batch << p.log2_tps;
  switch (dtype) {
    case UA_FLOAT32: cnot ? launch_special<float, true>(state, p, vec, s) : launch_special<float, false>(state, p, vec, s); break;
    case UA_FLOAT64: cnot ? launch_special<double, true>(state, p, vec, s) : launch_special<double, false>(state, p, vec, s); break;
    case UA_COMPLEX64: cnot ? launch_special<float2, true>(state, p, vec, s) : launch_special<float2, false>(state, p, vec, s); break;
    case UA_COMPLEX128: cnot ? launch_special<double2, true>(state, p, vec, s) : launch_special<double2, false>(state, p, vec, s); break;
  }
  return check_launch(cnot ? "cnot_kernel" : "cz_kernel");
}

}  // namespace ua

extern "C" int ua_apply_cz(void* state, int dtype, int num_qubits, int qubit_a, int qubit_b,
                           int64_t batch, ua_stream_t stream) {
  return ua::apply_special_impl(false, state, dtype, num_qubits, qubit_a, qubit_b, batch,
                         static_cast<hipStream_t>(stream));
}

extern "C" int ua_apply_cnot(void* state, int dtype, int num_qubits, int control, int target,
                             int64_t batch, ua_stream_t stream) {
  return ua::apply_special_impl(true, state, dtype, num_qubits, control, target, batch,
                         static_cast<hipStream_t>(stream));
}
