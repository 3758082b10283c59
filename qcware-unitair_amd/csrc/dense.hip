// Dense k-qubit gate application: ua_apply_operator.
//
// Replaces the reference's permute -> einsum -> inverse-permute sequence
// (unitair/simulation/operations.py:151-186 and :258-329), which moves the
// whole state three to five times per gate, with one streaming pass that
// addresses the target bits in place:
//
//   out[i] = sum_j U[r(i), j] * in[(i & ~targets) | scatter(j)]
//
// where r(i) gathers the target bits of i into a gate index whose most
// significant bit belongs to the first listed qubit (operations.py:78-86,
// unitair/utils.py:5-27).
//
// Fast path (k <= 5): one thread owns whole groups of 2**k amplitudes (the
// amplitudes that differ only in target bits), so a group is read once and
// written once and the kernel can run in place.  On the host the targets are
// sorted by state bit ("canonical" order) and the matrix is read through the
// matching index permutation, so the lowest target is always gate bit 0.
// That fixes the access pattern at compile time:
//   VEC  : lowest target bit >= 1 -> each thread takes two neighbouring groups
//          and moves 2 amplitudes per access (16 B for complex64);
//   PAIR : lowest target bit == 0 -> the two amplitudes that differ in bit 0
//          are adjacent and move as one 16 B access;
//   SCALAR: fallback for misaligned views.
// For k >= 3 the (permuted) matrix is staged in LDS once per workgroup.
#include "dense_kernels.h"
#include "fused_common.h"

namespace ua {

int g_unroll = 0;      // units per thread for single-operator launches (0 = auto)
extern bool g_use_mfma;  // 4/5-qubit complex gates on the matrix cores
int apply_dense_mfma(void* out, const void* in, const void* op, int dtype, int n, int k,
                     const int* tb, const int* gp, int64_t batch, int64_t in_bstride,
                     hipStream_t s);
bool g_use_xch = true; // lane-exchange path for 1-qubit gates on index bits 1..6

using namespace dense_detail;

namespace {

// One-qubit gate whose target is index bit b in [1, 6] (with 2 elements per
// lane: lane bit b-1).  Every lane moves one contiguous pair of amplitudes,
// so each wavefront access is one contiguous block; the partner amplitudes
// differ in lane bit b-1 and arrive by a cross-lane exchange.  The host
// guarantees 2**n >= 128, so every wavefront is full and inside one state.
template <typename T> __device__ __forceinline__ T shfl_xor_val(T v, int m);
template <> __device__ __forceinline__ float shfl_xor_val(float v, int m) { return __shfl_xor(v, m, 64); }
template <> __device__ __forceinline__ double shfl_xor_val(double v, int m) { return __shfl_xor(v, m, 64); }
template <> __device__ __forceinline__ float2 shfl_xor_val(float2 v, int m) {
  return make_float2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}
template <> __device__ __forceinline__ double2 shfl_xor_val(double2 v, int m) {
  return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

struct XchParams {
  int64_t total;        // units (pairs of elements) over the whole batch
  int64_t in_bstride;   // elements between input states
  int64_t out_bstride;  // elements between output states
  int log2_ups;         // log2(units per state)
  int lane_mask;        // 1 << (b - 1)
};

template <typename T, int UNR>
__global__ __launch_bounds__(kBlock) void dense_xch_kernel(
    T* out, const T* in, const T* __restrict__ op, XchParams p) {
  using A = Num<T>;
  const T u00 = op[0], u01 = op[1], u10 = op[2], u11 = op[3];
  const int64_t first = int64_t(blockIdx.x) * (kBlock * UNR) + threadIdx.x;
  Pair<T> v[UNR];
  T* dst[UNR];
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    const int64_t unit = first + int64_t(u) * kBlock;
    dst[u] = nullptr;
    if (unit >= p.total) continue;   // uniform over a wavefront
    const int64_t b = unit >> p.log2_ups;
    const int64_t e = (unit & ((int64_t(1) << p.log2_ups) - 1)) * 2;
    v[u] = *reinterpret_cast<const Pair<T>*>(in + b * p.in_bstride + e);
    dst[u] = out + b * p.out_bstride + e;
  }
  const bool upper = (threadIdx.x & p.lane_mask) != 0;
  // row of the gate this lane produces and the coefficients for (own, partner)
  const T c_own = upper ? u11 : u00;
  const T c_par = upper ? u10 : u01;
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    if (dst[u] == nullptr) continue;
    Pair<T> w;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const T partner = shfl_xor_val(v[u].v[h], p.lane_mask);
      w.v[h] = A::mac(A::mac(A::zero(), c_own, v[u].v[h]), c_par, partner);
    }
    *reinterpret_cast<Pair<T>*>(dst[u]) = w;
  }
}

// Any k (including 0 and k > 5): one thread per output amplitude, reading
// its group from memory.  Out of place only.
constexpr int kMaxQubits = 62;

struct GenericParams {
  int64_t total;  // batch * 2**n
  int64_t in_bstride, op_bstride;
  int n, k;
  uint64_t tmask;        // OR of all target bits
  int tbit[kMaxQubits];  // state bit of gate bit k-1-m, in `qubits` order
};

template <typename T>
__global__ __launch_bounds__(kBlock) void dense_generic_kernel(
    T* __restrict__ out, const T* __restrict__ in, const T* __restrict__ op,
    GenericParams p) {
  using A = Num<T>;
  const int64_t tid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (tid >= p.total) return;
  const int64_t b = tid >> p.n;
  const uint64_t x = uint64_t(tid) & ((uint64_t(1) << p.n) - 1);
  const int64_t D = int64_t(1) << p.k;
  int64_t r = 0;
  for (int m = 0; m < p.k; ++m) r = (r << 1) | int64_t((x >> p.tbit[m]) & 1);
  const uint64_t base = x & ~p.tmask;
  const T* src = in + b * p.in_bstride;
  const T* row = op + b * p.op_bstride + r * D;
  T acc = A::zero();
  for (int64_t j = 0; j < D; ++j) {
    uint64_t s = base;
    for (int m = 0; m < p.k; ++m)
      if ((j >> (p.k - 1 - m)) & 1) s |= uint64_t(1) << p.tbit[m];
    acc = A::mac(acc, row[j], src[s]);
  }
  out[tid] = acc;
}


template <typename T>
void launch_xch(void* out, const void* in, const void* op, const XchParams& p, int unr,
                hipStream_t s) {
  T* o = static_cast<T*>(out);
  const T* i = static_cast<const T*>(in);
  const T* m = static_cast<const T*>(op);
  switch (unr) {
    case 4: dense_xch_kernel<T, 4><<<grid_for(p.total, kBlock * 4), kBlock, 0, s>>>(o, i, m, p); break;
    case 2: dense_xch_kernel<T, 2><<<grid_for(p.total, kBlock * 2), kBlock, 0, s>>>(o, i, m, p); break;
    default: dense_xch_kernel<T, 1><<<grid_for(p.total, kBlock), kBlock, 0, s>>>(o, i, m, p);
  }
}

bool aligned(const void* ptr, int64_t stride_elems, size_t esize, size_t align) {
  return (reinterpret_cast<uintptr_t>(ptr) % align) == 0 &&
         ((stride_elems * int64_t(esize)) % int64_t(align)) == 0;
}

}  // namespace

// Single 6-qubit complex64 gates on the matrix cores (blocks.hip).
bool six_applies(int dtype, int n, int k, const void* out, const void* in, int64_t in_bstride);
int apply_dense_six(void* out, const void* in, const void* op, int n, const int* qubits,
                    int64_t batch, int64_t in_bstride, hipStream_t s);
// Single 7/8-qubit complex64 gates on the matrix cores, out of place (blocks.hip).
bool slice_applies(int dtype, int n, int k, const void* out, const void* in, int64_t in_bstride,
                   int64_t batch);
int apply_dense_slices(void* out, const void* in, const void* op, int n, int k,
                       const int* qubits, int64_t batch, int64_t in_bstride, hipStream_t s);

// Tiled dense gates on 6..kMaxWideGateK qubits (wide.hip).
constexpr int kMaxWideGateK = 10;
bool wide_applies(int n, int k);
int apply_dense_wide(void* out, const void* in, const void* op, int dtype, int n, int k,
                     const int* tb, int64_t batch, int64_t in_bstride, int64_t out_bstride,
                     int64_t op_bstride, hipStream_t s);

// Validated entry used by ua_apply_operator and by other host code.
int apply_operator_impl(void* out, const void* state, const void* op, int dtype,
                        int n, int k, const int* qubits, int64_t batch,
                        int64_t state_bstride, int64_t op_bstride, hipStream_t s) {
  const size_t es = elem_size(dtype);
  UA_CHECK_ARG(es != 0, "unsupported dtype code %d", dtype);
  UA_CHECK_ARG(n >= 0 && n <= kMaxQubits, "num_qubits=%d out of range", n);
  UA_CHECK_ARG(k >= 0 && k <= n, "cannot apply a %d-qubit operator to %d qubit(s)", k, n);
  UA_CHECK_ARG(batch >= 0, "batch=%lld is negative", (long long)batch);
  UA_CHECK_ARG(state_bstride >= 0 && op_bstride >= 0, "negative batch stride");
  UA_CHECK_ARG(k == 0 || qubits != nullptr, "qubits is NULL");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && state && op, "NULL data pointer");

  uint64_t seen = 0;
  for (int m = 0; m < k; ++m) {
    UA_CHECK_ARG(qubits[m] >= 0 && qubits[m] < n,
                 "qubit %d is not in range for a %d-qubit state", qubits[m], n);
    const uint64_t bit = uint64_t(1) << qubits[m];
    UA_CHECK_ARG(!(seen & bit), "qubit %d appears more than once", qubits[m]);
    seen |= bit;
  }
  const int64_t N = int64_t(1) << n;
  const bool in_place = (out == state);
  UA_CHECK_ARG(!in_place || state_bstride == N || batch == 1,
               "in-place application needs one state per batch entry");

  if (k >= 1 && k <= kMaxFastK) {
    FastParams p{};
    // canonical order: ascending state bit; gpos = caller's gate-index bit
    int tb[kMaxFastK], gp[kMaxFastK];
    for (int m = 0; m < k; ++m) {
      tb[m] = n - 1 - qubits[m];
      gp[m] = k - 1 - m;
    }
    for (int i = 1; i < k; ++i)  // insertion sort by tb
      for (int j = i; j > 0 && tb[j - 1] > tb[j]; --j) {
        int t0 = tb[j]; tb[j] = tb[j - 1]; tb[j - 1] = t0;
        int g0 = gp[j]; gp[j] = gp[j - 1]; gp[j - 1] = g0;
      }
    for (int m = 0; m < k; ++m) { p.tb[m] = tb[m]; p.gpos[m] = gp[m]; }
    // Matrix cores for the fused-block sizes where they measured faster on
    // MI355X (28-qubit complex64 / 27-qubit complex128 states): every 5-qubit
    // gate; complex128 4-qubit gates; complex64 4-qubit gates only when all
    // targets lie in the low 6 bits (the vector kernel streams the others at
    // HBM speed).
    const bool mfma_shape =
        (k == 5 && dtype != UA_FLOAT32 && dtype != UA_FLOAT64) ||
        (k == 4 && dtype == UA_COMPLEX128) || (k == 4 && dtype == UA_COMPLEX64 && tb[k - 1] < 6);
    if (g_use_mfma && mfma_shape && op_bstride == 0 && n - k >= 6)
      return apply_dense_mfma(out, state, op, dtype, n, k, tb, gp, batch, state_bstride, s);

    const size_t pair_align = 2 * es;
    const bool io_aligned = aligned(state, state_bstride, es, pair_align) &&
                            aligned(out, N, es, pair_align);
    // measured on MI355X: 1-qubit gates stream best with one group pair per
    // thread, wider gates with two units per thread
    const int unr = (op_bstride != 0) ? 1 : (g_unroll ? g_unroll : (k == 1 ? 1 : 2));
    if (g_use_xch && k == 1 && io_aligned && op_bstride == 0 && n >= 7 && tb[0] >= 1 &&
        tb[0] <= 6) {
      XchParams x{};
      x.log2_ups = n - 1;
      x.total = batch << x.log2_ups;
      x.in_bstride = state_bstride;
      x.out_bstride = N;
      x.lane_mask = 1 << (tb[0] - 1);
      switch (dtype) {
        case UA_FLOAT32: launch_xch<float>(out, state, op, x, unr, s); break;
        case UA_FLOAT64: launch_xch<double>(out, state, op, x, unr, s); break;
        case UA_COMPLEX64: launch_xch<float2>(out, state, op, x, unr, s); break;
        case UA_COMPLEX128: launch_xch<double2>(out, state, op, x, unr, s); break;
      }
      return check_launch("dense_xch_kernel");
    }
    int mode = SCALAR;
    if (io_aligned) mode = (tb[0] == 0) ? PAIR : VEC;
    const int log2_groups = n - k;  // groups per state
    const int log2_tps = log2_groups - (mode == VEC ? 1 : 0);
    p.log2_tps = log2_tps;
    p.total = batch << log2_tps;
    p.in_bstride = state_bstride;
    p.out_bstride = N;
    p.op_bstride = op_bstride;
    const bool lds = (k >= 3) && (op_bstride == 0 || log2_tps >= 8);
    switch (dtype) {
      case UA_FLOAT32: launch_fast_k_f32(k, mode, lds, out, state, op, p, unr, s); break;
      case UA_FLOAT64: launch_fast_k_f64(k, mode, lds, out, state, op, p, unr, s); break;
      case UA_COMPLEX64: launch_fast_k_c64(k, mode, lds, out, state, op, p, unr, s); break;
      case UA_COMPLEX128: launch_fast_k_c128(k, mode, lds, out, state, op, p, unr, s); break;
    }
    return check_launch("dense_fast_kernel");
  }

  if (g_use_mfma && op_bstride == 0 && six_applies(dtype, n, k, out, state, state_bstride))
    return apply_dense_six(out, state, op, n, qubits, batch, state_bstride, s);
  if (g_use_mfma && op_bstride == 0 &&
      slice_applies(dtype, n, k, out, state, state_bstride, batch))
    return apply_dense_slices(out, state, op, n, k, qubits, batch, state_bstride, s);
  if (wide_applies(n, k)) {  // 6..10 targets: tiled, in place or not
    int tbc[kMaxWideGateK];
    for (int m = 0; m < k; ++m) tbc[m] = n - 1 - qubits[m];
    return apply_dense_wide(out, state, op, dtype, n, k, tbc, batch, state_bstride, N,
                            op_bstride, s);
  }
  UA_CHECK_ARG(!in_place,
               "in-place application supports operators on 1..%d qubits (6..%d with at "
               "least 12 qubits)", kMaxFastK, kMaxWideGateK);
  UA_CHECK_ARG(k <= 20, "operators on more than 20 qubits are not supported");
  GenericParams g{};
  g.total = batch * N;
  g.in_bstride = state_bstride;
  g.op_bstride = op_bstride;
  g.n = n;
  g.k = k;
  g.tmask = 0;
  for (int m = 0; m < k; ++m) {
    g.tbit[m] = n - 1 - qubits[m];
    g.tmask |= uint64_t(1) << g.tbit[m];
  }
  const unsigned grid = grid_for(g.total);
  switch (dtype) {
    case UA_FLOAT32:
      dense_generic_kernel<float><<<grid, kBlock, 0, s>>>(
          (float*)out, (const float*)state, (const float*)op, g);
      break;
    case UA_FLOAT64:
      dense_generic_kernel<double><<<grid, kBlock, 0, s>>>(
          (double*)out, (const double*)state, (const double*)op, g);
      break;
    case UA_COMPLEX64:
      dense_generic_kernel<float2><<<grid, kBlock, 0, s>>>(
          (float2*)out, (const float2*)state, (const float2*)op, g);
      break;
    case UA_COMPLEX128:
      dense_generic_kernel<double2><<<grid, kBlock, 0, s>>>(
          (double2*)out, (const double2*)state, (const double2*)op, g);
      break;
  }
  return check_launch("dense_generic_kernel");
}

}  // namespace ua

extern "C" int ua_set_tuning(int key, int value) {
  switch (key) {
    case UA_TUNE_UNROLL:
      UA_CHECK_ARG(value == 0 || value == 1 || value == 2 || value == 4,
                   "unroll must be 0 (auto), 1, 2 or 4");
      ua::g_unroll = value;
      return UA_OK;
    case UA_TUNE_LANE_EXCHANGE:
      ua::g_use_xch = value != 0;
      return UA_OK;
    case UA_TUNE_MFMA:
      ua::g_use_mfma = value != 0;
      return UA_OK;
    case UA_TUNE_COMPILED_WAVES:
      return ua::set_compiled_waves(value);
  }
  ua::set_error("unknown tuning key %d", key);
  return UA_ERR_INVALID_ARGUMENT;
}

extern "C" int ua_apply_operator(void* out, const void* state, const void* op,
                                 int dtype, int num_qubits, int op_num_qubits,
                                 const int* qubits, int64_t batch,
                                 int64_t state_batch_stride,
                                 int64_t op_batch_stride, ua_stream_t stream) {
  return ua::apply_operator_impl(out, state, op, dtype, num_qubits, op_num_qubits,
                                 qubits, batch, state_batch_stride, op_batch_stride,
                                 static_cast<hipStream_t>(stream));
}
