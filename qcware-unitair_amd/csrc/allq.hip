// 1-qubit gate layers: ua_apply_all_qubits (the same 2x2 operator on every
// qubit, unitair/simulation/operations.py:332-413) and ua_apply_1q_layer (one
// operator per listed qubit, the fused form of apply_to_qubits,
// operations.py:416-503).
//
// The reference loops over qubits with swap (a transposed view) + einsum and
// finishes with a permute/roll and a contiguous copy: one pass over the state
// per qubit plus the copy.  Gates on distinct qubits commute, so here they are
// grouped into tiled passes.  A workgroup loads a 2^12-amplitude tile into
// LDS; the tile spans 12 index bits ("local bits").  In up to three stages
// every thread holds a cube of 16 amplitudes over 4 of the local bits and
// applies each target bit's 2x2 matrix in registers.
//   low pass:  local bits 0..11, gates on target bits among them;
//   high pass: local bits 0..3 (no gates there; they keep every access a
//              128 B run) plus 8 further bits carrying up to 8 gates.
// Low targets on bits 4..11 move into spare slots of the high passes, so a
// gate set takes the fewest passes and no pass carries more than 12 gates.
// 24 qubits of apply_all_qubits take 3 passes instead of 24.  The first pass
// reads the input and writes `out`; later passes run in place on `out`.
// Complex states with at least 12 qubits use the tiled passes; other cases
// apply the gates qubit by qubit with the dense kernels.
#include "ua_common.h"

namespace ua {
int apply_operator_impl(void* out, const void* state, const void* op, int dtype, int n,
                        int k, const int* qubits, int64_t batch, int64_t state_bstride,
                        int64_t op_bstride, hipStream_t s);

namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kCube = 16;
constexpr int kThreads = kTile / kCube;  // 256

// XOR swizzle of tile element e in LDS; spreads the 16-element strides that
// cube reads and writes make over distinct banks.
__device__ __forceinline__ int lds_pos(int e) { return e ^ ((e >> 5) & 31); }

template <typename T>
__device__ __forceinline__ T cmad(T acc, T a, T b) {  // acc + a*b
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}

template <typename T>
__device__ __forceinline__ T cmul(T a, T b) {
  T r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

template <int B, typename T>
__device__ __forceinline__ void cube_1q(T (&c)[kCube], const T u00, const T u01, const T u10,
                                        const T u11) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & (1 << B)) continue;
    const int k = j | (1 << B);
    const T a = c[j], b = c[k];
    c[j] = cmad(cmul(u00, a), u01, b);
    c[k] = cmad(cmul(u10, a), u11, b);
  }
}

struct AllqPass {
  int64_t total_tiles;       // batch * 2^(n-12)
  int64_t in_bstride;        // batch stride of the input (0: one state for every entry)
  int64_t op_bstride;        // 0 or 4
  int n;
  int log2_tiles;            // n - 12
  int local[kTileBits];      // ascending state bits of the tile
  int64_t target_stride;     // elements between consecutive targets' matrices (0: shared)
  int stage_bits[3][4];      // tile bits of each stage's cube
  int stage_apply[3];        // 4-bit mask: cube bits that get a gate
  int stage_slot[3][4];      // target index of each cube bit's gate
  int num_stages;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void allq_pass_kernel(T* out, const T* in,
                                                             const T* __restrict__ op,
                                                             AllqPass a) {
  __shared__ T tile[kTile];
  const int tid = threadIdx.x;
  const int64_t tg = blockIdx.x;  // one tile per workgroup
  if (tg >= a.total_tiles) return;
  const int64_t b = tg >> a.log2_tiles;
  const uint64_t tis = uint64_t(tg) & ((uint64_t(1) << a.log2_tiles) - 1);
  const uint64_t origin = insert_zero_bits(tis, a.local, kTileBits);
  const int64_t N = int64_t(1) << a.n;
  const T* src = in + b * a.in_bstride;
  T* dst = out + b * N;

  // ---- coalesced load: element pair f = 2*tid + 512*k; tile bits 0..3 are
  // state bits 0..3, so the pair is contiguous in memory.
  uint64_t dep_t = 0;
#pragma unroll
  for (int i = 1; i < 9; ++i) dep_t |= uint64_t(((2 * tid) >> i) & 1) << a.local[i];
  uint64_t dep_k[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    dep_k[k] = 0;
#pragma unroll
    for (int i = 9; i < 12; ++i) dep_k[k] |= uint64_t((k >> (i - 9)) & 1) << a.local[i];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = 2 * tid + 512 * k;
    const Pair<T> v = *reinterpret_cast<const Pair<T>*>(src + (origin | dep_t | dep_k[k]));
    tile[lds_pos(e)] = v.v[0];
    tile[lds_pos(e + 1)] = v.v[1];
  }

  const T* ops_b = op + b * a.op_bstride;

  for (int st = 0; st < a.num_stages; ++st) {
    __syncthreads();
    const int sb[4] = {a.stage_bits[st][0], a.stage_bits[st][1], a.stage_bits[st][2],
                       a.stage_bits[st][3]};
    const int e0 = int(insert_zero_bits(uint64_t(tid), sb, 4));
    const int m0 = 1 << sb[0], m1 = 1 << sb[1], m2 = 1 << sb[2], m3 = 1 << sb[3];
    T c[kCube];
#pragma unroll
    for (int j = 0; j < kCube; ++j) {
      const int e = e0 | ((j & 1) ? m0 : 0) | ((j & 2) ? m1 : 0) | ((j & 4) ? m2 : 0) |
                    ((j & 8) ? m3 : 0);
      c[j] = tile[lds_pos(e)];
    }
    const int act = a.stage_apply[st];
#define UA_LAYER_BIT(K)                                                       \
    if (act & (1 << K)) {                                                     \
      const T* u = ops_b + a.stage_slot[st][K] * a.target_stride;             \
      cube_1q<K>(c, u[0], u[1], u[2], u[3]);                                  \
    }
    UA_LAYER_BIT(0)
    UA_LAYER_BIT(1)
    UA_LAYER_BIT(2)
    UA_LAYER_BIT(3)
#undef UA_LAYER_BIT
#pragma unroll
    for (int j = 0; j < kCube; ++j) {
      const int e = e0 | ((j & 1) ? m0 : 0) | ((j & 2) ? m1 : 0) | ((j & 4) ? m2 : 0) |
                    ((j & 8) ? m3 : 0);
      tile[lds_pos(e)] = c[j];
    }
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = 2 * tid + 512 * k;
    Pair<T> v;
    v.v[0] = tile[lds_pos(e)];
    v.v[1] = tile[lds_pos(e + 1)];
    *reinterpret_cast<Pair<T>*>(dst + (origin | dep_t | dep_k[k])) = v;
  }
}

// Plan the passes for the gates on the state bits in `active` (slot[bit] =
// the bit's target index).  High targets (bits >= 12) go 8 to a pass; low
// targets on bits 4..11 fill the high passes' spare slots, highest first; the
// rest of the low targets share one pass over bits 0..11.  With no target at
// all, one pass without stages copies the input.  Returns the number of
// passes (at most 8).
int plan_layer(int n, uint64_t active, const int* slot, AllqPass* passes) {
  uint64_t low = active & 0xFFFull, high = active & ~0xFFFull;
  const int H = __builtin_popcountll(high);
  int spare = 8 * ((H + 7) / 8) - H;
  uint64_t movable = low & 0xFF0ull;
  while (spare > 0 && movable) {
    const int bit = 63 - __builtin_clzll(movable);
    movable &= ~(uint64_t(1) << bit);
    low &= ~(uint64_t(1) << bit);
    high |= uint64_t(1) << bit;
    --spare;
  }
  int np = 0;
  if (low || !high) {
    AllqPass& p = passes[np++];
    for (int i = 0; i < kTileBits; ++i) p.local[i] = i;
    p.num_stages = 0;
    for (int s = 0; s < 3; ++s) {
      int mask = 0;
      for (int k = 0; k < 4; ++k)
        if ((low >> (4 * s + k)) & 1) mask |= 1 << k;
      if (!mask) continue;
      for (int k = 0; k < 4; ++k) {
        p.stage_bits[p.num_stages][k] = 4 * s + k;
        p.stage_slot[p.num_stages][k] = (mask >> k) & 1 ? slot[4 * s + k] : 0;
      }
      p.stage_apply[p.num_stages++] = mask;
    }
  }
  while (high) {
    AllqPass& p = passes[np++];
    uint64_t chunk = 0;
    for (int c = 0; c < 8 && high; ++c) {
      const uint64_t lowest = high & (~high + 1);
      chunk |= lowest;
      high &= ~lowest;
    }
    // complete the eight high local bits with bits >= 4 that carry no gate here
    uint64_t local_hi = chunk;
    for (int bit = 4; __builtin_popcountll(local_hi) < 8; ++bit)
      local_hi |= uint64_t(1) << bit;
    for (int i = 0; i < 4; ++i) p.local[i] = i;
    for (int i = 4; i < kTileBits; ++i) {
      p.local[i] = __builtin_ctzll(local_hi);
      local_hi &= local_hi - 1;
    }
    p.num_stages = 0;
    for (int s = 0; s < 2; ++s) {
      int mask = 0;
      for (int k = 0; k < 4; ++k)
        if ((chunk >> p.local[4 + 4 * s + k]) & 1) mask |= 1 << k;
      if (!mask) continue;
      for (int k = 0; k < 4; ++k) {
        const int t = 4 + 4 * s + k;
        p.stage_bits[p.num_stages][k] = t;
        p.stage_slot[p.num_stages][k] = (mask >> k) & 1 ? slot[p.local[t]] : 0;
      }
      p.stage_apply[p.num_stages++] = mask;
    }
  }
  return np;
}

}  // namespace
}  // namespace ua

namespace ua {
namespace {

// Gates on distinct qubits: the matrix of target j for batch entry b is at
// op + b * op_bstride + j * target_stride.
int layer_impl(void* out, const void* state, const void* op, int dtype, int n,
               int num_targets, const int* qubits, int64_t batch, int64_t state_bstride,
               int64_t op_bstride, int64_t target_stride, hipStream_t s) {
  UA_CHECK_ARG(n >= 1 && n <= 62, "1-qubit layers need 1..62 qubits (got %d)", n);
  UA_CHECK_ARG(batch >= 0, "negative batch");
  UA_CHECK_ARG(batch == 0 || out != state, "1-qubit layers are out of place");
  UA_CHECK_ARG(num_targets >= 0 && num_targets <= n, "num_targets=%d out of range",
               num_targets);
  UA_CHECK_ARG(op_bstride >= 0 && target_stride >= 0, "negative operator stride");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && state && (op || num_targets == 0), "NULL pointer");
  UA_CHECK_ARG(num_targets == 0 || qubits, "NULL qubits");
  int slot[64];
  uint64_t active = 0;
  for (int j = 0; j < num_targets; ++j) {
    const int q = qubits[j];
    UA_CHECK_ARG(q >= 0 && q < n, "qubit %d out of range for %d qubits", q, n);
    const int bit = n - 1 - q;
    UA_CHECK_ARG(!((active >> bit) & 1), "qubit %d listed twice", q);
    active |= uint64_t(1) << bit;
    slot[bit] = j;
  }
  const int64_t N = int64_t(1) << n;
  const bool complex = dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128;
  const size_t es = complex ? elem_size(dtype) : 0;
  const bool aligned = complex && reinterpret_cast<uintptr_t>(out) % (2 * es) == 0 &&
                       reinterpret_cast<uintptr_t>(state) % (2 * es) == 0 &&
                       state_bstride % 2 == 0;
  if (n >= kTileBits && aligned) {
    AllqPass passes[8];
    const int np = plan_layer(n, active, slot, passes);
    for (int i = 0; i < np; ++i) {
      AllqPass& p = passes[i];
      p.n = n;
      p.log2_tiles = n - kTileBits;
      p.total_tiles = batch << p.log2_tiles;
      p.in_bstride = (i == 0) ? state_bstride : N;
      p.op_bstride = op_bstride;
      p.target_stride = target_stride;
      UA_CHECK_ARG(p.total_tiles <= 0x7fffffffLL, "too many tiles");
      const void* src = (i == 0) ? state : out;
      if (dtype == UA_COMPLEX64)
        allq_pass_kernel<float2><<<unsigned(p.total_tiles), kThreads, 0, s>>>(
            static_cast<float2*>(out), static_cast<const float2*>(src),
            static_cast<const float2*>(op), p);
      else
        allq_pass_kernel<double2><<<unsigned(p.total_tiles), kThreads, 0, s>>>(
            static_cast<double2*>(out), static_cast<const double2*>(src),
            static_cast<const double2*>(op), p);
      const int rc = check_launch("allq_pass_kernel");
      if (rc) return rc;
    }
    return UA_OK;
  }
  UA_CHECK_ARG(dtype >= UA_FLOAT32 && dtype <= UA_COMPLEX128, "bad dtype %d", dtype);
  if (num_targets == 0) {
    const size_t bytes = size_t(N) * elem_size(dtype);
    for (int64_t b = 0; b < batch; ++b) {
      const hipError_t e = hipMemcpyAsync(
          static_cast<char*>(out) + b * bytes,
          static_cast<const char*>(state) + b * state_bstride * elem_size(dtype), bytes,
          hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) {
        set_error("hipMemcpyAsync: %s", hipGetErrorString(e));
        return UA_ERR_HIP;
      }
    }
    return UA_OK;
  }
  const size_t mes = elem_size(dtype);
  for (int j = 0; j < num_targets; ++j) {
    const int qubit[1] = {qubits[j]};
    const void* src = (j == 0) ? state : out;
    const int64_t sstride = (j == 0) ? state_bstride : N;
    const void* u = static_cast<const char*>(op) + size_t(j * target_stride) * mes;
    const int rc = apply_operator_impl(out, src, u, dtype, n, 1, qubit, batch, sstride,
                                       op_bstride, s);
    if (rc) return rc;
  }
  return UA_OK;
}

}  // namespace
}  // namespace ua

extern "C" int ua_apply_all_qubits(void* out, const void* state, const void* op, int dtype,
                                   int num_qubits, int64_t batch,
                                   int64_t state_batch_stride, int64_t op_batch_stride,
                                   ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 1 && num_qubits <= 62, "apply_all_qubits needs 1..62 qubits");
  UA_CHECK_ARG(op_batch_stride == 0 || op_batch_stride == 4,
               "op_batch_stride must be 0 or 4");
  int qubits[62];
  for (int q = 0; q < num_qubits; ++q) qubits[q] = q;
  return layer_impl(out, state, op, dtype, num_qubits, num_qubits, qubits, batch,
                    state_batch_stride, op_batch_stride, 0, static_cast<hipStream_t>(stream));
}

extern "C" int ua_apply_1q_layer(void* out, const void* state, const void* ops, int dtype,
                                 int num_qubits, int num_targets, const int* qubits,
                                 int64_t batch, int64_t state_batch_stride,
                                 int64_t ops_batch_stride, ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(ops_batch_stride == 0 || ops_batch_stride == 4 * int64_t(num_targets),
               "ops_batch_stride must be 0 or 4 * num_targets");
  return layer_impl(out, state, ops, dtype, num_qubits, num_targets, qubits, batch,
                    state_batch_stride, ops_batch_stride, 4, static_cast<hipStream_t>(stream));
}
