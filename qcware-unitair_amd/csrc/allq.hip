// Same 2x2 operator on every qubit: ua_apply_all_qubits
// (unitair/simulation/operations.py:332-413).
//
// The reference loops over qubits with swap (a transposed view) + einsum and
// finishes with a roll and a contiguous copy: n passes over the state plus
// the copy.  Here the 1-qubit gates, which all commute, are grouped into
// tiled passes.  A workgroup loads a 2^12-amplitude tile into LDS; the tile
// spans 12 index bits ("local bits").  In three stages every thread holds a
// cube of 16 amplitudes over 4 of the local bits and applies U to each of
// those bits that the pass owns, in registers.
//   pass 0: local bits 0..11, U applied to bits 0..a0-1;
//   pass p: local bits 0..3 (not applied; they keep every access a 128 B
//           run) plus 8 further bits, applied.
// 24 qubits take 3 passes instead of 24.  The first pass reads the input and
// writes `out`; later passes run in place on `out`.  Complex states with at
// least 12 qubits use the tiled passes; other cases apply U qubit by qubit
// with the dense kernels.
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
  int stage_bits[3][4];      // tile bits of each stage's cube
  int stage_apply[3];        // 4-bit mask: cube bits that get U
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

  const T* u = op + b * a.op_bstride;
  const T u00 = u[0], u01 = u[1], u10 = u[2], u11 = u[3];

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
    if (act & 1) cube_1q<0>(c, u00, u01, u10, u11);
    if (act & 2) cube_1q<1>(c, u00, u01, u10, u11);
    if (act & 4) cube_1q<2>(c, u00, u01, u10, u11);
    if (act & 8) cube_1q<3>(c, u00, u01, u10, u11);
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

// Split the n state bits into tiled passes (see the file comment).  The
// fewest passes are P = 1 + ceil((n - 12) / 8); the gates are spread so that
// pass 0 applies bits 0..a0-1 (a0 = n - 8 (P - 1), between 4 and 12) and every
// later pass applies 8 bits.  Spreading them keeps each pass below its memory
// time (12 gates in one pass are VALU-bound without packed FP32).  Returns the
// number of passes written to `passes` (at most 8).
int plan_allq(int n, AllqPass* passes) {
  const int P = 1 + (n - kTileBits + 7) / 8;
  const int a0 = n - 8 * (P - 1);
  int np = 0;
  {
    AllqPass& p = passes[np++];
    for (int i = 0; i < kTileBits; ++i) p.local[i] = i;
    p.num_stages = 0;
    for (int s = 0; s < 3; ++s) {
      int mask = 0;
      for (int k = 0; k < 4; ++k)
        if (4 * s + k < a0) mask |= 1 << k;
      if (!mask) continue;
      for (int k = 0; k < 4; ++k) p.stage_bits[p.num_stages][k] = 4 * s + k;
      p.stage_apply[p.num_stages++] = mask;
    }
  }
  for (int hi0 = a0; hi0 < n; hi0 += 8) {
    AllqPass& p = passes[np++];
    for (int i = 0; i < 4; ++i) p.local[i] = i;
    for (int i = 0; i < 8; ++i) p.local[4 + i] = hi0 + i;
    p.num_stages = 2;
    for (int s = 0; s < 2; ++s) {
      for (int k = 0; k < 4; ++k) p.stage_bits[s][k] = 4 + 4 * s + k;
      p.stage_apply[s] = 0xF;
    }
  }
  return np;
}

}  // namespace
}  // namespace ua

extern "C" int ua_apply_all_qubits(void* out, const void* state, const void* op, int dtype,
                                   int num_qubits, int64_t batch,
                                   int64_t state_batch_stride, int64_t op_batch_stride,
                                   ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 1 && num_qubits <= 62, "apply_all_qubits needs 1..62 qubits");
  UA_CHECK_ARG(out != state, "apply_all_qubits is out of place");
  UA_CHECK_ARG(batch >= 0, "negative batch");
  UA_CHECK_ARG(op_batch_stride == 0 || op_batch_stride == 4,
               "op_batch_stride must be 0 or 4");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && state && op, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t N = int64_t(1) << num_qubits;
  const bool complex = dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128;
  const size_t es = complex ? elem_size(dtype) : 0;
  const bool aligned = complex &&
                       reinterpret_cast<uintptr_t>(out) % (2 * es) == 0 &&
                       reinterpret_cast<uintptr_t>(state) % (2 * es) == 0 &&
                       state_batch_stride % 2 == 0;
  if (num_qubits >= kTileBits && aligned) {
    AllqPass passes[8];
    const int np = plan_allq(num_qubits, passes);
    for (int i = 0; i < np; ++i) {
      AllqPass& p = passes[i];
      p.n = num_qubits;
      p.log2_tiles = num_qubits - kTileBits;
      p.total_tiles = batch << p.log2_tiles;
      p.in_bstride = (i == 0) ? state_batch_stride : N;
      p.op_bstride = op_batch_stride;
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
  for (int q = 0; q < num_qubits; ++q) {
    const int qubit[1] = {q};
    const void* src = (q == 0) ? state : out;
    const int64_t sstride = (q == 0) ? state_batch_stride : N;
    const int rc = apply_operator_impl(out, src, op, dtype, num_qubits, 1, qubit, batch,
                                       sstride, op_batch_stride, s);
    if (rc) return rc;
  }
  return UA_OK;
}
