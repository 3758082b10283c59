// 1-qubit gate layers: ua_apply_all_qubits (the same 2x2 operator on every
// qubit, unitair/simulation/operations.py:332-413) and ua_apply_1q_layer (one
// operator per listed qubit, the fused form of apply_to_qubits,
// operations.py:416-503).
//
// The reference loops over qubits with swap (a transposed view) + einsum and
// finishes with a permute/roll and a contiguous copy: one pass over the state
// per qubit plus the copy.  Gates on distinct qubits commute, so here they are
// grouped into tiled passes.  A workgroup loads a tile of 2^TB amplitudes into
// LDS (TB = 12, or 13 for complex64 where that saves a pass); the tile spans
// TB index bits ("local bits").  In up to four stages every thread holds a
// cube of 16 amplitudes over 4 of the local bits and applies each target
// bit's 2x2 matrix in registers.
//
// Every tile keeps the lowest R state bits (R = 2 for complex64, 1 for
// complex128: runs of at least 32 bytes) and takes up to TB - R target bits;
// the rest of its local bits are the lowest other state bits, so a pass whose
// targets are few or low reads long contiguous runs.  24 qubits of complex64
// apply_all_qubits take 2 passes of 13-bit tiles: bits 0..12 (contiguous
// tiles) and bits 0,1 + 13..23.  A pass whose tiles share 128-byte lines
// hands each XCD a contiguous range of tiles, so the tiles sharing a line
// meet in the same L2.  The first pass reads the input and writes `out`;
// later passes run in place on `out`.  Complex states with at least 12
// qubits use the tiled passes; other cases apply the gates qubit by qubit
// with the dense kernels.
//
// Cost: each gate is 8 FMAs per amplitude, so a layer of g gates on 2^n
// amplitudes is g * 8 * 2^n FMAs on the vector units; apply_all_qubits at 24
// qubits (3.2e9 FMAs) is bound by them (82 us at the FP32 vector peak)
// rather than by its two passes over HBM.
#include "ua_common.h"

namespace ua {
int apply_operator_impl(void* out, const void* state, const void* op, int dtype, int n,
                        int k, const int* qubits, int64_t batch, int64_t state_bstride,
                        int64_t op_bstride, hipStream_t s);

namespace {

constexpr int kMaxTileBits = 13;
constexpr int kMinTileBits = 12;
constexpr int kCube = 16;
constexpr int kMaxStages = 4;

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

// Two adjacent amplitudes as one 16-byte access (complex64) or two
// (complex128); plain element copies keep the values in registers.
__device__ __forceinline__ void load_pair(const float2* p, float2& a, float2& b) {
  const float4 w = *reinterpret_cast<const float4*>(p);
  a = make_float2(w.x, w.y);
  b = make_float2(w.z, w.w);
}
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load_pair(const double2* p, double2& a, double2& b) {
  const dvec2 x = reinterpret_cast<const dvec2*>(p)[0];
  const dvec2 y = reinterpret_cast<const dvec2*>(p)[1];
  a = make_double2(x.x, x.y);
  b = make_double2(y.x, y.y);
}
__device__ __forceinline__ void store_pair(float2* p, float2 a, float2 b) {
  *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ void store_pair(double2* p, double2 a, double2 b) {
  reinterpret_cast<dvec2*>(p)[0] = dvec2{a.x, a.y};
  reinterpret_cast<dvec2*>(p)[1] = dvec2{b.x, b.y};
}

struct AllqPass {
  int64_t total_tiles;       // batch * 2^(n-TB)
  int64_t in_bstride;        // batch stride of the input (0: one state for every entry)
  int64_t op_bstride;        // 0 or 4
  int n;
  int log2_tiles;            // n - TB
  int xcd_order;             // visit tiles in XCD-contiguous ranges (total_tiles % 8 == 0)
  int local[kMaxTileBits];   // ascending state bits of the tile; local[0] = 0
  int64_t target_stride;     // elements between consecutive targets' matrices (0: shared)
  int stage_bits[kMaxStages][4];   // tile bits of each stage's cube
  int stage_apply[kMaxStages];     // 4-bit mask: cube bits that get a gate
  int stage_slot[kMaxStages][4];   // target index of each cube bit's gate
  int num_stages;
};

template <typename T, int TB>
__global__ __launch_bounds__((1 << TB) / kCube) void allq_pass_kernel(T* out, const T* in,
                                                                      const T* __restrict__ op,
                                                                      AllqPass a) {
  constexpr int kTile = 1 << TB;
  constexpr int kThreads = kTile / kCube;
  constexpr int kTidBits = TB - 4;
  __shared__ T tile[kTile];
  const int tid = threadIdx.x;
  const int64_t N = int64_t(1) << a.n;

  // Element pair e = 2*tid + 2*kThreads*k of a tile (k = 0..7).  Tile bit 0
  // is state bit 0, so the pair is contiguous in memory; tile bits 1..TB-4
  // come from the thread index and the top three from k.
  uint64_t dep_t = 0;
#pragma unroll
  for (int i = 1; i <= kTidBits; ++i) dep_t |= uint64_t(((2 * tid) >> i) & 1) << a.local[i];
  uint64_t dep_k[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    dep_k[k] = 0;
#pragma unroll
    for (int i = TB - 3; i < TB; ++i)
      dep_k[k] |= uint64_t((k >> (i - (TB - 3))) & 1) << a.local[i];
  }
  // One tile per workgroup.  XCD order: workgroups are dealt to the 8 XCDs
  // round-robin, so workgroup w takes tile (w % 8) * (tiles / 8) + w / 8 and
  // each XCD walks a contiguous range of tiles.
  int64_t tg = blockIdx.x;
  if (tg >= a.total_tiles) return;
  if (a.xcd_order) tg = (tg & 7) * (a.total_tiles >> 3) + (tg >> 3);
  auto origin_of = [&](int64_t t, int64_t bstride) -> int64_t {
    const int64_t b = t >> a.log2_tiles;
    const uint64_t tis = uint64_t(t) & ((uint64_t(1) << a.log2_tiles) - 1);
    return b * bstride + int64_t(insert_zero_bits(tis, a.local, TB) | dep_t);
  };
  {
    const T* src = in + origin_of(tg, a.in_bstride);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      T v0, v1;
      load_pair(src + dep_k[k], v0, v1);
      const int e = 2 * tid + 2 * kThreads * k;
      tile[lds_pos(e)] = v0;
      tile[lds_pos(e + 1)] = v1;
    }
  }
  const T* ops_b = op + (tg >> a.log2_tiles) * a.op_bstride;
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
    if (act & (1 << K)) {                                                   \
      const T* u = ops_b + a.stage_slot[st][K] * a.target_stride;           \
      cube_1q<K>(c, u[0], u[1], u[2], u[3]);                                \
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
  T* dst = out + origin_of(tg, N);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = 2 * tid + 2 * kThreads * k;
    store_pair(dst + dep_k[k], tile[lds_pos(e)], tile[lds_pos(e + 1)]);
  }
}

// Plan the passes for the gates on the state bits in `active` (slot[bit] =
// the bit's target index) with tiles of TB bits that keep the lowest R state
// bits.  Target bits >= R go TB - R to a pass in ascending order; targets
// below R ride in the first pass.  Each tile is completed with the lowest
// state bits it lacks.  A pass's targets form stages of four tile bits (the
// last stage padded with bits that carry no gate there).  With no target at
// all, one pass without stages copies the input.  Returns the number of
// passes (at most 8; needs R < TB <= n).
int plan_layer(int n, uint64_t active, const int* slot, int TB, int R, AllqPass* passes) {
  const uint64_t low_mask = (uint64_t(1) << R) - 1;
  int rest[64], nrest = 0;
  for (int bit = R; bit < n; ++bit)
    if ((active >> bit) & 1) rest[nrest++] = bit;
  const int cap = TB - R;
  const int np = nrest == 0 ? 1 : (nrest + cap - 1) / cap;
  for (int pi = 0; pi < np; ++pi) {
    AllqPass& p = passes[pi];
    uint64_t gates = (pi == 0) ? (active & low_mask) : 0;
    for (int i = pi * cap; i < nrest && i < (pi + 1) * cap; ++i) gates |= uint64_t(1) << rest[i];
    uint64_t local = gates | low_mask;
    for (int bit = 0; __builtin_popcountll(local) < TB; ++bit) local |= uint64_t(1) << bit;
    int pos_of_gate[kMaxTileBits], ng = 0;
    for (int i = 0; i < TB; ++i) {
      p.local[i] = __builtin_ctzll(local);
      local &= local - 1;
      if ((gates >> p.local[i]) & 1) pos_of_gate[ng++] = i;
    }
    p.num_stages = 0;
    for (int g0 = 0; g0 < ng; g0 += 4) {
      const int st = p.num_stages++;
      int cube[4], nc = 0;
      for (int g = g0; g < ng && g < g0 + 4; ++g) cube[nc++] = pos_of_gate[g];
      const int applied = nc;
      for (int t = 0; nc < 4; ++t) {  // pad with tile bits outside this cube
        bool used = false;
        for (int c = 0; c < nc; ++c) used |= cube[c] == t;
        if (!used) cube[nc++] = t;
      }
      int on[4];
      for (int k = 0; k < 4; ++k) on[k] = k < applied;
      for (int i = 1; i < 4; ++i)  // ascending tile bits (insert_zero_bits needs it)
        for (int j = i; j > 0 && cube[j - 1] > cube[j]; --j) {
          const int c = cube[j]; cube[j] = cube[j - 1]; cube[j - 1] = c;
          const int o = on[j]; on[j] = on[j - 1]; on[j - 1] = o;
        }
      p.stage_apply[st] = 0;
      for (int k = 0; k < 4; ++k) {
        p.stage_bits[st][k] = cube[k];
        p.stage_apply[st] |= on[k] << k;
        p.stage_slot[st][k] = on[k] ? slot[p.local[cube[k]]] : 0;
      }
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
  if (n >= kMinTileBits && aligned) {
    // 12-bit tiles (32 KB of LDS for complex64: five workgroups per CU) keep
    // runs of at least 32 bytes: 4 amplitudes of complex64 (R = 2), 2 of
    // complex128 (R = 1).  Complex64 takes 13-bit tiles (64 KB, two per CU)
    // only where they save a pass: measured on MI355X, a 13-bit pass costs
    // about what a 12-bit one does plus its extra stage.
    const bool c64 = dtype == UA_COMPLEX64;
    const int R = c64 ? 2 : 1;
    const int line_bits = c64 ? 4 : 3;  // 128-byte lines
    AllqPass passes[8];
    int TB = kMinTileBits;
    int np = plan_layer(n, active, slot, TB, R, passes);
    if (c64 && n >= kMaxTileBits) {
      AllqPass wide[8];
      const int np13 = plan_layer(n, active, slot, kMaxTileBits, R, wide);
      if (np13 < np) {
        TB = kMaxTileBits;
        np = np13;
        for (int i = 0; i < np; ++i) passes[i] = wide[i];
      }
    }
    for (int i = 0; i < np; ++i) {
      AllqPass& p = passes[i];
      p.n = n;
      p.log2_tiles = n - TB;
      p.total_tiles = batch << p.log2_tiles;
      p.in_bstride = (i == 0) ? state_bstride : N;
      p.op_bstride = op_bstride;
      p.target_stride = target_stride;
      UA_CHECK_ARG(p.total_tiles <= 0x7fffffffLL, "too many tiles");
      // tiles share 128-byte lines when a bit below the line size is not
      // local: then each XCD takes a contiguous range of tiles, so the
      // tiles sharing a line meet in one L2 (measured: 1.6-2.4x faster)
      bool shares = false;
      for (int bit = 0; bit < line_bits; ++bit) {
        bool local = false;
        for (int j = 0; j < TB; ++j) local |= p.local[j] == bit;
        shares |= !local;
      }
      p.xcd_order = shares && p.total_tiles % 8 == 0;
      const void* src = (i == 0) ? state : out;
      const unsigned grid = unsigned(p.total_tiles);
      if (TB == kMaxTileBits)
        allq_pass_kernel<float2, 13><<<grid, (1 << 13) / kCube, 0, s>>>(
            static_cast<float2*>(out), static_cast<const float2*>(src),
            static_cast<const float2*>(op), p);
      else if (c64)
        allq_pass_kernel<float2, 12><<<grid, (1 << 12) / kCube, 0, s>>>(
            static_cast<float2*>(out), static_cast<const float2*>(src),
            static_cast<const float2*>(op), p);
      else
        allq_pass_kernel<double2, 12><<<grid, (1 << 12) / kCube, 0, s>>>(
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
