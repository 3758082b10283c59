// Fused local pass: many gates per read/write of the state.
//
// The reference applies every gate as its own pass over the state
// (unitair/simulation/operations.py:151-186); its only fusion is multiplying
// 1-qubit gates that hit the same qubit (unitair/gates/matrix_algebra.py:6-51,
// used by apply_to_qubits, operations.py:416-503).  Here a pass picks m <= 12
// "local" index bits (always including bits 0..3, so every tile is made of
// 128 B runs).  A workgroup loads one 2^m-amplitude tile into LDS, applies a
// list of gates, and writes the tile back:
//   - dense 1-/2-qubit gates whose targets are local bits,
//   - CNOT whose target is local (the control may be any bit),
//   - CZ, 1-qubit diagonals and full-length phases on any bits (they act
//     element by element, so they need no local bits at all).
// The gate list is cut into stages; in a stage every thread owns a "cube" of
// 16 amplitudes spanning 4 local bits, so the stage's gates run in registers
// and only stage boundaries go through LDS.
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kCube = 16;
constexpr int kThreads = kTile / kCube;  // 256
constexpr int kMatCache = UA_FUSED_MATRIX_CACHE;  // elements of gate data kept in LDS

// LDS position of tile element e: XOR-swizzled so that the 32 lanes of a
// half-wavefront reading one cube entry hit (nearly) distinct banks for every
// choice of cube bits (worst case 4-way instead of 16-way for 8 B elements).
__device__ __forceinline__ int lds_pos(int e) { return e ^ ((e >> 5) & 31); }

template <typename T> struct RealOf;
template <> struct RealOf<float2> { using type = float; };
template <> struct RealOf<double2> { using type = double; };

template <typename T>
__device__ __forceinline__ T cmul(T a, T b) {
  T r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T>
__device__ __forceinline__ T cmad(T acc, T a, T b) {  // acc + a*b
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ float2 expmi(float t) {
  float s, c;
  sincosf(t, &s, &c);
  return make_float2(c, -s);
}
__device__ __forceinline__ double2 expmi(double t) {
  double s, c;
  sincos(t, &s, &c);
  return make_double2(c, -s);
}

// 1-qubit dense gate on cube bit B.
template <int B, typename T>
__device__ __forceinline__ void cube_1q(T (&c)[kCube], const T* u) {
  const T u00 = u[0], u01 = u[1], u10 = u[2], u11 = u[3];
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & (1 << B)) continue;
    const int k = j | (1 << B);
    const T a = c[j], b = c[k];
    T z;
    z.x = 0;
    z.y = 0;
    c[j] = cmad(cmad(z, u00, a), u01, b);
    c[k] = cmad(cmad(z, u10, a), u11, b);
  }
}

// 2-qubit dense gate; gate index bit 1 = cube bit B1, bit 0 = cube bit B0.
template <int B1, int B0, typename T>
__device__ __forceinline__ void cube_2q(T (&c)[kCube], const T* u) {
  T m[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) m[i] = u[i];
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & ((1 << B1) | (1 << B0))) continue;
    const int idx[4] = {j, j | (1 << B0), j | (1 << B1), j | (1 << B1) | (1 << B0)};
    T a[4], o[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) a[s] = c[idx[s]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T acc;
      acc.x = 0;
      acc.y = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = cmad(acc, m[r * 4 + s], a[s]);
      o[r] = acc;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) c[idx[s]] = o[s];
  }
}

template <typename T>
__device__ __forceinline__ void dispatch_1q(T (&c)[kCube], int b, const T* u) {
  switch (b) {
    case 0: cube_1q<0>(c, u); break;
    case 1: cube_1q<1>(c, u); break;
    case 2: cube_1q<2>(c, u); break;
    default: cube_1q<3>(c, u); break;
  }
}

template <typename T>
__device__ __forceinline__ void dispatch_2q(T (&c)[kCube], int b1, int b0, const T* u) {
  switch (b1 * 4 + b0) {
    case 1: cube_2q<0, 1>(c, u); break;
    case 2: cube_2q<0, 2>(c, u); break;
    case 3: cube_2q<0, 3>(c, u); break;
    case 4: cube_2q<1, 0>(c, u); break;
    case 6: cube_2q<1, 2>(c, u); break;
    case 7: cube_2q<1, 3>(c, u); break;
    case 8: cube_2q<2, 0>(c, u); break;
    case 9: cube_2q<2, 1>(c, u); break;
    case 11: cube_2q<2, 3>(c, u); break;
    case 12: cube_2q<3, 0>(c, u); break;
    case 13: cube_2q<3, 1>(c, u); break;
    default: cube_2q<3, 2>(c, u); break;
  }
}

// swap cube entries that differ in cube bit B where `ctrl` has bit j set
template <int B, typename T>
__device__ __forceinline__ void cube_cnot(T (&c)[kCube], uint32_t ctrl) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & (1 << B)) continue;
    if ((ctrl >> j) & 1) {
      const T t = c[j];
      c[j] = c[j | (1 << B)];
      c[j | (1 << B)] = t;
    }
  }
}

struct PassArgs {
  int64_t total_tiles;      // batch * 2^(n-m)
  int n, m;
  int log2_tiles_per_state; // n - m
  int num_stages;
  int needs_index;          // some stage has an op that reads the global index
  int local[kTileBits];     // state bit of tile bit i (ascending; local[i] = i for i < 4)
  int stage_bits[UA_FUSED_MAX_STAGES][4];
  int stage_ops[UA_FUSED_MAX_STAGES + 1];
};

// Kernel variants by the op kinds a pass uses (register footprint, complex64):
//   kFull   every kind (229 VGPRs, 2 waves per SIMD);
//   kLight  no per-element phases and no 2-qubit matrices (164, 3 waves);
//   kLayers only 1-qubit layer ops (4 waves).
enum { kLayers = 0, kLight = 1, kFull = 2 };
template <typename T, int VARIANT>
__global__ __launch_bounds__(kThreads) void fused_pass_kernel(
    T* state, const ua_fused_op* __restrict__ ops, PassArgs a) {
  using R = typename RealOf<T>::type;
  __shared__ T tile[kTile];
  __shared__ T mcache[kMatCache];
  const int m = a.m;
  const int tiles_per_block = kTile >> m;
  const int tid = threadIdx.x;
  const int64_t N = int64_t(1) << a.n;

  // ---- coalesced load of this block's tiles into LDS (pairs of amplitudes).
  // Element f = 2*tid + 512*k of the block: its tile index bits 4.. split
  // into a per-thread part (bits 4..8 of f, from tid) and a per-k part.
  const int tiles_per_block_log2 = kTileBits - m;
  uint64_t dep_t = 0;  // deposit of f's tile-index bits 4..8 (from tid)
  for (int i = 4; i < 9 && i < m; ++i) dep_t |= uint64_t(((2 * tid) >> i) & 1) << a.local[i];
  auto tile_origin = [&](int64_t tg) -> uint64_t {
    const uint64_t tis = uint64_t(tg) & ((uint64_t(1) << a.log2_tiles_per_state) - 1);
    return insert_zero_bits(tis, a.local, m);
  };
  const int e_low = (2 * tid) & 15;
  // one tile per workgroup (m == 12): its origin is the same for every element
  const uint64_t origin1 = (tiles_per_block_log2 == 0 && int64_t(blockIdx.x) < a.total_tiles)
                               ? tile_origin(blockIdx.x) : 0;
#pragma unroll
  for (int k = 0; k < kTile / 2 / kThreads; ++k) {
    const int f = 2 * tid + k * 2 * kThreads;  // element index inside the block
    const int sub = f >> m;
    const int e = f & ((1 << m) - 1);
    const int64_t tg = int64_t(blockIdx.x) * tiles_per_block + sub;
    if (tg < a.total_tiles) {
      uint64_t dep_k = 0;  // deposit of e's bits 9.. (from k)
      for (int i = 9; i < m; ++i) dep_k |= uint64_t((e >> i) & 1) << a.local[i];
      const uint64_t g = (tiles_per_block_log2 == 0 ? origin1 : tile_origin(tg)) | dep_t |
                         dep_k | uint64_t(e_low);
      const int64_t b = tg >> a.log2_tiles_per_state;
      const Pair<T> v = *reinterpret_cast<const Pair<T>*>(state + b * N + g);
      tile[(sub << m) + lds_pos(e)] = v.v[0];
      tile[(sub << m) + lds_pos(e + 1)] = v.v[1];
    }
  }

  // ---- gate data shared by every state goes to LDS once per workgroup
  {
    const int first = a.stage_ops[0], last = a.stage_ops[a.num_stages];
    for (int o = first; o < last; ++o) {
      const ua_fused_op op = ops[o];
      if (op.cache_slot < 0) continue;
      const int len = op.kind == UA_FOP_DENSE1 ? 4 : op.kind == UA_FOP_DENSE2 ? 16 : 2;
      if (tid < len) mcache[op.cache_slot + tid] = static_cast<const T*>(op.data)[tid];
    }
  }

  // ---- this thread's tile and position in it
  const int threads_per_tile = 1 << (m - 4);
  const int sub = tid >> (m - 4);
  const int t_local = tid & (threads_per_tile - 1);
  const int64_t tg = int64_t(blockIdx.x) * tiles_per_block + sub;
  const bool active = tg < a.total_tiles;
  const int64_t b = active ? (tg >> a.log2_tiles_per_state) : 0;
  uint64_t tile_base = 0;
  if (active && a.needs_index) tile_base = tiles_per_block_log2 == 0 ? origin1 : tile_origin(tg);
  T* my_tile = tile + (sub << m);

  for (int st = 0; st < a.num_stages; ++st) {
    __syncthreads();
    const int s0 = a.stage_bits[st][0], s1 = a.stage_bits[st][1];
    const int s2 = a.stage_bits[st][2], s3 = a.stage_bits[st][3];
    const int sb[4] = {s0, s1, s2, s3};
    const int e0 = int(insert_zero_bits(uint64_t(t_local), sb, 4));
    const int m0 = 1 << s0, m1 = 1 << s1, m2 = 1 << s2, m3 = 1 << s3;
    T c[kCube];
#pragma unroll
    for (int j = 0; j < kCube; ++j) {
      const int e = e0 | ((j & 1) ? m0 : 0) | ((j & 2) ? m1 : 0) | ((j & 4) ? m2 : 0) |
                    ((j & 8) ? m3 : 0);
      c[j] = my_tile[lds_pos(e)];
    }
    // global index of cube entry j = gbase | sum of cube-bit masks
    uint64_t gbase = 0, g0 = 0, g1 = 0, g2 = 0, g3 = 0;
    if (a.needs_index) {
      gbase = tile_base;
      for (int i = 0; i < m; ++i) gbase |= uint64_t((e0 >> i) & 1) << a.local[i];
      g0 = uint64_t(1) << a.local[s0];
      g1 = uint64_t(1) << a.local[s1];
      g2 = uint64_t(1) << a.local[s2];
      g3 = uint64_t(1) << a.local[s3];
    }
    auto gidx = [&](int j) -> uint64_t {
      return gbase | ((j & 1) ? g0 : 0) | ((j & 2) ? g1 : 0) | ((j & 4) ? g2 : 0) |
             ((j & 8) ? g3 : 0);
    };
    // 16-bit mask of the cube entries whose state index has bit `bit` set
    const int cb0 = a.local[s0], cb1 = a.local[s1], cb2 = a.local[s2], cb3 = a.local[s3];
    auto sel = [&](int bit) -> uint32_t {
      if (bit == cb0) return 0xAAAAu;
      if (bit == cb1) return 0xCCCCu;
      if (bit == cb2) return 0xF0F0u;
      if (bit == cb3) return 0xFF00u;
      return ((gbase >> bit) & 1) ? 0xFFFFu : 0u;
    };
    if (active) {
      for (int o = a.stage_ops[st]; o < a.stage_ops[st + 1]; ++o) {
        const ua_fused_op op = ops[o];
        switch (op.kind) {
          case UA_FOP_DENSE1: {
            if constexpr (VARIANT != kLayers) {
            const T* u = op.cache_slot >= 0 ? mcache + op.cache_slot
                                            : static_cast<const T*>(op.data) + b * op.data_bstride;
            dispatch_1q(c, op.a, u);
            }
            break;
          }
          case UA_FOP_DENSE2: {
            if constexpr (VARIANT == kFull) {
              const T* u = op.cache_slot >= 0 ? mcache + op.cache_slot
                                              : static_cast<const T*>(op.data) + b * op.data_bstride;
              dispatch_2q(c, op.a, op.b, u);
            }
            break;
          }
          case UA_FOP_CNOT: {  // a: control state bit, b: target cube bit
            if constexpr (VARIANT != kLayers) {
            const uint32_t ctrl = sel(op.a);
            switch (op.b) {
              case 0: cube_cnot<0>(c, ctrl); break;
              case 1: cube_cnot<1>(c, ctrl); break;
              case 2: cube_cnot<2>(c, ctrl); break;
              default: cube_cnot<3>(c, ctrl); break;
            }
            }
            break;
          }
          case UA_FOP_CZ: {  // a, b: state bits
            if constexpr (VARIANT != kLayers) {
            const uint32_t neg = sel(op.a) & sel(op.b);
#pragma unroll
            for (int j = 0; j < kCube; ++j) {
              if ((neg >> j) & 1) {
                c[j].x = -c[j].x;
                c[j].y = -c[j].y;
              }
            }
            }
            break;
          }
          case UA_FOP_DIAG1: {  // a: state bit; data: (d0, d1)
            if constexpr (VARIANT != kLayers) {
            const T* d = op.cache_slot >= 0 ? mcache + op.cache_slot
                                            : static_cast<const T*>(op.data) + b * op.data_bstride;
            const T d0 = d[0], d1 = d[1];
            const uint32_t one = sel(op.a);
#pragma unroll
            for (int j = 0; j < kCube; ++j) c[j] = cmul(((one >> j) & 1) ? d1 : d0, c[j]);
            }
            break;
          }
          case UA_FOP_CZ_LAYER: {  // data: one uint64 mask per state
            if constexpr (VARIANT != kLayers) {
            // Sign of entry j: parity of popcount(g & (g >> 1) & mask) with
            // g = gbase + cube part x (disjoint bits).  Over GF(2) it splits
            // into P(gbase), a term linear in the cube bits (a cube bit at p
            // pairs with gbase bits p-1 and p+1) and a term for cube bits
            // that are neighbours; built as a 16-bit mask, no per-entry
            // 64-bit index.
            const uint64_t mask = static_cast<const uint64_t*>(op.data)[b * op.data_bstride];
            const int cbk[4] = {cb0, cb1, cb2, cb3};
            const uint32_t selk[4] = {0xAAAAu, 0xCCCCu, 0xF0F0u, 0xFF00u};
            uint32_t neg = (__popcll(gbase & (gbase >> 1) & mask) & 1) ? 0xFFFFu : 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int pk = cbk[k];
              uint64_t l = (gbase >> (pk + 1)) & (mask >> pk);
              if (pk > 0) l ^= (gbase >> (pk - 1)) & (mask >> (pk - 1));
              if (l & 1) neg ^= selk[k];
#pragma unroll
              for (int k2 = 0; k2 < 4; ++k2)
                if (cbk[k2] == pk + 1 && ((mask >> pk) & 1)) neg ^= selk[k] & selk[k2];
            }
#pragma unroll
            for (int j = 0; j < kCube; ++j) {
              if ((neg >> j) & 1) {
                c[j].x = -c[j].x;
                c[j].y = -c[j].y;
              }
            }
            }
            break;
          }
          case UA_FOP_LAYER1Q: {  // a: four 8-bit counts; DENSE1 sub-ops follow
            const uint32_t counts = uint32_t(op.a);
            int sub = o + 1;
#define UA_LAYER_BIT(B)                                                             \
            for (int t = 0, nb = int((counts >> (8 * B)) & 255u); t < nb; ++t, ++sub) { \
              const ua_fused_op so = ops[sub];                                      \
              const T* u = so.cache_slot >= 0                                       \
                               ? mcache + so.cache_slot                             \
                               : static_cast<const T*>(so.data) + b * so.data_bstride; \
              cube_1q<B>(c, u);                                                     \
            }
            UA_LAYER_BIT(0)
            UA_LAYER_BIT(1)
            UA_LAYER_BIT(2)
            UA_LAYER_BIT(3)
#undef UA_LAYER_BIT
            o = sub - 1;
            break;
          }
          case UA_FOP_PHASE: {  // data: 2^n real angles per state
            if constexpr (VARIANT == kFull) {
              const R* th = static_cast<const R*>(op.data) + b * op.data_bstride;
#pragma unroll
              for (int j = 0; j < kCube; ++j) c[j] = cmul(expmi(th[gidx(j)]), c[j]);
            }
            break;
          }
          default:
            break;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kCube; ++j) {
      const int e = e0 | ((j & 1) ? m0 : 0) | ((j & 2) ? m1 : 0) | ((j & 4) ? m2 : 0) |
                    ((j & 8) ? m3 : 0);
      my_tile[lds_pos(e)] = c[j];
    }
  }
  __syncthreads();

  // ---- coalesced store (same element order as the load)
#pragma unroll
  for (int k = 0; k < kTile / 2 / kThreads; ++k) {
    const int f = 2 * tid + k * 2 * kThreads;
    const int sub2 = f >> m;
    const int e = f & ((1 << m) - 1);
    const int64_t tg2 = int64_t(blockIdx.x) * tiles_per_block + sub2;
    if (tg2 < a.total_tiles) {
      uint64_t dep_k = 0;
      for (int i = 9; i < m; ++i) dep_k |= uint64_t((e >> i) & 1) << a.local[i];
      const uint64_t g = (tiles_per_block_log2 == 0 ? origin1 : tile_origin(tg2)) | dep_t |
                         dep_k | uint64_t(e_low);
      const int64_t b2 = tg2 >> a.log2_tiles_per_state;
      Pair<T> v;
      v.v[0] = tile[(sub2 << m) + lds_pos(e)];
      v.v[1] = tile[(sub2 << m) + lds_pos(e + 1)];
      *reinterpret_cast<Pair<T>*>(state + b2 * N + g) = v;
    }
  }
}

}  // namespace

int fused_pass_impl(void* state, int dtype, int n, int64_t batch, const ua_fused_pass* pass,
                    const ua_fused_op* ops, const ua_fused_op* host_ops, hipStream_t s) {
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128,
               "fused passes run on complex64/complex128 states");
  UA_CHECK_ARG(pass != nullptr, "NULL pass");
  const int m = pass->tile_bits;
  UA_CHECK_ARG(n >= 4 && n <= 62, "fused passes need 4..62 qubits (got %d)", n);
  UA_CHECK_ARG(m >= 4 && m <= kTileBits && m <= n, "tile_bits=%d invalid for %d qubits", m, n);
  UA_CHECK_ARG(pass->num_stages >= 0 && pass->num_stages <= UA_FUSED_MAX_STAGES,
               "num_stages=%d out of range", pass->num_stages);
  UA_CHECK_ARG(batch >= 0, "negative batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(state && (pass->num_stages == 0 || ops), "NULL pointer");
  UA_CHECK_ARG(reinterpret_cast<uintptr_t>(state) % (2 * elem_size(dtype)) == 0,
               "state must be aligned to two elements");
  UA_CHECK_ARG(pass->stage_ops[0] == 0, "stage_ops[0] must be 0");
  if (host_ops) {
    const int nops = pass->stage_ops[pass->num_stages];
    // a layer op's sub-ops must be DENSE1 on the counted cube bits, inside its stage
    for (int st = 0; st < pass->num_stages; ++st)
      for (int o = pass->stage_ops[st]; o < pass->stage_ops[st + 1]; ++o) {
        if (host_ops[o].kind != UA_FOP_LAYER1Q) continue;
        const uint32_t counts = uint32_t(host_ops[o].a);
        int sub = o + 1;
        for (int bit = 0; bit < 4; ++bit)
          for (uint32_t t = 0; t < ((counts >> (8 * bit)) & 255u); ++t, ++sub)
            UA_CHECK_ARG(sub < pass->stage_ops[st + 1] && host_ops[sub].kind == UA_FOP_DENSE1 &&
                             host_ops[sub].a == bit,
                         "op %d: layer sub-op %d is not DENSE1 on cube bit %d of its stage", o,
                         sub, bit);
        o = sub - 1;
      }
    for (int o = 0; o < nops; ++o) {
      const ua_fused_op& op = host_ops[o];
      UA_CHECK_ARG(op.kind >= UA_FOP_DENSE1 && op.kind <= UA_FOP_LAYER1Q, "op %d: bad kind", o);
      if (op.kind == UA_FOP_DENSE1 || op.kind == UA_FOP_CNOT)
        UA_CHECK_ARG(op.a >= 0 && op.a < (op.kind == UA_FOP_CNOT ? n : 4) &&
                         (op.kind == UA_FOP_DENSE1 || (op.b >= 0 && op.b < 4)),
                     "op %d: bad bits", o);
      if (op.kind == UA_FOP_DENSE2)
        UA_CHECK_ARG(op.a >= 0 && op.a < 4 && op.b >= 0 && op.b < 4 && op.a != op.b,
                     "op %d: bad cube bits", o);
      if (op.kind == UA_FOP_CZ)
        UA_CHECK_ARG(op.a >= 0 && op.a < n && op.b >= 0 && op.b < n, "op %d: bad bits", o);
      if (op.kind == UA_FOP_DIAG1) UA_CHECK_ARG(op.a >= 0 && op.a < n, "op %d: bad bit", o);
      if (op.cache_slot >= 0) {
        const int len = op.kind == UA_FOP_DENSE1 ? 4 : op.kind == UA_FOP_DENSE2 ? 16 : 2;
        UA_CHECK_ARG(op.kind != UA_FOP_CZ && op.kind != UA_FOP_CNOT && op.kind != UA_FOP_PHASE &&
                         op.kind != UA_FOP_CZ_LAYER && op.kind != UA_FOP_LAYER1Q &&
                         op.data_bstride == 0 && op.cache_slot + len <= kMatCache,
                     "op %d: bad cache slot %d", o, op.cache_slot);
      }
      UA_CHECK_ARG(op.data || op.kind == UA_FOP_CZ || op.kind == UA_FOP_CNOT ||
                       op.kind == UA_FOP_LAYER1Q,
                   "op %d: NULL data", o);
    }
  }
  PassArgs a{};
  a.n = n;
  a.m = m;
  a.log2_tiles_per_state = n - m;
  a.total_tiles = batch << (n - m);
  a.num_stages = pass->num_stages;
  a.needs_index = pass->needs_index ? 1 : 0;
  for (int i = 0; i < m; ++i) {
    a.local[i] = pass->local_bits[i];
    UA_CHECK_ARG(a.local[i] >= 0 && a.local[i] < n, "local bit %d out of range", a.local[i]);
    UA_CHECK_ARG(i >= 4 ? a.local[i] > a.local[i - 1] : a.local[i] == i,
                 "local bits must be ascending and start 0,1,2,3");
  }
  for (int st = 0; st < a.num_stages; ++st) {
    uint32_t seen = 0;
    for (int k = 0; k < 4; ++k) {
      const int t = pass->stage_bits[st][k];
      UA_CHECK_ARG(t >= 0 && t < m && !(seen & (1u << t)), "stage %d: bad cube bits", st);
      UA_CHECK_ARG(k == 0 || t > pass->stage_bits[st][k - 1], "stage %d: cube bits unsorted", st);
      seen |= 1u << t;
      a.stage_bits[st][k] = t;
    }
    a.stage_ops[st] = pass->stage_ops[st];
    UA_CHECK_ARG(pass->stage_ops[st + 1] >= pass->stage_ops[st], "stage %d: bad op range", st);
  }
  a.stage_ops[a.num_stages] = pass->stage_ops[a.num_stages];
  const int64_t tiles_per_block = int64_t(kTile) >> m;
  const int64_t blocks = (a.total_tiles + tiles_per_block - 1) / tiles_per_block;
  UA_CHECK_ARG(blocks <= 0x7fffffffLL, "too many tiles");
  // the variant follows the host op list; without it, the full kernel
  int variant = kFull;
  if (host_ops) {
    bool heavy = false, only_layers = true;
    for (int o = 0; o < pass->stage_ops[pass->num_stages]; ++o) {
      const int kind = host_ops[o].kind;
      heavy = heavy || kind == UA_FOP_PHASE || kind == UA_FOP_DENSE2;
      if (kind == UA_FOP_LAYER1Q) {
        const uint32_t counts = uint32_t(host_ops[o].a);
        o += int((counts & 255u) + ((counts >> 8) & 255u) + ((counts >> 16) & 255u) +
                 ((counts >> 24) & 255u));   // skip its sub-ops (validated above)
      } else {
        only_layers = false;
      }
    }
    variant = only_layers ? kLayers : (heavy ? kFull : kLight);
  }
#define UA_FUSED_LAUNCH(T, V)                                                     \
  fused_pass_kernel<T, V><<<unsigned(blocks), kThreads, 0, s>>>(static_cast<T*>(state), ops, a)
  if (dtype == UA_COMPLEX64) {
    if (variant == kLayers) UA_FUSED_LAUNCH(float2, kLayers);
    else if (variant == kLight) UA_FUSED_LAUNCH(float2, kLight);
    else UA_FUSED_LAUNCH(float2, kFull);
  } else {
    if (variant == kLayers) UA_FUSED_LAUNCH(double2, kLayers);
    else if (variant == kLight) UA_FUSED_LAUNCH(double2, kLight);
    else UA_FUSED_LAUNCH(double2, kFull);
  }
#undef UA_FUSED_LAUNCH
  return check_launch("fused_pass_kernel");
}

}  // namespace ua

extern "C" int ua_run_fused_pass(void* state, int dtype, int num_qubits, int64_t batch,
                                 const ua_fused_pass* pass, const ua_fused_op* ops,
                                 const ua_fused_op* host_ops, ua_stream_t stream) {
  return ua::fused_pass_impl(state, dtype, num_qubits, batch, pass, ops, host_ops,
                             static_cast<hipStream_t>(stream));
}
