// Fused local pass: many gates per read/write of the state.
//
// The reference applies every gate as its own pass over the state
// (unitair/simulation/operations.py:151-186); its only fusion is multiplying
// 1-qubit gates that hit the same qubit (unitair/gates/matrix_algebra.py:6-51,
// used by apply_to_qubits, operations.py:416-503).  Here a pass picks m <= 12
// "local" index bits (always including bits 0..3, so every tile is made of
// 128 B runs).  A workgroup loads one 2^m-amplitude tile, applies a list of
// gates, and writes the tile back:
//   - dense 1-/2-qubit gates whose targets are local bits,
//   - CNOT whose target is local (the control may be any bit),
//   - CZ, 1-qubit diagonals, CZ layers and full-length phases on any bits
//     (they act element by element, so they need no local bits at all).
// The gate list is cut into stages of two kinds:
//   cube stage:     every thread owns a "cube" of 16 amplitudes spanning 4
//                   tile bits, so the stage's gates run in registers;
//   diagonal stage: every thread owns the 16 amplitudes it loaded (8 pairs
//                   in tile order), so per-element data such as phase
//                   factors is read with the same coalesced pattern as the
//                   state.
// Only stage boundaries go through LDS; a diagonal stage at the start or the
// end of a pass runs on the registers of the global load or store.
//
// Gate data shared by the workgroup (matrices, angle rows of one state) is
// read through the constant address space at wave-uniform addresses, i.e.
// as scalar loads into SGPRs, so gate operands cost no VGPRs and no LDS.
// LDS addresses of a cube are one per-thread base XOR a wave-uniform offset
// (the swizzle is linear over GF(2)), one VALU op per element.
#include "ua_common.h"
#include "fused_common.h"

namespace ua {
namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kCube = 16;
constexpr int kThreads = kTile / kCube;        // 256
constexpr int kPairs = kTile / 2 / kThreads;   // 8 pairs per thread in tile order

#define UA_SCALAR __attribute__((address_space(4)))

// LDS position of tile element e: XOR-swizzled so that the 32 lanes of a
// half-wavefront reading one cube entry hit (nearly) distinct banks for every
// choice of cube bits.  Linear over GF(2): pos(x ^ y) = pos(x) ^ pos(y).
__device__ __forceinline__ int lds_pos(int e) { return e ^ ((e >> 5) & 31); }

// An op record read field by field through the constant address space.
__device__ __forceinline__ ua_fused_op load_op(const UA_SCALAR ua_fused_op* p) {
  ua_fused_op r;
  r.kind = p->kind;
  r.a = p->a;
  r.b = p->b;
  r.cache_slot = -1;
  r.data = p->data;
  r.data_bstride = p->data_bstride;
  return r;
}

// The value unchanged, but opaque to the optimiser (no CSE across it).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

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
template <typename T>
__device__ __forceinline__ T cmul2(T a, T b) {  // a*b as the first term of a sum
  T r;
  r.x = a.x * b.x;
  r.x = fma(-a.y, b.y, r.x);
  r.y = a.x * b.y;
  r.y = fma(a.y, b.x, r.y);
  return r;
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
template <typename T>
__device__ __forceinline__ T cneg(T a) {
  a.x = -a.x;
  a.y = -a.y;
  return a;
}

// Complex element i of gate data (interleaved re, im) through pointer p.
template <typename T, typename P>
__device__ __forceinline__ T ldc(P p, int i) {
  T r;
  r.x = p[2 * i];
  r.y = p[2 * i + 1];
  return r;
}

// 1-qubit dense gate on cube bit B.
template <int B, typename T>
__device__ __forceinline__ void cube_1q(T (&c)[kCube], const T u00, const T u01, const T u10,
                                        const T u11) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & (1 << B)) continue;
    const int k = j | (1 << B);
    const T x = c[j], y = c[k];
    c[j] = cmad(cmul2(u00, x), u01, y);
    c[k] = cmad(cmul2(u10, x), u11, y);
  }
}

template <typename T, typename P>
__device__ __forceinline__ void dispatch_1q(T (&c)[kCube], int bit, P u) {
  const T u00 = ldc<T>(u, 0), u01 = ldc<T>(u, 1), u10 = ldc<T>(u, 2), u11 = ldc<T>(u, 3);
  switch (bit) {
    case 0: cube_1q<0>(c, u00, u01, u10, u11); break;
    case 1: cube_1q<1>(c, u00, u01, u10, u11); break;
    case 2: cube_1q<2>(c, u00, u01, u10, u11); break;
    default: cube_1q<3>(c, u00, u01, u10, u11); break;
  }
}

// 2-qubit dense gate; gate index bit 1 = cube bit B1, bit 0 = cube bit B0.
template <int B1, int B0, typename T, typename P>
__device__ __forceinline__ void cube_2q(T (&c)[kCube], P u) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & ((1 << B1) | (1 << B0))) continue;
    const int idx[4] = {j, j | (1 << B0), j | (1 << B1), j | (1 << B1) | (1 << B0)};
    T x[4], o[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = c[idx[s]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T acc = cmul2(ldc<T>(u, r * 4), x[0]);
#pragma unroll
      for (int s = 1; s < 4; ++s) acc = cmad(acc, ldc<T>(u, r * 4 + s), x[s]);
      o[r] = acc;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) c[idx[s]] = o[s];
  }
}

template <typename T, typename P>
__device__ __forceinline__ void dispatch_2q(T (&c)[kCube], int b1, int b0, P u) {
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

// Component-wise s ? a : b (a select of two vector structs is lowered to a
// select of their addresses, which keeps the cube array out of registers).
template <typename T>
__device__ __forceinline__ T csel(bool s, T a, T b) {
  T r;
  r.x = s ? a.x : b.x;
  r.y = s ? a.y : b.y;
  return r;
}

// ---- sign, swap and diagonal steps specialised by where their qubits sit.
// Both qubits in the cube give a fixed pattern of entries (no per-entry test);
// a qubit outside the cube is one bit per thread, so the step is one factor
// or one select for the whole cube.

// negate the entries whose cube bits A and B are both set
template <int A, int B, typename T>
__device__ __forceinline__ void negate_cc(T (&c)[kCube]) {
#pragma unroll
  for (int j = 0; j < kCube; ++j)
    if (((j >> A) & 1) && ((j >> B) & 1)) c[j] = cneg(c[j]);
}
// multiply the entries whose cube bit A is set by s (+1 or -1)
template <int A, typename T, typename R>
__device__ __forceinline__ void scale_c(T (&c)[kCube], R s) {
#pragma unroll
  for (int j = 0; j < kCube; ++j)
    if ((j >> A) & 1) {
      c[j].x *= s;
      c[j].y *= s;
    }
}
template <typename T, typename R>
__device__ __forceinline__ void scale_all(T (&c)[kCube], R s) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    c[j].x *= s;
    c[j].y *= s;
  }
}
// CNOT with the control on cube bit C and the target on cube bit B: swap the
// pairs whose control bit is set
template <int C, int B, typename T>
__device__ __forceinline__ void cnot_cc(T (&c)[kCube]) {
#pragma unroll
  for (int j = 0; j < kCube; ++j)
    if (((j >> C) & 1) && !((j >> B) & 1)) {
      const T x = c[j];
      c[j] = c[j | (1 << B)];
      c[j | (1 << B)] = x;
    }
}
// CNOT with the control outside the cube (this thread's bit s): swap every
// pair along cube bit B, or none
template <int B, typename T>
__device__ __forceinline__ void cnot_lane(T (&c)[kCube], bool s) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) {
    if (j & (1 << B)) continue;
    const T x = c[j], y = c[j | (1 << B)];
    c[j] = csel(s, y, x);
    c[j | (1 << B)] = csel(s, x, y);
  }
}
// diag(d0, d1) on cube bit A
template <int A, typename T>
__device__ __forceinline__ void diag_c(T (&c)[kCube], T d0, T d1) {
#pragma unroll
  for (int j = 0; j < kCube; ++j) c[j] = cmul(((j >> A) & 1) ? d1 : d0, c[j]);
}

template <typename T>
__device__ __forceinline__ void negate_where(T (&c)[kCube], uint32_t neg) {
#pragma unroll
  for (int j = 0; j < kCube; ++j)
    if ((neg >> j) & 1) c[j] = cneg(c[j]);
}

struct PassArgs {
  int64_t total_tiles;      // batch * 2^(n-m)
  int n, m;
  int log2_tiles_per_state; // n - m
  int num_stages;
  int needs_index;          // some cube stage has an op that reads the global index
  int reg_head, reg_tail;   // the first / last stage is diagonal with PHASE_FACTORS ops
                            // only: it runs on the registers of the global load / store
  int local[kTileBits];     // state bit of tile bit i (ascending; local[i] = i for i < 4)
  int stage_bits[UA_FUSED_MAX_STAGES][4];  // -1: diagonal stage
  int stage_ops[UA_FUSED_MAX_STAGES + 1];
};

// Kernel variants by the op kinds a pass uses:
//   kFull   every kind;
//   kLight  no 2-qubit matrices;
//   kLayers only 1-qubit layer ops, no diagonal stages.
// MULTI: a workgroup holds several tiles (m < 12), so the state index and
// the addresses of per-state gate data differ between threads (vector loads);
// otherwise one tile per workgroup, every gate operand address is uniform.
enum { kLayers = 0, kLight = 1, kFull = 2 };

// Occupancy: a 2^12-amplitude tile takes 32 KiB of LDS (complex64), so up to
// five workgroups (waves per SIMD) fit a CU; the register budget asks for
// UA_FUSED_WAVES of them (at most 128 VGPRs for four).  complex128 tiles
// (64 KiB) allow two.
#ifndef UA_FUSED_WAVES
#define UA_FUSED_WAVES 4
#endif
template <typename T>
constexpr int waves_per_simd() { return sizeof(T) == 8 ? UA_FUSED_WAVES : 2; }

template <typename T, int VARIANT, bool MULTI>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(waves_per_simd<T>()))) void fused_pass_kernel(
    T* __restrict__ state, const ua_fused_op* __restrict__ ops_arg, PassArgs a) {
  using R = typename RealOf<T>::type;
  __shared__ T tile[kTile];
  const UA_SCALAR ua_fused_op* ops = (const UA_SCALAR ua_fused_op*)ops_arg;
  const int m = MULTI ? a.m : kTileBits;
  const int tpb_log2 = kTileBits - m;  // tiles per workgroup, log2
  const int tid = threadIdx.x;
  const int64_t N = int64_t(1) << a.n;

  auto tile_origin = [&](int64_t tg) -> uint64_t {
    const uint64_t tis = uint64_t(tg) & ((uint64_t(1) << a.log2_tiles_per_state) - 1);
    return insert_zero_bits(tis, a.local, m);
  };
  // Pair k of this thread in tile order: block element f = 2*t + 512*k,
  // tile sub = f >> m, element e = f & (2^m - 1).  The state index deposits
  // e's bits at the local bit positions: bits 0..3 stay, bits 4..8 come from
  // t (a per-thread part), bits 9.. from k (a wave-uniform part).  With one
  // tile per workgroup the LDS position splits the same way:
  // pos(2t + 512k) = pos(2t) ^ 512k ^ 16(k & 1).
  // `t` is the thread id; callers far apart pass it through opaque() so
  // that the compiler recomputes these few values instead of keeping every
  // pair's address live in registers across the whole pass.
  const uint64_t origin1 = MULTI ? 0 : tile_origin(blockIdx.x);
  struct PairAt {
    uint64_t g;  // state index of the pair's first element
    int64_t b;   // state of the tile
    int lds;     // LDS position of the first element (the second is lds ^ 1)
    bool valid;
  };
  struct ThreadPart {
    uint64_t dep;  // deposit of bits 0..8 of 2t
    int lds;       // pos(2t)
  };
  auto thread_part = [&](int t) -> ThreadPart {
    ThreadPart tp;
    tp.dep = uint64_t((2 * t) & 15);
    for (int i = 4; i < 9 && i < m; ++i) tp.dep |= uint64_t(((2 * t) >> i) & 1) << a.local[i];
    tp.lds = lds_pos(2 * t);
    return tp;
  };
  auto pair_at = [&](const ThreadPart& tp, int t, int k) -> PairAt {
    PairAt p;
    if constexpr (MULTI) {
      const int f = 2 * t + 2 * kThreads * k;
      const int sub = f >> m;
      const int e = f & ((1 << m) - 1);
      uint64_t dep_k = 0;
      for (int i = 9; i < m; ++i) dep_k |= uint64_t((e >> i) & 1) << a.local[i];
      const int64_t tg = (int64_t(blockIdx.x) << tpb_log2) + sub;
      p.valid = tg < a.total_tiles;
      p.g = (p.valid ? tile_origin(tg) : 0) | dep_k | tp.dep;
      p.b = p.valid ? (tg >> a.log2_tiles_per_state) : 0;
      p.lds = (sub << m) + lds_pos(e);
    } else {
      uint64_t dep_k = 0;
#pragma unroll
      for (int i = 9; i < kTileBits; ++i) dep_k |= uint64_t((k >> (i - 9)) & 1) << a.local[i];
      p.valid = true;
      p.g = origin1 | dep_k | tp.dep;
      p.b = int64_t(blockIdx.x) >> a.log2_tiles_per_state;
      p.lds = tp.lds ^ (2 * kThreads * k) ^ (16 * (k & 1));
    }
    return p;
  };

  // Gate data of state b (data_bstride complex elements apart), as reals.
  // MULTI: a per-thread (vector) pointer; otherwise a uniform constant-space
  // pointer (scalar loads).
  auto rows = [&](const ua_fused_op& op, int64_t b) {
    if constexpr (MULTI)
      return static_cast<const R*>(op.data) + 2 * b * op.data_bstride;
    else
      return (const UA_SCALAR R*)op.data + 2 * b * op.data_bstride;
  };

  // One diagonal op on one pair (x0, x1) of state indices g, g | 1.
  auto diag_pair = [&](const ua_fused_op& op, const PairAt& p, T& x0, T& x1) {
    const uint64_t g0 = p.g, g1 = p.g | 1;
    switch (op.kind) {
      case UA_FOP_PHASE_FACTORS: {  // data: 2^n factors per state, state dtype
        const T* f = static_cast<const T*>(op.data) + p.b * op.data_bstride;
        const Pair<T> v = *reinterpret_cast<const Pair<T>*>(f + p.g);
        x0 = cmul(v.v[0], x0);
        x1 = cmul(v.v[1], x1);
        break;
      }
      case UA_FOP_PHASE: {  // data: 2^n real angles per state, data_bstride in angles
        const R* th = static_cast<const R*>(op.data) + p.b * op.data_bstride;
        const Pair<R> t = *reinterpret_cast<const Pair<R>*>(th + p.g);
        x0 = cmul(expmi(t.v[0]), x0);
        x1 = cmul(expmi(t.v[1]), x1);
        break;
      }
      case UA_FOP_CZ: {
        if ((g0 >> op.a) & (g0 >> op.b) & 1) x0 = cneg(x0);
        if ((g1 >> op.a) & (g1 >> op.b) & 1) x1 = cneg(x1);
        break;
      }
      case UA_FOP_DIAG1: {
        const R* d = static_cast<const R*>(op.data) + 2 * p.b * op.data_bstride;
        const T d0 = ldc<T>(d, 0), d1 = ldc<T>(d, 1);
        x0 = cmul(((g0 >> op.a) & 1) ? d1 : d0, x0);
        x1 = cmul(((g1 >> op.a) & 1) ? d1 : d0, x1);
        break;
      }
      case UA_FOP_CZ_LAYER: {
        const uint64_t mask = static_cast<const uint64_t*>(op.data)[p.b * op.data_bstride];
        if (__popcll(g0 & (g0 >> 1) & mask) & 1) x0 = cneg(x0);
        if (__popcll(g1 & (g1 >> 1) & mask) & 1) x1 = cneg(x1);
        break;
      }
      default:
        break;
    }
  };

  // Diagonal stage in tile order, LDS to LDS, one pair at a time: rolled
  // loops keep the register footprint of sincos and index work small.
  auto diag_stage = [&](int o0, int o1) {
    if constexpr (VARIANT != kLayers) {
      const int t = opaque(tid);
      const ThreadPart tp = thread_part(t);
#pragma unroll 1
      for (int k = 0; k < kPairs; ++k) {
        const PairAt p = pair_at(tp, t, k);
        if (!p.valid) continue;
        T x0 = tile[p.lds], x1 = tile[p.lds ^ 1];
#pragma unroll 1
        for (int o = o0; o < o1; ++o) diag_pair(load_op(ops + o), p, x0, x1);
        tile[p.lds] = x0;
        tile[p.lds ^ 1] = x1;
      }
    }
  };

  // Diagonal stage of PHASE_FACTORS ops only, on the registers of the
  // global load or store (a.reg_head / a.reg_tail are set only then): the
  // factor pairs of an op are all in flight at once.
  auto diag_regs = [&](T (&c)[kCube], int t, int o0, int o1) {
    if constexpr (VARIANT != kLayers) {
      const ThreadPart tp = thread_part(t);
#pragma unroll 1
      for (int o = o0; o < o1; ++o) {
        const ua_fused_op op = load_op(ops + o);
        const T* f = static_cast<const T*>(op.data);
#pragma unroll
        for (int k = 0; k < kPairs; ++k) {
          const PairAt p = pair_at(tp, t, k);
          if (p.valid) {
            const Pair<T> v =
                *reinterpret_cast<const Pair<T>*>(f + p.b * op.data_bstride + p.g);
            c[2 * k] = cmul(v.v[0], c[2 * k]);
            c[2 * k + 1] = cmul(v.v[1], c[2 * k + 1]);
          }
        }
      }
    }
  };

  // ---- coalesced load of this workgroup's tiles (pairs of amplitudes)
  T c[kCube];
  {
    const ThreadPart tp = thread_part(tid);
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const PairAt p = pair_at(tp, tid, k);
      if (p.valid) {
        const Pair<T> v = *reinterpret_cast<const Pair<T>*>(state + p.b * N + p.g);
        c[2 * k] = v.v[0];
        c[2 * k + 1] = v.v[1];
      } else {
        c[2 * k].x = c[2 * k].y = 0;
        c[2 * k + 1] = c[2 * k];
      }
    }
  }
  auto store = [&](int t) {
    const ThreadPart tp = thread_part(t);
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const PairAt p = pair_at(tp, t, k);
      if (p.valid) {
        Pair<T> v;
        v.v[0] = c[2 * k];
        v.v[1] = c[2 * k + 1];
        *reinterpret_cast<Pair<T>*>(state + p.b * N + p.g) = v;
      }
    }
  };
  // a leading / trailing diagonal stage runs on these registers
  const int st_begin = a.reg_head ? 1 : 0;
  const int st_end = a.num_stages - ((a.reg_tail && a.num_stages > st_begin) ? 1 : 0);
  const bool tail = st_end < a.num_stages;
  if (a.reg_head) diag_regs(c, tid, a.stage_ops[0], a.stage_ops[1]);
  if (st_begin >= st_end) {  // nothing goes through LDS
    if (tail) diag_regs(c, tid, a.stage_ops[a.num_stages - 1], a.stage_ops[a.num_stages]);
    store(tid);
    return;
  }
  {
    const ThreadPart tp = thread_part(tid);
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const PairAt p = pair_at(tp, tid, k);
      tile[p.lds] = c[2 * k];
      tile[p.lds ^ 1] = c[2 * k + 1];
    }
  }

  // ---- this thread's place in its tile for cube stages
  const int t_local = MULTI ? (tid & ((1 << (m - 4)) - 1)) : tid;
  const int sub = MULTI ? (tid >> (m - 4)) : 0;
  const int64_t tg_c = (int64_t(blockIdx.x) << tpb_log2) + sub;
  const bool active = tg_c < a.total_tiles;
  const int64_t b_c = active ? (tg_c >> a.log2_tiles_per_state) : 0;
  uint64_t tile_base = 0;
  if (a.needs_index && active) tile_base = MULTI ? tile_origin(tg_c) : origin1;
  char* const tile_bytes = reinterpret_cast<char*>(tile);

  for (int st = st_begin; st < st_end; ++st) {
    __syncthreads();
    const int s0 = a.stage_bits[st][0];
    if (s0 < 0) {  // diagonal stage: the thread's own pairs, in tile order
      diag_stage(a.stage_ops[st], a.stage_ops[st + 1]);
      continue;
    }
    const int s1 = a.stage_bits[st][1], s2 = a.stage_bits[st][2], s3 = a.stage_bits[st][3];
    const int sb[4] = {s0, s1, s2, s3};
    const int e0 = int(insert_zero_bits(uint64_t(t_local), sb, 4));
    const int m0 = 1 << s0, m1 = 1 << s1, m2 = 1 << s2, m3 = 1 << s3;
    // byte address of cube entry j: base ^ off(j), off(j) wave-uniform
    const int base = ((sub << m) + lds_pos(e0)) * int(sizeof(T));
    auto off = [&](int j) -> int {
      return lds_pos(((j & 1) ? m0 : 0) | ((j & 2) ? m1 : 0) | ((j & 4) ? m2 : 0) |
                     ((j & 8) ? m3 : 0)) * int(sizeof(T));
    };
    T q[kCube];
#pragma unroll
    for (int j = 0; j < kCube; ++j) q[j] = *reinterpret_cast<const T*>(tile_bytes + (base ^ off(j)));
    // global index of cube entry j = gbase | (the cube part of j)
    uint64_t gbase = 0;
    if (a.needs_index) {
      gbase = tile_base;
      for (int i = 0; i < m; ++i) gbase |= uint64_t((e0 >> i) & 1) << a.local[i];
    }
    const int cb0 = a.local[s0], cb1 = a.local[s1], cb2 = a.local[s2], cb3 = a.local[s3];
    // cube position (0..3) of state bit `bit`, or -1 (wave-uniform)
    auto slot_of = [&](int bit) -> int {
      return bit == cb0 ? 0 : bit == cb1 ? 1 : bit == cb2 ? 2 : bit == cb3 ? 3 : -1;
    };
    // this thread's value of a state bit outside the cube
    auto bit_of = [&](int bit) -> bool { return (gbase >> bit) & 1; };
    if (active) {
      for (int o = a.stage_ops[st]; o < a.stage_ops[st + 1]; ++o) {
        const ua_fused_op op = load_op(ops + o);
        switch (op.kind) {
          case UA_FOP_DENSE1: {
            if constexpr (VARIANT != kLayers) dispatch_1q(q, op.a, rows(op, b_c));
            break;
          }
          case UA_FOP_DENSE2: {
            if constexpr (VARIANT == kFull) dispatch_2q(q, op.a, op.b, rows(op, b_c));
            break;
          }
          case UA_FOP_CNOT: {  // a: control state bit, b: target cube bit
            if constexpr (VARIANT != kLayers) {
              const int kc = slot_of(op.a);
              if (kc >= 0) {
                switch (kc * 4 + op.b) {
                  case 1: cnot_cc<0, 1>(q); break;
                  case 2: cnot_cc<0, 2>(q); break;
                  case 3: cnot_cc<0, 3>(q); break;
                  case 4: cnot_cc<1, 0>(q); break;
                  case 6: cnot_cc<1, 2>(q); break;
                  case 7: cnot_cc<1, 3>(q); break;
                  case 8: cnot_cc<2, 0>(q); break;
                  case 9: cnot_cc<2, 1>(q); break;
                  case 11: cnot_cc<2, 3>(q); break;
                  case 12: cnot_cc<3, 0>(q); break;
                  case 13: cnot_cc<3, 1>(q); break;
                  default: cnot_cc<3, 2>(q); break;
                }
              } else {
                const bool s = bit_of(op.a);
                switch (op.b) {
                  case 0: cnot_lane<0>(q, s); break;
                  case 1: cnot_lane<1>(q, s); break;
                  case 2: cnot_lane<2>(q, s); break;
                  default: cnot_lane<3>(q, s); break;
                }
              }
            }
            break;
          }
          case UA_FOP_CZ: {  // a, b: state bits
            if constexpr (VARIANT != kLayers) {
              const int ka = slot_of(op.a), kb = slot_of(op.b);
              if (ka >= 0 && kb >= 0) {
                switch ((1 << ka) | (1 << kb)) {
                  case 3: negate_cc<0, 1>(q); break;
                  case 5: negate_cc<0, 2>(q); break;
                  case 9: negate_cc<0, 3>(q); break;
                  case 6: negate_cc<1, 2>(q); break;
                  case 10: negate_cc<1, 3>(q); break;
                  default: negate_cc<2, 3>(q); break;
                }
              } else if (ka >= 0 || kb >= 0) {
                const int k = ka >= 0 ? ka : kb;
                const R sg = bit_of(ka >= 0 ? op.b : op.a) ? R(-1) : R(1);
                switch (k) {
                  case 0: scale_c<0>(q, sg); break;
                  case 1: scale_c<1>(q, sg); break;
                  case 2: scale_c<2>(q, sg); break;
                  default: scale_c<3>(q, sg); break;
                }
              } else {
                scale_all(q, (bit_of(op.a) && bit_of(op.b)) ? R(-1) : R(1));
              }
            }
            break;
          }
          case UA_FOP_DIAG1: {  // a: state bit; data: (d0, d1)
            if constexpr (VARIANT != kLayers) {
              const auto d = rows(op, b_c);
              const T d0 = ldc<T>(d, 0), d1 = ldc<T>(d, 1);
              const int k = slot_of(op.a);
              switch (k) {
                case 0: diag_c<0>(q, d0, d1); break;
                case 1: diag_c<1>(q, d0, d1); break;
                case 2: diag_c<2>(q, d0, d1); break;
                case 3: diag_c<3>(q, d0, d1); break;
                default: {
                  const T dd = csel(bit_of(op.a), d1, d0);
#pragma unroll
                  for (int j = 0; j < kCube; ++j) q[j] = cmul(dd, q[j]);
                }
              }
            }
            break;
          }
          case UA_FOP_CZ_LAYER: {  // data: one uint64 mask per state
            if constexpr (VARIANT != kLayers) {
              // Sign of entry j: parity of popcount(g & (g >> 1) & mask) with
              // g = gbase + cube part x (disjoint bits).  Over GF(2) it splits
              // into P(gbase), a term linear in the cube bits (a cube bit at p
              // pairs with gbase bits p-1 and p+1) and a term for cube bits
              // that are neighbours; built as a 16-bit mask.
              uint64_t mask;
              if constexpr (MULTI)
                mask = static_cast<const uint64_t*>(op.data)[b_c * op.data_bstride];
              else
                mask = ((const UA_SCALAR uint64_t*)op.data)[b_c * op.data_bstride];
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
              negate_where(q, neg);
            }
            break;
          }
          case UA_FOP_LAYER1Q: {  // a: four 8-bit counts; DENSE1 sub-ops follow
            const uint32_t counts = uint32_t(op.a);
            int so = o + 1;
#define UA_LAYER_BIT(B)                                                              \
            for (int t = 0, nb = int((counts >> (8 * B)) & 255u); t < nb; ++t, ++so) { \
              const ua_fused_op sop = load_op(ops + so);                                       \
              const auto u = rows(sop, b_c);                                         \
              cube_1q<B>(q, ldc<T>(u, 0), ldc<T>(u, 1), ldc<T>(u, 2), ldc<T>(u, 3)); \
            }
            UA_LAYER_BIT(0)
            UA_LAYER_BIT(1)
            UA_LAYER_BIT(2)
            UA_LAYER_BIT(3)
#undef UA_LAYER_BIT
            o = so - 1;
            break;
          }
          default:
            break;
        }
      }
    }
    const int base_w = opaque(base);  // recompute the addresses: no 16 live VGPRs
#pragma unroll
    for (int j = 0; j < kCube; ++j) *reinterpret_cast<T*>(tile_bytes + (base_w ^ off(j))) = q[j];
  }

  // ---- coalesced store (same element order as the load)
  __syncthreads();
  const int t_end = opaque(tid);
  {
    const ThreadPart tp = thread_part(t_end);
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const PairAt p = pair_at(tp, t_end, k);
      c[2 * k] = tile[p.lds];
      c[2 * k + 1] = tile[p.lds ^ 1];
    }
  }
  if (tail) diag_regs(c, t_end, a.stage_ops[a.num_stages - 1], a.stage_ops[a.num_stages]);
  store(t_end);
}

}  // namespace

bool is_diag_kind(int kind) {
  return kind == UA_FOP_CZ || kind == UA_FOP_DIAG1 || kind == UA_FOP_PHASE ||
         kind == UA_FOP_CZ_LAYER || kind == UA_FOP_PHASE_FACTORS;
}

int check_fused_pass(int dtype, int n, const ua_fused_pass* pass, const ua_fused_op* host_ops,
                     FusedPassInfo* info, int max_tile_bits) {
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128,
               "fused passes run on complex64/complex128 states");
  UA_CHECK_ARG(pass != nullptr, "NULL pass");
  const int m = pass->tile_bits;
  UA_CHECK_ARG(n >= 4 && n <= 62, "fused passes need 4..62 qubits (got %d)", n);
  UA_CHECK_ARG(m >= 4 && m <= max_tile_bits && m <= n, "tile_bits=%d invalid for %d qubits", m,
               n);
  UA_CHECK_ARG(pass->num_stages >= 0 && pass->num_stages <= UA_FUSED_MAX_STAGES,
               "num_stages=%d out of range", pass->num_stages);
  UA_CHECK_ARG(host_ops != nullptr || pass->num_stages == 0,
               "host_ops (the host copy of ops) is required");
  UA_CHECK_ARG(pass->stage_ops[0] == 0, "stage_ops[0] must be 0");
  const size_t rsize = elem_size(dtype) / 2;
  for (int i = 0; i < m; ++i) {
    const int l = pass->local_bits[i];
    UA_CHECK_ARG(l >= 0 && l < n, "local bit %d out of range", l);
    UA_CHECK_ARG(i >= 4 ? l > pass->local_bits[i - 1] : l == i,
                 "local bits must be ascending and start 0,1,2,3");
  }
  bool heavy = false, only_layers = true, needs_index = false;
  bool stage_not_factors[UA_FUSED_MAX_STAGES] = {};
  for (int st = 0; st < pass->num_stages; ++st) {
    UA_CHECK_ARG(pass->stage_ops[st + 1] >= pass->stage_ops[st], "stage %d: bad op range", st);
    const bool diag = pass->stage_bits[st][0] < 0;
    if (diag) {
      for (int k = 0; k < 4; ++k)
        UA_CHECK_ARG(pass->stage_bits[st][k] == -1, "stage %d: a diagonal stage has cube bits "
                     "{-1, -1, -1, -1}", st);
      only_layers = false;
    } else {
      uint32_t seen = 0;
      for (int k = 0; k < 4; ++k) {
        const int t = pass->stage_bits[st][k];
        UA_CHECK_ARG(t >= 0 && t < m && !(seen & (1u << t)), "stage %d: bad cube bits", st);
        UA_CHECK_ARG(k == 0 || t > pass->stage_bits[st][k - 1], "stage %d: cube bits unsorted",
                     st);
        seen |= 1u << t;
      }
    }
    for (int o = pass->stage_ops[st]; o < pass->stage_ops[st + 1]; ++o) {
      const ua_fused_op& op = host_ops[o];
      UA_CHECK_ARG(op.kind >= UA_FOP_DENSE1 && op.kind <= UA_FOP_PHASE_FACTORS,
                   "op %d: bad kind", o);
      if (diag) {
        UA_CHECK_ARG(is_diag_kind(op.kind), "op %d: a diagonal stage takes CZ, DIAG1, PHASE, "
                     "PHASE_FACTORS and CZ_LAYER ops", o);
        if (op.kind != UA_FOP_PHASE_FACTORS) stage_not_factors[st] = true;
      } else {
        UA_CHECK_ARG(op.kind != UA_FOP_PHASE && op.kind != UA_FOP_PHASE_FACTORS,
                     "op %d: PHASE ops run in diagonal stages", o);
        needs_index = needs_index || (op.kind != UA_FOP_DENSE1 && op.kind != UA_FOP_DENSE2 &&
                                      op.kind != UA_FOP_LAYER1Q);
      }
      if (op.kind == UA_FOP_LAYER1Q) {
        // its sub-ops: DENSE1 on the counted cube bits, inside its stage
        const uint32_t counts = uint32_t(op.a);
        int sub = o + 1;
        for (int bit = 0; bit < 4; ++bit)
          for (uint32_t t = 0; t < ((counts >> (8 * bit)) & 255u); ++t, ++sub)
            UA_CHECK_ARG(sub < pass->stage_ops[st + 1] && host_ops[sub].kind == UA_FOP_DENSE1 &&
                             host_ops[sub].a == bit,
                         "op %d: layer sub-op %d is not DENSE1 on cube bit %d of its stage", o,
                         sub, bit);
        o = sub - 1;  // sub-ops validated here, as DENSE1 ops
        for (int so = o - int((counts & 255u) + ((counts >> 8) & 255u) +
                              ((counts >> 16) & 255u) + ((counts >> 24) & 255u)) + 1;
             so <= o; ++so)
          UA_CHECK_ARG(host_ops[so].data != nullptr, "op %d: NULL data", so);
        continue;
      }
      only_layers = false;
      heavy = heavy || op.kind == UA_FOP_DENSE2;
      if (op.kind == UA_FOP_DENSE1 || op.kind == UA_FOP_CNOT)
        UA_CHECK_ARG(op.a >= 0 && op.a < (op.kind == UA_FOP_CNOT ? n : 4) &&
                         (op.kind == UA_FOP_DENSE1 || (op.b >= 0 && op.b < 4)),
                     "op %d: bad bits", o);
      if (op.kind == UA_FOP_DENSE2)
        UA_CHECK_ARG(op.a >= 0 && op.a < 4 && op.b >= 0 && op.b < 4 && op.a != op.b,
                     "op %d: bad cube bits", o);
      if (op.kind == UA_FOP_CZ)
        UA_CHECK_ARG(op.a >= 0 && op.a < n && op.b >= 0 && op.b < n && op.a != op.b,
                     "op %d: bad bits", o);
      if (op.kind == UA_FOP_DIAG1) UA_CHECK_ARG(op.a >= 0 && op.a < n, "op %d: bad bit", o);
      UA_CHECK_ARG(op.data_bstride >= 0, "op %d: negative data_bstride", o);
      UA_CHECK_ARG(op.data || op.kind == UA_FOP_CZ || op.kind == UA_FOP_CNOT,
                   "op %d: NULL data", o);
      if (op.kind == UA_FOP_PHASE)
        UA_CHECK_ARG(reinterpret_cast<uintptr_t>(op.data) % (2 * rsize) == 0 &&
                         op.data_bstride % 2 == 0,
                     "op %d: phase angles must be aligned to two angles", o);
      if (op.kind == UA_FOP_PHASE_FACTORS)
        UA_CHECK_ARG(reinterpret_cast<uintptr_t>(op.data) % (4 * rsize) == 0 &&
                         op.data_bstride % 2 == 0,
                     "op %d: phase factors must be aligned to two elements", o);
    }
  }
  const int ns = pass->num_stages;
  info->heavy = heavy;
  info->only_layers = only_layers;
  info->needs_index = needs_index;
  info->reg_head = ns > 0 && pass->stage_bits[0][0] < 0 && !stage_not_factors[0];
  info->reg_tail = ns > 0 && pass->stage_bits[ns - 1][0] < 0 && !stage_not_factors[ns - 1];
  return UA_OK;
}

int fused_pass_impl(void* state, int dtype, int n, int64_t batch, const ua_fused_pass* pass,
                    const ua_fused_op* ops, const ua_fused_op* host_ops, hipStream_t s) {
  FusedPassInfo info{};
  const int rc = check_fused_pass(dtype, n, pass, host_ops, &info, kTileBits);
  if (rc != UA_OK) return rc;
  UA_CHECK_ARG(batch >= 0, "negative batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(state && (pass->num_stages == 0 || ops), "NULL pointer");
  UA_CHECK_ARG(reinterpret_cast<uintptr_t>(state) % (2 * elem_size(dtype)) == 0,
               "state must be aligned to two elements");
  const int m = pass->tile_bits;
  PassArgs a{};
  a.n = n;
  a.m = m;
  a.log2_tiles_per_state = n - m;
  a.total_tiles = batch << (n - m);
  a.num_stages = pass->num_stages;
  for (int i = 0; i < m; ++i) a.local[i] = pass->local_bits[i];
  for (int st = 0; st < a.num_stages; ++st) {
    for (int k = 0; k < 4; ++k) a.stage_bits[st][k] = pass->stage_bits[st][k];
    a.stage_ops[st] = pass->stage_ops[st];
  }
  a.stage_ops[a.num_stages] = pass->stage_ops[a.num_stages];
  a.needs_index = info.needs_index ? 1 : 0;
  a.reg_head = info.reg_head ? 1 : 0;
  a.reg_tail = info.reg_tail ? 1 : 0;
  const bool heavy = info.heavy, only_layers = info.only_layers;
  const int64_t tiles_per_block = int64_t(kTile) >> m;
  const int64_t blocks = (a.total_tiles + tiles_per_block - 1) / tiles_per_block;
  UA_CHECK_ARG(blocks <= 0x7fffffffLL, "too many tiles");
  const bool multi = m < kTileBits;
  const int variant = only_layers ? kLayers : (heavy ? kFull : kLight);
#define UA_FUSED_LAUNCH(T, V, M)                                                     \
  fused_pass_kernel<T, V, M><<<unsigned(blocks), kThreads, 0, s>>>(static_cast<T*>(state), \
                                                                   ops, a)
#define UA_FUSED_BY_VARIANT(T, M)                         \
  do {                                                    \
    if (variant == kLayers) UA_FUSED_LAUNCH(T, kLayers, M); \
    else if (variant == kLight) UA_FUSED_LAUNCH(T, kLight, M); \
    else UA_FUSED_LAUNCH(T, kFull, M);                      \
  } while (0)
  if (dtype == UA_COMPLEX64) {
    if (multi) UA_FUSED_BY_VARIANT(float2, true);
    else UA_FUSED_BY_VARIANT(float2, false);
  } else {
    if (multi) UA_FUSED_BY_VARIANT(double2, true);
    else UA_FUSED_BY_VARIANT(double2, false);
  }
#undef UA_FUSED_BY_VARIANT
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
