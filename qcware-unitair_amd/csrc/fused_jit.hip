// Compiled fused passes: one kernel per pass structure.
//
// fused.hip runs any pass through one kernel that reads its op list at run
// time: a loop with a switch per op and nested switches over cube bits.  The
// 16-entry cube lives in 32 registers that every switch arm and the loop
// back edge must hand over in the same registers, which costs register moves
// and selects at each join (about as many VALU instructions as the gate
// arithmetic itself).  Here the pass's structure (tile bits, stages, op kinds,
// the bits each op acts on) is written out as straight-line HIP source, with
// every bit position a constant, and compiled with hiprtc for gfx950.  Gate
// data is not part of the source: the kernel reads matrices, diagonals,
// factors and angles through the op array at run time (scalar loads), so one
// compiled pass serves every op array of the same structure.
//
// Compiled code objects are cached for the life of the process, keyed by the
// generated source; a module is loaded on each device the first time the
// pass runs there.  The reference has no counterpart (every gate is its own
// pass, unitair/simulation/operations.py:151-186).
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "fused_common.h"
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kMinTileBits = 12;  // tiles below 12 bits stay with the interpreter
// largest tile: 2^14 complex64 amplitudes (128 KiB of LDS) or 2^13 complex128
int max_tile_bits(int dtype) { return dtype == UA_COMPLEX128 ? 13 : 14; }

// Waves per SIMD the compiled complex64 kernels ask the register allocator
// for (0: 4, the default; a complex64 tile takes 32 KiB of LDS, so at most
// five workgroups of four waves fit a CU).  complex128 kernels ask for 2.
std::atomic<int> g_compiled_waves{0};

// Device code shared by every compiled pass.  UA_DOUBLE selects complex128.
const char* kPrelude = R"UAJIT(
#if UA_DOUBLE
typedef double R;
typedef double2 T;
#define UA_SINCOS sincos
#else
typedef float R;
typedef float2 T;
#define UA_SINCOS sincosf
#endif
typedef unsigned long long u64;
#define SC __attribute__((address_space(4)))
#define GL __attribute__((address_space(1)))
struct OpRec { int kind, a, b, slot; const void* data; long long bstride; };
static_assert(sizeof(OpRec) == 32, "ua_fused_op layout");

__device__ __forceinline__ constexpr int lds_pos(int e) { return e ^ ((e >> 5) & 31); }
__device__ __forceinline__ T cmk(R x, R y) { T r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ T cmul(T a, T b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ T cmad(T acc, T a, T b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ T cmul2(T a, T b) {
  T r;
  r.x = a.x * b.x;
  r.x = fma(-a.y, b.y, r.x);
  r.y = a.x * b.y;
  r.y = fma(a.y, b.x, r.y);
  return r;
}
__device__ __forceinline__ T cneg(T a) { return cmk(-a.x, -a.y); }
__device__ __forceinline__ T csel(bool s, T a, T b) { return cmk(s ? a.x : b.x, s ? a.y : b.y); }
__device__ __forceinline__ T expmi(R t) { R s, c; UA_SINCOS(t, &s, &c); return cmk(c, -s); }
__device__ __forceinline__ T ldc(const SC R* p, int i) { return cmk(p[2 * i], p[2 * i + 1]); }

// exp(-i t) for phase angles.  complex128: sincos.  complex64: Cody-Waite
// reduction by pi/2 (pi/2 = P1 + P2 + P3; P1 has 8 significant bits, so k*P1
// is exact for the |k| <= 163 used here, and each FMA rounds once), then the
// Taylor polynomials of sin and cos on [-pi/4, pi/4] (truncation errors
// below 2e-9) in Horner form.  Angles beyond 256 in magnitude (or NaN) take
// sincosf.
#if UA_DOUBLE
__device__ __forceinline__ T expmi_phase(R t) { return expmi(t); }
#else
__device__ __forceinline__ T expmi_phase(float t) {
  if (!(fabsf(t) <= 256.0f)) return expmi(t);
  const float k = rintf(t * 0.636619772367581343f);
  float r = fmaf(-k, 1.5703125f, t);
  r = fmaf(-k, 4.83826792333275079e-4f, r);
  r = fmaf(-k, 2.56334406825708965e-12f, r);
  const float r2 = r * r;
  float s = fmaf(r2, 2.755731922e-6f, -1.984126984e-4f);   // 1/9!, -1/7!
  s = fmaf(r2, s, 8.333333333e-3f);                        // 1/5!
  s = fmaf(r2, s, -1.666666667e-1f);                       // -1/3!
  s = fmaf(r2 * r, s, r);
  float c = fmaf(r2, -2.755731922e-7f, 2.480158730e-5f);   // -1/10!, 1/8!
  c = fmaf(r2, c, -1.388888889e-3f);                       // -1/6!
  c = fmaf(r2, c, 4.166666667e-2f);                        // 1/4!
  c = fmaf(r2, c, -0.5f);
  c = fmaf(r2, c, 1.0f);
  // t = k pi/2 + r: sin t = (s, c, -s, -c)[k mod 4], cos t = (c, -s, -c, s)[k mod 4]
  const int q = int(k) & 3;
  float sn = (q & 1) ? c : s, cs = (q & 1) ? s : c;
  if (q & 2) sn = -sn;
  if ((q + 1) & 2) cs = -cs;
  return cmk(cs, -sn);
}
#endif

// Two consecutive elements (aligned to two elements), as native vectors.
#if UA_DOUBLE
typedef double v2r __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld2(const GL T* p, T& a, T& b) {
  const v2r x = *reinterpret_cast<const GL v2r*>(p), y = *reinterpret_cast<const GL v2r*>(p + 1);
  a = cmk(x.x, x.y); b = cmk(y.x, y.y);
}
__device__ __forceinline__ void st2(GL T* p, T a, T b) {
  v2r x, y; x.x = a.x; x.y = a.y; y.x = b.x; y.y = b.y;
  *reinterpret_cast<GL v2r*>(p) = x; *reinterpret_cast<GL v2r*>(p + 1) = y;
}
__device__ __forceinline__ void ld2r(const GL R* p, R& a, R& b) {
  const v2r x = *reinterpret_cast<const GL v2r*>(p); a = x.x; b = x.y;
}
#else
typedef float v4r __attribute__((ext_vector_type(4)));
typedef float v2r __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld2(const GL T* p, T& a, T& b) {
  const v4r x = *reinterpret_cast<const GL v4r*>(p);
  a = cmk(x.x, x.y); b = cmk(x.z, x.w);
}
__device__ __forceinline__ void st2(GL T* p, T a, T b) {
  v4r x; x.x = a.x; x.y = a.y; x.z = b.x; x.w = b.y;
  *reinterpret_cast<GL v4r*>(p) = x;
}
__device__ __forceinline__ void ld2r(const GL R* p, R& a, R& b) {
  const v2r x = *reinterpret_cast<const GL v2r*>(p); a = x.x; b = x.y;
}
#endif

// ---- gates on a thread's cube of 16 amplitudes (cube bits 0..3)
template <int B>
__device__ __forceinline__ void cube_1q(T (&c)[16], const SC R* u) {
  const T u00 = ldc(u, 0), u01 = ldc(u, 1), u10 = ldc(u, 2), u11 = ldc(u, 3);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j & (1 << B)) continue;
    const int k = j | (1 << B);
    const T x = c[j], y = c[k];
    c[j] = cmad(cmul2(u00, x), u01, y);
    c[k] = cmad(cmul2(u10, x), u11, y);
  }
}
// gate index bit 1 = cube bit B1, bit 0 = cube bit B0
template <int B1, int B0>
__device__ __forceinline__ void cube_2q(T (&c)[16], const SC R* u) {
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j & ((1 << B1) | (1 << B0))) continue;
    const int idx[4] = {j, j | (1 << B0), j | (1 << B1), j | (1 << B1) | (1 << B0)};
    T x[4], o[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = c[idx[s]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T acc = cmul2(ldc(u, r * 4), x[0]);
#pragma unroll
      for (int s = 1; s < 4; ++s) acc = cmad(acc, ldc(u, r * 4 + s), x[s]);
      o[r] = acc;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) c[idx[s]] = o[s];
  }
}
// CZ with both qubits on cube bits A, B
template <int A, int B>
__device__ __forceinline__ void negate_cc(T (&c)[16]) {
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (((j >> A) & 1) && ((j >> B) & 1)) c[j] = cneg(c[j]);
}
// CZ with one qubit on cube bit A and the other outside the cube (set: s)
template <int A>
__device__ __forceinline__ void negate_c(T (&c)[16], bool s) {
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if ((j >> A) & 1) c[j] = csel(s, cneg(c[j]), c[j]);
}
__device__ __forceinline__ void negate_all(T (&c)[16], bool s) {
#pragma unroll
  for (int j = 0; j < 16; ++j) c[j] = csel(s, cneg(c[j]), c[j]);
}
// CNOT, control on cube bit C, target on cube bit B
template <int C, int B>
__device__ __forceinline__ void cnot_cc(T (&c)[16]) {
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (((j >> C) & 1) && !((j >> B) & 1)) {
      const T x = c[j];
      c[j] = c[j | (1 << B)];
      c[j | (1 << B)] = x;
    }
}
// CNOT, control outside the cube (set: s), target on cube bit B
template <int B>
__device__ __forceinline__ void cnot_lane(T (&c)[16], bool s) {
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j & (1 << B)) continue;
    const T x = c[j], y = c[j | (1 << B)];
    c[j] = csel(s, y, x);
    c[j | (1 << B)] = csel(s, x, y);
  }
}
// diag(d0, d1) on cube bit A
template <int A>
__device__ __forceinline__ void diag_c(T (&c)[16], const SC R* d) {
  const T d0 = ldc(d, 0), d1 = ldc(d, 1);
#pragma unroll
  for (int j = 0; j < 16; ++j) c[j] = cmul(((j >> A) & 1) ? d1 : d0, c[j]);
}
// diag(d0, d1) on a qubit outside the cube (set: s)
__device__ __forceinline__ void diag_all(T (&c)[16], const SC R* d, bool s) {
  const T dd = csel(s, ldc(d, 1), ldc(d, 0));
#pragma unroll
  for (int j = 0; j < 16; ++j) c[j] = cmul(dd, c[j]);
}
// CZ layer: sign of entry j is the parity of popcount(g & (g >> 1) & mask),
// g = gb | (cube part of j) with state bits P0..P3 for cube bits 0..3.  Over
// GF(2) it is P(gb), a term linear in the cube bits (a cube bit at p pairs
// with gb's bits p-1 and p+1) and a term for cube bits that are neighbours.
template <int P0, int P1, int P2, int P3>
__device__ __forceinline__ void cz_layer(T (&c)[16], u64 gb, u64 mask) {
  constexpr int pk[4] = {P0, P1, P2, P3};
  constexpr unsigned selk[4] = {0xAAAAu, 0xCCCCu, 0xF0F0u, 0xFF00u};
  unsigned neg = (__popcll(gb & (gb >> 1) & mask) & 1) ? 0xFFFFu : 0u;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    u64 l = (gb >> (pk[k] + 1)) & (mask >> pk[k]);
    if (pk[k] > 0) l ^= (gb >> (pk[k] - 1)) & (mask >> (pk[k] - 1));
    if (l & 1) neg ^= selk[k];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
      if (pk[k2] == pk[k] + 1 && ((mask >> pk[k]) & 1)) neg ^= selk[k] & selk[k2];
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) c[j] = csel((neg >> j) & 1, cneg(c[j]), c[j]);
}

// Cube of tile bits S0 < S1 < S2 < S3: byte offset of entry j from the
// thread's base (XOR: the LDS swizzle is linear over GF(2)), and the tile
// element of entry 0 for thread t (t's bits spread around the cube bits).
template <int S0, int S1, int S2, int S3>
struct Cube {
  static constexpr int off(int j) {
    return lds_pos(((j & 1) ? (1 << S0) : 0) | ((j & 2) ? (1 << S1) : 0) |
                   ((j & 4) ? (1 << S2) : 0) | ((j & 8) ? (1 << S3) : 0)) * int(sizeof(T));
  }
  static __device__ __forceinline__ int e0(int t) {
    t = ((t >> S0) << (S0 + 1)) | (t & ((1 << S0) - 1));
    t = ((t >> S1) << (S1 + 1)) | (t & ((1 << S1) - 1));
    t = ((t >> S2) << (S2 + 1)) | (t & ((1 << S2) - 1));
    t = ((t >> S3) << (S3 + 1)) | (t & ((1 << S3) - 1));
    return t;
  }
  static __device__ __forceinline__ void load(T (&q)[16], const char* tb, int base) {
#pragma unroll
    for (int j = 0; j < 16; ++j) q[j] = *reinterpret_cast<const T*>(tb + (base ^ off(j)));
  }
  static __device__ __forceinline__ void store(const T (&q)[16], char* tb, int base) {
#pragma unroll
    for (int j = 0; j < 16; ++j) *reinterpret_cast<T*>(tb + (base ^ off(j))) = q[j];
  }
};
)UAJIT";

struct PassShape {
  int dtype, n;
  const ua_fused_pass* pass;
  const ua_fused_op* ops;
  FusedPassInfo info;
};

std::string hex64(uint64_t v) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "0x%llxull", static_cast<unsigned long long>(v));
  return buf;
}

// Source text of the kernel for one pass (validated by check_fused_pass,
// tile_bits == 12).
std::string generate(const PassShape& ps) {
  const ua_fused_pass& p = *ps.pass;
  const int n = ps.n;
  const int* local = p.local_bits;
  std::ostringstream s;
  const bool dbl = ps.dtype == UA_COMPLEX128;
  const int tbits = p.tile_bits;              // 12..14: 2^(tbits-4) threads of 16 amplitudes
  const int threads = 1 << (tbits - 4);
  s << "#define UA_DOUBLE " << (dbl ? 1 : 0) << "\n" << kPrelude;
  const int waves = dbl ? 2 : (g_compiled_waves > 0 ? int(g_compiled_waves) : 4);
  s << "extern \"C\" __global__ __launch_bounds__(" << threads
    << ") __attribute__((amdgpu_waves_per_eu("
    << waves << ")))\nvoid ua_compiled_pass(T* __restrict__ state, "
    << "const OpRec* __restrict__ ops_arg) {\n";
  s << "  const SC OpRec* ops = (const SC OpRec*)ops_arg;\n"
    << "  const int tid = threadIdx.x;\n"
    << "  const u64 tg = blockIdx.x;\n"
    << "  const long long b = (long long)(tg >> " << (n - tbits) << ");\n"
    << "  u64 origin = tg & " << hex64((uint64_t(1) << (n - tbits)) - 1) << ";\n";
  // tile origin: the tile index with zero bits inserted at the local bits
  for (int i = 0; i < tbits; ++i)
    s << "  origin = ((origin >> " << local[i] << ") << " << local[i] + 1 << ") | (origin & "
      << hex64((uint64_t(1) << local[i]) - 1) << ");\n";
  s << "  GL T* const st = (GL T*)state + b * " << hex64(uint64_t(1) << n) << ";\n"
    << "  const int t2 = 2 * tid;\n"
    << "  const u64 dep_t = u64(t2 & 15)";
  for (int i = 4; i < tbits - 3; ++i)
    s << " | (u64((t2 >> " << i << ") & 1) << " << local[i] << ")";
  s << ";\n  const int lds_t = lds_pos(t2);\n";
  // pair k of this thread (tile element 2t + 2 * threads * k: bit 0 the
  // pair, bits 1..tbits-4 the thread, the top 3 bits k): state index
  // origin | depk[k] | dep_t, LDS position lds_t ^ ldsk[k]
  uint64_t depk[8];
  int ldsk[8];
  for (int k = 0; k < 8; ++k) {
    depk[k] = 0;
    for (int i = tbits - 3; i < tbits; ++i)
      depk[k] |= uint64_t((k >> (i - (tbits - 3))) & 1) << local[i];
    const int e = 2 * threads * k;
    ldsk[k] = e ^ ((e >> 5) & 31);  // lds_pos, linear over GF(2)
  }
  auto gd = [&](int o) {
    std::ostringstream e;
    e << "((const SC R*)ops[" << o << "].data + 2 * b * ops[" << o << "].bstride)";
    return e.str();
  };
  // diagonal ops on the pair (x0, x1) at state indices g, g | 1
  // pk >= 0: the factors of PHASE_FACTORS ops were loaded ahead into
  // pf<o>[2 pk], pf<o>[2 pk + 1] (emit_prefetch)
  auto diag_ops = [&](int st, const char* x0, const char* x1, int pk = -1) {
    std::ostringstream e;
    for (int o = p.stage_ops[st]; o < p.stage_ops[st + 1]; ++o) {
      const ua_fused_op& op = ps.ops[o];
      switch (op.kind) {
        case UA_FOP_PHASE_FACTORS:
          if (pk >= 0)
            e << "      " << x0 << " = cmul(pf" << o << "[" << 2 * pk << "], " << x0 << "); "
              << x1 << " = cmul(pf" << o << "[" << 2 * pk + 1 << "], " << x1 << ");\n";
          else
            e << "      { T f0, f1; ld2((const GL T*)ops[" << o << "].data + b * ops[" << o
              << "].bstride + g, f0, f1); " << x0 << " = cmul(f0, " << x0 << "); " << x1
              << " = cmul(f1, " << x1 << "); }\n";
          break;
        case UA_FOP_PHASE:
          e << "      { R a0, a1; ld2r((const GL R*)ops[" << o << "].data + b * ops[" << o
            << "].bstride + g, a0, a1); " << x0 << " = cmul(expmi_phase(a0), " << x0 << "); "
            << x1 << " = cmul(expmi_phase(a1), " << x1 << "); }\n";
          break;
        case UA_FOP_CZ:
          e << "      if ((g >> " << op.a << ") & (g >> " << op.b << ") & 1) " << x0
            << " = cneg(" << x0 << ");\n"
            << "      if (((g | 1) >> " << op.a << ") & ((g | 1) >> " << op.b << ") & 1) " << x1
            << " = cneg(" << x1 << ");\n";
          break;
        case UA_FOP_DIAG1:
          e << "      { const SC R* d = " << gd(o) << "; const T d0 = ldc(d, 0), d1 = ldc(d, 1); "
            << x0 << " = cmul(((g >> " << op.a << ") & 1) ? d1 : d0, " << x0 << "); " << x1
            << " = cmul((((g | 1) >> " << op.a << ") & 1) ? d1 : d0, " << x1 << "); }\n";
          break;
        case UA_FOP_CZ_LAYER:
          e << "      { const u64 mask = ((const SC u64*)ops[" << o << "].data)[b * ops[" << o
            << "].bstride]; if (__popcll(g & (g >> 1) & mask) & 1) " << x0 << " = cneg(" << x0
            << "); if (__popcll((g | 1) & ((g | 1) >> 1) & mask) & 1) " << x1 << " = cneg("
            << x1 << "); }\n";
          break;
        default:
          break;
      }
    }
    return e.str();
  };

  // a diagonal stage on the registers of the global load / store
  auto diag_regs = [&](int st, bool prefetched = false) {
    for (int k = 0; k < 8; ++k) {
      s << "  {\n    const u64 g = origin | " << hex64(depk[k]) << " | dep_t;\n";
      std::string a = "c[" + std::to_string(2 * k) + "]", c1 = "c[" + std::to_string(2 * k + 1) + "]";
      s << diag_ops(st, a.c_str(), c1.c_str(), prefetched ? k : -1) << "  }\n";
    }
  };
  // Phase factors of a diagonal stage in tile (pair) order, loaded one stage
  // ahead: barriers wait for LDS only, so the loads stay in flight while the
  // stage before runs.
  auto has_factors = [&](int st) {
    for (int o = p.stage_ops[st]; o < p.stage_ops[st + 1]; ++o)
      if (ps.ops[o].kind == UA_FOP_PHASE_FACTORS) return true;
    return false;
  };
  auto emit_prefetch = [&](int st) {
    for (int o = p.stage_ops[st]; o < p.stage_ops[st + 1]; ++o) {
      if (ps.ops[o].kind != UA_FOP_PHASE_FACTORS) continue;
      s << "  T pf" << o << "[16];\n";
      for (int k = 0; k < 8; ++k)
        s << "  ld2((const GL T*)ops[" << o << "].data + b * ops[" << o << "].bstride + (origin | "
          << hex64(depk[k]) << " | dep_t), pf" << o << "[" << 2 * k << "], pf" << o << "["
          << 2 * k + 1 << "]);\n";
    }
  };

  const int ns = p.num_stages;
  // unlike the interpreter, any diagonal stage at either end runs on the
  // registers of the global load or store
  const bool head = ns > 0 && p.stage_bits[0][0] < 0;
  const bool tail = ns > (head ? 1 : 0) && p.stage_bits[ns - 1][0] < 0;
  const int st_begin = head ? 1 : 0, st_end = ns - (tail ? 1 : 0);
  // A cube stage that holds tile bit 0 (= state bit 0) takes its cube from
  // global memory, and writes it back, as 8 pairs of adjacent amplitudes, and
  // needs no LDS round trip on that side.  The lanes' index bits fill the
  // lowest tile bits outside the cube, so a 128 B line is read whole by one
  // load instruction of 8 lanes (tile bits 1..3 outside the cube) or by a
  // few consecutive loads of the same lanes.  Stages [lds_first, lds_end) go
  // through LDS.
  auto direct = [&](int st) { return p.stage_bits[st][0] == 0; };
  const bool first_direct = st_begin < st_end && direct(st_begin);
  const int lds_first = first_direct ? st_begin + 1 : st_begin;
  const bool last_direct = st_end - 1 >= lds_first && direct(st_end - 1);
  const int lds_end = last_direct ? st_end - 1 : st_end;
  const bool uses_lds = st_begin < st_end && !(first_direct && lds_first >= st_end);

  // one cube stage: its cube from LDS or global memory, its ops, and back to
  // LDS or global memory (diagonal head / tail stages on the registers)
  auto cube_stage = [&](int st, bool from_global, bool to_global) {
    const int* sb = p.stage_bits[st];
    int cb[4];  // state bits of the cube bits
    for (int k = 0; k < 4; ++k) cb[k] = local[sb[k]];
    auto slot_of = [&](int bit) {
      for (int k = 0; k < 4; ++k)
        if (cb[k] == bit) return k;
      return -1;
    };
    // this thread's value of a state bit outside the cube
    auto bit_of = [&](int bit) {
      std::ostringstream e;
      int tile_pos = -1;
      for (int i = 0; i < tbits; ++i)
        if (local[i] == bit) tile_pos = i;
      if (tile_pos >= 0)
        e << "(((e0 >> " << tile_pos << ") & 1) != 0)";
      else
        e << "(((origin >> " << bit << ") & 1) != 0)";
      return e.str();
    };
    // state-index offset of cube entry j
    auto dep_j = [&](int j) {
      uint64_t d = 0;
      for (int k = 0; k < 4; ++k)
        if ((j >> k) & 1) d |= uint64_t(1) << cb[k];
      return d;
    };
    bool need_gb = from_global || to_global;
    for (int o = p.stage_ops[st]; o < p.stage_ops[st + 1]; ++o)
      need_gb = need_gb || ps.ops[o].kind == UA_FOP_CZ_LAYER;
    s << "  {\n    using C = Cube<" << sb[0] << ", " << sb[1] << ", " << sb[2] << ", " << sb[3]
      << ">;\n    const int e0 = C::e0(tid);\n    T q[16];\n";
    if (need_gb) {  // state index of cube entry 0
      s << "    const u64 gb = origin";
      for (int i = 0; i < tbits; ++i)
        s << " | (u64((e0 >> " << i << ") & 1) << " << local[i] << ")";
      s << ";\n";
    }
    if (!from_global || !to_global)
      s << "    const int base = lds_pos(e0) * int(sizeof(T));\n";
    if (from_global) {
      for (int j = 0; j < 16; j += 2)
        s << "    ld2(st + (gb | " << hex64(dep_j(j)) << "), q[" << j << "], q[" << j + 1
          << "]);\n";
      if (head)
        for (int j = 0; j < 16; j += 2) {
          s << "    {\n      const u64 g = gb | " << hex64(dep_j(j)) << ";\n";
          std::string a = "q[" + std::to_string(j) + "]", c1 = "q[" + std::to_string(j + 1) + "]";
          s << diag_ops(0, a.c_str(), c1.c_str()) << "    }\n";
        }
    } else {
      s << "    C::load(q, tb, base);\n";
    }
    for (int o = p.stage_ops[st]; o < p.stage_ops[st + 1]; ++o) {
      const ua_fused_op& op = ps.ops[o];
      switch (op.kind) {
        case UA_FOP_DENSE1:
          s << "    cube_1q<" << op.a << ">(q, " << gd(o) << ");\n";
          break;
        case UA_FOP_DENSE2:
          s << "    cube_2q<" << op.a << ", " << op.b << ">(q, " << gd(o) << ");\n";
          break;
        case UA_FOP_CNOT: {
          const int kc = slot_of(op.a);
          if (kc >= 0)
            s << "    cnot_cc<" << kc << ", " << op.b << ">(q);\n";
          else
            s << "    cnot_lane<" << op.b << ">(q, " << bit_of(op.a) << ");\n";
          break;
        }
        case UA_FOP_CZ: {
          const int ka = slot_of(op.a), kb = slot_of(op.b);
          if (ka >= 0 && kb >= 0)
            s << "    negate_cc<" << std::min(ka, kb) << ", " << std::max(ka, kb) << ">(q);\n";
          else if (ka >= 0 || kb >= 0)
            s << "    negate_c<" << (ka >= 0 ? ka : kb) << ">(q, "
              << bit_of(ka >= 0 ? op.b : op.a) << ");\n";
          else
            s << "    negate_all(q, " << bit_of(op.a) << " && " << bit_of(op.b) << ");\n";
          break;
        }
        case UA_FOP_DIAG1: {
          const int k = slot_of(op.a);
          if (k >= 0)
            s << "    diag_c<" << k << ">(q, " << gd(o) << ");\n";
          else
            s << "    diag_all(q, " << gd(o) << ", " << bit_of(op.a) << ");\n";
          break;
        }
        case UA_FOP_CZ_LAYER:
          s << "    cz_layer<" << cb[0] << ", " << cb[1] << ", " << cb[2] << ", " << cb[3]
            << ">(q, gb, ((const SC u64*)ops[" << o << "].data)[b * ops[" << o
            << "].bstride]);\n";
          break;
        default:  // LAYER1Q heads: their DENSE1 sub-ops follow as ordinary ops
          break;
      }
    }
    if (to_global) {
      if (tail)
        for (int j = 0; j < 16; j += 2) {
          s << "    {\n      const u64 g = gb | " << hex64(dep_j(j)) << ";\n";
          std::string a = "q[" + std::to_string(j) + "]", c1 = "q[" + std::to_string(j + 1) + "]";
          s << diag_ops(ns - 1, a.c_str(), c1.c_str()) << "    }\n";
        }
      for (int j = 0; j < 16; j += 2)
        s << "    st2(st + (gb | " << hex64(dep_j(j)) << "), q[" << j << "], q[" << j + 1
          << "]);\n";
    } else {
      s << "    C::store(q, tb, base);\n";
    }
    s << "  }\n";
  };

  if (uses_lds)
    s << "  __shared__ T tile[" << (1 << tbits)
      << "];\n  char* const tb = reinterpret_cast<char*>(tile);\n";
  if (uses_lds && lds_first < lds_end && has_factors(lds_first)) emit_prefetch(lds_first);
  else if (uses_lds && lds_first >= lds_end && !last_direct && tail && has_factors(ns - 1))
    emit_prefetch(ns - 1);
  if (first_direct) {
    cube_stage(st_begin, true, !uses_lds);
  } else {
    s << "  T c[16];\n";
    for (int k = 0; k < 8; ++k)
      s << "  ld2(st + (origin | " << hex64(depk[k]) << " | dep_t), c[" << 2 * k << "], c["
        << 2 * k + 1 << "]);\n";
    if (head) diag_regs(0);
    if (uses_lds) {
      for (int k = 0; k < 8; ++k)
        s << "  tile[lds_t ^ " << ldsk[k] << "] = c[" << 2 * k << "]; tile[lds_t ^ " << ldsk[k]
          << " ^ 1] = c[" << 2 * k + 1 << "];\n";
    } else {  // diagonal stages only
      if (tail) diag_regs(ns - 1);
      for (int k = 0; k < 8; ++k)
        s << "  st2(st + (origin | " << hex64(depk[k]) << " | dep_t), c[" << 2 * k << "], c["
          << 2 * k + 1 << "]);\n";
    }
  }
  for (int st = lds_first; st < lds_end; ++st) {
    s << "  __syncthreads();\n";
    if (st + 1 < lds_end && has_factors(st + 1)) emit_prefetch(st + 1);
    else if (st + 1 == lds_end && !last_direct && tail && has_factors(ns - 1))
      emit_prefetch(ns - 1);
    if (p.stage_bits[st][0] < 0) {  // diagonal stage, LDS to LDS in tile order
      for (int k = 0; k < 8; ++k)
        s << "  {\n    const u64 g = origin | " << hex64(depk[k]) << " | dep_t;\n"
          << "    const int lp = lds_t ^ " << ldsk[k] << ";\n"
          << "    T x0 = tile[lp], x1 = tile[lp ^ 1];\n"
          << diag_ops(st, "x0", "x1", has_factors(st) ? k : -1)
          << "    tile[lp] = x0; tile[lp ^ 1] = x1;\n  }\n";
      continue;
    }
    cube_stage(st, false, false);
  }
  if (uses_lds) {
    s << "  __syncthreads();\n";
    if (last_direct) {
      cube_stage(st_end - 1, false, true);
    } else {
      s << "  {\n    T c[16];\n";
      for (int k = 0; k < 8; ++k)
        s << "    c[" << 2 * k << "] = tile[lds_t ^ " << ldsk[k] << "]; c[" << 2 * k + 1
          << "] = tile[lds_t ^ " << ldsk[k] << " ^ 1];\n";
      if (tail) diag_regs(ns - 1, has_factors(ns - 1));
      for (int k = 0; k < 8; ++k)
        s << "    st2(st + (origin | " << hex64(depk[k]) << " | dep_t), c[" << 2 * k
          << "], c[" << 2 * k + 1 << "]);\n";
      s << "  }\n";
    }
  }
  s << "}\n";
  return s.str();
}

// A compiled pass: its code object, and the function loaded on each device.
struct Compiled {
  int dtype = 0, n = 0, tile_bits = 0;
  std::vector<char> code;
  std::mutex mu;
  std::map<int, hipFunction_t> fn;
};

std::mutex g_cache_mu;
std::map<std::string, std::unique_ptr<Compiled>>& cache() {
  static auto* c = new std::map<std::string, std::unique_ptr<Compiled>>();  // never freed:
  return *c;  // modules stay loaded until the process ends
}

int check_shape(int dtype, int n, const ua_fused_pass* pass, const ua_fused_op* host_ops,
                PassShape* ps) {
  FusedPassInfo info{};
  const int rc = check_fused_pass(dtype, n, pass, host_ops, &info, max_tile_bits(dtype));
  if (rc != UA_OK) return rc;
  if (pass->tile_bits < kMinTileBits) {
    set_error("compiled passes need tile_bits of 12..%d (got %d)", max_tile_bits(dtype),
              pass->tile_bits);
    return UA_ERR_UNSUPPORTED;
  }
  *ps = PassShape{dtype, n, pass, host_ops, info};
  return UA_OK;
}

int compile(const std::string& src, std::vector<char>* code) {
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "ua_compiled_pass.hip", 0, nullptr,
                                       nullptr);
  if (r != HIPRTC_SUCCESS) {
    set_error("hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
    return UA_ERR_HIP;
  }
  // the library's own flags: gfx950, -O3, packed FP32 instructions disabled
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  r = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size + 1, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    set_error("hiprtc compile of a fused pass failed: %s: %.900s", hiprtcGetErrorString(r),
              log.c_str());
    return UA_ERR_HIP;
  }
  size_t size = 0;
  hiprtcGetCodeSize(prog, &size);
  code->resize(size);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return UA_OK;
}

}  // namespace

int set_compiled_waves(int waves) {
  UA_CHECK_ARG(waves == 0 || (waves >= 1 && waves <= 8),
               "compiled-pass waves must be 0 (default) or 1..8");
  g_compiled_waves = waves;
  return UA_OK;
}

}  // namespace ua

extern "C" int ua_fused_pass_source(int dtype, int num_qubits, const ua_fused_pass* pass,
                                    const ua_fused_op* host_ops, char* buf, int64_t capacity,
                                    int64_t* length) {
  ua::PassShape ps;
  const int rc = ua::check_shape(dtype, num_qubits, pass, host_ops, &ps);
  if (rc != UA_OK) return rc;
  UA_CHECK_ARG(length != nullptr, "NULL length");
  UA_CHECK_ARG(capacity >= 0 && (buf != nullptr || capacity == 0), "bad buffer");
  const std::string src = ua::generate(ps);
  *length = int64_t(src.size());
  if (capacity > 0) {
    const size_t k = std::min(src.size(), size_t(capacity - 1));
    std::memcpy(buf, src.data(), k);
    buf[k] = '\0';
  }
  return UA_OK;
}

extern "C" int ua_compile_fused_pass(int dtype, int num_qubits, const ua_fused_pass* pass,
                                     const ua_fused_op* host_ops, void** handle) {
  UA_CHECK_ARG(handle != nullptr, "NULL handle");
  *handle = nullptr;
  ua::PassShape ps;
  int rc = ua::check_shape(dtype, num_qubits, pass, host_ops, &ps);
  if (rc != UA_OK) return rc;
  const std::string src = ua::generate(ps);
  {
    std::lock_guard<std::mutex> lock(ua::g_cache_mu);
    auto it = ua::cache().find(src);
    if (it != ua::cache().end()) {
      *handle = it->second.get();
      return UA_OK;
    }
  }
  auto c = std::make_unique<ua::Compiled>();
  c->dtype = dtype;
  c->n = num_qubits;
  c->tile_bits = pass->tile_bits;
  rc = ua::compile(src, &c->code);  // outside the lock: compiles run in parallel
  if (rc != UA_OK) return rc;
  std::lock_guard<std::mutex> lock(ua::g_cache_mu);
  auto& slot = ua::cache()[src];
  if (!slot) slot = std::move(c);
  *handle = slot.get();
  return UA_OK;
}

namespace ua {
namespace {
// The pass's function on the current device, loading its code there on first
// use.
int function_of(Compiled* c, hipFunction_t* f) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("hipGetDevice: %s", hipGetErrorString(e));
    return UA_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(c->mu);
  auto it = c->fn.find(dev);
  if (it != c->fn.end()) {
    *f = it->second;
    return UA_OK;
  }
  hipModule_t mod;
  e = hipModuleLoadData(&mod, c->code.data());
  if (e == hipSuccess) e = hipModuleGetFunction(f, mod, "ua_compiled_pass");
  if (e != hipSuccess) {
    set_error("loading a compiled fused pass: %s", hipGetErrorString(e));
    return UA_ERR_HIP;
  }
  c->fn[dev] = *f;
  return UA_OK;
}
}  // namespace
}  // namespace ua

extern "C" int ua_load_compiled_pass(void* handle) {
  UA_CHECK_ARG(handle != nullptr, "NULL handle");
  hipFunction_t f = nullptr;
  return ua::function_of(static_cast<ua::Compiled*>(handle), &f);
}

extern "C" int ua_run_compiled_pass(void* handle, void* state, int64_t batch,
                                    const ua_fused_op* ops, ua_stream_t stream) {
  UA_CHECK_ARG(handle != nullptr, "NULL handle");
  auto* c = static_cast<ua::Compiled*>(handle);
  UA_CHECK_ARG(batch >= 0, "negative batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(state != nullptr && ops != nullptr, "NULL pointer");
  UA_CHECK_ARG(reinterpret_cast<uintptr_t>(state) % (2 * ua::elem_size(c->dtype)) == 0,
               "state must be aligned to two elements");
  const int64_t blocks = batch << (c->n - c->tile_bits);
  UA_CHECK_ARG(batch <= (int64_t(0x7fffffff) >> (c->n - c->tile_bits)), "too many tiles");
  hipFunction_t f = nullptr;
  const int rc = ua::function_of(c, &f);
  if (rc != UA_OK) return rc;
  void* args[] = {&state, &ops};
  const unsigned threads = 1u << (c->tile_bits - 4);
  const hipError_t e = hipModuleLaunchKernel(f, unsigned(blocks), 1, 1, threads, 1, 1, 0,
                                             static_cast<hipStream_t>(stream), args, nullptr);
  if (e != hipSuccess) {
    ua::set_error("launching a compiled fused pass: %s", hipGetErrorString(e));
    return UA_ERR_HIP;
  }
  return UA_OK;
}
