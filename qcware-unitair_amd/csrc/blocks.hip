// Runs of dense 4- and 5-qubit gates, several gates per pass over the state,
// on the matrix cores: ua_plan_dense_blocks / ua_run_dense_blocks.
//
// unitair applies each gate as its own contraction over the whole state
// (unitair/simulation/operations.py:151-186, 322), and so does
// ua_apply_operator (mfma.hip for these sizes): every 4/5-qubit gate reads and
// writes the state once, which at 28 qubits is ~4.3 GB of HBM traffic per
// gate for ~0.07 TFLOP of matrix work.  Here consecutive gates whose targets
// fit together in one 12-bit tile are applied in one pass:
//   - tile = 2^12 amplitudes spanning the union of the pass's targets, state
//     bits 0 and 1 (runs of at least 4 amplitudes), and the lowest other state
//     bits up to 12 ("local bits");
//   - each workgroup is persistent: it stages every gate of the pass as MFMA
//     A-operands in LDS once, then walks its tiles, prefetching the next
//     tile's amplitudes into registers while the current one is multiplied;
//   - per gate, the tile's 2^(12-k) groups are the GEMM columns: each
//     wavefront owns whole column tiles of 16 groups, reads their B-operands
//     from LDS and writes the results back in place (a wavefront only ever
//     touches its own groups, so one barrier per gate suffices);
//   - tiles that share 128-byte lines are visited by one XCD at about the
//     same time (workgroups are dealt to the 8 XCDs round-robin).
// A pass holds up to 3 gates in complex64 (80 KB of LDS, two workgroups per
// CU) and 2 in complex128 (128 KB); the union of two 5-qubit gates always
// fits (10 bits + 2).  Gates of other kinds or sizes, and states below 14
// qubits, run one by one as ua_run_circuit runs them.  The same kernel, with
// a 64 KB A-operand block, takes single 6-qubit complex64 gates of
// apply_operator (in place or not).
#include "mfma_common.h"

namespace ua {
int run_gate(const ua_gate& g, int index, void* dst, void* cur, int dtype, int n,
             int64_t batch, hipStream_t s);
int record_event(void* ev, hipStream_t s);

namespace {

constexpr int kTB = 12;
constexpr int kTile = 1 << kTB;
constexpr int kMaxPassGates = 3;
constexpr int kMinBlockQubits = 14;   // whole tiles and a few of them per state
constexpr int kFragWords = 4096;      // A-operand reals of a 5-qubit gate (64 x 64)
constexpr int kFragWords6 = 16384;    // of a 6-qubit gate (128 x 128)

struct BlockGate {
  const void* op;   // D x D matrix, row-major, in the caller's qubit order
  int k;            // 4, 5 (several per pass) or 6 (alone, complex64)
  int tpos[6];      // tile positions of the targets, ascending (canonical order)
  int gpos[6];      // gate-index bit of canonical target m
  int qpos[8];      // tile positions of the group bits, ascending (12 - k of them)
};

struct BlockPass {
  int64_t tiles;        // batch * 2^(n - 12)
  int64_t in_bstride;   // batch stride of the input (0: one state for every entry)
  int64_t bstride;      // batch stride of the output: 2^n
  int log2_tiles;       // n - 12
  int xcd;              // visit tiles in XCD-contiguous ranges
  int local[kTB];       // ascending state bits of the tile
  int num_gates;
  BlockGate g[kMaxPassGates];
};

// LDS position of tile element e: the high six bits are folded into the bank
// bits, which spreads the strided B-operand reads over the banks.
__device__ __forceinline__ int swz(int e) { return e ^ ((e >> 6) & 63); }

__device__ __forceinline__ int deposit(int x, const int* pos, int cnt) {
  int r = 0;
  for (int i = 0; i < cnt; ++i) r |= ((x >> i) & 1) << pos[i];
  return r;
}

// Hides a value from loop-invariant code motion: without it the compiler
// hoists every LDS address of every gate out of the tile loop and runs out
// of registers.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename R> struct Vec16;  // 16 bytes of A-operands
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };

// Multiplies every group of the tile by gate gi (K targets).  B-operands come
// from xr / xi, A-operands from af; results overwrite the wavefront's own
// groups.
template <typename T, int THREADS, int K>
__device__ __forceinline__ void apply_gate(typename RealOf<T>::type* xr,
                                           typename RealOf<T>::type* xi,
                                           const typename RealOf<T>::type* af,
                                           const BlockGate& g, int lane, int wave, int bl,
                                           const int (&wl)[4]) {
  using R = typename RealOf<T>::type;
  using M = Mfma<R>;
  using V = typename Vec16<R>::type;
  constexpr int W = THREADS / 64;
  constexpr int D = 1 << K;
  constexpr int RT = 2 * D / 16;           // row tiles of the real block
  constexpr int KS = 2 * D / 4;            // k-steps
  constexpr int AV = 16 / sizeof(R);       // A values per 16-byte read
  constexpr int CT = (kTile >> K) / 16;    // column tiles (16 groups each)
  // With fewer column tiles than wavefronts, WPC wavefronts share a column
  // tile, each taking RTW of its row tiles (then the results are written
  // only after every wavefront has read its B-operands).
  constexpr int WPC = W > CT ? W / CT : 1;
  constexpr int CPW = W > CT ? 1 : CT / W;  // column tiles per wavefront
  constexpr int RTW = RT / WPC;             // row tiles per wavefront
  static_assert(CT % (W / WPC) == 0 && RT % WPC == 0, "wavefront split");
  static_assert(WPC == 1 || RTW <= RT / 2, "a wavefront's rows stay in one half");
  const int nq = kTB - K;
  const int cw = wave % (W / WPC);          // column-tile lane of this wavefront
  const int t0 = (wave / (W / WPC)) * RTW;  // first row tile
  bl = opaque(bl);
  int cb[CPW];  // swizzled offsets of this wavefront's column tiles
#pragma unroll
  for (int c = 0; c < CPW; ++c) cb[c] = swz(deposit(16 * (cw + (W / WPC) * c), g.qpos, nq));
  typename M::acc_t acc[CPW][RTW];
#pragma unroll
  for (int c = 0; c < CPW; ++c)
#pragma unroll
    for (int ti = 0; ti < RTW; ++ti) acc[c][ti] = typename M::acc_t{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < KS / AV; ++kk) {
    V a[RTW];
#pragma unroll
    for (int ti = 0; ti < RTW; ++ti)
      a[ti] = *reinterpret_cast<const V*>(af + (((t0 + ti) * (KS / AV) + kk) * 64 + lane) * AV);
#pragma unroll
    for (int q = 0; q < AV; ++q) {
      const int ks = kk * AV + q;
      const R* src = (ks >= KS / 2) ? xi : xr;
      const int sj = swz(deposit(4 * (ks % (KS / 2)), g.tpos, K));
      R b[CPW];
#pragma unroll
      for (int c = 0; c < CPW; ++c) b[c] = src[bl ^ sj ^ cb[c]];
#pragma unroll
      for (int ti = 0; ti < RTW; ++ti)
#pragma unroll
        for (int c = 0; c < CPW; ++c) acc[c][ti] = M::mma(a[ti][q], b[c], acc[c][ti]);
    }
  }
  if constexpr (WPC > 1) __syncthreads();  // every B-operand of the shared columns is read
  int w[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) w[r] = opaque(wl[r]);
#pragma unroll
  for (int tl = 0; tl < RTW; ++tl) {
    const int ti = t0 + tl;
    R* dst = (ti >= RT / 2) ? xi : xr;
    const int sj = swz(deposit(16 * (ti % (RT / 2)), g.tpos, K));
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[w[r] ^ sj ^ cb[c]] = acc[c][tl][r];
  }
}

// FRAG: A-operand reals per gate (kFragWords for 4/5-qubit gates, kFragWords6
// for one 6-qubit gate).  `in` may equal `out` (in place): every tile is read
// before it is written, and tiles are disjoint.
template <typename T, int THREADS, int MAXG, int FRAG>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(FRAG == kFragWords6 ? 1 : 2))) void dense_blocks_kernel(
    T* out, const T* in, BlockPass p) {
  using R = typename RealOf<T>::type;
  constexpr int UPT = kTile / 2 / THREADS;  // element pairs per thread
  constexpr int AV = 16 / sizeof(R);
  __shared__ R xr[kTile], xi[kTile];
  __shared__ __attribute__((aligned(16))) R af[MAXG][FRAG];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // ---- A-operands of every gate, layout [ti][ks / AV][lane][ks % AV]
  for (int gi = 0; gi < p.num_gates; ++gi) {
    const BlockGate& g = p.g[gi];
    const int D = 1 << g.k, KS = D / 2, RT = D / 8;
    const T* U = static_cast<const T*>(g.op);
    for (int idx = tid; idx < RT * KS * 64; idx += THREADS) {
      const int q = idx % AV;
      const int l = (idx / AV) % 64;
      const int kq = (idx / AV / 64) % (KS / AV);
      const int ti = idx / AV / 64 / (KS / AV);
      const int ks = kq * AV + q;
      const int row = 16 * ti + (l & 15), col = 4 * ks + (l >> 4);
      const int rr = row & (D - 1), cc = col & (D - 1);
      int ur = 0, uc = 0;
      for (int m = 0; m < g.k; ++m) {
        ur |= ((rr >> m) & 1) << g.gpos[m];
        uc |= ((cc >> m) & 1) << g.gpos[m];
      }
      af[gi][idx] = real_block(U[ur * D + uc], row >= D, col >= D);
    }
  }
  // ---- per-gate lane offsets: B reads (bl) and result writes (wl)
  int bl[MAXG], wl[MAXG][4];
#pragma unroll
  for (int gi = 0; gi < MAXG; ++gi) {
    const BlockGate& g = p.g[gi];
    const int dq = deposit(lane & 15, g.qpos, 4);
    bl[gi] = swz(deposit(lane >> 4, g.tpos, 2) | dq);
#pragma unroll
    for (int r = 0; r < 4; ++r) wl[gi][r] = swz(deposit(Mfma<R>::row(lane, r), g.tpos, 4) | dq);
  }

  // ---- tiles: element pair e = 2 * (tid + THREADS * u), tile bit i = state bit local[i]
  int64_t dep_t = 0;
  for (int i = 1; i < kTB; ++i) dep_t |= int64_t(((2 * tid) >> i) & 1) << p.local[i];
  auto dep_u = [&](int u) {  // uniform part of pair u's offset
    int64_t d = 0;
    const int e = 2 * THREADS * u;
    for (int i = 1; i < kTB; ++i) d |= int64_t((e >> i) & 1) << p.local[i];
    return d;
  };
  const int swz_t = swz(2 * tid);
  const int64_t per = p.tiles >> 3;
  const int64_t g8 = gridDim.x >> 3;
  auto tile_at = [&](int64_t it) -> int64_t {  // -1 past the end
    if (p.xcd) {
      const int64_t t8 = (blockIdx.x >> 3) + it * g8;
      return t8 < per ? (blockIdx.x & 7) * per + t8 : -1;
    }
    const int64_t t = blockIdx.x + it * int64_t(gridDim.x);
    return t < p.tiles ? t : -1;
  };
  auto origin_of = [&](int64_t t, int64_t bstride) -> int64_t {
    const int64_t b = t >> p.log2_tiles;
    const uint64_t tin = uint64_t(t) & ((uint64_t(1) << p.log2_tiles) - 1);
    return b * bstride + int64_t(insert_zero_bits(tin, p.local, kTB));
  };

  Pair<T> v[UPT];
  int64_t tile = tile_at(0);
  if (tile >= 0) {
    const int64_t o = origin_of(tile, p.in_bstride) + dep_t;
#pragma unroll
    for (int u = 0; u < UPT; ++u) v[u] = *reinterpret_cast<const Pair<T>*>(in + o + dep_u(u));
  }
  for (int64_t it = 0; tile >= 0; ++it) {
    const int st = opaque(swz_t);
#pragma unroll
    for (int u = 0; u < UPT; ++u) {
      const int e = st ^ swz(2 * THREADS * u);
      xr[e] = v[u].v[0].x;
      xi[e] = v[u].v[0].y;
      xr[e ^ 1] = v[u].v[1].x;
      xi[e ^ 1] = v[u].v[1].y;
    }
    __syncthreads();
    const int64_t next = tile_at(it + 1);
    if (next >= 0) {
      const int64_t o = origin_of(next, p.in_bstride) + dep_t;
#pragma unroll
      for (int u = 0; u < UPT; ++u)
        v[u] = *reinterpret_cast<const Pair<T>*>(in + o + dep_u(u));
    }
#pragma unroll
    for (int gi = 0; gi < MAXG; ++gi) {
      if (gi < p.num_gates) {
        if constexpr (FRAG == kFragWords6)
          apply_gate<T, THREADS, 6>(xr, xi, af[gi], p.g[gi], lane, wave, bl[gi], wl[gi]);
        else if (p.g[gi].k == 5)
          apply_gate<T, THREADS, 5>(xr, xi, af[gi], p.g[gi], lane, wave, bl[gi], wl[gi]);
        else
          apply_gate<T, THREADS, 4>(xr, xi, af[gi], p.g[gi], lane, wave, bl[gi], wl[gi]);
        __syncthreads();
      }
    }
    const int64_t o = origin_of(tile, p.bstride) + dep_t;
    const int st2 = opaque(swz_t);
#pragma unroll
    for (int u = 0; u < UPT; ++u) {
      const int e = st2 ^ swz(2 * THREADS * u);
      Pair<T> w;
      w.v[0].x = xr[e];
      w.v[0].y = xi[e];
      w.v[1].x = xr[e ^ 1];
      w.v[1].y = xi[e ^ 1];
      *reinterpret_cast<Pair<T>*>(out + o + dep_u(u)) = w;
    }
    __syncthreads();
    tile = next;
  }
}

// Launch geometry: complex64 256 threads and 3 gates (80 KB LDS, two
// workgroups per CU); complex128 512 threads and 2 gates (128 KB, one).
constexpr int kThreads64 = 256, kThreads128 = 512;
constexpr int kThreads6 = 512;  // 6-qubit gates: 4 column tiles x 2 row halves
constexpr int kMaxGates64 = 3, kMaxGates128 = 2;
constexpr int kBlocksPerCu64 = 2, kBlocksPerCu128 = 1;
constexpr int kCus = 256;

bool block_gate(const ua_gate& g, int dtype, int n) {
  return g.kind == UA_GATE_DENSE && (g.num_targets == 4 || g.num_targets == 5) &&
         (dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128) && n >= kMinBlockQubits;
}

// Passes load element pairs as one 2 * sizeof(T) access.
bool pair_aligned(const void* state, int dtype) {
  return reinterpret_cast<uintptr_t>(state) % (2 * elem_size(dtype)) == 0;
}

uint64_t target_mask(const ua_gate& g, int n) {
  uint64_t m = 0;
  for (int j = 0; j < g.num_targets; ++j) m |= uint64_t(1) << (n - 1 - g.qubits[j]);
  return m;
}

// Units of work in order: a pass of block gates [first, first + count) or a
// single other gate (count = 1, pass = false).
struct Unit {
  int first, count;
  bool pass;
};

int plan_units(int dtype, int n, const ua_gate* gates, int num_gates, Unit* units) {
  const int maxg = dtype == UA_COMPLEX64 ? kMaxGates64 : kMaxGates128;
  int nu = 0;
  uint64_t uni = 0;
  for (int i = 0; i < num_gates; ++i) {
    const ua_gate& g = gates[i];
    if (!block_gate(g, dtype, n)) {
      units[nu++] = {i, 1, false};
      uni = 0;
      continue;
    }
    const uint64_t m = target_mask(g, n);
    Unit* last = nu ? &units[nu - 1] : nullptr;
    if (last && last->pass && last->count < maxg &&
        __builtin_popcountll(uni | m | 3) <= kTB) {
      ++last->count;
      uni |= m;
    } else {
      units[nu++] = {i, 1, true};
      uni = m;
    }
  }
  return nu;
}

int check_gates(int n, const ua_gate* gates, int num_gates) {
  for (int i = 0; i < num_gates; ++i) {
    const ua_gate& g = gates[i];
    if (g.kind != UA_GATE_DENSE && g.kind != UA_GATE_CZ && g.kind != UA_GATE_CNOT &&
        g.kind != UA_GATE_PHASE) {
      set_error("gate %d: unknown kind %d", i, g.kind);
      return UA_ERR_INVALID_ARGUMENT;
    }
    if (g.kind == UA_GATE_PHASE) continue;
    const int k = g.kind == UA_GATE_DENSE ? g.num_targets : 2;
    if (k < 1 || k > UA_MAX_GATE_QUBITS) {
      set_error("gate %d: %d targets", i, k);
      return UA_ERR_INVALID_ARGUMENT;
    }
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
      const int q = g.qubits[j];
      if (q < 0 || q >= n || ((seen >> q) & 1)) {
        set_error("gate %d: qubit %d repeated or out of range for %d qubits", i, q, n);
        return UA_ERR_INVALID_ARGUMENT;
      }
      seen |= uint64_t(1) << q;
    }
    if (g.kind == UA_GATE_DENSE && !g.data) {
      set_error("gate %d: NULL matrix", i);
      return UA_ERR_INVALID_ARGUMENT;
    }
  }
  return UA_OK;
}

// One pass of gates[0..count) (4/5-qubit gates, or a single 6-qubit complex64
// gate) from `in` (batch stride in_bstride) to `out` (stride 2^n; may be `in`).
int launch_pass(void* out, const void* in, int64_t in_bstride, int dtype, int n, int64_t batch,
                const ua_gate* gates, int count, hipStream_t s) {
  BlockPass p{};
  p.log2_tiles = n - kTB;
  p.tiles = batch << p.log2_tiles;
  p.in_bstride = in_bstride;
  p.bstride = int64_t(1) << n;
  uint64_t uni = 3;
  for (int i = 0; i < count; ++i) uni |= target_mask(gates[i], n);
  UA_CHECK_ARG(__builtin_popcountll(uni) <= kTB, "internal: %d tile bits",
               __builtin_popcountll(uni));
  for (int bit = 0; __builtin_popcountll(uni) < kTB; ++bit) uni |= uint64_t(1) << bit;
  for (int i = 0; i < kTB; ++i) {
    p.local[i] = __builtin_ctzll(uni);
    uni &= uni - 1;
  }
  p.num_gates = count;
  bool six = false;
  for (int gi = 0; gi < count; ++gi) {
    const ua_gate& g = gates[gi];
    BlockGate& b = p.g[gi];
    b.op = g.data;
    b.k = g.num_targets;
    six |= b.k == 6;
    int tb[6], gp[6];
    for (int m = 0; m < b.k; ++m) {
      tb[m] = n - 1 - g.qubits[m];
      gp[m] = b.k - 1 - m;
    }
    for (int i = 1; i < b.k; ++i)  // canonical order: ascending state bit
      for (int j = i; j > 0 && tb[j - 1] > tb[j]; --j) {
        int t0 = tb[j]; tb[j] = tb[j - 1]; tb[j - 1] = t0;
        int g0 = gp[j]; gp[j] = gp[j - 1]; gp[j - 1] = g0;
      }
    uint64_t tmask = 0;
    for (int m = 0; m < b.k; ++m) {
      int pos = -1;
      for (int i = 0; i < kTB; ++i)
        if (p.local[i] == tb[m]) pos = i;
      UA_CHECK_ARG(pos >= 0, "internal: target bit %d not in the tile", tb[m]);
      b.tpos[m] = pos;
      b.gpos[m] = gp[m];
      tmask |= uint64_t(1) << pos;
    }
    int nq = 0;
    for (int i = 0; i < kTB; ++i)
      if (!((tmask >> i) & 1)) b.qpos[nq++] = i;
  }
  const bool c64 = dtype == UA_COMPLEX64;
  UA_CHECK_ARG(!six || (c64 && count == 1), "internal: 6-qubit passes hold one complex64 gate");
  // 6-qubit gates: 96 KB of LDS (tile + the 128 x 128 block), one workgroup per CU
  const int per_cu = six ? 1 : (c64 ? kBlocksPerCu64 : kBlocksPerCu128);
  const int64_t want = int64_t(kCus) * per_cu;
  p.xcd = p.tiles % 8 == 0 && p.tiles >= want;
  int64_t grid = p.tiles < want ? p.tiles : want;
  if (p.xcd) grid &= ~int64_t(7);
  UA_CHECK_ARG(grid >= 1 && grid <= 0x7fffffffLL, "internal: grid %lld", (long long)grid);
  if (six)
    dense_blocks_kernel<float2, kThreads6, 1, kFragWords6><<<unsigned(grid), kThreads6, 0, s>>>(
        static_cast<float2*>(out), static_cast<const float2*>(in), p);
  else if (c64)
    dense_blocks_kernel<float2, kThreads64, kMaxGates64, kFragWords>
        <<<unsigned(grid), kThreads64, 0, s>>>(static_cast<float2*>(out),
                                               static_cast<const float2*>(in), p);
  else
    dense_blocks_kernel<double2, kThreads128, kMaxGates128, kFragWords>
        <<<unsigned(grid), kThreads128, 0, s>>>(static_cast<double2*>(out),
                                                static_cast<const double2*>(in), p);
  return check_launch("dense_blocks_kernel");
}

}  // namespace

// A single 6-qubit complex64 gate of apply_operator on the matrix cores (one
// operator for every batch entry).  qubits: the caller's order (first = gate
// MSB).  True when it applies: complex64, at least kMinBlockQubits qubits,
// pair-aligned buffers and strides.
bool six_applies(int dtype, int n, int k, const void* out, const void* in, int64_t in_bstride) {
  return k == 6 && dtype == UA_COMPLEX64 && n >= kMinBlockQubits && pair_aligned(out, dtype) &&
         pair_aligned(in, dtype) && in_bstride % 2 == 0;
}
int apply_dense_six(void* out, const void* in, const void* op, int n, const int* qubits,
                    int64_t batch, int64_t in_bstride, hipStream_t s) {
  ua_gate g{};
  g.kind = UA_GATE_DENSE;
  g.num_targets = 6;
  for (int m = 0; m < 6; ++m) g.qubits[m] = qubits[m];
  g.data = op;
  return launch_pass(out, in, in_bstride, UA_COMPLEX64, n, batch, &g, 1, s);
}

}  // namespace ua

extern "C" int ua_plan_dense_blocks(int dtype, int num_qubits, const ua_gate* gates,
                                    int num_gates, int* unit_of_gate, int* num_units) {
  using namespace ua;
  UA_CHECK_ARG(elem_size(dtype) != 0, "unsupported dtype code %d", dtype);
  UA_CHECK_ARG(num_qubits >= 1 && num_qubits <= 62, "num_qubits=%d out of range", num_qubits);
  UA_CHECK_ARG(num_gates >= 0 && (num_gates == 0 || gates), "bad gate list");
  UA_CHECK_ARG(num_units, "NULL num_units");
  const int rc = check_gates(num_qubits, gates, num_gates);
  if (rc) return rc;
  Unit* units = new Unit[num_gates > 0 ? num_gates : 1];
  const int nu = plan_units(dtype, num_qubits, gates, num_gates, units);
  if (unit_of_gate)
    for (int i = 0; i < nu; ++i)
      for (int j = 0; j < units[i].count; ++j) unit_of_gate[units[i].first + j] = i;
  delete[] units;
  *num_units = nu;
  return UA_OK;
}

extern "C" int ua_run_dense_blocks(void* state, int dtype, int num_qubits, int64_t batch,
                                   const ua_gate* gates, int num_gates, void* const* events,
                                   ua_stream_t stream) {
  using namespace ua;
  hipStream_t s = static_cast<hipStream_t>(stream);
  UA_CHECK_ARG(elem_size(dtype) != 0, "unsupported dtype code %d", dtype);
  UA_CHECK_ARG(num_qubits >= 1 && num_qubits <= 62, "num_qubits=%d out of range", num_qubits);
  UA_CHECK_ARG(num_gates >= 0 && (num_gates == 0 || gates), "bad gate list");
  UA_CHECK_ARG(batch >= 0, "negative batch");
  int rc = check_gates(num_qubits, gates, num_gates);
  if (rc) return rc;
  if (batch == 0 || num_gates == 0) return UA_OK;
  UA_CHECK_ARG(state, "NULL state");
  Unit* units = new Unit[num_gates];
  const int nu = plan_units(dtype, num_qubits, gates, num_gates, units);
  for (int i = 0; i < nu && rc == UA_OK; ++i) {
    // a start event that is the previous unit's end event is already recorded
    if (events && !(i > 0 && events[2 * i] == events[2 * i - 1])) rc = record_event(events[2 * i], s);
    if (rc) break;
    const Unit& u = units[i];
    if (u.pass && pair_aligned(state, dtype)) {
      rc = launch_pass(state, state, int64_t(1) << num_qubits, dtype, num_qubits, batch,
                       gates + u.first, u.count, s);
    } else {
      for (int j = 0; j < u.count && rc == UA_OK; ++j)
        rc = run_gate(gates[u.first + j], u.first + j, state, state, dtype, num_qubits, batch,
                      s);
    }
    if (rc) break;
    if (events) rc = record_event(events[2 * i + 1], s);
  }
  delete[] units;
  return rc;
}
