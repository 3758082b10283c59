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
// apply_operator (in place or not); single 7- and 8-qubit complex64 gates
// (out of place) run on dense_slice_kernel, which keeps one row slice of the
// gate's real block in LDS.
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

// How apply_gate splits a K-target gate over the workgroup's wavefronts.
template <typename T, int THREADS, int K> struct GateShape {
  using R = typename RealOf<T>::type;
  static constexpr int W = THREADS / 64;
  static constexpr int D = 1 << K;
  static constexpr int RT = 2 * D / 16;           // row tiles of the real block
  static constexpr int KS = 2 * D / 4;            // k-steps
  static constexpr int AV = 16 / sizeof(R);       // A values per 16-byte read
  static constexpr int CT = (kTile >> K) / 16;    // column tiles (16 groups each)
  // With fewer column tiles than wavefronts, WPC wavefronts share a column
  // tile, each taking RTW of its row tiles (then the results are written
  // only after every wavefront has read its B-operands).
  static constexpr int WPC = W > CT ? W / CT : 1;
  static constexpr int CPW = W > CT ? 1 : CT / W;  // column tiles per wavefront
  static constexpr int RTW = RT / WPC;             // row tiles per wavefront
  static constexpr int NPOS = KS * CPW;            // B-operand reads per wavefront
  static_assert(CT % (W / WPC) == 0 && RT % WPC == 0, "wavefront split");
  static_assert(WPC == 1 || RTW <= RT / 2, "a wavefront's rows stay in one half");
};

// Swizzled offset of column tile c of this wavefront for gate g.
template <typename T, int THREADS, int K>
__device__ __forceinline__ int column_offset(const BlockGate& g, int wave, int c) {
  using Sh = GateShape<T, THREADS, K>;
  const int cw = wave % (Sh::W / Sh::WPC);
  return swz(deposit(16 * (cw + (Sh::W / Sh::WPC) * c), g.qpos, kTB - K));
}

// B-operand positions of gate g read by this wavefront, in apply_gate's
// order (k-step ks, column tile c at ks * CPW + c), as offsets from the Re
// half of the tile (the Im half starts at kTile).  They are the same for
// every tile, so the kernel forms them once and keeps them in registers:
// recomputed per tile, each read waited on its own address chain.
template <typename T, int THREADS, int K>
__device__ __forceinline__ void gate_positions(const BlockGate& g, int lane, int wave, int* pos) {
  using Sh = GateShape<T, THREADS, K>;
  const int bl = swz(deposit(lane >> 4, g.tpos, 2) | deposit(lane & 15, g.qpos, 4));
#pragma unroll
  for (int ks = 0; ks < Sh::KS; ++ks) {
    const int sj = swz(deposit(4 * (ks % (Sh::KS / 2)), g.tpos, K));
#pragma unroll
    for (int c = 0; c < Sh::CPW; ++c)
      pos[ks * Sh::CPW + c] = opaque((ks >= Sh::KS / 2 ? kTile : 0) +
                                     (bl ^ sj ^ column_offset<T, THREADS, K>(g, wave, c)));
  }
}

// Multiplies every group of the tile by gate gi (K targets).  B-operands come
// from the tile x (Re half, then Im half) at the positions pos, A-operands
// from af; results overwrite the wavefront's own groups.
template <typename T, int THREADS, int K>
__device__ __forceinline__ void apply_gate(typename RealOf<T>::type* x,
                                           const typename RealOf<T>::type* af,
                                           const BlockGate& g, int lane, int wave,
                                           const int* pos, const int (&wl)[4]) {
  using Sh = GateShape<T, THREADS, K>;
  using R = typename RealOf<T>::type;
  using M = Mfma<R>;
  using V = typename Vec16<R>::type;
  constexpr int RT = Sh::RT, KS = Sh::KS, AV = Sh::AV, W = Sh::W, WPC = Sh::WPC;
  constexpr int CPW = Sh::CPW, RTW = Sh::RTW;
  const int t0 = (wave / (W / WPC)) * RTW;  // first row tile
  int cb[CPW];  // swizzled offsets of this wavefront's column tiles
#pragma unroll
  for (int c = 0; c < CPW; ++c) cb[c] = column_offset<T, THREADS, K>(g, wave, c);
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
      R b[CPW];
#pragma unroll
      for (int c = 0; c < CPW; ++c) b[c] = x[pos[ks * CPW + c]];
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
    R* dst = (ti >= RT / 2) ? x + kTile : x;
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
  __shared__ R xs[2 * kTile];  // the tile: Re parts, then Im parts
  __shared__ __attribute__((aligned(16))) R af[MAXG][FRAG];
  R* const xr = xs;
  R* const xi = xs + kTile;
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
  // ---- per-gate B-operand positions and result-write lane offsets (wl)
  constexpr int NPOS = FRAG == kFragWords6 ? GateShape<T, THREADS, 6>::NPOS
                                           : GateShape<T, THREADS, 5>::NPOS;
  static_assert(FRAG == kFragWords6 || GateShape<T, THREADS, 4>::NPOS == NPOS, "positions");
  int pos[MAXG][NPOS], wl[MAXG][4];
#pragma unroll
  for (int gi = 0; gi < MAXG; ++gi) {
    const BlockGate& g = p.g[gi];
    if constexpr (FRAG == kFragWords6)
      gate_positions<T, THREADS, 6>(g, lane, wave, pos[gi]);
    else if (g.k == 5)
      gate_positions<T, THREADS, 5>(g, lane, wave, pos[gi]);
    else
      gate_positions<T, THREADS, 4>(g, lane, wave, pos[gi]);
    const int dq = deposit(lane & 15, g.qpos, 4);
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
          apply_gate<T, THREADS, 6>(xs, af[gi], p.g[gi], lane, wave, pos[gi], wl[gi]);
        else if (p.g[gi].k == 5)
          apply_gate<T, THREADS, 5>(xs, af[gi], p.g[gi], lane, wave, pos[gi], wl[gi]);
        else
          apply_gate<T, THREADS, 4>(xs, af[gi], p.g[gi], lane, wave, pos[gi], wl[gi]);
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

// ---- single 7- and 8-qubit complex64 gates on the matrix cores (out of place)
//
// The real block of a 7-qubit gate (256 x 256 floats, 256 KB) or an 8-qubit
// one (512 x 512, 1 MB) does not fit in LDS.  It is cut by output rows into S
// slices of D / S complex rows (the Re and the Im real rows of each), 128 KB
// of A-operands per slice: S = 2 for 7 targets, 8 for 8.  A persistent
// workgroup keeps one slice in LDS and walks tiles: it stages the whole
// 4096-amplitude tile (every column's inputs), multiplies its rows, and
// writes back only the tile's amplitudes whose top log2(S) canonical target
// bits equal its slice.  The S workgroups of a tile sit on one XCD and walk
// in step, so the tile comes from HBM once and from that XCD's L2 S - 1
// times.  Slices read all of a tile but write disjoint parts of it, so the
// input must not overlap the output.
constexpr int kSliceThreads = 512;
constexpr int kSliceWords = 32768;  // A-operand reals per slice (128 KB)
constexpr int kSliceGrid = 256;     // one workgroup per CU (160 KB of LDS each)

template <int K, int THREADS> struct SliceShape {
  static constexpr int W = THREADS / 64;              // wavefronts
  static constexpr int D = 1 << K;
  static constexpr int LS = K == 7 ? 1 : 3;          // log2 of the slice count
  static constexpr int S = 1 << LS;
  static constexpr int CR = D / S;                    // complex rows per slice
  static constexpr int RTS = CR / 8;                  // real row tiles per slice (Re, then Im)
  static constexpr int CT = (kTile >> K) / 16;        // column tiles of 16 groups
  static constexpr int KS = 2 * D / 4;                // k-steps
  static constexpr int KSPLIT = (K == 7 ? 1 : 2) * W / 8;  // k-ranges of a row tile
  static constexpr int WG = W / KSPLIT;               // wavefronts per k-range
  static constexpr int RPW = RTS * CT / WG;           // row tiles per wavefront
  static constexpr int NACC = RPW == 1 ? 2 : 1;       // accumulators per row tile
  static_assert(RTS * KS * 64 == kSliceWords, "slice size");
  static_assert(RPW >= 1 && (RTS * CT) % WG == 0 && KS % (4 * KSPLIT * NACC) == 0, "split");
  static_assert(KSPLIT == 1 || (KS / 2) % (KS / KSPLIT) == 0, "a k-range reads xr or xi only");
};

struct SlicePass {
  const void* op;       // D x D, row-major, in the caller's qubit order
  int64_t tiles;        // batch * 2^(n - 12)
  int64_t in_bstride;   // batch stride of the input (0: one state for every entry)
  int64_t bstride;      // batch stride of the output: 2^n
  int log2_tiles;       // n - 12
  int xcd;              // 256 workgroups, a tile's slices on one XCD
  int local[kTB];       // ascending state bits of the tile
  int tpos[8];          // tile positions of the targets, ascending (canonical order)
  int gpos[8];          // gate-index bit of canonical target m
  int qpos[8];          // tile positions of the group bits, ascending (12 - K of them)
};

template <int K, int THREADS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(1))) void
dense_slice_kernel(float2* out, const float2* in, SlicePass p) {
  using Sh = SliceShape<K, THREADS>;
  using M = Mfma<float>;
  using V = typename Vec16<float>::type;
  constexpr int D = Sh::D, RTS = Sh::RTS, CT = Sh::CT, KS = Sh::KS, RPW = Sh::RPW;
  constexpr int NACC = Sh::NACC, KW = KS / Sh::KSPLIT;  // k-steps per wavefront
  constexpr int UPT = kTile / 2 / THREADS;          // tile element pairs per thread
  constexpr int SP = kTile / 2 / Sh::S;                   // element pairs of a slice
  constexpr int UPS = (SP + THREADS - 1) / THREADS;
  constexpr int NQ = kTB - K;
  __shared__ float xs[2 * kTile];  // the tile: Re parts, then Im parts
  __shared__ __attribute__((aligned(16))) float af[kSliceWords];
  float* const xr = xs;
  float* const xi = xs + kTile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // ---- this workgroup's slice and tile walk
  int s, lg, lanes;
  int64_t base, span;
  if (p.xcd) {  // workgroup b runs on XCD b % 8
    const int j = blockIdx.x >> 3;
    s = j % Sh::S;
    lg = j / Sh::S;
    lanes = (kSliceGrid / 8) / Sh::S;
    span = p.tiles >> 3;
    base = (blockIdx.x & 7) * span;
  } else {
    s = blockIdx.x % Sh::S;
    lg = blockIdx.x / Sh::S;
    lanes = gridDim.x / Sh::S;
    span = p.tiles;
    base = 0;
  }

  // ---- A-operands of the slice, layout [lti][ks / 4][lane][ks % 4] (lane l
  // of k-step ks holds real row 16 * lti + l % 16, real column 4 * ks + l / 16).
  // Local row tile lti < RTS / 2 holds the Re rows of complex rows CR * s + 16 *
  // lti + (0..15), the others the Im rows of CR * s + 16 * (lti - RTS / 2) + ....
  // Each complex entry u of the slice's rows is read once and gives four reals:
  // Re rows [Re u, -Im u], Im rows [Im u, Re u] (real columns c and D + c).
  {
    const float2* U = static_cast<const float2*>(p.op);
    constexpr int NE = Sh::CR * D;
    static_assert(NE % (4 * THREADS) == 0, "unrolled staging");
    auto put = [&](int lti, int r, int col, float val) {
      const int ks = col >> 2;
      af[((lti * (KS / 4) + (ks >> 2)) * 64 + r + 16 * (col & 3)) * 4 + (ks & 3)] = val;
    };
#pragma unroll 4
    for (int idx = tid; idx < NE; idx += THREADS) {
      const int cc = idx & (D - 1), crl = idx >> K;
      const int cr = Sh::CR * s + crl;
      int ur = 0, uc = 0;
      for (int m = 0; m < K; ++m) {
        ur |= ((cr >> m) & 1) << p.gpos[m];
        uc |= ((cc >> m) & 1) << p.gpos[m];
      }
      const float2 u = U[ur * D + uc];
      const int lt = crl >> 4, r = crl & 15;
      put(lt, r, cc, u.x);
      put(lt, r, D + cc, -u.y);
      put(lt + RTS / 2, r, cc, u.y);
      put(lt + RTS / 2, r, D + cc, u.x);
    }
  }

  // ---- this wavefront's work: k-range kh, column tile ct, row tiles rg + (RTS / RPW) * j
  const int kh = wave / Sh::WG;
  const int ct = (wave % Sh::WG) % CT;
  const int rg = (wave % Sh::WG) / CT;
  const int dq = deposit(lane & 15, p.qpos, 4);
  const int bl = swz(deposit(lane >> 4, p.tpos, 2) | dq);
  int wl[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) wl[r] = swz(deposit(M::row(lane, r), p.tpos, 4) | dq);
  const int cb = swz(deposit(16 * ct, p.qpos, NQ));
  // B-operand positions in xs of this wavefront's k-steps, the same for every
  // tile (kept in registers): complex column 4 * ks walks the target bits
  // 2..K-1 of the tile as a subset of tmask (d <- ((d | ~tmask) + 1) & tmask,
  // wrapping to 0 after the last column); real columns D.. read the Im half
  int bpos[KW];
  {
    int tmask = 0;
    for (int m = 2; m < K; ++m) tmask |= 1 << p.tpos[m];
    int d = deposit(4 * ((kh * KW) % (KS / 2)), p.tpos, K);
    const int b0 = bl ^ cb;
#pragma unroll
    for (int t = 0; t < KW; ++t) {
      const bool im = Sh::KSPLIT == 1 ? t >= KS / 2 : kh * KW >= KS / 2;
      bpos[t] = opaque((im ? kTile : 0) + (b0 ^ swz(d)));
      d = ((d | ~tmask) + 1) & tmask;
    }
  }

  // ---- tile element pairs: loads cover the tile, stores the slice's part
  int64_t dep_t = 0;
  for (int i = 1; i < kTB; ++i) dep_t |= int64_t(((2 * tid) >> i) & 1) << p.local[i];
  auto dep_e = [&](int e) {
    int64_t d = 0;
    for (int i = 1; i < kTB; ++i) d |= int64_t((e >> i) & 1) << p.local[i];
    return d;
  };
  // slice bits: the top LS canonical targets (tile positions >= 5, so never
  // bit 0: an element pair lies in one slice)
  const int* spos = p.tpos + (K - Sh::LS);
  int sbits = 0;
  for (int j = 0; j < Sh::LS; ++j) sbits |= ((s >> j) & 1) << spos[j];
  int64_t st_off[UPS];
  int st_lds[UPS];
#pragma unroll
  for (int u = 0; u < UPS; ++u) {
    const int pi = tid + THREADS * u;
    const int e = (pi < SP ? int(insert_zero_bits(uint64_t(2 * pi), spos, Sh::LS)) : 0) | sbits;
    st_off[u] = dep_e(e);
    st_lds[u] = swz(e);
  }
  const int swz_t = swz(2 * tid);
  auto tile_at = [&](int64_t it) -> int64_t {  // -1 past the end
    const int64_t t = lg + it * lanes;
    return t < span ? base + t : -1;
  };
  auto origin_of = [&](int64_t t, int64_t bstride) -> int64_t {
    const int64_t b = t >> p.log2_tiles;
    const uint64_t tin = uint64_t(t) & ((uint64_t(1) << p.log2_tiles) - 1);
    return b * bstride + int64_t(insert_zero_bits(tin, p.local, kTB));
  };

  Pair<float2> v[UPT];
  int64_t tile = tile_at(0);
  if (tile >= 0) {
    const int64_t o = origin_of(tile, p.in_bstride) + dep_t;
#pragma unroll
    for (int u = 0; u < UPT; ++u)
      v[u] = *reinterpret_cast<const Pair<float2>*>(in + o + dep_e(2 * THREADS * u));
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
    __syncthreads();  // also: the A-operands are staged (first tile)
    const int64_t next = tile_at(it + 1);
    if (next >= 0) {
      const int64_t o = origin_of(next, p.in_bstride) + dep_t;
#pragma unroll
      for (int u = 0; u < UPT; ++u)
        v[u] = *reinterpret_cast<const Pair<float2>*>(in + o + dep_e(2 * THREADS * u));
    }
    // ---- products: acc[j][a] over this wavefront's k-steps (a: even / odd k-step chains)
    typename M::acc_t acc[RPW][NACC];
#pragma unroll
    for (int j = 0; j < RPW; ++j)
#pragma unroll
      for (int a = 0; a < NACC; ++a) acc[j][a] = typename M::acc_t{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KW / 4; ++kk) {
      V a[RPW];
#pragma unroll
      for (int j = 0; j < RPW; ++j) {
        const int lti = rg + (RTS / RPW) * j;
        a[j] = *reinterpret_cast<const V*>(
            af + ((lti * (KS / 4) + kh * (KW / 4) + kk) * 64 + lane) * 4);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float b = xs[bpos[4 * kk + q]];
#pragma unroll
        for (int j = 0; j < RPW; ++j)
          acc[j][q % NACC] = M::mma(a[j][q], b, acc[j][q % NACC]);
      }
    }
    __syncthreads();  // every B-operand is read: results may replace the inputs
    // ---- results into the tile: the last k-range stores, the others add in turn
#pragma unroll
    for (int round = Sh::KSPLIT - 1; round >= 0; --round) {
      if (kh == round) {
#pragma unroll
        for (int j = 0; j < RPW; ++j) {
          const int lti = rg + (RTS / RPW) * j;
          float* dst = lti < RTS / 2 ? xr : xi;
          const int sj = swz(deposit(Sh::CR * s + 16 * (lti % (RTS / 2)), p.tpos, K)) ^ cb;
          typename M::acc_t r4 = acc[j][0];
          if constexpr (NACC == 2) r4 += acc[j][1];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int e = opaque(wl[r]) ^ sj;
            dst[e] = round == Sh::KSPLIT - 1 ? r4[r] : dst[e] + r4[r];
          }
        }
      }
      __syncthreads();
    }
    // ---- the slice's amplitudes back to the state
    const int64_t o = origin_of(tile, p.bstride);
#pragma unroll
    for (int u = 0; u < UPS; ++u) {
      if (tid + THREADS * u < SP) {
        const int e = opaque(st_lds[u]);
        Pair<float2> w;
        w.v[0].x = xr[e];
        w.v[0].y = xi[e];
        w.v[1].x = xr[e ^ 1];
        w.v[1].y = xi[e ^ 1];
        *reinterpret_cast<Pair<float2>*>(out + o + st_off[u]) = w;
      }
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

// A single 7- or 8-qubit complex64 gate of apply_operator on the matrix cores
// (one operator for every batch entry), out of place.  True when it applies:
// complex64, at least kMinBlockQubits qubits, pair-aligned buffers and
// strides, and an input that does not overlap the output (batch >= 1).
bool slice_applies(int dtype, int n, int k, const void* out, const void* in, int64_t in_bstride,
                   int64_t batch) {
  if (!((k == 7 || k == 8) && dtype == UA_COMPLEX64 && n >= kMinBlockQubits && batch >= 1 &&
        pair_aligned(out, dtype) && pair_aligned(in, dtype) && in_bstride % 2 == 0))
    return false;
  const uintptr_t o0 = reinterpret_cast<uintptr_t>(out);
  const uintptr_t o1 = o0 + uintptr_t(batch << n) * sizeof(float2);
  const uintptr_t i0 = reinterpret_cast<uintptr_t>(in);
  const uintptr_t i1 = i0 + uintptr_t((batch - 1) * in_bstride + (int64_t(1) << n)) * sizeof(float2);
  return o1 <= i0 || i1 <= o0;
}

int apply_dense_slices(void* out, const void* in, const void* op, int n, int k,
                       const int* qubits, int64_t batch, int64_t in_bstride, hipStream_t s) {
  UA_CHECK_ARG(slice_applies(UA_COMPLEX64, n, k, out, in, in_bstride, batch),
               "internal: sliced matrix-core path for %d targets on %d qubits", k, n);
  SlicePass p{};
  p.op = op;
  p.log2_tiles = n - kTB;
  p.tiles = batch << p.log2_tiles;
  p.in_bstride = in_bstride;
  p.bstride = int64_t(1) << n;
  int tb[8], gp[8];
  for (int m = 0; m < k; ++m) {
    tb[m] = n - 1 - qubits[m];
    gp[m] = k - 1 - m;
  }
  for (int i = 1; i < k; ++i)  // canonical order: ascending state bit
    for (int j = i; j > 0 && tb[j - 1] > tb[j]; --j) {
      int t0 = tb[j]; tb[j] = tb[j - 1]; tb[j - 1] = t0;
      int g0 = gp[j]; gp[j] = gp[j - 1]; gp[j - 1] = g0;
    }
  // tile: the targets, state bits 0 and 1, then the lowest other bits
  uint64_t uni = 3, tmask_state = 0;
  for (int m = 0; m < k; ++m) tmask_state |= uint64_t(1) << tb[m];
  uni |= tmask_state;
  for (int bit = 0; __builtin_popcountll(uni) < kTB; ++bit) uni |= uint64_t(1) << bit;
  for (int i = 0; i < kTB; ++i) {
    p.local[i] = __builtin_ctzll(uni);
    uni &= uni - 1;
  }
  int nq = 0, nt = 0;
  for (int i = 0; i < kTB; ++i) {
    if ((tmask_state >> p.local[i]) & 1) {
      p.tpos[nt] = i;
      p.gpos[nt] = gp[nt];
      UA_CHECK_ARG(tb[nt] == p.local[i], "internal: target order");
      ++nt;
    } else {
      p.qpos[nq++] = i;
    }
  }
  UA_CHECK_ARG(nt == k && nq == kTB - k, "internal: tile of %d targets", nt);
  const int S = k == 7 ? SliceShape<7, kSliceThreads>::S : SliceShape<8, kSliceThreads>::S;
  // 256 workgroups with each tile's S slices on one XCD, or (few tiles, or a
  // tile count that is not a multiple of 8) up to 256 dealt in slice order
  p.xcd = p.tiles % 8 == 0 && p.tiles * S >= kSliceGrid;
  const int64_t grid = p.xcd ? kSliceGrid : (p.tiles * S < kSliceGrid ? p.tiles * S : kSliceGrid);
  UA_CHECK_ARG(grid >= S && grid % S == 0, "internal: grid %lld", (long long)grid);
  float2* o = static_cast<float2*>(out);
  const float2* i = static_cast<const float2*>(in);
  if (k == 7)
    dense_slice_kernel<7, kSliceThreads><<<unsigned(grid), kSliceThreads, 0, s>>>(o, i, p);
  else
    dense_slice_kernel<8, kSliceThreads><<<unsigned(grid), kSliceThreads, 0, s>>>(o, i, p);
  return check_launch("dense_slice_kernel");
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
