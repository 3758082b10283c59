// Dense gates on 6..10 qubits, in place or out of place.
//
// unitair applies any k-qubit operator by permuting the targets to the front
// and contracting with einsum (unitair/simulation/operations.py:151-186,
// 258-329).  For D = 2^k >= 64 every output amplitude is a D-term sum, so
// reading a group's D inputs straight from HBM per output (2^k scattered
// reads) is far from both rooflines.  Here a workgroup owns a 4096-amplitude
// tile: the k target bits plus the 12-k lowest other bits ("local bits",
// W = 4096 / D groups).  The tile is loaded with every access in ascending
// address order of its bits (so runs are as long as the low local bits
// allow), each thread then forms 16 outputs (RT rows x GT groups, RT * GT =
// 16) from the tile in LDS and the gate's columns staged through LDS in
// chunks, and the results go back through LDS to a store with the load's
// order.  A group's inputs are all read before any output is written, so the
// kernel may run in place.
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kWideTileBits = 12;
constexpr int kWideTile = 1 << kWideTileBits;   // amplitudes per workgroup
constexpr int kWideThreads = 256;
constexpr int kPerThread = kWideTile / kWideThreads;  // 16
constexpr int kMaxWideK = 10;

struct WideParams {
  int64_t tiles;              // batch * 2^(n - 12)
  int64_t in_bstride, out_bstride, op_bstride;
  int n, k;
  int log2_tiles_per_state;   // n - 12
  int local[kWideTileBits];   // ascending state bits of the tile
  int is_target[kWideTileBits];  // local bit i is target number is_target[i] (canonical), or -1
  int w_pos[kWideTileBits];   // local bit i is group bit w_pos[i] (or -1)
  int gpos[kMaxWideK];        // gate-index bit of canonical target m
};

template <typename T>
__device__ __forceinline__ T cmad(T acc, T a, T b) {  // acc + a*b
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}

// Real states: the same kernel with a real multiply-add.
__device__ __forceinline__ float cmad(float acc, float a, float b) { return fmaf(a, b, acc); }
__device__ __forceinline__ double cmad(double acc, double a, double b) { return fma(a, b, acc); }

template <typename T, int K>
__global__ __launch_bounds__(kWideThreads) void dense_wide_kernel(T* out, const T* in,
                                                                 const T* __restrict__ op,
                                                                 WideParams p) {
  constexpr int D = 1 << K;
  constexpr int W = kWideTile / D;              // groups per tile
  constexpr int RB = D < kWideThreads ? D : kWideThreads;  // distinct rows per pass of threads
  constexpr int RT = D / RB;                    // rows per thread
  constexpr int GT = kPerThread / RT;           // groups per thread
  constexpr int GS = kWideThreads / RB;         // group stride between a thread's groups
  constexpr int CH = kWideTile / D;             // gate columns per staged chunk
  __shared__ T xs[kWideTile];                   // [group][canonical column]
  __shared__ T us[kWideTile];                   // [column in chunk][canonical row]
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int64_t b = tile >> p.log2_tiles_per_state;
  // origin of the tile: its index bits go to the non-local state bits
  uint64_t origin = insert_zero_bits(uint64_t(tile) & ((uint64_t(1) << p.log2_tiles_per_state) - 1),
                                     p.local, kWideTileBits);
  const T* src = in + b * p.in_bstride;
  T* dst = out + b * p.out_bstride;
  const T* U = op + b * p.op_bstride;

  // element e of the tile (bits of e = local bits in ascending order):
  // state offset, and its (group, canonical column) place in xs
  auto place = [&](int e, uint64_t& off, int& lds) {
    uint64_t o = 0;
    int w = 0, j = 0;
#pragma unroll
    for (int i = 0; i < kWideTileBits; ++i) {
      const int bit = (e >> i) & 1;
      o |= uint64_t(bit) << p.local[i];
      if (p.is_target[i] >= 0) j |= bit << p.is_target[i];
      else w |= bit << p.w_pos[i];
    }
    off = o;
    lds = w * D + j;
  };
  // ---- load the tile (consecutive threads take consecutive elements)
  T v[kPerThread];
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    uint64_t off;
    int l;
    place(tid + kWideThreads * i, off, l);
    v[i] = src[origin | off];
  }
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    uint64_t off;
    int l;
    place(tid + kWideThreads * i, off, l);
    xs[l] = v[i];
  }
  // gate-matrix index of canonical target pattern c
  auto orig = [&](int c) {
    int r = 0;
#pragma unroll
    for (int m = 0; m < K; ++m) r |= ((c >> m) & 1) << p.gpos[m];
    return r;
  };
  const int r0 = tid % RB, g0 = tid / RB;
  T acc[RT][GT];
#pragma unroll
  for (int h = 0; h < RT; ++h)
#pragma unroll
    for (int q = 0; q < GT; ++q) {
      acc[h][q] = T{};
    }
  for (int c0 = 0; c0 < D; c0 += CH) {
    __syncthreads();  // xs written (first chunk) / previous chunk consumed
    // stage columns c0 .. c0+CH-1 of U: us[cc * D + r] = U[orig(r)][orig(c0 + cc)]
#pragma unroll 4
    for (int e = tid; e < kWideTile; e += kWideThreads) {
      const int cc = e % CH, r = e / CH;
      us[cc * D + r] = U[orig(r) * D + orig(c0 + cc)];
    }
    __syncthreads();
#pragma unroll 2
    for (int cc = 0; cc < CH; ++cc) {
      T u[RT];
#pragma unroll
      for (int h = 0; h < RT; ++h) u[h] = us[cc * D + r0 + RB * h];
#pragma unroll
      for (int q = 0; q < GT; ++q) {
        const T x = xs[(g0 + GS * q) * D + c0 + cc];
#pragma unroll
        for (int h = 0; h < RT; ++h) acc[h][q] = cmad(acc[h][q], u[h], x);
      }
    }
  }
  __syncthreads();  // every read of xs is done: results replace the inputs
#pragma unroll
  for (int h = 0; h < RT; ++h)
#pragma unroll
    for (int q = 0; q < GT; ++q) xs[(g0 + GS * q) * D + r0 + RB * h] = acc[h][q];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    uint64_t off;
    int l;
    place(tid + kWideThreads * i, off, l);
    dst[origin | off] = xs[l];
  }
}

template <typename T>
int launch_wide(void* out, const void* in, const void* op, const WideParams& p, hipStream_t s) {
  const unsigned g = unsigned(p.tiles);
  T* o = static_cast<T*>(out);
  const T* i = static_cast<const T*>(in);
  const T* m = static_cast<const T*>(op);
  switch (p.k) {
    case 6: dense_wide_kernel<T, 6><<<g, kWideThreads, 0, s>>>(o, i, m, p); break;
    case 7: dense_wide_kernel<T, 7><<<g, kWideThreads, 0, s>>>(o, i, m, p); break;
    case 8: dense_wide_kernel<T, 8><<<g, kWideThreads, 0, s>>>(o, i, m, p); break;
    case 9: dense_wide_kernel<T, 9><<<g, kWideThreads, 0, s>>>(o, i, m, p); break;
    default: dense_wide_kernel<T, 10><<<g, kWideThreads, 0, s>>>(o, i, m, p); break;
  }
  return check_launch("dense_wide_kernel");
}

}  // namespace

// True when apply_dense_wide takes the gate: 6..10 targets and a state of at
// least 12 qubits (whole 4096-amplitude tiles within each state).
bool wide_applies(int n, int k) { return k >= 6 && k <= kMaxWideK && n >= kWideTileBits; }

// tb: target state bits in the caller's order (gate-index MSB first).
int apply_dense_wide(void* out, const void* in, const void* op, int dtype, int n, int k,
                     const int* tb, int64_t batch, int64_t in_bstride, int64_t out_bstride,
                     int64_t op_bstride, hipStream_t s) {
  UA_CHECK_ARG(wide_applies(n, k), "wide path: %d targets on %d qubits", k, n);
  WideParams p{};
  p.n = n;
  p.k = k;
  p.in_bstride = in_bstride;
  p.out_bstride = out_bstride;
  p.op_bstride = op_bstride;
  p.log2_tiles_per_state = n - kWideTileBits;
  p.tiles = batch << p.log2_tiles_per_state;
  UA_CHECK_ARG(p.tiles <= 0x7fffffffLL, "too many tiles");
  // local bits: the targets plus the lowest 12-k other bits, ascending
  uint64_t tmask = 0;
  for (int m = 0; m < k; ++m) tmask |= uint64_t(1) << tb[m];
  uint64_t lmask = tmask;
  for (int bit = 0, need = kWideTileBits - k; bit < n && need > 0; ++bit)
    if (!((tmask >> bit) & 1)) {
      lmask |= uint64_t(1) << bit;
      --need;
    }
  // canonical target order: ascending state bit; gpos = caller's gate bit
  int canon[kMaxWideK], gp[kMaxWideK];
  for (int m = 0; m < k; ++m) {
    canon[m] = tb[m];
    gp[m] = k - 1 - m;
  }
  for (int i = 1; i < k; ++i)
    for (int j = i; j > 0 && canon[j - 1] > canon[j]; --j) {
      int t0 = canon[j]; canon[j] = canon[j - 1]; canon[j - 1] = t0;
      int g0 = gp[j]; gp[j] = gp[j - 1]; gp[j - 1] = g0;
    }
  for (int m = 0; m < k; ++m) p.gpos[m] = gp[m];
  int li = 0, wi = 0;
  for (int bit = 0; bit < n; ++bit) {
    if (!((lmask >> bit) & 1)) continue;
    p.local[li] = bit;
    p.is_target[li] = -1;
    p.w_pos[li] = -1;
    for (int m = 0; m < k; ++m)
      if (canon[m] == bit) p.is_target[li] = m;
    if (p.is_target[li] < 0) p.w_pos[li] = wi++;
    ++li;
  }
  switch (dtype) {
    case UA_FLOAT32: return launch_wide<float>(out, in, op, p, s);
    case UA_FLOAT64: return launch_wide<double>(out, in, op, p, s);
    case UA_COMPLEX64: return launch_wide<float2>(out, in, op, p, s);
    default: return launch_wide<double2>(out, in, op, p, s);
  }
}

}  // namespace ua
