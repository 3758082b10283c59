// Dense 4- and 5-qubit gates on the matrix cores (MFMA), one gate per launch.
//
// For a k-qubit gate (D = 2^k) every group of D amplitudes that differ only
// in the target bits is multiplied by U (operations.py:322 contracts the
// same way).  In real arithmetic a batch of 64 groups is a (2D x 2D) x
// (2D x 64) GEMM (mfma_common.h): 8 D^2 flops per group, for 16 D bytes moved
// (complex64).  At D = 32 that is 16 flop/B, more than the vector units
// sustain without packed FP32, so these gates go to the matrix cores.
//
// Workgroup = 4 wavefronts; per chunk it stages 64 groups in LDS as real and
// imaginary rows, each wavefront multiplies 16 of them, and the results go
// back through LDS so that global reads and writes stay coalesced.  The U
// block lives in registers as MFMA A-operands for the whole (persistent)
// launch.
#include "mfma_common.h"

namespace ua {

bool g_use_mfma = true;

namespace {

constexpr int kGroupsPerChunk = 64;
constexpr int kMaxK = 5;

struct MfmaParams {
  int64_t total_groups;  // batch * 2^(n-k)
  int64_t in_bstride, out_bstride;
  int log2_groups;       // n - k
  int tb[kMaxK];         // ascending target bits
  int gpos[kMaxK];       // gate-index bit of canonical bit p
  int low_targets;       // lowest target bit < 6: stage groups with j fastest
};

template <int K>
__device__ __forceinline__ int orig_index(int c, const MfmaParams& p) {
  int r = 0;
#pragma unroll
  for (int q = 0; q < K; ++q) r |= ((c >> q) & 1) << p.gpos[q];
  return r;
}

template <int K>
__device__ __forceinline__ uint64_t offset_of(int j, const MfmaParams& p) {
  uint64_t o = 0;
#pragma unroll
  for (int q = 0; q < K; ++q)
    if ((j >> q) & 1) o += uint64_t(1) << p.tb[q];
  return o;
}

// LDS position of (row, column) in a chunk tile of 64 columns.  The column
// is XORed with a row-dependent mask so that every access pattern below is
// free of bank conflicts on 64 four-byte banks: rows differing in bits 0-1
// land in different 16-column blocks (MFMA B reads), rows differing in bits
// 2-4 land on different banks within a block (j-fastest staging, where 16 or
// 32 consecutive rows are written in one instruction).
__device__ __forceinline__ int tile_pos(int row, int col) {
  const int s = (((row ^ (row >> 2)) & 3) << 4) | (row & 12) | ((row >> 3) & 2);
  return row * kGroupsPerChunk + (col ^ s);
}

template <typename T, int K>
__global__ __launch_bounds__(kBlock) void dense_mfma_kernel(T* out, const T* in,
                                                            const T* __restrict__ op,
                                                            MfmaParams p) {
  using R = typename RealOf<T>::type;
  using M = Mfma<R>;
  constexpr int D = 1 << K;
  constexpr int R2 = 2 * D;        // real block size
  constexpr int RT = R2 / 16;      // row tiles
  constexpr int KS = R2 / 4;       // k-steps
  __shared__ R xs[R2 * kGroupsPerChunk];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  // ---- U as MFMA A-operands: a[ti][ks] = Ublock[16 ti + lane%16][4 ks + lane/16]
  R a[RT][KS];
#pragma unroll
  for (int ti = 0; ti < RT; ++ti)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int row = 16 * ti + (lane & 15), col = 4 * ks + (lane >> 4);
      const int rr = row & (D - 1), cc = col & (D - 1);
      const T u = op[orig_index<K>(rr, p) * D + orig_index<K>(cc, p)];
      a[ti][ks] = real_block(u, row >= D, col >= D);
    }

  const int64_t chunks = p.total_groups / kGroupsPerChunk;  // whole chunks (host-checked)
  const uint64_t gmask = (uint64_t(1) << p.log2_groups) - 1;
  // Staging order.  Each thread moves D/4 elements per chunk.
  //  gl-fastest (targets all >= bit 6): gl = tid % 64 is fixed, j = tid/64 + 4i
  //  j-fastest  (a target below bit 6): j = tid % D is fixed, gl = tid/D + (256/D) i
  // Chunks start at multiples of 64 groups and a state holds a multiple of
  // 64 groups, so the bits of (chunk start) and gl are disjoint and the
  // address of element i is base(chunk) + pre[i] with
  //   pre[i] = deposit(gl_i) + offset(j_i),
  // computed once: no per-element index arithmetic in the loop.
  constexpr int PER = D / 4;
  constexpr int GSTEP = kBlock / D;
  int64_t pre[PER];
  int jj[PER], glv[PER];
  const bool jfast = p.low_targets != 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    jj[i] = jfast ? (tid & (D - 1)) : ((tid >> 6) + 4 * i);
    glv[i] = jfast ? (tid / D) + GSTEP * i : (tid & (kGroupsPerChunk - 1));
    pre[i] = int64_t(insert_zero_bits(uint64_t(glv[i]), p.tb, K) + offset_of<K>(jj[i], p));
  }
  auto base_of = [&](int64_t chunk, int64_t bstride) -> int64_t {
    const int64_t g0 = chunk * kGroupsPerChunk;
    return (g0 >> p.log2_groups) * bstride +
           int64_t(insert_zero_bits(uint64_t(g0) & gmask, p.tb, K));
  };

  T cur[PER];
  int64_t chunk = blockIdx.x;
  if (chunk < chunks) {
    const int64_t base = base_of(chunk, p.in_bstride);
#pragma unroll
    for (int i = 0; i < PER; ++i) cur[i] = in[base + pre[i]];
  }
  for (; chunk < chunks; chunk += gridDim.x) {
    // ---- stage this chunk: xs[j][gl] = Re, xs[D + j][gl] = Im
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      xs[tile_pos(jj[i], glv[i])] = cur[i].x;
      xs[tile_pos(D + jj[i], glv[i])] = cur[i].y;
    }
    __syncthreads();
    // ---- issue the next chunk's reads (loads overlap the matrix products)
    const int64_t next = chunk + gridDim.x;
    T nxt[PER];
    if (next < chunks) {
      const int64_t base = base_of(next, p.in_bstride);
#pragma unroll
      for (int i = 0; i < PER; ++i) nxt[i] = in[base + pre[i]];
    }
    // ---- 16 groups per wavefront
    typename M::acc_t acc[RT];
#pragma unroll
    for (int ti = 0; ti < RT; ++ti) acc[ti] = typename M::acc_t{0, 0, 0, 0};
    const int col = 16 * wave + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const R bfrag = xs[tile_pos(4 * ks + (lane >> 4), col)];
#pragma unroll
      for (int ti = 0; ti < RT; ++ti) acc[ti] = M::mma(a[ti][ks], bfrag, acc[ti]);
    }
    __syncthreads();
#pragma unroll
    for (int ti = 0; ti < RT; ++ti)
#pragma unroll
      for (int r = 0; r < 4; ++r) xs[tile_pos(16 * ti + M::row(lane, r), col)] = acc[ti][r];
    __syncthreads();
    // ---- store (same element order as the load)
    const int64_t obase = base_of(chunk, p.out_bstride);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      T v;
      v.x = xs[tile_pos(jj[i], glv[i])];
      v.y = xs[tile_pos(D + jj[i], glv[i])];
      out[obase + pre[i]] = v;
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) cur[i] = nxt[i];
    __syncthreads();
  }
}

template <typename T>
int launch(int k, void* out, const void* in, const void* op, const MfmaParams& p,
           hipStream_t s) {
  const int64_t chunks = p.total_groups / kGroupsPerChunk;
  int64_t blocks = chunks < 2048 ? chunks : 2048;
  if (blocks < 1) blocks = 1;
  if (k == 4)
    dense_mfma_kernel<T, 4><<<unsigned(blocks), kBlock, 0, s>>>(
        static_cast<T*>(out), static_cast<const T*>(in), static_cast<const T*>(op), p);
  else
    dense_mfma_kernel<T, 5><<<unsigned(blocks), kBlock, 0, s>>>(
        static_cast<T*>(out), static_cast<const T*>(in), static_cast<const T*>(op), p);
  return check_launch("dense_mfma_kernel");
}

}  // namespace

// Called by apply_operator_impl for complex 4/5-qubit gates with one operator.
// tb/gp: targets in canonical order (ascending state bit) and their gate bits.
int apply_dense_mfma(void* out, const void* in, const void* op, int dtype, int n, int k,
                     const int* tb, const int* gp, int64_t batch, int64_t in_bstride,
                     hipStream_t s) {
  MfmaParams p{};
  p.log2_groups = n - k;
  p.total_groups = batch << p.log2_groups;
  p.in_bstride = in_bstride;
  p.out_bstride = int64_t(1) << n;
  for (int m = 0; m < k; ++m) {
    p.tb[m] = tb[m];
    p.gpos[m] = gp[m];
  }
  p.low_targets = tb[0] < 6 ? 1 : 0;
  // whole chunks of 64 groups, none across two states
  UA_CHECK_ARG(n - k >= 6, "matrix-core path needs n - k >= 6");
  if (dtype == UA_COMPLEX64) return launch<float2>(k, out, in, op, p, s);
  return launch<double2>(k, out, in, op, p, s);
}

}  // namespace ua
