// Kernels for adjoint differentiation of Pauli-rotation circuits with the
// observable H = sum_i Z_i.
//
// unitair has no gradient code: its circuits are differentiated by torch
// autograd through apply_operator / apply_phase (unitair/simulation/
// operations.py:41-42,322), the gate builders (unitair/gates/gates.py:35-91)
// and the loss diag_expectation_value (unitair/states/innerprod.py:49-59),
// which keeps one saved state per gate.  The adjoint method keeps two states
// (psi and the co-state lambda = H psi pulled back through the gates) and
// needs, per rotation exp(-i t/2 P_q):
//     dL/dt = Im <lambda| P_q |psi>
// evaluated at the point right after that gate.  Same-kind rotations on
// different qubits commute, so one (lambda, psi) point serves every qubit of a
// layer: ua_pauli_gradients returns all n values of a layer in one call.
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kLowBits = 12;
constexpr int kLowTile = 1 << kLowBits;
constexpr int kMaxN = 40;
constexpr int kZChunkConst = 4096;

template <typename T> struct D2;
template <> struct D2<float2> {
  __device__ static double2 get(float2 v) { return make_double2(v.x, v.y); }
};
template <> struct D2<double2> {
  __device__ static double2 get(double2 v) { return v; }
};

// Re(conj(a) * b)
__device__ __forceinline__ double re_cj(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
// Im(conj(a) * b)
__device__ __forceinline__ double im_cj(double2 a, double2 b) { return a.x * b.y - a.y * b.x; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sums v over the workgroup (256 threads); result in thread 0.
__device__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) r += scratch[w];
  return r;
}

// Im <lambda| P_b |psi> summed over one amplitude pair (x0 with bit b clear,
// x1 = x0 | 2^b), for P = X or Y:
//   X: (P psi)_x = psi_{x^b}             -> Im(l0* p1) + Im(l1* p0)
//   Y: (P psi)_{x0} = -i psi_{x1},
//      (P psi)_{x1} =  i psi_{x0}        -> Re(l1* p0) - Re(l0* p1)
template <int P>
__device__ __forceinline__ double pair_term(double2 l0, double2 l1, double2 p0, double2 p1) {
  if constexpr (P == UA_PAULI_X) return im_cj(l0, p1) + im_cj(l1, p0);
  else return re_cj(l1, p0) - re_cj(l0, p1);
}

// X / Y gradients for index bits 0..min(n,12)-1.  One workgroup per (state,
// tile of 2^m contiguous amplitudes); lambda and psi tiles sit in LDS.
// partial[(b * tiles + tile) * n + bit] = sum over the tile's pairs.
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void grad_low_kernel(
    double* __restrict__ partial, const T* __restrict__ lam, const T* __restrict__ psi,
    int n, int m, int64_t tiles_per_state) {
  __shared__ T sl[kLowTile];
  __shared__ T sp[kLowTile];
  __shared__ double scratch[kBlock / 64];
  const int64_t tile = blockIdx.x;
  const int64_t b = tile / tiles_per_state;
  const int64_t t = tile % tiles_per_state;
  const int64_t N = int64_t(1) << n;
  const int size = 1 << m;
  const T* L = lam + b * N + t * size;
  const T* Ps = psi + b * N + t * size;
  for (int i = threadIdx.x; i < size; i += kBlock) {
    sl[i] = L[i];
    sp[i] = Ps[i];
  }
  __syncthreads();
  for (int bit = 0; bit < m; ++bit) {
    double acc = 0.0;
    const int half = size >> 1;
    for (int pidx = threadIdx.x; pidx < half; pidx += kBlock) {
      const int x0 = ((pidx >> bit) << (bit + 1)) | (pidx & ((1 << bit) - 1));
      const int x1 = x0 | (1 << bit);
      acc += pair_term<P>(D2<T>::get(sl[x0]), D2<T>::get(sl[x1]), D2<T>::get(sp[x0]),
                          D2<T>::get(sp[x1]));
    }
    const double s = block_sum(acc, scratch);
    if (threadIdx.x == 0) partial[(b * tiles_per_state + t) * n + bit] = s;
  }
}

// pair_term in the state's precision, returned as double (the sums are
// float64): complex64 terms are formed in float32 (relative rounding 6e-8
// per term, far inside complex64 state error), complex128 in double.
template <int P>
__device__ __forceinline__ double pair_term_of(float2 l0, float2 l1, float2 p0, float2 p1) {
  float r;
  if constexpr (P == UA_PAULI_X)
    r = (l0.x * p1.y - l0.y * p1.x) + (l1.x * p0.y - l1.y * p0.x);
  else
    r = (l1.x * p0.x + l1.y * p0.y) - (l0.x * p1.x + l0.y * p1.y);
  return double(r);
}
template <int P>
__device__ __forceinline__ double pair_term_of(double2 l0, double2 l1, double2 p0, double2 p1) {
  return pair_term<P>(l0, l1, p0, p1);
}

// The same sums for a full 4096-amplitude tile (m = 12, every state of 12 or
// more qubits) with the work in registers: after a coalesced copy of both
// tiles into LDS, each thread reads three 16-amplitude cubes (tile bits 0..3,
// 4..7, 8..11) and adds the pair terms of the cube's 4 bits; the 12 sums stay
// in registers until one reduction at the end.  LDS position of element e:
// e ^ ((e >> 4) & 15), so the copy and all three cube reads touch distinct
// banks across 16 lanes.
__device__ __forceinline__ int low_swz(int e) { return e ^ ((e >> 4) & 15); }

template <int P, int B0, typename T>
__device__ __forceinline__ void low_cube(double (&acc)[kLowBits], const T* sl, const T* sp,
                                         int e0) {
  T l[16], p[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    l[j] = sl[low_swz(e0 | (j << B0))];
    p[j] = sp[low_swz(e0 | (j << B0))];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if (j & (1 << k)) continue;
      const int j1 = j | (1 << k);
      acc[B0 + k] += pair_term_of<P>(l[j], l[j1], p[j], p[j1]);
    }
}

// complex64 tiles (64 KiB of LDS) leave room for two workgroups per CU:
// the register budget asks for both
template <typename T>
constexpr int low_full_waves() { return sizeof(T) == 8 ? 2 : 1; }

template <typename T, int P>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(low_full_waves<T>()))) void grad_low_full_kernel(
    double* __restrict__ partial, const T* __restrict__ lam, const T* __restrict__ psi, int n,
    int64_t tiles_per_state) {
  __shared__ T sl[kLowTile];
  __shared__ T sp[kLowTile];
  __shared__ double red[kBlock / 64][kLowBits];
  const int64_t tile = blockIdx.x;
  const int64_t b = tile / tiles_per_state;
  const int64_t t = tile % tiles_per_state;
  const int64_t N = int64_t(1) << n;
  const int tid = threadIdx.x;
  const T* L = lam + b * N + t * kLowTile;
  const T* Ps = psi + b * N + t * kLowTile;
#pragma unroll
  for (int k = 0; k < kLowTile / kBlock; ++k) {
    const int i = tid + k * kBlock;
    sl[low_swz(i)] = L[i];
    sp[low_swz(i)] = Ps[i];
  }
  __syncthreads();
  double acc[kLowBits];
#pragma unroll
  for (int i = 0; i < kLowBits; ++i) acc[i] = 0.0;
  // thread tid's cube entry 0: tid's 8 bits around the cube's 4
  // (the empty asm keeps each cube's LDS reads after the previous cube's
  // arithmetic: with all three cubes' reads hoisted the registers overflow)
  low_cube<P, 0>(acc, sl, sp, tid << 4);
  asm volatile("" ::: "memory");
  low_cube<P, 4>(acc, sl, sp, (tid & 15) | ((tid >> 4) << 8));
  asm volatile("" ::: "memory");
  low_cube<P, 8>(acc, sl, sp, tid);
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < kLowBits; ++i) {
    const double v = wave_sum(acc[i]);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (tid < kLowBits) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r += red[w][tid];
    partial[(b * tiles_per_state + t) * n + tid] = r;
  }
}

// X / Y gradients for a window of H index bits w0..w0+H-1 (w0 >= 12).  Each
// thread owns one low index (bits 0..11: 256 consecutive per workgroup, so
// every load is coalesced) and, for each value r of the other high bits in
// its block's range, the 2^H amplitudes over the window.
// partial[((b * 16 + blockIdx.y) * gridDim.z + blockIdx.z) * n + bit].
template <typename T, int P, int H>
__global__ __launch_bounds__(kBlock) void grad_high_kernel(
    double* __restrict__ partial, const T* __restrict__ lam, const T* __restrict__ psi, int n,
    int w0, int64_t rest_per_block) {
  __shared__ double scratch[kBlock / 64];
  constexpr int HN = 1 << H;
  const int64_t b = blockIdx.x;
  const int low = blockIdx.y * kBlock + threadIdx.x;  // < 4096
  const int64_t N = int64_t(1) << n;
  const int below = w0 - kLowBits;                    // other high bits under the window
  double acc[H];
#pragma unroll
  for (int hb = 0; hb < H; ++hb) acc[hb] = 0.0;
  const int64_t r0 = int64_t(blockIdx.z) * rest_per_block;
  for (int64_t r = r0; r < r0 + rest_per_block; ++r) {
    const uint64_t rest = ((uint64_t(r) & ((uint64_t(1) << below) - 1)) << kLowBits) |
                          ((uint64_t(r) >> below) << (w0 + H));
    const int64_t base = b * N + int64_t(rest) + low;
    double2 l[HN], p[HN];
#pragma unroll
    for (int h = 0; h < HN; ++h) {
      l[h] = D2<T>::get(lam[base + (int64_t(h) << w0)]);
      p[h] = D2<T>::get(psi[base + (int64_t(h) << w0)]);
    }
#pragma unroll
    for (int hb = 0; hb < H; ++hb) {
#pragma unroll
      for (int h = 0; h < HN; ++h) {
        if (h & (1 << hb)) continue;
        const int h1 = h | (1 << hb);
        acc[hb] += pair_term<P>(l[h], l[h1], p[h], p[h1]);
      }
    }
  }
  const int64_t part = (b * gridDim.y + blockIdx.y) * gridDim.z + blockIdx.z;
#pragma unroll
  for (int hb = 0; hb < H; ++hb) {
    const double s = block_sum(acc[hb], scratch);
    if (threadIdx.x == 0) partial[part * n + w0 + hb] = s;
  }
}

// Z gradients (and, with want_loss, the loss sum_x |psi_x|^2 (n-2 popcount x)).
// A workgroup covers a chunk of 4096 amplitudes of one state (or the whole
// state when it is smaller); thread t visits i = chunk_base + t + 256*k.
// Index bits 0..7 of i are fixed per thread, bits 8..11 follow k and bits
// >= 12 are fixed per chunk, so each thread keeps 16 partial sums (one per k)
// and the per-qubit signs are applied afterwards.
// partial[(b * chunks + c) * (n + 1) + q]; slot n holds the loss part.
template <typename T>
__global__ __launch_bounds__(kBlock) void grad_z_kernel(
    double* __restrict__ partial, const T* __restrict__ lam, const T* __restrict__ psi, int n,
    int64_t chunks, int want_loss) {
  constexpr int K = kZChunkConst / kBlock;  // 16
  const int64_t b = blockIdx.x / chunks;
  const int64_t c = blockIdx.x % chunks;
  const int64_t N = int64_t(1) << n;
  const int64_t base = c * kZChunkConst;
  double vk[K];
  double loss = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    vk[k] = 0.0;
    const int64_t i = base + threadIdx.x + int64_t(k) * kBlock;
    if (i < N) {
      const double2 pv = D2<T>::get(psi[b * N + i]);
      if (want_loss) {
        const double w = pv.x * pv.x + pv.y * pv.y;
        loss += w * double(n - 2 * int(__popcll(uint64_t(i))));
      }
      if (lam) vk[k] = im_cj(D2<T>::get(lam[b * N + i]), pv);
    }
  }
  double total = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) total += vk[k];
  // this thread's share of each low bit's signed sum (bits >= 12 are fixed
  // per chunk: plus or minus the chunk's total), the total and the loss,
  // reduced together once
  constexpr int kVals = kLowBits + 2;
  double vals[kVals];
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) vals[bit] = ((threadIdx.x >> bit) & 1) ? -total : total;
#pragma unroll
  for (int bit = 8; bit < kLowBits; ++bit) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) v += ((k >> (bit - 8)) & 1) ? -vk[k] : vk[k];
    vals[bit] = v;
  }
  vals[kLowBits] = total;
  vals[kLowBits + 1] = loss;
  __shared__ double red[kBlock / 64][kVals];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kVals; ++i) {
    const double v = wave_sum(vals[i]);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < kVals) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r += red[w][threadIdx.x];
    red[0][threadIdx.x] = r;  // each thread reads and writes only its own column
  }
  __syncthreads();
  double* out = partial + (b * chunks + c) * (n + 1);
  if (want_loss && threadIdx.x == 0) out[n] = red[0][kLowBits + 1];
  if (!lam) return;
  for (int bit = threadIdx.x; bit < n; bit += kBlock) {
    const double sum = bit < kLowBits ? red[0][bit]
                                      : (((base >> bit) & 1) ? -red[0][kLowBits]
                                                             : red[0][kLowBits]);
    out[n - 1 - bit] = sum;
  }
}

// out[b * out_stride + out_q0 + out_dir * q] = sum over `parts` rows of
// partial[(b*parts + r)*stride + q0 + q], q < n_out (out_dir -1: index bit
// q0 + q is qubit out_q0 - q)
__global__ __launch_bounds__(kBlock) void sum_parts_kernel(
    double* __restrict__ out, const double* __restrict__ partial, int64_t batch, int64_t parts,
    int stride, int q0, int n_out, int out_stride, int out_q0, int accumulate, int out_dir) {
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= batch * n_out) return;
  const int64_t b = idx / n_out;
  const int q = int(idx % n_out);
  double s = 0.0;
  for (int64_t r = 0; r < parts; ++r) s += partial[(b * parts + r) * stride + q0 + q];
  double* o = out + b * out_stride + out_q0 + out_dir * q;
  *o = accumulate ? *o + s : s;
}

// lambda_x = (n - 2 popcount(x)) psi_x
template <typename T>
__global__ __launch_bounds__(kBlock) void sum_z_apply_kernel(T* __restrict__ out,
                                                             const T* __restrict__ psi, int n,
                                                             int64_t total) {
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  const uint64_t mask = (uint64_t(1) << n) - 1;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
    const auto w = decltype(psi[0].x)(n - 2 * int(__popcll(uint64_t(i) & mask)));
    T v = psi[i];
    v.x *= w;
    v.y *= w;
    out[i] = v;
  }
}

int64_t y_low_parts(int n) { return n > kLowBits ? (int64_t(1) << (n - kLowBits)) : 1; }
constexpr int64_t kHighParts = kLowTile / kBlock;  // 16 low chunks per state
constexpr int kHighWindow = 4;                      // index bits per high-bit launch
// Splits of the other high bits of a window launch across blockIdx.z: enough
// workgroups to fill the GPU (about 2048 per launch) with at least one value
// of the other bits each.
int64_t high_z(int n, int64_t batch) {
  const int rest_bits = n - kLowBits - kHighWindow;
  if (rest_bits <= 0) return 1;
  int64_t z = (2048 + batch * kHighParts - 1) / (batch * kHighParts);
  const int64_t rest = int64_t(1) << rest_bits;
  if (z > rest) z = rest;
  int64_t p2 = 1;
  while (p2 * 2 <= z) p2 *= 2;   // a power of two divides the rest count
  return p2;
}
int64_t z_parts(int n) {
  return ((int64_t(1) << n) + kZChunkConst - 1) / kZChunkConst;
}

}  // namespace
}  // namespace ua

using namespace ua;

extern "C" int64_t ua_pauli_gradients_workspace(int num_qubits, int64_t batch) {
  const int n = num_qubits;
  if (n < 1 || n > kMaxN || batch < 0) return -1;
  int64_t a = y_low_parts(n) * n;
  const int64_t h = kHighParts * high_z(n, batch < 1 ? 1 : batch) * n;
  const int64_t z = z_parts(n) * (n + 1);
  if (h > a) a = h;
  if (z > a) a = z;
  return (batch < 1 ? 1 : batch) * a * int64_t(sizeof(double));
}

extern "C" int ua_pauli_gradients(double* out, const void* lambda, const void* psi, int dtype,
                                  int num_qubits, int64_t batch, int pauli, void* workspace,
                                  ua_stream_t stream) {
  const int n = num_qubits;
  hipStream_t s = static_cast<hipStream_t>(stream);
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128, "complex states only");
  UA_CHECK_ARG(n >= 1 && n <= kMaxN, "num_qubits=%d out of range", n);
  UA_CHECK_ARG(pauli == UA_PAULI_X || pauli == UA_PAULI_Y || pauli == UA_PAULI_Z,
               "pauli must be X, Y or Z");
  UA_CHECK_ARG(batch >= 0, "negative batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && lambda && psi && workspace, "NULL pointer");
  double* part = static_cast<double*>(workspace);
  const bool c64 = dtype == UA_COMPLEX64;
  if (pauli == UA_PAULI_Z) {
    const int64_t chunks = z_parts(n);
    UA_CHECK_ARG(batch * chunks <= 0x7fffffffLL, "grid too large");
    const unsigned g = unsigned(batch * chunks);
    if (c64)
      grad_z_kernel<float2><<<g, kBlock, 0, s>>>(part, (const float2*)lambda, (const float2*)psi,
                                                 n, chunks, 0);
    else
      grad_z_kernel<double2><<<g, kBlock, 0, s>>>(part, (const double2*)lambda,
                                                  (const double2*)psi, n, chunks, 0);
    int rc = check_launch("grad_z_kernel");
    if (rc) return rc;
    // out[b*n + q] = sum over chunks (partials are stored by qubit)
    sum_parts_kernel<<<grid_for(batch * n), kBlock, 0, s>>>(out, part, batch, chunks, n + 1, 0, n,
                                                            n, 0, 0, 1);
    return check_launch("sum_parts_kernel");
  }
  // X / Y: index bit i belongs to qubit n-1-i; partials are kept per index
  // bit and mapped to qubits when summed.
  const int m = n < kLowBits ? n : kLowBits;
  const int64_t tiles = y_low_parts(n);
  UA_CHECK_ARG(batch * tiles <= 0x7fffffffLL, "grid too large");
  const unsigned glow = unsigned(batch * tiles);
#define UA_LOW(TT, PP)                                                                   \
  do {                                                                                   \
    if (m == kLowBits)                                                                   \
      grad_low_full_kernel<TT, PP><<<glow, kBlock, 0, s>>>(part, (const TT*)lambda,      \
                                                           (const TT*)psi, n, tiles);     \
    else                                                                                 \
      grad_low_kernel<TT, PP><<<glow, kBlock, 0, s>>>(part, (const TT*)lambda,           \
                                                      (const TT*)psi, n, m, tiles);       \
  } while (0)
  if (pauli == UA_PAULI_X) {
    if (c64) UA_LOW(float2, UA_PAULI_X); else UA_LOW(double2, UA_PAULI_X);
  } else {
    if (c64) UA_LOW(float2, UA_PAULI_Y); else UA_LOW(double2, UA_PAULI_Y);
  }
#undef UA_LOW
  int rc = check_launch("grad_low_kernel");
  if (rc) return rc;
  // bits 0..m-1 -> qubits n-1 .. n-m, one launch
  sum_parts_kernel<<<grid_for(batch * m), kBlock, 0, s>>>(out, part, batch, tiles, n, 0, m, n,
                                                          n - 1, 0, -1);
  rc = check_launch("sum_parts_kernel");
  if (rc) return rc;
  // high bits, kHighWindow at a time; each launch reads both states once
  UA_CHECK_ARG(batch <= 0x7fffffffLL, "grid too large");
  const int64_t z = high_z(n, batch);
  for (int w0 = kLowBits; w0 < n; w0 += kHighWindow) {
    const int H = (n - w0) < kHighWindow ? (n - w0) : kHighWindow;
    const int rest_bits = n - kLowBits - H;
    const int64_t zz = rest_bits > 0 ? z : 1;
    const int64_t per_block = (int64_t(1) << (rest_bits > 0 ? rest_bits : 0)) / zz;
    const dim3 grid = dim3(unsigned(batch), unsigned(kHighParts), unsigned(zz));
#define UA_HIGH(HH)                                                                             \
  if (H == HH) {                                                                               \
    if (pauli == UA_PAULI_X) {                                                                 \
      if (c64)                                                                                 \
        grad_high_kernel<float2, UA_PAULI_X, HH><<<grid, kBlock, 0, s>>>(                      \
            part, (const float2*)lambda, (const float2*)psi, n, w0, per_block);                \
      else                                                                                     \
        grad_high_kernel<double2, UA_PAULI_X, HH><<<grid, kBlock, 0, s>>>(                     \
            part, (const double2*)lambda, (const double2*)psi, n, w0, per_block);              \
    } else {                                                                                   \
      if (c64)                                                                                 \
        grad_high_kernel<float2, UA_PAULI_Y, HH><<<grid, kBlock, 0, s>>>(                      \
            part, (const float2*)lambda, (const float2*)psi, n, w0, per_block);                \
      else                                                                                     \
        grad_high_kernel<double2, UA_PAULI_Y, HH><<<grid, kBlock, 0, s>>>(                     \
            part, (const double2*)lambda, (const double2*)psi, n, w0, per_block);              \
    }                                                                                          \
  }
    UA_HIGH(1) UA_HIGH(2) UA_HIGH(3) UA_HIGH(4)
#undef UA_HIGH
    rc = check_launch("grad_high_kernel");
    if (rc) return rc;
    // bits w0..w0+H-1 -> qubits n-1-w0 .. n-w0-H, one launch
    sum_parts_kernel<<<grid_for(batch * H), kBlock, 0, s>>>(out, part, batch, kHighParts * zz, n,
                                                            w0, H, n, n - 1 - w0, 0, -1);
    rc = check_launch("sum_parts_kernel");
    if (rc) return rc;
  }
  return UA_OK;
}

namespace ua {
namespace {
// out_x = d_x psi_x with a real diagonal d (one row, or one row per state)
template <typename T, typename D>
__global__ __launch_bounds__(kBlock) void diag_scale_kernel(T* __restrict__ out,
                                                            const T* __restrict__ psi,
                                                            const D* __restrict__ diag, int n,
                                                            int64_t total, int64_t dstride) {
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  const uint64_t mask = (uint64_t(1) << n) - 1;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
    const auto w = decltype(psi[0].x)(diag[(i >> n) * dstride + int64_t(uint64_t(i) & mask)]);
    T v = psi[i];
    v.x *= w;
    v.y *= w;
    out[i] = v;
  }
}
}  // namespace
}  // namespace ua

extern "C" int ua_apply_diagonal(void* out, const void* psi, int dtype, const void* diag,
                                 int diag_dtype, int64_t diag_batch_stride, int num_qubits,
                                 int64_t batch, ua_stream_t stream) {
  const int n = num_qubits;
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128, "complex states only");
  UA_CHECK_ARG(diag_dtype == UA_FLOAT32 || diag_dtype == UA_FLOAT64, "real diagonal only");
  UA_CHECK_ARG(n >= 0 && n <= 62 && batch >= 0 && diag_batch_stride >= 0, "bad sizes");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && psi && diag, "NULL pointer");
  const int64_t total = batch << n;
  int64_t blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = unsigned(blocks);
#define UA_DS(T, D)                                                                     \
  diag_scale_kernel<T, D><<<g, kBlock, 0, s>>>((T*)out, (const T*)psi, (const D*)diag, n, \
                                               total, diag_batch_stride)
  if (dtype == UA_COMPLEX64) {
    if (diag_dtype == UA_FLOAT32) UA_DS(float2, float); else UA_DS(float2, double);
  } else {
    if (diag_dtype == UA_FLOAT32) UA_DS(double2, float); else UA_DS(double2, double);
  }
#undef UA_DS
  return check_launch("diag_scale_kernel");
}

extern "C" int ua_apply_sum_z(void* out, const void* psi, int dtype, int num_qubits,
                              int64_t batch, ua_stream_t stream) {
  const int n = num_qubits;
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128, "complex states only");
  UA_CHECK_ARG(n >= 1 && n <= 62 && batch >= 0, "bad sizes");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && psi, "NULL pointer");
  const int64_t total = batch << n;
  int64_t blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == UA_COMPLEX64)
    sum_z_apply_kernel<float2><<<unsigned(blocks), kBlock, 0, s>>>((float2*)out,
                                                                   (const float2*)psi, n, total);
  else
    sum_z_apply_kernel<double2><<<unsigned(blocks), kBlock, 0, s>>>((double2*)out,
                                                                    (const double2*)psi, n, total);
  return check_launch("sum_z_apply_kernel");
}

extern "C" int ua_expect_sum_z(double* out, const void* psi, int dtype, int num_qubits,
                               int64_t batch, void* workspace, ua_stream_t stream) {
  const int n = num_qubits;
  hipStream_t s = static_cast<hipStream_t>(stream);
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128, "complex states only");
  UA_CHECK_ARG(n >= 1 && n <= kMaxN && batch >= 0, "bad sizes");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && psi && workspace, "NULL pointer");
  double* part = static_cast<double*>(workspace);
  const int64_t chunks = z_parts(n);
  UA_CHECK_ARG(batch * chunks <= 0x7fffffffLL, "grid too large");
  const unsigned g = unsigned(batch * chunks);
  if (dtype == UA_COMPLEX64)
    grad_z_kernel<float2><<<g, kBlock, 0, s>>>(part, nullptr, (const float2*)psi, n, chunks, 1);
  else
    grad_z_kernel<double2><<<g, kBlock, 0, s>>>(part, nullptr, (const double2*)psi, n, chunks, 1);
  int rc = check_launch("grad_z_kernel");
  if (rc) return rc;
  sum_parts_kernel<<<grid_for(batch), kBlock, 0, s>>>(out, part, batch, chunks, n + 1, n, 1, 1, 0,
                                                      0, 1);
  return check_launch("sum_parts_kernel");
}

// ---------------------------------------------------------------------------
// Gradient of apply_operator with respect to the operator (backward of
// unitair/simulation/operations.py:322, out = U @ in on the target bits):
//   grad_U[o][r][c] = sum over batch entries b using operator o, and over
//                     groups g:  grad_out[b][g|off(r)] * conj(in[b][g|off(c)])
// i.e. grad_U = G^T conj(X) with G, X the (groups x 2^k) views of grad_out
// and in: a tall-skinny product whose reduction runs over 2^(n-k) groups.
//
// op_grad_partial_kernel: a workgroup owns one operator o, one slice ("part")
// of that operator's groups and one RB x CB block of the 2^k x 2^k result.
// Each thread walks the slice with stride 256 (consecutive threads take
// consecutive groups, so every load is coalesced), keeps the block's RB*CB
// products in float64 registers, and the workgroup reduces them with
// wavefront shuffles and LDS.  With one part the sums go straight to grad_op;
// otherwise to the float64 workspace, and op_grad_finish_kernel adds the
// parts (one workgroup per result entry).  k = 1, 2 take the whole result in one block (each amplitude read
// once); larger k tile it 4 x 4.
namespace ua {
namespace {

struct OpGradParams {
  int n, k;
  int64_t batch, op_count;   // op_count: distinct operators (1 or batch)
  int64_t in_bstride;        // 0 when one input state serves the whole batch
  int64_t parts;             // slices of each operator's groups
  int64_t units;             // groups per operator: (batch / op_count) << (n - k)
  int64_t units_per_part;
  int nblk_c;                // column blocks: 2^k / CB
  int tbit[UA_MAX_GRAD_QUBITS];    // state bit of gate-index bit k-1-m (qubits order)
  int sorted[UA_MAX_GRAD_QUBITS];  // target state bits, ascending
};

template <typename T> __device__ __forceinline__ double2 as_d2(T v);
template <> __device__ __forceinline__ double2 as_d2(float v) { return make_double2(v, 0.0); }
template <> __device__ __forceinline__ double2 as_d2(double v) { return make_double2(v, 0.0); }
template <> __device__ __forceinline__ double2 as_d2(float2 v) { return make_double2(v.x, v.y); }
template <> __device__ __forceinline__ double2 as_d2(double2 v) { return v; }

template <typename T> __device__ __forceinline__ T from_d2(double2 v);
template <> __device__ __forceinline__ float from_d2(double2 v) { return float(v.x); }
template <> __device__ __forceinline__ double from_d2(double2 v) { return v.x; }
template <> __device__ __forceinline__ float2 from_d2(double2 v) { return make_float2(float(v.x), float(v.y)); }
template <> __device__ __forceinline__ double2 from_d2(double2 v) { return v; }

__device__ __forceinline__ uint64_t gate_offset(int idx, const OpGradParams& p) {
  uint64_t o = 0;
  for (int m = 0; m < p.k; ++m)
    if ((idx >> (p.k - 1 - m)) & 1) o |= uint64_t(1) << p.tbit[m];
  return o;
}

template <typename T, int RB, int CB, int UNR>
__global__ __launch_bounds__(kBlock) void op_grad_partial_kernel(
    T* __restrict__ grad_op, double* __restrict__ part_out, const T* __restrict__ gout,
    const T* __restrict__ in, OpGradParams p) {
  __shared__ double2 red[kBlock / 64][RB * CB];
  const int64_t D = int64_t(1) << p.k;
  const int blk = int(blockIdx.x % unsigned(p.nblk_c * (D / RB)));
  const int64_t rest = int64_t(blockIdx.x) / (p.nblk_c * (D / RB));
  const int64_t part = rest % p.parts;
  const int64_t o = rest / p.parts;
  const int r0 = (blk / p.nblk_c) * RB, c0 = (blk % p.nblk_c) * CB;
  uint64_t offr[RB], offc[CB];
#pragma unroll
  for (int i = 0; i < RB; ++i) offr[i] = gate_offset(r0 + i, p);
#pragma unroll
  for (int j = 0; j < CB; ++j) offc[j] = gate_offset(c0 + j, p);

  const int gbits = p.n - p.k;
  const int64_t N = int64_t(1) << p.n;
  const int64_t u_begin = part * p.units_per_part;
  int64_t u_end = u_begin + p.units_per_part;
  if (u_end > p.units) u_end = p.units;
  double2 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) acc[i][j] = make_double2(0.0, 0.0);
  // UNR groups per iteration, all their loads issued before any product
  for (int64_t u0 = u_begin + threadIdx.x; u0 < u_end; u0 += UNR * kBlock) {
    T gr[UNR][RB], xr[UNR][CB];
#pragma unroll
    for (int q = 0; q < UNR; ++q) {
      const int64_t u = u0 + q * kBlock;
      if (u >= u_end) continue;
      const int64_t b = (p.op_count == 1) ? (u >> gbits) : o;
      const uint64_t gi = uint64_t(u) & ((uint64_t(1) << gbits) - 1);
      const uint64_t base = insert_zero_bits(gi, p.sorted, p.k);
      const T* gb = gout + b * N + base;
      const T* xb = in + b * p.in_bstride + base;
#pragma unroll
      for (int i = 0; i < RB; ++i) gr[q][i] = gb[offr[i]];
#pragma unroll
      for (int j = 0; j < CB; ++j) xr[q][j] = xb[offc[j]];
    }
#pragma unroll
    for (int q = 0; q < UNR; ++q) {
      if (u0 + q * kBlock >= u_end) continue;
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          // g * conj(x)
          const double2 g = as_d2(gr[q][i]), x = as_d2(xr[q][j]);
          acc[i][j].x = fma(g.x, x.x, fma(g.y, x.y, acc[i][j].x));
          acc[i][j].y = fma(g.y, x.x, fma(-g.x, x.y, acc[i][j].y));
        }
    }
  }
  // workgroup reduction: shuffles inside each wavefront, then LDS
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      double vr = acc[i][j].x, vi = acc[i][j].y;
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) {
        vr += __shfl_xor(vr, d, 64);
        vi += __shfl_xor(vi, d, 64);
      }
      if (lane == 0) red[wave][i * CB + j] = make_double2(vr, vi);
    }
  __syncthreads();
  if (threadIdx.x < RB * CB) {
    double2 t = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) {
      t.x += red[w][threadIdx.x].x;
      t.y += red[w][threadIdx.x].y;
    }
    const int64_t e = int64_t(r0 + threadIdx.x / CB) * D + (c0 + threadIdx.x % CB);
    if (p.parts == 1) {
      grad_op[o * D * D + e] = from_d2<T>(t);
    } else {
      double* dst = part_out + ((o * p.parts + part) * D * D + e) * 2;
      dst[0] = t.x;
      dst[1] = t.y;
    }
  }
}

// One workgroup per (operator, entry): the threads split the parts, then
// reduce with wavefront shuffles and LDS.
template <typename T>
__global__ __launch_bounds__(kBlock) void op_grad_finish_kernel(T* __restrict__ grad_op,
                                                                const double* __restrict__ part,
                                                                int64_t parts, int64_t DD) {
  __shared__ double2 red[kBlock / 64];
  const int64_t t = blockIdx.x;  // o * DD + e
  const int64_t o = t / DD, e = t % DD;
  double vr = 0.0, vi = 0.0;
  for (int64_t q = threadIdx.x; q < parts; q += kBlock) {
    const double* src = part + ((o * parts + q) * DD + e) * 2;
    vr += src[0];
    vi += src[1];
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    vr += __shfl_xor(vr, d, 64);
    vi += __shfl_xor(vi, d, 64);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double2(vr, vi);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 sum = red[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) {
      sum.x += red[w].x;
      sum.y += red[w].y;
    }
    grad_op[t] = from_d2<T>(sum);
  }
}

// Launch shape shared by the workspace query and the entry point.
struct OpGradShape {
  int rb, cb;
  int64_t nblk, parts, units, units_per_part;
};

OpGradShape op_grad_shape(int n, int k, int64_t batch, int64_t op_count) {
  OpGradShape sh{};
  const int64_t D = int64_t(1) << k;
  sh.rb = sh.cb = int(D < 4 ? D : 4);
  sh.nblk = (D / sh.rb) * (D / sh.cb);
  sh.units = (batch / op_count) << (n - k);
  // about 4096 workgroups in all, each thread with at least 16 groups
  const int64_t want = (4096 + op_count * sh.nblk - 1) / (op_count * sh.nblk);
  const int64_t most = (sh.units + kBlock * 16 - 1) / (kBlock * 16);
  sh.parts = want < most ? want : most;
  if (sh.parts < 1) sh.parts = 1;
  sh.units_per_part = (sh.units + sh.parts - 1) / sh.parts;
  return sh;
}

template <typename T>
void launch_op_grad(int rb, unsigned grid, T* grad, double* part, const T* g, const T* x,
                    const OpGradParams& p, hipStream_t s) {
  if (rb == 1)
    op_grad_partial_kernel<T, 1, 1, 4><<<grid, kBlock, 0, s>>>(grad, part, g, x, p);
  else if (rb == 2)
    op_grad_partial_kernel<T, 2, 2, 4><<<grid, kBlock, 0, s>>>(grad, part, g, x, p);
  else
    op_grad_partial_kernel<T, 4, 4, 2><<<grid, kBlock, 0, s>>>(grad, part, g, x, p);
}

}  // namespace
}  // namespace ua

extern "C" int64_t ua_operator_gradient_workspace_bytes(int num_qubits, int op_num_qubits,
                                                       int64_t batch, int64_t op_count) {
  const int n = num_qubits, k = op_num_qubits;
  if (n < 1 || n > 62 || k < 0 || k > n || k > UA_MAX_GRAD_QUBITS || batch < 0 ||
      !(op_count == 1 || op_count == batch))
    return -1;
  if (batch == 0) return 0;
  const ua::OpGradShape sh = ua::op_grad_shape(n, k, batch, op_count);
  if (sh.parts == 1) return 0;
  return op_count * sh.parts * (int64_t(1) << (2 * k)) * 2 * int64_t(sizeof(double));
}

extern "C" int ua_operator_gradient(void* grad_op, const void* grad_out, const void* state,
                                    int dtype, int num_qubits, int op_num_qubits,
                                    const int* qubits, int64_t batch, int64_t op_count,
                                    int64_t state_batch_stride, void* workspace,
                                    ua_stream_t stream) {
  using namespace ua;
  const int n = num_qubits, k = op_num_qubits;
  UA_CHECK_ARG(elem_size(dtype) != 0, "bad dtype");
  UA_CHECK_ARG(n >= 1 && n <= 62 && k >= 0 && k <= n && k <= UA_MAX_GRAD_QUBITS,
               "bad qubit counts (n=%d, k=%d)", n, k);
  UA_CHECK_ARG(batch >= 0 && (op_count == 1 || op_count == batch), "op_count must be 1 or batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(grad_op && grad_out && state, "NULL pointer");
  OpGradParams p{};
  p.n = n;
  p.k = k;
  p.batch = batch;
  p.op_count = op_count;
  p.in_bstride = state_batch_stride;
  uint64_t seen = 0;
  for (int m = 0; m < k; ++m) {
    UA_CHECK_ARG(qubits[m] >= 0 && qubits[m] < n, "qubit out of range");
    p.tbit[m] = n - 1 - qubits[m];
    UA_CHECK_ARG(!((seen >> p.tbit[m]) & 1), "duplicate qubit");
    seen |= uint64_t(1) << p.tbit[m];
  }
  for (int b = 0, m = 0; b < n; ++b)
    if ((seen >> b) & 1) p.sorted[m++] = b;
  const OpGradShape sh = op_grad_shape(n, k, batch, op_count);
  p.parts = sh.parts;
  p.units = sh.units;
  p.units_per_part = sh.units_per_part;
  p.nblk_c = int((int64_t(1) << k) / sh.cb);
  const int64_t blocks = op_count * sh.parts * sh.nblk;
  UA_CHECK_ARG(blocks <= 0x7fffffffLL, "grid too large");
  UA_CHECK_ARG(sh.parts == 1 || (workspace && reinterpret_cast<uintptr_t>(workspace) % 8 == 0),
               "workspace (ua_operator_gradient_workspace_bytes) missing or misaligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(workspace);
  const unsigned grid = unsigned(blocks);
  switch (dtype) {
    case UA_FLOAT32: launch_op_grad(sh.rb, grid, (float*)grad_op, part, (const float*)grad_out, (const float*)state, p, s); break;
    case UA_FLOAT64: launch_op_grad(sh.rb, grid, (double*)grad_op, part, (const double*)grad_out, (const double*)state, p, s); break;
    case UA_COMPLEX64: launch_op_grad(sh.rb, grid, (float2*)grad_op, part, (const float2*)grad_out, (const float2*)state, p, s); break;
    case UA_COMPLEX128: launch_op_grad(sh.rb, grid, (double2*)grad_op, part, (const double2*)grad_out, (const double2*)state, p, s); break;
  }
  if (int rc = check_launch("op_grad_partial_kernel")) return rc;
  if (sh.parts == 1) return UA_OK;
  const int64_t DD = int64_t(1) << (2 * k);
  UA_CHECK_ARG(op_count * DD <= 0x7fffffffLL, "grid too large");
  const unsigned fgrid = unsigned(op_count * DD);
  switch (dtype) {
    case UA_FLOAT32: op_grad_finish_kernel<float><<<fgrid, kBlock, 0, s>>>((float*)grad_op, part, sh.parts, DD); break;
    case UA_FLOAT64: op_grad_finish_kernel<double><<<fgrid, kBlock, 0, s>>>((double*)grad_op, part, sh.parts, DD); break;
    case UA_COMPLEX64: op_grad_finish_kernel<float2><<<fgrid, kBlock, 0, s>>>((float2*)grad_op, part, sh.parts, DD); break;
    case UA_COMPLEX128: op_grad_finish_kernel<double2><<<fgrid, kBlock, 0, s>>>((double2*)grad_op, part, sh.parts, DD); break;
  }
  return check_launch("op_grad_finish_kernel");
}
