// Qubit permutations and sign/flip gates (unitair/simulation/operations.py:
// 506-806): permute_qubits / swap / roll_qubits move amplitudes between index
// bits; multi_cz and multi_controlled_z multiply amplitudes by +-1;
// multi_controlled_x exchanges amplitudes that differ in the target bit.
//
// permute: output qubit q is input qubit perm[q] (torch.permute on the qubit
// dims, operations.py:626-654).  A tile of 2^10 amplitudes is built from the
// input bits that feed the output's 5 lowest bits, the input's own 5 lowest
// bits, and further low input bits up to 10; the tile is read with the input
// bits ascending and written with the output bits ascending, so both sides
// move runs of at least 32 consecutive amplitudes whatever the permutation.
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kPermTileMax = 10;
constexpr int kPermThreads = 256;

struct PermArgs {
  int64_t total_tiles;     // batch * 2^(n - t)
  int64_t in_bstride;
  int n, t;                // state bits, tile bits
  int log2_tiles;          // n - t
  int in_tile[kPermTileMax];   // input bits of the tile, ascending
  int out_tile[kPermTileMax];  // output bits of the tile, ascending
  int e_of_f[kPermTileMax];    // tile position (input order) of output tile bit j
  int out_of_in[64];           // output bit fed by each input bit
};

__device__ __forceinline__ uint64_t deposit(uint32_t x, const int* bits, int t) {
  uint64_t r = 0;
#pragma unroll
  for (int j = 0; j < kPermTileMax; ++j)
    if (j < t) r |= uint64_t((x >> j) & 1) << bits[j];
  return r;
}

template <typename T>
__global__ __launch_bounds__(kPermThreads) void permute_kernel(T* out, const T* in,
                                                               PermArgs a) {
  __shared__ T tile[1 << kPermTileMax];
  const int64_t tg = blockIdx.x;
  if (tg >= a.total_tiles) return;
  const int tid = threadIdx.x;
  const int size = 1 << a.t;
  const int64_t b = tg >> a.log2_tiles;
  const uint64_t rest = uint64_t(tg) & ((uint64_t(1) << a.log2_tiles) - 1);
  const uint64_t in_origin = insert_zero_bits(rest, a.in_tile, a.t);
  uint64_t out_origin = 0;
  for (int bit = 0; bit < a.n; ++bit)
    if ((in_origin >> bit) & 1) out_origin |= uint64_t(1) << a.out_of_in[bit];
  const T* src = in + b * a.in_bstride;
  T* dst = out + (b << a.n);

  for (int e = tid; e < size; e += kPermThreads) tile[e] = src[in_origin | deposit(e, a.in_tile, a.t)];
  __syncthreads();
  for (int f = tid; f < size; f += kPermThreads) {
    uint32_t e = 0;
#pragma unroll
    for (int j = 0; j < kPermTileMax; ++j)
      if (j < a.t) e |= ((uint32_t(f) >> j) & 1) << a.e_of_f[j];
    dst[out_origin | deposit(f, a.out_tile, a.t)] = tile[e];
  }
}

// Sign gates: out[i] = s(i) in[i].  With the CZ pairs stored once per lower
// bit a as a partner mask P_a, the number of pairs with both bits set in i is
//   sum_a x_a popcount(i & P_a) = popcount(i & F(i))  (mod 2),
// F(i) = XOR of P_a over the set bits a of i, which is linear in i over
// GF(2): F(i) = T_0[byte 0 of i] ^ T_1[byte 1 of i] ^ ...  The byte tables
// live in LDS, built by each workgroup.  multi_controlled_z adds -1 where
// every bit of all_mask is set (use_all).
constexpr int kSignChunks = 8;   // bytes of a 62-bit index
struct SignArgs {
  int64_t total;           // batch * 2^n
  int64_t in_bstride;
  int n;
  int chunks;              // index bytes that carry CZ rows
  int use_all;
  uint64_t all_mask;
  uint64_t partners[64];   // P_a for index bit a (zero for most bits)
};

template <typename T>
__device__ __forceinline__ T negate(T v);
template <> __device__ __forceinline__ float negate(float v) { return -v; }
template <> __device__ __forceinline__ double negate(double v) { return -v; }
template <> __device__ __forceinline__ float2 negate(float2 v) { return make_float2(-v.x, -v.y); }
template <> __device__ __forceinline__ double2 negate(double2 v) {
  return make_double2(-v.x, -v.y);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void sign_kernel(T* out, const T* in, SignArgs a) {
  __shared__ uint64_t table[kSignChunks][256];
  for (int c = 0; c < a.chunks; ++c) {
    uint64_t f = 0;
    for (int j = 0; j < 8; ++j)
      if ((threadIdx.x >> j) & 1) f ^= a.partners[8 * c + j];
    table[c][threadIdx.x] = f;
  }
  __syncthreads();
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const uint64_t imask = (uint64_t(1) << a.n) - 1;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < a.total; g += stride) {
    const uint64_t i = uint64_t(g) & imask;
    const int64_t b = g >> a.n;
    uint64_t f = 0;
    for (int c = 0; c < a.chunks; ++c) f ^= table[c][(i >> (8 * c)) & 255];
    uint32_t parity = __popcll(i & f);
    if (a.use_all && (i & a.all_mask) == a.all_mask) parity ^= 1;
    const T v = in[b * a.in_bstride + int64_t(i)];
    out[g] = (parity & 1) ? negate(v) : v;
  }
}

// The same over adjacent pairs (i, i | 1), i even: F is linear, so
// F(i | 1) = F(i) ^ P_0 and one table walk serves both entries, which move as
// one 2-element access.  Needs 2-element alignment of the buffers and stride.
template <typename T> struct PairIO;
template <> struct PairIO<float> {
  __device__ static void load(const float* p, float& a, float& b) {
    const float2 w = *reinterpret_cast<const float2*>(p);
    a = w.x; b = w.y;
  }
  __device__ static void store(float* p, float a, float b) {
    *reinterpret_cast<float2*>(p) = make_float2(a, b);
  }
};
template <> struct PairIO<double> {
  __device__ static void load(const double* p, double& a, double& b) {
    const double2 w = *reinterpret_cast<const double2*>(p);
    a = w.x; b = w.y;
  }
  __device__ static void store(double* p, double a, double b) {
    *reinterpret_cast<double2*>(p) = make_double2(a, b);
  }
};
template <> struct PairIO<float2> {
  __device__ static void load(const float2* p, float2& a, float2& b) {
    const float4 w = *reinterpret_cast<const float4*>(p);
    a = make_float2(w.x, w.y); b = make_float2(w.z, w.w);
  }
  __device__ static void store(float2* p, float2 a, float2 b) {
    *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
  }
};
template <> struct PairIO<double2> {
  __device__ static void load(const double2* p, double2& a, double2& b) {
    a = make_double2(p[0].x, p[0].y); b = make_double2(p[1].x, p[1].y);
  }
  __device__ static void store(double2* p, double2 a, double2 b) {
    p[0] = a; p[1] = b;
  }
};

template <typename T>
__global__ __launch_bounds__(kBlock) void sign_pair_kernel(T* out, const T* in, SignArgs a) {
  __shared__ uint64_t table[kSignChunks][256];
  for (int c = 0; c < a.chunks; ++c) {
    uint64_t f = 0;
    for (int j = 0; j < 8; ++j)
      if ((threadIdx.x >> j) & 1) f ^= a.partners[8 * c + j];
    table[c][threadIdx.x] = f;
  }
  __syncthreads();
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const uint64_t imask = (uint64_t(1) << a.n) - 1;
  const uint64_t p0 = a.partners[0];
  for (int64_t h = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; 2 * h < a.total;
       h += stride) {
    const int64_t g = 2 * h;
    const uint64_t i = uint64_t(g) & imask;  // even
    const int64_t b = g >> a.n;
    uint64_t f = 0;
    for (int c = 0; c < a.chunks; ++c) f ^= table[c][(i >> (8 * c)) & 255];
    uint32_t par0 = __popcll(i & f);
    uint32_t par1 = __popcll((i | 1) & (f ^ p0));
    if (a.use_all) {
      par0 ^= (i & a.all_mask) == a.all_mask;
      par1 ^= ((i | 1) & a.all_mask) == a.all_mask;
    }
    T v0, v1;
    PairIO<T>::load(in + b * a.in_bstride + int64_t(i), v0, v1);
    PairIO<T>::store(out + g, (par0 & 1) ? negate(v0) : v0, (par1 & 1) ? negate(v1) : v1);
  }
}

// Controlled bit flip: out[i] = in[i ^ flip] when every bit of ctrl is set in
// i, else in[i] (flip is not in ctrl, so the map is an involution).
template <typename T>
__global__ __launch_bounds__(kBlock) void cflip_kernel(T* out, const T* in, int64_t total,
                                                      int64_t in_bstride, int n, uint64_t ctrl,
                                                      uint64_t flip) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const uint64_t imask = (uint64_t(1) << n) - 1;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += stride) {
    const uint64_t i = uint64_t(g) & imask;
    const int64_t b = g >> n;
    const uint64_t j = ((i & ctrl) == ctrl) ? (i ^ flip) : i;
    out[g] = in[b * in_bstride + int64_t(j)];
  }
}

unsigned capped_grid(int64_t total) {
  const int64_t blocks = (total + kBlock - 1) / kBlock;
  return unsigned(blocks < 65536 ? (blocks < 1 ? 1 : blocks) : 65536);
}

int common_checks(void* out, const void* in, int dtype, int n, int64_t batch,
                  int64_t in_bstride) {
  UA_CHECK_ARG(n >= 1 && n <= 62, "num_qubits=%d out of range", n);
  UA_CHECK_ARG(dtype >= UA_FLOAT32 && dtype <= UA_COMPLEX128, "bad dtype %d", dtype);
  UA_CHECK_ARG(batch >= 0 && in_bstride >= 0, "negative batch or stride");
  UA_CHECK_ARG(batch == 0 || (out && in), "NULL pointer");
  UA_CHECK_ARG(batch == 0 || out != in, "out must not alias the input");
  return UA_OK;
}

}  // namespace
}  // namespace ua

extern "C" int ua_permute_qubits(void* out, const void* state, int dtype, int num_qubits,
                                 const int* permutation, int64_t batch,
                                 int64_t state_batch_stride, ua_stream_t stream) {
  using namespace ua;
  const int n = num_qubits;
  if (int rc = common_checks(out, state, dtype, n, batch, state_batch_stride)) return rc;
  UA_CHECK_ARG(permutation, "NULL permutation");
  // output bit ob = n-1-q takes input bit n-1-perm[q]
  int src_of_out[64], out_of_in[64];
  uint64_t seen = 0;
  for (int q = 0; q < n; ++q) {
    const int p = permutation[q];
    UA_CHECK_ARG(p >= 0 && p < n && !((seen >> p) & 1), "permutation is not a permutation of "
                 "0..%d (entry %d = %d)", n - 1, q, p);
    seen |= uint64_t(1) << p;
    src_of_out[n - 1 - q] = n - 1 - p;
    out_of_in[n - 1 - p] = n - 1 - q;
  }
  if (batch == 0) return UA_OK;
  PermArgs a{};
  a.n = n;
  a.t = n < kPermTileMax ? n : kPermTileMax;
  a.in_bstride = state_batch_stride;
  // input tile bits: sources of output bits 0..4, input bits 0..4, then the
  // lowest remaining input bits
  uint64_t tb = 0;
  const int c = n < 5 ? n : 5;
  for (int ob = 0; ob < c; ++ob) tb |= uint64_t(1) << src_of_out[ob];
  for (int ib = 0; ib < c; ++ib) tb |= uint64_t(1) << ib;
  for (int ib = 0; __builtin_popcountll(tb) < a.t; ++ib) tb |= uint64_t(1) << ib;
  uint64_t ob_mask = 0;
  {
    uint64_t x = tb;
    for (int j = 0; j < a.t; ++j) {
      a.in_tile[j] = __builtin_ctzll(x);
      x &= x - 1;
      ob_mask |= uint64_t(1) << out_of_in[a.in_tile[j]];
    }
    x = ob_mask;
    for (int j = 0; j < a.t; ++j) {
      a.out_tile[j] = __builtin_ctzll(x);
      x &= x - 1;
      const int ib = src_of_out[a.out_tile[j]];
      int pos = 0;
      while (a.in_tile[pos] != ib) ++pos;
      a.e_of_f[j] = pos;
    }
  }
  for (int ib = 0; ib < n; ++ib) a.out_of_in[ib] = out_of_in[ib];
  a.log2_tiles = n - a.t;
  a.total_tiles = batch << a.log2_tiles;
  UA_CHECK_ARG(a.total_tiles <= 0x7fffffffLL, "too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned grid = unsigned(a.total_tiles);
  switch (elem_size(dtype)) {
    case 4:
      permute_kernel<float><<<grid, kPermThreads, 0, s>>>(
          static_cast<float*>(out), static_cast<const float*>(state), a);
      break;
    case 8:
      permute_kernel<float2><<<grid, kPermThreads, 0, s>>>(
          static_cast<float2*>(out), static_cast<const float2*>(state), a);
      break;
    default:
      permute_kernel<double2><<<grid, kPermThreads, 0, s>>>(
          static_cast<double2*>(out), static_cast<const double2*>(state), a);
      break;
  }
  return check_launch("permute_kernel");
}

extern "C" int ua_apply_sign_gates(void* out, const void* state, int dtype, int num_qubits,
                                   int num_pairs, const int* pair_bits, int use_all,
                                   uint64_t all_mask, int64_t batch,
                                   int64_t state_batch_stride, ua_stream_t stream) {
  using namespace ua;
  const int n = num_qubits;
  if (int rc = common_checks(out, state, dtype, n, batch, state_batch_stride)) return rc;
  UA_CHECK_ARG(num_pairs >= 0 && (num_pairs == 0 || pair_bits), "bad pair list");
  UA_CHECK_ARG(!use_all || (all_mask >> n) == 0, "all_mask beyond the state");
  SignArgs a{};
  a.n = n;
  a.in_bstride = state_batch_stride;
  a.use_all = use_all ? 1 : 0;
  a.all_mask = all_mask;
  for (int p = 0; p < num_pairs; ++p) {
    const int x = pair_bits[2 * p], y = pair_bits[2 * p + 1];
    UA_CHECK_ARG(x >= 0 && x < n && y >= 0 && y < n && x != y, "pair %d: bits (%d, %d) invalid",
                 p, x, y);
    const int lo = x < y ? x : y, hi = x < y ? y : x;
    a.partners[lo] ^= uint64_t(1) << hi;  // a repeated pair cancels
  }
  for (int bit = 0; bit < n; ++bit)
    if (a.partners[bit]) a.chunks = bit / 8 + 1;
  if (batch == 0) return UA_OK;
  a.total = batch << n;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // a few elements per thread so that the table set-up is amortised
  const size_t es = elem_size(dtype);
  const bool pairs = n >= 1 && reinterpret_cast<uintptr_t>(out) % (2 * es) == 0 &&
                     reinterpret_cast<uintptr_t>(state) % (2 * es) == 0 &&
                     state_batch_stride % 2 == 0;
  const int64_t units = pairs ? a.total / 2 : a.total;
  const int64_t want = (units + 16 * kBlock - 1) / (16 * kBlock);
  const unsigned grid = unsigned(want < 1 ? 1 : (want > 8192 ? 8192 : want));
#define UA_SIGN_LAUNCH(T)                                                                   \
  do {                                                                                      \
    if (pairs)                                                                              \
      sign_pair_kernel<T><<<grid, kBlock, 0, s>>>(static_cast<T*>(out),                     \
                                                  static_cast<const T*>(state), a);         \
    else                                                                                    \
      sign_kernel<T><<<grid, kBlock, 0, s>>>(static_cast<T*>(out),                          \
                                             static_cast<const T*>(state), a);              \
  } while (0)
  switch (dtype) {
    case UA_FLOAT32: UA_SIGN_LAUNCH(float); break;
    case UA_FLOAT64: UA_SIGN_LAUNCH(double); break;
    case UA_COMPLEX64: UA_SIGN_LAUNCH(float2); break;
    default: UA_SIGN_LAUNCH(double2); break;
  }
#undef UA_SIGN_LAUNCH
  return check_launch("sign_kernel");
}

extern "C" int ua_apply_controlled_flip(void* out, const void* state, int dtype,
                                        int num_qubits, uint64_t control_mask, int target_bit,
                                        int64_t batch, int64_t state_batch_stride,
                                        ua_stream_t stream) {
  using namespace ua;
  const int n = num_qubits;
  if (int rc = common_checks(out, state, dtype, n, batch, state_batch_stride)) return rc;
  UA_CHECK_ARG(target_bit >= 0 && target_bit < n, "target bit %d out of range", target_bit);
  UA_CHECK_ARG((control_mask >> n) == 0, "control mask beyond the state");
  UA_CHECK_ARG(!((control_mask >> target_bit) & 1), "target is also a control");
  if (batch == 0) return UA_OK;
  const int64_t total = batch << n;
  const uint64_t flip = uint64_t(1) << target_bit;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned grid = capped_grid(total);
  switch (elem_size(dtype)) {
    case 4:
      cflip_kernel<float><<<grid, kBlock, 0, s>>>(static_cast<float*>(out),
                                                  static_cast<const float*>(state), total,
                                                  state_batch_stride, n, control_mask, flip);
      break;
    case 8:
      cflip_kernel<float2><<<grid, kBlock, 0, s>>>(static_cast<float2*>(out),
                                                   static_cast<const float2*>(state), total,
                                                   state_batch_stride, n, control_mask, flip);
      break;
    default:
      cflip_kernel<double2><<<grid, kBlock, 0, s>>>(static_cast<double2*>(out),
                                                    static_cast<const double2*>(state), total,
                                                    state_batch_stride, n, control_mask, flip);
      break;
  }
  return check_launch("cflip_kernel");
}
