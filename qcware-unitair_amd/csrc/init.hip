// State constructors on the device (unitair/initializations.py:46-128).
//
// ua_rand_state: Gaussian entries normalised per state, as rand_state
// (initializations.py:71-100: torch.randn, then divide by the norm).  The
// normals come from Philox4x32-10 (Salmon et al., SC'11) with counter
// (element pair, 0, 0, 0) and key (seed), through the Box-Muller transform
// (Box and Muller, 1958), in float64 for float64/complex128 states and in
// float32 (to a few ulp of the float64 transform) for float32/complex64.
// The generation pass writes the normals unscaled and sums |x|^2 in float64
// per block; after the per-state norms, a streaming pass scales in place.
// (Generating is compute-bound, so one generation plus a read-modify-write
// pass beats generating twice.)
// ua_fill_state: uniform_superposition (2^(-n/2) everywhere, :103-128) and
// unit_vector (one 1, :46-68).
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kInitThreads = 256;

__device__ __forceinline__ void philox(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(M0) * c[0];
    const uint64_t p1 = uint64_t(M1) * c[2];
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
    c[1] = uint32_t(p1);
    c[3] = uint32_t(p0);
    c[0] = n0;
    c[2] = n2;
    k0 += W0;
    k1 += W1;
  }
}

// Four standard normals for counter `ctr`, in the arithmetic of R (float64
// for float64/complex128 states, float32 for float32/complex64 ones).
template <typename R>
__device__ __forceinline__ void normals4(uint64_t ctr, uint64_t seed, R (&z)[4]);

template <>
__device__ __forceinline__ void normals4<double>(uint64_t ctr, uint64_t seed, double (&z)[4]) {
  uint32_t c[4] = {uint32_t(ctr), uint32_t(ctr >> 32), 0u, 0u};
  philox(c, uint32_t(seed), uint32_t(seed >> 32));
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const double u1 = (double(c[2 * h]) + 1.0) * 0x1.0p-32;  // (0, 1]
    const double u2 = double(c[2 * h + 1]) * 0x1.0p-32;      // [0, 1)
    const double r = sqrt(-2.0 * log(u1));
    double s, co;
    sincospi(2.0 * u2, &s, &co);
    z[2 * h] = r * co;
    z[2 * h + 1] = r * s;
  }
}

// The same transform in float32.  u1 itself is not representable to 32 bits
// in float32, so ln(u1) is taken from whichever of u1 and 1 - u1 is exact to
// float32 rounding: ln(u1) for u1 <= 1/2, log1p(-(1 - u1)) above (1 - u1 =
// (2^32 - 1 - x) 2^-32 is an exact integer times 2^-32).  Every normal then
// carries a float32 relative error of a few ulp of the float64 transform.
template <>
__device__ __forceinline__ void normals4<float>(uint64_t ctr, uint64_t seed, float (&z)[4]) {
  uint32_t c[4] = {uint32_t(ctr), uint32_t(ctr >> 32), 0u, 0u};
  philox(c, uint32_t(seed), uint32_t(seed >> 32));
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t x = c[2 * h];
    const float l = (x < 0x80000000u)
                        ? logf((float(x) + 1.0f) * 0x1.0p-32f)
                        : log1pf(-float(0xFFFFFFFFu - x) * 0x1.0p-32f);
    const float r = sqrtf(-2.0f * l);
    // angle 2 pi u2: the quadrant from the top two bits of y (exact), the
    // rest (< pi/2) from the low 30 bits
    const uint32_t y = c[2 * h + 1];
    float s0, c0;
    sincospif(float(y & 0x3FFFFFFFu) * 0x1.0p-31f, &s0, &c0);
    const uint32_t quad = y >> 30;
    const float sq0 = (quad & 1) ? c0 : s0, cq0 = (quad & 1) ? s0 : c0;
    const float s = (quad >= 2) ? -sq0 : sq0;                 // sin(a + quad pi/2)
    const float co = (quad == 1 || quad == 2) ? -cq0 : cq0;   // cos(a + quad pi/2)
    z[2 * h] = r * co;
    z[2 * h + 1] = r * s;
  }
}

// Element e of the flat (batch * 2^n) array gets normals from counter
// e / per (per = 4 elements per counter for real states, 2 for complex).
template <typename T> struct Gen;
template <> struct Gen<float> {
  using R = float;
  static constexpr int per = 4;
  __device__ static float make(const float* z, int j, double scale) { return float(z[j] * scale); }
  __device__ static double sq(const float* z, int j) { return double(z[j]) * z[j]; }
};
template <> struct Gen<double> {
  using R = double;
  static constexpr int per = 4;
  __device__ static double make(const double* z, int j, double scale) { return z[j] * scale; }
  __device__ static double sq(const double* z, int j) { return z[j] * z[j]; }
};
template <> struct Gen<float2> {
  using R = float;
  static constexpr int per = 2;
  __device__ static float2 make(const float* z, int j, double scale) {
    return make_float2(float(z[2 * j] * scale), float(z[2 * j + 1] * scale));
  }
  __device__ static double sq(const float* z, int j) {
    return double(z[2 * j]) * z[2 * j] + double(z[2 * j + 1]) * z[2 * j + 1];
  }
};
template <> struct Gen<double2> {
  using R = double;
  static constexpr int per = 2;
  __device__ static double2 make(const double* z, int j, double scale) {
    return make_double2(z[2 * j] * scale, z[2 * j + 1] * scale);
  }
  __device__ static double sq(const double* z, int j) {
    return z[2 * j] * z[2 * j] + z[2 * j + 1] * z[2 * j + 1];
  }
};

// Pass 1: block k generates counters [k * G, (k + 1) * G) (G counters per
// block; blocks never straddle two states), writes the normals unscaled and
// sums their |x|^2 in float64.
template <typename T>
__global__ __launch_bounds__(kInitThreads) void rand_gen_kernel(T* out, int64_t G, uint64_t seed,
                                                                double* partial) {
  __shared__ double red[kInitThreads / 64];
  double acc = 0;
  const int64_t c0 = int64_t(blockIdx.x) * G;
  for (int64_t c = c0 + threadIdx.x; c < c0 + G; c += kInitThreads) {
    typename Gen<T>::R z[4];
    normals4(uint64_t(c), seed, z);
#pragma unroll
    for (int j = 0; j < Gen<T>::per; ++j) {
      acc += Gen<T>::sq(z, j);
      out[c * Gen<T>::per + j] = Gen<T>::make(z, j, 1.0);
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kInitThreads / 64; ++w) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// One workgroup per state: 1/sqrt(sum of its blocks' partials).
__global__ __launch_bounds__(kInitThreads) void rand_scale_kernel(const double* partial,
                                                                  int64_t blocks_per_state,
                                                                  double* inv_norm) {
  __shared__ double red[kInitThreads / 64];
  double acc = 0;
  const double* p = partial + int64_t(blockIdx.x) * blocks_per_state;
  for (int64_t i = threadIdx.x; i < blocks_per_state; i += kInitThreads) acc += p[i];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kInitThreads / 64; ++w) t += red[w];
    inv_norm[blockIdx.x] = 1.0 / sqrt(t);
  }
}

// Pass 2: scale in place, two elements per thread (states hold >= 4).  The
// product is formed in float64 and rounded once, as if the normals had been
// scaled before they were first stored.
template <typename T> __device__ __forceinline__ T scaled(T v, double f);
template <> __device__ __forceinline__ float scaled(float v, double f) { return float(v * f); }
template <> __device__ __forceinline__ double scaled(double v, double f) { return v * f; }
template <> __device__ __forceinline__ float2 scaled(float2 v, double f) {
  return make_float2(float(v.x * f), float(v.y * f));
}
template <> __device__ __forceinline__ double2 scaled(double2 v, double f) {
  return make_double2(v.x * f, v.y * f);
}

template <typename T>
__global__ __launch_bounds__(kInitThreads) void rand_apply_kernel(T* out, int64_t pairs, int n,
                                                                  const double* inv_norm) {
  const int64_t stride = int64_t(gridDim.x) * kInitThreads;
  for (int64_t q = int64_t(blockIdx.x) * kInitThreads + threadIdx.x; q < pairs; q += stride) {
    const double f = inv_norm[(q * 2) >> n];
    Pair<T> v = reinterpret_cast<Pair<T>*>(out)[q];
    v.v[0] = scaled(v.v[0], f);
    v.v[1] = scaled(v.v[1], f);
    reinterpret_cast<Pair<T>*>(out)[q] = v;
  }
}

template <typename T> __device__ __forceinline__ T real_value(double v);
template <> __device__ __forceinline__ float real_value(double v) { return float(v); }
template <> __device__ __forceinline__ double real_value(double v) { return v; }
template <> __device__ __forceinline__ float2 real_value(double v) {
  return make_float2(float(v), 0.f);
}
template <> __device__ __forceinline__ double2 real_value(double v) { return make_double2(v, 0.0); }

template <typename T>
__global__ __launch_bounds__(kBlock) void fill_kernel(T* out, int64_t total, int n, int kind,
                                                     int64_t index, double value) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t mask = (int64_t(1) << n) - 1;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += stride) {
    const double v = kind == 0 ? value : ((g & mask) == index ? 1.0 : 0.0);
    out[g] = real_value<T>(v);
  }
}

int64_t counters_per_block(int dtype, int n) {
  const int per = (dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128) ? 2 : 4;
  const int64_t ctr_per_state = (int64_t(1) << n) / per;
  return ctr_per_state < 2048 ? ctr_per_state : 2048;
}

}  // namespace
}  // namespace ua

extern "C" int64_t ua_rand_state_workspace_bytes(int num_qubits, int64_t batch, int dtype) {
  if (num_qubits < 2 || num_qubits > 40 || batch < 0) return -1;
  const int per = (dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128) ? 2 : 4;
  const int64_t ctr_per_state = (int64_t(1) << num_qubits) / per;
  const int64_t G = ua::counters_per_block(dtype, num_qubits);
  return 8 * (batch * (ctr_per_state / G) + batch);
}

extern "C" int ua_rand_state(void* out, int dtype, int num_qubits, int64_t batch, uint64_t seed,
                             void* workspace, ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 2 && num_qubits <= 40, "rand_state on the device needs 2..40 "
               "qubits (got %d)", num_qubits);
  UA_CHECK_ARG(dtype >= UA_FLOAT32 && dtype <= UA_COMPLEX128, "bad dtype %d", dtype);
  UA_CHECK_ARG(batch >= 0, "negative batch");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && workspace, "NULL pointer");
  const int per = (dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128) ? 2 : 4;
  const int64_t ctr_per_state = (int64_t(1) << num_qubits) / per;
  const int64_t G = counters_per_block(dtype, num_qubits);
  const int64_t blocks_per_state = ctr_per_state / G;
  const int64_t blocks = batch * blocks_per_state;
  UA_CHECK_ARG(blocks <= 0x7fffffffLL && batch <= 0x7fffffffLL, "state too large");
  double* partial = static_cast<double*>(workspace);
  double* inv_norm = partial + blocks;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t pairs = (batch << num_qubits) / 2;
  const int64_t ab = (pairs + kInitThreads - 1) / kInitThreads;
  const unsigned agrid = unsigned(ab < 65536 ? ab : 65536);
#define UA_RAND(T)                                                                     \
  rand_gen_kernel<T><<<unsigned(blocks), kInitThreads, 0, s>>>(static_cast<T*>(out), G, \
                                                               seed, partial);         \
  if (int rc = check_launch("rand_gen_kernel")) return rc;                             \
  rand_scale_kernel<<<unsigned(batch), kInitThreads, 0, s>>>(partial, blocks_per_state, \
                                                             inv_norm);                \
  if (int rc = check_launch("rand_scale_kernel")) return rc;                           \
  rand_apply_kernel<T><<<agrid, kInitThreads, 0, s>>>(static_cast<T*>(out), pairs,      \
                                                      num_qubits, inv_norm);
  switch (dtype) {
    case UA_FLOAT32: UA_RAND(float); break;
    case UA_FLOAT64: UA_RAND(double); break;
    case UA_COMPLEX64: UA_RAND(float2); break;
    default: UA_RAND(double2); break;
  }
#undef UA_RAND
  return check_launch("rand_apply_kernel");
}

extern "C" int ua_fill_state(void* out, int dtype, int num_qubits, int64_t batch, int kind,
                             int64_t index, ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 0 && num_qubits <= 62, "num_qubits=%d out of range", num_qubits);
  UA_CHECK_ARG(dtype >= UA_FLOAT32 && dtype <= UA_COMPLEX128, "bad dtype %d", dtype);
  UA_CHECK_ARG(kind == 0 || kind == 1, "kind must be 0 (uniform) or 1 (basis vector)");
  UA_CHECK_ARG(batch >= 0, "negative batch");
  UA_CHECK_ARG(kind == 0 || (index >= 0 && index < (int64_t(1) << num_qubits)),
               "index %lld out of range", (long long)index);
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out, "NULL pointer");
  const int64_t total = batch << num_qubits;
  const double value = exp2(-0.5 * num_qubits);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t b = (total + kBlock - 1) / kBlock;
  const unsigned grid = unsigned(b < 65536 ? b : 65536);
  switch (dtype) {
    case UA_FLOAT32:
      fill_kernel<float><<<grid, kBlock, 0, s>>>(static_cast<float*>(out), total, num_qubits,
                                                 kind, index, value);
      break;
    case UA_FLOAT64:
      fill_kernel<double><<<grid, kBlock, 0, s>>>(static_cast<double*>(out), total, num_qubits,
                                                  kind, index, value);
      break;
    case UA_COMPLEX64:
      fill_kernel<float2><<<grid, kBlock, 0, s>>>(static_cast<float2*>(out), total, num_qubits,
                                                  kind, index, value);
      break;
    default:
      fill_kernel<double2><<<grid, kBlock, 0, s>>>(static_cast<double2*>(out), total,
                                                   num_qubits, kind, index, value);
      break;
  }
  return check_launch("fill_kernel");
}
