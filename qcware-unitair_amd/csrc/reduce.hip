// Row reductions with float64 accumulation (unitair/states/innerprod.py).
//
// Two deterministic passes (no atomics): pass 1 splits every row into
// chunks, one workgroup per chunk, and writes one float64 (re, im) partial per
// chunk into the caller's workspace; pass 2 sums each row's partials in a
// fixed order and converts to the requested output dtype.
#include "ua_common.h"

#include <mutex>

namespace ua {
namespace {

constexpr int kChunk = kBlock * 16;  // elements per pass-1 workgroup

enum Kind { NORM = 0, INNER = 1, DIAG = 2 };

struct RedParams {
  int64_t batch, length, chunks;
  int64_t a_stride, b_stride;  // row strides of the two inputs (elements)
};

__device__ __forceinline__ double2 as_d2(float x) { return make_double2(x, 0.0); }
__device__ __forceinline__ double2 as_d2(double x) { return make_double2(x, 0.0); }
__device__ __forceinline__ double2 as_d2(float2 x) { return make_double2(x.x, x.y); }
__device__ __forceinline__ double2 as_d2(double2 x) { return x; }

__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  return v;
}

// Sum over a workgroup; result valid in thread 0.
__device__ __forceinline__ double2 block_sum(double2 v) {
  __shared__ double2 part[kBlock / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = v;
  __syncthreads();
  double2 r = make_double2(0.0, 0.0);
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) {
      r.x += part[w].x;
      r.y += part[w].y;
    }
  return r;
}

// Contribution of one element.  NORM: |a|^2.  INNER: conj(a) * b.
// DIAG: |b|^2 * a  (a = diagonal value, b = state element).
template <int KIND, typename TA, typename TB>
__device__ __forceinline__ double2 term(TA a, TB b) {
  const double2 x = as_d2(a);
  if constexpr (KIND == NORM) {
    return make_double2(x.x * x.x + x.y * x.y, 0.0);
  } else if constexpr (KIND == INNER) {
    const double2 y = as_d2(b);
    return make_double2(x.x * y.x + x.y * y.y, x.x * y.y - x.y * y.x);
  } else {
    const double2 y = as_d2(b);
    const double m = y.x * y.x + y.y * y.y;
    return make_double2(m * x.x, m * x.y);
  }
}

template <int KIND, typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void reduce_pass1(
    double2* __restrict__ partial, const TA* __restrict__ a, const TB* __restrict__ b,
    RedParams p) {
  const int64_t row = blockIdx.x / p.chunks;
  const int64_t chunk = blockIdx.x % p.chunks;
  const TA* ar = a + row * p.a_stride;
  const TB* br = b + row * p.b_stride;
  const int64_t lo = chunk * kChunk;
  const int64_t hi = (lo + kChunk < p.length) ? lo + kChunk : p.length;
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double2 t = term<KIND>(ar[i], KIND == NORM ? TB{} : br[i]);
    acc.x += t.x;
    acc.y += t.y;
  }
  const double2 r = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <typename TO> __device__ __forceinline__ void store_out(TO* o, double2 v);
template <> __device__ __forceinline__ void store_out(float* o, double2 v) { *o = float(v.x); }
template <> __device__ __forceinline__ void store_out(double* o, double2 v) { *o = v.x; }
template <> __device__ __forceinline__ void store_out(float2* o, double2 v) {
  *o = make_float2(float(v.x), float(v.y));
}
template <> __device__ __forceinline__ void store_out(double2* o, double2 v) { *o = v; }

template <typename TO>
__global__ __launch_bounds__(kBlock) void reduce_pass2(
    TO* __restrict__ out, const double2* __restrict__ partial, int64_t chunks) {
  const double2* pr = partial + int64_t(blockIdx.x) * chunks;
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t i = threadIdx.x; i < chunks; i += kBlock) {
    acc.x += pr[i].x;
    acc.y += pr[i].y;
  }
  const double2 r = block_sum(acc);
  if (threadIdx.x == 0) store_out(out + blockIdx.x, r);
}

int64_t chunks_for(int64_t length) { return length <= 0 ? 1 : (length + kChunk - 1) / kChunk; }

template <int KIND, typename TA, typename TB>
int run(void* out, int odt, const void* a, const void* b, const RedParams& p,
        void* ws, hipStream_t s) {
  UA_CHECK_ARG(p.batch * p.chunks <= 0x7fffffffLL, "reduction grid too large");
  double2* partial = static_cast<double2*>(ws);
  reduce_pass1<KIND, TA, TB><<<unsigned(p.batch * p.chunks), kBlock, 0, s>>>(
      partial, static_cast<const TA*>(a), static_cast<const TB*>(b), p);
  int rc = check_launch("reduce_pass1");
  if (rc) return rc;
  UA_CHECK_ARG(p.batch <= 0x7fffffffLL, "reduction batch too large");
  const unsigned g = unsigned(p.batch);
  switch (odt) {
    case UA_FLOAT32: reduce_pass2<float><<<g, kBlock, 0, s>>>((float*)out, partial, p.chunks); break;
    case UA_FLOAT64: reduce_pass2<double><<<g, kBlock, 0, s>>>((double*)out, partial, p.chunks); break;
    case UA_COMPLEX64: reduce_pass2<float2><<<g, kBlock, 0, s>>>((float2*)out, partial, p.chunks); break;
    case UA_COMPLEX128: reduce_pass2<double2><<<g, kBlock, 0, s>>>((double2*)out, partial, p.chunks); break;
    default: set_error("bad output dtype %d", odt); return UA_ERR_INVALID_ARGUMENT;
  }
  return check_launch("reduce_pass2");
}

template <int KIND, typename TA>
int run_b(int bdt, void* out, int odt, const void* a, const void* b, const RedParams& p,
          void* ws, hipStream_t s) {
  switch (bdt) {
    case UA_FLOAT32: return run<KIND, TA, float>(out, odt, a, b, p, ws, s);
    case UA_FLOAT64: return run<KIND, TA, double>(out, odt, a, b, p, ws, s);
    case UA_COMPLEX64: return run<KIND, TA, float2>(out, odt, a, b, p, ws, s);
    case UA_COMPLEX128: return run<KIND, TA, double2>(out, odt, a, b, p, ws, s);
  }
  set_error("bad dtype %d", bdt);
  return UA_ERR_INVALID_ARGUMENT;
}

template <int KIND>
int run_ab(int adt, int bdt, void* out, int odt, const void* a, const void* b,
           const RedParams& p, void* ws, hipStream_t s) {
  switch (adt) {
    case UA_FLOAT32: return run_b<KIND, float>(bdt, out, odt, a, b, p, ws, s);
    case UA_FLOAT64: return run_b<KIND, double>(bdt, out, odt, a, b, p, ws, s);
    case UA_COMPLEX64: return run_b<KIND, float2>(bdt, out, odt, a, b, p, ws, s);
    case UA_COMPLEX128: return run_b<KIND, double2>(bdt, out, odt, a, b, p, ws, s);
  }
  set_error("bad dtype %d", adt);
  return UA_ERR_INVALID_ARGUMENT;
}

int check_common(void* out, int odt, int64_t batch, int64_t length, void* ws) {
  UA_CHECK_ARG(elem_size(odt) != 0, "bad output dtype %d", odt);
  UA_CHECK_ARG(batch >= 0 && length >= 0, "negative size");
  if (batch == 0) return UA_OK;
  UA_CHECK_ARG(out && ws, "NULL output or workspace pointer");
  return UA_OK;
}

}  // namespace

int64_t reduce_workspace_bytes(int64_t batch, int64_t length) {
  return (batch < 1 ? 1 : batch) * chunks_for(length) * int64_t(sizeof(double2));
}

int norm_squared_impl(void* out, int odt, const void* st, int dt, int64_t batch,
                      int64_t length, int64_t stride, void* ws, hipStream_t s) {
  int rc = check_common(out, odt, batch, length, ws);
  if (rc || batch == 0) return rc;
  RedParams p{batch, length, chunks_for(length), stride, 0};
  return run_ab<NORM>(dt, dt, out, odt, st, st, p, ws, s);
}

}  // namespace ua

// ---------------------------------------------------------------------------
// ua_norm_squared_noalloc: ua_norm_squared with its float64 partials in a
// static device array instead of a caller workspace, for callers that cannot
// allocate (code under graph capture, tight serving loops).  Every GPU holds
// its own copy of the array for the life of the process; calls on one GPU
// share it, so calls that overlap on one GPU race.  Nothing in it needs to
// start at zero: pass 1 writes every partial that pass 2 reads.
namespace {
constexpr int64_t kScratchPartials = int64_t(1) << 16;  // 1 MiB of (re, im) float64 pairs
constexpr int kScratchDevices = 64;
}  // namespace

static __device__ __align__(256) double2 scratch[kScratchPartials];

namespace {
std::mutex g_scratch_mu;
double2* g_scratch_ptr[kScratchDevices] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current (the name itself is a host-side stand-in).
int scratch_of_current_device(double2** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    ua::set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    return UA_ERR_HIP;
  }
  UA_CHECK_ARG(dev >= 0 && dev < kScratchDevices, "device %d beyond the %d supported", dev,
               kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      ua::set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return UA_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double2*>(p);
  }
  *out = g_scratch_ptr[dev];
  return UA_OK;
}
}  // namespace

extern "C" int ua_norm_squared_noalloc(void* out, int out_dtype, const void* state, int dtype,
                                       int64_t batch, int64_t length, int64_t row_stride,
                                       ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(elem_size(out_dtype) != 0, "bad output dtype %d", out_dtype);
  UA_CHECK_ARG(elem_size(dtype) != 0, "bad dtype %d", dtype);
  UA_CHECK_ARG(batch >= 0 && length >= 0, "negative size");
  if (batch == 0) return UA_OK;
  const int64_t chunks = chunks_for(length);
  UA_CHECK_ARG(chunks <= kScratchPartials && batch <= kScratchPartials / chunks,
               "batch %lld x %lld chunks of %d elements exceeds the static scratch of %lld "
               "partials (at most 2^28 elements per call)",
               (long long)batch, (long long)chunks, kChunk, (long long)kScratchPartials);
  UA_CHECK_ARG(out && state, "NULL output or state pointer");
  double2* ws = nullptr;
  if (int rc = scratch_of_current_device(&ws)) return rc;
  return norm_squared_impl(out, out_dtype, state, dtype, batch, length, row_stride, ws,
                           static_cast<hipStream_t>(stream));
}

extern "C" int64_t ua_reduce_workspace_bytes(int64_t batch, int64_t length) {
  return ua::reduce_workspace_bytes(batch, length);
}

extern "C" int ua_norm_squared(void* out, int out_dtype, const void* state, int dtype,
                               int64_t batch, int64_t length, int64_t row_stride,
                               void* workspace, ua_stream_t stream) {
  return ua::norm_squared_impl(out, out_dtype, state, dtype, batch, length, row_stride,
                               workspace, static_cast<hipStream_t>(stream));
}

extern "C" int ua_inner_product(void* out, int out_dtype, const void* a, int64_t a_row_stride,
                                const void* b, int64_t b_row_stride, int dtype,
                                int64_t batch, int64_t length, void* workspace,
                                ua_stream_t stream) {
  using namespace ua;
  int rc = check_common(out, out_dtype, batch, length, workspace);
  if (rc || batch == 0) return rc;
  UA_CHECK_ARG(a && b, "NULL input pointer");
  RedParams p{batch, length, chunks_for(length), a_row_stride, b_row_stride};
  return run_ab<INNER>(dtype, dtype, out, out_dtype, a, b, p, workspace,
                       static_cast<hipStream_t>(stream));
}

extern "C" int ua_diag_expectation(void* out, int out_dtype, const void* diag, int diag_dtype,
                                   int64_t diag_row_stride, const void* state, int dtype,
                                   int64_t state_row_stride, int64_t batch, int64_t length,
                                   void* workspace, ua_stream_t stream) {
  using namespace ua;
  int rc = check_common(out, out_dtype, batch, length, workspace);
  if (rc || batch == 0) return rc;
  UA_CHECK_ARG(diag && state, "NULL input pointer");
  RedParams p{batch, length, chunks_for(length), diag_row_stride, state_row_stride};
  return run_ab<DIAG>(diag_dtype, dtype, out, out_dtype, diag, state, p, workspace,
                      static_cast<hipStream_t>(stream));
}

namespace ua {
namespace {
template <typename T, typename R>
__global__ __launch_bounds__(kBlock) void abs_squared_kernel(R* __restrict__ out,
                                                             const T* __restrict__ x,
                                                             int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    if constexpr (Num<T>::complex) {
      const T v = x[i];
      out[i] = v.x * v.x + v.y * v.y;
    } else {
      out[i] = x[i] * x[i];
    }
  }
}
}  // namespace
}  // namespace ua

extern "C" int ua_abs_squared(void* out, const void* state, int dtype, int64_t numel,
                              ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(numel >= 0, "negative size");
  if (numel == 0) return UA_OK;
  UA_CHECK_ARG(out && state, "NULL data pointer");
  int64_t blocks = (numel + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = unsigned(blocks);
  switch (dtype) {
    case UA_FLOAT32: abs_squared_kernel<float, float><<<g, kBlock, 0, s>>>((float*)out, (const float*)state, numel); break;
    case UA_FLOAT64: abs_squared_kernel<double, double><<<g, kBlock, 0, s>>>((double*)out, (const double*)state, numel); break;
    case UA_COMPLEX64: abs_squared_kernel<float2, float><<<g, kBlock, 0, s>>>((float*)out, (const float2*)state, numel); break;
    case UA_COMPLEX128: abs_squared_kernel<double2, double><<<g, kBlock, 0, s>>>((double*)out, (const double2*)state, numel); break;
    default: set_error("bad dtype %d", dtype); return UA_ERR_INVALID_ARGUMENT;
  }
  return check_launch("abs_squared_kernel");
}
