// Diagonal phase: ua_apply_phase, out = exp(-1j * angles) * state with torch
// broadcasting (unitair/simulation/operations.py:15-42).
//
// The reference runs exp() into a full-size complex temporary and then a
// broadcast multiply (two kernels, three state-sized passes).  Here one pass
// reads each amplitude and each angle once, forms cos/sin in registers and
// writes the product.  `out` may equal `state` (in place) when both are
// contiguous with the output's shape.
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kMaxRank = 8;

enum PhaseMode { CONTIG = 0, TILE = 1, GENERAL = 2 };

struct PhaseParams {
  int64_t total;
  int64_t tile;  // TILE: angles repeat every `tile` elements
  int rank;
  int64_t shape[kMaxRank];
  int64_t sstride[kMaxRank];
  int64_t astride[kMaxRank];
};

// exp(-1j * a) in the precision of the angle type.
__device__ __forceinline__ float2 phase_of(float a) {
  float s, c;
  sincosf(a, &s, &c);
  return make_float2(c, -s);
}
__device__ __forceinline__ double2 phase_of(double a) {
  double s, c;
  sincos(a, &s, &c);
  return make_double2(c, -s);
}
// Complex angle a = x + iy: exp(-1j*a) = exp(y) * (cos x - i sin x).
__device__ __forceinline__ float2 phase_of(float2 a) {
  float s, c;
  sincosf(a.x, &s, &c);
  const float m = expf(a.y);
  return make_float2(m * c, -(m * s));
}
__device__ __forceinline__ double2 phase_of(double2 a) {
  double s, c;
  sincos(a.x, &s, &c);
  const double m = exp(a.y);
  return make_double2(m * c, -(m * s));
}

template <typename TO> struct Widen;
template <> struct Widen<float2> {
  __device__ static float2 from(float x) { return make_float2(x, 0.f); }
  __device__ static float2 from(float2 x) { return x; }
};
template <> struct Widen<double2> {
  __device__ static double2 from(float x) { return make_double2(x, 0.0); }
  __device__ static double2 from(double x) { return make_double2(x, 0.0); }
  __device__ static double2 from(float2 x) { return make_double2(x.x, x.y); }
  __device__ static double2 from(double2 x) { return x; }
};

template <typename TO>
__device__ __forceinline__ TO cmul(TO a, TO b) {
  TO r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

template <typename TO, typename TS, typename TA, int MODE>
__global__ __launch_bounds__(kBlock) void phase_kernel(
    TO* out, const TS* state, const TA* __restrict__ angles, PhaseParams p) {
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < p.total; i += stride) {
    int64_t si, ai;
    if constexpr (MODE == CONTIG) {
      si = i;
      ai = i;
    } else if constexpr (MODE == TILE) {
      si = i;
      ai = i % p.tile;
    } else {
      int64_t rem = i;
      si = 0;
      ai = 0;
      for (int d = p.rank - 1; d >= 0; --d) {
        const int64_t c = rem % p.shape[d];
        rem /= p.shape[d];
        si += c * p.sstride[d];
        ai += c * p.astride[d];
      }
    }
    const TO ph = Widen<TO>::from(phase_of(angles[ai]));
    const TO x = Widen<TO>::from(state[si]);
    out[i] = cmul(ph, x);
  }
}

// CONTIG layout, two neighbouring elements per access and UNR pairs per
// thread: a full wavefront moves contiguous 16 B (complex64) pieces of state
// and 8 B pieces of angles.
template <typename TO, typename TS, typename TA, int UNR>
__global__ __launch_bounds__(kBlock) void phase_pair_kernel(
    TO* out, const TS* state, const TA* __restrict__ angles, int64_t pairs) {
  const int64_t first = int64_t(blockIdx.x) * (kBlock * UNR) + threadIdx.x;
  Pair<TS> x[UNR];
  Pair<TA> a[UNR];
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    const int64_t i = first + int64_t(u) * kBlock;
    if (i < pairs) {
      x[u] = reinterpret_cast<const Pair<TS>*>(state)[i];
      a[u] = reinterpret_cast<const Pair<TA>*>(angles)[i];
    }
  }
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    const int64_t i = first + int64_t(u) * kBlock;
    if (i >= pairs) continue;
    Pair<TO> y;
#pragma unroll
    for (int h = 0; h < 2; ++h)
      y.v[h] = cmul(Widen<TO>::from(phase_of(a[u].v[h])), Widen<TO>::from(x[u].v[h]));
    reinterpret_cast<Pair<TO>*>(out)[i] = y;
  }
}

template <typename TO, typename TS, typename TA>
void launch_phase(int mode, void* out, const void* st, const void* an,
                  const PhaseParams& p, hipStream_t s) {
  int64_t blocks = (p.total + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  const unsigned g = static_cast<unsigned>(blocks);
  TO* o = static_cast<TO*>(out);
  const TS* x = static_cast<const TS*>(st);
  const TA* a = static_cast<const TA*>(an);
  const bool pair_ok = (p.total % 2 == 0) &&
                       reinterpret_cast<uintptr_t>(out) % (2 * sizeof(TO)) == 0 &&
                       reinterpret_cast<uintptr_t>(st) % (2 * sizeof(TS)) == 0 &&
                       reinterpret_cast<uintptr_t>(an) % (2 * sizeof(TA)) == 0;
  if (mode == CONTIG && pair_ok) {
    const int64_t pairs = p.total / 2;
    phase_pair_kernel<TO, TS, TA, 2><<<grid_for(pairs, kBlock * 2), kBlock, 0, s>>>(o, x, a, pairs);
    return;
  }
  switch (mode) {
    case CONTIG: phase_kernel<TO, TS, TA, CONTIG><<<g, kBlock, 0, s>>>(o, x, a, p); break;
    case TILE: phase_kernel<TO, TS, TA, TILE><<<g, kBlock, 0, s>>>(o, x, a, p); break;
    default: phase_kernel<TO, TS, TA, GENERAL><<<g, kBlock, 0, s>>>(o, x, a, p); break;
  }
}

template <typename TO, typename TS>
int dispatch_angle(int adt, int mode, void* out, const void* st, const void* an,
                   const PhaseParams& p, hipStream_t s) {
  switch (adt) {
    case UA_FLOAT32: launch_phase<TO, TS, float>(mode, out, st, an, p, s); return UA_OK;
    case UA_COMPLEX64: launch_phase<TO, TS, float2>(mode, out, st, an, p, s); return UA_OK;
  }
  if constexpr (sizeof(TO) == 16) {
    switch (adt) {
      case UA_FLOAT64: launch_phase<TO, TS, double>(mode, out, st, an, p, s); return UA_OK;
      case UA_COMPLEX128: launch_phase<TO, TS, double2>(mode, out, st, an, p, s); return UA_OK;
    }
  }
  set_error("angle dtype %d cannot produce a result of this precision", adt);
  return UA_ERR_INVALID_ARGUMENT;
}

template <typename TO>
int dispatch_state(int sdt, int adt, int mode, void* out, const void* st,
                   const void* an, const PhaseParams& p, hipStream_t s) {
  switch (sdt) {
    case UA_FLOAT32: return dispatch_angle<TO, float>(adt, mode, out, st, an, p, s);
    case UA_COMPLEX64: return dispatch_angle<TO, float2>(adt, mode, out, st, an, p, s);
  }
  if constexpr (sizeof(TO) == 16) {
    switch (sdt) {
      case UA_FLOAT64: return dispatch_angle<TO, double>(adt, mode, out, st, an, p, s);
      case UA_COMPLEX128: return dispatch_angle<TO, double2>(adt, mode, out, st, an, p, s);
    }
  }
  set_error("state dtype %d cannot produce a result of this precision", sdt);
  return UA_ERR_INVALID_ARGUMENT;
}

bool is_row_major(int rank, const int64_t* shape, const int64_t* stride, int first) {
  int64_t expect = 1;
  for (int d = rank - 1; d >= first; --d) {
    if (shape[d] != 1 && stride[d] != expect) return false;
    expect *= shape[d];
  }
  return true;
}

}  // namespace

int apply_phase_impl(void* out, int odt, const void* state, int sdt, const void* angles,
                     int adt, int rank, const int64_t* shape, const int64_t* sstr,
                     const int64_t* astr, hipStream_t s) {
  UA_CHECK_ARG(odt == UA_COMPLEX64 || odt == UA_COMPLEX128,
               "phase output dtype must be complex (got %d)", odt);
  UA_CHECK_ARG(elem_size(sdt) && elem_size(adt), "unsupported dtype");
  UA_CHECK_ARG(rank >= 0 && rank <= kMaxRank, "rank %d exceeds %d", rank, kMaxRank);
  PhaseParams p{};
  p.rank = rank;
  p.total = 1;
  for (int d = 0; d < rank; ++d) {
    UA_CHECK_ARG(shape[d] >= 0, "negative size");
    p.shape[d] = shape[d];
    p.sstride[d] = sstr[d];
    p.astride[d] = astr[d];
    p.total *= shape[d];
  }
  if (p.total == 0) return UA_OK;
  UA_CHECK_ARG(out && state && angles, "NULL data pointer");

  int mode = GENERAL;
  if (is_row_major(rank, p.shape, p.sstride, 0)) {
    if (is_row_major(rank, p.shape, p.astride, 0)) {
      mode = CONTIG;
    } else {
      // angles contiguous over trailing dims [f, rank) and broadcast before
      for (int f = 0; f <= rank; ++f) {
        bool lead_zero = true;
        for (int d = 0; d < f; ++d) lead_zero &= (p.astride[d] == 0 || p.shape[d] == 1);
        if (lead_zero && is_row_major(rank, p.shape, p.astride, f)) {
          int64_t t = 1;
          for (int d = f; d < rank; ++d) t *= p.shape[d];
          p.tile = t;
          mode = TILE;
          break;
        }
      }
    }
  }
  int rc;
  if (odt == UA_COMPLEX64)
    rc = dispatch_state<float2>(sdt, adt, mode, out, state, angles, p, s);
  else
    rc = dispatch_state<double2>(sdt, adt, mode, out, state, angles, p, s);
  if (rc != UA_OK) return rc;
  return check_launch("phase_kernel");
}

}  // namespace ua

extern "C" int ua_apply_phase(void* out, int out_dtype, const void* state, int state_dtype,
                              const void* angles, int angle_dtype, int rank,
                              const int64_t* shape, const int64_t* state_strides,
                              const int64_t* angle_strides, ua_stream_t stream) {
  return ua::apply_phase_impl(out, out_dtype, state, state_dtype, angles, angle_dtype,
                              rank, shape, state_strides, angle_strides,
                              static_cast<hipStream_t>(stream));
}

// ---- angle gradient of apply_phase (real angles)
//
// torch differentiates exp(-1j*angles) * state with grad_angles =
// Re(1j * conj(out) * g) summed over the broadcast rows:
// Im(out) Re(g) - Re(out) Im(g).  One thread per angle column walks the rows
// (consecutive threads read consecutive amplitudes) and sums in float64.
namespace ua {
namespace {

template <typename T, typename G>
__global__ __launch_bounds__(kBlock) void phase_angle_grad_kernel(G* grad, const T* g,
                                                                  const T* out, int64_t rows,
                                                                  int64_t cols) {
  const int64_t x = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (x >= cols) return;
  double acc = 0.0;
  for (int64_t r = 0; r < rows; ++r) {
    const T o = out[r * cols + x], d = g[r * cols + x];
    acc += double(o.y) * double(d.x) - double(o.x) * double(d.y);
  }
  grad[x] = G(acc);
}

template <typename T>
int launch_angle_grad(void* grad, int grad_dtype, const void* g, const void* out, int64_t rows,
                      int64_t cols, hipStream_t s) {
  const unsigned grid = grid_for(cols);
  if (grad_dtype == UA_FLOAT32)
    phase_angle_grad_kernel<T, float><<<grid, kBlock, 0, s>>>(
        static_cast<float*>(grad), static_cast<const T*>(g), static_cast<const T*>(out), rows,
        cols);
  else
    phase_angle_grad_kernel<T, double><<<grid, kBlock, 0, s>>>(
        static_cast<double*>(grad), static_cast<const T*>(g), static_cast<const T*>(out), rows,
        cols);
  return check_launch("phase_angle_grad_kernel");
}

}  // namespace
}  // namespace ua

extern "C" int ua_phase_angle_grad(void* grad, int grad_dtype, const void* g, const void* out,
                                   int dtype, int64_t rows, int64_t cols, ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(dtype == UA_COMPLEX64 || dtype == UA_COMPLEX128,
               "phase angle gradient: complex states only (dtype %d)", dtype);
  UA_CHECK_ARG(grad_dtype == UA_FLOAT32 || grad_dtype == UA_FLOAT64,
               "phase angle gradient: real angles only (dtype %d)", grad_dtype);
  UA_CHECK_ARG(rows >= 0 && cols >= 0, "negative size");
  if (cols == 0) return UA_OK;
  UA_CHECK_ARG(grad && (rows == 0 || (g && out)), "NULL pointer");
  UA_CHECK_ARG(cols / kBlock < 0x7fffffffLL, "too many angles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == UA_COMPLEX64) return launch_angle_grad<float2>(grad, grad_dtype, g, out, rows, cols, s);
  return launch_angle_grad<double2>(grad, grad_dtype, g, out, rows, cols, s);
}
