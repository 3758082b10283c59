// Vector kernels for dense k-qubit gates (k <= 5) and their launch
// helpers, shared by dense.hip (dispatch) and the per-dtype instantiation
// files dense_f32/f64/c64/c128.hip and dense_c64/c128_k45.hip, which compile
// in parallel.  See dense.hip
// for the algorithm.
#pragma once
#include "ua_common.h"

namespace ua {
namespace dense_detail {

constexpr int kMaxFastK = 5;

enum Mode { SCALAR = 0, VEC = 1, PAIR = 2 };

struct FastParams {
  int64_t total;        // threads that own work
  int64_t in_bstride;   // elements between input states
  int64_t out_bstride;  // elements between output states
  int64_t op_bstride;   // elements between operators
  int log2_tps;         // log2(threads per state)
  int tb[kMaxFastK];    // ascending target bit positions
  int gpos[kMaxFastK];  // gate-index bit of canonical bit p
};

// Original (caller's) gate index of canonical index c.
template <int K>
__device__ __forceinline__ int orig_index(int c, const FastParams& p) {
  int r = 0;
#pragma unroll
  for (int q = 0; q < K; ++q) r |= ((c >> q) & 1) << p.gpos[q];
  return r;
}

// State offset of canonical gate index j (j is a compile-time constant at
// every call site, so this folds to a few adds).
template <int K>
__device__ __forceinline__ uint64_t offset_of(int j, const FastParams& p) {
  uint64_t o = 0;
#pragma unroll
  for (int q = 0; q < K; ++q)
    if ((j >> q) & 1) o += uint64_t(1) << p.tb[q];
  return o;
}

template <typename T, int K, int MODE, bool LDS, int UNR>
__global__ __launch_bounds__(kBlock) void dense_fast_kernel(
    T* out, const T* in, const T* __restrict__ op, FastParams p) {
  using A = Num<T>;
  constexpr int D = 1 << K;
  constexpr int G = (MODE == VEC) ? 2 : 1;  // groups per unit
  // UNR > 1 only with one operator for the whole launch (op_bstride == 0).
  const int64_t first = int64_t(blockIdx.x) * (kBlock * UNR) + threadIdx.x;

  __shared__ T smat[LDS ? D * D : 1];
  if constexpr (LDS) {
    // The host only stages in LDS when every thread of the workgroup uses the
    // same operator.
    const int64_t b0 = (int64_t(blockIdx.x) * (kBlock * UNR)) >> p.log2_tps;
    const T* opb = op + b0 * p.op_bstride;
    for (int e = threadIdx.x; e < D * D; e += kBlock) {
      const int r = e >> K, c = e & (D - 1);
      smat[e] = opb[orig_index<K>(r, p) * D + orig_index<K>(c, p)];
    }
    __syncthreads();
  }

  // ---- load UNR units (each G groups of D amplitudes) before any math
  T a[UNR][D][G];
  T* dst[UNR];
  const T* opb = op;
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    const int64_t unit = first + int64_t(u) * kBlock;
    dst[u] = nullptr;
    if (unit >= p.total) continue;
    const int64_t b = unit >> p.log2_tps;
    const uint64_t t = uint64_t(unit) & ((uint64_t(1) << p.log2_tps) - 1);
    const uint64_t base = insert_zero_bits(t * G, p.tb, K);
    const T* src = in + b * p.in_bstride + base;
    dst[u] = out + b * p.out_bstride + base;
    if (UNR == 1) opb = op + b * p.op_bstride;
    if constexpr (MODE == VEC) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const Pair<T> v = *reinterpret_cast<const Pair<T>*>(src + offset_of<K>(j, p));
        a[u][j][0] = v.v[0];
        a[u][j][1] = v.v[1];
      }
    } else if constexpr (MODE == PAIR) {
#pragma unroll
      for (int j = 0; j < D; j += 2) {
        const Pair<T> v = *reinterpret_cast<const Pair<T>*>(src + offset_of<K>(j, p));
        a[u][j][0] = v.v[0];
        a[u][j + 1][0] = v.v[1];
      }
    } else {
#pragma unroll
      for (int j = 0; j < D; ++j) a[u][j][0] = src[offset_of<K>(j, p)];
    }
  }

  // ---- matrix element accessor (canonical order)
  T ureg[(LDS || K > 2) ? 1 : D * D];
  if constexpr (!LDS && K <= 2) {
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int c = 0; c < D; ++c)
        ureg[r * D + c] = opb[orig_index<K>(r, p) * D + orig_index<K>(c, p)];
  }
  auto U = [&](int r, int c) -> T {
    if constexpr (LDS) {
      return smat[r * D + c];
    } else if constexpr (K <= 2) {
      return ureg[r * D + c];
    } else {
      return opb[orig_index<K>(r, p) * D + orig_index<K>(c, p)];
    }
  };

  // ---- multiply and store
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    if (dst[u] == nullptr) continue;
    if constexpr (MODE == PAIR) {
#pragma unroll
      for (int r = 0; r < D; r += 2) {
        T acc0 = A::zero(), acc1 = A::zero();
#pragma unroll
        for (int j = 0; j < D; ++j) {
          acc0 = A::mac(acc0, U(r, j), a[u][j][0]);
          acc1 = A::mac(acc1, U(r + 1, j), a[u][j][0]);
        }
        Pair<T> v;
        v.v[0] = acc0;
        v.v[1] = acc1;
        *reinterpret_cast<Pair<T>*>(dst[u] + offset_of<K>(r, p)) = v;
      }
    } else {
#pragma unroll
      for (int r = 0; r < D; ++r) {
        T acc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] = A::zero();
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const T uu = U(r, j);
#pragma unroll
          for (int g = 0; g < G; ++g) acc[g] = A::mac(acc[g], uu, a[u][j][g]);
        }
        if constexpr (MODE == VEC) {
          Pair<T> v;
          v.v[0] = acc[0];
          v.v[1] = acc[1];
          *reinterpret_cast<Pair<T>*>(dst[u] + offset_of<K>(r, p)) = v;
        } else {
          dst[u][offset_of<K>(r, p)] = acc[0];
        }
      }
    }
  }
}

template <typename T, int K, int MODE, bool LDS>
void launch_fast(void* out, const void* in, const void* op, const FastParams& p, int unr,
                 hipStream_t s) {
  T* o = static_cast<T*>(out);
  const T* i = static_cast<const T*>(in);
  const T* m = static_cast<const T*>(op);
  switch (unr) {
    case 4:
      dense_fast_kernel<T, K, MODE, LDS, 4><<<grid_for(p.total, kBlock * 4), kBlock, 0, s>>>(o, i, m, p);
      break;
    case 2:
      dense_fast_kernel<T, K, MODE, LDS, 2><<<grid_for(p.total, kBlock * 2), kBlock, 0, s>>>(o, i, m, p);
      break;
    default:
      dense_fast_kernel<T, K, MODE, LDS, 1><<<grid_for(p.total, kBlock), kBlock, 0, s>>>(o, i, m, p);
  }
}

template <typename T, int K, int MODE>
void launch_fast_lds(bool lds, void* out, const void* in, const void* op,
                     const FastParams& p, int unr, hipStream_t s) {
  if constexpr (K >= 3) {
    if (lds) return launch_fast<T, K, MODE, true>(out, in, op, p, unr, s);
  }
  launch_fast<T, K, MODE, false>(out, in, op, p, unr, s);
}

template <typename T, int K>
void launch_fast_mode(int mode, bool lds, void* out, const void* in,
                      const void* op, const FastParams& p, int unr, hipStream_t s) {
  switch (mode) {
    case VEC: return launch_fast_lds<T, K, VEC>(lds, out, in, op, p, unr, s);
    case PAIR: return launch_fast_lds<T, K, PAIR>(lds, out, in, op, p, unr, s);
    default: return launch_fast_lds<T, K, SCALAR>(lds, out, in, op, p, unr, s);
  }
}

// k = 1..3 and k = 4..5 instantiate separately (separate translation units
// for the complex dtypes, whose 4/5-qubit kernels take longest to compile).
template <typename T>
void launch_fast_k123(int k, int mode, bool lds, void* out, const void* in,
                      const void* op, const FastParams& p, int unr, hipStream_t s) {
  switch (k) {
    case 1: return launch_fast_mode<T, 1>(mode, lds, out, in, op, p, unr, s);
    case 2: return launch_fast_mode<T, 2>(mode, lds, out, in, op, p, unr, s);
    default: return launch_fast_mode<T, 3>(mode, lds, out, in, op, p, unr, s);
  }
}

template <typename T>
void launch_fast_k45(int k, int mode, bool lds, void* out, const void* in,
                     const void* op, const FastParams& p, int unr, hipStream_t s) {
  if (k == 4) return launch_fast_mode<T, 4>(mode, lds, out, in, op, p, unr, s);
  launch_fast_mode<T, 5>(mode, lds, out, in, op, p, unr, s);
}

// One instantiation per dtype (complex: and per k range), each in its own
// translation unit.
void launch_fast_k45_c64(int k, int mode, bool lds, void* out, const void* in, const void* op,
                         const FastParams& p, int unr, hipStream_t s);
void launch_fast_k45_c128(int k, int mode, bool lds, void* out, const void* in, const void* op,
                          const FastParams& p, int unr, hipStream_t s);
void launch_fast_k_f32(int k, int mode, bool lds, void* out, const void* in, const void* op,
                       const FastParams& p, int unr, hipStream_t s);
void launch_fast_k_f64(int k, int mode, bool lds, void* out, const void* in, const void* op,
                       const FastParams& p, int unr, hipStream_t s);
void launch_fast_k_c64(int k, int mode, bool lds, void* out, const void* in, const void* op,
                       const FastParams& p, int unr, hipStream_t s);
void launch_fast_k_c128(int k, int mode, bool lds, void* out, const void* in, const void* op,
                        const FastParams& p, int unr, hipStream_t s);

}  // namespace dense_detail
}  // namespace ua
