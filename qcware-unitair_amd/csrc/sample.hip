// Sampling measurement outcomes: ua_sample replaces the probability and
// draw steps of unitair.simulation.measure (unitair/simulation/measurement.py:
// 9-71: abs_squared, then torch.distributions.Categorical(probs).sample).
//
// 1. scan_chunks_kernel: |psi_x|^2 in float64, inclusive prefix sums inside
//    chunks of 4096 amplitudes (cum), the prefix at the end of every row of
//    16 (row_cum) and of every group of 16 rows (group_cum), and each
//    chunk's total (chunk_sum).
// 2. scan_totals_kernel: one workgroup turns chunk_sum into inclusive chunk
//    prefixes (chunk_prefix).
// 3. draw_kernel: sample s takes u = U(s) * total, U(s) a 53-bit uniform in
//    [0, 1) from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
//    numbers: as easy as 1, 2, 3", SC'11) with counter (s + offset, 0, 0, 0)
//    and key (seed lo, seed hi); the outcome is the first index whose
//    inclusive prefix exceeds u: a binary search over chunk_prefix, then over
//    that chunk's group_cum, row_cum and cum (branch-free, four samples per
//    thread in lockstep).
// Categorical(probs) normalises by the sum of probs; so does drawing u in
// [0, total).
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kChunkBits = 12;
constexpr int kChunk = 1 << kChunkBits;       // amplitudes per chunk
constexpr int kScanThreads = 256;
constexpr int kPerThread = kChunk / kScanThreads;  // 16 amplitudes per row
constexpr int kFan = 16;                           // rows per group
constexpr int kGroups = kScanThreads / kFan;       // groups per chunk

template <typename T> __device__ __forceinline__ double prob(T v);
template <> __device__ __forceinline__ double prob(float v) { return double(v) * double(v); }
template <> __device__ __forceinline__ double prob(double v) { return v * v; }
template <> __device__ __forceinline__ double prob(float2 v) {
  return double(v.x) * double(v.x) + double(v.y) * double(v.y);
}
template <> __device__ __forceinline__ double prob(double2 v) { return v.x * v.x + v.y * v.y; }

// The chunk passes through LDS so that global loads and stores stay
// coalesced while each thread scans 16 contiguous entries; rows of 16 are
// padded to 17 doubles to spread the threads' LDS accesses over the banks.
constexpr int kRow = kPerThread + 1;

__device__ __forceinline__ int lds_slot(int i) { return (i / kPerThread) * kRow + i % kPerThread; }

template <typename T>
__global__ __launch_bounds__(kScanThreads) void scan_chunks_kernel(const T* state, int64_t N,
                                                                   double* cum, double* row_cum,
                                                                   double* group_cum,
                                                                   double* chunk_sum) {
  __shared__ double vals[kScanThreads * kRow];
  __shared__ double wave_tot[kScanThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = int64_t(blockIdx.x) * kChunk;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int i = j * kScanThreads + tid;
    vals[lds_slot(i)] = (c0 + i < N) ? prob(state[c0 + i]) : 0.0;
  }
  __syncthreads();
  double local[kPerThread];
  double run = 0;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    run += vals[tid * kRow + j];
    local[j] = run;
  }
  // inclusive scan of the thread totals across the wavefront
  double x = run;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wave_tot[wave] = x;
  __syncthreads();
  double before = x - run;  // exclusive within the wavefront
  for (int w = 0; w < wave; ++w) before += wave_tot[w];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) vals[tid * kRow + j] = before + local[j];
  // the row's last in-chunk prefix, bit for bit the cum entry it ends on
  row_cum[int64_t(blockIdx.x) * kScanThreads + tid] = before + local[kPerThread - 1];
  // and every 16th row's, for the level above
  if (tid % kFan == kFan - 1)
    group_cum[int64_t(blockIdx.x) * kGroups + tid / kFan] = before + local[kPerThread - 1];
  if (tid == kScanThreads - 1) chunk_sum[blockIdx.x] = before + run;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int i = j * kScanThreads + tid;
    if (c0 + i < N) cum[c0 + i] = vals[lds_slot(i)];
  }
}

// One workgroup: inclusive prefix of the chunk totals, 1024 at a time.
__global__ __launch_bounds__(1024) void scan_totals_kernel(const double* chunk_sum,
                                                          int64_t chunks, double* prefix) {
  __shared__ double wave_tot[16];
  __shared__ double carry_s;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int64_t start = 0; start < chunks; start += 1024) {
    const int64_t i = start + tid;
    const double v = i < chunks ? chunk_sum[i] : 0.0;
    double x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wave_tot[wave] = x;
    __syncthreads();
    double before = carry_s;
    for (int w = 0; w < wave; ++w) before += wave_tot[w];
    if (i < chunks) prefix[i] = before + x;
    __syncthreads();
    if (tid == 1023) carry_s = before + x;
    __syncthreads();
  }
}

// Philox4x32-10 as published (SC'11): 10 rounds of two 32x32->64 multiplies
// with the Weyl key schedule.
__device__ __forceinline__ void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(M0) * c[0];
    const uint64_t p1 = uint64_t(M1) * c[2];
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

// LDS_PREFIX: the chunk prefixes (at most kLdsChunks) are staged in LDS once
// per workgroup, so the first half of every search stays on chip.
constexpr int kLdsChunks = 4096;  // 32 KiB of doubles (states up to 24 qubits)
// Samples per thread, searched in lockstep: their dependent loads overlap.
constexpr int kLanesPerThread = 4;

// First index i in [0, m) with a[i] > u, or m - 1 if there is none; m is a
// power of two.  Branch-free: log2(m) steps for every u.
// The S loads of a step are all issued before any is compared, so that the
// S searches' memory latencies overlap.
template <int S>
__device__ __forceinline__ void search_pow2(const double* __restrict__ a, int64_t m,
                                            const double (&u)[S], int64_t (&at)[S]) {
  for (int64_t half = m >> 1; half >= 1; half >>= 1) {
    double v[S];
#pragma unroll
    for (int i = 0; i < S; ++i) v[i] = a[at[i] + half - 1];
#pragma unroll
    for (int i = 0; i < S; ++i) at[i] += (v[i] <= u[i]) ? half : 0;
  }
}

template <bool LDS_PREFIX>
__global__ __launch_bounds__(kBlock) void draw_kernel(int64_t* out, int64_t num_samples,
                                                     const double* cum, const double* row_cum,
                                                     const double* group_cum,
                                                     const double* gprefix, int64_t chunks,
                                                     int64_t chunk_len, uint64_t seed,
                                                     uint64_t offset) {
  extern __shared__ double sprefix[];
  const double* prefix = gprefix;
  if constexpr (LDS_PREFIX) {
    // four loads in flight per thread before their LDS stores
    for (int64_t i0 = threadIdx.x; i0 < chunks; i0 += 4 * kBlock) {
      double v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t i = i0 + k * kBlock;
        v[k] = i < chunks ? gprefix[i] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < chunks) sprefix[i] = v[k];
      }
    }
    __syncthreads();
    prefix = sprefix;
  }
  constexpr int S = kLanesPerThread;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const double total = prefix[chunks - 1];
  for (int64_t s0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; s0 < num_samples;
       s0 += S * stride) {
    double u[S];
    int64_t lo[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const uint64_t ctr = uint64_t(s0 + i * stride) + offset;
      uint32_t c[4] = {uint32_t(ctr), uint32_t(ctr >> 32), 0u, 0u};
      philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
      const uint64_t bits53 = (uint64_t(c[0]) << 21) ^ (uint64_t(c[1]) >> 11);
      u[i] = double(bits53 & ((uint64_t(1) << 53) - 1)) * 0x1.0p-53 * total;
      lo[i] = 0;
    }
    // first chunk whose inclusive prefix exceeds u, then the first index of
    // that chunk whose inclusive (in-chunk) prefix exceeds what is left of
    // u, down a tree of fan-out 16: the group of 16 rows by the groups'
    // last prefixes, the row by the rows' last prefixes, the entry inside
    // the row -- one 128-byte line per level.  Rounding can leave u at a
    // chunk's total: every level ends on its last entry then.
    search_pow2<S>(prefix, chunks, u, lo);
    const int64_t row_len = chunk_len < kPerThread ? chunk_len : kPerThread;
    const int64_t rows = chunk_len / row_len;
    const int64_t groups = rows >= kFan ? rows / kFan : 1;
    const int64_t rows_per_group = rows / groups;
    double ul[S];
    int64_t at[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      ul[i] = u[i] - (lo[i] > 0 ? prefix[lo[i] - 1] : 0.0);
      at[i] = lo[i] * kGroups;
    }
    search_pow2<S>(group_cum, groups, ul, at);
#pragma unroll
    for (int i = 0; i < S; ++i)
      at[i] = lo[i] * kScanThreads + (at[i] - lo[i] * kGroups) * rows_per_group;
    search_pow2<S>(row_cum, rows_per_group, ul, at);
#pragma unroll
    for (int i = 0; i < S; ++i) at[i] = lo[i] * chunk_len + (at[i] - lo[i] * kScanThreads) * row_len;
    search_pow2<S>(cum, row_len, ul, at);
#pragma unroll
    for (int i = 0; i < S; ++i)
      if (s0 + i * stride < num_samples) out[s0 + i * stride] = at[i];
  }
}

}  // namespace
}  // namespace ua

extern "C" int64_t ua_sample_workspace_bytes(int num_qubits) {
  if (num_qubits < 0 || num_qubits > 40) return -1;
  const int64_t N = int64_t(1) << num_qubits;
  const int64_t chunks = (N + ua::kChunk - 1) / ua::kChunk;
  return 8 * (N + chunks * (2 + ua::kScanThreads + ua::kGroups));
}

extern "C" int ua_sample(int64_t* out, const void* state, int dtype, int num_qubits,
                         int64_t num_samples, uint64_t seed, uint64_t offset, void* workspace,
                         ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 0 && num_qubits <= 40, "num_qubits=%d out of range", num_qubits);
  UA_CHECK_ARG(dtype >= UA_FLOAT32 && dtype <= UA_COMPLEX128, "bad dtype %d", dtype);
  UA_CHECK_ARG(num_samples >= 0, "negative num_samples");
  if (num_samples == 0) return UA_OK;
  UA_CHECK_ARG(out && state && workspace, "NULL pointer");
  UA_CHECK_ARG(reinterpret_cast<uintptr_t>(workspace) % 8 == 0, "workspace must be 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t N = int64_t(1) << num_qubits;
  const int64_t chunks = (N + kChunk - 1) / kChunk;
  double* cum = static_cast<double*>(workspace);
  double* chunk_sum = cum + N;
  double* prefix = chunk_sum + chunks;
  double* row_cum = prefix + chunks;              // kScanThreads rows per chunk
  double* group_cum = row_cum + chunks * kScanThreads;  // kGroups groups per chunk
  switch (dtype) {
    case UA_FLOAT32:
      scan_chunks_kernel<float><<<unsigned(chunks), kScanThreads, 0, s>>>(
          static_cast<const float*>(state), N, cum, row_cum, group_cum, chunk_sum);
      break;
    case UA_FLOAT64:
      scan_chunks_kernel<double><<<unsigned(chunks), kScanThreads, 0, s>>>(
          static_cast<const double*>(state), N, cum, row_cum, group_cum, chunk_sum);
      break;
    case UA_COMPLEX64:
      scan_chunks_kernel<float2><<<unsigned(chunks), kScanThreads, 0, s>>>(
          static_cast<const float2*>(state), N, cum, row_cum, group_cum, chunk_sum);
      break;
    default:
      scan_chunks_kernel<double2><<<unsigned(chunks), kScanThreads, 0, s>>>(
          static_cast<const double2*>(state), N, cum, row_cum, group_cum, chunk_sum);
      break;
  }
  if (int rc = check_launch("scan_chunks_kernel")) return rc;
  scan_totals_kernel<<<1, 1024, 0, s>>>(chunk_sum, chunks, prefix);
  if (int rc = check_launch("scan_totals_kernel")) return rc;
  // chunks and chunk_len are powers of two (N is)
  const int64_t chunk_len = N < kChunk ? N : kChunk;
  const int64_t per_block = int64_t(kBlock) * kLanesPerThread;
  const int64_t blocks = (num_samples + per_block - 1) / per_block;
  if (chunks <= kLdsChunks) {
    // enough workgroups to fill the GPU, each drawing many samples
    const unsigned grid = unsigned(blocks < 1024 ? blocks : 1024);
    draw_kernel<true><<<grid, kBlock, size_t(chunks) * sizeof(double), s>>>(
        out, num_samples, cum, row_cum, group_cum, prefix, chunks, chunk_len, seed, offset);
  } else {
    draw_kernel<false><<<unsigned(blocks < 8192 ? blocks : 8192), kBlock, 0, s>>>(
        out, num_samples, cum, row_cum, group_cum, prefix, chunks, chunk_len, seed, offset);
  }
  return check_launch("draw_kernel");
}
