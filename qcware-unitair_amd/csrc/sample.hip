// Sampling measurement outcomes: ua_sample replaces the probability and
// draw steps of unitair.simulation.measure (unitair/simulation/measurement.py:
// 9-71: abs_squared, then torch.distributions.Categorical(probs).sample).
//
// Sample s takes u = U(s) * total, U(s) a 53-bit uniform in [0, 1) from
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) with counter (s + offset, 0, 0, 0) and key (seed
// lo, seed hi).  The outcome is the first index whose inclusive prefix of
// |psi|^2 exceeds u, with prefixes formed in float64 inside chunks of 4096
// amplitudes and then across chunks:
// 1. scan_chunks_kernel: each chunk's total (chunk_sum).
// 2. scan_totals_kernel: one workgroup turns chunk_sum into inclusive chunk
//    prefixes.
// 3. bucket_kernel: per sample, the first chunk whose prefix exceeds u (the
//    chunk prefixes staged in LDS), and a slot in that chunk's bucket
//    (vector atomic on one of the chunk's kSets sub-counts, picked by the
//    workgroup, so that fewer atomics meet on one address).
// 4. count_scan_kernel / key_start_kernel / scatter_kernel: bucket starts
//    per chunk, then per sub-bucket (a chunk's sub-buckets one after
//    another), then the sample ids in bucket order.
// 5. draw_chunk_kernel: one workgroup per chunk that holds samples streams
//    the chunk from the state, rebuilds its in-chunk prefixes in LDS with the
//    same code as step 1 (so the values agree bit for bit), and searches
//    there for each of its samples.
// The state is read twice in order; no prefix array goes through HBM and no
// sample makes a scattered read.  Categorical(probs) normalises by the sum
// of probs; so does drawing u in [0, total).
#include "ua_common.h"

namespace ua {
namespace {

constexpr int kChunkBits = 12;
constexpr int kChunk = 1 << kChunkBits;       // amplitudes per chunk
constexpr int kScanThreads = 256;
constexpr int kPerThread = kChunk / kScanThreads;  // 16 amplitudes per thread
constexpr int kSets = 8;                           // sub-counts per chunk

template <typename T> __device__ __forceinline__ double prob(T v);
template <> __device__ __forceinline__ double prob(float v) { return double(v) * double(v); }
template <> __device__ __forceinline__ double prob(double v) { return v * v; }
template <> __device__ __forceinline__ double prob(float2 v) {
  return double(v.x) * double(v.x) + double(v.y) * double(v.y);
}
template <> __device__ __forceinline__ double prob(double2 v) { return v.x * v.x + v.y * v.y; }

// The chunk passes through LDS so that global loads stay coalesced while
// each thread scans 16 contiguous entries; rows of 16 are padded to 17
// doubles to spread the threads' LDS accesses over the banks.
constexpr int kRow = kPerThread + 1;

__device__ __forceinline__ int lds_slot(int i) { return (i / kPerThread) * kRow + i % kPerThread; }

// Chunk `c` of the state: leaves the inclusive in-chunk prefix of entry i
// at vals[lds_slot(i)] (entries at or beyond N count as 0) and returns the
// calling thread's inclusive total; thread kScanThreads-1 holds the chunk's.
template <typename T>
__device__ double chunk_prefix_lds(const T* state, int64_t N, int64_t c, double* vals,
                                   double* wave_tot) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = c * kChunk;
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
  __syncthreads();
  return before + run;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void scan_chunks_kernel(const T* state, int64_t N,
                                                                   double* chunk_sum) {
  __shared__ double vals[kScanThreads * kRow];
  __shared__ double wave_tot[kScanThreads / 64];
  const double t = chunk_prefix_lds(state, N, blockIdx.x, vals, wave_tot);
  if (threadIdx.x == kScanThreads - 1) chunk_sum[blockIdx.x] = t;
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

// u = U(s) * total, kept below total: the product can round up to it, and
// the first prefix above a value just below total is the last outcome with
// nonzero probability (an outcome of probability 0 adds nothing to the
// prefix, so it is never the first to exceed anything).
__device__ __forceinline__ double draw_u(int64_t s, uint64_t seed, uint64_t offset,
                                         double total) {
  const uint64_t ctr = uint64_t(s) + offset;
  uint32_t c[4] = {uint32_t(ctr), uint32_t(ctr >> 32), 0u, 0u};
  philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
  const uint64_t bits53 = (uint64_t(c[0]) << 21) ^ (uint64_t(c[1]) >> 11);
  const double u = double(bits53 & ((uint64_t(1) << 53) - 1)) * 0x1.0p-53 * total;
  return fmin(u, nextafter(total, 0.0));
}

// First index i in [0, m) with a[i] > u, or m - 1 if there is none; m is a
// power of two.  Branch-free: log2(m) steps for every u.
__device__ __forceinline__ int64_t first_above(const double* a, int64_t m, double u) {
  int64_t at = 0;
  for (int64_t half = m >> 1; half >= 1; half >>= 1) at += (a[at + half - 1] <= u) ? half : 0;
  return at;
}

__global__ __launch_bounds__(kBlock) void zero_counts_kernel(uint32_t* count, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) count[i] = 0u;
}

// LDS_PREFIX: the chunk prefixes (at most kLdsChunks) are staged in LDS once
// per workgroup.
constexpr int kLdsChunks = 4096;  // 32 KiB of doubles (states up to 24 qubits)

template <bool LDS_PREFIX>
__global__ __launch_bounds__(kBlock) void bucket_kernel(int64_t num_samples,
                                                       const double* gprefix, int64_t chunks,
                                                       uint64_t seed, uint64_t offset,
                                                       uint32_t* cid, uint32_t* slot,
                                                       uint32_t* count) {
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
  const double total = prefix[chunks - 1];
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  for (int64_t s = int64_t(blockIdx.x) * kBlock + threadIdx.x; s < num_samples; s += stride) {
    const int64_t c = first_above(prefix, chunks, draw_u(s, seed, offset, total));
    const uint32_t key = uint32_t(c * kSets + (blockIdx.x & (kSets - 1)));
    cid[s] = key;
    slot[s] = atomicAdd(&count[key], 1u);
  }
}

// One workgroup: exclusive prefix over chunks of the bucket sizes (each the
// sum of the chunk's kSets sub-counts), 1024 chunks at a time.
__global__ __launch_bounds__(1024) void count_scan_kernel(const uint32_t* count, int64_t chunks,
                                                         uint32_t* start) {
  __shared__ uint32_t wave_tot[16];
  __shared__ uint32_t carry_s;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int64_t base = 0; base < chunks; base += 1024) {
    const int64_t i = base + tid;
    uint32_t v = 0;
    if (i < chunks) {
#pragma unroll
      for (int k = 0; k < kSets; ++k) v += count[i * kSets + k];
    }
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wave_tot[wave] = x;
    __syncthreads();
    uint32_t before = carry_s;
    for (int w = 0; w < wave; ++w) before += wave_tot[w];
    if (i < chunks) start[i] = before + x - v;
    __syncthreads();
    if (tid == 1023) carry_s = before + x;
    __syncthreads();
  }
}

// A chunk's bucket holds its sub-buckets in order.  One thread per chunk
// turns the chunk's kSets sub-counts into the sub-buckets' starts, in place.
__global__ __launch_bounds__(kBlock) void key_start_kernel(uint32_t* count, const uint32_t* start,
                                                          int64_t chunks) {
  const int64_t c = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (c >= chunks) return;
  uint32_t at = start[c];
#pragma unroll
  for (int k = 0; k < kSets; ++k) {
    const uint32_t n = count[c * kSets + k];
    count[c * kSets + k] = at;
    at += n;
  }
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(int64_t num_samples, const uint32_t* cid,
                                                        const uint32_t* slot,
                                                        const uint32_t* key_start,
                                                        uint32_t* bucket) {
  const int64_t s = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (s < num_samples) bucket[key_start[cid[s]] + slot[s]] = uint32_t(s);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void draw_chunk_kernel(
    int64_t* out, const T* state, int64_t N, int64_t chunk_len, const double* prefix,
    int64_t chunks, int64_t num_samples, const uint32_t* start, const uint32_t* bucket,
    uint64_t seed, uint64_t offset) {
  __shared__ double vals[kScanThreads * kRow];
  __shared__ double wave_tot[kScanThreads / 64];
  const int64_t c = blockIdx.x;
  const uint32_t first = start[c];
  const uint32_t n_c = uint32_t(c + 1 < chunks ? start[c + 1] : num_samples) - first;
  if (n_c == 0) return;  // the whole workgroup leaves together
  chunk_prefix_lds(state, N, c, vals, wave_tot);
  const double total = prefix[chunks - 1];
  const double below = c > 0 ? prefix[c - 1] : 0.0;
  const double below_chunk_total = nextafter(vals[lds_slot(int(chunk_len - 1))], 0.0);
  for (uint32_t j = threadIdx.x; j < n_c; j += kScanThreads) {
    const int64_t s = bucket[first + j];
    // what is left of u inside the chunk, kept below the chunk's total (the
    // subtraction can round up to it) for the same reason as u itself
    const double ul = fmin(draw_u(s, seed, offset, total) - below, below_chunk_total);
    int64_t at = 0;
    for (int64_t half = chunk_len >> 1; half >= 1; half >>= 1)
      at += (vals[lds_slot(int(at + half - 1))] <= ul) ? half : 0;
    out[s] = c * chunk_len + at;
  }
}

struct SampleLayout {
  int64_t chunks;
  int64_t bytes;
};

SampleLayout sample_layout(int num_qubits, int64_t num_samples) {
  const int64_t N = int64_t(1) << num_qubits;
  SampleLayout l;
  l.chunks = (N + kChunk - 1) / kChunk;
  // chunk_sum, prefix (float64); count (uint32, kSets per chunk); start
  // (uint32 per chunk); cid, slot, bucket (uint32 per sample)
  const int64_t b = 8 * 2 * l.chunks + 4 * (kSets + 1) * l.chunks + 4 * 3 * num_samples;
  l.bytes = (b + 7) / 8 * 8;
  return l;
}

}  // namespace
}  // namespace ua

extern "C" int64_t ua_sample_workspace_bytes(int num_qubits, int64_t num_samples) {
  if (num_qubits < 0 || num_qubits > 40 || num_samples < 0 || num_samples > 0x7fffffffLL)
    return -1;
  return ua::sample_layout(num_qubits, num_samples).bytes;
}

extern "C" int64_t ua_sample_total_offset(int num_qubits, int64_t num_samples) {
  if (num_qubits < 0 || num_qubits > 40 || num_samples < 0 || num_samples > 0x7fffffffLL)
    return -1;
  // chunk_sum[chunks], then prefix[chunks]: the total is the last prefix
  const int64_t chunks = ua::sample_layout(num_qubits, num_samples).chunks;
  return 8 * (2 * chunks - 1);
}

extern "C" int ua_sample(int64_t* out, const void* state, int dtype, int num_qubits,
                         int64_t num_samples, uint64_t seed, uint64_t offset, void* workspace,
                         ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 0 && num_qubits <= 40, "num_qubits=%d out of range", num_qubits);
  UA_CHECK_ARG(dtype >= UA_FLOAT32 && dtype <= UA_COMPLEX128, "bad dtype %d", dtype);
  UA_CHECK_ARG(num_samples >= 0 && num_samples <= 0x7fffffffLL,
               "num_samples=%lld out of range (0 .. 2^31 - 1)", (long long)num_samples);
  if (num_samples == 0) return UA_OK;
  UA_CHECK_ARG(out && state && workspace, "NULL pointer");
  UA_CHECK_ARG(reinterpret_cast<uintptr_t>(workspace) % 8 == 0, "workspace must be 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t N = int64_t(1) << num_qubits;
  const int64_t chunks = sample_layout(num_qubits, num_samples).chunks;
  const int64_t chunk_len = N < kChunk ? N : kChunk;  // chunks and chunk_len are powers of two
  double* chunk_sum = static_cast<double*>(workspace);
  double* prefix = chunk_sum + chunks;
  const int64_t keys = chunks * kSets;  // (chunk, sub-count) pairs, chunk-major
  uint32_t* count = reinterpret_cast<uint32_t*>(prefix + chunks);
  uint32_t* start = count + keys;      // one per chunk
  uint32_t* cid = start + chunks;
  uint32_t* slot = cid + num_samples;
  uint32_t* bucket = slot + num_samples;

  zero_counts_kernel<<<grid_for(keys), kBlock, 0, s>>>(count, keys);
  if (int rc = check_launch("zero_counts_kernel")) return rc;
#define UA_SCAN(T) \
  scan_chunks_kernel<T><<<unsigned(chunks), kScanThreads, 0, s>>>(static_cast<const T*>(state), N, chunk_sum)
  switch (dtype) {
    case UA_FLOAT32: UA_SCAN(float); break;
    case UA_FLOAT64: UA_SCAN(double); break;
    case UA_COMPLEX64: UA_SCAN(float2); break;
    default: UA_SCAN(double2); break;
  }
#undef UA_SCAN
  if (int rc = check_launch("scan_chunks_kernel")) return rc;
  scan_totals_kernel<<<1, 1024, 0, s>>>(chunk_sum, chunks, prefix);
  if (int rc = check_launch("scan_totals_kernel")) return rc;
  const int64_t blocks = (num_samples + kBlock - 1) / kBlock;
  if (chunks <= kLdsChunks) {
    const unsigned grid = unsigned(blocks < 2048 ? blocks : 2048);
    bucket_kernel<true><<<grid, kBlock, size_t(chunks) * sizeof(double), s>>>(
        num_samples, prefix, chunks, seed, offset, cid, slot, count);
  } else {
    bucket_kernel<false><<<unsigned(blocks < 8192 ? blocks : 8192), kBlock, 0, s>>>(
        num_samples, prefix, chunks, seed, offset, cid, slot, count);
  }
  if (int rc = check_launch("bucket_kernel")) return rc;
  count_scan_kernel<<<1, 1024, 0, s>>>(count, chunks, start);
  if (int rc = check_launch("count_scan_kernel")) return rc;
  key_start_kernel<<<grid_for(chunks), kBlock, 0, s>>>(count, start, chunks);
  if (int rc = check_launch("key_start_kernel")) return rc;
  scatter_kernel<<<unsigned(blocks), kBlock, 0, s>>>(num_samples, cid, slot, count, bucket);
  if (int rc = check_launch("scatter_kernel")) return rc;
#define UA_DRAW(T)                                                                           \
  draw_chunk_kernel<T><<<unsigned(chunks), kScanThreads, 0, s>>>(                            \
      out, static_cast<const T*>(state), N, chunk_len, prefix, chunks, num_samples, start,   \
      bucket,                                                                                \
      seed, offset)
  switch (dtype) {
    case UA_FLOAT32: UA_DRAW(float); break;
    case UA_FLOAT64: UA_DRAW(double); break;
    case UA_COMPLEX64: UA_DRAW(float2); break;
    default: UA_DRAW(double2); break;
  }
#undef UA_DRAW
  return check_launch("draw_chunk_kernel");
}
