/*
 * Closed-form CPU restatement of unitair's dense gate application, complex128.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, built by
 * oracle/Makefile, loaded by oracle/closed_form.py from tests/, smoke() and
 * bench.py's cpu_baseline leg.  Never linked into the product.
 *
 * It states apply_operator (unitair/simulation/operations.py:45-148) by its
 * index definition instead of the reference's permute/einsum sequence, so it
 * is an independent check of oracle/unitair_np.py:
 *
 *   out[i] = sum_j U[r(i)][j] * in[(i & ~mask) | scatter(j)]
 *
 * with qubit q = index bit n-1-q and qubits[0] the most significant bit of
 * the gate index r (operations.py:78-86, unitair/utils.py:5-27).  Groups of
 * 2^k amplitudes are read before they are written, so out may equal in.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define CF_MAXK 12

/* Returns 0 on success, -1 on bad arguments. */
int cf_apply_dense_c128(double* out, const double* in, const double* op, int n, int k,
                        const int* qubits, int64_t batch, int64_t in_bstride,
                        int64_t op_bstride) {
  if (n < 0 || n > 40 || k < 0 || k > n || k > CF_MAXK) return -1;
  int bits[CF_MAXK];   /* state bit of gate bit k-1-m, in `qubits` order */
  int sorted[CF_MAXK]; /* ascending state bits, for zero insertion */
  for (int m = 0; m < k; ++m) {
    if (qubits[m] < 0 || qubits[m] >= n) return -1;
    bits[m] = n - 1 - qubits[m];
    sorted[m] = bits[m];
  }
  for (int i = 1; i < k; ++i)
    for (int j = i; j > 0 && sorted[j - 1] > sorted[j]; --j) {
      int t = sorted[j]; sorted[j] = sorted[j - 1]; sorted[j - 1] = t;
    }
  for (int i = 1; i < k; ++i)
    if (sorted[i] == sorted[i - 1]) return -1;
  const int64_t D = (int64_t)1 << k;
  const int64_t N = (int64_t)1 << n;
  const int64_t groups = N >> k;
  /* offsets of gate index j */
  int64_t* off = malloc(sizeof(int64_t) * D);
  for (int64_t j = 0; j < D; ++j) {
    int64_t o = 0;
    for (int m = 0; m < k; ++m)
      if ((j >> (k - 1 - m)) & 1) o |= (int64_t)1 << bits[m];
    off[j] = o;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t b = 0; b < batch; ++b) {
    for (int64_t g = 0; g < groups; ++g) {
      double complex a[1 << CF_MAXK];
      const double complex* src = (const double complex*)in + b * in_bstride;
      double complex* dst = (double complex*)out + b * N;
      const double complex* U = (const double complex*)op + b * op_bstride;
      uint64_t base = (uint64_t)g;
      for (int p = 0; p < k; ++p) {
        const uint64_t low = base & ((1ULL << sorted[p]) - 1);
        base = ((base >> sorted[p]) << (sorted[p] + 1)) | low;
      }
      for (int64_t j = 0; j < D; ++j) a[j] = src[base | off[j]];
      for (int64_t r = 0; r < D; ++r) {
        double complex acc = 0;
        for (int64_t j = 0; j < D; ++j) acc += U[r * D + j] * a[j];
        dst[base | off[r]] = acc;
      }
    }
  }
  free(off);
  return 0;
}

/* apply_phase (operations.py:15-42) for float64 angles:
 * out[b][x] = exp(-i angles[b*angle_bstride + x]) * in[b][x]. */
void cf_apply_phase_c128(double* out, const double* in, const double* angles, int64_t len,
                         int64_t batch, int64_t angle_bstride) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t b = 0; b < batch; ++b)
    for (int64_t x = 0; x < len; ++x) {
      const double t = angles[b * angle_bstride + x];
      const double complex f = cos(t) - I * sin(t);
      ((double complex*)out)[b * len + x] = f * ((const double complex*)in)[b * len + x];
    }
}
