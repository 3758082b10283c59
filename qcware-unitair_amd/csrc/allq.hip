// Same 2x2 operator on every qubit: ua_apply_all_qubits
// (unitair/simulation/operations.py:332-413).
//
// The reference loops over qubits with swap (a transposed view) + einsum and
// finishes with a roll and a contiguous copy.  Here: the first pass reads the
// input and writes `out`, later passes run in place on `out`.
#include "ua_common.h"

namespace ua {
int apply_operator_impl(void* out, const void* state, const void* op, int dtype, int n,
                        int k, const int* qubits, int64_t batch, int64_t state_bstride,
                        int64_t op_bstride, hipStream_t s);
}

extern "C" int ua_apply_all_qubits(void* out, const void* state, const void* op, int dtype,
                                   int num_qubits, int64_t batch,
                                   int64_t state_batch_stride, int64_t op_batch_stride,
                                   ua_stream_t stream) {
  using namespace ua;
  UA_CHECK_ARG(num_qubits >= 1, "apply_all_qubits needs at least one qubit");
  UA_CHECK_ARG(out != state, "apply_all_qubits is out of place");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t N = int64_t(1) << num_qubits;
  for (int q = 0; q < num_qubits; ++q) {
    const int qubit[1] = {q};
    const void* src = (q == 0) ? state : out;
    const int64_t sstride = (q == 0) ? state_batch_stride : N;
    const int rc = apply_operator_impl(out, src, op, dtype, num_qubits, 1, qubit, batch,
                                       sstride, op_batch_stride, s);
    if (rc) return rc;
  }
  return UA_OK;
}
