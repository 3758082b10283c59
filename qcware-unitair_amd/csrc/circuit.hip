// Native circuit executor: ua_run_circuit.
//
// unitair has no circuit object; a circuit is the user's Python loop of
// apply_operator / apply_phase calls (unitair README.rst:11-16,
// unitair/simulation/operations.py:15-148).  That loop pays Python and launch
// overhead per gate, which at 2^24 amplitudes is the same order as the gate
// itself.  Here the whole gate list is handed over once and every gate is
// launched from C onto one stream, so the GPU runs back to back; the same
// call can be captured into a HIP graph and replayed.
#include "ua_common.h"

namespace ua {
int apply_operator_impl(void* out, const void* state, const void* op, int dtype, int n,
                        int k, const int* qubits, int64_t batch, int64_t state_bstride,
                        int64_t op_bstride, hipStream_t s);
int apply_phase_impl(void* out, int odt, const void* state, int sdt, const void* angles,
                     int adt, int rank, const int64_t* shape, const int64_t* sstr,
                     const int64_t* astr, hipStream_t s);
int apply_special_impl(bool cnot, void* state, int dtype, int n, int qa, int qb,
                       int64_t batch, hipStream_t s);
}  // namespace ua

namespace ua {

// One gate of a ua_gate list: reads `cur`, writes `dst` (the same buffer for
// in-place application).  Shared by ua_run_circuit and ua_run_dense_blocks.
int run_gate(const ua_gate& g, int index, void* dst, void* cur, int dtype, int n,
             int64_t batch, hipStream_t s) {
  const int64_t N = int64_t(1) << n;
  int rc = UA_OK;
  switch (g.kind) {
    case UA_GATE_DENSE:
      UA_CHECK_ARG(g.num_targets >= 1 && g.num_targets <= UA_MAX_GATE_QUBITS,
                   "gate %d: %d targets", index, g.num_targets);
      rc = apply_operator_impl(dst, cur, g.data, dtype, n, g.num_targets, g.qubits, batch, N,
                               0, s);
      break;
    case UA_GATE_PHASE: {
      const int64_t shape[2] = {batch, N}, sstr[2] = {N, 1}, astr[2] = {0, 1};
      rc = apply_phase_impl(dst, dtype, cur, dtype, g.data, g.data_dtype, 2, shape, sstr, astr,
                            s);
      break;
    }
    case UA_GATE_CZ:
    case UA_GATE_CNOT:
      if (dst != cur) {
        const hipError_t e = hipMemcpyAsync(dst, cur, size_t(batch) * N * elem_size(dtype),
                                            hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
          set_error("gate %d: copy failed: %s", index, hipGetErrorString(e));
          return UA_ERR_HIP;
        }
      }
      rc = apply_special_impl(g.kind == UA_GATE_CNOT, dst, dtype, n, g.qubits[0], g.qubits[1],
                              batch, s);
      break;
    default:
      set_error("gate %d: unknown kind %d", index, g.kind);
      return UA_ERR_INVALID_ARGUMENT;
  }
  if (rc != UA_OK) {
    char msg[1024];
    snprintf(msg, sizeof(msg), "gate %d: %s", index, get_error());
    set_error("%s", msg);
  }
  return rc;
}

int record_event(void* ev, hipStream_t s) {
  const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(ev), s);
  if (e != hipSuccess) {
    set_error("hipEventRecord: %s", hipGetErrorString(e));
    return UA_ERR_HIP;
  }
  return UA_OK;
}

}  // namespace ua

extern "C" int ua_run_circuit(void* state, void* work, int dtype, int num_qubits,
                              int64_t batch, const ua_gate* gates, int num_gates,
                              void* const* events, int* result_in_work,
                              ua_stream_t stream) {
  using namespace ua;
  hipStream_t s = static_cast<hipStream_t>(stream);
  UA_CHECK_ARG(elem_size(dtype) != 0, "unsupported dtype code %d", dtype);
  UA_CHECK_ARG(num_qubits >= 1 && num_qubits <= 62, "num_qubits=%d out of range", num_qubits);
  UA_CHECK_ARG(num_gates >= 0 && (num_gates == 0 || gates), "bad gate list");
  UA_CHECK_ARG(batch >= 0, "negative batch");
  void* cur = state;  // buffer holding the current state
  for (int i = 0; i < num_gates; ++i) {
    void* dst = (work == nullptr) ? cur : (cur == state ? work : state);
    // a start event that is the previous gate's end event is already recorded
    if (events && !(i > 0 && events[2 * i] == events[2 * i - 1])) {
      const int rc = record_event(events[2 * i], s);
      if (rc) return rc;
    }
    const int rc = run_gate(gates[i], i, dst, cur, dtype, num_qubits, batch, s);
    if (rc != UA_OK) return rc;
    if (events) {
      const int rc2 = record_event(events[2 * i + 1], s);
      if (rc2) return rc2;
    }
    cur = dst;
  }
  if (result_in_work) *result_in_work = (cur != state) ? 1 : 0;
  return UA_OK;
}
