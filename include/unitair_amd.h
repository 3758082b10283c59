/*
 * unitair_amd.h -- C ABI of libunitair_amd.so, the MI355X (gfx950) gate
 * application library behind the `unitair_amd.simulation` Python surface.
 *
 * Plain pointers and sizes only.  Every pointer argument that names state,
 * operator, angle or result data is a device pointer (HBM) on the current
 * HIP device; `stream` is a hipStream_t (NULL = the legacy default stream).
 * No entry point synchronises, allocates device memory, or keeps a pointer
 * after it returns; every launch goes onto `stream` and can be captured into
 * a HIP graph.  Return value: UA_OK (0) or an error code, with a message in
 * ua_last_error() (thread-local).
 *
 * Data layout (unitair "vector layout", unitair/states/shapes.py:13-27,
 * unitair/states/conversions.py:5-92): a batch of states is `batch`
 * contiguous vectors of 2**num_qubits elements.  Complex elements are
 * interleaved (re, im) pairs, as torch.complex64/complex128 store them.
 * Qubit q is index bit num_qubits-1-q (qubit 0 = most significant bit).
 *
 * Operator layout (unitair/simulation/operations.py:45-148): a (2**k, 2**k)
 * row-major matrix acting as out = U @ in; for qubits=(q_1, ..., q_k) the
 * first listed qubit is the most significant bit of the gate index.
 */
#ifndef UNITAIR_AMD_H
#define UNITAIR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UA_ABI_VERSION 2

/* Element types (torch.float32, float64, complex64, complex128). */
enum {
  UA_FLOAT32 = 0,
  UA_FLOAT64 = 1,
  UA_COMPLEX64 = 2,
  UA_COMPLEX128 = 3
};

/* Status codes. */
enum {
  UA_OK = 0,
  UA_ERR_INVALID_ARGUMENT = 1, /* shapes, qubits, dtypes or sizes rejected */
  UA_ERR_UNSUPPORTED = 2,      /* valid request the library does not handle */
  UA_ERR_HIP = 3               /* a HIP runtime call or launch failed */
};

typedef void* ua_stream_t; /* hipStream_t */

/* ABI version of the loaded library (UA_ABI_VERSION at build time). */
int ua_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char* ua_last_error(void);

/* Launch-shape knobs (no effect on results; process-wide, not thread-safe):
 * UA_TUNE_UNROLL        amplitude groups each thread moves per launch, 1/2/4,
 *                       or 0 (default: 1 for 1-qubit, 2 for wider gates); used
 *                       when one operator serves the whole batch
 * UA_TUNE_LANE_EXCHANGE 1 (default): 1-qubit gates on index bits 1..6 move
 *                       contiguous amplitude pairs and exchange partners
 *                       across lanes; 0: strided pair access.
 * UA_TUNE_MFMA          1 (default): complex gates with one operator for the
 *                       batch run on the matrix cores (v_mfma_f32/f64_16x16x4)
 *                       when k = 5, when k = 4 in complex128, and when k = 4
 *                       in complex64 with every target on index bits 0..5;
 *                       0: vector units for every gate.
 * UA_TUNE_COMPILED_WAVES waves per SIMD that passes compiled afterwards by
 *                       ua_compile_fused_pass ask the register allocator for
 *                       (complex64; 0 = default 4, else 1..8); complex128
 *                       passes always ask for 2. */
enum { UA_TUNE_UNROLL = 0, UA_TUNE_LANE_EXCHANGE = 1, UA_TUNE_MFMA = 2,
       UA_TUNE_COMPILED_WAVES = 3 };
int ua_set_tuning(int key, int value);

/*
 * ua_apply_operator -- replaces unitair.simulation.apply_operator
 * (unitair/simulation/operations.py:45-148), i.e. its tensor-layout core
 * apply_operator_tensor (operations.py:151-186) and act_first_qubits_tensor
 * (operations.py:258-329); act_first_qubits (operations.py:236-255) is the
 * call with qubits = {0, ..., k-1}.
 *
 *   out[b] = (op[b] acting on `qubits`) state[b]      for b in [0, batch)
 *
 * out       : batch * 2**num_qubits elements, contiguous.  May equal `state`
 *             (in place) when state_batch_stride is 2**num_qubits or batch 1.
 * state     : input states; state b starts at state + b*state_batch_stride
 *             (stride 0: one state broadcast against a batch of operators,
 *             the expand() case of operations.py:304-309).
 * op        : operator b starts at op + b*op_batch_stride; stride 0 = one
 *             operator for the whole batch.  Same dtype as the state
 *             (operations.py:322 einsum requires it).
 * qubits    : op_num_qubits distinct qubit indices in [0, num_qubits), in the
 *             order of the gate index bits (first = most significant).
 * Strides are in elements.
 */
int ua_apply_operator(void* out, const void* state, const void* op, int dtype,
                      int num_qubits, int op_num_qubits, const int* qubits,
                      int64_t batch, int64_t state_batch_stride,
                      int64_t op_batch_stride, ua_stream_t stream);

/*
 * ua_apply_phase -- replaces unitair.simulation.apply_phase
 * (unitair/simulation/operations.py:15-42):
 *
 *   out = exp(-1j * angles) * state          (torch broadcasting)
 *
 * The output has `rank` dims of sizes `shape` and is written contiguously.
 * state_strides / angle_strides give, per output dim, the element stride of
 * each input (0 where that input is broadcast).  The phase factor is formed
 * in the precision of the angles (complex64 for float32 angles, as torch's
 * exp(-1j*angles) is), then both factors are promoted to out_dtype and
 * multiplied.  out_dtype must be complex; rank <= 8.
 */
int ua_apply_phase(void* out, int out_dtype, const void* state, int state_dtype,
                   const void* angles, int angle_dtype, int rank,
                   const int64_t* shape, const int64_t* state_strides,
                   const int64_t* angle_strides, ua_stream_t stream);

/*
 * ua_apply_all_qubits -- replaces unitair.simulation.apply_all_qubits
 * (unitair/simulation/operations.py:332-366) and apply_all_qubits_tensor
 * (operations.py:369-413): the same 2x2 operator on every qubit,
 * out[b] = (op[b] (x) ... (x) op[b]) state[b].  Strides as ua_apply_operator.
 */
int ua_apply_all_qubits(void* out, const void* state, const void* op, int dtype,
                        int num_qubits, int64_t batch,
                        int64_t state_batch_stride, int64_t op_batch_stride,
                        ua_stream_t stream);

/*
 * ua_permute_qubits -- replaces unitair.simulation.permute_qubits
 * (unitair/simulation/operations.py:600-623, permute_qubits_tensor :626-654),
 * and with the matching permutations swap (:506-537) and roll_qubits
 * (:540-597): output qubit q is input qubit permutation[q] (torch.permute on
 * the qubit dims of the tensor layout).  Out of place; any dtype.
 */
int ua_permute_qubits(void* out, const void* state, int dtype, int num_qubits,
                      const int* permutation, int64_t batch, int64_t state_batch_stride,
                      ua_stream_t stream);

/*
 * ua_apply_sign_gates -- replaces unitair.simulation.multi_cz
 * (operations.py:657-748) and multi_controlled_z (:751-783): out = s * state
 * with s(i) = (-1)^(number of pairs whose two index bits are set in i),
 * times -1 where every bit of all_mask is set when use_all is nonzero.
 * pair_bits holds num_pairs (bit, bit) pairs of INDEX bits (qubit q is bit
 * num_qubits-1-q).  Out of place; any dtype.
 */
int ua_apply_sign_gates(void* out, const void* state, int dtype, int num_qubits,
                        int num_pairs, const int* pair_bits, int use_all, uint64_t all_mask,
                        int64_t batch, int64_t state_batch_stride, ua_stream_t stream);

/*
 * ua_apply_controlled_flip -- replaces unitair.simulation.multi_controlled_x
 * (operations.py:786-806, H . multi_controlled_z . H on the target): out[i] =
 * state[i ^ 2^target_bit] where every bit of control_mask is set in i, else
 * state[i].  Exact (no Hadamard round trip).  Out of place; any dtype.
 */
int ua_apply_controlled_flip(void* out, const void* state, int dtype, int num_qubits,
                             uint64_t control_mask, int target_bit, int64_t batch,
                             int64_t state_batch_stride, ua_stream_t stream);

/*
 * ua_sample -- replaces the probability and draw steps of
 * unitair.simulation.measure (unitair/simulation/measurement.py:9-71:
 * abs_squared, then torch.distributions.Categorical(probs).sample): draws
 * num_samples outcomes x with probability |state[x]|^2 / sum |state|^2 into
 * out (int64).  Float64 prefix sums; uniforms from Philox4x32-10 with key
 * `seed` and counters offset .. offset + num_samples - 1 (num_samples below
 * 2^31).  workspace: ua_sample_workspace_bytes(num_qubits, num_samples)
 * bytes, 8-byte aligned.  One state.
 */
int64_t ua_sample_workspace_bytes(int num_qubits, int64_t num_samples);
/* Byte offset in that workspace of the float64 total sum |state|^2 once
 * ua_sample has run on the stream (-1 for invalid sizes).  A total that is
 * zero, negative or not finite is the case in which Categorical(probs)
 * raises ValueError (measurement.py:45-48); the draws are then meaningless
 * and the caller reports the error. */
int64_t ua_sample_total_offset(int num_qubits, int64_t num_samples);
int ua_sample(int64_t* out, const void* state, int dtype, int num_qubits, int64_t num_samples,
              uint64_t seed, uint64_t offset, void* workspace, ua_stream_t stream);

/*
 * ua_rand_state -- replaces unitair.initializations.rand_state on a device
 * (unitair/initializations.py:71-100: torch.randn, then divide each state by
 * its norm): batch states of 2^num_qubits Gaussian entries (real and
 * imaginary parts independent for complex dtypes), each normalised in
 * float64.  Philox4x32-10 keyed by `seed`, Box-Muller normals; the same seed
 * gives the same states.  2 <= num_qubits <= 40.  workspace:
 * ua_rand_state_workspace_bytes(num_qubits, batch, dtype) bytes.
 */
int64_t ua_rand_state_workspace_bytes(int num_qubits, int64_t batch, int dtype);
int ua_rand_state(void* out, int dtype, int num_qubits, int64_t batch, uint64_t seed,
                  void* workspace, ua_stream_t stream);

/*
 * ua_fill_state -- replaces uniform_superposition (kind 0: every entry
 * 2^(-num_qubits/2), initializations.py:103-128) and unit_vector (kind 1: 1 at
 * `index`, 0 elsewhere, initializations.py:46-68) on a device, for batch
 * states of 2^num_qubits entries.
 */
int ua_fill_state(void* out, int dtype, int num_qubits, int64_t batch, int kind, int64_t index,
                  ua_stream_t stream);

/*
 * ua_apply_1q_layer -- replaces unitair.simulation.apply_to_qubits
 * (unitair/simulation/operations.py:416-446) and apply_to_qubits_tensor
 * (operations.py:449-503) after its per-qubit fusion
 * (unitair/gates/matrix_algebra.py:6-51): one 2x2 operator on each of
 * num_targets distinct qubits.  ops holds num_targets row-major 2x2 matrices
 * per batch entry (target j at ops + 4 j); ops_batch_stride is 0 (shared) or
 * 4 * num_targets.  With num_targets = 0 the state is copied.  Out of place.
 */
int ua_apply_1q_layer(void* out, const void* state, const void* ops, int dtype,
                      int num_qubits, int num_targets, const int* qubits, int64_t batch,
                      int64_t state_batch_stride, int64_t ops_batch_stride,
                      ua_stream_t stream);

/*
 * Reductions over the last dimension with float64 accumulation; replace
 * unitair/states/innerprod.py.  `length` elements per row, `batch` rows.
 * Each input row r starts at ptr + r*row_stride (0 = broadcast).  The result
 * (batch values, contiguous) is written in out_dtype.  `workspace` must hold
 * ua_reduce_workspace_bytes(batch, length) bytes of device memory.
 */
int64_t ua_reduce_workspace_bytes(int64_t batch, int64_t length);

/* norm_squared (innerprod.py:29-46): sum_x |state_x|^2.  out_dtype real. */
int ua_norm_squared(void* out, int out_dtype, const void* state, int dtype,
                    int64_t batch, int64_t length, int64_t row_stride,
                    void* workspace, ua_stream_t stream);

/*
 * ua_norm_squared_noalloc -- ua_norm_squared with no workspace argument and no
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * The float64 partials live in a static device array of 2^16 (re, im) pairs
 * (1 MiB, held by every GPU the library runs on until the process ends), so
 * batch * ceil(length / 4096) must be at most 65536 (2^28 elements per call);
 * larger calls fail with UA_ERR_INVALID_ARGUMENT.  The array's address on the
 * current device is looked up on the first call on that device.  Calls on one
 * GPU share the array: calls that overlap on one GPU (different streams) race.
 */
int ua_norm_squared_noalloc(void* out, int out_dtype, const void* state, int dtype,
                            int64_t batch, int64_t length, int64_t row_stride,
                            ua_stream_t stream);

/* inner_product (innerprod.py:62-65): sum_x conj(a_x) b_x.  out_dtype
 * complex; a and b share `dtype`. */
int ua_inner_product(void* out, int out_dtype, const void* a, int64_t a_row_stride,
                     const void* b, int64_t b_row_stride, int dtype,
                     int64_t batch, int64_t length, void* workspace,
                     ua_stream_t stream);

/* diag_expectation_value (innerprod.py:49-59): sum_x |state_x|^2 diag_x.
 * diag is real (float32/float64) or complex; out_dtype follows torch's
 * promotion of abs_squared(state) * diag. */
int ua_diag_expectation(void* out, int out_dtype, const void* diag, int diag_dtype,
                        int64_t diag_row_stride, const void* state, int dtype,
                        int64_t state_row_stride, int64_t batch, int64_t length,
                        void* workspace, ua_stream_t stream);

/* abs_squared (innerprod.py:4-26): out_x = |state_x|^2, real, numel values. */
int ua_abs_squared(void* out, const void* state, int dtype, int64_t numel,
                   ua_stream_t stream);

/*
 * In-place circuit gates.  `batch` contiguous states of 2**num_qubits
 * elements are updated in place; only the amplitudes the gate changes are
 * read and written.  Results are bit-identical to ua_apply_operator with the
 * reference's matrices.
 *
 * ua_apply_cz   -- apply_operator(cz(), (qubit_a, qubit_b), state)
 *                  (unitair/gates/gates.py:195-213); also one pair of
 *                  unitair.simulation.multi_cz (operations.py:657-748).
 * ua_apply_cnot -- apply_operator(cnot(), (control, target), state)
 *                  (unitair/gates/gates.py:170-188, control first).
 */
int ua_apply_cz(void* state, int dtype, int num_qubits, int qubit_a, int qubit_b,
                int64_t batch, ua_stream_t stream);
int ua_apply_cnot(void* state, int dtype, int num_qubits, int control, int target,
                  int64_t batch, ua_stream_t stream);

/*
 * ua_fuse_1q_ops -- products of 1-qubit gates per qubit
 * (fuse_single_qubit_operators, unitair/gates/matrix_algebra.py:6-51):
 * out[b, s] = U_m ... U_2 U_1 over the operators i with slot[i] == s, in list
 * order (later operators on the left), for b < batch.  ops: num_ops device
 * pointers (a host array), each to row-major 2x2 matrices, op_batch_stride
 * elements apart per batch entry (0: one matrix for every entry).  out:
 * (batch, num_slots, 2, 2) contiguous -- the operand layout of ua_apply_1q_layer.
 * accumulate = 1 starts from out's current matrices instead of the identity, so
 * longer lists are fused UA_FUSE_MAX_OPS at a time.  Any of the four dtypes.
 */
#define UA_FUSE_MAX_OPS 64
int ua_fuse_1q_ops(void* out, const void* const* ops, int dtype, int num_ops,
                   const int* slot, int num_slots, int64_t batch, int64_t op_batch_stride,
                   int accumulate, ua_stream_t stream);

/*
 * ua_phase_angle_grad -- the angle gradient of apply_phase for real angles
 * (torch autograd through exp(-1j * angles) * state, operations.py:15-42):
 * grad[x] = sum_{r < rows} Im(out[r, x]) Re(g[r, x]) - Re(out[r, x]) Im(g[r, x]),
 * with out = apply_phase's result and g the incoming gradient, both (rows, cols)
 * contiguous complex64 / complex128; grad: cols float32 or float64 (grad_dtype).
 * rows = 1 is the elementwise case; angles broadcast over leading dims sum them.
 */
int ua_phase_angle_grad(void* grad, int grad_dtype, const void* g, const void* out, int dtype,
                        int64_t rows, int64_t cols, ua_stream_t stream);

/*
 * Circuits.  unitair has no circuit object: a circuit is the user's loop of
 * apply_operator / apply_phase calls (unitair README.rst:11-16).  ua_run_circuit
 * takes that loop as data and launches every gate from C onto `stream`.
 */
#define UA_MAX_GATE_QUBITS 5

enum {
  UA_GATE_DENSE = 0, /* apply_operator(data, qubits[:num_targets], state) */
  UA_GATE_PHASE = 1, /* apply_phase(data, state), data = 2**n angles     */
  UA_GATE_CZ = 2,    /* apply_operator(cz(), qubits[:2], state)          */
  UA_GATE_CNOT = 3   /* apply_operator(cnot(), qubits[:2], state)        */
};

typedef struct {
  int32_t kind;                        /* UA_GATE_* */
  int32_t num_targets;                 /* dense: k in 1..UA_MAX_GATE_QUBITS */
  int32_t qubits[UA_MAX_GATE_QUBITS];  /* unitair qubit order (first = gate MSB) */
  int32_t data_dtype;                  /* phase: dtype code of the angles */
  const void* data;                    /* device pointer: (2^k)x(2^k) matrix in the
                                          state dtype, or 2**n angles */
} ua_gate;

/*
 * ua_run_circuit -- apply gates[0..num_gates) in order to `batch` contiguous
 * states of 2**num_qubits elements.  work == NULL: in place on `state`.
 * Otherwise every gate reads one buffer and writes the other (the work of one
 * out-of-place apply_operator call per gate); *result_in_work (if not NULL)
 * tells which buffer holds the result.  events: NULL, or 2*num_gates
 * hipEvent_t handles recorded on `stream` before and after each gate; where
 * gate i's start handle equals gate i-1's end handle it is recorded once.
 */
int ua_run_circuit(void* state, void* work, int dtype, int num_qubits, int64_t batch,
                   const ua_gate* gates, int num_gates, void* const* events,
                   int* result_in_work, ua_stream_t stream);

/*
 * Dense-block passes.  unitair applies every gate as its own contraction over
 * the state (unitair/simulation/operations.py:151-186, 322; one
 * apply_operator call per gate).  ua_run_dense_blocks applies a gate list in
 * place like ua_run_circuit with work == NULL, but consecutive dense gates on
 * 4 or 5 qubits (complex64 / complex128 states of at least 14 qubits) whose
 * targets fit together in a 12-bit tile (their union plus index bits 0 and 1)
 * share one pass over the state on the matrix cores: up to 3 gates per pass
 * in complex64, 2 in complex128.  Every other gate is one launch, as in
 * ua_run_circuit.  Results equal gate-by-gate application to rounding.
 *
 * ua_plan_dense_blocks -- the launch units of that schedule, without running
 * it: *num_units units; unit_of_gate (NULL, or num_gates ints) receives each
 * gate's unit.  Validates the gate list like ua_run_dense_blocks.
 * ua_run_dense_blocks -- events: NULL, or 2 * num_units hipEvent_t handles
 * recorded on `stream` before and after each unit.
 */
int ua_plan_dense_blocks(int dtype, int num_qubits, const ua_gate* gates, int num_gates,
                         int* unit_of_gate, int* num_units);
int ua_run_dense_blocks(void* state, int dtype, int num_qubits, int64_t batch,
                        const ua_gate* gates, int num_gates, void* const* events,
                        ua_stream_t stream);

/*
 * Fused passes: several gates per read and write of the state.  The
 * reference's only fusion is multiplying 1-qubit gates that hit the same qubit
 * (unitair/gates/matrix_algebra.py:6-51, apply_to_qubits operations.py:
 * 416-503); a fused pass applies a whole run of gates to each 2^m-amplitude
 * tile of the state (m = tile_bits <= 12 "local" index bits) while it sits in
 * LDS.  In place, complex64 or complex128.
 *
 * A pass is cut into stages of two kinds.  In a cube stage each thread holds
 * the 16 amplitudes spanned by the stage's 4 tile bits (its "cube"); dense
 * gates address cube bits 0..3, element-wise gates address state index bits.
 * A diagonal stage (stage_bits {-1, -1, -1, -1}) holds only CZ, DIAG1, PHASE,
 * PHASE_FACTORS and CZ_LAYER ops and visits the amplitudes in the order of the state's
 * coalesced load, so per-element data (phase angles) is read contiguously.
 * PHASE and PHASE_FACTORS ops belong in diagonal stages.
 */
#define UA_FUSED_MAX_STAGES 32
#define UA_FUSED_MAX_TILE_BITS 16

enum {
  UA_FOP_DENSE1 = 0, /* a: cube bit; data: 2x2 matrix (state dtype)          */
  UA_FOP_DENSE2 = 1, /* a: cube bit of gate-index bit 1, b: of bit 0; 4x4     */
  UA_FOP_CNOT = 2,   /* a: control state index bit, b: target cube bit       */
  UA_FOP_CZ = 3,     /* a, b: state index bits                               */
  UA_FOP_DIAG1 = 4,  /* a: state index bit; data: (d0, d1), state dtype      */
  UA_FOP_PHASE = 5,  /* data: 2**n real angles, exp(-1j*angle) per element
                        (float32 for complex64 states, float64 for complex128) */
  UA_FOP_CZ_LAYER = 6, /* data: uint64 mask(s); bit j set = CZ between index bits
                        j and j+1 (qubits n-2-j and n-1-j); data_bstride in
                        masks (1 = one mask per state) */
  UA_FOP_LAYER1Q = 7, /* a: four 8-bit counts c0..c3 (c_k in bits 8k..8k+7); the
                        next c0+c1+c2+c3 ops are UA_FOP_DENSE1 on cube bit 0
                        (c0 of them, in order), then cube bit 1, 2, 3: 1-qubit
                        gates on distinct bits commute, so a run of them is
                        applied grouped by bit with no per-op dispatch */
  UA_FOP_PHASE_FACTORS = 8 /* data: 2**n complex factors (state dtype), element x
                        multiplied by data[x]: apply_phase with exp(-1j*angle)
                        formed once (aligned to two elements, data_bstride even) */
};

typedef struct {
  int32_t kind; /* UA_FOP_* */
  int32_t a, b;
  int32_t cache_slot;   /* unused (kept for the record layout); set -1 */
  const void* data;     /* device pointer; read-only while the pass runs (gate data
                           is read through scalar loads) */
  int64_t data_bstride; /* elements between the data of consecutive states; PHASE
                           angles must be aligned to two angles and data_bstride
                           even */
} ua_fused_op;

typedef struct {
  int32_t tile_bits;           /* m <= num_qubits: 4..12 for ua_run_fused_pass; compiled
                                  passes 12..14 (complex64) or 12..13 (complex128) */
  int32_t local_bits[UA_FUSED_MAX_TILE_BITS]; /* state bit of tile bit i (i < m):
                                  ascending, [0..4) = 0,1,2,3 */
  int32_t num_stages;          /* <= UA_FUSED_MAX_STAGES */
  int32_t needs_index;         /* unused (derived from the ops); kept for the layout */
  int32_t stage_bits[UA_FUSED_MAX_STAGES][4]; /* ascending tile bits of each cube, or
                                                 all -1 for a diagonal stage */
  int32_t stage_ops[UA_FUSED_MAX_STAGES + 1]; /* stage s runs ops [stage_ops[s], stage_ops[s+1]) */
} ua_fused_pass;

/* Run one fused pass on `batch` contiguous states.  `ops` is the device copy
 * of the pass's ops; `host_ops` is the same array in host memory (required
 * when num_stages > 0): it is validated before launch and selects the kernel
 * variant (1-qubit layers only / no DENSE2 / every kind), so it must be an
 * exact copy of ops. */
int ua_run_fused_pass(void* state, int dtype, int num_qubits, int64_t batch,
                      const ua_fused_pass* pass, const ua_fused_op* ops,
                      const ua_fused_op* host_ops, ua_stream_t stream);

/*
 * Compiled fused passes.  No counterpart in the reference (each of its gates
 * is a pass of its own, unitair/simulation/operations.py:151-186).
 * ua_run_fused_pass interprets a pass's op list inside one kernel;
 * ua_compile_fused_pass writes a kernel for one pass structure (tile bits,
 * stages, op kinds and the bits each op acts on, all constants) and compiles
 * it with hiprtc for gfx950; ua_run_compiled_pass launches it.  Gate data is
 * not compiled in: the kernel reads it through `ops` at run time, so a handle
 * serves every op array with the same kinds, bits and stage split as the
 * host_ops it was compiled from.  Handles are cached by structure and stay
 * valid until the process ends (each GPU loads the code the first time the
 * pass runs there).  Compilation takes of the order of a second and runs on
 * the host only; calls on different threads compile in parallel.
 * Tiles of 12..14 bits for complex64 (a 2^14-amplitude tile is 128 KiB of
 * LDS, 1024 threads) and 12..13 for complex128 (UA_ERR_UNSUPPORTED
 * otherwise): larger tiles hold more gates per pass, so a circuit needs
 * fewer passes.  Every other argument check of ua_run_fused_pass applies.
 */
int ua_compile_fused_pass(int dtype, int num_qubits, const ua_fused_pass* pass,
                          const ua_fused_op* host_ops, void** handle);
/* Loads a compiled pass's code on the current GPU (what the first
 * ua_run_compiled_pass there does otherwise: call it before capturing that
 * call into a HIP graph).  Allocates the code object's device memory. */
int ua_load_compiled_pass(void* handle);
/* Runs a compiled pass on `batch` contiguous states of the dtype and qubit
 * count it was compiled for; `ops` is a device op array of the compiled
 * structure. */
int ua_run_compiled_pass(void* handle, void* state, int64_t batch, const ua_fused_op* ops,
                         ua_stream_t stream);
/* The generated source of a pass's kernel (for inspection): *length gets its
 * size in bytes; up to capacity - 1 bytes and a terminating NUL go to buf. */
int ua_fused_pass_source(int dtype, int num_qubits, const ua_fused_pass* pass,
                         const ua_fused_op* host_ops, char* buf, int64_t capacity,
                         int64_t* length);

/*
 * Differentiation.  unitair relies on torch autograd through its op sequence
 * (operations.py:41-42,322) and keeps one saved state per gate.
 *
 * ua_operator_gradient -- backward of apply_operator w.r.t. the operator:
 *   grad_op[o][r][c] = sum over batch entries b using operator o and over the
 *   2^(n-k) groups:  grad_out[b][.. r ..] * conj(state[b][.. c ..])
 * op_count is 1 (one operator for the whole batch) or batch.  grad_out is
 * contiguous (batch, 2^n); state rows start at state + b*state_batch_stride.
 * k <= UA_MAX_GRAD_QUBITS.  The gradient w.r.t. the state is
 * ua_apply_operator with the conjugate-transposed operator.  The reduction
 * over groups is split across many workgroups; the float64 partial sums go to
 * `workspace` (ua_operator_gradient_workspace_bytes, 8-byte aligned; 0 bytes,
 * and NULL allowed, when one workgroup per result block suffices).
 */
#define UA_MAX_GRAD_QUBITS 16
int64_t ua_operator_gradient_workspace_bytes(int num_qubits, int op_num_qubits, int64_t batch,
                                             int64_t op_count);
int ua_operator_gradient(void* grad_op, const void* grad_out, const void* state, int dtype,
                         int num_qubits, int op_num_qubits, const int* qubits,
                         int64_t batch, int64_t op_count, int64_t state_batch_stride,
                         void* workspace, ua_stream_t stream);

/*
 * Adjoint differentiation of Pauli-rotation circuits for H = sum_i Z_i (the
 * loss unitair computes as diag_expectation_value(n - 2*popcount(x), psi),
 * innerprod.py:49-59).  No reference counterpart beyond that loss.
 *
 * ua_expect_sum_z   out[b] = sum_x |psi_b(x)|^2 (n - 2 popcount x)   (float64)
 * ua_apply_sum_z    out = H psi (the co-state at the end of the circuit)
 * ua_pauli_gradients out[b*n + q] = Im <lambda_b| P_q |psi_b> for every qubit
 *                   q, P = X, Y or Z: the derivative of <psi|H|psi> w.r.t. the
 *                   angle t of a rotation exp(-i t/2 P_q) sitting where (lambda,
 *                   psi) were taken (twice that for unitair's exp_x/exp_y/exp_z(t)
 *                   = exp(-i t P), gates.py:7-91).
 * Workspace: ua_pauli_gradients_workspace(num_qubits, batch) bytes (both
 * ua_expect_sum_z and ua_pauli_gradients).  complex64/complex128 states.
 */
enum { UA_PAULI_X = 0, UA_PAULI_Y = 1, UA_PAULI_Z = 2 };
int64_t ua_pauli_gradients_workspace(int num_qubits, int64_t batch);
int ua_expect_sum_z(double* out, const void* psi, int dtype, int num_qubits, int64_t batch,
                    void* workspace, ua_stream_t stream);
int ua_apply_sum_z(void* out, const void* psi, int dtype, int num_qubits, int64_t batch,
                   ua_stream_t stream);
int ua_pauli_gradients(double* out, const void* lambda, const void* psi, int dtype,
                       int num_qubits, int64_t batch, int pauli, void* workspace,
                       ua_stream_t stream);

/* ua_apply_diagonal  out = d * psi elementwise for a real diagonal d (float32
 *                    or float64; diag_batch_stride 0 = one row for every state,
 *                    else elements between rows): the co-state H psi of a
 *                    diagonal observable (the loss diag_expectation_value(d,
 *                    psi), innerprod.py:49-59). */
int ua_apply_diagonal(void* out, const void* psi, int dtype, const void* diag, int diag_dtype,
                      int64_t diag_batch_stride, int num_qubits, int64_t batch,
                      ua_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* UNITAIR_AMD_H */
