"""ctypes binding of libunitair_amd.so (declarations: include/unitair_amd.h).

torch is imported before the library is opened so that the HIP runtime
already loaded by torch (soname libamdhip64.so.7) is the one the library
binds to: device pointers and streams handed over from torch are then valid
inside the library.  There is no fallback: if the library is missing or
fails to load, every GPU entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (must precede the library load, see above)

LIB_ENV = "UNITAIR_AMD_LIB"
_DEFAULT_PATH = Path(__file__).resolve().parent / "libunitair_amd.so"

UA_OK = 0
UA_ERR_INVALID_ARGUMENT = 1
UA_ERR_UNSUPPORTED = 2
UA_ERR_HIP = 3
ABI_VERSION = 2

UA_FLOAT32 = 0
UA_FLOAT64 = 1
UA_COMPLEX64 = 2
UA_COMPLEX128 = 3

DTYPE_CODES = {
    torch.float32: UA_FLOAT32,
    torch.float64: UA_FLOAT64,
    torch.complex64: UA_COMPLEX64,
    torch.complex128: UA_COMPLEX128,
}

_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_p = ctypes.c_void_p
_pi32 = ctypes.POINTER(ctypes.c_int)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); mirrors include/unitair_amd.h
SIGNATURES = {
    "ua_abi_version": (_i32, []),
    "ua_last_error": (ctypes.c_char_p, []),
    "ua_set_tuning": (_i32, [_i32, _i32]),
    "ua_apply_operator": (_i32, [_p, _p, _p, _i32, _i32, _i32, _pi32, _i64, _i64, _i64, _p]),
    "ua_apply_phase": (_i32, [_p, _i32, _p, _i32, _p, _i32, _i32, _pi64, _pi64, _pi64, _p]),
    "ua_apply_all_qubits": (_i32, [_p, _p, _p, _i32, _i32, _i64, _i64, _i64, _p]),
    "ua_apply_1q_layer": (_i32, [_p, _p, _p, _i32, _i32, _i32, _pi32, _i64, _i64, _i64, _p]),
    "ua_rand_state_workspace_bytes": (_i64, [_i32, _i64, _i32]),
    "ua_rand_state": (_i32, [_p, _i32, _i32, _i64, _u64, _p, _p]),
    "ua_fill_state": (_i32, [_p, _i32, _i32, _i64, _i32, _i64, _p]),
    "ua_sample_workspace_bytes": (_i64, [_i32, _i64]),
    "ua_sample_total_offset": (_i64, [_i32, _i64]),
    "ua_sample": (_i32, [_p, _p, _i32, _i32, _i64, _u64, _u64, _p, _p]),
    "ua_permute_qubits": (_i32, [_p, _p, _i32, _i32, _pi32, _i64, _i64, _p]),
    "ua_apply_sign_gates": (_i32, [_p, _p, _i32, _i32, _i32, _pi32, _i32, _u64, _i64, _i64, _p]),
    "ua_apply_controlled_flip": (_i32, [_p, _p, _i32, _i32, _u64, _i32, _i64, _i64, _p]),
    "ua_reduce_workspace_bytes": (_i64, [_i64, _i64]),
    "ua_norm_squared": (_i32, [_p, _i32, _p, _i32, _i64, _i64, _i64, _p, _p]),
    "ua_norm_squared_noalloc": (_i32, [_p, _i32, _p, _i32, _i64, _i64, _i64, _p]),
    "ua_inner_product": (_i32, [_p, _i32, _p, _i64, _p, _i64, _i32, _i64, _i64, _p, _p]),
    "ua_diag_expectation": (_i32, [_p, _i32, _p, _i32, _i64, _p, _i32, _i64, _i64, _i64, _p, _p]),
    "ua_abs_squared": (_i32, [_p, _p, _i32, _i64, _p]),
    "ua_operator_gradient_workspace_bytes": (_i64, [_i32, _i32, _i64, _i64]),
    "ua_operator_gradient": (_i32, [_p, _p, _p, _i32, _i32, _i32, _pi32, _i64, _i64, _i64, _p,
                                    _p]),
    "ua_pauli_gradients_workspace": (_i64, [_i32, _i64]),
    "ua_expect_sum_z": (_i32, [_p, _p, _i32, _i32, _i64, _p, _p]),
    "ua_apply_sum_z": (_i32, [_p, _p, _i32, _i32, _i64, _p]),
    "ua_pauli_gradients": (_i32, [_p, _p, _p, _i32, _i32, _i64, _i32, _p, _p]),
    "ua_apply_diagonal": (_i32, [_p, _p, _i32, _p, _i32, _i64, _i32, _i64, _p]),
    "ua_apply_cz": (_i32, [_p, _i32, _i32, _i32, _i32, _i64, _p]),
    "ua_apply_cnot": (_i32, [_p, _i32, _i32, _i32, _i32, _i64, _p]),
}

UA_TUNE_UNROLL = 0
UA_TUNE_LANE_EXCHANGE = 1
UA_TUNE_MFMA = 2
UA_TUNE_COMPILED_WAVES = 3

UA_PAULI_X = 0
UA_PAULI_Y = 1
UA_PAULI_Z = 2
UA_MAX_GRAD_QUBITS = 16

UA_GATE_DENSE = 0
UA_GATE_PHASE = 1
UA_GATE_CZ = 2
UA_GATE_CNOT = 3
UA_MAX_GATE_QUBITS = 5


class Gate(ctypes.Structure):
    """ua_gate (include/unitair_amd.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("num_targets", ctypes.c_int32),
                ("qubits", ctypes.c_int32 * UA_MAX_GATE_QUBITS),
                ("data_dtype", ctypes.c_int32), ("data", ctypes.c_void_p)]


UA_FUSED_MAX_STAGES = 32
UA_FUSED_MAX_TILE_BITS = 16


class FusedOp(ctypes.Structure):
    """ua_fused_op (include/unitair_amd.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("a", ctypes.c_int32), ("b", ctypes.c_int32),
                ("cache_slot", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("data_bstride", ctypes.c_int64)]


class FusedPass(ctypes.Structure):
    """ua_fused_pass (include/unitair_amd.h)."""
    _fields_ = [("tile_bits", ctypes.c_int32),
                ("local_bits", ctypes.c_int32 * UA_FUSED_MAX_TILE_BITS),
                ("num_stages", ctypes.c_int32), ("needs_index", ctypes.c_int32),
                ("stage_bits", (ctypes.c_int32 * 4) * UA_FUSED_MAX_STAGES),
                ("stage_ops", ctypes.c_int32 * (UA_FUSED_MAX_STAGES + 1))]


SIGNATURES["ua_run_fused_pass"] = (_i32, [_p, _i32, _i32, _i64, ctypes.POINTER(FusedPass),
                                          _p, _p, _p])
SIGNATURES["ua_run_circuit"] = (_i32, [_p, _p, _i32, _i32, _i64, ctypes.POINTER(Gate), _i32,
                                       ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int), _p])
SIGNATURES["ua_compile_fused_pass"] = (_i32, [_i32, _i32, ctypes.POINTER(FusedPass), _p,
                                              ctypes.POINTER(ctypes.c_void_p)])
SIGNATURES["ua_run_compiled_pass"] = (_i32, [_p, _p, _i64, _p, _p])
SIGNATURES["ua_load_compiled_pass"] = (_i32, [_p])
SIGNATURES["ua_fused_pass_source"] = (_i32, [_i32, _i32, ctypes.POINTER(FusedPass), _p,
                                             ctypes.c_char_p, _i64, _pi64])
UA_FUSE_MAX_OPS = 64
SIGNATURES["ua_fuse_1q_ops"] = (_i32, [_p, ctypes.POINTER(ctypes.c_void_p), _i32, _i32,
                                       ctypes.POINTER(ctypes.c_int), _i32, _i64, _i64, _i32, _p])
SIGNATURES["ua_phase_angle_grad"] = (_i32, [_p, _i32, _p, _p, _i32, _i64, _i64, _p])
SIGNATURES["ua_plan_dense_blocks"] = (_i32, [_i32, _i32, ctypes.POINTER(Gate), _i32,
                                             ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)])
SIGNATURES["ua_run_dense_blocks"] = (_i32, [_p, _i32, _i32, _i64, ctypes.POINTER(Gate), _i32,
                                            ctypes.POINTER(ctypes.c_void_p), _p])

_lib = None


class UnitairAmdError(RuntimeError):
    """A call into libunitair_amd.so failed."""


def library_path() -> Path:
    return Path(os.environ.get(LIB_ENV, str(_DEFAULT_PATH)))


def load() -> ctypes.CDLL:
    """Open the library once and bind every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise ImportError(
            f"{path} not found: build it with `python -m unitair_amd._build` "
            f"(or __graft_entry__.build()); unitair_amd has no non-HIP fallback")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.ua_abi_version() != ABI_VERSION:
        raise ImportError(f"{path}: ABI version {lib.ua_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    """Raise the Python exception matching a library status code."""
    if rc == UA_OK:
        return
    msg = load().ua_last_error().decode(errors="replace")
    if rc == UA_ERR_INVALID_ARGUMENT:
        raise ValueError(f"{what}: {msg}")
    if rc == UA_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise UnitairAmdError(f"{what}: {msg}")


def stream_handle(device: torch.device) -> ctypes.c_void_p:
    """hipStream_t of torch's current stream on `device`."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def int_array(values) -> ctypes.Array:
    values = list(values)
    return (ctypes.c_int * max(1, len(values)))(*values)


def i64_array(values) -> ctypes.Array:
    values = list(values)
    return (ctypes.c_int64 * max(1, len(values)))(*values)
