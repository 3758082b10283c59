"""ctypes wrapper of liboracle_cf.so (closed_form.c).  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path
from typing import Sequence

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB_PATH = _DIR / "liboracle_cf.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_DIR)], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        lib.cf_apply_dense_c128.restype = ctypes.c_int
        lib.cf_apply_dense_c128.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(ctypes.c_int), ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        lib.cf_apply_phase_c128.restype = None
        lib.cf_apply_phase_c128.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
            ctypes.c_int64, ctypes.c_int64]
        _lib = lib
    return _lib


def apply_dense(op: np.ndarray, qubits: Sequence[int], state: np.ndarray,
                out: np.ndarray = None) -> np.ndarray:
    """out[b] = (op[b] on qubits) state[b], complex128.  op: (D, D) or
    (batch, D, D); state: (2**n,) or (batch, 2**n).  out may be state."""
    state = np.ascontiguousarray(state, dtype=np.complex128)
    op = np.ascontiguousarray(op, dtype=np.complex128)
    N = state.shape[-1]
    n = N.bit_length() - 1
    k = len(qubits)
    batch = state.size // N
    if op.ndim == 3 and state.ndim == 1:
        batch = op.shape[0]
        sstride = 0
    else:
        sstride = N
    if out is None:
        out = np.empty((batch, N) if (op.ndim == 3 or state.ndim == 2) else (N,),
                       dtype=np.complex128)
    ostride = op.shape[-1] ** 2 if op.ndim == 3 else 0
    q = (ctypes.c_int * max(1, k))(*qubits)
    rc = _load().cf_apply_dense_c128(out.ctypes.data, state.ctypes.data, op.ctypes.data,
                                     n, k, q, batch, sstride, ostride)
    if rc != 0:
        raise ValueError("cf_apply_dense_c128 rejected its arguments")
    return out


def apply_phase(angles: np.ndarray, state: np.ndarray, out: np.ndarray = None) -> np.ndarray:
    """exp(-1j*angles) * state for float64 angles of shape (N,) or state's shape."""
    state = np.ascontiguousarray(state, dtype=np.complex128)
    angles = np.ascontiguousarray(angles, dtype=np.float64)
    N = state.shape[-1]
    batch = state.size // N
    astride = 0 if angles.size == N else N
    if out is None:
        out = np.empty_like(state)
    _load().cf_apply_phase_c128(out.ctypes.data, state.ctypes.data, angles.ctypes.data,
                                N, batch, astride)
    return out
