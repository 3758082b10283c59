"""Gate lists and their execution on HBM-resident states.

unitair has no circuit object: users loop over apply_operator / apply_phase
calls (unitair README.rst:11-16).  A `Circuit` is that loop written down as
data, so it can run without per-gate allocation: gate matrices and phase
angles are uploaded once, states stay in HBM, and each gate is one launch of
the C ABI on torch's current stream (hence capturable in a CUDA/HIP graph).

Two execution forms:
  * out of place (`in_place=False`): every gate reads one buffer and writes
    the other, exactly the work of one `apply_operator` call per gate;
  * in place (default): dense gates update their amplitude groups in place,
    CZ/CNOT touch only the amplitudes they change, phases update in place.
Both produce the same values (CZ/CNOT are exact either way).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib

KINDS = ("dense", "phase", "cz", "cnot")


@dataclass
class Gate:
    kind: str                               # one of KINDS
    qubits: Tuple[int, ...] = ()
    matrix: Optional[np.ndarray] = None     # dense: (2^k, 2^k) complex, row = output
    angles: Optional[np.ndarray] = None     # phase: (2^n,) real, exp(-1j*angles)

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError(f"unknown gate kind {self.kind!r}")
        self.qubits = tuple(int(q) for q in self.qubits)

    def dense_matrix(self) -> np.ndarray:
        """The gate as a (2^k, 2^k) complex128 matrix (dense/cz/cnot)."""
        if self.kind == "dense":
            return np.asarray(self.matrix, dtype=np.complex128)
        if self.kind == "cz":
            return np.diag([1, 1, 1, -1]).astype(np.complex128)
        if self.kind == "cnot":
            return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]],
                            dtype=np.complex128)
        raise ValueError("phase gates have no small matrix")

    def amplitude_updates(self, num_amplitudes: int) -> int:
        return num_amplitudes


@dataclass
class Circuit:
    num_qubits: int
    gates: List[Gate] = field(default_factory=list)

    def __len__(self):
        return len(self.gates)

    def counts(self):
        c = {}
        for g in self.gates:
            key = g.kind if g.kind != "dense" else f"dense{len(g.qubits)}"
            c[key] = c.get(key, 0) + 1
        return c


_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64,
         torch.float32: torch.float32, torch.float64: torch.float64}


class CircuitRunner:
    """Runs a Circuit on a (batch, 2^n) state tensor on one GPU."""

    def __init__(self, circuit: Circuit, dtype: torch.dtype = torch.complex64,
                 device="cuda", in_place: bool = True):
        self.lib = _lib.load()
        self.circuit = circuit
        self.n = circuit.num_qubits
        self.dtype = dtype
        self.code = _lib.DTYPE_CODES[dtype]
        self.device = torch.device(device)
        self.in_place = in_place
        self._buffers = []          # keeps uploaded operands alive
        self._ops = []
        real = _REAL[dtype]
        for g in circuit.gates:
            if g.kind == "dense":
                m = torch.as_tensor(np.ascontiguousarray(g.matrix)).to(self.device, dtype)
                self._buffers.append(m)
                self._ops.append(("dense", g.qubits, m))
            elif g.kind == "phase":
                a = torch.as_tensor(np.ascontiguousarray(g.angles)).to(self.device, real)
                self._buffers.append(a)
                self._ops.append(("phase", g.qubits, a))
            else:
                self._ops.append((g.kind, g.qubits, None))
        self._work = None

    def _check_state(self, state: torch.Tensor):
        if state.device != self.device or state.dtype != self.dtype:
            raise ValueError(f"state must be {self.dtype} on {self.device}")
        if state.shape[-1] != 2 ** self.n or not state.is_contiguous():
            raise ValueError("state must be contiguous (batch, 2**n)")

    def _launch(self, op, src: torch.Tensor, dst: torch.Tensor, batch: int, stream):
        kind, qubits, t = op
        lib, n, code = self.lib, self.n, self.code
        if kind == "dense":
            rc = lib.ua_apply_operator(_lib.ptr(dst), _lib.ptr(src), _lib.ptr(t), code, n,
                                       len(qubits), _lib.int_array(qubits), batch,
                                       2 ** n, 0, stream)
        elif kind == "phase":
            N = 2 ** n
            shape = _lib.i64_array([batch, N])
            rc = lib.ua_apply_phase(_lib.ptr(dst), code, _lib.ptr(src), code, _lib.ptr(t),
                                    _lib.DTYPE_CODES[t.dtype], 2, shape,
                                    _lib.i64_array([N, 1]), _lib.i64_array([0, 1]), stream)
        else:
            if dst.data_ptr() != src.data_ptr():
                dst.copy_(src)
            fn = lib.ua_apply_cz if kind == "cz" else lib.ua_apply_cnot
            rc = fn(_lib.ptr(dst), code, n, qubits[0], qubits[1], batch, stream)
        if rc:
            _lib.check(rc, f"circuit gate {kind}{qubits}")

    def run(self, state: torch.Tensor, gate_events: Optional[list] = None) -> torch.Tensor:
        """Apply every gate.  In place: `state` is updated and returned.  Out of
        place: `state` is left unchanged and the result is returned (a work
        buffer owned by the runner is reused between calls).

        gate_events, if given, receives one (start, end) pair of timing events
        per gate, recorded on the current stream around each launch."""
        self._check_state(state)
        batch = state.numel() // (2 ** self.n)
        stream = _lib.stream_handle(self.device)
        cur = torch.cuda.current_stream(self.device)
        if self.in_place:
            bufs = [state, state]
        else:
            if self._work is None or self._work.shape != state.shape:
                self._work = [torch.empty_like(state), torch.empty_like(state)]
            bufs = self._work
        src = state
        for i, op in enumerate(self._ops):
            dst = bufs[i % 2]
            if gate_events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(cur)
            self._launch(op, src, dst, batch, stream)
            if gate_events is not None:
                e1.record(cur)
                gate_events.append((e0, e1))
            src = dst
        return src
