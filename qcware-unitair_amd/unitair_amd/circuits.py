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
    """Runs a Circuit on a (batch, 2^n) state tensor on one GPU.

    The gate list is converted once into the C ABI's ua_gate array (device
    pointers to the uploaded matrices/angles); run() is then a single call of
    ua_run_circuit, which launches every gate from C on torch's current
    stream.  capture() records that call into a CUDA/HIP graph for replay.

    dense_blocks=True (in place only) runs the list through
    ua_run_dense_blocks instead: consecutive 4/5-qubit dense gates that fit
    one 12-bit tile share a matrix-core pass over the state.  `kinds` then
    names the launch units ("blocks" + the gates' sizes for a pass)."""

    def __init__(self, circuit: Circuit, dtype: torch.dtype = torch.complex64,
                 device="cuda", in_place: bool = True, dense_blocks: bool = False):
        self.lib = _lib.load()
        self.circuit = circuit
        self.n = circuit.num_qubits
        self.dtype = dtype
        self.code = _lib.DTYPE_CODES[dtype]
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.in_place = in_place
        self._buffers = []          # keeps uploaded operands alive
        real = _REAL[dtype]
        gates = (_lib.Gate * max(1, len(circuit.gates)))()
        kinds = []
        # all dense matrices go to the device in one copy
        dense = [np.asarray(g.matrix, dtype=np.complex128).reshape(-1)
                 for g in circuit.gates if g.kind == "dense"]
        pool = None
        if dense:
            pool = torch.from_numpy(np.concatenate(dense)).to(self.device, dtype)
            self._buffers.append(pool)
        esize = pool.element_size() if pool is not None else 0
        offset = 0
        for i, g in enumerate(circuit.gates):
            c = gates[i]
            if g.kind == "dense":
                if not 1 <= len(g.qubits) <= _lib.UA_MAX_GATE_QUBITS:
                    raise ValueError(f"dense gates act on 1..{_lib.UA_MAX_GATE_QUBITS} qubits")
                c.kind, c.data = _lib.UA_GATE_DENSE, pool.data_ptr() + offset * esize
                offset += 4 ** len(g.qubits)
                kinds.append(f"dense{len(g.qubits)}q")
            elif g.kind == "phase":
                if isinstance(g.angles, torch.Tensor):
                    a = g.angles.to(self.device, real).contiguous()
                else:
                    a = torch.as_tensor(np.ascontiguousarray(g.angles)).to(self.device, real)
                self._buffers.append(a)
                c.kind, c.data, c.data_dtype = _lib.UA_GATE_PHASE, a.data_ptr(), \
                    _lib.DTYPE_CODES[real]
                kinds.append("phase")
            else:
                c.kind = _lib.UA_GATE_CZ if g.kind == "cz" else _lib.UA_GATE_CNOT
                kinds.append(g.kind)
            c.num_targets = len(g.qubits)
            for j, q in enumerate(g.qubits):
                c.qubits[j] = q
        self._gates = gates
        self.kinds = kinds
        self._work = None
        self._graph = None
        self.dense_blocks = dense_blocks
        if dense_blocks:
            if not in_place:
                raise ValueError("dense_blocks runs in place")
            ng = len(circuit.gates)
            unit_of = (ctypes.c_int * max(1, ng))()
            nu = ctypes.c_int(0)
            _lib.check(self.lib.ua_plan_dense_blocks(self.code, self.n, gates, ng, unit_of,
                                                     ctypes.byref(nu)), "ua_plan_dense_blocks")
            units = [[] for _ in range(nu.value)]
            for i in range(ng):
                units[unit_of[i]].append(i)
            blocked = dtype.is_complex and self.n >= 14

            def unit_kind(u):
                g = circuit.gates[u[0]]
                if blocked and g.kind == "dense" and len(g.qubits) in (4, 5):
                    return "blocks" + "".join(str(len(circuit.gates[i].qubits)) for i in u)
                return kinds[u[0]]
            self.kinds = [unit_kind(u) for u in units]
            self.units = units

    def _check_state(self, state: torch.Tensor):
        if state.device != self.device or state.dtype != self.dtype:
            raise ValueError(f"state must be {self.dtype} on {self.device}")
        if state.shape[-1] != 2 ** self.n or not state.is_contiguous():
            raise ValueError("state must be contiguous (batch, 2**n)")

    def run(self, state: torch.Tensor, events=None) -> torch.Tensor:
        """Apply every gate.  In place: `state` is updated and returned.  Out of
        place: `state` is left unchanged and the result (a work buffer owned by
        the runner, reused between calls) is returned.

        events: optional _hip.EventPool of len(circuit) gates; each gate is
        bracketed by its two events on the current stream."""
        self._check_state(state)
        batch = state.numel() // (2 ** self.n)
        if self.dense_blocks:
            ev = None if events is None else ctypes.cast(events.handles,
                                                         ctypes.POINTER(ctypes.c_void_p))
            rc = self.lib.ua_run_dense_blocks(_lib.ptr(state), self.code, self.n, batch,
                                              self._gates, len(self.circuit.gates), ev,
                                              _lib.stream_handle(self.device))
            _lib.check(rc, "ua_run_dense_blocks")
            return state
        if self.in_place:
            src, work = state, None
        else:
            if self._work is None or self._work[0].shape != state.shape:
                self._work = [torch.empty_like(state), torch.empty_like(state)]
            self._work[0].copy_(state)
            src, work = self._work
        flag = ctypes.c_int(0)
        ev = None if events is None else ctypes.cast(events.handles,
                                                     ctypes.POINTER(ctypes.c_void_p))
        rc = self.lib.ua_run_circuit(
            _lib.ptr(src), None if work is None else _lib.ptr(work), self.code, self.n,
            batch, self._gates, len(self.circuit.gates), ev, ctypes.byref(flag),
            _lib.stream_handle(self.device))
        _lib.check(rc, "ua_run_circuit")
        if self.in_place:
            return state
        return work if flag.value else src

    def capture(self, state: torch.Tensor):
        """Record run(state) into a graph; replay() re-runs it on the same
        buffers.  In place only (the state tensor is the graph's operand)."""
        if not self.in_place:
            raise ValueError("graph capture is supported for in-place runners")
        self._check_state(state)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.run(state)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graph = (g, state)
        return g

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph[0].replay()
        return self._graph[1]
