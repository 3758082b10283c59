"""Products of gate matrices (unitair/gates/matrix_algebra.py)."""
from __future__ import annotations

import ctypes
import math
from typing import Dict, Iterable, List, Sequence

import torch

_NATIVE_DTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)


def fuse_1q_native(operators: Sequence[torch.Tensor], slots: Sequence[int], num_slots: int,
                   out: torch.Tensor) -> torch.Tensor:
    """out[..., s, :, :] = product of the operators with slot s (later ones on
    the left), computed by ua_fuse_1q_ops.  operators: equal-shape
    (*batch, 2, 2) device tensors of out's dtype, contiguous; out:
    (*batch, num_slots, 2, 2) contiguous."""
    from .. import _lib
    from .._device import on_device
    lib = _lib.load()
    batch = math.prod(out.shape[:-3])
    per_op = 4 if operators[0].dim() > 2 else 0
    code = _lib.DTYPE_CODES[out.dtype]
    with on_device(out.device):
        for c0 in range(0, len(operators), _lib.UA_FUSE_MAX_OPS):
            chunk = operators[c0:c0 + _lib.UA_FUSE_MAX_OPS]
            ptrs = (ctypes.c_void_p * len(chunk))(*[_lib.ptr(o) for o in chunk])
            sl = _lib.int_array(list(slots[c0:c0 + len(chunk)]))
            rc = lib.ua_fuse_1q_ops(_lib.ptr(out), ptrs, code, len(chunk), sl, num_slots, batch,
                                    per_op, 1 if c0 > 0 else 0, _lib.stream_handle(out.device))
            _lib.check(rc, "fuse_1q_ops")
    return out


def _native_ok(ops: List[torch.Tensor]) -> bool:
    first = ops[0]
    return (first.device.type == "cuda" and first.dtype in _NATIVE_DTYPES and
            first.dim() >= 2 and tuple(first.shape[-2:]) == (2, 2) and
            not any(o.requires_grad for o in ops) and
            all(o.shape == first.shape and o.dtype == first.dtype and o.device == first.device
                for o in ops))


def fuse_single_qubit_operators(qubits: Iterable[int],
                                operators: Iterable[torch.Tensor]) -> Dict[int, torch.Tensor]:
    """Multiply the 1-qubit gates that act on the same qubit, in order
    (unitair/gates/matrix_algebra.py:6-51).

    For qubits [2, 4, 2, 7] and operators [A, B, C, D] the result is
    {2: C @ A, 4: B, 7: D}.  Operators are (*batch_dims, 2, 2).  Qubits with one
    operator map to that operator itself.  On a ROCm device, equal-shape
    operators without autograd are multiplied by ua_fuse_1q_ops (one launch for
    every repeated qubit); otherwise, and for matrices on the host, with
    torch.matmul as in the reference.
    """
    qubits, operators = list(qubits), list(operators)
    order: Dict[int, List[int]] = {}
    for i, q in enumerate(qubits[:len(operators)]):
        order.setdefault(q, []).append(i)
    repeated = [q for q, idx in order.items() if len(idx) > 1]
    if repeated and _native_ok(operators):
        from .._device import prepare
        first = operators[0]
        out = torch.empty(tuple(first.shape[:-2]) + (len(repeated), 2, 2), dtype=first.dtype,
                          device=first.device)
        ops, slots = [], []
        for s, q in enumerate(repeated):
            for i in order[q]:
                ops.append(prepare(operators[i]).contiguous())
                slots.append(s)
        fuse_1q_native(ops, slots, len(repeated), out)
        products = {q: out[..., s, :, :] for s, q in enumerate(repeated)}
        return {q: products[q] if q in products else operators[idx[0]]
                for q, idx in order.items()}
    fused: Dict[int, torch.Tensor] = {}
    for q, op in zip(qubits, operators):
        fused[q] = torch.matmul(op, fused[q]) if q in fused else op
    return fused
