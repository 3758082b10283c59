"""Products of gate matrices (unitair/gates/matrix_algebra.py)."""
from __future__ import annotations

from typing import Dict, Iterable

import torch


def fuse_single_qubit_operators(qubits: Iterable[int],
                                operators: Iterable[torch.Tensor]) -> Dict[int, torch.Tensor]:
    """Multiply the 1-qubit gates that act on the same qubit, in order
    (unitair/gates/matrix_algebra.py:6-51).

    For qubits [2, 4, 2, 7] and operators [A, B, C, D] the result is
    {2: C @ A, 4: B, 7: D}.  Operators are (*batch_dims, 2, 2); the products
    run where the tensors live (torch.matmul, as in the reference).
    """
    fused: Dict[int, torch.Tensor] = {}
    for q, op in zip(qubits, operators):
        fused[q] = torch.matmul(op, fused[q]) if q in fused else op
    return fused
