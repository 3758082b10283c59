"""Operator batch helpers (unitair/simulation/utils.py)."""
import torch


def count_gate_batch_dims(gate: torch.Tensor) -> int:
    """Number of batch dims of a (*batch, 2^k, 2^k) gate
    (unitair/simulation/utils.py:4-20)."""
    out = gate.dim() - 2
    if out < 0:
        raise RuntimeError(
            f"Gate with size {gate.size()} is incorrectly shaped for an operator batch. "
            f"Expected size is\n  (*optional_batch_dims, 2^k, 2^k)\nwith k the number of "
            f"qubits on which the gate acts.")
    return out


def gate_batch_size(gate: torch.Tensor) -> torch.Size:
    """The batch dims of a gate (unitair/simulation/utils.py:23-31)."""
    return gate.size()[:count_gate_batch_dims(gate)]
