"""Small host helpers of unitair/utils.py."""
from __future__ import annotations

from typing import List, Sequence

import torch


def permutation_to_front(num_qubits: int, entries: Sequence[int]) -> List[int]:
    """The permutation that moves `entries` to the front in order, the other
    indices after them in order (unitair/utils.py:5-27)."""
    entries = list(entries)
    return entries + [i for i in range(num_qubits) if i not in entries]


def inverse_list_permutation(permutation: Sequence[int]) -> List[int]:
    """inv with inv[permutation[i]] = i (unitair/utils.py:30-49)."""
    inv = [0] * len(permutation)
    for i, p in enumerate(permutation):
        inv[p] = i
    return inv


def int_to_bin(x: int, num_bits: int, str_output: bool = False,
               device: torch.device = torch.device("cpu")):
    """Big-endian binary digits of x as a bool tensor, or as a '0'/'1'
    string (unitair/utils.py:50-69)."""
    if x >= 2 ** num_bits:
        raise ValueError(f"Insufficient bits to store integer {x}.")
    if x < 0:
        raise ValueError(f"Expected a nonnegative integer, found {x}.")
    digits = [int(d) for d in format(x, "b")] if x else [0]
    digits = [0] * (num_bits - len(digits)) + digits
    if str_output:
        return "".join(str(d) for d in digits)
    return torch.tensor(digits, dtype=torch.bool, device=device)


def bool_to_strings(bool_data) -> str:
    """[True, False, False, True] -> '1001' (unitair/utils.py:72-80)."""
    return "".join(str(int(e.item() if isinstance(e, torch.Tensor) else e)) for e in bool_data)
