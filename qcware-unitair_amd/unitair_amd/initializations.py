"""State constructors (mirror of unitair/initializations.py).

`rand_state` draws like the reference (torch.randn of the target dtype,
normalised per state, initializations.py:71-100).  `seeded_state` is the
reproducible variant used by tests and benchmarks: numpy PCG64 Gaussian real
and imaginary parts of variance 1/2, normalised in float64, then cast.
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np
import torch


def unit_vector(index: int, num_qubits: Optional[int] = None, dim: Optional[int] = None,
                device: torch.device = torch.device("cpu"),
                dtype: torch.dtype = torch.complex64) -> torch.Tensor:
    """Basis vector |index> (initializations.py:46-68)."""
    if dim is None:
        if num_qubits is None:
            raise TypeError("To specify a unit vector, provide either `num_qubits` or `dim`.")
        dim = 2 ** num_qubits
    elif num_qubits is not None:
        raise TypeError("Unit vector can be specified by `dim` or `num_qubits` but not both.")
    v = torch.zeros(dim, device=device, dtype=dtype)
    v[index] = 1
    return v


def unit_vector_from_bitstring(bitstring: Union[str, Sequence[int]],
                               device: torch.device = torch.device("cpu"),
                               dtype: torch.dtype = torch.complex64) -> torch.Tensor:
    """|b_0 b_1 ... b_{n-1}>, b_0 the most significant bit (initializations.py:8-43)."""
    bits = [int(b) for b in bitstring]
    index = 0
    for b in bits:
        index = 2 * index + b
    return unit_vector(index, num_qubits=len(bits), device=device, dtype=dtype)


def rand_state(num_qubits: int, batch_dims: Optional[Sequence] = None,
               device: torch.device = torch.device("cpu"),
               dtype: torch.dtype = torch.complex64,
               requires_grad: bool = False) -> torch.Tensor:
    """Random normalised state(s), Gaussian entries (initializations.py:71-100)."""
    size = torch.Size(batch_dims or ()) + torch.Size((2 ** num_qubits,))
    state = torch.randn(size, device=device, dtype=dtype)
    norm = (state.conj() * state).real.sum(-1).sqrt()
    state /= norm.unsqueeze(-1)
    return state.requires_grad_(requires_grad)


def uniform_superposition(num_qubits: int, batch_dims: Optional[Sequence] = None,
                          device: torch.device = torch.device("cpu"),
                          dtype: torch.dtype = torch.complex64,
                          requires_grad: bool = False) -> torch.Tensor:
    """|+...+> (initializations.py:103-128)."""
    state = 2 ** (-num_qubits / 2.0) * torch.ones(2 ** num_qubits, device=device, dtype=dtype)
    if batch_dims is not None:
        state = state.repeat(torch.Size(batch_dims) + (1,))
    return state.requires_grad_(requires_grad)


def seeded_state_np(num_qubits: int, batch: Sequence[int] = (), seed: int = 0) -> np.ndarray:
    """Reproducible complex128 state(s) of shape (*batch, 2**n), unit norm."""
    rng = np.random.default_rng(seed)
    shape = tuple(batch) + (2 ** num_qubits,)
    x = np.empty(shape, dtype=np.complex128)
    x.real = rng.standard_normal(shape) * np.sqrt(0.5)
    x.imag = rng.standard_normal(shape) * np.sqrt(0.5)
    x /= np.sqrt(np.sum(np.abs(x) ** 2, axis=-1, keepdims=True))
    return x


def seeded_state(num_qubits: int, batch: Sequence[int] = (), seed: int = 0,
                 dtype: torch.dtype = torch.complex64,
                 device: Union[str, torch.device] = "cpu") -> torch.Tensor:
    """seeded_state_np cast to `dtype` on `device`."""
    return torch.from_numpy(seeded_state_np(num_qubits, batch, seed)).to(
        device=device, dtype=dtype)
