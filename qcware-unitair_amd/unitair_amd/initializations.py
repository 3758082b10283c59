"""State constructors (mirror of unitair/initializations.py).

On a ROCm device the states are made there by the library: `rand_state`
(Gaussian entries from Philox4x32-10 + Box-Muller, normalised per state in
float64; the generator key comes from torch's default CPU generator, so
torch.manual_seed repeats a draw, or from `seed=`), `uniform_superposition`
and `unit_vector` (a fill kernel).  The distribution is the reference's
(torch.randn normalised, initializations.py:71-100), not its bits.  On the
CPU the constructors are the reference's torch calls.

`seeded_state_np` / `seeded_state` are the host-side reproducible states
used by tests and the checker: numpy PCG64 Gaussian real and imaginary parts
of variance 1/2, normalised in float64, then cast.
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np
import torch

from . import _lib


def _on_gpu(device) -> bool:
    return torch.device(device).type == "cuda"


def _gpu_fill(shape, num_qubits, device, dtype, kind, index=0):
    from ._device import on_device
    out = torch.empty(shape, device=device, dtype=dtype)
    batch = out.numel() >> num_qubits
    lib = _lib.load()
    with on_device(out.device):
        rc = lib.ua_fill_state(_lib.ptr(out), _lib.DTYPE_CODES[dtype], num_qubits, batch, kind,
                               index, _lib.stream_handle(out.device))
    _lib.check(rc, "fill_state")
    return out


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
    index = int(index)
    n = dim.bit_length() - 1
    if (_on_gpu(device) and dim == 2 ** n and dtype in _lib.DTYPE_CODES
            and -dim <= index < dim):
        return _gpu_fill((dim,), n, device, dtype, 1, index % dim)
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
               requires_grad: bool = False, *, seed: Optional[int] = None) -> torch.Tensor:
    """Random normalised state(s), Gaussian entries (initializations.py:71-100).
    `seed` (this package only) keys the device generator directly."""
    size = torch.Size(batch_dims or ()) + torch.Size((2 ** num_qubits,))
    if _on_gpu(device) and dtype in _lib.DTYPE_CODES and 2 <= num_qubits <= 40:
        from ._device import on_device
        state = torch.empty(size, device=device, dtype=dtype)
        batch = state.numel() >> num_qubits
        if seed is None:
            seed = int(torch.randint(0, 2 ** 62, (1,)).item())
        lib = _lib.load()
        code = _lib.DTYPE_CODES[dtype]
        ws = torch.empty(max(8, lib.ua_rand_state_workspace_bytes(num_qubits, batch, code)),
                         dtype=torch.uint8, device=state.device)
        with on_device(state.device):
            rc = lib.ua_rand_state(_lib.ptr(state), code, num_qubits, batch, int(seed),
                                   _lib.ptr(ws), _lib.stream_handle(state.device))
        _lib.check(rc, "rand_state")
        return state.requires_grad_(requires_grad)
    gen = None
    if seed is not None and not _on_gpu(device):
        gen = torch.Generator().manual_seed(int(seed))
    state = torch.randn(size, device=device, dtype=dtype, generator=gen)
    norm = (state.conj() * state).real.sum(-1).sqrt()
    state /= norm.unsqueeze(-1)
    return state.requires_grad_(requires_grad)


def uniform_superposition(num_qubits: int, batch_dims: Optional[Sequence] = None,
                          device: torch.device = torch.device("cpu"),
                          dtype: torch.dtype = torch.complex64,
                          requires_grad: bool = False) -> torch.Tensor:
    """|+...+> (initializations.py:103-128)."""
    if _on_gpu(device) and dtype in _lib.DTYPE_CODES:
        shape = torch.Size(batch_dims or ()) + (2 ** num_qubits,)
        return _gpu_fill(shape, num_qubits, device, dtype, 0).requires_grad_(requires_grad)
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
