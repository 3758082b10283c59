"""Gate matrices in unitair's conventions (unitair/gates/gates.py).

Row index = output, and for a multi-qubit gate the first qubit it is applied
to is the most significant bit of the matrix index, so `cnot()` has its
control first (gates.py:170-188).  exp_x/y/z(a) = e^{-i a X/Y/Z} with no
factor 1/2 (gates.py:6-91): the textbook RY(t) is exp_y(t/2).

These are small host-built tensors.  Unlike the reference, whose exp_x/y/z
return complex64 whatever `dtype` is passed, the builders here honour
`dtype` (needed for complex128 circuits).
"""
from __future__ import annotations

import math
from typing import Optional, Union

import numpy as np
import torch

_C64 = torch.complex64


def _const(rows, device, dtype):
    t = torch.tensor(rows, dtype=torch.complex128)
    return t.to(device=device or "cpu", dtype=dtype or _C64)


def hadamard(device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None):
    r = 2.0 ** -0.5
    return _const([[r, r], [r, -r]], device, dtype)


def pauli_x(device=None, dtype=None):
    return _const([[0, 1], [1, 0]], device, dtype)


def pauli_y(device=None, dtype=None):
    return _const([[0, -1j], [1j, 0]], device, dtype)


def pauli_z(device=None, dtype=None):
    return _const([[1, 0], [0, -1]], device, dtype)


def cnot(device=None, dtype=None):
    """CNOT with the control as the first qubit (gates.py:170-188)."""
    return _const([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], device, dtype)


cx = cnot


def cz(device=None, dtype=None):
    return _const([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]], device, dtype)


def _angles(angle, dtype):
    real = torch.float64 if dtype == torch.complex128 else torch.float32
    if isinstance(angle, torch.Tensor):
        return angle.to(real)
    return torch.tensor(float(angle), dtype=real)


def exp_x(angle: Union[torch.Tensor, float], dtype: torch.dtype = _C64):
    """e^{-i a X}; batch shape = shape of `angle`."""
    a = _angles(angle, dtype)
    c, s = torch.cos(a).to(dtype), torch.sin(a).to(dtype)
    return torch.stack([torch.stack([c, -1j * s], -1), torch.stack([-1j * s, c], -1)], -2)


def exp_y(angle: Union[torch.Tensor, float], dtype: torch.dtype = _C64):
    """e^{-i a Y}."""
    a = _angles(angle, dtype)
    c, s = torch.cos(a).to(dtype), torch.sin(a).to(dtype)
    return torch.stack([torch.stack([c, -s], -1), torch.stack([s, c], -1)], -2)


def exp_z(angle: Union[torch.Tensor, float], dtype: torch.dtype = _C64):
    """e^{-i a Z} = diag(e^{-ia}, e^{ia})."""
    a = _angles(angle, dtype)
    e = torch.polar(torch.ones_like(a), -a).to(dtype)
    z = torch.zeros_like(e)
    return torch.stack([torch.stack([e, z], -1), torch.stack([z, e.conj()], -1)], -2)


def haar_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random unitary (complex128) by QR of a complex Gaussian matrix with
    the phases of R's diagonal moved into Q (Mezzadri, Notices AMS 54, 2007)."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[None, :]
