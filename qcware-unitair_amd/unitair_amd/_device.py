"""Device checks shared by the host-side entry points."""
from __future__ import annotations

import torch


def require_gpu(t: torch.Tensor, what: str) -> None:
    """Raise unless `t` lives on a ROCm GPU: this package has no CPU path."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"unitair_amd.{what}: expected a torch.Tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise RuntimeError(
            f"unitair_amd.{what}: tensors must be on a ROCm GPU device, got {t.device} "
            f"(unitair_amd runs only the HIP kernels; there is no CPU path)")


def same_device(a: torch.Tensor, b: torch.Tensor, what: str) -> None:
    if a.device != b.device:
        raise RuntimeError(
            f"unitair_amd.{what}: Expected all tensors to be on the same device, "
            f"but found {a.device} and {b.device}")


def prepare(t: torch.Tensor) -> torch.Tensor:
    """Materialise lazy conjugate/negative views so data_ptr() holds the values."""
    if t.is_conj():
        t = t.resolve_conj()
    if t.is_neg():
        t = t.resolve_neg()
    return t


class on_device:
    """Make `device` current for the duration of a launch (cheap when it
    already is)."""

    def __init__(self, device: torch.device):
        self.index = device.index if device.index is not None else torch.cuda.current_device()
        self.prev = None

    def __enter__(self):
        cur = torch.cuda.current_device()
        if cur != self.index:
            self.prev = cur
            torch.cuda.set_device(self.index)
        return self

    def __exit__(self, *exc):
        if self.prev is not None:
            torch.cuda.set_device(self.prev)
        return False
