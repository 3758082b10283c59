"""Computational-basis sampling (unitair/simulation/measurement.py).

`measure` draws the samples on the GPU (ua_sample: float64 prefix sums of
|psi|^2 and Philox4x32-10 uniforms); the histogram is then assembled on the
host in the reference's format.  The draws are random, so they match the
reference in distribution, not sample for sample.  The Philox key comes
from torch's default CPU generator, so torch.manual_seed makes a run
repeatable.
"""
from __future__ import annotations

from collections import OrderedDict
from typing import Optional

import torch

from . import _lib
from ._device import on_device, prepare, require_gpu
from .states import count_qubits
from .utils import int_to_bin

__all__ = ["measure", "MeasurementHistogram", "sample_indices"]


def sample_indices(state: torch.Tensor, num_samples: int,
                   generator: Optional[torch.Generator] = None) -> torch.Tensor:
    """num_samples basis-state indices (int64, on the state's device) drawn
    with probability |state[x]|^2 / ||state||^2.  One state (vector layout)."""
    require_gpu(state, "measure")
    if state.dtype not in _lib.DTYPE_CODES:
        raise TypeError(f"measure: dtype {state.dtype} is not supported")
    n = count_qubits(state)
    if state.dim() != 1:
        # the reference's histogram step hashes sample.tolist() rows and
        # fails on batched states
        raise TypeError("measure: one state in vector layout, size (2^n,), is expected")
    num_samples = int(num_samples)
    if not 0 <= num_samples < 2 ** 31:
        # sample ids and bucket slots are 32-bit on the device
        raise ValueError(f"measure: num_samples={num_samples} out of range (0 .. 2^31 - 1)")
    state = prepare(state).contiguous()
    out = torch.empty(num_samples, dtype=torch.int64, device=state.device)
    if num_samples == 0:
        return out
    lib = _lib.load()
    ws = torch.empty(lib.ua_sample_workspace_bytes(n, num_samples), dtype=torch.uint8,
                     device=state.device)
    seed = int(torch.randint(0, 2 ** 62, (1,), generator=generator).item())
    with on_device(state.device):
        rc = lib.ua_sample(_lib.ptr(out), _lib.ptr(state), _lib.DTYPE_CODES[state.dtype], n,
                           num_samples, seed, 0, _lib.ptr(ws),
                           _lib.stream_handle(state.device))
    _lib.check(rc, "measure")
    # Categorical(probs) (measurement.py:45-48) rejects a distribution whose
    # total is zero or not finite; the draws above are then meaningless
    off = lib.ua_sample_total_offset(n, num_samples)
    total = ws[off:off + 8].view(torch.float64)
    if not bool(((total > 0) & torch.isfinite(total)).item()):
        raise ValueError(
            f"Expected parameter probs (Tensor of shape {tuple(state.shape)}) of distribution "
            f"Categorical to satisfy the constraint Simplex(), but found invalid values: "
            f"sum of |state|^2 = {float(total.item())}")
    return out


def measure(state: torch.Tensor, num_samples: int, raw_output: bool = False):
    """Draw samples of the computational-basis distribution of `state`
    (measurement.py:9-71).  The state is unchanged.  Returns a
    MeasurementHistogram sorted by count (descending), or with raw_output a
    dict {outcome int: count}."""
    num_qubits = count_qubits(state)
    samples = sample_indices(state, num_samples)
    configs, counts = torch.unique(samples, return_counts=True)
    configs, counts = configs.tolist(), counts.tolist()
    if raw_output:
        return dict(zip(configs, counts))
    order = sorted(range(len(configs)), key=lambda i: -counts[i])
    histogram = MeasurementHistogram(num_qubits=num_qubits)
    for i in order:
        histogram[int_to_bin(configs[i], num_bits=num_qubits, str_output=True)] = counts[i]
    return histogram


class MeasurementHistogram:
    """Record of a measurement of a given number of qubits
    (measurement.py:74-140)."""

    def __init__(self, num_qubits: int, histogram: Optional[OrderedDict] = None):
        self.num_qubits = num_qubits
        self.histogram = OrderedDict() if histogram is None else histogram
        self._int_key_histogram = None

    def __getitem__(self, item: str):
        try:
            return self.histogram[item]
        except KeyError:
            if set(item).issubset({"0", "1"}) and len(item) == self.num_qubits:
                return 0
            raise KeyError(f"Given key {item} is not a valid binary string for "
                           f"{self.num_qubits} bits.")

    def __setitem__(self, key, value):
        self.histogram[key] = value

    @property
    def num_distinct_samples(self):
        return len(self.histogram)

    @property
    def observed_samples(self):
        return set(self.histogram)

    def int_key_histogram(self):
        if self._int_key_histogram is None:
            self._int_key_histogram = OrderedDict(
                (int(config, 2) if config else 0, count)
                for config, count in self.histogram.items())
        return self._int_key_histogram

    def __repr__(self):
        items = list(self.histogram.items())
        shortened = len(items) >= 50
        lines = [f"{c}: {k}" for c, k in (items[:15] if shortened else items)]
        out = "\n".join(lines)
        if shortened:
            out += f"\n ... ({len(items) - 15} lines omitted)"
        return out
