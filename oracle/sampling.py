"""Philox4x32-10, the draw rule of ua_sample and the states of
ua_rand_state, in Python integers and numpy.

TEST INFRASTRUCTURE ONLY (the checker of the sampling kernels).
Philox4x32-10 follows the published algorithm (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of
  (c0, c1, c2, c3) -> (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0))
with the key bumped by the Weyl constants between rounds.  The draw rule
(first index whose inclusive prefix of |psi|^2 exceeds U * total) is the
inverse-CDF form of torch.distributions.Categorical(probs).sample used by
unitair/simulation/measurement.py:47-48.
"""
from __future__ import annotations

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, \
            ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return c0, c1, c2, c3


def uniform53(s: int, seed: int, offset: int = 0) -> float:
    """The uniform in [0, 1) that ua_sample uses for sample s."""
    ctr = s + offset
    c = philox4x32_10((ctr & MASK, (ctr >> 32) & MASK, 0, 0), (seed & MASK, (seed >> 32) & MASK))
    bits = ((c[0] << 21) ^ (c[1] >> 11)) & ((1 << 53) - 1)
    return bits * 2.0 ** -53


def draw(probs: np.ndarray, u: float) -> int:
    """First index whose inclusive prefix sum exceeds u * sum(probs)."""
    cum = np.cumsum(probs)
    return int(np.searchsorted(cum, u * cum[-1], side="right"))


def rand_state_np(num_qubits: int, batch: int, seed: int, complex_dtype: bool) -> np.ndarray:
    """The states ua_rand_state makes, in float64: counter c gives four
    Box-Muller normals (u1 = (x + 1) 2^-32, u2 = y 2^-32; r = sqrt(-2 ln u1);
    (r cos 2 pi u2, r sin 2 pi u2)) from Philox4x32-10 with key `seed`;
    complex entries take two normals (re, im), real entries one; each state
    is divided by its float64 norm."""
    N = 2 ** num_qubits
    per_ctr = 2 if complex_dtype else 4
    z = np.empty(batch * N * (2 if complex_dtype else 1))
    for c in range(batch * N // per_ctr):
        x = philox4x32_10((c & MASK, (c >> 32) & MASK, 0, 0), (seed & MASK, (seed >> 32) & MASK))
        for h in range(2):
            u1 = (x[2 * h] + 1.0) * 2.0 ** -32
            u2 = x[2 * h + 1] * 2.0 ** -32
            r = np.sqrt(-2.0 * np.log(u1))
            z[4 * c + 2 * h] = r * np.cos(2 * np.pi * u2)
            z[4 * c + 2 * h + 1] = r * np.sin(2 * np.pi * u2)
    v = z[0::2] + 1j * z[1::2] if complex_dtype else z
    v = v.reshape(batch, N)
    return v / np.sqrt(np.sum(np.abs(v) ** 2, axis=-1, keepdims=True))
