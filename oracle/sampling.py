"""Philox4x32-10 and the draw rule of ua_sample, in Python integers.

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
