"""Seeded synthetic circuits for the benchmark workloads.

Each builder returns a Circuit whose gates are drawn with numpy PCG64 from
`seed`, so the GPU run, the CPU oracle and the CPU baseline all see the same
gates.  Haar-random unitaries use the QR construction in gates.haar_unitary.
"""
from __future__ import annotations

import numpy as np

from .circuits import Circuit, Gate
from .gates import haar_unitary

_CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)


class SeededAngles:
    """Full-length phase angles drawn on first use from their own seed:
    uniform [0, 2pi) float32 values, 2**n of them.  Keeps circuit construction
    cheap when the angles are never looked at (planning, tests)."""

    def __init__(self, n: int, seed: int):
        self.n, self.seed = n, int(seed)
        self._values = None

    def values(self) -> np.ndarray:
        if self._values is None:
            rng = np.random.default_rng(self.seed)
            self._values = rng.random(2 ** self.n, dtype=np.float32) * np.float32(2 * np.pi)
        return self._values

    def __array__(self, dtype=None, copy=None):
        v = self.values()
        return v if dtype is None else v.astype(dtype)

    def __len__(self):
        return 2 ** self.n


def single_qubit_phase_angles(n: int, qubit: int, phi0: float, phi1: float) -> np.ndarray:
    """Full-length apply_phase angles of a diagonal gate on one qubit:
    angle[x] = phi0 if that qubit's bit of x is 0 else phi1."""
    x = np.arange(2 ** n, dtype=np.int64)
    bit = (x >> (n - 1 - qubit)) & 1
    return np.where(bit == 1, phi1, phi0).astype(np.float64)


def layered_circuit_12q(seed: int = 1, n: int = 12, depth: int = 20) -> Circuit:
    """12 qubits, depth 20: per layer a Haar 2x2 on every qubit, CNOT(i, i+1)
    for i of the layer's parity, and apply_phase with uniform [0, 2pi) phases
    on one random qubit."""
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(depth):
        for q in range(n):
            gates.append(Gate("dense", (q,), matrix=haar_unitary(2, rng)))
        for i in range(layer % 2, n - 1, 2):
            gates.append(Gate("cnot", (i, i + 1)))
        q = int(rng.integers(n))
        phi = rng.uniform(0.0, 2 * np.pi, size=2)
        gates.append(Gate("phase", (q,), angles=single_qubit_phase_angles(n, q, *phi)))
    return Circuit(n, gates)


def random_circuit_24q(seed: int = 2, n: int = 24, n_1q: int = 120, n_2q_each: int = 20,
                       n_phase: int = 20) -> Circuit:
    """One state, 200 gates: 120 Haar 1-qubit gates on uniform targets that
    include qubits 0-5 and n-6..n-1 (the largest and the smallest strides),
    20 CNOT + 20 CZ + 20 Haar 4x4 on random distinct ordered pairs, and 20
    apply_phase with full-length uniform [0, 2pi) float32 angles (the complex128
    runs use the same values widened); shuffled."""
    rng = np.random.default_rng(seed)
    gates = []
    forced = list(range(6)) + list(range(n - 6, n))
    targets = forced + list(rng.integers(0, n, size=n_1q - len(forced)))
    for q in targets:
        gates.append(Gate("dense", (int(q),), matrix=haar_unitary(2, rng)))

    def pair():
        a, b = rng.choice(n, size=2, replace=False)
        return int(a), int(b)

    for _ in range(n_2q_each):
        gates.append(Gate("cnot", pair()))
    for _ in range(n_2q_each):
        gates.append(Gate("cz", pair()))
    for _ in range(n_2q_each):
        gates.append(Gate("dense", pair(), matrix=haar_unitary(4, rng)))
    for _ in range(n_phase):
        gates.append(Gate("phase", (), angles=SeededAngles(n, rng.integers(2 ** 62))))
    order = rng.permutation(len(gates))
    return Circuit(n, [gates[i] for i in order])


def fused_dense_circuit(seed: int = 4, n: int = 28, n5: int = 64, n4: int = 16) -> Circuit:
    """One state, Haar-random dense blocks: n5 32x32 unitaries on random
    5-qubit subsets (the first two are qubits 0-4 and n-5..n-1, the largest
    and the smallest strides) and n4 16x16 unitaries on random 4-qubit
    subsets; shuffled.  The same builder with a smaller n gives the slice
    used for the complex128 CPU comparison."""
    rng = np.random.default_rng(seed)
    gates = []
    subsets5 = [tuple(range(5)), tuple(range(n - 5, n))]
    while len(subsets5) < n5:
        subsets5.append(tuple(int(q) for q in rng.choice(n, 5, replace=False)))
    for qs in subsets5:
        gates.append(Gate("dense", qs, matrix=haar_unitary(32, rng)))
    for _ in range(n4):
        qs = tuple(int(q) for q in rng.choice(n, 4, replace=False))
        gates.append(Gate("dense", qs, matrix=haar_unitary(16, rng)))
    order = rng.permutation(len(gates))
    return Circuit(n, [gates[i] for i in order])
