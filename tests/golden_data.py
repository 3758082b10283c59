"""Loader for tests/golden/doc_known_answers.json (shared by CPU and GPU tests)."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "doc_known_answers.json"

R = 2.0 ** -0.5
NAMED = {
    "hadamard": np.array([[R, R], [R, -R]], dtype=np.complex128),
    "pauli_x": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "pauli_z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
    "cnot": np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128),
    "cz": np.diag([1, 1, 1, -1]).astype(np.complex128),
}


def cplx(x):
    """[re, im] pairs (nested) -> complex128 array."""
    a = np.asarray(x, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


def load():
    return json.loads(GOLDEN.read_text())


def exp_x(t):
    """e^{-i t X} in closed form, batched over t."""
    t = np.asarray(t, dtype=np.float64)
    c, s = np.cos(t), np.sin(t)
    m = np.empty(t.shape + (2, 2), dtype=np.complex128)
    m[..., 0, 0] = c
    m[..., 1, 1] = c
    m[..., 0, 1] = -1j * s
    m[..., 1, 0] = -1j * s
    return m
