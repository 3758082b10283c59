"""Run a gate list on the CPU in complex128 (TEST INFRASTRUCTURE ONLY).

Gates are duck-typed: `.kind` in {"dense", "phase", "cz", "cnot"}, `.qubits`,
`.matrix` (dense), `.angles` (phase).  Each gate is what the user loop of
unitair would call: apply_operator(matrix | cz() | cnot(), qubits, state)
(operations.py:45-148, gates/gates.py:170-213) or apply_phase(angles, state)
(operations.py:15-42).
"""
from __future__ import annotations

import numpy as np

from . import closed_form, unitair_np

_CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
_CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)


def _matrix(g):
    if g.kind == "dense":
        return np.asarray(g.matrix, dtype=np.complex128)
    return _CZ if g.kind == "cz" else _CNOT


def run_closed_form(gates, state: np.ndarray) -> np.ndarray:
    """Fast path: the C closed form, in place on a complex128 copy."""
    psi = np.array(state, dtype=np.complex128, copy=True)
    for g in gates:
        if g.kind == "phase":
            closed_form.apply_phase(np.asarray(g.angles, np.float64), psi, out=psi)
        else:
            closed_form.apply_dense(_matrix(g), g.qubits, psi, out=psi)
    return psi


def run_restatement(gates, state: np.ndarray) -> np.ndarray:
    """Slow path: the numpy restatement of the reference's permute/einsum."""
    psi = np.asarray(state, dtype=np.complex128)
    for g in gates:
        if g.kind == "phase":
            psi = unitair_np.apply_phase(np.asarray(g.angles, np.float64), psi)
        else:
            psi = unitair_np.apply_operator(_matrix(g), g.qubits, psi)
    return psi
