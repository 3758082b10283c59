"""Shapes, counting and layout errors (unitair/states/shapes.py)."""
from . import (StateLayout, StateShapeError, count_batch_dims_tensor, count_qubits,  # noqa: F401
               count_qubits_gate_matrix, count_qubits_tensor, get_qubit_indices,
               hilbert_space_dim, subset_roll_to_back, subset_roll_to_front)
