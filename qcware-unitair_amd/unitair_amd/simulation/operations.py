"""Gate application (unitair/simulation/operations.py); the functions live in
the package, this module gives them the reference's import path."""
from . import (act_first_qubits, act_last_qubit, apply_all_qubits, apply_operator,  # noqa: F401
               apply_phase, apply_to_qubits, multi_controlled_x, multi_controlled_z, multi_cz,
               permute_qubits, roll_qubits, swap)
