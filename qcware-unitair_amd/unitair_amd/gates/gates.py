"""Gate matrices (unitair/gates/gates.py); defined in the package."""
from . import cnot, cx, cz, exp_x, exp_y, exp_z, hadamard, pauli_x, pauli_y, pauli_z  # noqa: F401
