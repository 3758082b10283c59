"""Vector <-> tensor layout (unitair/states/conversions.py)."""
from . import to_tensor_layout, to_vector_layout  # noqa: F401
