"""Reductions over states (unitair/states/innerprod.py), float64 accumulation on
the GPU."""
from . import abs_squared, diag_expectation_value, inner_product, norm_squared  # noqa: F401
