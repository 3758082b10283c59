"""Sampling (unitair/simulation/measurement.py)."""
from ..measurement import MeasurementHistogram, measure  # noqa: F401
