"""unitair_amd: MI355X-native gate application with unitair's call surface.

Mirrors the top-level API of unitair (unitair/__init__.py:1-12) for the
gate-application path: states stay plain torch tensors in unitair's vector
layout, qubit 0 is the most significant index bit, and the arithmetic runs in
hand-written HIP kernels for gfx950 (libunitair_amd.so, C ABI in
include/unitair_amd.h).  Tensors must live on a ROCm GPU.
"""
from . import states
from . import simulation
from . import gates
from . import initializations
from . import utils

from .states import StateLayout, StateShapeError
from .states import count_qubits, hilbert_space_dim
from .states import diag_expectation_value, inner_product, norm_squared, abs_squared
from .initializations import unit_vector, uniform_superposition, rand_state

VECTOR_LAYOUT = StateLayout.VECTOR
TENSOR_LAYOUT = StateLayout.TENSOR

__version__ = "0.1.0"
