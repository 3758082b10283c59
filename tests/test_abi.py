"""The C-ABI library loads and exports what include/unitair_amd.h declares.

No kernel is launched here (no GPU): only calls that the library rejects
during argument validation, before any HIP call.
"""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "unitair_amd.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ua_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from unitair_amd import _lib
    if not _lib.library_path().exists():
        pytest.fail("libunitair_amd.so is not built (run __graft_entry__.build())")
    return _lib.load()


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("ua_apply_operator", "ua_apply_phase", "ua_apply_all_qubits",
                 "ua_norm_squared", "ua_inner_product", "ua_diag_expectation",
                 "ua_abs_squared", "ua_apply_cz", "ua_apply_cnot", "ua_last_error"):
        assert name in syms


def test_every_declared_symbol_is_exported(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert missing == []


def test_python_binding_covers_header():
    from unitair_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_abi_version(lib):
    from unitair_amd import _lib
    assert lib.ua_abi_version() == _lib.ABI_VERSION


def test_validation_errors_before_launch(lib):
    from unitair_amd import _lib
    q = _lib.int_array([0, 0])
    buf = ctypes.c_void_p(0x1000)   # never dereferenced: validation fails first
    rc = lib.ua_apply_operator(buf, buf, buf, _lib.UA_COMPLEX64, 4, 2, q, 1, 16, 0, None)
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    assert "more than once" in lib.ua_last_error().decode()
    q = _lib.int_array([5])
    rc = lib.ua_apply_operator(buf, buf, buf, _lib.UA_COMPLEX64, 4, 1, q, 1, 16, 0, None)
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    rc = lib.ua_apply_operator(buf, buf, buf, 17, 4, 1, q, 1, 16, 0, None)
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    rc = lib.ua_apply_cz(buf, _lib.UA_COMPLEX64, 4, 1, 1, 1, None)
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    # zero-sized work is a no-op that succeeds without touching memory
    rc = lib.ua_apply_operator(None, None, None, _lib.UA_COMPLEX64, 4, 1,
                               _lib.int_array([1]), 0, 16, 0, None)
    assert rc == _lib.UA_OK


def test_validation_errors_of_layer_permutation_sampling_and_constructors(lib):
    from unitair_amd import _lib
    bad = _lib.UA_ERR_INVALID_ARGUMENT
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)   # never dereferenced
    c64 = _lib.UA_COMPLEX64
    # 1-qubit layer: repeated qubit, qubit out of range, bad operator stride, aliasing
    assert lib.ua_apply_1q_layer(a, b, a, c64, 12, 2, _lib.int_array([3, 3]), 1, 4096, 0,
                                 None) == bad
    assert lib.ua_apply_1q_layer(a, b, a, c64, 12, 1, _lib.int_array([12]), 1, 4096, 0,
                                 None) == bad
    assert lib.ua_apply_1q_layer(a, b, a, c64, 12, 2, _lib.int_array([1, 2]), 1, 4096, 5,
                                 None) == bad
    assert lib.ua_apply_1q_layer(a, a, b, c64, 12, 1, _lib.int_array([1]), 1, 4096, 0,
                                 None) == bad
    assert lib.ua_apply_all_qubits(a, b, a, c64, 12, 1, 4096, 3, None) == bad
    # permutation: not a permutation, aliasing
    assert lib.ua_permute_qubits(a, b, c64, 4, _lib.int_array([0, 1, 1, 3]), 1, 16, None) == bad
    assert "not a permutation" in lib.ua_last_error().decode()
    assert lib.ua_permute_qubits(a, a, c64, 4, _lib.int_array([0, 1, 2, 3]), 1, 16, None) == bad
    # sign gates: pair with equal bits, mask beyond the state
    assert lib.ua_apply_sign_gates(a, b, c64, 4, 1, _lib.int_array([2, 2]), 0, 0, 1, 16,
                                   None) == bad
    assert lib.ua_apply_sign_gates(a, b, c64, 4, 0, None, 1, 1 << 4, 1, 16, None) == bad
    # controlled flip: target among the controls, target out of range
    assert lib.ua_apply_controlled_flip(a, b, c64, 4, 0b0101, 2, 1, 16, None) == bad
    assert lib.ua_apply_controlled_flip(a, b, c64, 4, 0, 4, 1, 16, None) == bad
    # sampling and constructors
    assert lib.ua_sample(None, a, c64, 4, -1, 0, 0, b, None) == bad
    assert lib.ua_sample(None, a, c64, 41, 5, 0, 0, b, None) == bad
    assert lib.ua_sample(None, a, c64, 4, 1 << 31, 0, 0, b, None) == bad
    assert lib.ua_sample_workspace_bytes(41, 5) == -1
    assert lib.ua_sample_total_offset(41, 5) == -1
    assert lib.ua_sample_total_offset(12, 0) == 8     # one chunk: prefix[0] after chunk_sum[0]
    assert lib.ua_sample_total_offset(20, 7) == 8 * (2 * 256 - 1)
    assert lib.ua_sample_workspace_bytes(12, 1 << 31) == -1
    # chunk total and prefix (float64), 8 bucket sub-counts and a bucket
    # start (uint32) per 4096-amplitude chunk; chunk key, slot and bucket
    # entry per sample
    assert lib.ua_sample_workspace_bytes(12, 0) == 56   # 52, rounded up to 8
    assert lib.ua_sample_workspace_bytes(20, 101) == 52 * 256 + 12 * 101 + 4
    # operator gradient: 2048 slices of 2^23 pairs for a 1-qubit operator, 16
    # slices x 256 blocks for a 6-qubit one, none when k == n
    assert lib.ua_operator_gradient_workspace_bytes(24, 1, 1, 1) == 2048 * 4 * 16
    assert lib.ua_operator_gradient_workspace_bytes(24, 6, 1, 1) == 16 * 4096 * 16
    assert lib.ua_operator_gradient_workspace_bytes(10, 10, 1, 1) == 0
    assert lib.ua_operator_gradient_workspace_bytes(4, 5, 1, 1) == -1
    assert lib.ua_operator_gradient_workspace_bytes(4, 1, 3, 2) == -1
    assert lib.ua_operator_gradient(a, b, a, c64, 24, 1, _lib.int_array([0]), 1, 1, 0, None,
                                    None) == bad
    assert "workspace" in lib.ua_last_error().decode()
    assert lib.ua_rand_state(a, c64, 1, 1, 0, b, None) == bad
    assert lib.ua_fill_state(a, c64, 4, 1, 1, 16, None) == bad
    assert lib.ua_fill_state(a, c64, 4, 1, 2, 0, None) == bad
    # zero-sized work succeeds without touching memory
    assert lib.ua_apply_1q_layer(None, None, None, c64, 12, 0, None, 0, 4096, 0, None) == _lib.UA_OK
    assert lib.ua_sample(None, None, c64, 4, 0, 0, 0, None, None) == _lib.UA_OK


def test_norm_squared_noalloc_rejects_sizes_beyond_its_scratch(lib):
    """ua_norm_squared_noalloc checks its size limit (batch x ceil(length/4096)
    partials <= 2^16) before it touches any device: sizes beyond it fail with
    UA_ERR_INVALID_ARGUMENT and a message naming the static scratch."""
    from unitair_amd import _lib
    bad = _lib.UA_ERR_INVALID_ARGUMENT
    c64 = _lib.UA_COMPLEX64
    f64 = _lib.UA_FLOAT64
    assert lib.ua_norm_squared_noalloc(1, f64, 1, c64, 17, 2 ** 24, 2 ** 24, None) == bad
    assert b"static scratch" in lib.ua_last_error()
    assert lib.ua_norm_squared_noalloc(1, f64, 1, c64, 1, 2 ** 28 + 1, 0, None) == bad
    assert lib.ua_norm_squared_noalloc(1, 99, 1, c64, 1, 16, 16, None) == bad
    assert lib.ua_norm_squared_noalloc(1, f64, 1, c64, -1, 16, 16, None) == bad
    assert lib.ua_norm_squared_noalloc(None, f64, None, c64, 0, 16, 16, None) == _lib.UA_OK
