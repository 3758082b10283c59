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
