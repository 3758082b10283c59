"""CPU checks of the compiled fused passes (ua_compile_fused_pass): every pass
of the 24-qubit benchmark circuit generates and compiles with hiprtc for
gfx950 (no GPU is needed to compile), identical structures share one
compiled kernel, and the argument checks hold.  Results are checked on the
GPU (tests/test_gpu_fused.py)."""
import concurrent.futures
import ctypes

import pytest

from unitair_amd import _lib
from unitair_amd.fusion import FusedOp, _items_from_circuit, _merge_1q, _pass_structs
from unitair_amd.workloads import random_circuit_24q

from test_fusion_plan import random_mixed_circuit


def _structs(circ, n, tile_bits=12):
    steps, ops, pending = _pass_structs(_merge_1q(_items_from_circuit(circ)), n, tile_bits)
    for idx, _ in pending:
        ops[idx].data = 0x10000        # any aligned non-NULL address: nothing runs
    arr = (FusedOp * max(1, len(ops)))(*ops)
    passes = [(st[1], ctypes.addressof(arr) + st[2] * ctypes.sizeof(FusedOp))
              for st in steps if st[0] == "fused"]
    return passes, arr


def _compile(lib, code, n, fp, host):
    h = ctypes.c_void_p()
    rc = lib.ua_compile_fused_pass(code, n, ctypes.byref(fp), ctypes.c_void_p(host),
                                   ctypes.byref(h))
    return rc, h.value


@pytest.mark.parametrize("code,tile_bits", [(_lib.UA_COMPLEX64, 12), (_lib.UA_COMPLEX128, 12),
                                           (_lib.UA_COMPLEX64, 14), (_lib.UA_COMPLEX128, 13)])
def test_every_pass_of_the_24q_circuit_compiles(code, tile_bits):
    lib = _lib.load()
    n = 24
    passes, arr = _structs(random_circuit_24q(seed=0, n=n), n, tile_bits)
    assert len(passes) >= 8 and all(fp.tile_bits == tile_bits for fp, _ in passes)
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        res = list(ex.map(lambda p: _compile(lib, code, n, *p), passes))
    for rc, h in res:
        _lib.check(rc, "compile")
        assert h
    # the same structure again: the cached kernel, no new compile
    rc, h0 = _compile(lib, code, n, *passes[0])
    assert rc == 0 and h0 == res[0][1]


def test_source_names_the_pass_structure():
    lib = _lib.load()
    n = 16
    passes, arr = _structs(random_mixed_circuit(n, 40, seed=3), n)
    fp, host = passes[0]
    length = ctypes.c_int64()
    _lib.check(lib.ua_fused_pass_source(_lib.UA_COMPLEX64, n, ctypes.byref(fp),
                                        ctypes.c_void_p(host), None, 0, ctypes.byref(length)),
               "source")
    buf = ctypes.create_string_buffer(length.value + 1)
    _lib.check(lib.ua_fused_pass_source(_lib.UA_COMPLEX64, n, ctypes.byref(fp),
                                        ctypes.c_void_p(host), buf, length.value + 1,
                                        ctypes.byref(length)), "source")
    src = buf.value.decode()
    assert len(src) == length.value
    assert 'extern "C" __global__' in src and "ua_compiled_pass" in src
    assert "#define UA_DOUBLE 0" in src
    # a truncated copy is NUL-terminated at capacity - 1
    small = ctypes.create_string_buffer(16)
    _lib.check(lib.ua_fused_pass_source(_lib.UA_COMPLEX64, n, ctypes.byref(fp),
                                        ctypes.c_void_p(host), small, 16,
                                        ctypes.byref(length)), "source")
    assert small.value == src[:15].encode()


def test_argument_checks():
    lib = _lib.load()
    # fewer than 12 tile bits: the interpreter's case only
    n = 10
    passes, arr = _structs(random_mixed_circuit(n, 20, seed=1), n)
    fp, host = passes[0]
    rc, h = _compile(lib, _lib.UA_COMPLEX64, n, fp, host)
    assert rc == _lib.UA_ERR_UNSUPPORTED and not h
    with pytest.raises(NotImplementedError, match="tile_bits of 12..14"):
        _lib.check(rc, "compile")
    # above the largest tile: 14 bits for complex128
    n = 16
    passes, arr16 = _structs(random_mixed_circuit(n, 30, seed=4), n, 14)
    rc, _ = _compile(lib, _lib.UA_COMPLEX128, n, *passes[0])
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    # the interpreter's validation applies (here: a float dtype)
    n = 14
    passes, arr = _structs(random_mixed_circuit(n, 20, seed=2), n)
    fp, host = passes[0]
    rc, _ = _compile(lib, _lib.UA_FLOAT32, n, fp, host)
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    # a CZ of a qubit with itself is refused
    bad = _pass_with_self_cz(n)
    rc, _ = _compile(lib, _lib.UA_COMPLEX64, n, *bad[:2])
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    assert b"bad bits" in lib.ua_last_error()
    # run: NULL handle, then batch 0 returns before touching a device
    assert lib.ua_run_compiled_pass(None, None, 1, None, None) == _lib.UA_ERR_INVALID_ARGUMENT
    rc, h = _compile(lib, _lib.UA_COMPLEX64, n, fp, host)
    _lib.check(rc, "compile")
    assert lib.ua_run_compiled_pass(ctypes.c_void_p(h), None, 0, None, None) == _lib.UA_OK
    assert lib.ua_run_compiled_pass(ctypes.c_void_p(h), None, -1, None, None) == \
        _lib.UA_ERR_INVALID_ARGUMENT


def _pass_with_self_cz(n):
    fp = _lib.FusedPass()
    fp.tile_bits = 12
    for i in range(12):
        fp.local_bits[i] = i
    fp.num_stages = 1
    for k in range(4):
        fp.stage_bits[0][k] = k
    fp.stage_ops[0], fp.stage_ops[1] = 0, 1
    op = FusedOp()
    op.kind, op.a, op.b = 3, 5, 5          # UA_FOP_CZ on state bits 5 and 5
    arr = (FusedOp * 1)(op)
    return fp, ctypes.addressof(arr), arr


def test_compile_in_background_fills_the_table():
    """compile_in_background plans with placeholder data and compiles on a
    host thread (no GPU); the table then holds every 12-bit pass of the plan
    made from the real items, under the keys FusedProgram looks up."""
    import numpy as np
    from unitair_amd import fusion
    from unitair_amd.fusion import PlanTemplate, _Item, _structure_key
    n = 13
    rng = np.random.default_rng(0)
    items = []
    for layer in range(3):
        for q in range(n):
            items.append(_Item("dense1", (n - 1 - q,), rng.standard_normal((2, 2)), False))
        items.append(_Item("cz_layer", (), np.array([0b1010101], dtype=np.int64), False))
    done = fusion.compile_in_background(items, n, reorder=False).result(timeout=300)
    assert done >= 1
    t = PlanTemplate(items, n, reorder=False)
    fused = [st for st in t.steps if st[0] == "fused" and st[1].tile_bits == 12]
    assert fused
    for _, fp, first, count in fused:
        arr = (FusedOp * count)(*t.op_structs[first:first + count])
        assert _structure_key(_lib.UA_COMPLEX64, n, fp, ctypes.addressof(arr)) in fusion._COMPILED
    # a second request for the same structure compiles nothing
    assert fusion.compile_in_background(items, n, reorder=False).result(timeout=300) == 0
