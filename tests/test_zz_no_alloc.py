"""ua_norm_squared_noalloc through the C ABI against ua_norm_squared (the
workspace entry point): the same kernels, so the same bits, on two calls in
a row (the second on a smaller input, which reuses the static partials the
first one wrote), at the maximum size, and on a row-strided view."""
import pytest
import torch

from unitair_amd import _lib

pytestmark = pytest.mark.gpu


def _norms(lib, st, rows, length, stride, noalloc):
    dev = st.device
    out = torch.full((rows,), float("nan"), dtype=torch.float64, device=dev)
    code = _lib.DTYPE_CODES[st.dtype]
    stream = _lib.stream_handle(dev)
    if noalloc:
        rc = lib.ua_norm_squared_noalloc(_lib.ptr(out), _lib.UA_FLOAT64, _lib.ptr(st), code,
                                         rows, length, stride, stream)
    else:
        ws = torch.empty(lib.ua_reduce_workspace_bytes(rows, length), dtype=torch.uint8,
                         device=dev)
        rc = lib.ua_norm_squared(_lib.ptr(out), _lib.UA_FLOAT64, _lib.ptr(st), code, rows,
                                 length, stride, _lib.ptr(ws), stream)
    _lib.check(rc, "norm_squared")
    return out


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128, torch.float32])
def test_noalloc_equals_workspace_entry_point(device, dt):
    lib = _lib.load()
    torch.manual_seed(0)
    for rows, n in ((8, 20), (3, 13), (1, 5)):     # each call smaller than the one before
        st = torch.randn(rows, 2 ** n, dtype=dt, device=device)
        got = _norms(lib, st, rows, 2 ** n, 2 ** n, noalloc=True)
        ref = _norms(lib, st, rows, 2 ** n, 2 ** n, noalloc=False)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (rows, n)
        exact = (st.to(torch.complex128).abs() ** 2).sum(-1) if dt.is_complex else \
            (st.double() ** 2).sum(-1)
        assert torch.allclose(got, exact.real.double(), rtol=1e-12)


def test_noalloc_at_its_maximum_and_on_a_strided_view(device):
    lib = _lib.load()
    torch.manual_seed(1)
    # 4 x 2^24 complex64 elements: 4 x 4096 = 2^14 partials; then a
    # batch x chunks product exactly at the 2^16 limit, via a row stride
    base = torch.randn(16, 2 ** 24, dtype=torch.complex64, device=device)
    for rows, length, stride in ((4, 2 ** 24, 2 ** 24), (16, 2 ** 24, 2 ** 24),
                                 (8, 2 ** 23, 2 ** 25)):
        got = _norms(lib, base, rows, length, stride, noalloc=True)
        ref = _norms(lib, base, rows, length, stride, noalloc=False)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (rows, length, stride)
    rc = lib.ua_norm_squared_noalloc(_lib.ptr(got), _lib.UA_FLOAT64, _lib.ptr(base),
                                     _lib.UA_COMPLEX64, 17, 2 ** 24, 2 ** 24,
                                     _lib.stream_handle(device))
    assert rc == _lib.UA_ERR_INVALID_ARGUMENT
    assert b"static scratch" in lib.ua_last_error()
