"""sqdk_dots_noalloc (no workspace, no allocation: the partial sums live in
the library's static device array) against sqdk_dots with a caller
workspace, through the C ABI: the same kernels in the same order, so the
results must be bit-identical, on two calls in a row, the second smaller."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pairs(vecs):
    arr = (C.c_void_p * len(vecs))(*[v.data_ptr() for v in vecs])
    return arr


def _dots_ws(L, lib, n, prec, a, b, stream):
    npairs = len(a)
    nb = C.c_size_t(0)
    lib.check(L.sqdk_dots_workspace(n, npairs, C.byref(nb)), "sqdk_dots_workspace")
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    out = torch.empty(2 * npairs, dtype=torch.float64, device="cuda")
    lib.check(L.sqdk_dots(n, prec, npairs, _pairs(a), _pairs(b), out.data_ptr(),
                          work.data_ptr(), nb.value, stream), "sqdk_dots")
    return out


def _dots_noalloc(L, lib, n, prec, a, b, stream):
    out = torch.empty(2 * len(a), dtype=torch.float64, device="cuda")
    lib.check(L.sqdk_dots_noalloc(n, prec, len(a), _pairs(a), _pairs(b), out.data_ptr(),
                                  stream), "sqdk_dots_noalloc")
    return out


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_dots_noalloc_matches_workspace_entry_point(gpu, precision):
    from rqutils_amd import _lib
    L = _lib.lib()
    prec = _lib.C128 if precision == "c128" else _lib.C64
    dtype = torch.complex128 if precision == "c128" else torch.complex64
    stream = _lib.stream_ptr()
    g = torch.Generator(device="cpu").manual_seed(7)
    # a large call with all 16 pairs, then a smaller one with fewer pairs: the
    # second must not see anything the first left in the static array
    for n, npairs in ((3_000_001, 16), (1_000, 3)):
        a = [torch.randn(n, dtype=dtype, generator=g).cuda() for _ in range(npairs)]
        b = [torch.randn(n, dtype=dtype, generator=g).cuda() for _ in range(npairs)]
        got = _dots_noalloc(L, _lib, n, prec, a, b, stream)
        want = _dots_ws(L, _lib, n, prec, a, b, stream)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (n, npairs)
        # and against the plain sums in FP64
        ref = torch.stack([torch.vdot(x.to(torch.complex128), y.to(torch.complex128))
                           for x, y in zip(a, b)])
        got_c = got.view(-1, 2)
        assert torch.allclose(got_c[:, 0], ref.real, rtol=1e-9, atol=1e-9)
        assert torch.allclose(got_c[:, 1], ref.imag, rtol=1e-9, atol=1e-9)


def test_dots_noalloc_rejects_more_pairs_than_the_scratch_holds(gpu):
    from rqutils_amd import _lib
    L = _lib.lib()
    a = [torch.zeros(8, dtype=torch.complex128, device="cuda") for _ in range(17)]
    out = torch.empty(34, dtype=torch.float64, device="cuda")
    rc = L.sqdk_dots_noalloc(8, _lib.C128, 17, _pairs(a), _pairs(a), out.data_ptr(),
                             _lib.stream_ptr())
    assert rc == _lib.EINVAL
    assert b"static scratch" in L.sqdk_last_error()
