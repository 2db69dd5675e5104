"""The no-allocation VJP entry point (mm_hermite_batch_vjp_static), called
through the C ABI: bit-identical to mm_hermite_batch_vjp (same kernels, same
partial-sum layout) on two calls in a row, the second with a smaller input,
so that the static buffer's leftover contents from the first call must not
leak into the second; a request above the buffer's size is refused on the
host with MM_ERR_UNSUPPORTED."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _call(L, fn, G, c, g, shape, D, ws=None):
    import torch

    B = G.shape[0]
    dA = torch.empty((B, D, D), dtype=torch.complex128, device=G.device)
    db = torch.empty((B, D), dtype=torch.complex128, device=G.device)
    dc = torch.empty((B,), dtype=torch.complex128, device=G.device)
    s = np.asarray(shape, np.int64)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    if ws is None:
        rc = fn(ptr(G), ptr(c), ptr(g), B, s.ctypes.data, D, ptr(dA), ptr(db), ptr(dc), stream)
    else:
        rc = fn(ptr(G), ptr(c), ptr(g), B, s.ctypes.data, D, ptr(dA), ptr(db), ptr(dc), ptr(ws), ws.numel(), stream)
    torch.cuda.synchronize()
    return rc, dA.cpu().numpy(), db.cpu().numpy(), dc.cpu().numpy()


def test_static_vjp_matches_workspace_vjp():
    import torch

    from mrmustard_amd import _lib

    torch.cuda.set_device(0)
    L = _lib.load()
    assert L.mm_init(0) == 0
    rng = np.random.default_rng(31)
    # first a larger input, then a smaller one: the buffer keeps the first call's partial sums
    for B, shape in ((40, (4,) * 8), (6, (9, 8, 7)), (2, (5, 4))):
        D = len(shape)
        print(f"static VJP B={B} shape={shape}", flush=True)
        G = torch.tensor(rng.standard_normal((B, *shape)) + 1j * rng.standard_normal((B, *shape)), device="cuda")
        g = torch.tensor(rng.standard_normal((B, *shape)) + 1j * rng.standard_normal((B, *shape)), device="cuda")
        c = torch.tensor(rng.standard_normal(B) + 1j * rng.standard_normal(B) + 2.0, device="cuda")
        nbytes = L.mm_hermite_batch_vjp_workspace(B, np.asarray(shape, np.int64).ctypes.data, D)
        assert 0 < nbytes <= (8 << 20)
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        rc0, *ref = _call(L, L.mm_hermite_batch_vjp, G, c, g, shape, D, ws=ws)
        rc1, *got = _call(L, L.mm_hermite_batch_vjp_static, G, c, g, shape, D)
        assert rc0 == 0 and rc1 == 0, L.mm_last_error()
        for x, r in zip(got, ref):
            assert np.array_equal(x, r)


def test_static_vjp_refuses_more_than_its_buffer():
    import torch

    from mrmustard_amd import _lib

    torch.cuda.set_device(0)
    L = _lib.load()
    assert L.mm_init(0) == 0
    shape, D, B = (2,) * 8, 8, 20000  # 20000 x 54 partial sums x 16 B > 8 MiB
    assert L.mm_hermite_batch_vjp_workspace(B, np.asarray(shape, np.int64).ctypes.data, D) > (8 << 20)
    t = torch.zeros(4, dtype=torch.complex128, device="cuda")  # never dereferenced: refused on the host
    s = np.asarray(shape, np.int64)
    p = ctypes.c_void_p(t.data_ptr())
    rc = L.mm_hermite_batch_vjp_static(p, p, p, B, s.ctypes.data, D, p, p, p, None)
    assert rc == _lib.MM_ERR_UNSUPPORTED
    assert b"static buffer" in L.mm_last_error()
