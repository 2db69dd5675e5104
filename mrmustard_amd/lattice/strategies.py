"""GPU strategies of the Bargmann->Fock path (names follow the reference's
``mrmustard.math.lattice.strategies``, with ``_hip`` for ``_numba``).

Each function takes torch complex128 tensors on a ROCm device and calls the C
ABI of libmmhip.so on the current stream:

- ``vanilla_hip``            <- vanilla/core.py:25-124  ``vanilla_numba``
- ``stable_hip``             <- vanilla/core.py:127-213 ``stable_numba``
- ``vanilla_batch_hip``      <- vanilla/batch.py:27-61  ``vanilla_batch_numba``
- ``vanilla_vjp_hip``        <- vanilla/gradients.py:25-82  ``vanilla_vjp_numba``
- ``vanilla_batch_vjp_hip``  <- vanilla/gradients.py:85-116 ``vanilla_batch_vjp_numba``
- ``vanilla_fused_hip`` / ``vanilla_batch_fused_hip``: forward and VJP in one
  sweep when the cotangent is known up front (no reference counterpart; it
  equals the two calls in sequence).
- ``vanilla_vjp_recompute_hip``: the VJP from (A, b, c) without G, recomputing
  the lattice slab by slab in a bounded window.
- ``binomial_hip``           <- binomial.py:30-74  ``binomial`` (fill by weight,
  stop on the squared norm).

The VJPs use the reference's holomorphic convention: ``dLdG`` is not
conjugated, and dA is returned symmetrised, (M + M^T)/2.
"""

from __future__ import annotations

import ctypes
import math as pymath

import numpy as np
import torch

from .. import _lib

__all__ = [
    "binomial_hip",
    "stable_hip",
    "vanilla_vjp_recompute_hip",
    "vanilla_batch_fused_hip",
    "vanilla_batch_hip",
    "vanilla_batch_vjp_hip",
    "vanilla_fused_hip",
    "vanilla_hip",
    "vanilla_vjp_hip",
]

C128 = torch.complex128


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("mrmustard_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def as_device_c128(x) -> torch.Tensor:
    """complex128, contiguous, on the current GPU (copies only when needed)."""
    dev = _device()
    if isinstance(x, torch.Tensor):
        t = x
        if t.dtype != C128:
            t = t.to(C128)
        if t.device != dev:
            t = t.to(dev)
        if t.is_conj():
            t = t.resolve_conj()
        if t.is_neg():
            t = t.resolve_neg()
        return t.contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.complex128), device=dev).contiguous()


def _shape_arr(shape) -> np.ndarray:
    return np.ascontiguousarray(np.asarray([int(s) for s in shape], dtype=np.int64))


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _workspace(nbytes: int) -> torch.Tensor | None:
    if nbytes <= 0:
        return None
    return torch.empty(int(nbytes), dtype=torch.uint8, device=_device())


def _check_shape(shape, D: int):
    shape = tuple(int(s) for s in shape)
    if len(shape) != D:
        raise ValueError(f"shape {shape} has {len(shape)} indices but b has {D} entries")
    if any(s < 1 for s in shape):
        raise ValueError(f"shape {shape} must contain positive integers")
    return shape


def _out_tensor(out, full_shape) -> torch.Tensor:
    if out is None:
        return torch.empty(full_shape, dtype=C128, device=_device())
    if not isinstance(out, torch.Tensor) or out.dtype != C128 or not out.is_contiguous() or out.device != _device():
        raise ValueError("out must be a contiguous complex128 tensor on the current GPU")
    if tuple(out.shape) != tuple(full_shape):
        if out.numel() != pymath.prod(full_shape):
            raise ValueError(f"out of shape {tuple(out.shape)} cannot hold an array of shape {tuple(full_shape)}")
    return out


def _fwd(shape, A, b, c, stable: bool, out=None) -> torch.Tensor:
    L = _lib.load()
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c)
    D = int(b.shape[-1])
    if tuple(A.shape) != (D, D) or c.numel() != 1:
        raise ValueError(f"A must be ({D}, {D}) and c a scalar; got A{tuple(A.shape)}, c{tuple(c.shape)}")
    shape = _check_shape(shape, D)
    G = _out_tensor(out, shape)
    s = _shape_arr(shape)
    ws = _workspace(L.mm_hermite_fwd_workspace(s.ctypes.data, D, int(stable)))
    rc = L.mm_hermite_fwd(
        _ptr(A), _ptr(b), _ptr(c), s.ctypes.data, D, int(stable), _ptr(G), _ptr(ws), 0 if ws is None else ws.numel(), _stream()
    )
    _lib.check(rc, "mm_hermite_fwd")
    return G


def vanilla_hip(shape, A, b, c, out=None) -> torch.Tensor:
    """First-pivot recurrence (vanilla_numba, core.py:25-124) on the GPU."""
    return _fwd(shape, A, b, c, False, out)


def stable_hip(shape, A, b, c, out=None) -> torch.Tensor:
    """All-pivot average recurrence (stable_numba, core.py:127-213) on the GPU."""
    return _fwd(shape, A, b, c, True, out)


def vanilla_batch_hip(shape, A, b, c, stable: bool = False, out=None) -> torch.Tensor:
    """Batch of triples sharing ``shape`` (vanilla_batch_numba, batch.py:27-61)."""
    L = _lib.load()
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c).reshape(-1)
    B, D = int(b.shape[0]), int(b.shape[-1])
    if tuple(A.shape) != (B, D, D) or c.numel() != B:
        raise ValueError(f"batched A must be ({B}, {D}, {D}) and c ({B},)")
    shape = _check_shape(shape, D)
    G = _out_tensor(out, (B, *shape))
    if B == 0:
        return G
    s = _shape_arr(shape)
    ws = _workspace(L.mm_hermite_batch_fwd_workspace(B, s.ctypes.data, D, int(stable)))
    rc = L.mm_hermite_batch_fwd(
        _ptr(A), _ptr(b), _ptr(c), B, s.ctypes.data, D, int(stable), _ptr(G), _ptr(ws), 0 if ws is None else ws.numel(), _stream()
    )
    _lib.check(rc, "mm_hermite_batch_fwd")
    return G


def vanilla_vjp_hip(G, c, dLdG):
    """(dL/dA, dL/db, dL/dc) of the forward (vanilla_vjp_numba, gradients.py:25-82)."""
    L = _lib.load()
    G, g, c = as_device_c128(G), as_device_c128(dLdG), as_device_c128(c).reshape(())
    if tuple(g.shape) != tuple(G.shape):
        raise ValueError("dLdG must have the shape of G")
    D = G.dim()
    s = _shape_arr(G.shape)
    dA = torch.empty((D, D), dtype=C128, device=G.device)
    db = torch.empty((D,), dtype=C128, device=G.device)
    dc = torch.empty((), dtype=C128, device=G.device)
    ws = _workspace(L.mm_hermite_vjp_workspace(s.ctypes.data, D))
    rc = L.mm_hermite_vjp(
        _ptr(G), _ptr(c), _ptr(g), s.ctypes.data, D, _ptr(dA), _ptr(db), _ptr(dc), _ptr(ws), 0 if ws is None else ws.numel(), _stream()
    )
    _lib.check(rc, "mm_hermite_vjp")
    return dA, db, dc


def vanilla_batch_vjp_hip(G, c, dLdG):
    """Batched VJP (vanilla_batch_vjp_numba, gradients.py:85-116)."""
    L = _lib.load()
    G, g, c = as_device_c128(G), as_device_c128(dLdG), as_device_c128(c).reshape(-1)
    if tuple(g.shape) != tuple(G.shape):
        raise ValueError("dLdG must have the shape of G")
    B, D = int(G.shape[0]), G.dim() - 1
    s = _shape_arr(G.shape[1:])
    dA = torch.empty((B, D, D), dtype=C128, device=G.device)
    db = torch.empty((B, D), dtype=C128, device=G.device)
    dc = torch.empty((B,), dtype=C128, device=G.device)
    if B == 0:
        return dA, db, dc
    ws = _workspace(L.mm_hermite_batch_vjp_workspace(B, s.ctypes.data, D))
    rc = L.mm_hermite_batch_vjp(
        _ptr(G), _ptr(c), _ptr(g), B, s.ctypes.data, D, _ptr(dA), _ptr(db), _ptr(dc), _ptr(ws), 0 if ws is None else ws.numel(),
        _stream(),
    )
    _lib.check(rc, "mm_hermite_batch_vjp")
    return dA, db, dc


def vanilla_fused_hip(shape, A, b, c, dLdG, out=None):
    """G and its VJP in one sweep: (G, dA, db, dc)."""
    L = _lib.load()
    A, b, c, g = as_device_c128(A), as_device_c128(b), as_device_c128(c), as_device_c128(dLdG)
    D = int(b.shape[-1])
    shape = _check_shape(shape, D)
    if tuple(g.shape) != shape:
        raise ValueError("dLdG must have the requested shape")
    G = _out_tensor(out, shape)
    s = _shape_arr(shape)
    dA = torch.empty((D, D), dtype=C128, device=G.device)
    db = torch.empty((D,), dtype=C128, device=G.device)
    dc = torch.empty((), dtype=C128, device=G.device)
    ws = _workspace(L.mm_hermite_fused_workspace(s.ctypes.data, D))
    rc = L.mm_hermite_fused(
        _ptr(A), _ptr(b), _ptr(c), _ptr(g), s.ctypes.data, D, _ptr(G), _ptr(dA), _ptr(db), _ptr(dc), _ptr(ws),
        0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_fused")
    return G, dA, db, dc


def vanilla_batch_fused_hip(shape, A, b, c, dLdG, out=None):
    """Batched (uniform shape) forward + VJP in one sweep: (G, dA, db, dc)."""
    L = _lib.load()
    A, b, c, g = as_device_c128(A), as_device_c128(b), as_device_c128(c).reshape(-1), as_device_c128(dLdG)
    B, D = int(b.shape[0]), int(b.shape[-1])
    shape = _check_shape(shape, D)
    if tuple(g.shape) != (B, *shape):
        raise ValueError("dLdG must have shape (B, *shape)")
    G = _out_tensor(out, (B, *shape))
    s = _shape_arr(shape)
    dA = torch.empty((B, D, D), dtype=C128, device=G.device)
    db = torch.empty((B, D), dtype=C128, device=G.device)
    dc = torch.empty((B,), dtype=C128, device=G.device)
    if B == 0:
        return G, dA, db, dc
    ws = _workspace(L.mm_hermite_batch_fused_workspace(B, s.ctypes.data, D))
    rc = L.mm_hermite_batch_fused(
        _ptr(A), _ptr(b), _ptr(c), _ptr(g), B, s.ctypes.data, D, _ptr(G), _ptr(dA), _ptr(db), _ptr(dc), _ptr(ws),
        0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_batch_fused")
    return G, dA, db, dc


def vanilla_vjp_recompute_hip(shape, A, b, c, dLdG):
    """(dL/dA, dL/db, dL/dc) of the forward without G: the lattice is recomputed
    slab by slab in a three-slab window (mm_hermite_vjp_recompute).  Equal to
    ``vanilla_vjp_hip(vanilla_hip(shape, A, b, c), c, dLdG)``.

    The windowed sweep does not rerun an overflowing triple in the reference's
    arithmetic, but any inf or NaN in G or dL/dG makes dL/dc non-finite (it is
    the sum of G * dL/dG over the lattice, divided by c).  The call therefore
    reads dL/dc back (one stream synchronisation) and, when it is not finite,
    takes the stored-lattice path, whose results are the reference's values,
    NaN and +-inf included; only then is the whole lattice held."""
    L = _lib.load()
    A, b, c, g = as_device_c128(A), as_device_c128(b), as_device_c128(c), as_device_c128(dLdG)
    D = int(b.shape[-1])
    shape = _check_shape(shape, D)
    if tuple(g.shape) != shape:
        raise ValueError("dLdG must have the requested shape")
    s = _shape_arr(shape)
    dA = torch.empty((D, D), dtype=C128, device=g.device)
    db = torch.empty((D,), dtype=C128, device=g.device)
    dc = torch.empty((), dtype=C128, device=g.device)
    ws = _workspace(L.mm_hermite_vjp_recompute_workspace(s.ctypes.data, D))
    rc = L.mm_hermite_vjp_recompute(
        _ptr(A), _ptr(b), _ptr(c), _ptr(g), s.ctypes.data, D, _ptr(dA), _ptr(db), _ptr(dc), _ptr(ws),
        0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_vjp_recompute")
    if not bool(torch.isfinite(dc)):
        return vanilla_vjp_hip(vanilla_hip(shape, A, b, c), c, g)
    return dA, db, dc


def binomial_hip(local_cutoffs, A, b, c, max_l2, global_cutoff):
    """``binomial`` (binomial.py:30-74): (G, norm).  G has shape
    ``local_cutoffs`` and holds the amplitudes of total photon number up to the
    first weight at which |c|^2 plus the squared norms of weights 1..w exceed
    ``max_l2`` (None: never), or up to ``global_cutoff - 1``; zeros beyond.
    ``norm`` is that accumulated squared norm (a float)."""
    shape = tuple(int(s) for s in local_cutoffs)
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c)
    D = len(shape)
    if tuple(A.shape) != (D, D) or b.numel() != D or c.numel() != 1:
        raise ValueError(f"A must be ({D}, {D}), b ({D},) and c scalar for shape {shape}")
    G = torch.empty(shape, dtype=C128, device=A.device)
    norm = torch.empty((), dtype=torch.float64, device=A.device)
    s = _shape_arr(shape)
    L = _lib.load()
    ws = _workspace(L.mm_hermite_binomial_workspace(s.ctypes.data, D))
    rc = L.mm_hermite_binomial(
        _ptr(A), _ptr(b), _ptr(c), s.ctypes.data, D, float("inf") if max_l2 is None else float(max_l2), int(global_cutoff),
        _ptr(G), _ptr(norm), _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_binomial")
    return G, float(norm)
