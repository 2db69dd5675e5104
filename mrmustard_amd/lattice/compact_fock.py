"""Photon-number-resolved amplitudes without the full lattice (the reference's
``mrmustard.math.lattice.strategies.compactFock``), on the GPU.

- ``diagonal_hip`` <- compactFock/inputValidation.py:61-79
  ``hermite_multidimensional_diagonal`` (-> diagonal_amps.py:200-248): the
  diagonal G[k0, k0, k1, k1, ...] of a 2M-index triple, shape ``cutoffs``
  (``cutoffs + (nb,)`` for a batch of B's).
- ``leftover_hip`` <- compactFock/inputValidation.py:103-118
  ``hermite_multidimensional_1leftoverMode`` (-> singleLeftoverMode_amps.py
  :369-421), as called by hermite_renormalized_1leftoverMode_jax
  (jax_vjps/hermite.py:343-387): the conditional density matrices of mode 0,
  shape ``(output_cutoff + 1, output_cutoff + 1, *(pnr_cutoffs + 1))``.
"""

from __future__ import annotations

from collections.abc import Iterable

import numpy as np
import torch

from .. import _lib
from .strategies import C128, _ptr, _shape_arr, _stream, _workspace, as_device_c128

__all__ = ["diagonal_hip", "diagonal_vjp_hip", "input_validation", "leftover_hip", "leftover_vjp_hip"]


def input_validation(A, rtol: float = 1e-05, atol: float = 1e-08) -> bool:
    """A square, NaN-free and symmetric within (rtol, atol) (inputValidation.py:25-58).
    A is at most 20 x 20: one copy to the host, then numpy's own checks."""
    Ah = A.detach().cpu().numpy() if isinstance(A, torch.Tensor) else np.asarray(A)
    if Ah.ndim != 2 or Ah.shape[0] != Ah.shape[1]:
        raise ValueError("Input matrix must be square.")
    if np.isnan(Ah).any():
        raise ValueError("Input matrix must not contain NaNs.")
    if not np.allclose(Ah, Ah.T, rtol=rtol, atol=atol):
        raise ValueError("Input matrix must be symmetric.")
    return True


def diagonal_hip(A, B, G0, cutoffs, reordered: bool = False, rtol: float = 1e-05, atol: float = 1e-08) -> torch.Tensor:
    """The photon-number diagonal of the lattice of (A, B, G0).

    ``A`` (2M x 2M) and ``B`` (2M, or 2M x nb for a batch of B's sharing A) are
    in the [mode0, mode0, mode1, mode1, ...] order unless ``reordered``, which
    takes them in the [modes..., modes...] order and reorders on the device
    (backend_numpy.py:367-375).  ``G0`` is a scalar, or nb values with a batch.
    Raises ValueError as inputValidation.py:61-79 does."""
    A, B, G0 = as_device_c128(A), as_device_c128(B), as_device_c128(G0)
    input_validation(A, rtol=rtol, atol=atol)
    if B.ndim > 2:
        raise ValueError("B should be either unbactched or two dimensional (vector and batch dimension)")
    if A.shape[0] != B.shape[0]:
        raise ValueError("The matrix A and vector B have incompatible dimensions")
    if not isinstance(cutoffs, Iterable):
        raise ValueError("cutoffs should be array like of length M")
    cutoffs = tuple(int(k) for k in cutoffs)
    M = len(cutoffs)
    if A.shape[0] // 2 != M or A.shape[0] != 2 * M:
        raise ValueError("The matrix A and cutoffs have incompatible dimensions")
    nb = 1 if B.ndim == 1 else int(B.shape[1])
    if G0.numel() not in (1, nb):
        raise ValueError(f"G0 must hold 1 or {nb} values, got {G0.numel()}")
    out_shape = cutoffs if B.ndim == 1 else cutoffs + (nb,)
    G = torch.empty(out_shape, dtype=C128, device=A.device)
    if nb == 0:
        return G
    s = _shape_arr(cutoffs)
    L = _lib.load()
    ws = _workspace(L.mm_hermite_diagonal_workspace(s.ctypes.data, M, nb))
    rc = L.mm_hermite_diagonal(
        _ptr(A), _ptr(B), _ptr(G0), int(G0.numel() > 1), s.ctypes.data, M, nb, int(bool(reordered)),
        _ptr(G), _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_diagonal")
    return G


def diagonal_vjp_hip(A, B, G0, cutoffs, dLdG, reordered: bool = False):
    """(G, dA, dB, dG0): the diagonal strategy's forward and its holomorphic VJP
    for the cotangent ``dLdG`` (shape ``cutoffs``, not conjugated), every entry
    of A independent, in the order A and B are given in.  The reference forms
    the Jacobians (compactFock/diagonal_grad.py) and contracts them with the
    cotangent (jax_vjps/hermite.py:292-330); this computes the contraction
    directly with a reverse sweep.  Unbatched B only."""
    A, B, G0 = as_device_c128(A), as_device_c128(B), as_device_c128(G0)
    cutoffs = tuple(int(k) for k in cutoffs)
    M = len(cutoffs)
    if tuple(A.shape) != (2 * M, 2 * M) or tuple(B.shape) != (2 * M,) or G0.numel() != 1:
        raise ValueError(f"A must be ({2 * M}, {2 * M}), B ({2 * M},) and G0 scalar for cutoffs {cutoffs}")
    g = as_device_c128(dLdG)
    if tuple(g.shape) != cutoffs:
        raise ValueError(f"dLdG has shape {tuple(g.shape)}, expected {cutoffs}")
    G = torch.empty(cutoffs, dtype=C128, device=A.device)
    dA = torch.empty((2 * M, 2 * M), dtype=C128, device=A.device)
    dB = torch.empty((2 * M,), dtype=C128, device=A.device)
    dc = torch.empty((), dtype=C128, device=A.device)
    s = _shape_arr(cutoffs)
    L = _lib.load()
    ws = _workspace(L.mm_hermite_diagonal_vjp_workspace(s.ctypes.data, M))
    rc = L.mm_hermite_diagonal_vjp(
        _ptr(A), _ptr(B), _ptr(G0), s.ctypes.data, M, int(bool(reordered)), _ptr(g), _ptr(G), _ptr(dA), _ptr(dB), _ptr(dc),
        _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_diagonal_vjp")
    return G, dA, dB, dc


def leftover_hip(A, b, c, output_cutoff: int, pnr_cutoffs, reordered: bool = True, rtol: float = 1e-05,
                 atol: float = 1e-08) -> torch.Tensor:
    """out[m, n, k] = G[(m, k), (n, k)]: mode 0's density matrix (m, n <=
    ``output_cutoff``) with the other modes detected with k_j <= pnr_cutoffs[j]
    photons.  (A, b, c) in the [modes..., modes...] order when ``reordered``
    (reorder_AB_bargmann, backend_numpy.py:367-375), else [mode0, mode0, mode1,
    mode1, ...].  Raises ValueError as inputValidation.py:103-118 does."""
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c)
    input_validation(A, rtol=rtol, atol=atol)
    if A.shape[0] != b.shape[0]:
        raise ValueError("The matrix A and vector B have incompatible dimensions")
    if not isinstance(pnr_cutoffs, Iterable):
        raise ValueError("cutoffs should be array like of length M")
    pnr = tuple(int(k) for k in pnr_cutoffs)
    M = len(pnr) + 1
    if A.shape[0] // 2 != M or A.shape[0] != 2 * M:
        raise ValueError("The matrix A and cutoffs have incompatible dimensions")
    if M <= 1:
        raise ValueError("The number of modes should be greater than 1.")
    if c.numel() != 1:
        raise ValueError("c must be a scalar")
    o = int(output_cutoff)
    out = torch.empty((o + 1, o + 1, *(k + 1 for k in pnr)), dtype=C128, device=A.device)
    s = _shape_arr(pnr)
    L = _lib.load()
    ws = _workspace(L.mm_hermite_1leftover_workspace(o, s.ctypes.data, M - 1))
    rc = L.mm_hermite_1leftover(
        _ptr(A), _ptr(b), _ptr(c), o, s.ctypes.data, M - 1, int(bool(reordered)),
        _ptr(out), _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_1leftover")
    return out


def leftover_vjp_hip(A, b, c, output_cutoff: int, pnr_cutoffs, dLdout, reordered: bool = True):
    """(out, dA, db, dc): the one-leftover-mode forward and its holomorphic VJP
    for the cotangent ``dLdout`` (shape of out, not conjugated), every entry of
    A independent, in the order A and b are given in.  The reference forms the
    Jacobians (compactFock/singleLeftoverMode_grad.py:695-724, via
    inputValidation.py:121-141) and contracts them with the cotangent
    (jax_vjps/hermite.py:397-440); this computes the contraction directly with
    a reverse sweep."""
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c)
    pnr = tuple(int(k) for k in pnr_cutoffs)
    M = len(pnr) + 1
    if M <= 1:
        raise ValueError("The number of modes should be greater than 1.")
    if tuple(A.shape) != (2 * M, 2 * M) or tuple(b.shape) != (2 * M,) or c.numel() != 1:
        raise ValueError(f"A must be ({2 * M}, {2 * M}), b ({2 * M},) and c scalar for {M} modes")
    o = int(output_cutoff)
    shape = (o + 1, o + 1, *(k + 1 for k in pnr))
    g = as_device_c128(dLdout)
    if tuple(g.shape) != shape:
        raise ValueError(f"dLdout has shape {tuple(g.shape)}, expected {shape}")
    out = torch.empty(shape, dtype=C128, device=A.device)
    dA = torch.empty((2 * M, 2 * M), dtype=C128, device=A.device)
    db = torch.empty((2 * M,), dtype=C128, device=A.device)
    dc = torch.empty((), dtype=C128, device=A.device)
    s = _shape_arr(pnr)
    L = _lib.load()
    ws = _workspace(L.mm_hermite_1leftover_vjp_workspace(o, s.ctypes.data, M - 1))
    rc = L.mm_hermite_1leftover_vjp(
        _ptr(A), _ptr(b), _ptr(c), o, s.ctypes.data, M - 1, int(bool(reordered)), _ptr(g), _ptr(out), _ptr(dA), _ptr(db),
        _ptr(dc), _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_1leftover_vjp")
    return out, dA, db, dc
