"""Photon-number-resolved amplitudes without the full lattice (the reference's
``mrmustard.math.lattice.strategies.compactFock``), on the GPU.

- ``diagonal_hip`` <- compactFock/inputValidation.py:61-79
  ``hermite_multidimensional_diagonal`` (-> diagonal_amps.py:200-248): the
  diagonal G[k0, k0, k1, k1, ...] of a 2M-index triple, shape ``cutoffs``
  (``cutoffs + (nb,)`` for a batch of B's).
"""

from __future__ import annotations

from collections.abc import Iterable

import numpy as np
import torch

from .. import _lib
from .strategies import C128, _ptr, _shape_arr, _stream, _workspace, as_device_c128

__all__ = ["diagonal_hip", "input_validation"]


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
