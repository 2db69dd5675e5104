"""Fock cutoffs of Gaussian states on the GPU.

``autoshape_hip`` mirrors ``autoshape_numba(A, b, c, max_prob, max_shape,
min_shape)`` (mrmustard/math/lattice/autoshape.py:23-154): A is the 2M x 2M
Bargmann matrix of a density matrix with indices (bra modes, ket modes), b
its 2M vector, c the vacuum amplitude; the result is an int64 array of M
cutoffs.  ``autoshape_batch_hip`` does B states of the same M in one call.
Both call ``mm_autoshape`` of libmmhip.so and copy the M (or B x M) integers
back to the host, as the caller needs them to size the Fock array.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .strategies import as_device_c128

__all__ = ["autoshape_hip", "autoshape_batch_hip"]


def autoshape_batch_hip(A, b, c, max_prob, max_shape, min_shape) -> np.ndarray:
    """Cutoffs of B states: A (B, 2M, 2M), b (B, 2M), c (B,) -> (B, M) int64."""
    A = as_device_c128(A)
    b = as_device_c128(b)
    c = as_device_c128(c).reshape(-1)
    if b.dim() != 2 or b.shape[-1] % 2:
        raise ValueError(f"b must be (B, 2M), got {tuple(b.shape)}")
    B, M = int(b.shape[0]), int(b.shape[-1]) // 2
    if tuple(A.shape) != (B, 2 * M, 2 * M) or c.numel() != B:
        raise ValueError(f"A must be ({B}, {2 * M}, {2 * M}) and c ({B},)")
    out = torch.empty((B, M), dtype=torch.int64, device=A.device)
    L = _lib.load()
    nbytes = int(L.mm_autoshape_workspace(B, M))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=A.device)
    rc = L.mm_autoshape(
        ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(c.data_ptr()), B, M,
        float(max_prob), int(max_shape), int(min_shape), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
        nbytes, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
    )
    _lib.check(rc, "mm_autoshape")
    res = out.cpu().numpy()
    if (res < 0).any():
        raise ValueError("singular matrix in the reduced-state Schur complement")
    return res


def autoshape_hip(A, b, c, max_prob, max_shape, min_shape) -> np.ndarray:
    """Cutoffs of one state: A (2M, 2M), b (2M,), c scalar -> (M,) int64."""
    b_ = np.asarray(b) if not isinstance(b, torch.Tensor) else b
    if len(b_.shape) != 1 or b_.shape[0] % 2:
        raise ValueError(f"b must have an even length 2M, got shape {tuple(b_.shape)}")
    A_ = as_device_c128(A)
    return autoshape_batch_hip(A_.unsqueeze(0), as_device_c128(b).unsqueeze(0), as_device_c128(c).reshape(1),
                               max_prob, max_shape, min_shape)[0]
