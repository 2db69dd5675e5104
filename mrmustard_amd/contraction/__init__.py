"""Contraction of Fock-basis tensors on the FP64 matrix cores.

The step after the Bargmann->Fock conversion in a circuit contraction is
``ArrayAnsatz.contract`` -> ``math.einsum`` over the converted arrays
(mrmustard/physics/ansatz/array_ansatz.py:159+,
mrmustard/math/backend_manager.py:526-548).  When the contracted indices can
be grouped, that einsum is one complex matrix product; ``zgemm_hip`` runs it
with ``mm_zgemm`` of libmmhip.so (v_mfma_f64_16x16x4_f64 tiles).

``dm_operator_product`` is that contraction for an n-mode density matrix in
the reference's wire order (all bra indices, then all ket indices;
wires.py:101-106): the (bra multi-index) x (ket multi-index) matrix is
multiplied on the left by an operator on the same number basis.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..lattice.strategies import as_device_c128

__all__ = ["zgemm_hip", "dm_operator_product"]


def zgemm_hip(A, B, out: torch.Tensor | None = None) -> torch.Tensor:
    """C = A @ B for complex128 matrices on the GPU (row-major, any strides
    whose last dimension is contiguous are copied to contiguous first)."""
    A = as_device_c128(A)
    B = as_device_c128(B)
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError(f"zgemm_hip needs matrices, got shapes {tuple(A.shape)} and {tuple(B.shape)}")
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError(f"inner dimensions differ: {tuple(A.shape)} @ {tuple(B.shape)}")
    if out is None:
        out = torch.empty((M, N), dtype=torch.complex128, device=A.device)
    elif out.shape != (M, N) or out.dtype != torch.complex128 or not out.is_contiguous() or out.device != A.device:
        raise ValueError("out must be a contiguous complex128 tensor of shape (M, N) on the inputs' device")
    if K == 0:
        out.zero_()
        return out
    L = _lib.load()
    rc = L.mm_zgemm(
        ctypes.c_void_p(A.data_ptr()), K, ctypes.c_void_p(B.data_ptr()), N, ctypes.c_void_p(out.data_ptr()), N, M, N, K,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
    )
    _lib.check(rc, "mm_zgemm")
    return out


def dm_operator_product(P, rho, out: torch.Tensor | None = None) -> torch.Tensor:
    """P @ rho for an n-mode density matrix ``rho`` of shape (d_1..d_n, d_1..d_n)
    (bra indices first) viewed as a (prod d) x (prod d) matrix, and an operator
    ``P`` given as that matrix or with the same 2n-index shape.  Returns the
    product with rho's shape."""
    rho = as_device_c128(rho)
    if rho.dim() % 2:
        raise ValueError(f"a density matrix has an even number of indices, got shape {tuple(rho.shape)}")
    n = rho.dim() // 2
    if tuple(rho.shape[:n]) != tuple(rho.shape[n:]):
        raise ValueError(f"bra and ket dimensions differ: {tuple(rho.shape)}")
    d = 1
    for s in rho.shape[:n]:
        d *= int(s)
    P = as_device_c128(P).reshape(d, d)
    C = zgemm_hip(P, rho.reshape(d, d), out=None if out is None else out.reshape(d, d))
    return C.reshape(rho.shape)
