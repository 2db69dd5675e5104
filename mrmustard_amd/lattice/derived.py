"""Fock arrays of Bargmann triples with derived variables, on the GPU.

A ``PolyExpAnsatz`` with n continuous and m derived variables has A, b over
n + m variables (batch first) and c of shape ``batch + shape_derived_vars``.

- ``fock_array_derived`` <- CircuitComponent.fock_array's ``num_derived_vars > 0``
  branch (mrmustard/lab/circuit_components.py:516-528): the lattice on
  ``shape + shape_derived_vars`` with c = 1, its derived indices flattened and
  contracted with c.
- ``decompose_ansatz_triple`` <- PolyExpAnsatz.decompose_ansatz
  (mrmustard/physics/ansatz/polyexp_ansatz.py:415-469): the equivalent triple
  with n derived variables of size sum(shape_derived_vars) each, for m >= n.

Both run the batched Hermite kernels and one batched product on the FP64
matrix cores (``zgemm_batched_hip``) for the contraction with c.  For large
lattices ``fock_array_derived`` fuses the conversion with the contraction
(``mm_hermite_batch_fwd_contract``, every triple of the batch in one call):
the lattice is computed one slab of its first axis at a time in a three-slab
window and each slab is contracted with c as soon as it exists, so the
lattice on ``shape + shape_derived_vars`` is never stored whole.
"""

from __future__ import annotations

import math as pymath

import torch

from .. import _lib
from ..contraction import zgemm_batched_hip
from ..settings import settings
from .strategies import C128, _ptr, _shape_arr, _stream, _workspace, as_device_c128

FUSE_MIN_AMPS = 1 << 20  # lattice entries per triple from which fock_array_derived fuses by default

__all__ = ["decompose_ansatz_triple", "fock_array_derived"]


def _split(A, b, c):
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c)
    batch_shape = tuple(A.shape[:-2])
    if tuple(b.shape[:-1]) != batch_shape or tuple(c.shape[: len(batch_shape)]) != batch_shape:
        raise ValueError(f"batch shapes of A {tuple(A.shape)}, b {tuple(b.shape)} and c {tuple(c.shape)} differ")
    shape_derived = tuple(int(k) for k in c.shape[len(batch_shape):])
    n = int(A.shape[-1]) - len(shape_derived)
    if n < 0:
        raise ValueError("c has more derived axes than A has variables")
    return A, b, c, batch_shape, shape_derived, n


def _lattice(A, b, batch_shape, shape):
    from .. import math as mm

    ones = torch.ones(batch_shape, dtype=torch.complex128, device=A.device)
    return mm.hermite_renormalized(A, b, ones, shape=tuple(shape))


def _contract_derived(G, c, batch_shape, head_shape):
    """sum_k G[..., head, k] c[..., k] with the derived axes flattened into k."""
    nb = pymath.prod(batch_shape) if batch_shape else 1
    rows = pymath.prod(head_shape) if head_shape else 1
    K = pymath.prod(c.shape[len(batch_shape):])
    G3 = G.reshape(nb, rows, K)
    c3 = c.reshape(nb, K, 1)
    return zgemm_batched_hip(G3, c3).reshape(tuple(batch_shape) + tuple(head_shape))


def _fused_batch(A3, b3, c3, full_shape, n):
    """All triples through one mm_hermite_batch_fwd_contract call: the lattice
    of (A_t, b_t, 1) on ``full_shape`` contracted with c_t over its last axes
    as it is made, one slab at a time."""
    L = _lib.load()
    s = _shape_arr(full_shape)
    D = len(full_shape)
    nb = A3.shape[0]
    ones = torch.ones(nb, dtype=C128, device=A3.device)
    K = pymath.prod(full_shape[n:])
    w = c3.reshape(nb, K).contiguous()
    out = torch.empty((nb,) + tuple(full_shape[:n]), dtype=C128, device=A3.device)
    ws = _workspace(L.mm_hermite_batch_fwd_contract_workspace(s.ctypes.data, D))
    rc = L.mm_hermite_batch_fwd_contract(
        _ptr(A3.contiguous()), _ptr(b3.contiguous()), _ptr(ones), nb, s.ctypes.data, D, n, _ptr(w), 1, K, _ptr(out),
        _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
    )
    _lib.check(rc, "mm_hermite_batch_fwd_contract")
    return out


def fock_array_derived(A, b, c, shape, fused: bool | None = None) -> torch.Tensor:
    """The Fock array of shape ``batch + shape`` of a triple whose c carries the
    derived variables (circuit_components.py:516-528): the lattice of (A, b, 1)
    on ``shape + shape_derived_vars``, contracted over the derived indices
    with c.  ``shape`` has one entry per continuous variable.  ``fused``
    (default: from FUSE_MIN_AMPS lattice entries per triple) computes each
    triple with the conversion fused with the contraction, never storing the
    lattice whole; otherwise (and always with gradients tracked or
    settings.STABLE_FOCK_CONVERSION) all triples go through the batched kernels
    and one batched product."""
    A, b, c, batch_shape, shape_derived, n = _split(A, b, c)
    shape = tuple(int(s) for s in shape)
    if len(shape) != n:
        raise ValueError(f"shape {shape} has {len(shape)} entries for {n} continuous variables")
    full = shape + shape_derived
    if fused is None:
        fused = pymath.prod(full) >= FUSE_MIN_AMPS
    # the fused kernel computes the vanilla lattice and has no backward: the
    # stable conversion and tracked gradients go through the lattice + product
    tracked = torch.is_grad_enabled() and any(t.requires_grad for t in (A, b, c))
    if fused and n >= 1 and not tracked and not settings.STABLE_FOCK_CONVERSION:
        nb = pymath.prod(batch_shape) if batch_shape else 1
        if nb == 0:
            return torch.empty(tuple(batch_shape) + shape, dtype=C128, device=A.device)
        A3 = A.reshape(nb, *A.shape[-2:])
        b3 = b.reshape(nb, b.shape[-1])
        c3 = c.reshape(nb, *shape_derived)
        return _fused_batch(A3, b3, c3, full, n).reshape(tuple(batch_shape) + shape)
    G = _lattice(A, b, batch_shape, full)
    return _contract_derived(G, c, batch_shape, shape)


def decompose_ansatz_triple(A, b, c):
    """(A', b', c') of PolyExpAnsatz.decompose_ansatz (polyexp_ansatz.py:415-469):
    for m derived variables with m >= n continuous ones,
      A' = [[A_nn, I], [I, 0]],  b' = [b_n, 0],
      c'[..., l_1..l_n] = sum_k H[..., l, k] c[..., k],  l_i < sum(shape_derived_vars),
    with H the lattice of ([[0, A_nm], [A_mn, A_mm]], [0, b_m], 1).  Returns
    the triple unchanged when m < n."""
    A, b, c, batch_shape, shape_derived, n = _split(A, b, c)
    m = len(shape_derived)
    if m < n:
        return A, b, c
    L = sum(shape_derived)
    pulled = (L,) * n
    A_core = A.clone()
    A_core[..., :n, :n] = 0
    b_core = b.clone()
    b_core[..., :n] = 0
    H = _lattice(A_core, b_core, batch_shape, pulled + shape_derived)
    c_prime = _contract_derived(H, c, batch_shape, pulled)
    eye = torch.eye(n, dtype=torch.complex128, device=A.device).expand(*batch_shape, n, n)
    zero = torch.zeros(*batch_shape, n, n, dtype=torch.complex128, device=A.device)
    A_dec = torch.cat([torch.cat([A[..., :n, :n], eye], dim=-1), torch.cat([eye, zero], dim=-1)], dim=-2).contiguous()
    b_dec = torch.cat([b[..., :n], torch.zeros(*batch_shape, n, dtype=torch.complex128, device=A.device)], dim=-1).contiguous()
    return A_dec, b_dec, c_prime
