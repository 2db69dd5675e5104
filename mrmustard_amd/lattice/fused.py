"""A Bargmann triple converted to Fock and contracted with a Fock array, the
lattice never stored whole.

The reference contracts a Bargmann component with a Fock one by converting
both to Fock and contracting the arrays: ``CircuitComponent.contract``
(mrmustard/lab/circuit_components.py:445-458) calls ``to_fock``
(mrmustard/physics/mm_einsum.py:272-292 -> ``math.hermite_renormalized``) and
then ``ArrayAnsatz.contract`` (mrmustard/physics/ansatz/array_ansatz.py:
159-225 -> ``math.einsum``).  ``fock_contract`` does both in one pass: the
triple's variables are reordered so that the kept wires come first and the
contracted ones last, the lattice is computed one slab of its first axis at a
time (``mm_hermite_batch_fwd_contract``, a three-slab window for lattices
beyond LDS), and every slab is multiplied by the array, reshaped to a
(contracted x remaining) matrix, as soon as it exists.  Resident memory is the
window and the output instead of the whole lattice.

The fused pass covers the common case: core wires of the triple either kept
(at least one) or contracted, its batch axes kept and absent from the array,
no gradients tracked, the vanilla recurrence.  Anything else goes through
``hermite_renormalized`` and ``contract``, which is also the reference's
order of operations.  Reordering the variables changes the rounding of the
lattice (the recurrence pivots on the first axis), so fused results agree
with the unfused ones to rounding, not bit for bit.
"""

from __future__ import annotations

import math as pymath

import torch

from .. import _lib
from ..contraction import contract
from ..settings import settings
from .strategies import C128, _ptr, _shape_arr, _stream, _workspace, as_device_c128

__all__ = ["fock_contract"]


def _unfused(A, b, c, shape, idx1, array, idx2, idx_out):
    from .. import math as mm

    G = mm.hermite_renormalized(A, b, c, shape=tuple(shape))
    return contract(G, array, idx1, idx2, idx_out)


def fock_contract(A, b, c, shape, idx1, array, idx2, idx_out, fused: bool | None = None) -> torch.Tensor:
    """``contract(hermite_renormalized(A, b, c, shape), array, idx1, idx2,
    idx_out)`` with the conversion fused with the contraction.

    ``A``, ``b``, ``c`` may carry batch axes (A: batch + (D, D)); ``idx1``
    labels the axes of the lattice (batch axes, then the D core axes) and
    ``idx2`` those of ``array``.  A contracted label's lattice axis is cut to
    the array's size when that is smaller, as ``ArrayAnsatz.contract`` cuts
    shared axes.  ``fused=None`` fuses whenever the case is covered (see the
    module docstring), ``False`` never does."""
    A, b, c = as_device_c128(A), as_device_c128(b), as_device_c128(c)
    W = as_device_c128(array)
    shape = [int(s) for s in shape]
    idx1, idx2, idx_out = list(idx1), list(idx2), list(idx_out)
    D = int(A.shape[-1])
    batch_shape = tuple(A.shape[:-2])
    nbatch = len(batch_shape)
    if len(shape) != D or len(idx1) != nbatch + D or len(idx2) != W.dim():
        raise ValueError(f"labels {idx1}/{idx2} do not match the lattice {batch_shape + tuple(shape)} and array {tuple(W.shape)}")
    tracked = torch.is_grad_enabled() and any(t.requires_grad for t in (A, b, c, W))
    batch_labs, core_labs = idx1[:nbatch], idx1[nbatch:]
    shared = set(core_labs) & set(idx2)
    kept = [lab for lab in idx_out if lab in core_labs and lab not in shared]
    summed = [lab for lab in core_labs if lab in shared and lab not in idx_out]
    covered = (
        fused is not False
        and not tracked
        and not settings.STABLE_FOCK_CONVERSION
        and len(set(idx1)) == len(idx1)
        and len(set(idx2)) == len(idx2)
        and len(kept) >= 1
        and len(kept) + len(summed) == D
        and all(lab in idx_out and lab not in idx2 for lab in batch_labs)
        and set(idx_out) <= set(idx1) | set(idx2)
        and (b.dim() == nbatch + 1 and (nbatch == 0 or tuple(c.shape) == batch_shape))
    )
    if not covered:
        return _unfused(A, b, c, shape, idx1, W, idx2, idx_out)

    # the array: contracted axes (cut to the common size) first, then the rest
    cut = {}
    for lab in summed:
        cut[lab] = min(shape[core_labs.index(lab)], int(W.shape[idx2.index(lab)]))
    sl = [slice(None)] * W.dim()
    for lab in summed:
        sl[idx2.index(lab)] = slice(0, cut[lab])
    W = W[tuple(sl)]
    lone = [lab for lab in idx2 if lab not in shared and lab not in idx_out]
    if lone:  # labels of the array alone and not in the output are summed first
        W = W.sum(dim=[idx2.index(lab) for lab in lone])
        idx2 = [lab for lab in idx2 if lab not in lone]
    cols = [lab for lab in idx2 if lab not in shared]
    K = pymath.prod(cut[lab] for lab in summed)
    M = pymath.prod(int(W.shape[idx2.index(lab)]) for lab in cols)
    Wm = W.permute(tuple(idx2.index(lab) for lab in summed + cols)).reshape(K, max(M, 1)).contiguous()
    M = max(M, 1)

    # the triple: kept variables first, contracted ones last
    perm = [core_labs.index(lab) for lab in kept + summed]
    p = torch.tensor(perm, device=A.device)
    nb = pymath.prod(batch_shape) if batch_shape else 1
    Ap = A.reshape(nb, D, D)[:, p][:, :, p].contiguous()
    bp = b.reshape(nb, D)[:, p].contiguous()
    cp = c.reshape(nb).contiguous()
    shp = [shape[core_labs.index(lab)] for lab in kept] + [cut[lab] for lab in summed]
    rows = pymath.prod(shp[: len(kept)])
    out = torch.empty((nb, rows, M), dtype=C128, device=A.device)
    if nb > 0:
        L = _lib.load()
        s = _shape_arr(shp)
        ws = _workspace(L.mm_hermite_batch_fwd_contract_workspace(s.ctypes.data, D))
        rc = L.mm_hermite_batch_fwd_contract(
            _ptr(Ap), _ptr(bp), _ptr(cp), nb, s.ctypes.data, D, len(kept), _ptr(Wm), M, 0, _ptr(out),
            _ptr(ws), 0 if ws is None else ws.numel(), _stream(),
        )
        _lib.check(rc, "mm_hermite_batch_fwd_contract")
    dims = list(batch_shape) + shp[: len(kept)] + [int(W.shape[idx2.index(lab)]) for lab in cols]
    order = batch_labs + kept + cols
    res = out.reshape(dims)
    if not idx_out:
        return res
    return res.permute(tuple(order.index(lab) for lab in idx_out)).contiguous()
