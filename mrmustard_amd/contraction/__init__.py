"""Contraction of Fock-basis tensors on the FP64 matrix cores.

The step after the Bargmann->Fock conversion in a circuit contraction is
``ArrayAnsatz.contract`` -> ``math.einsum`` over the converted arrays
(mrmustard/physics/ansatz/array_ansatz.py:159+,
mrmustard/math/backend_manager.py:526-548).  When the contracted indices can
be grouped, that einsum is one complex matrix product; ``zgemm_hip`` runs it
with ``mm_zgemm`` of libmmhip.so (v_mfma_f64_16x16x4_f64 tiles).

``contract`` is ``ArrayAnsatz.contract`` (array_ansatz.py:159-225) for two
arrays: labels shared by both operands and the output form a batch, labels
shared by the operands only are summed, the rest are the rows and columns;
one permutation per operand, one batched product (``mm_zgemm_batched``) and a
permutation of the result.  ``einsum`` is the two-operand case of
``math.einsum`` (backend_manager.py:526-548) on the same path.

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

__all__ = ["contract", "dm_operator_product", "einsum", "zgemm_batched_hip", "zgemm_hip"]


def zgemm_hip(A, B, out: torch.Tensor | None = None) -> torch.Tensor:
    """C = A @ B for complex128 matrices on the GPU (row-major, any strides
    whose last dimension is contiguous are copied to contiguous first).
    Differentiable when an operand requires grad (then without ``out``)."""
    A = as_device_c128(A)
    B = as_device_c128(B)
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError(f"zgemm_hip needs matrices, got shapes {tuple(A.shape)} and {tuple(B.shape)}")
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError(f"inner dimensions differ: {tuple(A.shape)} @ {tuple(B.shape)}")
    if torch.is_grad_enabled() and (A.requires_grad or B.requires_grad):
        if out is not None:
            raise ValueError("out is not supported when gradients are tracked")
        return zgemm_batched_hip(A[None], B[None])[0]
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


class _ZgemmBatched(torch.autograd.Function):
    """C = A @ B batched, differentiable as the JAX backend's einsum is
    (backend_jax.py einsum under jax.grad): with torch's convention for complex
    tensors, dL/dA = dL/dC @ B^H and dL/dB = A^H @ dL/dC, summed over the
    batch for an operand that was broadcast.  Both are products on the same
    matrix-core kernel."""

    @staticmethod
    def forward(ctx, A, B):
        ctx.save_for_backward(A, B)
        return _zgemm_batched(A, B)

    @staticmethod
    def backward(ctx, gC):
        A, B = ctx.saved_tensors
        gA = gB = None
        if ctx.needs_input_grad[0]:
            gA = _zgemm_batched(gC, B.transpose(1, 2).conj())
            if A.shape[0] == 1 and gA.shape[0] != 1:
                gA = gA.sum(0, keepdim=True)
        if ctx.needs_input_grad[1]:
            gB = _zgemm_batched(A.transpose(1, 2).conj(), gC)
            if B.shape[0] == 1 and gB.shape[0] != 1:
                gB = gB.sum(0, keepdim=True)
        return gA, gB


def zgemm_batched_hip(A, B) -> torch.Tensor:
    """C[b] = A[b] @ B[b] for (batch, M, K) and (batch, K, N) complex128
    tensors; an operand with batch 1 is broadcast.  Differentiable when an
    operand requires grad."""
    A = as_device_c128(A)
    B = as_device_c128(B)
    if A.dim() != 3 or B.dim() != 3:
        raise ValueError(f"zgemm_batched_hip needs (batch, M, K) and (batch, K, N), got {tuple(A.shape)} and {tuple(B.shape)}")
    ba, M, K = A.shape
    bb, K2, N = B.shape
    if K != K2:
        raise ValueError(f"inner dimensions differ: {tuple(A.shape)} @ {tuple(B.shape)}")
    if ba != bb and 1 not in (ba, bb):
        raise ValueError(f"batch sizes differ: {ba} and {bb}")
    if torch.is_grad_enabled() and (A.requires_grad or B.requires_grad):
        return _ZgemmBatched.apply(A, B)
    return _zgemm_batched(A, B)


def _zgemm_batched(A, B) -> torch.Tensor:
    """The kernel call of zgemm_batched_hip (shapes already checked); lazy
    conjugate views are materialised before their pointers are taken."""
    A = as_device_c128(A)
    B = as_device_c128(B)
    ba, M, K = A.shape
    bb, _, N = B.shape
    nb = max(ba, bb)
    out = torch.empty((nb, M, N), dtype=torch.complex128, device=A.device)
    if K == 0 or out.numel() == 0:
        return out.zero_()
    L = _lib.load()
    rc = L.mm_zgemm_batched(
        ctypes.c_void_p(A.data_ptr()), K, 0 if ba == 1 else M * K, ctypes.c_void_p(B.data_ptr()), N, 0 if bb == 1 else K * N,
        ctypes.c_void_p(out.data_ptr()), N, M * N, M, N, K, nb, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
    )
    _lib.check(rc, "mm_zgemm_batched")
    return out


def _middle_block(x, lx, y, ly, idx_out):
    """x[lead, s, trail] contracted with y[s, free] when the output is
    (lead, free, trail): one batched product per lead index, C = y^T x_lead,
    with x read in place (batch = prod(lead), K = prod(s), N = prod(trail))
    and only y permuted.  None when the labels do not have that form, when
    trail is empty (the generic product then reads x in place too) or when
    the products would be narrow (N < 32 with more than one)."""
    shared = set(lx) & set(ly)
    if not shared or any(lab in idx_out for lab in shared):
        return None
    pos = [k for k, lab in enumerate(lx) if lab in shared]
    if pos != list(range(pos[0], pos[0] + len(pos))):
        return None
    summed = [lx[k] for k in pos]
    lead, trail = lx[: pos[0]], lx[pos[-1] + 1 :]
    if not trail:  # x is then already (rows, summed): the generic product reads it in place
        return None
    nl, nt = len(lead), len(trail)
    if list(idx_out[:nl]) != lead or list(idx_out[len(idx_out) - nt :]) != trail:
        return None
    mid = list(idx_out[nl : len(idx_out) - nt])
    free = [lab for lab in ly if lab not in shared]
    if sorted(map(str, mid)) != sorted(map(str, free)) or len(mid) != len(free):
        return None

    def prod(t, labs, lab_list):
        n = 1
        for lab in labs:
            n *= int(t.shape[lab_list.index(lab)])
        return n

    nb, K, N = prod(x, lead, lx), prod(x, summed, lx), prod(x, trail, lx)
    M = prod(y, mid, ly)
    if N < 32 and nb > 1:  # narrow products: one permutation and a single wide product do better
        return None
    yt = y.permute(tuple(ly.index(lab) for lab in mid + summed)).reshape(1, M, K).contiguous()
    C = zgemm_batched_hip(yt, x.contiguous().reshape(nb, K, N))
    dims = [int(x.shape[lx.index(lab)]) for lab in lead] + [int(y.shape[ly.index(lab)]) for lab in mid] + \
        [int(x.shape[lx.index(lab)]) for lab in trail]
    return C.reshape(dims)


def contract(array1, array2, idx1, idx2, idx_out) -> torch.Tensor:
    """``ArrayAnsatz.contract`` (array_ansatz.py:159-225) of two arrays on the
    GPU: labels (str or int) name the axes of each array; shared labels whose
    sizes differ are cut to the smaller size, as the reference does; the result
    has the axes ``idx_out``.  A label repeated within an operand takes that
    operand's diagonal (numpy's einsum rule); a label of one operand that is
    not in the output is summed over."""
    a = as_device_c128(array1)
    b = as_device_c128(array2)
    idx1, idx2, idx_out = list(idx1), list(idx2), list(idx_out)
    if len(idx1) != a.dim():
        raise ValueError(f"expected len(idx1)={a.dim()} got {len(idx1)}")
    if len(idx2) != b.dim():
        raise ValueError(f"expected len(idx2)={b.dim()} got {len(idx2)}")
    if not set(idx_out).issubset(set(idx1) | set(idx2)):
        raise ValueError("Output labels must be present in input labels.")
    if len(set(idx_out)) != len(idx_out):
        raise ValueError(f"output labels {idx_out} repeat")
    shared = set(idx1) & set(idx2)
    sl1 = [slice(None)] * a.dim()
    sl2 = [slice(None)] * b.dim()
    for lab in shared:  # the first axis of a label is cut, as the reference does
        p1, p2 = idx1.index(lab), idx2.index(lab)
        d = min(a.shape[p1], b.shape[p2])
        sl1[p1] = slice(0, d)
        sl2[p2] = slice(0, d)
    a = a[tuple(sl1)]
    b = b[tuple(sl2)]
    # a label repeated within an operand takes that operand's diagonal first
    for k, (t, idx) in enumerate(((a, idx1), (b, idx2))):
        if len(set(idx)) == len(idx):
            continue
        for lab in set(idx):
            dims = {int(n) for x, n in zip(idx, t.shape) if x == lab}
            if len(dims) > 1:
                raise ValueError(f"the repeated axis {lab!r} of operand {k + 1} has sizes {sorted(dims)}")
        t, idx = _take_diagonals(t, idx)
        if k == 0:
            a, idx1 = t, idx
        else:
            b, idx2 = t, idx
    # a label of one operand only, absent from the output, is summed first
    for arr, idx in ((0, idx1), (1, idx2)):
        lone = [lab for lab in idx if lab not in shared and lab not in idx_out]
        if lone:
            t = a if arr == 0 else b
            t = t.sum(dim=[idx.index(lab) for lab in lone])
            for lab in lone:
                idx.remove(lab)
            if arr == 0:
                a = t
            else:
                b = t
    # an operand whose summed axes form one block between axes the output
    # keeps in place is read where it lies (no permutation of it or of the result)
    for x, lx, y, ly in sorted(((a, idx1, b, idx2), (b, idx2, a, idx1)), key=lambda t: -t[0].numel()):
        r = _middle_block(x, lx, y, ly, idx_out)
        if r is not None:
            return r
    batch = [lab for lab in idx_out if lab in shared]
    summed = [lab for lab in idx1 if lab in shared and lab not in idx_out]
    rows = [lab for lab in idx_out if lab in idx1 and lab not in shared]
    cols = [lab for lab in idx_out if lab in idx2 and lab not in shared]

    def size(t, idx, labs):
        n = 1
        for lab in labs:
            n *= int(t.shape[idx.index(lab)])
        return n

    A3 = a.permute(tuple(idx1.index(lab) for lab in batch + rows + summed)).reshape(
        size(a, idx1, batch), size(a, idx1, rows), size(a, idx1, summed))
    B3 = b.permute(tuple(idx2.index(lab) for lab in batch + summed + cols)).reshape(
        size(b, idx2, batch), size(b, idx2, summed), size(b, idx2, cols))
    C3 = zgemm_batched_hip(A3.contiguous(), B3.contiguous())
    dims = [int(a.shape[idx1.index(lab)]) for lab in batch + rows] + [int(b.shape[idx2.index(lab)]) for lab in cols]
    order = batch + rows + cols
    C = C3.reshape(dims)
    if not idx_out:
        return C
    return C.permute(tuple(order.index(lab) for lab in idx_out)).contiguous()


def einsum_labels(string: str, ndims) -> tuple[list[list[str]], list[str]]:
    """Per-operand label lists and the output labels of an einsum string, with
    numpy's rules: ``...`` stands for an operand's leading axes not named
    explicitly (right-aligned across operands, given labels of its own), and
    without ``->`` the output is the labels that appear once, sorted, after the
    ellipsis axes."""
    string = string.replace(" ", "")
    ins, out = string.split("->") if "->" in string else (string, None)
    parts = ins.split(",")
    if len(parts) != len(ndims):
        raise ValueError(f"einsum string {string!r} names {len(parts)} operands, got {len(ndims)}")
    n_ell = 0
    for part, nd in zip(parts, ndims):
        if "..." in part:
            n_ell = max(n_ell, nd - len(part.replace("...", "")))
    ell = [f"...{k}" for k in range(n_ell)]  # labels of the ellipsis axes, right-aligned

    def expand(part, nd):
        if "..." not in part:
            if len(part) != nd:
                raise ValueError(f"operand with {nd} axes labelled {part!r}")
            return list(part)
        head, tail = part.split("...")
        k = nd - len(head) - len(tail)
        if k < 0:
            raise ValueError(f"operand with {nd} axes labelled {part!r}")
        return list(head) + ell[n_ell - k:] + list(tail)

    labels = [expand(part, nd) for part, nd in zip(parts, ndims)]
    if out is None:
        counts = {}
        for labs in labels:
            for lab in labs:
                counts[lab] = counts.get(lab, 0) + 1
        outl = ell + sorted(lab for lab, n in counts.items() if n == 1 and not lab.startswith("..."))
    else:
        outl = list(out.split("...")[0]) + (ell if "..." in out else []) + (list(out.split("...")[1]) if "..." in out else [])
    return labels, outl


def _take_diagonals(t, labs):
    """Collapse a label repeated within one operand to its diagonal."""
    labs = list(labs)
    while len(set(labs)) != len(labs):
        lab = next(x for x in labs if labs.count(x) > 1)
        i = labs.index(lab)
        j = labs.index(lab, i + 1)
        t = torch.diagonal(t, dim1=i, dim2=j)  # the diagonal axis goes last
        labs = [x for k, x in enumerate(labs) if k not in (i, j)] + [lab]
    return t, labs


def greedy_path(labels, output, sizes):
    """Pairwise contraction order of opt_einsum's "greedy" optimizer
    (opt_einsum >= 3.4, ``paths.greedy`` with its default memory-removed
    cost: the path ``math.einsum`` asks for, backend_manager.py:526-548),
    restated on label sets:
      1. operands with equal label sets are multiplied first, left to right;
      2. every pair of operands sharing a label that the output lacks is a
         candidate, scored when it is found by (size of its result - sizes
         of the two operands, later operand id, earlier operand id); the best
         candidate whose operands are both still alive is contracted, and the
         result's pairs are added (scores are not revised afterwards, as in
         opt_einsum's lazy heap);
      3. what is left is multiplied as outer products, smallest first.
    A result keeps the labels of the output and those another operand still
    carries.  Returns steps (i, j, keep) over SSA ids: inputs are 0..n-1 and
    every product gets the next id.
    """
    import heapq

    output = frozenset(output)
    n = len(labels)
    if n <= 1:
        return []
    first = {lab: k for k, lab in enumerate(dict.fromkeys(x for lb in labels for x in lb))}

    def ordered(k):
        return sorted(k, key=first.__getitem__)

    def size(k):
        m = 1
        for lab in k:
            m *= sizes[lab]
        return m

    steps, nxt = [], n
    alive = {}  # label set -> id of the operand holding it
    for sid, lb in enumerate(labels):
        k = frozenset(lb)
        if k in alive:
            steps.append((alive[k], sid, ordered(k)))
            alive[k] = nxt
            nxt += 1
        else:
            alive[k] = sid
    holders = {}  # label outside the output -> label sets holding it
    for k in alive:
        for lab in k - output:
            holders.setdefault(lab, set()).add(k)

    def count(lab):
        return len(holders.get(lab, ()))

    def candidate(k1, k2):
        both, either = k1 & k2, k1 | k2
        k12 = frozenset(lab for lab in either if lab in output or count(lab) >= (3 if lab in both else 2))
        i1, i2 = alive[k1], alive[k2]
        if i1 > i2:
            k1, k2, i1, i2 = k2, k1, i2, i1
        return (size(k12) - size(k1) - size(k2), i2, i1), k1, k2, k12

    queue = []
    for lab, ks in holders.items():
        ks = sorted(ks, key=alive.__getitem__)
        for a in range(len(ks) - 1):
            for kb in ks[a + 1 :]:
                heapq.heappush(queue, candidate(ks[a], kb))
    while queue:
        _, k1, k2, k12 = heapq.heappop(queue)
        if k1 not in alive or k2 not in alive:
            continue  # an operand of this candidate is gone
        i1, i2 = alive.pop(k1), alive.pop(k2)
        for lab in k1 - output:
            holders[lab].discard(k1)
        for lab in k2 - output:
            holders[lab].discard(k2)
        steps.append((i1, i2, ordered(k12)))
        if k12 in alive:
            steps.append((alive[k12], nxt, ordered(k12)))
            nxt += 1
        else:
            for lab in k12 - output:
                holders[lab].add(k12)
        alive[k12] = nxt
        nxt += 1
        partners = {k for lab in k12 - output for k in holders[lab]} - {k12}
        for k in sorted(partners, key=alive.__getitem__):
            heapq.heappush(queue, candidate(k12, k))
    heap = [(size(k & output), sid, k) for k, sid in alive.items()]
    heapq.heapify(heap)
    _, s1, k1 = heapq.heappop(heap)
    while heap:
        _, s2, k2 = heapq.heappop(heap)
        k12 = (k1 | k2) & output
        steps.append((min(s1, s2), max(s1, s2), ordered(k12)))
        _, s1, k1 = heapq.heappushpop(heap, (size(k12), nxt, k12))
        nxt += 1
    return steps


def einsum(string: str, *tensors) -> torch.Tensor:
    """``math.einsum`` (backend_manager.py:526-548) on the GPU, with numpy's
    string rules (``...``, implicit output, a label repeated within an operand
    for its diagonal, size-1 axes broadcast).  Operands are contracted
    pairwise, each pair as one batched product on the matrix cores
    (``contract``), in the order of opt_einsum's "greedy" path
    (``greedy_path``); an intermediate keeps the labels the output or another
    operand still needs.  Axes of one label must have equal
    sizes or size 1 (ValueError otherwise, as numpy)."""
    if not tensors:
        raise ValueError("einsum needs at least one operand")
    ts = [as_device_c128(t) for t in tensors]
    labels, outl = einsum_labels(string, [t.dim() for t in ts])
    if len(set(outl)) != len(outl):
        raise ValueError(f"output labels of {string!r} repeat")
    every = set().union(*labels)
    if not set(outl).issubset(every):
        raise ValueError("Output labels must be present in input labels.")
    for k, (t, labs) in enumerate(zip(ts, labels)):
        for lab in set(labs):
            dims = {int(n) for x, n in zip(labs, t.shape) if x == lab}
            if len(dims) > 1:
                raise ValueError(f"einsum {string!r}: the repeated axis {lab!r} of operand {k} has sizes {sorted(dims)}")
    pairs = [_take_diagonals(t, labs) for t, labs in zip(ts, labels)]
    sizes = {}
    for t, labs in pairs:
        for lab, n in zip(labs, t.shape):
            n = int(n)
            m = sizes.setdefault(lab, n)
            if m != n:
                if 1 not in (m, n):
                    raise ValueError(f"einsum {string!r}: axis {lab!r} has sizes {m} and {n}")
                sizes[lab] = max(m, n)
    # size-1 axes broadcast against the others, as numpy's do
    pairs = [(t.expand(*[sizes[lab] for lab in labs]) if any(int(n) != sizes[lab] for lab, n in zip(labs, t.shape)) else t, labs)
             for t, labs in pairs]
    # the contraction order of opt_einsum's "greedy" path (the reference's
    # default, settings.EINSUM_OPTIMIZE), each pair one batched product
    ops = dict(enumerate(pairs))
    nxt = len(pairs)
    for i, j, keep in greedy_path([lb for _, lb in pairs], outl, sizes):
        (ti, li), (tj, lj) = ops.pop(i), ops.pop(j)
        ops[nxt] = (contract(ti, tj, li, lj, keep), keep)
        nxt += 1
    pairs = list(ops.values())
    res, rl = pairs[0]
    # one operand left: sum what the output does not name, then order the axes
    lone = [i for i, lab in enumerate(rl) if lab not in outl]
    if lone:
        res = res.sum(dim=lone)
        rl = [lab for lab in rl if lab in outl]
    if not outl:
        return res.reshape(())
    return res.permute(tuple(rl.index(lab) for lab in outl)).contiguous()
