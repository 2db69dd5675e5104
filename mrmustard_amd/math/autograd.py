"""torch.autograd wiring of the Hermite forward and its closed-form VJP.

Counterpart of the reference's ``jax.custom_vjp`` pair
(math/jax_vjps/hermite.py:47-102 unbatched, :113-175 batched): the forward
keeps (G, c) and the backward calls the VJP kernel, whatever ``stable`` was,
exactly as ``hermite_renormalized_jax_bwd`` always calls ``vanilla_vjp_numba``.

Convention.  The reference VJP is holomorphic: it returns sum_k g_k dG_k/dtheta
for the cotangent g, unconjugated.  torch hands a backward function
grad = dL/d(conj G) and expects dL/d(conj theta), which for a holomorphic map
is conj( sum_k conj(grad_k) dG_k/dtheta ).  So the backward conjugates the
incoming gradient, runs the reference VJP, and conjugates the result.
"""

from __future__ import annotations

import torch

from ..lattice import strategies as st


class HermiteRenormalized(torch.autograd.Function):
    """G = hermite_renormalized(A, b, c, shape, stable) with a custom VJP."""

    @staticmethod
    def forward(ctx, A, b, c, shape, stable):
        G = st.stable_hip(shape, A, b, c) if stable else st.vanilla_hip(shape, A, b, c)
        ctx.save_for_backward(G, c)
        ctx.a_shape, ctx.b_shape, ctx.c_shape = A.shape, b.shape, c.shape
        return G

    @staticmethod
    def backward(ctx, grad):
        G, c = ctx.saved_tensors
        dA, db, dc = st.vanilla_vjp_hip(G, c, grad.conj())
        return (
            dA.conj().resolve_conj().reshape(ctx.a_shape),
            db.conj().resolve_conj().reshape(ctx.b_shape),
            dc.conj().resolve_conj().reshape(ctx.c_shape),
            None,
            None,
        )


class HermiteRenormalizedBatched(torch.autograd.Function):
    """Batched G[k] = hermite_renormalized(A[k], b[k], c[k], shape, stable)."""

    @staticmethod
    def forward(ctx, A, b, c, shape, stable):
        G = st.vanilla_batch_hip(shape, A, b, c, stable=stable)
        ctx.save_for_backward(G, c)
        ctx.a_shape, ctx.b_shape, ctx.c_shape = A.shape, b.shape, c.shape
        return G

    @staticmethod
    def backward(ctx, grad):
        G, c = ctx.saved_tensors
        dA, db, dc = st.vanilla_batch_vjp_hip(G, c, grad.conj())
        return (
            dA.conj().resolve_conj().reshape(ctx.a_shape),
            db.conj().resolve_conj().reshape(ctx.b_shape),
            dc.conj().resolve_conj().reshape(ctx.c_shape),
            None,
            None,
        )


class HermiteRenormalizedRecompute(torch.autograd.Function):
    """As HermiteRenormalized, but autograd keeps only (A, b, c): the backward
    recomputes the lattice slab by slab in a three-slab window
    (``vanilla_vjp_recompute_hip``), so G need not stay resident between the
    forward and the backward.  The VJP applies to the forward's own lattice,
    as the reference's backward does (jax_vjps/hermite.py:78-101): for the
    all-pivot average the backward recomputes that lattice whole with
    ``stable_hip`` (the windowed sweep runs the first-pivot recurrence only)
    and frees it once the VJP is taken."""

    @staticmethod
    def forward(ctx, A, b, c, shape, stable):
        G = st.stable_hip(shape, A, b, c) if stable else st.vanilla_hip(shape, A, b, c)
        ctx.save_for_backward(A, b, c)
        ctx.shape = tuple(shape)
        ctx.stable = bool(stable)
        return G

    @staticmethod
    def backward(ctx, grad):
        A, b, c = ctx.saved_tensors
        if ctx.stable:
            G = st.stable_hip(ctx.shape, A, b, c)
            dA, db, dc = st.vanilla_vjp_hip(G, c, grad.conj())
            del G
        else:
            dA, db, dc = st.vanilla_vjp_recompute_hip(ctx.shape, A, b, c, grad.conj())
        return (
            dA.conj().resolve_conj().reshape(A.shape),
            db.conj().resolve_conj().reshape(b.shape),
            dc.conj().resolve_conj().reshape(c.shape),
            None,
            None,
        )


class HermiteDiagonal(torch.autograd.Function):
    """G = hermite_renormalized_diagonal(A, B, c, cutoffs, reorderedAB) with the
    reverse-sweep VJP (the reference's jax custom VJP pair,
    jax_vjps/hermite.py:248-336).  The backward recomputes the forward together
    with the VJP (``diagonal_vjp_hip``), so only (A, B, c) are kept."""

    @staticmethod
    def forward(ctx, A, B, c, cutoffs, reordered):
        from ..lattice import compact_fock as cf

        G = cf.diagonal_hip(A, B, c, cutoffs, reordered=reordered)
        ctx.save_for_backward(A, B, c)
        ctx.cutoffs, ctx.reordered = tuple(cutoffs), bool(reordered)
        return G

    @staticmethod
    def backward(ctx, grad):
        from ..lattice import compact_fock as cf

        A, B, c = ctx.saved_tensors
        _, dA, dB, dc = cf.diagonal_vjp_hip(A, B, c, ctx.cutoffs, grad.conj().resolve_conj(), reordered=ctx.reordered)
        return (
            dA.conj().resolve_conj().reshape(A.shape),
            dB.conj().resolve_conj().reshape(B.shape),
            dc.conj().resolve_conj().reshape(c.shape),
            None,
            None,
        )


class HermiteLeftover(torch.autograd.Function):
    """out = hermite_renormalized_1leftoverMode(A, b, c, output_cutoff,
    pnr_cutoffs) with the reverse-sweep VJP (the reference's jax custom VJP
    pair, jax_vjps/hermite.py:343-445).  The backward recomputes the forward
    together with the VJP (``leftover_vjp_hip``), so only (A, b, c) are kept."""

    @staticmethod
    def forward(ctx, A, b, c, output_cutoff, pnr_cutoffs, reordered):
        from ..lattice import compact_fock as cf

        out = cf.leftover_hip(A, b, c, output_cutoff, pnr_cutoffs, reordered=reordered)
        ctx.save_for_backward(A, b, c)
        ctx.args = (int(output_cutoff), tuple(int(k) for k in pnr_cutoffs), bool(reordered))
        return out

    @staticmethod
    def backward(ctx, grad):
        from ..lattice import compact_fock as cf

        A, b, c = ctx.saved_tensors
        o, pnr, reordered = ctx.args
        _, dA, db, dc = cf.leftover_vjp_hip(A, b, c, o, pnr, grad.conj().resolve_conj(), reordered=reordered)
        return (
            dA.conj().resolve_conj().reshape(A.shape),
            db.conj().resolve_conj().reshape(b.shape),
            dc.conj().resolve_conj().reshape(c.shape),
            None,
            None,
            None,
        )
