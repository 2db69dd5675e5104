"""``BackendHip``: the backend object behind ``mrmustard_amd.math``.

It plays the role of ``BackendNumpy`` (math/backend_numpy.py) for the
Bargmann->Fock path: ``hermite_renormalized`` and
``hermite_renormalized_batched`` have the signatures of
backend_numpy.py:381-403, but run the HIP kernels on the current GPU and
return torch tensors.  The few generic tensor ops that the manager's
dispatch uses (``astensor``, ``reshape``, ``broadcast_to``, ...) are torch
operations.

``out`` semantics follow the numpy backend: the result is written into
``out`` and ``out`` itself is returned (test_vanilla.py:185).  ``out`` may be
a GPU complex128 tensor (written in place) or a numpy array (filled by a
device-to-host copy).  When any of A, b, c requires grad, the call goes
through the autograd Function and ``out`` is rejected, as the JAX backend
rejects it (backend_jax.py:485-486).
"""

from __future__ import annotations

import numpy as np
import torch

from ..lattice import compact_fock as cf
from ..lattice import strategies as st
from ..settings import settings
from .autograd import (
    HermiteDiagonal,
    HermiteLeftover,
    HermiteRenormalized,
    HermiteRenormalizedBatched,
    HermiteRenormalizedRecompute,
)


def _needs_grad(*xs) -> bool:
    return torch.is_grad_enabled() and any(isinstance(x, torch.Tensor) and x.requires_grad for x in xs)


class BackendHip:
    """Torch/ROCm backend whose Hermite methods call libmmhip.so."""

    name = "hip"

    int32 = torch.int32
    int64 = torch.int64
    float32 = torch.float32
    float64 = torch.float64
    complex64 = torch.complex64
    complex128 = torch.complex128

    def __repr__(self) -> str:
        return "BackendHip()"

    # ------------------------------------------------------------ generic ops
    def astensor(self, array, dtype=None):
        if isinstance(array, torch.Tensor):
            return array if dtype is None else array.to(dtype)
        arr = np.asarray(array)
        t = torch.as_tensor(arr)
        return t if dtype is None else t.to(dtype)

    def asnumpy(self, tensor) -> np.ndarray:
        if isinstance(tensor, torch.Tensor):
            return tensor.detach().resolve_conj().cpu().numpy()
        return np.asarray(tensor)

    def reshape(self, array, shape):
        if isinstance(array, torch.Tensor):
            return array.reshape(tuple(shape))
        return np.reshape(array, tuple(shape))

    def broadcast_to(self, array, shape):
        if isinstance(array, torch.Tensor):
            return array.broadcast_to(tuple(shape))
        return np.broadcast_to(array, tuple(shape))

    def zeros(self, shape, dtype=torch.complex128):
        return torch.zeros(tuple(shape), dtype=dtype)

    def conj(self, array):
        return array.conj() if isinstance(array, torch.Tensor) else np.conj(array)

    def sum(self, array, axis=None):
        if isinstance(array, torch.Tensor):
            return array.sum() if axis is None else array.sum(dim=axis)
        return np.sum(array, axis=axis)

    def allclose(self, a, b, rtol=1e-5, atol=1e-8) -> bool:
        return bool(np.allclose(self.asnumpy(a), self.asnumpy(b), rtol=rtol, atol=atol))

    # ---------------------------------------------------- hermite_renormalized
    @staticmethod
    def _deliver(G: torch.Tensor, out):
        """Write G into a numpy `out` and return `out` itself (numpy backend semantics)."""
        if out is None or isinstance(out, torch.Tensor):
            return G if out is None else out
        host = G.detach().cpu().numpy()
        if host.size != np.size(out):
            raise ValueError(f"out of shape {np.shape(out)} cannot hold an array of shape {host.shape}")
        np.copyto(out, host.reshape(np.shape(out)))
        return out

    def hermite_renormalized(self, A, b, c, shape: tuple[int], stable: bool = False, out=None):
        """Unbatched path (backend_numpy.py:381-392 -> core.py:25 / :127)."""
        if _needs_grad(A, b, c):
            if out is not None:
                raise ValueError("out is not supported when gradients are tracked")
            A, b, c = st.as_device_c128(A), st.as_device_c128(b), st.as_device_c128(c)
            fn = HermiteRenormalizedRecompute if settings.HERMITE_RECOMPUTE_VJP else HermiteRenormalized
            return fn.apply(A, b, c, tuple(int(s) for s in shape), bool(stable))
        dev_out = out if isinstance(out, torch.Tensor) else None
        G = (st.stable_hip if stable else st.vanilla_hip)(shape, A, b, c, out=dev_out)
        return self._deliver(G, out)

    def hermite_renormalized_binomial(self, A, B, C, shape: tuple[int], max_l2, global_cutoff):
        """Fill by weight with an L2 stop (backend_numpy.py:405-421 -> binomial.py:30-74)."""
        return st.binomial_hip(
            tuple(shape), A, B, C,
            max_l2=max_l2 or settings.AUTOSHAPE_PROBABILITY,
            global_cutoff=global_cutoff or sum(shape) - len(shape) + 1,
        )[0]

    def hermite_renormalized_diagonal(self, A, B, C, cutoffs: tuple[int], reorderedAB: bool):
        """Photon-number diagonal (backend_numpy.py:423-432 -> inputValidation.py:61-79);
        the [modes..., modes...] -> [mode0, mode0, ...] reorder happens in the kernel.
        With gradients tracked (unbatched B) it goes through the autograd
        Function, as the JAX backend's custom VJP does (backend_jax.py:513-522)."""
        if _needs_grad(A, B, C):
            A, B, C = st.as_device_c128(A), st.as_device_c128(B), st.as_device_c128(C)
            if B.ndim != 1:
                raise ValueError("gradients of hermite_renormalized_diagonal need an unbatched B")
            cf.input_validation(A)
            return HermiteDiagonal.apply(A, B, C, tuple(int(k) for k in cutoffs), bool(reorderedAB))
        return cf.diagonal_hip(A, B, C, cutoffs, reordered=reorderedAB)

    def hermite_renormalized_1leftoverMode(self, A, b, c, output_cutoff: int, pnr_cutoffs: tuple[int, ...],
                                           reorderedAB: bool = True):
        """Conditional density matrices of mode 0 (backend_jax.py:524-534 ->
        jax_vjps/hermite.py:343-387; the numpy backend's fast_diagonal,
        backend_numpy.py:434-446, returns the same amplitudes).  With gradients
        tracked it goes through the autograd Function, as the JAX backend's
        custom VJP does (jax_vjps/hermite.py:390-445)."""
        if _needs_grad(A, b, c):
            A, b, c = st.as_device_c128(A), st.as_device_c128(b), st.as_device_c128(c)
            return HermiteLeftover.apply(A, b, c, int(output_cutoff), tuple(int(k) for k in pnr_cutoffs), bool(reorderedAB))
        return cf.leftover_hip(A, b, c, output_cutoff, pnr_cutoffs, reordered=reorderedAB)

    def hermite_renormalized_batched(self, A, b, c, shape: tuple[int], stable: bool = False, out=None):
        """Batched path (backend_numpy.py:394-403 -> batch.py:27)."""
        if _needs_grad(A, b, c):
            if out is not None:
                raise ValueError("out is not supported when gradients are tracked")
            A, b, c = st.as_device_c128(A), st.as_device_c128(b), st.as_device_c128(c)
            return HermiteRenormalizedBatched.apply(A, b, c, tuple(int(s) for s in shape), bool(stable))
        dev_out = out if isinstance(out, torch.Tensor) else None
        G = st.vanilla_batch_hip(shape, A, b, c, stable=stable, out=dev_out)
        return self._deliver(G, out)
