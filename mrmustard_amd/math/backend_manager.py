"""``BackendManager``: the plugin surface of the path (``mrmustard_amd.math``).

Mirrors mrmustard/math/backend_manager.py:
- singleton (``__new__``, :132-138) holding a backend object in ``_backend``;
- ``_apply`` (:89-116) resolves a method by name on the backend and raises
  ``NotImplementedError`` when it is missing;
- ``_bind`` (:118-130) copies the backend's dtype attributes;
- ``change_backend`` / ``get_backend`` (:167-209), here for the ``"hip"``
  backend;
- ``hermite_renormalized`` (:643-727) with the same three-branch dispatch,
  ``out`` shape check and ``ValueError`` messages, and ``stable`` ORed with
  ``settings.STABLE_FOCK_CONVERSION`` (:682).

Inputs may be numpy arrays, Python scalars or torch tensors; they are turned
into tensors before the dispatch so that ``.ndim`` exists on all of them.
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import numpy as np

from ..settings import settings
from .backend_hip import BackendHip

_BACKENDS = {"hip": BackendHip}


class BackendManager:
    """Singleton dispatching by method name to the backend in use."""

    _backend = BackendHip()

    def __new__(cls):
        try:
            return cls.instance
        except AttributeError:
            cls.instance = super().__new__(cls)
            return cls.instance

    def __init__(self) -> None:
        self._bind()

    def __repr__(self) -> str:
        return f"Backend({self.backend_name})"

    # ---------------------------------------------------------------- plumbing
    def _apply(self, fn: str, args: Sequence[Any] | None = (), kwargs: dict | None = None, backend_name: str | None = None):
        kwargs = kwargs or {}
        backend = self.get_backend(backend_name) if backend_name else self.backend
        try:
            attr = getattr(backend, fn)
        except AttributeError:
            raise NotImplementedError(f"Function ``{fn}`` not implemented for backend ``{backend.name}``.") from None
        return attr(*args, **kwargs)

    def _bind(self) -> None:
        for name in ["int32", "int64", "float32", "float64", "complex64", "complex128"]:
            setattr(self, name, getattr(self._backend, name))

    @property
    def backend(self):
        return self._backend

    @property
    def backend_name(self) -> str:
        return self._backend.name

    def change_backend(self, name: str) -> None:
        if self.backend_name != name:
            self._backend = self.get_backend(name)
            self._bind()

    def get_backend(self, name: str | None = None):
        if name not in _BACKENDS:
            raise ValueError("Backend must be ``hip``.")
        if self.backend_name != name:
            return _BACKENDS[name]()
        return self.backend

    # ----------------------------------------------------------- generic ops
    def astensor(self, array, dtype=None):
        return self._apply("astensor", (array,), {"dtype": dtype})

    def asnumpy(self, tensor):
        return self._apply("asnumpy", (tensor,))

    def reshape(self, array, shape):
        return self._apply("reshape", (array, shape))

    def broadcast_to(self, array, shape):
        return self._apply("broadcast_to", (array, shape))

    def __getattr__(self, name: str):
        # any other backend method, by name (the reference lists ~120 explicitly)
        if name.startswith("_"):
            raise AttributeError(name)
        return lambda *args, **kwargs: self._apply(name, args, kwargs)

    # --------------------------------------------------- the hot path boundary
    def hermite_renormalized(self, A, b, c, shape: tuple[int], stable: bool = False, out=None):
        r"""Renormalized multidimensional Hermite polynomial: the Taylor
        coefficients of :math:`c\,\exp(bx + \tfrac12 x^T A x)` at zero with
        :math:`\sqrt{n!}` in the denominator, on the box ``shape``.

        1. A.ndim = 2, b.ndim = 1, c scalar: unbatched;
        2. A.ndim = 2, b.ndim > 1: A and c broadcast over b's batch;
        3. A.ndim > 2, b.ndim > 1, c.ndim > 0: fully batched.

        Batch axes come first in the result.
        """
        A, b, c = self._as_input(A), self._as_input(b), self._as_input(c)

        def check_out_shape(batch_shape):
            if out is not None and any(d_out < d for d_out, d in zip(np.shape(out), batch_shape + tuple(shape))):
                raise ValueError(f"batch+shape {batch_shape + tuple(shape)} is too large for out.shape={tuple(np.shape(out))}")

        stable = stable or settings.STABLE_FOCK_CONVERSION
        if A.ndim > 2 and b.ndim > 1 and c.ndim > 0:
            batch_shape = tuple(A.shape[:-2])
            check_out_shape(batch_shape)
            if tuple(b.shape[:-1]) != batch_shape:
                raise ValueError(f"b.shape={tuple(b.shape)} must match batch_shape={batch_shape}")
            if tuple(c.shape[: len(batch_shape)]) != batch_shape:
                raise ValueError(f"c.shape={tuple(c.shape)} must match batch_shape={batch_shape}")
            D = int(np.prod(batch_shape))
            A = self.reshape(A, (D, *A.shape[-2:]))
            b = self.reshape(b, (D, *b.shape[-1:]))
            c = self.reshape(c, (D,))
            result = self._apply(
                "hermite_renormalized_batched",
                (A, b, c),
                {"shape": tuple(shape), "stable": stable, "out": self.reshape(out, (D, *shape)) if out is not None else None},
            )
            return self.reshape(result, batch_shape + tuple(shape))
        if A.ndim == 2 and b.ndim > 1:  # b-batched case
            batch_shape = tuple(b.shape[:-1])
            check_out_shape(batch_shape)
            D = int(np.prod(batch_shape))
            b = self.reshape(b, (D, *b.shape[-1:]))
            A_broadcast = self.broadcast_to(A, (D, *A.shape))
            c_broadcast = self.broadcast_to(c, (D,))
            result = self._apply(
                "hermite_renormalized_batched",
                (A_broadcast, b, c_broadcast),
                {"shape": tuple(shape), "stable": stable, "out": self.reshape(out, (D, *shape)) if out is not None else None},
            )
            return self.reshape(result, batch_shape + tuple(shape))
        check_out_shape(())
        return self._apply("hermite_renormalized", (A, b, c), {"shape": tuple(shape), "stable": stable, "out": out})

    def hermite_renormalized_binomial(self, A, B, C, shape: tuple[int], max_l2, global_cutoff):
        r"""The amplitudes of ``hermite_renormalized`` filled by total photon
        number, up to the weight at which the squared L2 norm exceeds ``max_l2``
        (default ``settings.AUTOSHAPE_PROBABILITY``) or up to ``global_cutoff``
        (default ``sum(shape) - len(shape) + 1``) (backend_manager.py:775-800)."""
        return self._apply("hermite_renormalized_binomial", (A, B, C, shape, max_l2, global_cutoff))

    def hermite_renormalized_diagonal(self, A, b, c, cutoffs: tuple[int], reorderedAB: bool = True):
        r"""The diagonal G[k0, k0, k1, k1, ...] of the lattice of (A, b, c): the
        photon-number amplitudes of every mode, computed without the full
        lattice; b may carry a batch of B's as its second axis
        (backend_manager.py:729-757)."""
        return self._apply("hermite_renormalized_diagonal", (A, b, c, cutoffs, reorderedAB))

    def hermite_renormalized_1leftoverMode(self, A, b, c, output_cutoff: int, pnr_cutoffs: tuple[int, ...],
                                           reorderedAB: bool = True):
        r"""The density matrix of mode 0 when all the other modes are detected
        with PNR detectors up to the given photon numbers: shape
        ``(output_cutoff + 1, output_cutoff + 1, *pnr_cutoffs + 1)``
        (backend_manager.py:759-786)."""
        return self._apply("hermite_renormalized_1leftoverMode", (A, b, c, output_cutoff, pnr_cutoffs, reorderedAB))

    def einsum(self, string: str, *tensors, optimize="greedy", backend=None):
        r"""Einstein summation of Fock-basis arrays on the FP64 matrix cores
        (backend_manager.py:526-548), as ``ArrayAnsatz.contract`` and the
        derived-variables branch of ``fock_array`` call it: numpy's string
        rules, operands contracted pairwise in the order of opt_einsum's
        greedy path (the reference's default ``optimize="greedy"``).  Other
        ``optimize`` values and ``backend`` are accepted for the reference's
        signature and take the same order."""
        from ..contraction import einsum as _einsum

        return _einsum(string, *tensors)

    def _as_input(self, x):
        if hasattr(x, "ndim") and hasattr(x, "shape"):
            return x
        return self.astensor(np.asarray(x, dtype=np.complex128))
