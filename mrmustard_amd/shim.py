"""Install the GPU path into an existing MrMustard ``BackendManager``.

MrMustard resolves every math method by name on ``BackendManager._backend``
(mrmustard/math/backend_manager.py:89-116) and only knows the "numpy" and
"jax" backends (:193).  ``install(math)`` swaps ``_backend`` for a wrapper
whose ``hermite_renormalized`` / ``hermite_renormalized_batched``
(backend_numpy.py:381-403), ``hermite_renormalized_binomial`` (:405-421),
``hermite_renormalized_diagonal`` (:423-432) and
``hermite_renormalized_1leftoverMode`` (:434-446) run libmmhip.so and whose
other attributes are those of the wrapped backend.  Inputs and outputs stay numpy arrays, as the
numpy backend expects: the wrapper copies to the GPU, runs, and copies the
result back (into ``out`` when given, which is then returned).
"""

from __future__ import annotations

import numpy as np


class HipBackendWrapper:
    """A numpy-style backend whose Hermite methods run on the GPU."""

    def __init__(self, base):
        self._base = base
        self.name = getattr(base, "name", "numpy")

    def __getattr__(self, name):
        return getattr(self._base, name)

    @staticmethod
    def _to_numpy(G, out):
        host = G.detach().cpu().numpy()
        if out is None:
            return host
        np.copyto(out, host.reshape(np.shape(out)))
        return out

    def hermite_renormalized(self, A, b, c, shape, stable=False, out=None):
        from .lattice import strategies as st

        G = (st.stable_hip if stable else st.vanilla_hip)(tuple(shape), A, b, c)
        return self._to_numpy(G, out)

    def hermite_renormalized_batched(self, A, b, c, shape, stable=False, out=None):
        from .lattice import strategies as st

        G = st.vanilla_batch_hip(tuple(shape), A, b, c, stable=stable)
        return self._to_numpy(G, out)

    def hermite_renormalized_binomial(self, A, B, C, shape, max_l2, global_cutoff):
        from .lattice import strategies as st
        from .settings import settings

        G, _ = st.binomial_hip(
            tuple(shape), A, B, C,
            max_l2=max_l2 or settings.AUTOSHAPE_PROBABILITY,
            global_cutoff=global_cutoff or sum(shape) - len(shape) + 1,
        )
        return self._to_numpy(G, None)

    def hermite_renormalized_diagonal(self, A, B, C, cutoffs, reorderedAB):
        from .lattice import compact_fock as cf

        return self._to_numpy(cf.diagonal_hip(A, B, C, cutoffs, reordered=reorderedAB), None)

    def hermite_renormalized_1leftoverMode(self, A, b, c, output_cutoff, pnr_cutoffs, stable=False, reorderedAB=True):
        """The numpy backend's signature (backend_numpy.py:434-446): its
        fast_diagonal reads (A, b) in the [modes..., modes...] order whatever
        the flags, and so does this wrapper; ``stable`` changes rounding only."""
        from .lattice import compact_fock as cf

        return self._to_numpy(cf.leftover_hip(A, b, c, output_cutoff, pnr_cutoffs, reordered=True), None)


def install(manager) -> HipBackendWrapper:
    """Wrap ``manager``'s current backend (instance attribute, so the class
    default is untouched and ``uninstall`` restores it)."""
    current = manager._backend
    if isinstance(current, HipBackendWrapper):
        return current
    wrapper = HipBackendWrapper(current)
    manager._backend = wrapper
    return wrapper


def uninstall(manager) -> None:
    current = manager.__dict__.get("_backend")
    if isinstance(current, HipBackendWrapper):
        manager._backend = current._base
