"""Global settings that the Bargmann->Fock path reads.

Mirrors the hot-path flags of the reference's ``Settings`` singleton
(mrmustard/utils/settings.py:30-160): attribute access, and use as a context
manager that restores the previous values on exit
(``with settings(STABLE_FOCK_CONVERSION=True): ...``).
"""

from __future__ import annotations

import numpy as np


class Settings:
    """Singleton holding the flags consulted by ``math.hermite_renormalized``."""

    _instance = None

    def __new__(cls):
        if cls._instance is None:
            inst = super().__new__(cls)
            inst._stack = []
            # settings.py:101 -- ORed into `stable` by backend_manager.py:682
            inst.STABLE_FOCK_CONVERSION = False
            # not in the reference: autograd keeps only (A, b, c) and the backward
            # recomputes the lattice slab by slab (bounded memory)
            inst.HERMITE_RECOMPUTE_VJP = False
            # settings.py:79, :67-74 (consumed by callers that pick a shape)
            inst.DEFAULT_FOCK_SIZE = 50
            inst.AUTOSHAPE_PROBABILITY = 0.99999
            inst.AUTOSHAPE_MAX = 50
            inst.AUTOSHAPE_MIN = 1
            inst._seed = None
            inst.rng = np.random.default_rng()
            cls._instance = inst
        return cls._instance

    @property
    def SEED(self):
        return self._seed

    @SEED.setter
    def SEED(self, value):
        self._seed = value
        self.rng = np.random.default_rng(value)

    def __call__(self, **kwargs):
        for k in kwargs:
            if not hasattr(self, k):
                raise ValueError(f"unknown setting {k!r}")
        self._stack.append({k: getattr(self, k) for k in kwargs})
        for k, v in kwargs.items():
            setattr(self, k, v)
        return self

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        old = self._stack.pop()
        for k, v in old.items():
            setattr(self, k, v)
        return False


settings = Settings()
