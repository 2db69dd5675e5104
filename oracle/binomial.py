"""CPU oracle of the binomial strategy.  TEST INFRASTRUCTURE ONLY.

Restates mrmustard/math/lattice/strategies/binomial.py:30-74: the tensor
starts as zeros with G[0] = c and norm = |c|^2; for photons = 1 ..
global_cutoff - 1 the subspace of that weight is filled and its squared norm
added, stopping after the first weight with norm > max_l2 (max_l2 None:
never).  The values of a subspace are vanilla_step (steps.py:37-67), the
first-pivot step of the vanilla recurrence, whose terms and order are those
of vanilla_numba (the extra terms of lower_neighbors for axes below the pivot
are A_ij * SQRT[0] * G[...] = +-0), so they are taken from the pinned vanilla
oracle.  Only ``tests/`` may import this module.
"""

from __future__ import annotations

import numpy as np

from . import vanilla


def binomial(local_cutoffs, A, b, c, max_l2, global_cutoff):
    shape = tuple(int(s) for s in local_cutoffs)
    full = vanilla(shape, A, b, c)
    weight = np.sum(np.indices(shape), axis=0)
    G = np.zeros(shape, dtype=np.complex128)
    G.flat[0] = c
    norm = abs(complex(c)) ** 2
    for photons in range(1, global_cutoff):
        sel = weight == photons
        G[sel] = full[sel]
        norm += float(np.sum(np.abs(full[sel]) ** 2))
        if max_l2 is not None and norm > max_l2:
            break
    return G, norm
