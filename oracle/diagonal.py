"""CPU oracle of the diagonal strategy.  TEST INFRASTRUCTURE ONLY.

Restates mrmustard/math/lattice/strategies/compactFock/diagonal_amps.py
(use_diag_pivot :108-150, use_offDiag_pivot :19-105, the weight loop
:153-197, the initialisation :200-248) in plain Python over numpy arrays, for
small cutoffs.  A and B are in the compact ordering [mode0, mode0, mode1,
mode1, ...] (backend_numpy.py:367-375 reorders from [modes..., modes...]).
Only ``tests/`` may import it.  Pinned in ``tests/test_diagonal.py`` against
the diagonal of the full Fock tensor from the pinned vanilla oracle, as the
reference's own test_compactFock.py:19-44 does.
"""

from __future__ import annotations

import numpy as np

SQRT = np.sqrt(np.arange(1000))


def reorder_ab(A, B):
    """[modes..., modes...] -> [mode0, mode0, mode1, mode1, ...] (backend_numpy.py:367-375)."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    order = np.arange(A.shape[0]).reshape(2, -1).T.flatten()
    return A[order][:, order], B[order]


def diagonal(A, B, G0, cutoffs):
    """arr0 of fock_representation_diagonal_amps: the photon-number diagonal, shape ``cutoffs``."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    cutoffs = tuple(int(x) for x in cutoffs)
    M = len(cutoffs)
    arr0 = np.zeros(cutoffs, np.complex128)
    arr2 = np.zeros((M, *cutoffs), np.complex128)
    arr1 = np.zeros((2 * M, *cutoffs), np.complex128)
    if M == 1:
        arr1010 = np.zeros((1, 1, 1), np.complex128)
        arr1001 = np.zeros((1, 1, 1), np.complex128)
    else:
        arr1010 = np.zeros((M, M - 1, *cutoffs), np.complex128)
        arr1001 = np.zeros((M, M - 1, *cutoffs), np.complex128)
    arr0[(0,) * M] = G0

    def dec(p, i):
        q = list(p)
        q[i] -= 1
        return tuple(q)

    by_weight = {}
    for p in np.ndindex(cutoffs):
        by_weight.setdefault(sum(p), []).append(p)
    for w in range(sum(cutoffs)):
        for p in by_weight.get(w, []):
            pivot = np.repeat(np.array(p), 2)
            if cutoffs[0] == 1 or p[0] < cutoffs[0] - 1:
                K_l, K_i = SQRT[pivot], SQRT[pivot + 1]
                G_in = np.zeros(2 * M, np.complex128)
                GB = arr0[p] * B
                for i in range(2 * M):
                    if p[i // 2] > 0:
                        G_in[i] = arr1[(i + 1 - 2 * (i % 2), *dec(p, i // 2))]
                G_in = K_l * G_in
                for i in range(2 * M):
                    if p[i // 2] + 1 < cutoffs[i // 2] and (i != 1 or p[0] + 2 < cutoffs[0]):
                        arr1[(i, *p)] = (GB[i] + A[i] @ G_in) / K_i[i]
            for d in range(M):
                if all(x == 0 for x in p[:d]) and p[d] < cutoffs[d] - 1:
                    pv = pivot.copy()
                    pv[2 * d] += 1
                    K_l, K_i = SQRT[pv], SQRT[pv + 1]
                    G_in = np.zeros(2 * M, np.complex128)
                    GB = arr1[(2 * d, *p)] * B
                    G_in[2 * d] = arr0[p]
                    if p[d] > 0:
                        G_in[2 * d + 1] = arr2[(d, *dec(p, d))]
                    for i in range(d + 1, M):
                        if p[i] > 0:
                            G_in[2 * i] = arr1001[(d, i - d - 1, *dec(p, i))]
                            G_in[2 * i + 1] = arr1010[(d, i - d - 1, *dec(p, i))]
                    G_in = K_l * G_in
                    up = list(p)
                    up[d] += 1
                    arr0[tuple(up)] = (GB[2 * d + 1] + A[2 * d + 1] @ G_in) / K_i[2 * d + 1]
                    if p[d] + 2 < cutoffs[d]:
                        arr2[(d, *p)] = (GB[2 * d] + A[2 * d] @ G_in) / K_i[2 * d]
                    for i in range(d + 1, M):
                        if p[i] + 1 < cutoffs[i]:
                            arr1010[(d, i - d - 1, *p)] = (GB[2 * i] + A[2 * i] @ G_in) / K_i[2 * i]
                            arr1001[(d, i - d - 1, *p)] = (GB[2 * i + 1] + A[2 * i + 1] @ G_in) / K_i[2 * i + 1]
    return arr0
