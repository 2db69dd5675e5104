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


def diagonal_jacobian(A, B, G0, cutoffs):
    """arr0 with its derivatives, as fock_representation_diagonal_grad computes
    them (compactFock/diagonal_grad.py:19-38 calc_dA_dB, :41-258, :357-384):
    every stored value carries [value, d/dA (2M x 2M, entries independent),
    d/dB (2M)] through the same recursion.  Returns (arr0, dA, dB) with shapes
    cutoffs, cutoffs + (2M, 2M), cutoffs + (2M,)."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    cutoffs = tuple(int(x) for x in cutoffs)
    M = len(cutoffs)
    G = 2 * M
    N = 1 + G * G + G
    arr0 = np.zeros(cutoffs + (N,), np.complex128)
    arr2 = np.zeros((M, *cutoffs, N), np.complex128)
    arr1 = np.zeros((G, *cutoffs, N), np.complex128)
    arr1010 = np.zeros((M, max(M - 1, 1), *cutoffs, N), np.complex128)
    arr1001 = np.zeros((M, max(M - 1, 1), *cutoffs, N), np.complex128)
    arr0[(0,) * M + (0,)] = G0

    def out(i, g, G_in, K):
        """(g B_i + A_i . G_in) / K on duals, plus d/dB_i = g and d/dA_il = G_in_l."""
        o = g * B[i] + A[i] @ G_in
        o[1 + i * G: 1 + (i + 1) * G] += G_in[:, 0]
        o[1 + G * G + i] += g[0]
        return o / K

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
                G_in = np.zeros((G, N), np.complex128)
                for i in range(G):
                    if p[i // 2] > 0:
                        G_in[i] = arr1[(i + 1 - 2 * (i % 2), *dec(p, i // 2))]
                G_in = K_l[:, None] * G_in
                for i in range(G):
                    if p[i // 2] + 1 < cutoffs[i // 2] and (i != 1 or p[0] + 2 < cutoffs[0]):
                        arr1[(i, *p)] = out(i, arr0[p], G_in, K_i[i])
            for d in range(M):
                if all(x == 0 for x in p[:d]) and p[d] < cutoffs[d] - 1:
                    pv = pivot.copy()
                    pv[2 * d] += 1
                    K_l, K_i = SQRT[pv], SQRT[pv + 1]
                    G_in = np.zeros((G, N), np.complex128)
                    g = arr1[(2 * d, *p)]
                    G_in[2 * d] = arr0[p]
                    if p[d] > 0:
                        G_in[2 * d + 1] = arr2[(d, *dec(p, d))]
                    for i in range(d + 1, M):
                        if p[i] > 0:
                            G_in[2 * i] = arr1001[(d, i - d - 1, *dec(p, i))]
                            G_in[2 * i + 1] = arr1010[(d, i - d - 1, *dec(p, i))]
                    G_in = K_l[:, None] * G_in
                    up = list(p)
                    up[d] += 1
                    arr0[tuple(up)] = out(2 * d + 1, g, G_in, K_i[2 * d + 1])
                    if p[d] + 2 < cutoffs[d]:
                        arr2[(d, *p)] = out(2 * d, g, G_in, K_i[2 * d])
                    for i in range(d + 1, M):
                        if p[i] + 1 < cutoffs[i]:
                            arr1010[(d, i - d - 1, *p)] = out(2 * i, g, G_in, K_i[2 * i])
                            arr1001[(d, i - d - 1, *p)] = out(2 * i + 1, g, G_in, K_i[2 * i + 1])
    val = arr0[..., 0]
    dA = arr0[..., 1:1 + G * G].reshape(cutoffs + (G, G))
    dB = arr0[..., 1 + G * G:]
    return val, dA, dB


def diagonal_vjp(A, B, G0, cutoffs, cot):
    """(dA, dB, dG0) = sum_p cot[p] d arr0[p] / d(A, B, G0), holomorphic, in the
    order A and B are given in; dG0 = sum cot * arr0 / G0 as
    grad_hermite_multidimensional_diagonal (inputValidation.py:82-100) forms it
    and hermite_renormalized_diagonal_jax_bwd (jax_vjps/hermite.py:292-330)
    contracts it."""
    val, dA, dB = diagonal_jacobian(A, B, G0, cutoffs)
    cot = np.asarray(cot, np.complex128)
    M = len(cutoffs)
    ax = tuple(range(M))
    return (np.tensordot(cot, dA, axes=(ax, ax)), np.tensordot(cot, dB, axes=(ax, ax)), np.sum(cot * val) / complex(G0))
