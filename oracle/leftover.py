"""CPU oracle of the one-leftover-mode strategy.  TEST INFRASTRUCTURE ONLY.

Restates mrmustard/math/lattice/strategies/compactFock/singleLeftoverMode_amps.py
(write_block :20-72, use_offDiag_pivot :99-221, use_diag_pivot :224-285, the
leftover block at zero detections and the weight loop :288-366, the
initialisation :369-421) in plain Python over numpy arrays, for small cutoffs.
Mode 0 keeps a full (c0, c0) block; modes 1.. are PNR-detected.  A and B are
in the compact order [mode0, mode0, mode1, mode1, ...].  ``leftover`` applies
the [modes..., modes...] -> compact reorder and the (output_cutoff + 1,
pnr_cutoffs + 1) shape convention of hermite_renormalized_1leftoverMode
(jax_vjps/hermite.py:343-387, backend_numpy.py:434-446).  Only ``tests/`` may
import it.  Pinned in ``tests/test_leftover.py`` against the full lattice of the
pinned vanilla oracle (test_compactFock.py:46-69, test_fast_diagonal.py:24-45)
and against product-state closed forms.  ``leftover_jacobian`` / ``leftover_vjp``
restate the forward-mode gradient (singleLeftoverMode_grad.py) on dual numbers
and are pinned there by central differences of ``leftover_compact``, per entry
of A, B and c.
"""

from __future__ import annotations

import numpy as np

from .diagonal import reorder_ab

SQRT = np.sqrt(np.arange(1000))


def _block(i, pivot, G_in, GB, A, K):
    """write_block (singleLeftoverMode_amps.py:20-72): output index i >= 2 of
    every (m, n) of the leftover block; ``pivot`` is the pivot's own block."""
    c0 = pivot.shape[0]
    out = np.zeros((c0, c0), np.complex128)
    for m in range(c0):
        for n in range(c0):
            acc = GB[m, n, i] + A[i, 2:] @ G_in[m, n]
            if m > 0:
                acc += A[i, 0] * pivot[m - 1, n] * np.sqrt(m)
            if n > 0:
                acc += A[i, 1] * pivot[m, n - 1] * np.sqrt(n)
            out[m, n] = acc / K
    return out


def leftover_compact(A, B, G0, cutoffs):
    """arr0 of fock_representation_1leftoverMode_amps: shape (c0, c0, *cutoffs[1:])."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    cutoffs = tuple(int(x) for x in cutoffs)
    c0, tail = cutoffs[0], cutoffs[1:]
    Mt = len(tail)
    arr0 = np.zeros((c0, c0, *tail), np.complex128)
    arr2 = np.zeros((c0, c0, Mt, *tail), np.complex128)
    arr1 = np.zeros((c0, c0, 2 * Mt, *tail), np.complex128)
    arr1010 = np.zeros((c0, c0, Mt, max(Mt - 1, 1), *tail), np.complex128)
    arr1001 = np.zeros((c0, c0, Mt, max(Mt - 1, 1), *tail), np.complex128)
    z = (0,) * Mt
    arr0[(0, 0) + z] = G0
    for m in range(c0 - 1):
        prev = arr0[(m - 1, 0) + z] if m > 0 else 0.0
        arr0[(m + 1, 0) + z] = (arr0[(m, 0) + z] * B[0] + np.sqrt(m) * A[0, 0] * prev) / np.sqrt(m + 1)
    for m in range(c0):
        for n in range(c0 - 1):
            pm = arr0[(m - 1, n) + z] if m > 0 else 0.0
            pn = arr0[(m, n - 1) + z] if n > 0 else 0.0
            arr0[(m, n + 1) + z] = (arr0[(m, n) + z] * B[1] + np.sqrt(m) * A[1, 0] * pm + np.sqrt(n) * A[1, 1] * pn) / np.sqrt(n + 1)

    def dec(p, i):
        q = list(p)
        q[i] -= 1
        return tuple(q)

    by_weight = {}
    for p in np.ndindex(tail):
        by_weight.setdefault(sum(p), []).append(p)
    for w in range(sum(tail)):
        for p in by_weight.get(w, []):
            pivot = np.repeat(np.array(p), 2)
            if tail[0] == 1 or p[0] < tail[0] - 1:
                K_l, K_i = SQRT[pivot], SQRT[pivot + 1]
                G_in = np.zeros((c0, c0, 2 * Mt), np.complex128)
                piv = arr0[(slice(None), slice(None)) + p]
                GB = piv[:, :, None] * B[None, None, :]
                for i in range(2 * Mt):
                    if p[i // 2] > 0:
                        G_in[:, :, i] = arr1[(slice(None), slice(None), i + 1 - 2 * (i % 2)) + dec(p, i // 2)]
                G_in = G_in * K_l
                for i in range(2 * Mt):
                    if p[i // 2] + 1 < tail[i // 2] and (i != 1 or p[0] + 2 < tail[0]):
                        arr1[(slice(None), slice(None), i) + p] = _block(i + 2, piv, G_in, GB, A, K_i[i])
            for d in range(Mt):
                if all(x == 0 for x in p[:d]) and p[d] < tail[d] - 1:
                    pv = pivot.copy()
                    pv[2 * d] += 1
                    K_l, K_i = SQRT[pv], SQRT[pv + 1]
                    piv = arr1[(slice(None), slice(None), 2 * d) + p]
                    GB = piv[:, :, None] * B[None, None, :]
                    G_in = np.zeros((c0, c0, 2 * Mt), np.complex128)
                    G_in[:, :, 2 * d] = arr0[(slice(None), slice(None)) + p]
                    if p[d] > 0:
                        G_in[:, :, 2 * d + 1] = arr2[(slice(None), slice(None), d) + dec(p, d)]
                    for i in range(d + 1, Mt):
                        if p[i] > 0:
                            G_in[:, :, 2 * i] = arr1001[(slice(None), slice(None), d, i - d - 1) + dec(p, i)]
                            G_in[:, :, 2 * i + 1] = arr1010[(slice(None), slice(None), d, i - d - 1) + dec(p, i)]
                    G_in = G_in * K_l
                    up = list(p)
                    up[d] += 1
                    arr0[(slice(None), slice(None)) + tuple(up)] = _block(2 * d + 3, piv, G_in, GB, A, K_i[2 * d + 1])
                    if p[d] + 2 < tail[d]:
                        arr2[(slice(None), slice(None), d) + p] = _block(2 * d + 2, piv, G_in, GB, A, K_i[2 * d])
                    for i in range(d + 1, Mt):
                        if p[i] + 1 < tail[i]:
                            arr1010[(slice(None), slice(None), d, i - d - 1) + p] = _block(2 * i + 2, piv, G_in, GB, A, K_i[2 * i])
                            arr1001[(slice(None), slice(None), d, i - d - 1) + p] = _block(2 * i + 3, piv, G_in, GB, A, K_i[2 * i + 1])
    return arr0


def leftover(A, b, c, output_cutoff, pnr_cutoffs):
    """hermite_renormalized_1leftoverMode with reorderedAB: shape
    (output_cutoff + 1, output_cutoff + 1, *(pnr_cutoffs + 1))."""
    Ar, br = reorder_ab(A, b)
    cut = (int(output_cutoff) + 1,) + tuple(int(k) + 1 for k in pnr_cutoffs)
    return leftover_compact(Ar, br, complex(c), cut)


def leftover_jacobian(A, B, G0, cutoffs):
    """arr0 with its derivatives, as fock_representation_1leftoverMode_grad
    computes them (compactFock/singleLeftoverMode_grad.py:19-46 calc_dA_dB,
    :50-225 write_block_grad, :560-692 including the zero-detection block):
    every value carries [value, d/dA (2L x 2L, entries independent), d/dB
    (2L)] through leftover_compact's recursion.  Returns (arr0, dA, dB)."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    cutoffs = tuple(int(x) for x in cutoffs)
    c0, tail = cutoffs[0], cutoffs[1:]
    Mt = len(tail)
    G = A.shape[0]
    N = 1 + G * G + G
    arr0 = np.zeros((c0, c0, *tail, N), np.complex128)
    arr2 = np.zeros((c0, c0, Mt, *tail, N), np.complex128)
    arr1 = np.zeros((c0, c0, 2 * Mt, *tail, N), np.complex128)
    arr1010 = np.zeros((c0, c0, Mt, max(Mt - 1, 1), *tail, N), np.complex128)
    arr1001 = np.zeros((c0, c0, Mt, max(Mt - 1, 1), *tail, N), np.complex128)
    z = (0,) * Mt
    dA = lambda r, l: 1 + r * G + l  # noqa: E731
    dB = lambda r: 1 + G * G + r  # noqa: E731

    def term(coef, r, l, x):
        """coef * A[r, l] * x on duals, plus d/dA[r, l] = coef * x."""
        t = coef * A[r, l] * x
        t[dA(r, l)] += coef * x[0]
        return t

    arr0[(0, 0) + z + (0,)] = G0
    for m in range(c0 - 1):
        t = arr0[(m, 0) + z] * B[0]
        t[dB(0)] += arr0[(m, 0) + z][0]
        if m > 0:
            t = t + term(np.sqrt(m), 0, 0, arr0[(m - 1, 0) + z])
        arr0[(m + 1, 0) + z] = t / np.sqrt(m + 1)
    for m in range(c0):
        for n in range(c0 - 1):
            t = arr0[(m, n) + z] * B[1]
            t[dB(1)] += arr0[(m, n) + z][0]
            if m > 0:
                t = t + term(np.sqrt(m), 1, 0, arr0[(m - 1, n) + z])
            if n > 0:
                t = t + term(np.sqrt(n), 1, 1, arr0[(m, n - 1) + z])
            arr0[(m, n + 1) + z] = t / np.sqrt(n + 1)

    def block(i, pivot, G_in, K):
        """write_block on duals: output index i >= 2 of every (m, n)."""
        out = np.zeros((c0, c0, N), np.complex128)
        for m in range(c0):
            for n in range(c0):
                t = pivot[m, n] * B[i]
                t[dB(i)] += pivot[m, n][0]
                for l in range(2 * Mt):
                    t = t + term(1.0, i, l + 2, G_in[m, n, l])
                if m > 0:
                    t = t + term(np.sqrt(m), i, 0, pivot[m - 1, n])
                if n > 0:
                    t = t + term(np.sqrt(n), i, 1, pivot[m, n - 1])
                out[m, n] = t / K
        return out

    def dec(p, i):
        q = list(p)
        q[i] -= 1
        return tuple(q)

    S_ = (slice(None), slice(None))
    by_weight = {}
    for p in np.ndindex(tail):
        by_weight.setdefault(sum(p), []).append(p)
    for w in range(sum(tail)):
        for p in by_weight.get(w, []):
            pivot = np.repeat(np.array(p), 2)
            if tail[0] == 1 or p[0] < tail[0] - 1:
                K_l, K_i = SQRT[pivot], SQRT[pivot + 1]
                G_in = np.zeros((c0, c0, 2 * Mt, N), np.complex128)
                piv = arr0[S_ + p]
                for i in range(2 * Mt):
                    if p[i // 2] > 0:
                        G_in[:, :, i] = arr1[S_ + (i + 1 - 2 * (i % 2),) + dec(p, i // 2)]
                G_in = G_in * K_l[None, None, :, None]
                for i in range(2 * Mt):
                    if p[i // 2] + 1 < tail[i // 2] and (i != 1 or p[0] + 2 < tail[0]):
                        arr1[S_ + (i,) + p] = block(i + 2, piv, G_in, K_i[i])
            for d in range(Mt):
                if all(x == 0 for x in p[:d]) and p[d] < tail[d] - 1:
                    pv = pivot.copy()
                    pv[2 * d] += 1
                    K_l, K_i = SQRT[pv], SQRT[pv + 1]
                    piv = arr1[S_ + (2 * d,) + p]
                    G_in = np.zeros((c0, c0, 2 * Mt, N), np.complex128)
                    G_in[:, :, 2 * d] = arr0[S_ + p]
                    if p[d] > 0:
                        G_in[:, :, 2 * d + 1] = arr2[S_ + (d,) + dec(p, d)]
                    for i in range(d + 1, Mt):
                        if p[i] > 0:
                            G_in[:, :, 2 * i] = arr1001[S_ + (d, i - d - 1) + dec(p, i)]
                            G_in[:, :, 2 * i + 1] = arr1010[S_ + (d, i - d - 1) + dec(p, i)]
                    G_in = G_in * K_l[None, None, :, None]
                    up = list(p)
                    up[d] += 1
                    arr0[S_ + tuple(up)] = block(2 * d + 3, piv, G_in, K_i[2 * d + 1])
                    if p[d] + 2 < tail[d]:
                        arr2[S_ + (d,) + p] = block(2 * d + 2, piv, G_in, K_i[2 * d])
                    for i in range(d + 1, Mt):
                        if p[i] + 1 < tail[i]:
                            arr1010[S_ + (d, i - d - 1) + p] = block(2 * i + 2, piv, G_in, K_i[2 * i])
                            arr1001[S_ + (d, i - d - 1) + p] = block(2 * i + 3, piv, G_in, K_i[2 * i + 1])
    val = arr0[..., 0]
    return val, arr0[..., 1:1 + G * G].reshape(val.shape + (G, G)), arr0[..., 1 + G * G:]


def leftover_vjp(A, B, G0, cutoffs, cot):
    """(dA, dB, dG0) = sum cot * d arr0 / d(A, B, G0), holomorphic, in the
    compact order, as the JAX backward contracts the Jacobian
    (jax_vjps/hermite.py:397-440); dG0 = sum cot * arr0 / G0."""
    val, dA, dB = leftover_jacobian(A, B, G0, cutoffs)
    cot = np.asarray(cot, np.complex128)
    ax = tuple(range(val.ndim))
    return np.tensordot(cot, dA, axes=(ax, ax)), np.tensordot(cot, dB, axes=(ax, ax)), np.sum(cot * val) / complex(G0)
