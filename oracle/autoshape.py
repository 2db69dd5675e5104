"""CPU oracle of ``autoshape_numba``.  TEST INFRASTRUCTURE ONLY.

Restates mrmustard/math/lattice/autoshape.py:74-154 with numpy (the
reference's own ``np.linalg.inv`` / ``np.linalg.det`` for the Schur
complements, plain Python loops for the diagonal recursion).  Only ``tests/``
may import it.  Pinned in ``tests/test_autoshape.py`` by the Poisson and
squeezed-vacuum photon-number distributions (closed forms), by the cutoffs
the reference's own tests assert (test_ket.py:98-104, test_dm.py:82-86,
test_branch_and_bound.py:40-54, test_coherent.py:75), and by a brute-force
marginal: the Fock density matrix of the whole state from the pinned forward
oracle, traced down to one mode.
"""

from __future__ import annotations

import numpy as np

SQRT = np.sqrt(np.arange(100000))


def single_mode_marginals(A, b, c):
    """(A_m, b_m, c_m) of the single-mode reduced density matrix of every mode
    m, as autoshape.py:77-108 forms them: A is 2M x 2M over (bra modes, ket
    modes), b has 2M entries, c is the vacuum amplitude."""
    A = np.asarray(A, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    M = len(b) // 2
    A4 = A.reshape((2, M, 2, M)).transpose((1, 3, 0, 2))  # (M, M, 2, 2)
    b2 = b.reshape((2, M)).transpose()  # (M, 2)
    zero = np.zeros((M - 1, M - 1), dtype=np.complex128)
    ident = np.eye(M - 1, dtype=np.complex128)
    X = np.vstack((np.hstack((zero, ident)), np.hstack((ident, zero))))
    out = []
    for m in range(M):
        idx_m = np.array([m])
        idx_n = np.delete(np.arange(M), m)
        A_mm = np.ascontiguousarray(A4[idx_m, :][:, idx_m].transpose((2, 0, 3, 1))).reshape((2, 2))
        A_nn = np.ascontiguousarray(A4[idx_n, :][:, idx_n].transpose((2, 0, 3, 1))).reshape((2 * M - 2, 2 * M - 2))
        A_mn = np.ascontiguousarray(A4[idx_m, :][:, idx_n].transpose((2, 0, 3, 1))).reshape((2, 2 * M - 2))
        A_nm = np.transpose(A_mn)
        b_m = np.ascontiguousarray(b2[idx_m].transpose()).reshape((2,))
        b_n = np.ascontiguousarray(b2[idx_n].transpose()).reshape((2 * M - 2,))
        R_inv = np.linalg.inv(A_nn - X)
        A_ = A_mm - A_mn @ R_inv @ A_nm
        b_ = b_m - A_mn @ R_inv @ b_n
        c_ = c * np.exp(-0.5 * b_n @ R_inv @ b_n) / np.sqrt(np.linalg.det(A_nn - X))
        out.append((A_, b_, c_))
    return out


def diagonal_steps(A_, b_, c_, max_prob, max_shape):
    """The early-stopped diagonal recursion of autoshape.py:110-153 for one
    mode; returns (k, norm): the step count and the probability captured."""
    buf2 = np.zeros((2, 2), dtype=np.complex128)
    buf3 = np.zeros((2, 3), dtype=np.complex128)
    buf3[0, 1] = c_
    norm = np.abs(c_)
    k = 0
    while norm < max_prob and k < max_shape:
        buf2[(k + 1) % 2] = (b_ * buf3[k % 2, 1] + A_ @ buf2[k % 2] * SQRT[k]) / SQRT[k + 1]
        buf3[(k + 1) % 2, 0] = (
            b_[0] * buf2[(k + 1) % 2, 0] + A_[0, 0] * buf3[k % 2, 1] * SQRT[k + 1] + A_[0, 1] * buf3[k % 2, 0] * SQRT[k]
        ) / SQRT[k + 2]
        buf3[(k + 1) % 2, 1] = (
            b_[1] * buf2[(k + 1) % 2, 0] + A_[1, 0] * buf3[k % 2, 1] * SQRT[k + 1] + A_[1, 1] * buf3[k % 2, 0] * SQRT[k]
        ) / SQRT[k + 1]
        buf3[(k + 1) % 2, 2] = (
            b_[1] * buf2[(k + 1) % 2, 1] + A_[1, 0] * buf3[k % 2, 2] * SQRT[k] + A_[1, 1] * buf3[k % 2, 1] * SQRT[k + 1]
        ) / SQRT[k + 2]
        norm += np.abs(buf3[(k + 1) % 2, 1])
        k += 1
    return k, norm


def autoshape(A, b, c, max_prob, max_shape, min_shape):
    """autoshape_numba (autoshape.py:23-154): per-mode cutoffs, clipped."""
    M = len(b) // 2
    shape = np.ones(M, dtype=np.int64)
    for m, (A_, b_, c_) in enumerate(single_mode_marginals(A, b, c)):
        shape[m] = diagonal_steps(A_, b_, c_, max_prob, max_shape)[0]
    return np.clip(shape, min_shape, max_shape)


def margins(A, b, c, max_prob, max_shape):
    """For each mode, how far the captured probability is from max_prob at
    the last two steps (small values mean a cutoff that rounding could move)."""
    out = []
    for A_, b_, c_ in single_mode_marginals(A, b, c):
        k, norm = diagonal_steps(A_, b_, c_, max_prob, max_shape)
        km, normm = diagonal_steps(A_, b_, c_, max_prob, max(k - 1, 0)) if k > 0 else (0, norm)
        out.append(min(abs(norm - max_prob), abs(normm - max_prob)))
    return np.array(out)
