"""Displacement-gate matrix elements from the generalized Laguerre recurrence.
TEST INFRASTRUCTURE ONLY (the checker of the large-cutoff stable tests).

Restates ``displacement`` and ``laguerre`` of
mrmustard/math/lattice/strategies/displacement.py:24-80: for each diagonal
n - m the first generalized Laguerre polynomials L_m^{(n-m)}(|alpha|^2) by
their three-term recurrence, then
    D[n, m] = exp( (log m! - log n!)/2 + (n-m) log|alpha| - |alpha|^2/2
                   + i (n-m) arg(alpha) + log L_m^{(n-m)}(|alpha|^2) ),
    D[m, n] = (-1)^(n-m) conj(D[n, m]),
in logarithms so that no intermediate overflows at cutoffs of 1000.  The
recurrence runs over all diagonals at once (numpy vectors over n - m) instead
of one diagonal at a time; each element sees the same operations.

Pinned in tests/test_oracle.py by the closed form
    <n|D(alpha)|m> = sqrt(m!/n!) alpha^(n-m) e^(-|alpha|^2/2) L_m^(n-m)(|alpha|^2)
in 50-digit arithmetic, and by unitarity of the truncated matrix's
well-converged block.
"""

from __future__ import annotations

import numpy as np


def laguerre_table(x: float, n_max: int, alphas: np.ndarray) -> np.ndarray:
    """L[a, m] = L_m^{(alphas[a])}(x) for m < n_max (displacement.py:63-80)."""
    alphas = np.asarray(alphas, dtype=np.float64)
    L = np.zeros((len(alphas), max(n_max, 1)), np.complex128)
    L[:, 0] = 1.0
    if n_max > 1:
        L[:, 1] = (1.0 + alphas - x) * L[:, 0]
        # entries with alpha + m beyond the matrix are never read and may overflow
        with np.errstate(over="ignore", invalid="ignore"):
            for m in range(1, n_max - 1):
                L[:, m + 1] = ((2 * m + 1 + alphas - x) * L[:, m] - (m + alphas) * L[:, m - 1]) / (m + 1)
    return L


def displacement(cutoffs, alpha) -> np.ndarray:
    """Matrix elements <n|D(alpha)|m>, n < cutoffs[0], m < cutoffs[1]
    (displacement.py:24-60)."""
    r = abs(alpha)
    phi = np.angle(alpha)
    N, M = int(cutoffs[0]), int(cutoffs[1])
    flipped = N < M
    if flipped:
        N, M = M, N
    D = np.zeros((N, M), np.complex128)
    rng = np.arange(max(N, M), dtype=np.float64)
    rng[0] = 1.0
    log_k_fac = np.cumsum(np.log(rng))
    L = laguerre_table(r**2, M, np.arange(N))  # L[n - m, m]
    with np.errstate(divide="ignore", invalid="ignore"):
        logL = np.log(L)
        log_r = np.log(r)
    for d in range(N):  # d = n - m
        m_max = min(M, N - d)
        m = np.arange(m_max)
        n = d + m
        sign = np.where(flipped & (n > m) & (d % 2 == 1), -1.0, 1.0)
        conj = np.where(flipped & (n > m), -1.0, 1.0)
        with np.errstate(invalid="ignore"):
            D[n, m] = sign * np.exp(
                0.5 * (log_k_fac[m] - log_k_fac[n]) + d * log_r - r**2 / 2.0 + conj * 1j * phi * d + logL[d, :m_max]
            )
        low = n < M
        D[m[low], n[low]] = (-1.0) ** d * np.conj(D[n[low], m[low]])
    return D.T if flipped else D
