"""Independent high-precision reference for tiny lattices (test helper).

The renormalized Hermite tensor is defined by its generating function
    sum_k G_k z^k / sqrt(k!) = c * exp(1/2 z^T A z + b^T z),
(docstring of backend_manager.py:652-655 and core.py:40-43).  This module
expands the right-hand side as a truncated multivariate power series in
mpmath at 50 significant digits, so it shares no code or operation order with
the recurrence it checks.  ``vanilla_numba`` reads only the upper triangle of
``A`` (core.py:101-103), so the symmetric matrix used here is
``triu(A) + triu(A, 1).T``.
"""

from __future__ import annotations

import itertools
import math as pymath

import mpmath
import numpy as np

mpmath.mp.dps = 50


def _mul(p, q, shape):
    out = {}
    for ka, va in p.items():
        for kb, vb in q.items():
            k = tuple(x + y for x, y in zip(ka, kb))
            if all(ki < si for ki, si in zip(k, shape)):
                out[k] = out.get(k, 0) + va * vb
    return out


def hermite_series(A, b, c, shape, upper_only: bool = True) -> np.ndarray:
    """G on the box ``shape`` from the power series of c*exp(z^T A z/2 + b^T z)."""
    A = np.asarray(A, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    D = len(shape)
    if upper_only:
        A = np.triu(A) + np.triu(A, 1).T
    mpc = lambda x: mpmath.mpc(complex(x).real, complex(x).imag)  # noqa: E731
    P = {}
    unit = lambda *axes: tuple(sum(1 for a in axes if a == i) for i in range(D))  # noqa: E731
    for i in range(D):
        if shape[i] > 1:
            P[unit(i)] = P.get(unit(i), 0) + mpc(b[i])
        if shape[i] > 2:
            P[unit(i, i)] = P.get(unit(i, i), 0) + mpc(A[i, i]) / 2
        for j in range(i + 1, D):
            if shape[i] > 1 and shape[j] > 1:
                P[unit(i, j)] = P.get(unit(i, j), 0) + mpc(A[i, j])
    total = sum(s - 1 for s in shape)
    series = {tuple([0] * D): mpmath.mpc(1)}
    term = {tuple([0] * D): mpmath.mpc(1)}
    for n in range(1, total + 1):
        term = _mul(term, P, shape)
        term = {k: v / n for k, v in term.items()}
        for k, v in term.items():
            series[k] = series.get(k, 0) + v
    cc = mpc(np.asarray(c).reshape(()))
    G = np.zeros(shape, dtype=np.complex128)
    for k in itertools.product(*[range(s) for s in shape]):
        fact = 1
        for ki in k:
            fact *= pymath.factorial(ki)
        v = cc * series.get(k, 0) * mpmath.sqrt(fact)
        G[k] = complex(v)
    return G


def stable_tuple_recursion(A, b, c, shape) -> np.ndarray:
    """The all-pivot average of stable_numba (core.py:164-213) written on index
    tuples in mpmath at 50 digits: for every k != 0,
        G_k = mean over i with k_i > 0 of
              ( b_i G_{k-e_i} + sum_j A[i, j] sqrt(k_j - delta_ij) G_{k-e_i-e_j} ) / sqrt(k_i),
    reading row i of A as given (both triangles).  Terms whose sqrt factor is
    zero are left out (the reference multiplies them by SQRT[0] = 0).  No
    strides, flat indices or operation order are shared with the oracle."""
    A = np.asarray(A, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    D = len(shape)
    mpc = lambda x: mpmath.mpc(complex(x).real, complex(x).imag)  # noqa: E731
    Am = [[mpc(A[i, j]) for j in range(D)] for i in range(D)]
    bm = [mpc(b[i]) for i in range(D)]
    G = {tuple([0] * D): mpc(np.asarray(c).reshape(()))}

    def get(k):
        return G[k] if all(x >= 0 for x in k) else mpmath.mpc(0)

    for k in itertools.product(*[range(s) for s in shape]):
        if not any(k):
            continue
        acc, n = mpmath.mpc(0), 0
        for i in range(D):
            if k[i] == 0:
                continue
            n += 1
            p = tuple(x - (a == i) for a, x in enumerate(k))
            val = bm[i] * get(p)
            for j in range(D):
                w = k[j] - (1 if j == i else 0)
                if w <= 0:
                    continue
                q = tuple(x - (a == j) for a, x in enumerate(p))
                val += Am[i][j] * mpmath.sqrt(w) * get(q)
            acc += val / mpmath.sqrt(k[i])
        G[k] = acc / n
    out = np.zeros(shape, dtype=np.complex128)
    for k, v in G.items():
        out[k] = complex(v)
    return out
