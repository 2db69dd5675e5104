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


# ------------------------------------------------------------------------
# CPython complex arithmetic on Python lists (non-finite values).
#
# numba compiles the reference's complex products and quotients to CPython's
# own algorithms (_Py_c_prod, _Py_c_quot, a float promoted to complex(x, 0.0))
# and wraps negative array indices as numpy does.  The two functions below
# run the reference's loops (core.py:64-124, gradients.py:41-82) in plain
# Python complex numbers on lists, where these rules are the interpreter's
# own rather than the C oracle's restatement of them, so that the oracle's
# non-finite results (0 * inf and inf - inf giving NaN, the NaN part of an
# infinite number divided by a float) can be checked on small lattices.


def _strides(shape):
    st = [1] * len(shape)
    for i in range(len(shape) - 1, 0, -1):
        st[i - 1] = st[i] * shape[i]
    return st


def python_complex_vanilla(A, b, c, shape) -> np.ndarray:
    A = [[complex(x) for x in row] for row in np.asarray(A, dtype=np.complex128)]
    b = [complex(x) for x in np.asarray(b, dtype=np.complex128)]
    D = len(shape)
    st = _strides(shape)
    N = int(np.prod(shape))
    SQRT = [pymath.sqrt(n) for n in range(max(shape) + 2)]
    SQRT_m1 = pymath.sqrt(99999)  # SQRT[-1] of the reference's table
    sq = lambda n: SQRT[n] if n >= 0 else SQRT_m1  # noqa: E731
    G = [0j] * N
    G[0] = complex(c)
    for flat, index in enumerate(itertools.product(*[range(s) for s in shape])):
        if flat == 0:
            continue
        i = next(a for a in range(D) if flat - st[a] >= 0)
        pivot = flat - st[i]
        v = b[i] * G[pivot]
        v += A[i][i] * sq(index[i] - 1) * G[pivot - st[i]]
        for j in range(i + 1, D):
            v += A[i][j] * sq(index[j]) * G[pivot - st[j]]
        G[flat] = v / sq(index[i])
    return np.array(G, dtype=np.complex128).reshape(shape)


def python_complex_vjp(G, c, g):
    shape = G.shape
    D = len(shape)
    st = _strides(shape)
    Gl = [complex(x) for x in G.ravel()]
    gl = [complex(x) for x in g.ravel()]
    SQRT = [pymath.sqrt(n) for n in range(max(shape) + 2)]
    N = len(Gl)
    # an index below -N (a leading axis of size 1) makes numba read before
    # the array, under a zero weight; it is taken modulo N here as the oracle does
    Gw = lambda x: Gl[x % N]  # noqa: E731
    dLdA = [[0j] * D for _ in range(D)]
    dLdb = [0j] * D
    dA = [[0j] * D for _ in range(D)]
    db = [0j] * D
    for flat, index in enumerate(itertools.product(*[range(s) for s in shape])):
        if flat == 0:
            continue
        for i in range(D):
            pivot = flat - st[i]
            db[i] = SQRT[index[i]] * Gw(pivot)
            dA[i][i] = 0.5 * SQRT[index[i]] * SQRT[index[i] - 1] * Gw(pivot - st[i]) if index[i] > 1 else 0j
            for j in range(i + 1, D):
                dA[i][j] = SQRT[index[i]] * SQRT[index[j]] * Gw(pivot - st[j])
        for i in range(D):
            for j in range(D):
                dLdA[i][j] += dA[i][j] * gl[flat]
            dLdb[i] += db[i] * gl[flat]
    s = 0j
    for x, y in zip(Gl, gl):
        s += x * y
    dc = s / complex(c)
    out = np.array([[(dLdA[i][j] + dLdA[j][i]) / 2 for j in range(D)] for i in range(D)], dtype=np.complex128)
    return out, np.array(dLdb, dtype=np.complex128), complex(dc)


def python_complex_stable(A, b, c, shape) -> np.ndarray:
    """stable_numba (core.py:164-213) in Python complex numbers on a list; the
    list starts at zero as the reference's G does, and negative indices wrap
    as numpy's do (modulo N where numba would read before the array)."""
    A = [[complex(x) for x in row] for row in np.asarray(A, dtype=np.complex128)]
    b = [complex(x) for x in np.asarray(b, dtype=np.complex128)]
    D = len(shape)
    st = _strides(shape)
    N = int(np.prod(shape))
    SQRT = [pymath.sqrt(n) for n in range(max(shape) + 2)]
    G = [0j] * N
    G[0] = complex(c)
    Gw = lambda x: G[x % N]  # noqa: E731
    for flat, nd in enumerate(itertools.product(*[range(s) for s in shape])):
        if flat == 0:
            continue
        num, vals = 0, 0
        for i in range(D):
            if nd[i] == 0:
                continue
            num += 1
            pivot = flat - st[i]
            val = b[i] * G[pivot]
            for j in range(i):
                val += A[i][j] * SQRT[nd[j]] * Gw(pivot - st[j])
            val += A[i][i] * SQRT[nd[i] - 1] * Gw(pivot - st[i])
            for j in range(i + 1, D):
                val += A[i][j] * SQRT[nd[j]] * Gw(pivot - st[j])
            vals += val / SQRT[nd[i]]
        G[flat] = vals / num
    return np.array(G, dtype=np.complex128).reshape(shape)
